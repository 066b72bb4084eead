"""Snapshot files in numpy: read and write the format mpm_snapshot_save and
mpm_snapshot_load use (layout in csrc/mpm_io.hip), so fixtures, checkpoints
on disk and resumed runs share one format.

    read(path)  -> {"step", "scenes": [{"params": {...}, "n", "colliders":
                   (k, 13) f32, "x", "v", "B", "FE", "FP", "m", "V0", "E0"}]}
    write(path, scenes, step=0)   scenes in the same form
"""
from __future__ import annotations

import numpy as np

MAGIC = b"MPMSNAP1"
VERSION = 1
NFIELDS = 36
RECORD = 88

HEADER = np.dtype([("magic", "S8"), ("version", "<u4"), ("nscenes", "<u4"), ("step", "<i8"), ("nfields", "<u4"),
                   ("record", "<u4"), ("total", "<u8"), ("zero", "u1", 24)])
SCENE = np.dtype([("dims", "<i4", 3), ("compose", "<i4"), ("bc_mode", "<i4"), ("flags", "<i4"), ("h", "<f4"),
                  ("E", "<f4"), ("nu", "<f4"), ("xi", "<f4"), ("theta_c", "<f8"), ("theta_s", "<f8"),
                  ("gravity", "<f4", 3), ("svd_gap_fp64", "<f4"), ("resort_every", "<i4"), ("n", "<i4"),
                  ("ncoll", "<i4"), ("pad", "<i4")])
assert HEADER.itemsize == 64 and SCENE.itemsize == RECORD
PARAM_KEYS = ("dims", "compose", "bc_mode", "flags", "h", "E", "nu", "xi", "theta_c", "theta_s", "gravity",
              "svd_gap_fp64", "resort_every")
# particle blocks in file order: name, components
BLOCKS = (("x", 3), ("v", 3), ("B", 9), ("FE", 9), ("FP", 9), ("m", 1), ("V0", 1), ("E0", 1))


def params_dict(p):
    """mpm_amd.Params (ctypes) -> plain dict of the record's fields."""
    out = {}
    for k in PARAM_KEYS:
        v = getattr(p, k)
        out[k] = list(v) if k in ("dims", "gravity") else v
    return out


def params_struct(d):
    """dict from read() -> mpm_amd.Params."""
    from . import Params

    p = Params()
    for k in PARAM_KEYS:
        if k in ("dims", "gravity"):
            getattr(p, k)[:] = [type(getattr(p, k)[0])(v) for v in d[k]]
        else:
            setattr(p, k, d[k])
    return p


def read(path, particles=True):
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(HEADER.itemsize), HEADER)
        if hdr.shape[0] != 1 or hdr["magic"][0] != MAGIC:
            raise ValueError(f"{path}: not a snapshot file")
        hdr = hdr[0]
        if hdr["version"] != VERSION or hdr["nfields"] != NFIELDS or hdr["record"] != RECORD:
            raise ValueError(f"{path}: unsupported snapshot version or layout")
        ns = int(hdr["nscenes"])
        recs = np.frombuffer(f.read(RECORD * ns), SCENE)
        if recs.shape[0] != ns:
            raise ValueError(f"{path}: truncated scene records")
        scenes = []
        for r in recs:
            prm = {k: (r[k].tolist() if k in ("dims", "gravity") else r[k].item()) for k in PARAM_KEYS}
            scenes.append(dict(params=prm, n=int(r["n"]), ncoll=int(r["ncoll"])))
        for sc in scenes:
            k = sc.pop("ncoll")
            sc["colliders"] = np.frombuffer(f.read(52 * k), "<f4").reshape(k, 13).copy()
        if particles:
            for sc in scenes:
                n = sc["n"]
                for name, c in BLOCKS:
                    a = np.frombuffer(f.read(4 * n * c), "<f4")
                    if a.shape[0] != n * c:
                        raise ValueError(f"{path}: truncated particle data")
                    sc[name] = a.reshape(n, c).copy() if c > 1 else a.copy()
    return dict(step=int(hdr["step"]), scenes=scenes)


def write(path, scenes, step=0):
    """scenes: list of dicts with params (dict or mpm_amd.Params), the
    particle arrays of BLOCKS (B, FE, FP default to 0, I, I; V0 to 0; E0 to
    params E) and optional colliders ((k, 13) or mpm_amd.Collider list)."""
    recs = np.zeros(len(scenes), SCENE)
    colls, blocks = [], []
    for i, sc in enumerate(scenes):
        prm = sc["params"] if isinstance(sc["params"], dict) else params_dict(sc["params"])
        n = int(np.asarray(sc["x"]).shape[0])
        for k in PARAM_KEYS:
            recs[i][k] = prm[k]
        recs[i]["n"] = n
        c = sc.get("colliders", np.zeros((0, 13), np.float32))
        if not isinstance(c, np.ndarray):
            c = np.array([list(o.translation) + list(o.rotation) + list(o.scale) + list(o.velocity) for o in c],
                         np.float32).reshape(-1, 13)
        recs[i]["ncoll"] = c.shape[0]
        colls.append(np.ascontiguousarray(c, "<f4"))
        eye = np.tile(np.float32([1, 0, 0, 0, 1, 0, 0, 0, 1]), (n, 1))
        defaults = dict(B=np.zeros((n, 9), np.float32), FE=eye, FP=eye, V0=np.zeros(n, np.float32),
                        E0=np.full(n, prm["E"], np.float32))
        for name, comps in BLOCKS:
            a = sc.get(name, defaults.get(name))
            a = np.ascontiguousarray(a, "<f4").reshape(n, comps) if comps > 1 else np.ascontiguousarray(a, "<f4")
            if a.shape[0] != n:
                raise ValueError(f"scene {i}: {name} has {a.shape[0]} rows, expected {n}")
            blocks.append(a)
    hdr = np.zeros(1, HEADER)
    hdr["magic"], hdr["version"], hdr["nscenes"], hdr["step"] = MAGIC, VERSION, len(scenes), step
    hdr["nfields"], hdr["record"], hdr["total"] = NFIELDS, RECORD, int(recs["n"].sum())
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(recs.tobytes())
        for c in colls:
            f.write(c.tobytes())
        for a in blocks:
            f.write(a.tobytes())
