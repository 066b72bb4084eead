"""Python view of libmpm_hip.so through its C ABI (include/mpm_c.h).

Used by the tests, the benchmark and the autograd wrapper.  It loads the
in-tree library and fails loudly when it is missing: there is no CPU
fallback.  Matrices are 9 floats per particle in glm column-major order
(element (row r, col c) at index 3c + r), as in the reference's glm::mat3.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libmpm_hip.so")

MPM_OK = 0
MPM_ERR_ARG, MPM_ERR_HIP, MPM_ERR_BOUNDS, MPM_ERR_NONFINITE, MPM_ERR_STATE, MPM_ERR_NOMEM = 1, 2, 3, 4, 5, 6
MPM_ERR_CFL = 7  # warning: reported by scene_status only
NOALLOC_MAX_PARTICLES = 262144  # MPM_NOALLOC_MAX_PARTICLES
ERRORS = {1: "MPM_ERR_ARG", 2: "MPM_ERR_HIP", 3: "MPM_ERR_BOUNDS", 4: "MPM_ERR_NONFINITE", 5: "MPM_ERR_STATE",
          6: "MPM_ERR_NOMEM", 7: "MPM_ERR_CFL"}
COMPOSE_LITERAL, COMPOSE_TEXTBOOK = 0, 1
BC_COLLIDERS, BC_DOMAIN = 1, 2
STAGE_P2G, STAGE_VOLUMES, STAGE_FORCES, STAGE_GRID, STAGE_COLLIDE, STAGE_FUPDATE, STAGE_G2P, STAGE_ADVECT = range(8)


class MpmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("dims", C.c_int32 * 3),
        ("compose", C.c_int32),
        ("bc_mode", C.c_int32),
        ("flags", C.c_int32),
        ("h", C.c_float),
        ("E", C.c_float),
        ("nu", C.c_float),
        ("xi", C.c_float),
        ("theta_c", C.c_double),
        ("theta_s", C.c_double),
        ("gravity", C.c_float * 3),
        ("svd_gap_fp64", C.c_float),
        ("resort_every", C.c_int32),
    ]


class Collider(C.Structure):
    _fields_ = [
        ("translation", C.c_float * 3),
        ("rotation", C.c_float * 4),
        ("scale", C.c_float * 3),
        ("velocity", C.c_float * 3),
    ]


class Generator(C.Structure):
    """mpm_generator: seeded device placement of a scene's particles."""
    _fields_ = [
        ("kind", C.c_int32),
        ("reserved", C.c_int32),
        ("seed", C.c_uint64),
        ("centre", C.c_float * 3),
        ("radius", C.c_float),
        ("cell_lo", C.c_int32 * 3),
        ("cells", C.c_int32 * 3),
        ("ppc_lo", C.c_int32),
        ("ppc_hi", C.c_int32),
        ("velocity", C.c_float * 3),
        ("mass", C.c_float),
        ("E_lo", C.c_float),
        ("E_hi", C.c_float),
    ]


GEN_BALL, GEN_CELLS, GEN_LATTICE = 0, 1, 2

_F = C.POINTER(C.c_float)
_lib = None

# Every symbol include/mpm_c.h declares (checked by the CPU tests).
EXPORTS = [
    "mpm_default_params", "mpm_create", "mpm_destroy", "mpm_last_error", "mpm_num_particles",
    "mpm_set_particles", "mpm_set_state", "mpm_set_colliders", "mpm_init_volumes", "mpm_stage",
    "mpm_substeps", "mpm_sync", "mpm_get_particles", "mpm_get_grid", "mpm_profile_begin", "mpm_profile_end",
    "mpm_count_active", "mpm_debug_bin", "mpm_debug_blocks", "mpm_copy_probe", "mpm_use_graphs",
    "mpm_set_reproducible", "mpm_get_reproducible",
    "mpm_create_batch", "mpm_num_scenes", "mpm_scene_set_particles", "mpm_scene_set_state",
    "mpm_scene_set_colliders", "mpm_scene_get_particles", "mpm_scene_get_grid", "mpm_scene_status",
    "mpm_set_tuning", "mpm_get_tuning", "mpm_autotune", "mpm_adjoint_forward", "mpm_adjoint_interval", "mpm_adjoint_backward",
    "mpm_scene_generate", "mpm_generate", "mpm_scene_get_material", "mpm_get_positions", "mpm_positions_begin",
    "mpm_positions_wait", "mpm_snapshot_save", "mpm_snapshot_load", "mpm_create_from_snapshot",
    "mpm_scene_get_particles_noalloc",
]
KERNELS = {"p2g": 0, "grid": 1, "g2p": 2, "bin": 3, "fused": 4}
NKERNELS = 8


def lib():
    """Load libmpm_hip.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with realtime-deformations_amd/build.py")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.mpm_default_params.argtypes = [C.POINTER(Params), C.c_int32, C.c_int32, C.c_int32, C.c_float]
        L.mpm_create.argtypes = [C.POINTER(Params), C.c_int32, C.POINTER(vp)]
        L.mpm_destroy.argtypes = [vp]
        L.mpm_last_error.argtypes = [vp]
        L.mpm_last_error.restype = C.c_char_p
        L.mpm_num_particles.argtypes = [vp]
        L.mpm_set_particles.argtypes = [vp, _F, _F, _F, _F]
        L.mpm_set_state.argtypes = [vp, _F, _F, _F, _F]
        L.mpm_set_colliders.argtypes = [vp, C.POINTER(Collider), C.c_int32]
        L.mpm_init_volumes.argtypes = [vp]
        L.mpm_stage.argtypes = [vp, C.c_int32, C.c_float]
        L.mpm_substeps.argtypes = [vp, C.c_float, C.c_int32]
        L.mpm_sync.argtypes = [vp]
        L.mpm_get_particles.argtypes = [vp] + [_F] * 6
        L.mpm_get_grid.argtypes = [vp, _F, _F, _F, C.POINTER(C.c_uint8)]
        L.mpm_profile_begin.argtypes = [vp, C.c_int32]
        L.mpm_use_graphs.argtypes = [vp, C.c_int32]
        L.mpm_set_reproducible.argtypes = [vp, C.c_int32]
        L.mpm_get_reproducible.argtypes = [vp, C.POINTER(C.c_int32)]
        L.mpm_create_batch.argtypes = [C.POINTER(Params), C.POINTER(C.c_int32), C.c_int32, C.POINTER(vp)]
        L.mpm_num_scenes.argtypes = [vp]
        L.mpm_scene_set_particles.argtypes = [vp, C.c_int32, _F, _F, _F, _F]
        L.mpm_scene_set_state.argtypes = [vp, C.c_int32, _F, _F, _F, _F]
        L.mpm_scene_set_colliders.argtypes = [vp, C.c_int32, C.POINTER(Collider), C.c_int32]
        L.mpm_scene_get_particles.argtypes = [vp, C.c_int32] + [_F] * 6
        L.mpm_scene_get_particles_noalloc.argtypes = [vp, C.c_int32] + [_F] * 6
        L.mpm_scene_get_grid.argtypes = [vp, C.c_int32, _F, _F, _F, C.POINTER(C.c_uint8)]
        L.mpm_scene_status.argtypes = [vp, C.POINTER(C.c_int32)]
        L.mpm_set_tuning.argtypes = [vp, C.c_int32, C.c_int32]
        L.mpm_get_tuning.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.mpm_autotune.argtypes = [vp, C.c_float, C.c_int32, C.POINTER(C.c_double)]
        L.mpm_adjoint_forward.argtypes = [vp, C.c_float, C.c_int32, C.c_int32]
        L.mpm_adjoint_interval.argtypes = [vp, C.POINTER(C.c_int32)]
        L.mpm_adjoint_backward.argtypes = [vp, _F, _F, _F, _F, _F]
        L.mpm_profile_end.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_double)]
        L.mpm_count_active.argtypes = [vp, C.POINTER(C.c_int64)]
        L.mpm_debug_bin.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_int32)]
        L.mpm_debug_blocks.argtypes = [vp, C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_int32)]
        L.mpm_copy_probe.argtypes = [vp, C.c_int64]
        L.mpm_scene_generate.argtypes = [vp, C.c_int32, C.POINTER(Generator), C.POINTER(C.c_int64)]
        L.mpm_generate.argtypes = [vp, C.POINTER(Generator), C.POINTER(C.c_int64)]
        L.mpm_scene_get_material.argtypes = [vp, C.c_int32, _F, _F]
        L.mpm_get_positions.argtypes = [vp, _F]
        L.mpm_positions_begin.argtypes = [vp]
        L.mpm_positions_wait.argtypes = [vp, C.POINTER(_F)]
        L.mpm_snapshot_save.argtypes = [vp, C.c_char_p, C.c_int64]
        L.mpm_snapshot_load.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int64)]
        L.mpm_create_from_snapshot.argtypes = [C.c_char_p, C.POINTER(vp), C.POINTER(C.c_int64)]
        _lib = L
    return _lib


def _fp(a):
    return None if a is None else a.ctypes.data_as(_F)


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def default_params(dims, h, **kw) -> Params:
    p = Params()
    lib().mpm_default_params(C.byref(p), int(dims[0]), int(dims[1]), int(dims[2]), float(h))
    for k, v in kw.items():
        if k == "gravity":
            p.gravity[:] = list(v)
        else:
            setattr(p, k, v)
    return p


def collider(translation, rotation=(1.0, 0.0, 0.0, 0.0), scale=(1.0, 1.0, 1.0), velocity=(0.0, 0.0, 0.0)):
    c = Collider()
    c.translation[:] = list(translation)
    c.rotation[:] = list(rotation)
    c.scale[:] = list(scale)
    c.velocity[:] = list(velocity)
    return c


class Simulation:
    """One device context (mpm_ctx): one scene, or a batch of scenes when
    `params` and `n` are sequences (scene s: params[s], n[s] particles)."""

    def __init__(self, params, n):
        self.L = lib()
        h = C.c_void_p()
        if isinstance(params, (list, tuple)):
            self.scene_params = list(params)
            self.scene_n = [int(k) for k in n]
            arr = (Params * len(params))(*params)
            cnt = (C.c_int32 * len(params))(*self.scene_n)
            self._check(self.L.mpm_create_batch(arr, cnt, len(params), C.byref(h)), h)
        else:
            self.scene_params = [params]
            self.scene_n = [int(n)]
            self._check(self.L.mpm_create(C.byref(params), int(n), C.byref(h)), h)
        self.ctx = h
        self.n = sum(self.scene_n)
        self.params = self.scene_params[0]
        self.dims = tuple(self.params.dims)
        self.offsets = np.r_[0, np.cumsum(self.scene_n)].astype(np.int64)

    @classmethod
    def from_snapshot(cls, path):
        """-> (Simulation, step): a context with the snapshot's scenes and state."""
        from . import snapshot as _snap

        meta = _snap.read(path, particles=False)
        self = cls.__new__(cls)
        self.L = lib()
        h = C.c_void_p()
        step = C.c_int64()
        self._check(self.L.mpm_create_from_snapshot(os.fsencode(path), C.byref(h), C.byref(step)), h)
        self.ctx = h
        self.scene_params = [_snap.params_struct(sc["params"]) for sc in meta["scenes"]]
        self.scene_n = [sc["n"] for sc in meta["scenes"]]
        self.n = sum(self.scene_n)
        self.params = self.scene_params[0]
        self.dims = tuple(self.params.dims)
        self.offsets = np.r_[0, np.cumsum(self.scene_n)].astype(np.int64)
        return self, step.value

    def _check(self, rc, h=None):
        """Raise MpmError for a failed call.  h: the handle of a failed create,
        destroyed here once its message is read."""
        if rc != MPM_OK:
            ctx = h if h is not None else self.ctx
            msg = self.L.mpm_last_error(ctx).decode() if ctx is not None and ctx.value else ""
            if h is not None and h.value:
                self.L.mpm_destroy(h)
                h.value = None
            raise MpmError(rc, msg)

    def close(self):
        if getattr(self, "ctx", None) is not None and self.ctx.value:
            self.L.mpm_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_particles(self, x, v, m, E0=None):
        x, v, m, E0 = _f32(x), _f32(v), _f32(m), _f32(E0)
        self._check(self.L.mpm_set_particles(self.ctx, _fp(x), _fp(v), _fp(m), _fp(E0)))

    def set_state(self, B=None, FE=None, FP=None, V0=None):
        B, FE, FP, V0 = _f32(B), _f32(FE), _f32(FP), _f32(V0)
        self._check(self.L.mpm_set_state(self.ctx, _fp(B), _fp(FE), _fp(FP), _fp(V0)))

    def set_colliders(self, colliders):
        arr = (Collider * max(1, len(colliders)))(*colliders)
        self._check(self.L.mpm_set_colliders(self.ctx, arr, len(colliders)))

    def init_volumes(self):
        self._check(self.L.mpm_init_volumes(self.ctx))

    def stage(self, st, dt=0.0):
        self._check(self.L.mpm_stage(self.ctx, int(st), float(dt)))

    def substeps(self, dt, n):
        self._check(self.L.mpm_substeps(self.ctx, float(dt), int(n)))

    def sync(self):
        self._check(self.L.mpm_sync(self.ctx))

    def use_graphs(self, enable=True):
        """Replay runs of substeps from captured HIP graphs (the default)."""
        self._check(self.L.mpm_use_graphs(self.ctx, int(bool(enable))))

    def set_reproducible(self, enable=True):
        """Reproducible mode: P2G contributions rounded to per-scene quanta,
        so the grid sums, and every result, repeat bit for bit whatever the
        order of the atomics, the binning interval or the batch."""
        self._check(self.L.mpm_set_reproducible(self.ctx, int(bool(enable))))

    @property
    def reproducible(self):
        out = C.c_int32()
        self._check(self.L.mpm_get_reproducible(self.ctx, C.byref(out)))
        return bool(out.value)

    def profile_begin(self, per_kernel=True):
        """Start a profiled region.  per_kernel brackets every launch with HIP
        events (and launches directly, bypassing the substep graphs)."""
        self._check(self.L.mpm_profile_begin(self.ctx, int(bool(per_kernel))))

    def profile_end(self):
        """-> (total_ms, {kernel: (ms_sum, launches)})"""
        ms = (C.c_double * NKERNELS)()
        cnt = (C.c_int32 * NKERNELS)()
        tot = C.c_double()
        self._check(self.L.mpm_profile_end(self.ctx, ms, cnt, C.byref(tot)))
        return tot.value, {k: (ms[i], cnt[i]) for k, i in KERNELS.items()}

    def copy_probe(self, nbytes):
        self._check(self.L.mpm_copy_probe(self.ctx, int(nbytes)))

    def count_active(self):
        v = C.c_int64()
        self._check(self.L.mpm_count_active(self.ctx, C.byref(v)))
        return v.value

    def debug_bin(self):
        """-> (sorted keys, stable permutation, caller index per slot after reorder)"""
        n = self.n
        k, p, s = np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.zeros(n, np.int32)
        self._check(self.L.mpm_debug_bin(self.ctx, k.ctypes.data_as(C.POINTER(C.c_uint32)),
                                         p.ctypes.data_as(C.POINTER(C.c_uint32)),
                                         s.ctypes.data_as(C.POINTER(C.c_int32))))
        return k, p, s

    def debug_blocks(self):
        """-> (nblocks, 4) int32: home blocks of the last binning in claim
        order (grid block, first slot, end slot, scene)"""
        cnt = C.c_int32(0)
        self._check(self.L.mpm_debug_blocks(self.ctx, None, 0, C.byref(cnt)))
        out = np.zeros((cnt.value, 4), np.int32)
        if cnt.value:
            self._check(self.L.mpm_debug_blocks(self.ctx, out.ctypes.data_as(C.POINTER(C.c_int32)), cnt.value,
                                                C.byref(cnt)))
        return out

    def particles(self):
        n = self.n
        out = dict(x=np.zeros((n, 3), np.float32), v=np.zeros((n, 3), np.float32), B=np.zeros((n, 9), np.float32),
                   FE=np.zeros((n, 9), np.float32), FP=np.zeros((n, 9), np.float32), V0=np.zeros(n, np.float32))
        self._check(self.L.mpm_get_particles(self.ctx, *[_fp(out[k]) for k in ("x", "v", "B", "FE", "FP", "V0")]))
        return out

    # ---- device generators, readback, snapshots
    def generate(self, kind, seed=0, scene=0, centre=(0.5, 0.5, 0.5), radius=0.1, cell_lo=(0, 0, 0),
                 cells=(1, 1, 1), ppc=(4, 12), velocity=(0.0, 0.0, 0.0), mass=1e-4, E=(1.4e5, 1.4e5)):
        """Seeded placement of a scene's particles on the device (GEN_BALL,
        GEN_CELLS, GEN_LATTICE); -> candidates kept (LATTICE) or n."""
        g = Generator()
        g.kind, g.seed = int(kind), int(seed)
        g.centre[:] = [float(c) for c in centre]
        g.radius = float(radius)
        g.cell_lo[:] = [int(c) for c in cell_lo]
        g.cells[:] = [int(c) for c in cells]
        g.ppc_lo, g.ppc_hi = int(ppc[0]), int(ppc[1])
        g.velocity[:] = [float(c) for c in velocity]
        g.mass = float(mass)
        g.E_lo, g.E_hi = float(E[0]), float(E[1])
        kept = C.c_int64()
        self._check(self.L.mpm_scene_generate(self.ctx, int(scene), C.byref(g), C.byref(kept)))
        return kept.value

    def scene_material(self, scene=0):
        """-> (m, E0) of one scene in caller order."""
        n = self.scene_n[scene]
        m, E0 = np.zeros(n, np.float32), np.zeros(n, np.float32)
        self._check(self.L.mpm_scene_get_material(self.ctx, int(scene), _fp(m), _fp(E0)))
        return m, E0

    def positions(self):
        """All positions (n, 3) in caller order through the pinned mirror."""
        x = np.zeros((self.n, 3), np.float32)
        self._check(self.L.mpm_get_positions(self.ctx, _fp(x)))
        return x

    def positions_begin(self):
        """Enqueue a positions readback that overlaps later substeps."""
        self._check(self.L.mpm_positions_begin(self.ctx))

    def positions_wait(self):
        """-> (n, 3) view of the pinned mirror (valid until the next begin)."""
        p = _F()
        self._check(self.L.mpm_positions_wait(self.ctx, C.byref(p)))
        return np.ctypeslib.as_array(p, shape=(self.n, 3))

    def save(self, path, step=0):
        self._check(self.L.mpm_snapshot_save(self.ctx, os.fsencode(path), int(step)))

    def load(self, path):
        """Replace the state from a snapshot of the same scenes; -> step."""
        step = C.c_int64()
        self._check(self.L.mpm_snapshot_load(self.ctx, os.fsencode(path), C.byref(step)))
        return step.value

    # ---- batched scenes
    @property
    def nscenes(self):
        return len(self.scene_n)

    def scene_set_particles(self, scene, x, v, m, E0=None):
        x, v, m, E0 = _f32(x), _f32(v), _f32(m), _f32(E0)
        self._check(self.L.mpm_scene_set_particles(self.ctx, int(scene), _fp(x), _fp(v), _fp(m), _fp(E0)))

    def scene_set_colliders(self, scene, colliders):
        arr = (Collider * max(1, len(colliders)))(*colliders)
        self._check(self.L.mpm_scene_set_colliders(self.ctx, int(scene), arr, len(colliders)))

    def scene_particles(self, scene):
        n = self.scene_n[scene]
        out = dict(x=np.zeros((n, 3), np.float32), v=np.zeros((n, 3), np.float32), B=np.zeros((n, 9), np.float32),
                   FE=np.zeros((n, 9), np.float32), FP=np.zeros((n, 9), np.float32), V0=np.zeros(n, np.float32))
        self._check(self.L.mpm_scene_get_particles(self.ctx, int(scene),
                                                   *[_fp(out[k]) for k in ("x", "v", "B", "FE", "FP", "V0")]))
        return out

    def scene_status(self):
        """Per-scene MPM_OK / MPM_ERR_* since the last status read
        (MPM_ERR_NONFINITE, then MPM_ERR_BOUNDS, then the MPM_ERR_CFL
        warning); clears every flag."""
        st = (C.c_int32 * self.nscenes)()
        self._check(self.L.mpm_scene_status(self.ctx, st))
        return list(st)

    def set_tuning(self, resort_every, block_grid):
        self._check(self.L.mpm_set_tuning(self.ctx, int(resort_every), int(block_grid)))

    def tuning(self):
        r, g = C.c_int32(), C.c_int32()
        self._check(self.L.mpm_get_tuning(self.ctx, C.byref(r), C.byref(g)))
        return r.value, g.value

    def autotune(self, dt, steps=16):
        """-> best ms per substep; keeps the fastest (resort_every, block_grid)."""
        ms = C.c_double()
        self._check(self.L.mpm_autotune(self.ctx, float(dt), int(steps), C.byref(ms)))
        return ms.value

    # ---- reverse mode
    def adjoint_forward(self, dt, nsteps, checkpoint_every=0):
        """nsteps substeps from the current state, keeping checkpoints every
        `checkpoint_every` substeps (1: every substep's record; 0: the library
        chooses 1 when the records fit in device memory, else 20)."""
        self._check(self.L.mpm_adjoint_forward(self.ctx, float(dt), int(nsteps), int(checkpoint_every)))
        # a context holds one reverse-mode record at a time: each forward
        # replaces it, and callers that keep a record for a later backward
        # (autograd.Rollout) compare this counter to detect that
        self.adjoint_generation = getattr(self, "adjoint_generation", 0) + 1

    def adjoint_interval(self):
        """Checkpoint interval the last adjoint_forward used (0 before any)."""
        e = C.c_int32()
        self._check(self.L.mpm_adjoint_interval(self.ctx, C.byref(e)))
        return e.value

    def adjoint_backward(self, gx=None, gv=None):
        """dL/dx, dL/dv of the final particles -> (dL/dx0, dL/dv0, dL/dE0);
        leaves the context at the initial state."""
        n = self.n
        gx, gv = _f32(gx), _f32(gv)
        gx0, gv0, gE0 = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32), np.zeros(n, np.float32)
        self._check(self.L.mpm_adjoint_backward(self.ctx, _fp(gx), _fp(gv), _fp(gx0), _fp(gv0), _fp(gE0)))
        return gx0, gv0, gE0

    def scene_grid(self, scene, force=False):
        """Dense grid of one scene: m, vel (3), active (and force, stage
        path only)."""
        cells = int(np.prod(tuple(self.scene_params[scene].dims)))
        m, vel = np.zeros(cells, np.float32), np.zeros((cells, 3), np.float32)
        f = np.zeros((cells, 3), np.float32) if force else None
        act = np.zeros(cells, np.uint8)
        self._check(self.L.mpm_scene_get_grid(self.ctx, int(scene), _fp(m), _fp(vel), _fp(f),
                                              act.ctypes.data_as(C.POINTER(C.c_uint8))))
        out = dict(m=m, vel=vel, active=act.astype(bool))
        if force:
            out["force"] = f
        return out

    def grid(self, force=False):
        cells = int(np.prod(self.dims))
        m, vel = np.zeros(cells, np.float32), np.zeros((cells, 3), np.float32)
        f = np.zeros((cells, 3), np.float32) if force else None
        act = np.zeros(cells, np.uint8)
        self._check(self.L.mpm_get_grid(self.ctx, _fp(m), _fp(vel), _fp(f), act.ctypes.data_as(C.POINTER(C.c_uint8))))
        out = dict(m=m, vel=vel, active=act.astype(bool))
        if force:
            out["force"] = f
        return out
