"""The snapshot format's numpy side (mpm_amd.snapshot): write -> read round
trip of every field, the documented header and record sizes, defaults for
missing state, and rejection of a truncated or foreign file.  The library's
own writer and reader are checked against this module on the GPU."""
import numpy as np
import pytest

import mpm_amd as M
from mpm_amd import snapshot as S


def _scene(rng, n, dims=(32, 32, 32), h=1 / 32):
    p = M.default_params(dims, h, bc_mode=M.BC_DOMAIN | M.BC_COLLIDERS, theta_s=7.5e-3)
    return dict(params=p, x=rng.uniform(0.2, 0.8, (n, 3)).astype(np.float32),
                v=rng.normal(size=(n, 3)).astype(np.float32), B=rng.normal(size=(n, 9)).astype(np.float32),
                FE=rng.normal(size=(n, 9)).astype(np.float32), FP=rng.normal(size=(n, 9)).astype(np.float32),
                m=rng.uniform(1, 2, n).astype(np.float32), V0=rng.uniform(1, 2, n).astype(np.float32),
                E0=rng.uniform(1e5, 2e5, n).astype(np.float32),
                colliders=[M.collider((0.5, 0.1, 0.5), (0.92, 0, 0, 0.38), (0.3, 0.05, 0.3), (0, 1, 0))])


def test_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    scs = [_scene(rng, 100), _scene(rng, 0), _scene(rng, 7, dims=(16, 24, 20), h=0.05)]
    p = tmp_path / "s.snap"
    S.write(p, scs, step=1234)
    back = S.read(p)
    assert back["step"] == 1234 and len(back["scenes"]) == 3
    size = p.stat().st_size
    assert size == 64 + 3 * 88 + 3 * 52 + 4 * 36 * 107
    for a, b in zip(scs, back["scenes"]):
        assert b["n"] == a["x"].shape[0]
        for k, _ in S.BLOCKS:
            np.testing.assert_array_equal(b[k].reshape(a[k].shape), a[k])
        assert b["params"]["dims"] == list(a["params"].dims)
        assert b["params"]["theta_s"] == a["params"].theta_s
        assert b["colliders"].shape == (1, 13) and b["colliders"][0, 3] == np.float32(0.92)
        q = S.params_struct(b["params"])
        assert S.params_dict(q) == b["params"]


def test_defaults_and_rejections(tmp_path):
    p = M.default_params((32, 32, 32), 1 / 32)
    x = np.full((5, 3), 0.5, np.float32)
    path = tmp_path / "d.snap"
    S.write(path, [dict(params=p, x=x, v=np.zeros_like(x), m=np.ones(5, np.float32))])
    sc = S.read(path)["scenes"][0]
    np.testing.assert_array_equal(sc["FE"], np.tile(np.float32([1, 0, 0, 0, 1, 0, 0, 0, 1]), (5, 1)))
    np.testing.assert_array_equal(sc["B"], 0)
    np.testing.assert_array_equal(sc["E0"], np.float32(p.E))
    raw = path.read_bytes()
    (tmp_path / "t.snap").write_bytes(raw[:-4])
    with pytest.raises(ValueError):
        S.read(tmp_path / "t.snap")
    (tmp_path / "f.snap").write_bytes(b"NOTASNAP" + raw[8:])
    with pytest.raises(ValueError):
        S.read(tmp_path / "f.snap")
