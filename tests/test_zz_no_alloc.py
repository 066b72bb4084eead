"""mpm_scene_get_particles_noalloc, the readback that allocates no device
memory (it stages through a file-scope device array), called through the C
ABI and held bit for bit to mpm_scene_get_particles, the entry point it
varies, on two calls in a row, the second with a smaller scene; a scene
beyond the array's capacity is rejected with MPM_ERR_ARG."""
import ctypes as C

import numpy as np
import pytest

import mpm_amd as M
from mpm_amd import scenes

pytestmark = pytest.mark.gpu

_F = C.POINTER(C.c_float)


def _fields(n):
    return [np.full((n, k), np.nan, np.float32) for k in (3, 3, 9, 9, 9, 1)]


def _read(L, ctx, scene, n, fn):
    out = _fields(n)
    rc = fn(ctx, scene, *[a.ctypes.data_as(_F) for a in out])
    return rc, out


def test_noalloc_readback_equals_scene_readback():
    L = M.lib()
    big = scenes.snowball(seed=41, n=M.NOALLOC_MAX_PARTICLES, grid=128, radius=0.12, v0=(0.5, -1.0, 0.2))
    small = scenes.snowball(seed=42, n=5000, grid=64, radius=0.05, v0=(-1.0, 0.0, 0.0))
    over = scenes.snowball(seed=43, n=M.NOALLOC_MAX_PARTICLES + 1, grid=128, radius=0.12)
    scs = [big, small, over]
    prm = [M.default_params(s["dims"], s["h"], bc_mode=M.BC_DOMAIN, theta_s=7.5e-3) for s in scs]
    b = M.Simulation(prm, [s["x"].shape[0] for s in scs])
    for i, s in enumerate(scs):
        b.scene_set_particles(i, s["x"], s["v"], s["m"], s["E0"])
    b.init_volumes()
    b.substeps(1e-4, 5)  # binned, reordered, B and F off their initial values
    b.sync()
    for scene in (0, 1):  # the second call reads a smaller scene
        n = scs[scene]["x"].shape[0]
        rc_ref, ref = _read(L, b.ctx, scene, n, L.mpm_scene_get_particles)
        rc, got = _read(L, b.ctx, scene, n, L.mpm_scene_get_particles_noalloc)
        assert rc_ref == 0 and rc == 0, (rc_ref, rc)
        for g, r in zip(got, ref):
            np.testing.assert_array_equal(g, r)
        assert not np.isnan(got[4]).any()
    # NULL fields are skipped
    n = scs[1]["x"].shape[0]
    x = np.zeros((n, 3), np.float32)
    assert L.mpm_scene_get_particles_noalloc(b.ctx, 1, x.ctypes.data_as(_F), None, None, None, None, None) == 0
    np.testing.assert_array_equal(x, b.scene_particles(1)["x"])
    # beyond the staging array: a clear error, nothing written
    n = scs[2]["x"].shape[0]
    rc, got = _read(L, b.ctx, 2, n, L.mpm_scene_get_particles_noalloc)
    assert rc == M.MPM_ERR_ARG
    assert "MPM_NOALLOC_MAX_PARTICLES" in L.mpm_last_error(b.ctx).decode()
    assert np.isnan(got[0]).all()
    b.close()
