"""Edge cases of the C ABI on the device: rejected arguments, grids beyond
the 32-bit sort keys, empty contexts and empty scenes inside a batch,
particles whose stencil leaves the grid, calls out of order."""
import numpy as np
import pytest

import mpm_amd as M
from helpers import rel_err
from mpm_amd import scenes

pytestmark = pytest.mark.gpu


def _raises(code, fn, *args, match=None):
    with pytest.raises(M.MpmError) as e:
        fn(*args)
    assert e.value.code == code, (e.value.code, str(e.value))
    if match:
        assert match in str(e.value), str(e.value)


def test_invalid_parameters_are_rejected():
    good = M.default_params((32, 32, 32), 1 / 32)
    _raises(M.MPM_ERR_ARG, M.Simulation, M.default_params((5, 32, 32), 1 / 32), 10)  # under 6 nodes
    _raises(M.MPM_ERR_ARG, M.Simulation, M.default_params((32, 32, 32), 0.0), 10)
    _raises(M.MPM_ERR_ARG, M.Simulation, M.default_params((32, 32, 32), -1.0), 10)
    _raises(M.MPM_ERR_ARG, M.Simulation, good, -1)
    other = M.default_params((32, 32, 32), 1 / 32, compose=M.COMPOSE_TEXTBOOK)
    _raises(M.MPM_ERR_ARG, M.Simulation, [good, other], [10, 10])  # one composition per batch


def test_grid_beyond_32_bit_keys_is_rejected_before_allocating():
    # 256 x 256 x 1025 blocks of 64 nodes: 2^32 + 2^22 nodes
    p = M.default_params((1024, 1024, 4100), 1 / 1024)
    _raises(M.MPM_ERR_ARG, M.Simulation, p, 16, match="32-bit sort keys")


def test_empty_context_steps_and_reads_back_nothing():
    g = M.Simulation(M.default_params((32, 32, 32), 1 / 32, bc_mode=M.BC_DOMAIN), 0)
    z = np.zeros((0, 3), np.float32)
    g.set_particles(z, z, np.zeros(0, np.float32))
    g.init_volumes()
    g.substeps(1e-4, 5)
    g.sync()
    out = g.particles()
    assert all(v.shape[0] == 0 for v in out.values())
    g.close()


def test_empty_scene_inside_a_batch():
    """A scene of zero particles steps with the others and changes nothing:
    the other scenes equal their single-scene runs."""
    a = scenes.snowball(seed=1, n=3000, grid=32, radius=0.08)
    b = scenes.snowball(seed=2, n=2000, grid=48, radius=0.07, v0=(0.5, -1.0, 0.0))

    def params(sc):
        return M.default_params(sc["dims"], sc["h"], bc_mode=M.BC_DOMAIN, theta_s=7.5e-3)

    empty = M.default_params((40, 40, 40), 1 / 40, bc_mode=M.BC_DOMAIN)
    batch = M.Simulation([params(a), empty, params(b)], [3000, 0, 2000])
    batch.scene_set_particles(0, a["x"], a["v"], a["m"], a["E0"])
    z = np.zeros((0, 3), np.float32)
    batch.scene_set_particles(1, z, z, np.zeros(0, np.float32))
    batch.scene_set_particles(2, b["x"], b["v"], b["m"], b["E0"])
    batch.init_volumes()
    batch.substeps(1e-4, 5)
    assert batch.scene_status() == [0, 0, 0]
    assert batch.scene_particles(1)["x"].shape == (0, 3)
    for i, sc in ((0, a), (2, b)):
        s = M.Simulation(params(sc), sc["x"].shape[0])
        s.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"])
        s.init_volumes()
        s.substeps(1e-4, 5)
        want, got = s.particles(), batch.scene_particles(i)
        for k in ("x", "v", "FE"):
            assert rel_err(got[k], want[k]) < 1e-5, (i, k, rel_err(got[k], want[k]))
        s.close()
    batch.close()


def test_stencil_outside_the_grid_is_rejected_with_the_particle():
    n = 100
    rng = np.random.default_rng(1)
    x = rng.uniform(0.3, 0.7, (n, 3)).astype(np.float32)
    x[37, 2] = 1.5 / 32  # base cell 1: the stencil would reach node -1
    g = M.Simulation(M.default_params((32, 32, 32), 1 / 32), n)
    _raises(M.MPM_ERR_BOUNDS, g.set_particles, x, np.zeros_like(x), np.ones(n, np.float32), match="particle 37")
    # nothing was uploaded: stepping is a state error until valid particles arrive
    _raises(M.MPM_ERR_STATE, g.substeps, 1e-4, 1)
    x[37, 2] = 0.5
    g.set_particles(x, np.zeros_like(x), np.ones(n, np.float32))
    g.init_volumes()
    g.substeps(1e-4, 2)
    g.close()


def test_call_order_and_step_counts():
    n = 64
    x = np.random.default_rng(2).uniform(0.4, 0.6, (n, 3)).astype(np.float32)
    g = M.Simulation(M.default_params((32, 32, 32), 1 / 32), n)
    _raises(M.MPM_ERR_STATE, g.substeps, 1e-4, 1)  # no particles yet
    g.set_particles(x, np.zeros_like(x), np.ones(n, np.float32))
    g.init_volumes()
    _raises(M.MPM_ERR_ARG, g.substeps, 1e-4, -1)
    before = g.particles()["x"]
    g.substeps(1e-4, 0)  # zero substeps change nothing
    np.testing.assert_array_equal(g.particles()["x"], before)
    g.close()
