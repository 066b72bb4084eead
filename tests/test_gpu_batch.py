"""Batched scenes: one context stepping several independent scenes (each
with its own grid, spacing, material, velocity and colliders) against the
same scenes run one by one in single-scene contexts, and against the oracle.

Batched and single-scene runs execute the same kernels, but the grid
accumulation order of float atomics differs between launches, so results
agree to rounding (1e-6 after one substep, 1e-5 after 20), not bit for bit.
"""
import numpy as np
import pytest

import mpm_amd as M
import oracle as O
from helpers import rel_err
from mpm_amd import scenes

pytestmark = pytest.mark.gpu


def _params(sc, **kw):
    prm = dict(sc["params"])
    prm.update(kw)
    return M.default_params(sc["dims"], sc["h"], compose=prm.pop("compose", 0), bc_mode=prm["bc_mode"],
                            E=prm["E"], nu=prm["nu"], xi=prm["xi"], theta_c=prm["theta_c"],
                            theta_s=prm["theta_s"], resort_every=prm.get("resort_every", 0))


def _mixed_batch():
    out = scenes.scene_batch(seed=11, nscenes=5, n_range=(2000, 40000), split=12000, small_grid=32,
                             large_grid=64, dt=1e-4)
    out.append(scenes.snowball(seed=4, n=3000, grid=48, radius=0.06, v0=(1.0, -2.0, 0.5)))
    return out


def _run_batch(scs, dt, steps, colliders=None, sync=True):
    b = M.Simulation([_params(s) for s in scs], [s["x"].shape[0] for s in scs])
    for i, s in enumerate(scs):
        b.scene_set_particles(i, s["x"], s["v"], s["m"], s["E0"])
    if colliders:
        for i, c in colliders.items():
            b.scene_set_colliders(i, c)
    b.init_volumes()
    b.substeps(dt, steps)
    if sync:
        b.sync()
    return b, [b.scene_particles(i) for i in range(len(scs))]


def _run_single(s, dt, steps, colliders=None):
    g = M.Simulation(_params(s), s["x"].shape[0])
    g.set_particles(s["x"], s["v"], s["m"], s["E0"])
    if colliders:
        g.set_colliders(colliders)
    g.init_volumes()
    g.substeps(dt, steps)
    g.sync()
    out = g.particles()
    g.close()
    return out


@pytest.mark.parametrize("steps,tol,with_collider", [(1, 1e-6, True), (20, 1e-5, False)])
def test_batch_equals_scenes_run_one_by_one(steps, tol, with_collider):
    """A collider's response branches on the sign of v.n, so over many
    substeps rounding-level differences in contact grow past any fixed
    tolerance; the long comparison runs the scenes with domain walls only."""
    scs = _mixed_batch()
    for s in scs:
        s["params"]["bc_mode"] = M.BC_DOMAIN | M.BC_COLLIDERS
    cols = {}
    if with_collider:
        cols = {2: [M.collider(translation=(0.5, 0.2, 0.5), rotation=(0.9238795, 0, 0, 0.3826834),
                               scale=(0.3, 0.05, 0.3))]}
    dt = 1e-4
    b, got = _run_batch(scs, dt, steps, cols)
    assert b.scene_status() == [0] * len(scs)
    for i, s in enumerate(scs):
        ref = _run_single(s, dt, steps, cols.get(i))
        for k in ("x", "v", "FE", "FP", "V0"):
            assert rel_err(got[i][k], ref[k]) < tol, (i, k, rel_err(got[i][k], ref[k]))
    b.close()


def test_batch_scene_matches_oracle():
    """Two scenes of a batch (different grids and spacing) against the oracle
    run on each scene alone: one substep at 1e-5."""
    scs = _mixed_batch()[4:6]
    dt = 1e-4
    _, got = _run_batch(scs, dt, 1)
    for i, s in enumerate(scs):
        prm = s["params"]
        p = O.make_params(s["dims"], s["h"], bc_mode=prm["bc_mode"], E=prm["E"], nu=prm["nu"], xi=prm["xi"],
                          theta_c=prm["theta_c"], theta_s=prm["theta_s"], threads=8)
        o = O.Sim(p, s["x"].shape[0])
        o.set_particles(s["x"], s["v"], s["m"], s["E0"])
        o.stage(O.STAGE_P2G)
        o.stage(O.STAGE_VOLUMES)
        o.substeps(dt, 1)
        ref = o.particles()
        for k in ("x", "v", "FE", "FP", "V0"):
            assert rel_err(got[i][k], ref[k]) < 1e-5, (i, k, rel_err(got[i][k], ref[k]))


def test_nonfinite_scene_is_flagged_alone():
    """A scene whose particle carries a non-finite velocity raises its own
    status flag; the other scenes finish exactly as they do alone."""
    scs = _mixed_batch()[:3]
    bad = scs[1]
    bad["v"] = bad["v"].copy()
    bad["v"][0] = np.nan
    dt = 1e-4
    b, got = _run_batch(scs, dt, 3, sync=False)
    st = b.scene_status()
    assert st[1] == M.MPM_ERR_NONFINITE and st[0] == 0 and st[2] == 0
    for i in (0, 2):
        ref = _run_single(scs[i], dt, 3)
        assert rel_err(got[i]["x"], ref["x"]) < 1e-6
    b.close()


def test_cfl_warning_for_stiff_and_fast_scenes():
    """Per-substep stability flags on the device: a scene of E0 = 5e7 at 128^3
    and dt = 1e-4 (elastic wave dt c / h ~ 4.7, over the bound of 2) and a
    scene at 150 m/s (|v| dt / h ~ 1.9, over the bound of 1) raise the CFL
    warning for their own scene only; a scene of the stiffest configured
    modulus (E0 = 5e5 at 128^3, ~0.47) does not.  The warning does not fail synchronising calls; scene_status
    reports it once and clears it."""
    dt = 1e-4
    ok = scenes.snowball(seed=5, n=4000, grid=128, radius=0.04, v0=(0.0, -3.0, 0.0), E=5e5)
    stiff = scenes.snowball(seed=6, n=4000, grid=128, radius=0.04, v0=(0.0, -1.0, 0.0), E=5e7)
    fast = scenes.snowball(seed=7, n=4000, grid=128, radius=0.04, centre=(0.3, 0.5, 0.5), v0=(150.0, 0.0, 0.0))
    b, got = _run_batch([ok, stiff, fast], dt, 2)
    assert b.scene_status() == [0, M.MPM_ERR_CFL, M.MPM_ERR_CFL]
    assert b.scene_status() == [0, 0, 0]
    assert all(np.isfinite(g["x"]).all() for g in got)
    b.close()


def test_autotune_keeps_state_and_picks_a_candidate():
    scs = _mixed_batch()
    b = M.Simulation([_params(s) for s in scs], [s["x"].shape[0] for s in scs])
    for i, s in enumerate(scs):
        b.scene_set_particles(i, s["x"], s["v"], s["m"], s["E0"])
    b.init_volumes()
    b.substeps(1e-4, 5)
    b.sync()
    before = b.particles()
    ms = b.autotune(1e-4, steps=8)
    assert ms > 0.0
    r, g = b.tuning()
    assert r in (8, 16, 32, 64) and g >= 1
    after = b.particles()
    for k in before:
        np.testing.assert_array_equal(before[k], after[k])
    b.substeps(1e-4, 10)
    b.sync()
    b.close()


def test_batch_block_list_covers_each_scene():
    """After a batched binning every listed home block lies in its scene's
    grid blocks and its slot range in that scene's particle slots, and the
    listed ranges tile all slots exactly once."""
    scs = _mixed_batch()
    b, _ = _run_batch(scs, 1e-4, 3)
    b.debug_bin()
    blocks = b.debug_blocks()
    nb = [int(np.prod([(d + 3) // 4 for d in s["dims"]])) for s in scs]
    blk0 = np.concatenate([[0], np.cumsum(nb)])
    ns = [s["x"].shape[0] for s in scs]
    p0 = np.concatenate([[0], np.cumsum(ns)])
    sc = blocks[:, 3]
    assert sc.min() >= 0 and sc.max() < len(scs)
    assert np.all((blocks[:, 0] >= blk0[sc]) & (blocks[:, 0] < blk0[sc + 1]))
    assert np.all((blocks[:, 1] >= p0[sc]) & (blocks[:, 2] <= p0[sc + 1]) & (blocks[:, 1] < blocks[:, 2]))
    assert np.unique(blocks[:, 0]).size == blocks.shape[0]
    r = blocks[np.argsort(blocks[:, 1])]
    assert r[0, 1] == 0 and r[-1, 2] == p0[-1]
    np.testing.assert_array_equal(r[1:, 1], r[:-1, 2])
    b.close()
