"""GPU parity: the HIP library (through its C ABI) against the oracle.

Tolerances: grid mass and momentum and particle x, v, C, F_E,
F_P and J_P within 1e-5 norm-relative after one substep, within 1e-3 after
100 substeps.  The deformation-gradient update on identical inputs is checked
bit-exactly (same FP32 operation order, correctly rounded division and sqrt).
"""
import numpy as np
import pytest

import mpm_amd as M
import oracle as O
from helpers import eye9, fp64_textbook_update, max_abs, near_tied_spectra, rel_err, to_math, warm_state
from mpm_amd import scenes

pytestmark = pytest.mark.gpu


def _pair(sc, compose=0, colliders=None, gap64=0.0, threads=8):
    prm = dict(sc["params"])
    if colliders:
        prm["bc_mode"] |= O.BC_COLLIDERS
    p_gpu = M.default_params(sc["dims"], sc["h"], compose=compose, bc_mode=prm["bc_mode"], E=prm["E"], nu=prm["nu"],
                             xi=prm["xi"], theta_c=prm["theta_c"], theta_s=prm["theta_s"], svd_gap_fp64=gap64)
    n = sc["x"].shape[0]
    g = M.Simulation(p_gpu, n)
    g.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"])
    p_or = O.make_params(sc["dims"], sc["h"], compose=compose, bc_mode=prm["bc_mode"], E=prm["E"], nu=prm["nu"],
                         xi=prm["xi"], theta_c=prm["theta_c"], theta_s=prm["theta_s"], threads=1)
    o = O.Sim(p_or, n)
    o.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"])
    if colliders:
        g.set_colliders([M.collider(**c) for c in colliders])
        o.set_colliders(np.array([list(c["translation"]) + list(c["rotation"]) + list(c["scale"]) +
                                  list(c.get("velocity", (0, 0, 0))) for c in colliders], np.float32))
    return g, o


def _init(g, o):
    g.init_volumes()
    o.stage(O.STAGE_P2G)
    o.stage(O.STAGE_VOLUMES)


def _jp(FP):
    return np.linalg.det(to_math(FP))


def _compare_particles(gp, op, tol, h=None, fields=("x", "v", "B", "FE", "FP", "V0")):
    """Norm-relative error per field.  B (and C = B D^-1, same relative
    error) is measured against max(||B_ref||, ||v_ref|| h): in a rigid or
    free-fall motion B is rounding noise and its own norm is no scale."""
    errs = {k: rel_err(gp[k], op[k]) for k in fields}
    if h is not None and "B" in fields:
        scale = max(np.linalg.norm(op["B"]), np.linalg.norm(op["v"]) * h)
        errs["B"] = float(np.linalg.norm(gp["B"].astype(np.float64) - op["B"]) / scale)
    errs["JP"] = rel_err(_jp(gp["FP"]), _jp(op["FP"]))
    bad = {k: e for k, e in errs.items() if e > tol}
    assert not bad, f"fields over {tol}: {bad} (all: {errs})"
    return errs


def test_textbook_fp64_resolve_on_the_device():
    """svd_gap_fp64 > 0 (TEXTBOOK): particles whose singular values nearly tie
    re-solve the SVD and clamp in FP64 on the device.  On near-tied pairs
    that straddle a clamp bound the device F_E is within one FP32 ulp of the
    oracle's FP64 update; with the re-solve off it is not."""
    rng = np.random.default_rng(21)
    n = 4096
    B, FE, FP = near_tied_spectra(rng, n)
    ref = fp64_textbook_update(B, FE, FP)
    x = rng.uniform(0.3, 0.7, (n, 3)).astype(np.float32)
    errs = {}
    for gap in (0.0, 1e-3):
        pg = M.default_params((64, 64, 64), 1 / 64, compose=M.COMPOSE_TEXTBOOK, theta_s=7.5e-3, svd_gap_fp64=gap)
        g = M.Simulation(pg, n)
        g.set_particles(x, np.zeros_like(x), np.ones(n, np.float32))
        g.set_state(B=B, FE=FE, FP=FP)
        g.stage(M.STAGE_FUPDATE, 1e-4)
        out = g.particles()
        g.close()
        errs[gap] = float(np.abs(out["FE"] - ref["FE"]).max())
        if gap:
            np.testing.assert_allclose(out["FP"], ref["FP"], rtol=0, atol=4e-7)
    assert errs[1e-3] < 1.2e-7, errs
    assert errs[0.0] > 4 * errs[1e-3], errs


def test_fused_textbook_with_fp64_resolve_matches_fp64_oracle():
    """The fused path with the FP64 re-solve on (TEXTBOOK, svd_gap_fp64 =
    1e-3) on a warm two-snowball state: 1 and 10 substeps against the
    oracle's FP64 build within the FP32 tolerances."""
    rng = np.random.default_rng(22)
    sc = scenes.two_snowballs(n_total=16384, grid=64, dt=1e-4)
    n = sc["x"].shape[0]
    Bw, FEw, FPw = warm_state(rng, n, bscale=0.01, scale=5e-3)
    g, _ = _pair(sc, compose=1, gap64=1e-3)
    g.set_state(B=Bw, FE=FEw, FP=FPw)
    prm = sc["params"]
    p64 = O.make_params(sc["dims"], sc["h"], compose=1, bc_mode=prm["bc_mode"], theta_s=prm["theta_s"], threads=8)
    o = O.Sim(p64, n, double=True)
    o.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"], B=Bw, FE=FEw, FP=FPw)
    _init(g, o)
    for steps, tol in ((1, 1e-5), (9, 1e-4)):
        g.substeps(1e-4, steps)
        g.sync()
        o.substeps(1e-4, steps)
        _compare_particles(g.particles(), o.particles(), tol, h=sc["h"])
    g.close()


def test_fupdate_bit_exact_literal_and_textbook():
    rng = np.random.default_rng(7)
    n = 4096
    B, FE, FP = warm_state(rng, n, bscale=3.0)
    x = rng.uniform(0.3, 0.7, (n, 3)).astype(np.float32)
    for compose in (0, 1):
        pg = M.default_params((64, 64, 64), 1 / 64, compose=compose, theta_s=7.5e-3)
        g = M.Simulation(pg, n)
        g.set_particles(x, np.zeros_like(x), np.ones(n, np.float32))
        g.set_state(B=B, FE=FE, FP=FP)
        g.stage(M.STAGE_FUPDATE, 1e-4)
        out = g.particles()
        for p in range(0, n, 97):
            rc, FEo, FPo, _ = O.fupdate(compose, 1e-4, 1 / 64, 2.5e-2, 7.5e-3, B[p], FE[p], FP[p])
            assert rc == 0
            np.testing.assert_array_equal(out["FE"][p], FEo)
            np.testing.assert_array_equal(out["FP"][p], FPo)
        g.close()


def test_stages_reference_scene_one_substep():
    """Each legacy stage on the reference demo scene (20^3, h=0.05, 2,147
    particles, main.cpp's colliders) from identical inputs."""
    sc = scenes.reference_scene()
    g, o = _pair(sc, colliders=scenes.reference_colliders())
    _init(g, o)
    np.testing.assert_allclose(g.particles()["V0"], o.particles()["V0"], rtol=2e-6)
    dt = sc["dt"]
    for st in (O.STAGE_P2G, O.STAGE_FORCES, O.STAGE_GRID, O.STAGE_COLLIDE):
        g.stage(st, dt)
        o.stage(st, dt)
    gg, og = g.grid(force=True), o.grid()
    assert np.array_equal(gg["active"], og["active"])
    assert rel_err(gg["m"], og["m"]) < 1e-6
    assert rel_err(gg["vel"], og["vel"]) < 1e-5
    assert rel_err(gg["force"], og["force"]) < 1e-5
    for st in (O.STAGE_FUPDATE, O.STAGE_G2P, O.STAGE_ADVECT):
        g.stage(st, dt)
        o.stage(st, dt)
    _compare_particles(g.particles(), o.particles(), 1e-5, h=sc["h"])


def test_snowball_fused_1_and_100_substeps():
    """8,192-particle snowball on 64^3 through the fused substep."""
    sc = scenes.snowball()
    g, o = _pair(sc)
    _init(g, o)
    g.substeps(sc["dt"], 1)
    o.substeps(sc["dt"], 1)
    g.sync()
    _compare_particles(g.particles(), o.particles(), 1e-5, h=sc["h"])
    gg, og = g.grid(), o.grid()
    assert rel_err(gg["m"], og["m"]) < 1e-5
    assert rel_err(gg["vel"], og["vel"]) < 1e-5
    g.substeps(sc["dt"], 99)
    o.substeps(sc["dt"], 99)
    g.sync()
    _compare_particles(g.particles(), o.particles(), 1e-3, h=sc["h"])


def test_snowball_free_fall_closed_form_on_the_device():
    """The 8,192-particle snowball never touches a boundary in 100 substeps:
    uniform v, B = 0, F = I, so the fused path itself must give
    x_n = x_0 + n dt v0 + dt^2 g n(n+1)/2 and v_n = v0 + n dt g (the
    mathematics, independent of the oracle; FP32 tolerances as the oracle's
    own check)."""
    sc = scenes.snowball()
    g, _ = _pair(sc)
    g.init_volumes()
    n, dt = 100, sc["dt"]
    g.substeps(dt, n)
    P = g.particles()
    grav = np.array([0.0, -9.8, 0.0])
    expect = sc["x"].astype(np.float64) + n * dt * sc["v"] + dt * dt * grav * n * (n + 1) / 2
    assert np.abs(P["x"] - expect).max() < 2e-6
    np.testing.assert_allclose(P["v"], sc["v"] + n * dt * grav, atol=2e-5)
    np.testing.assert_allclose(P["FE"], eye9(P["FE"].shape[0]), atol=1e-5)
    g.close()


def test_nodes_that_lost_their_mass_read_zero():
    """The falling snowball leaves grid blocks behind every few substeps.
    After 150 substeps the dense grid holds mass and velocity only where
    the particles of substep 149 deposit: nodes that had mass at substep 50
    and have none now read zero mass and zero velocity (the grid pass clears
    blocks that dropped out of the active list), and the mass equals the
    oracle's P2G of the positions the last substep started from."""
    sc = scenes.snowball()
    g, _ = _pair(sc)
    g.init_volumes()
    dt = sc["dt"]
    g.substeps(dt, 50)
    early = g.grid()
    g.substeps(dt, 99)
    P = g.particles()
    g.substeps(dt, 1)
    late = g.grid()
    o = O.Sim(O.make_params(sc["dims"], sc["h"], bc_mode=sc["params"]["bc_mode"], theta_s=sc["params"]["theta_s"],
                            threads=1), P["x"].shape[0])
    o.set_particles(P["x"], P["v"], sc["m"], sc["E0"])
    o.stage(O.STAGE_P2G)
    want = o.grid()["m"]
    empty = want == 0
    left = (early["m"] > 0) & empty
    assert left.sum() > 100  # blocks did drop out between substeps 50 and 150
    assert np.all(late["m"][empty] == 0.0)
    assert np.all(late["vel"][empty] == 0.0)
    assert rel_err(late["m"], want) < 1e-5
    g.close()


def test_warm_state_one_substep():
    """A warm state (non-zero B, F_E off identity, F_P != I) and one fused
    substep: exercises stress, hardening and the clamp."""
    sc = scenes.snowball(n=8192)
    rng = np.random.default_rng(3)
    B, FE, FP = warm_state(rng, 8192, bscale=0.5)
    g, o = _pair(sc)
    _init(g, o)
    V0 = o.particles()["V0"].astype(np.float32)
    g.set_state(B=B, FE=FE, FP=FP, V0=V0)
    o.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"], B=B, FE=FE, FP=FP, V0=V0)
    g.substeps(sc["dt"], 1)
    o.substeps(sc["dt"], 1)
    g.sync()
    _compare_particles(g.particles(), o.particles(), 1e-5, h=sc["h"])


@pytest.mark.parametrize("which", ["reference", "snowball", "two_snowballs", "random", "ragged_slab", "clumped"])
def test_binning_matches_cpu_stable_sort(which):
    """Keys and permutation of the radix sort equal numpy's stable argsort,
    and the particles take their slots in that order (a clumped cell of 300
    particles included)."""
    if which == "reference":
        sc = scenes.reference_scene()
    elif which == "snowball":
        sc = scenes.snowball()
    elif which == "two_snowballs":
        sc = scenes.two_snowballs()
    elif which == "ragged_slab":
        sc = scenes.slab(n_total=2 ** 16, cells=(32, 8, 32), grid=128)
    elif which == "clumped":
        rng = np.random.default_rng(12)
        n = 40000
        x = rng.uniform(0.2, 0.8, (n, 3)).astype(np.float32)
        x[:300] = (np.float32([20.5, 21.5, 22.5]) + rng.uniform(-0.45, 0.45, (300, 3)).astype(np.float32)) / 64
        sc = dict(x=x, v=np.zeros_like(x), m=np.ones(n, np.float32), E0=np.full(n, 1.4e5, np.float32),
                  dims=(64, 64, 64), h=1 / 64, params=scenes.material())
    else:
        rng = np.random.default_rng(11)
        n = 100003
        x = rng.uniform(0.05, 0.7, (n, 3)).astype(np.float32)
        sc = dict(x=x, v=np.zeros_like(x), m=np.ones(n, np.float32), E0=np.full(n, 1.4e5, np.float32),
                  dims=(96, 80, 72), h=1 / 96, params=scenes.material())
    n = sc["x"].shape[0]
    g = M.Simulation(M.default_params(sc["dims"], sc["h"]), n)
    g.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"])
    keys, perm, slot = g.debug_bin()
    rk, rp, rs = O.bin_reference(sc["x"], sc["h"], sc["dims"])
    np.testing.assert_array_equal(keys, rk)
    np.testing.assert_array_equal(perm, rp)
    np.testing.assert_array_equal(slot, rs)
    # the home-block list: every non-empty block once, with its slot range in
    # the sorted keys; claimed in a spread order, not key order
    blocks = g.debug_blocks()
    ublk, first, cnt = np.unique(rk >> 6, return_index=True, return_counts=True)
    assert blocks.shape == (ublk.size, 4)
    b = blocks[np.argsort(blocks[:, 0])]
    np.testing.assert_array_equal(b[:, 0], ublk)
    np.testing.assert_array_equal(b[:, 1], first)
    np.testing.assert_array_equal(b[:, 2], first + cnt)
    np.testing.assert_array_equal(b[:, 3], 0)
    if ublk.size > 2:
        assert not np.all(np.diff(blocks[:, 0]) > 0)
    # particles come back in the caller's order after binning
    np.testing.assert_array_equal(g.particles()["x"], sc["x"])


def test_two_snowballs_one_and_ten_substeps():
    """262,144 particles on 128^3 (two colliding snowballs), fused path."""
    sc = scenes.two_snowballs()
    g, o = _pair(sc)
    o.set_threads(8)
    _init(g, o)
    g.substeps(sc["dt"], 1)
    o.substeps(sc["dt"], 1)
    g.sync()
    _compare_particles(g.particles(), o.particles(), 1e-5, h=sc["h"])
    g.substeps(sc["dt"], 9)
    o.substeps(sc["dt"], 9)
    g.sync()
    _compare_particles(g.particles(), o.particles(), 1e-4, h=sc["h"])


def test_resort_interval_does_not_change_results():
    """Re-binning every substep vs every 16, 32 or 64, with the block pass on its
    default width, on one workgroup that claims every home block in turn, or
    on more workgroups than there are home blocks: same physics (float
    atomics aside), including particles that drift across block boundaries."""
    sc = scenes.two_snowballs(n_total=32768, grid=64, dt=2e-4)
    out = []
    for every, width in ((1, None), (16, None), (32, 1), (32, 4096), (64, 1024)):
        prm = sc["params"]
        p = M.default_params(sc["dims"], sc["h"], bc_mode=prm["bc_mode"], theta_s=prm["theta_s"], resort_every=every)
        g = M.Simulation(p, sc["x"].shape[0])
        if width is not None:
            g.set_tuning(every, width)
        g.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"])
        g.init_volumes()
        g.substeps(sc["dt"], 40)
        g.sync()
        out.append(g.particles())
        g.close()
    for o in out[1:]:
        for k in ("x", "v", "FE", "FP"):
            assert rel_err(out[0][k], o[k]) < 1e-5, k


def test_graph_replay_matches_direct_launches():
    """Substeps replayed from HIP graphs (re-bin + resort_every substeps per
    segment, one graph per particle-buffer parity) against the same calls
    launched directly: runs of odd lengths so segments start and end
    mid-call, a dt change and a collider change that force re-capture."""
    sc = scenes.two_snowballs(n_total=32768, grid=64, dt=2e-4)
    cols = [M.collider(translation=(0.5, 0.1, 0.5), rotation=(1, 0, 0, 0), scale=(0.3, 0.05, 0.3)),
            M.collider(translation=(0.2, 0.3, 0.5), rotation=(0.9238795, 0, 0, 0.3826834), scale=(0.05, 0.05, 0.2))]
    out = []
    for graphs in (True, False):
        prm = sc["params"]
        p = M.default_params(sc["dims"], sc["h"], bc_mode=prm["bc_mode"] | M.BC_COLLIDERS, theta_s=prm["theta_s"],
                             resort_every=4)
        g = M.Simulation(p, sc["x"].shape[0])
        g.use_graphs(graphs)
        g.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"])
        g.set_colliders(cols[:1])
        g.init_volumes()
        g.substeps(sc["dt"], 1)
        g.substeps(sc["dt"], 9)
        g.substeps(sc["dt"] * 0.5, 13)
        g.set_colliders(cols)
        g.substeps(sc["dt"] * 0.5, 17)
        g.sync()
        out.append(g.particles())
    for k in ("x", "v", "FE", "FP"):
        assert rel_err(out[0][k], out[1][k]) < 1e-5, k
