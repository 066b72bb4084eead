"""Reproducible mode (mpm_set_reproducible): every P2G contribution is
rounded to a per-scene quantum (2^(e-40) of mass, 2^(e-38) of momentum, 2^e
the power of two just above the scene's largest particle mass) before the
FP64 atomics, so each node's sum is exact and does not depend on the order
in which the atomics land.

- A batch of scenes gives, scene by scene, the same bits as each scene run
  alone, for every particle field (x, v, B, F_E, F_P, V0) and the dense grid
  (mass, velocity, active set), although the single runs re-bin at another
  interval and run another block-pass width.  The scenes include free
  flight, a compressing (plastically deforming) snowball and a snowball
  landing on the sticky floor.
- Repeated runs give the same bits.
- The mode holds the parity bars against the oracle: 1e-5 after one substep
  and 1e-3 after 100 (norm-relative; grid momentum m v rather than grid
  velocity, since nodes reached only by stencil weights near 1e-12 carry a
  few quanta), and stays within 1e-5 of the default mode.
"""
import numpy as np
import pytest

import mpm_amd as M
import oracle as O
from helpers import rel_err
from mpm_amd import scenes

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "B", "FE", "FP", "V0")


def _params(sc, compose=M.COMPOSE_LITERAL):
    prm = sc["params"]
    return M.default_params(sc["dims"], sc["h"], compose=compose, bc_mode=prm["bc_mode"], E=prm["E"], nu=prm["nu"],
                            xi=prm["xi"], theta_c=prm["theta_c"], theta_s=prm["theta_s"])


def _mixed_scenes():
    out = scenes.scene_batch(seed=11, nscenes=3, n_range=(2000, 20000), split=12000, small_grid=32, large_grid=64)
    out.append(scenes.compressing(scenes.snowball(seed=4, n=6000, grid=48, radius=0.07, v0=(0.5, -1.0, 0.0)), 5.0))
    h = 1.0 / 64
    out.append(scenes.snowball(seed=5, n=5000, grid=64, radius=0.06, centre=(0.5, 4 * h + 0.065, 0.5),
                               v0=(0.0, -2.0, 0.0)))
    return out


def _run(scs, steps, exact=True, tuning=None, batch=True, dt=1e-4, compose=M.COMPOSE_LITERAL):
    """-> ([particles per scene], [grid per scene], [status per scene])."""
    if batch:
        runs = [(list(range(len(scs))),
                 M.Simulation([_params(s, compose) for s in scs], [s["x"].shape[0] for s in scs]))]
    else:
        runs = [([i], M.Simulation(_params(s, compose), s["x"].shape[0])) for i, s in enumerate(scs)]
    parts, grids, status = [None] * len(scs), [None] * len(scs), [None] * len(scs)
    for idx, sim in runs:
        sim.set_reproducible(exact)
        assert sim.reproducible == exact
        for k, i in enumerate(idx):
            sim.scene_set_particles(k, scs[i]["x"], scs[i]["v"], scs[i]["m"], scs[i]["E0"])
        sim.init_volumes()
        if tuning:
            sim.set_tuning(*tuning)
        sim.substeps(dt, steps)
        sim.sync()
        st = sim.scene_status()
        for k, i in enumerate(idx):
            parts[i] = sim.scene_particles(k)
            grids[i] = sim.scene_grid(k)
            status[i] = st[k]
        sim.close()
    return parts, grids, status


def _max_rel(a, b):
    """(largest norm-relative difference, its scene and field)."""
    return max((rel_err(a[i][f], b[i][f]), i, f) for i in range(len(a)) for f in FIELDS)


def test_batch_equals_single_scenes_bit_for_bit():
    scs = _mixed_scenes()
    steps = 80
    bp, bg, bs = _run(scs, steps, tuning=(16, 2048))
    sp, sg, ss = _run(scs, steps, tuning=(8, 512), batch=False)
    assert all(s in (0, M.MPM_ERR_CFL) for s in bs + ss), (bs, ss)
    # the runs deform: the compressing ball flows plastically, the landing
    # ball meets the floor
    FE3 = bp[3]["FE"].reshape(-1, 3, 3)
    assert (np.abs(FE3 - np.eye(3)).max(axis=(1, 2)) > 1e-3).mean() > 0.5
    JP3 = np.linalg.det(bp[3]["FP"].reshape(-1, 3, 3))
    assert (np.abs(JP3 - 1.0) > 1e-4).mean() > 0.05
    assert float(bp[4]["x"][:, 1].min()) < 3.5 / 64
    for i in range(len(scs)):
        for f in FIELDS:
            assert np.array_equal(bp[i][f], sp[i][f]), (i, f, rel_err(bp[i][f], sp[i][f]))
        for f in ("m", "vel", "active"):
            assert np.array_equal(bg[i][f], sg[i][f]), (i, f)
    # for the record: the default mode's sums depend on the atomics' order
    dp, _, _ = _run(scs, steps, exact=False, tuning=(16, 2048))
    dq, _, _ = _run(scs, steps, exact=False, tuning=(8, 512), batch=False)
    e, i, f = _max_rel(dp, dq)
    ex = max(rel_err(dp[k]["x"], dq[k]["x"]) for k in range(len(scs)))
    print(f"default mode, batch vs single after {steps} substeps: largest norm-relative difference {e:.2e} "
          f"(scene {i}, {f}); positions {ex:.2e}")


def test_repeated_runs_are_bit_for_bit():
    sc = scenes.compressing(scenes.snowball(seed=6, n=8192, grid=64, v0=(1.0, -1.0, 0.5)), 3.0)
    a, ga, _ = _run([sc], 60, tuning=(16, 2048), batch=False)
    b, gb, _ = _run([sc], 60, tuning=(32, 1024), batch=False)
    for f in FIELDS:
        assert np.array_equal(a[0][f], b[0][f]), f
    assert np.array_equal(ga[0]["m"], gb[0]["m"]) and np.array_equal(ga[0]["vel"], gb[0]["vel"])


def _grid_errs(gg, og):
    m_g, m_o = gg["m"].astype(np.float64), og["m"]
    return {"grid_m": rel_err(m_g, m_o),
            "grid_p": rel_err(m_g[:, None] * gg["vel"].astype(np.float64), m_o[:, None] * og["vel"]),
            "grid_vel": rel_err(gg["vel"], og["vel"])}


@pytest.mark.parametrize("compose", [M.COMPOSE_LITERAL, M.COMPOSE_TEXTBOOK], ids=["literal", "textbook"])
def test_reproducible_mode_matches_the_oracle(compose):
    """The snowball (8,192 particles, 64^3) in reproducible mode: 1 substep
    within 1e-5, 100 within 1e-3, against the oracle (FP32)."""
    sc = scenes.snowball()
    n = sc["x"].shape[0]
    g = M.Simulation(_params(sc, compose), n)
    g.set_reproducible(True)
    g.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"])
    prm = sc["params"]
    o = O.Sim(O.make_params(sc["dims"], sc["h"], compose=compose, bc_mode=prm["bc_mode"], E=prm["E"], nu=prm["nu"],
                            xi=prm["xi"], theta_c=prm["theta_c"], theta_s=prm["theta_s"], threads=8), n)
    o.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"])
    g.init_volumes()
    o.stage(O.STAGE_P2G)
    o.stage(O.STAGE_VOLUMES)
    for steps, tol in ((1, 1e-5), (99, 1e-3)):
        g.substeps(sc["dt"], steps)
        g.sync()
        o.substeps(sc["dt"], steps)
        gp, op = g.particles(), o.particles()
        errs = {f: rel_err(gp[f], op[f]) for f in ("x", "v", "FE", "FP", "V0")}
        errs["B"] = float(np.linalg.norm(gp["B"].astype(np.float64) - op["B"]) /
                          max(np.linalg.norm(op["B"]), np.linalg.norm(op["v"]) * sc["h"]))
        ge = _grid_errs(g.grid(), o.grid())
        print(f"reproducible, {steps} substep(s): " + " ".join(f"{k} {v:.2e}" for k, v in {**errs, **ge}.items()))
        errs.update({k: ge[k] for k in ("grid_m", "grid_p")})
        bad = {k: e for k, e in errs.items() if e > tol}
        assert not bad, (steps, bad, errs)
    g.close()


def test_reproducible_and_default_modes_agree():
    """The same deforming scene 40 substeps in both modes, TEXTBOOK (whose
    F update is continuous; LITERAL's basis choice turns last-bit
    differences into branch flips): within 1e-5."""
    sc = scenes.compressing(scenes.snowball(seed=7, n=8192, grid=64), 2.0)
    a, _, _ = _run([sc], 40, exact=True, batch=False, compose=M.COMPOSE_TEXTBOOK)
    b, _, _ = _run([sc], 40, exact=False, batch=False, compose=M.COMPOSE_TEXTBOOK)
    errs = {f: rel_err(a[0][f], b[0][f]) for f in FIELDS if f != "B"}
    print("reproducible vs default after 40 substeps: " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert max(errs.values()) < 1e-5, errs


def test_reverse_mode_in_reproducible_mode():
    """The reverse mode is not covered by the mode: its recording passes
    sum in the default format, and the grid pass reads that format while
    they run.  With the mode on, the gradients (kept records and checkpoints
    every 5) equal those of the default mode within the spread of two
    checkpoint intervals (1e-5 for x0 and v0, 1e-3 for E0)."""
    sc = scenes.compressing(scenes.snowball(seed=21, n=2000, grid=32, radius=0.12, centre=(0.5, 0.5, 0.5),
                                            v0=(0.0, -1.0, 0.0), E=1.2e5), 5.0)
    rng = np.random.default_rng(5)
    a = rng.normal(size=sc["x"].shape).astype(np.float32)
    b = (rng.normal(size=sc["x"].shape) * 1e-3).astype(np.float32)

    def grads(exact, every):
        p = M.default_params(sc["dims"], sc["h"], compose=M.COMPOSE_TEXTBOOK, bc_mode=M.BC_DOMAIN, theta_s=7.5e-3)
        g = M.Simulation(p, sc["x"].shape[0])
        g.set_reproducible(exact)
        g.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"])
        g.init_volumes()
        g.adjoint_forward(1e-4, 12, every)
        out = g.adjoint_backward(a, b)
        g.close()
        return out

    for every in (1, 5):
        ref, got = grads(False, every), grads(True, every)
        for x, r, tol in zip(got, ref, (1e-5, 1e-5, 1e-3)):
            assert np.isfinite(x).all()
            err = np.linalg.norm(x.astype(np.float64) - r) / np.linalg.norm(r.astype(np.float64))
            assert err < tol, (every, err, tol)


def test_profiled_launches_match_graph_replays():
    """The per-kernel profile (mpm_profile_begin with per_kernel = 1) launches
    every kernel directly between HIP events instead of replaying the
    substep graphs.  In reproducible mode it gives the same bits as the
    graph replays; it reports one first-substep P2G pass, a fused pass for
    each later substep, a grid pass per substep, the closing G2P pass and at
    least one re-bin, with positive times; and mpm_count_active counts the
    grid nodes that hold mass."""
    scs = _mixed_scenes()
    steps, dt = 40, 1e-4
    ref, ref_grids, _ = _run(scs, steps, tuning=(16, 1024))
    sim = M.Simulation([_params(s) for s in scs], [s["x"].shape[0] for s in scs])
    sim.set_reproducible(True)
    for k, s in enumerate(scs):
        sim.scene_set_particles(k, s["x"], s["v"], s["m"], s["E0"])
    sim.init_volumes()
    sim.set_tuning(16, 1024)
    sim.profile_begin(per_kernel=True)
    sim.substeps(dt, steps)
    total_ms, kern = sim.profile_end()
    sim.sync()
    for k in range(len(scs)):
        P = sim.scene_particles(k)
        for f in FIELDS:
            assert np.array_equal(P[f], ref[k][f]), (k, f)
        G = sim.scene_grid(k)
        assert np.array_equal(G["m"], ref_grids[k]["m"]), k
    assert kern["p2g"][1] == 1 and kern["fused"][1] == steps - 1
    assert kern["grid"][1] == steps and kern["g2p"][1] == 1
    assert kern["bin"][1] >= 1
    for name in ("p2g", "fused", "grid", "g2p", "bin"):
        assert kern[name][0] > 0.0, name
    assert total_ms >= kern["fused"][0]
    # the dense grids of all scenes, counted on the device and on the host
    with_mass = sum(int(np.count_nonzero(sim.scene_grid(k)["m"])) for k in range(len(scs)))
    assert sim.count_active() == with_mass > 0
    sim.close()
