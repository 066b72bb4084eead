"""The 64-scene batch at full size (scenes.scene_batch(): 5,010,213
particles, N_s log-uniform in [4096, 262144], 42 scenes on 128^3 and 22 on
64^3, E0 log-uniform in [5e4, 5e5], v0 uniform in [-3, 3]^3, dt 1e-4, LITERAL
composition), as the bench runs it.

- One substep: two scenes (the smallest and the largest) against the oracle
  within 1e-5.
- 100 substeps in reproducible mode: every scene finite with no error, and
  six scenes of every size class bit-equal to their single-scene runs (every
  particle field), at another re-bin interval and block-pass width.
"""
import numpy as np
import pytest

import mpm_amd as M
import oracle as O
from helpers import rel_err
from mpm_amd import scenes

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "B", "FE", "FP", "V0")


def _params(sc):
    prm = sc["params"]
    return M.default_params(sc["dims"], sc["h"], bc_mode=prm["bc_mode"], E=prm["E"], nu=prm["nu"], xi=prm["xi"],
                            theta_c=prm["theta_c"], theta_s=prm["theta_s"])


@pytest.fixture(scope="module")
def batch():
    scs = scenes.scene_batch()
    assert len(scs) == 64 and sum(s["x"].shape[0] for s in scs) == 5010213
    return scs


def _batched(scs, exact, steps, tuning=None):
    b = M.Simulation([_params(s) for s in scs], [s["x"].shape[0] for s in scs])
    b.set_reproducible(exact)
    for i, s in enumerate(scs):
        b.scene_set_particles(i, s["x"], s["v"], s["m"], s["E0"])
    b.init_volumes()
    if tuning:
        b.set_tuning(*tuning)
    b.substeps(scs[0]["dt"], steps)
    b.sync()
    return b


def test_one_substep_against_the_oracle(batch):
    b = _batched(batch, False, 1)
    ns = [s["x"].shape[0] for s in batch]
    for i in (int(np.argmin(ns)), int(np.argmax(ns))):
        sc = batch[i]
        prm = sc["params"]
        o = O.Sim(O.make_params(sc["dims"], sc["h"], bc_mode=prm["bc_mode"], E=prm["E"], nu=prm["nu"], xi=prm["xi"],
                                theta_c=prm["theta_c"], theta_s=prm["theta_s"], threads=16), sc["x"].shape[0])
        o.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"])
        o.stage(O.STAGE_P2G)
        o.stage(O.STAGE_VOLUMES)
        o.substeps(sc["dt"], 1)
        gp, op = b.scene_particles(i), o.particles()
        errs = {f: rel_err(gp[f], op[f]) for f in ("x", "v", "FE", "FP", "V0")}
        errs["B"] = float(np.linalg.norm(gp["B"].astype(np.float64) - op["B"]) /
                          max(np.linalg.norm(op["B"]), np.linalg.norm(op["v"]) * sc["h"]))
        print(f"scene {i} ({sc['x'].shape[0]} particles, {sc['dims'][0]}^3): " +
              " ".join(f"{k} {v:.2e}" for k, v in errs.items()))
        assert max(errs.values()) < 1e-5, (i, errs)
    assert all(s in (0, M.MPM_ERR_CFL) for s in b.scene_status())
    b.close()


def test_reproducible_batch_equals_single_scenes(batch):
    steps = 100
    b = _batched(batch, True, steps, tuning=(16, 2048))
    status = b.scene_status()
    assert all(s in (0, M.MPM_ERR_CFL) for s in status), status
    got = {}
    ns = np.array([s["x"].shape[0] for s in batch])
    order = np.argsort(ns)
    picks = sorted({int(order[0]), int(order[len(order) // 3]), int(order[len(order) // 2]),
                    int(order[2 * len(order) // 3]), int(order[-1]),
                    int(np.flatnonzero(np.array([s["dims"][0] for s in batch]) == 64)[-1])})
    for i in range(len(batch)):
        P = b.scene_particles(i)
        assert all(np.isfinite(P[f]).all() for f in FIELDS), i
        if i in picks:
            got[i] = P
    b.close()
    for i in picks:
        sc = batch[i]
        g = M.Simulation(_params(sc), sc["x"].shape[0])
        g.set_reproducible(True)
        g.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"])
        g.init_volumes()
        g.set_tuning(8, 512)
        g.substeps(sc["dt"], steps)
        g.sync()
        P = g.particles()
        g.close()
        for f in FIELDS:
            assert np.array_equal(got[i][f], P[f]), (i, f, rel_err(got[i][f], P[f]))
    print(f"bit-equal after {steps} substeps: scenes {picks} ({[int(ns[i]) for i in picks]} particles)")
