"""The headline workload at its full size: the 4,194,304-particle snow slab
on the 256^3 grid (128 x 32 x 128 cells of U{4..12} particles at dx = 1/256,
E0 in [1e5, 2e5], dt = 2.5e-5, LITERAL composition, domain walls).

- One fused substep against the oracle (run multithreaded on the host), every
  particle field within 1e-5 norm-relative, as at the small sizes.
- Over 400 substeps (25 re-binnings), a property that does not depend on
  size: the slab falls freely (0.1 m/s at the end, 0.13 cells of drift), far
  from every wall, so internal forces cancel and the APIC transfers conserve
  momentum; total momentum equals M g t.  (Positions near
  y = 0.5 have an FP32 spacing of 6e-8, above the first substeps'
  displacements, so the centre of mass is no check at this size.)
"""
import numpy as np
import pytest

import mpm_amd as M
import oracle as O
from helpers import rel_err
from mpm_amd import scenes

pytestmark = pytest.mark.gpu


def _gpu_sim(sc):
    prm = sc["params"]
    p = M.default_params(sc["dims"], sc["h"], compose=M.COMPOSE_LITERAL, bc_mode=prm["bc_mode"], E=prm["E"],
                         nu=prm["nu"], xi=prm["xi"], theta_c=prm["theta_c"], theta_s=prm["theta_s"])
    g = M.Simulation(p, sc["x"].shape[0])
    g.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"])
    g.init_volumes()
    return g


def _oracle_sim(sc, threads=16):
    prm = sc["params"]
    p = O.make_params(sc["dims"], sc["h"], compose=O.LITERAL, bc_mode=prm["bc_mode"], E=prm["E"], nu=prm["nu"],
                      xi=prm["xi"], theta_c=prm["theta_c"], theta_s=prm["theta_s"], threads=threads)
    o = O.Sim(p, sc["x"].shape[0])
    o.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"])
    o.stage(O.STAGE_P2G)
    o.stage(O.STAGE_VOLUMES)
    return o


def test_slab_4m_one_substep_matches_oracle():
    sc = scenes.slab()
    n = sc["x"].shape[0]
    assert n == 4194304
    print(f"full-size slab: {n} particles, one substep against the oracle", flush=True)
    g = _gpu_sim(sc)
    o = _oracle_sim(sc)
    assert rel_err(g.particles()["V0"], o.particles()["V0"]) < 1e-5
    g.substeps(sc["dt"], 1)
    o.substeps(sc["dt"], 1)
    gp, op = g.particles(), o.particles()
    errs = {k: rel_err(gp[k], op[k]) for k in ("x", "v", "FE", "FP")}
    # B is rounding noise in a free fall: measured against |v| h
    errs["B"] = float(np.linalg.norm(gp["B"].astype(np.float64) - op["B"]) /
                      max(np.linalg.norm(op["B"]), np.linalg.norm(op["v"]) * sc["h"]))
    assert all(e < 1e-5 for e in errs.values()), errs
    g.close()


def test_slab_4m_momentum_over_400_substeps():
    sc = scenes.slab()
    dt, k = sc["dt"], 400
    g = _gpu_sim(sc)
    g.substeps(dt, k)
    p = g.particles()
    m = sc["m"].astype(np.float64)
    mom = (m[:, None] * p["v"].astype(np.float64)).sum(0)
    want = m.sum() * np.array([0.0, -9.8, 0.0]) * (k * dt)
    assert np.linalg.norm(mom - want) < 1e-4 * np.linalg.norm(want), (mom, want)
    assert np.isfinite(p["x"]).all() and np.isfinite(p["FE"]).all()
    g.close()
