"""Deforming-regime parity: the HIP path against the oracle on states where
the elastic response, the SVD clamp, hardening and the boundary act, which
free fall (the other long comparisons) never exercises.

Each case runs the GPU from a scene's start to just before contact, reads
that state back (x, v, B, F_E, F_P, V0), loads it into a fresh GPU context
and into the oracle (FP32, same operation order), and compares both over the
next 100 substeps of contact.  Fields: grid mass and momentum (dense grid of
the last substep), particle x, v, B (against max(||B||, ||v|| h)), F_E, F_P
and J_P = det F_P; norm-relative errors, tolerance 1e-3 after 100 substeps.

LITERAL's F_E = U^T S V (material_point_method.cpp:323-326) depends on which
basis the Jacobi SVD returns, a discontinuous function of F_hat, and in
plastic flow two singular values sit on the same clamp bound, where that
basis is arbitrary: where the FP32 grid sums of the two paths differ in
their last bits (different summation order), such a particle takes the
other branch and its F_E jumps.  The count of such particles ("branch
flips": max |dF_E| > 1e-3) is reported.  TEXTBOOK's V S U^T is continuous
and holds every field to 1e-3.  LITERAL holds positions to 1e-4; its other
fields carry the flips (measured after 100 substeps of the two-snowball
collision: 17% of particles flipped, v 3.4e-3, F_E 6.1e-3, grid momentum
4.6e-3; of the landing: 14%, v 1.1e-3, F_E 1.8e-3), held to 2e-2 with
fewer than 25% of particles flipped.
"""
import numpy as np
import pytest

import mpm_amd as M
import oracle as O
from helpers import rel_err, to_math
from mpm_amd import scenes

pytestmark = pytest.mark.gpu


def _params(sc, compose):
    prm = sc["params"]
    kw = dict(compose=compose, bc_mode=prm["bc_mode"], E=prm["E"], nu=prm["nu"], xi=prm["xi"],
              theta_c=prm["theta_c"], theta_s=prm["theta_s"])
    return M.default_params(sc["dims"], sc["h"], **kw), O.make_params(sc["dims"], sc["h"], threads=16, **kw)


def _gpu_state(sc, compose, steps):
    """The scene on the GPU from its start for `steps` substeps -> state."""
    pg, _ = _params(sc, compose)
    g = M.Simulation(pg, sc["x"].shape[0])
    g.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"])
    g.init_volumes()
    g.substeps(sc["dt"], steps)
    g.sync()
    st = g.particles()
    assert g.scene_status()[0] in (0, M.MPM_ERR_CFL)
    g.close()
    return st


def _grid_errs(gg, og):
    m_g, m_o = gg["m"].astype(np.float64), og["m"]
    p_g = m_g[:, None] * gg["vel"].astype(np.float64)
    p_o = m_o[:, None] * og["vel"]
    return {"grid_m": rel_err(m_g, m_o), "grid_p": rel_err(p_g, p_o)}


def _from_state(sc, st, compose, steps):
    """GPU and oracle from the same state for `steps` substeps; -> errors,
    branch flips, and a measure of how much the state deformed."""
    n = sc["x"].shape[0]
    pg, po = _params(sc, compose)
    f32 = {k: np.ascontiguousarray(st[k], np.float32) for k in ("x", "v", "B", "FE", "FP", "V0")}
    g = M.Simulation(pg, n)
    g.set_particles(f32["x"], f32["v"], sc["m"], sc["E0"])
    g.set_state(B=f32["B"], FE=f32["FE"], FP=f32["FP"], V0=f32["V0"])
    o = O.Sim(po, n)
    o.set_particles(f32["x"], f32["v"], sc["m"], sc["E0"], B=f32["B"], FE=f32["FE"], FP=f32["FP"], V0=f32["V0"])
    g.substeps(sc["dt"], steps)
    g.sync()
    o.substeps(sc["dt"], steps)
    gp, op = g.particles(), o.particles()
    errs = {k: rel_err(gp[k], op[k]) for k in ("x", "v", "FE", "FP")}
    scale = max(np.linalg.norm(op["B"]), np.linalg.norm(op["v"]) * sc["h"])
    errs["B"] = float(np.linalg.norm(gp["B"].astype(np.float64) - op["B"]) / scale)
    errs["JP"] = rel_err(np.linalg.det(to_math(gp["FP"])), np.linalg.det(to_math(op["FP"])))
    errs.update(_grid_errs(g.grid(), o.grid()))
    status = g.scene_status()
    g.close()
    flips = int((np.abs(gp["FE"].astype(np.float64) - op["FE"]).max(1) > 1e-3).sum())
    FEo = to_math(op["FE"])
    deformed = float((np.abs(FEo - np.eye(3)).max(axis=(1, 2)) > 1e-3).mean())
    plastic = float((np.abs(np.linalg.det(to_math(op["FP"])) - 1.0) > 1e-4).mean())
    return errs, flips, deformed, plastic, status


def _check(errs, flips, n, compose, tol=1e-3):
    if compose == M.COMPOSE_TEXTBOOK:
        bad = {k: e for k, e in errs.items() if e > tol}
        assert not bad, (bad, errs)
    else:
        # the measured limit of the reference's composition (module docstring)
        assert errs["x"] < 1e-4, errs
        bad = {k: e for k, e in errs.items() if e > 2e-2}
        assert not bad, (bad, errs)
        assert flips < 0.25 * n, (flips, errs)


@pytest.mark.parametrize("compose", [M.COMPOSE_TEXTBOOK, M.COMPOSE_LITERAL], ids=["textbook", "literal"])
def test_two_snowballs_through_contact(compose):
    """The two-snowball collision: 262,144 particles in two snowballs at +-3 m/s on 128^3,
    dt 5e-5.  The GPU runs to substep 620 (the spheres touch near substep
    667), then GPU and oracle run 100 substeps of the collision from that
    state."""
    sc = scenes.two_snowballs()
    st = _gpu_state(sc, compose, 620)
    errs, flips, deformed, plastic, status = _from_state(sc, st, compose, 100)
    print(f"two snowballs 620->720 {'textbook' if compose else 'literal'}: deformed {deformed:.3f} plastic "
          f"{plastic:.3f} flips {flips} status {status} errs " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert deformed > 0.05, deformed  # the comparison window is in contact
    assert status[0] in (0, M.MPM_ERR_CFL)  # no error; a CFL warning is reported above
    _check(errs, flips, sc["x"].shape[0], compose)


@pytest.mark.parametrize("compose", [M.COMPOSE_TEXTBOOK, M.COMPOSE_LITERAL], ids=["textbook", "literal"])
def test_snowball_landing_on_the_sticky_floor(compose):
    """The 8,192-particle snowball (64^3, dt 1e-4) started three
    cells above the sticky floor band at 2 m/s: its bottom reaches the band's
    stencils (y < 4h) near substep 151, so the GPU runs 140 substeps, then GPU
    and oracle run the 100 substeps of the landing from that state (the floor
    stops the bottom layers, the ball compresses and spreads)."""
    h = 1.0 / 64
    sc = scenes.snowball(centre=(0.5, 6 * h + 0.08, 0.5), v0=(0.0, -2.0, 0.0))
    st = _gpu_state(sc, compose, 140)
    errs, flips, deformed, plastic, status = _from_state(sc, st, compose, 100)
    print(f"landing 140->240 {'textbook' if compose else 'literal'}: deformed {deformed:.3f} plastic "
          f"{plastic:.3f} flips {flips} status {status} errs " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert deformed > 0.05, deformed
    assert status[0] in (0, M.MPM_ERR_CFL)
    _check(errs, flips, sc["x"].shape[0], compose)
