"""Reverse mode: gradients of a loss on the final particles with respect to
the initial positions, velocities and Young's moduli, from the HIP adjoint
(mpm_adjoint_forward / mpm_adjoint_backward), against central finite
differences of the oracle's FP64 build run on the same inputs (relative
step 1e-4 for E0 and v0; V0 held fixed as the adjoint holds it).

The scene is a compressing snowball in flight: strains stay inside the
clamp band and nothing touches a wall, so the loss is smooth where it is
differentiated.  Tolerance: 1e-3 relative on directional derivatives (the
adjoint follows the FP32 trajectory, the differences the FP64 one; measured
9e-5 for E0 and 2e-6 for v0), and on each spot-checked dL/dx0, whose
dynamic part (through the transfers and the stress) is also held to 1%.
"""
import numpy as np
import pytest

import mpm_amd as M
import oracle as O
from mpm_amd import scenes

pytestmark = pytest.mark.gpu

DT, STEPS, EVERY = 1e-4, 12, 5


def _scene():
    sc = scenes.snowball(seed=21, n=2000, grid=32, radius=0.12, centre=(0.5, 0.5, 0.5), v0=(0.0, -1.0, 0.0), E=1.2e5)
    c = np.float32([0.5, 0.5, 0.5])
    sc["v"] = (sc["v"] - np.float32(5.0) * (sc["x"] - c)).astype(np.float32)  # compress at 5 / s
    return sc


def _gpu(sc):
    p = M.default_params(sc["dims"], sc["h"], compose=M.COMPOSE_TEXTBOOK, bc_mode=M.BC_DOMAIN, theta_s=7.5e-3)
    g = M.Simulation(p, sc["x"].shape[0])
    g.rollout_mass = sc["m"]
    return g


def _oracle_final(sc, x, v, E, V0):
    p = O.make_params(sc["dims"], sc["h"], compose=O.TEXTBOOK, bc_mode=O.BC_DOMAIN, theta_s=7.5e-3, threads=8)
    o = O.Sim(p, x.shape[0], double=True)
    o.set_particles(x, v, sc["m"], E, V0=V0)
    o.substeps(DT, STEPS)
    return o.particles()


def _loss(out, a, b):
    return float(np.sum(a * out["x"]) + np.sum(b * out["v"]))


@pytest.fixture(scope="module")
def case():
    sc = _scene()
    n = sc["x"].shape[0]
    rng = np.random.default_rng(5)
    a = rng.normal(size=(n, 3)).astype(np.float32)
    b = rng.normal(size=(n, 3)).astype(np.float32) * np.float32(1e-3)
    g = _gpu(sc)
    g.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"])
    g.init_volumes()
    V0 = g.particles()["V0"]
    g.adjoint_forward(DT, STEPS, EVERY)
    gx0, gv0, gE0 = g.adjoint_backward(a, b)
    # the context is back at its initial state
    np.testing.assert_array_equal(g.particles()["x"], sc["x"])
    g.close()
    return dict(sc=sc, a=a, b=b, V0=V0, gx0=gx0, gv0=gv0, gE0=gE0)


def _fd(case, xp, xm, vp, vm, Ep, Em):
    sc, a, b, V0 = case["sc"], case["a"], case["b"], case["V0"]
    return _loss(_oracle_final(sc, xp, vp, Ep, V0), a, b) - _loss(_oracle_final(sc, xm, vm, Em, V0), a, b)


def test_gradient_wrt_E0(case):
    sc = case["sc"]
    E = sc["E0"].astype(np.float64)
    d = 1e-4
    Ep, Em = (E * (1 + d)).astype(np.float32), (E * (1 - d)).astype(np.float32)
    fd = _fd(case, sc["x"], sc["x"], sc["v"], sc["v"], Ep, Em) / float(np.sum(Ep.astype(np.float64) - Em))
    ad = float(np.sum(case["gE0"].astype(np.float64) * (Ep.astype(np.float64) - Em))) / float(
        np.sum(Ep.astype(np.float64) - Em))
    print(f"E0 directional: adjoint {ad:.6e} fd {fd:.6e}")
    assert abs(ad - fd) <= 1e-3 * abs(fd), (ad, fd)


def test_gradient_wrt_v0(case):
    sc = case["sc"]
    u = np.random.default_rng(6).normal(size=sc["v"].shape)
    d = 1e-4
    vp, vm = (sc["v"] + d * u).astype(np.float32), (sc["v"] - d * u).astype(np.float32)
    step = (vp.astype(np.float64) - vm)
    fd = _fd(case, sc["x"], sc["x"], vp, vm, sc["E0"], sc["E0"])
    ad = float(np.sum(case["gv0"].astype(np.float64) * step))
    print(f"v0 directional: adjoint {ad:.6e} fd {fd:.6e}")
    assert abs(ad - fd) <= 1e-3 * abs(fd), (ad, fd)


def test_gradient_wrt_x0_spot_checks(case):
    sc = case["sc"]
    n = sc["x"].shape[0]
    h = sc["h"]
    picks = [(0, 0), (17, 1), (512, 2), (n - 1, 1), (1000, 0)]
    for p, k in picks:
        d = np.float32(1e-3 * h)
        xp, xm = sc["x"].copy(), sc["x"].copy()
        xp[p, k] += d
        xm[p, k] -= d
        step = float(xp[p, k]) - float(xm[p, k])
        fd = _fd(case, xp, xm, sc["v"], sc["v"], sc["E0"], sc["E0"]) / step
        ad = float(case["gx0"][p, k])
        # dL/dx0 = a (the particle carried along) + the dynamic part through
        # the grid transfers and the stress: the whole within 1e-3, the
        # dynamic part within 1%
        dyn_fd, dyn_ad = fd - float(case["a"][p, k]), ad - float(case["a"][p, k])
        print(f"x0[{p},{k}]: adjoint {ad:.6e} fd {fd:.6e} dynamic part adjoint {dyn_ad:.4e} fd {dyn_fd:.4e}")
        assert abs(ad - fd) <= 1e-3 * abs(fd), (p, k, ad, fd)
        assert abs(ad - fd) <= 1e-2 * abs(dyn_fd), (p, k, ad, fd, dyn_fd)


def test_keep_every_substep_matches_checkpoints(case):
    """checkpoint_every = 1 keeps every substep's record (state, P2G sums,
    slot order) and skips the segment re-runs: the same gradients as
    checkpoints every 5, and the context is back at its initial state.
    Different checkpoint intervals bin at different substeps, so the FP32
    grid sums add in another order; measured across intervals 2, 5, 6, 12:
    dL/dx0 and dL/dv0 agree to ~4e-7, the small per-particle dL/dE0 values to
    ~3e-4 and their sum (the E0 direction the finite differences check, to
    2e-3) to ~2e-4.  Tolerances: 1e-5, 1e-3 and 1e-3."""
    sc = case["sc"]
    g = _gpu(sc)
    g.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"])
    g.init_volumes()
    g.adjoint_forward(DT, STEPS, 1)
    gx0, gv0, gE0 = g.adjoint_backward(case["a"], case["b"])
    np.testing.assert_array_equal(g.particles()["x"], sc["x"])
    for got, ref, tol in ((gx0, case["gx0"], 1e-5), (gv0, case["gv0"], 1e-5), (gE0, case["gE0"], 1e-3)):
        err = np.linalg.norm(got.astype(np.float64) - ref) / np.linalg.norm(ref.astype(np.float64))
        assert err < tol, (err, tol)
    dE, dE_ref = float(np.sum(gE0.astype(np.float64))), float(np.sum(case["gE0"].astype(np.float64)))
    assert abs(dE - dE_ref) <= 1e-3 * abs(dE_ref), (dE, dE_ref)
    # and a second backward from the same records gives the same result
    g.adjoint_forward(DT, STEPS, 1)
    gx1, _, _ = g.adjoint_backward(case["a"], case["b"])
    assert np.linalg.norm(gx1.astype(np.float64) - gx0) <= 1e-5 * np.linalg.norm(gx0.astype(np.float64))
    g.close()


def test_automatic_interval_keeps_every_substep_when_it_fits(case):
    """checkpoint_every = 0 lets the library choose: at this size every
    substep's record fits in device memory, so it keeps them all
    (mpm_adjoint_interval reports 1) and the gradients are those of the
    checkpointed sweep, to the tolerances measured across intervals; an
    explicit interval is reported as given."""
    sc = case["sc"]
    g = _gpu(sc)
    assert g.adjoint_interval() == 0
    g.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"])
    g.init_volumes()
    g.adjoint_forward(DT, STEPS, 0)
    assert g.adjoint_interval() == 1
    gx0, gv0, gE0 = g.adjoint_backward(case["a"], case["b"])
    for got, ref, tol in ((gx0, case["gx0"], 1e-5), (gv0, case["gv0"], 1e-5), (gE0, case["gE0"], 1e-3)):
        err = np.linalg.norm(got.astype(np.float64) - ref) / np.linalg.norm(ref.astype(np.float64))
        assert err < tol, (err, tol)
    g.adjoint_forward(DT, STEPS, 5)
    assert g.adjoint_interval() == 5
    g.adjoint_backward(case["a"], case["b"])
    g.close()


def test_automatic_interval_falls_back_to_checkpoints_when_memory_is_short():
    """checkpoint_every = 0 when the records do not fit: a compressing ball of
    2,000 particles on a 128^3 grid for 40 substeps (34 MB of records per
    substep, 1.46 GB for the run) with all but 1.5 GB of device memory held
    by a torch allocation.  Keeping every substep would need more than half
    of what is free, so the library must take checkpoints every 20
    (mpm_adjoint_interval reports 20; their workspace, 20 substeps of
    records, still fits) and give the gradients of an explicit interval of
    20 run without the pressure (1e-5 for x0 and v0, 1e-3 for E0, as across
    intervals)."""
    torch = pytest.importorskip("torch")
    sc = scenes.snowball(seed=21, n=2000, grid=128, radius=0.03, centre=(0.5, 0.5, 0.5), v0=(0.0, -1.0, 0.0), E=1.2e5)
    sc["v"] = (sc["v"] - np.float32(5.0) * (sc["x"] - np.float32([0.5, 0.5, 0.5]))).astype(np.float32)
    rng = np.random.default_rng(6)
    a = rng.normal(size=sc["x"].shape).astype(np.float32)
    b = (rng.normal(size=sc["x"].shape) * 1e-3).astype(np.float32)

    def run(every):
        g = _gpu(sc)
        g.set_particles(sc["x"], sc["v"], sc["m"], sc["E0"])
        g.init_volumes()
        g.adjoint_forward(DT, 40, every)
        chosen = g.adjoint_interval()
        out = g.adjoint_backward(a, b)
        g.close()
        return chosen, out

    chosen_ref, ref = run(20)
    assert chosen_ref == 20
    free, _ = torch.cuda.mem_get_info()
    hold = free - (1536 << 20)
    assert hold > 0
    block = torch.empty(hold, dtype=torch.uint8, device="cuda")
    try:
        chosen, got = run(0)
    finally:
        del block
        torch.cuda.empty_cache()
    assert chosen == 20, chosen
    for x, r, tol in zip(got, ref, (1e-5, 1e-5, 1e-3)):
        err = np.linalg.norm(x.astype(np.float64) - r) / np.linalg.norm(r.astype(np.float64))
        assert err < tol, (err, tol)


def _scene_b():
    sc = scenes.snowball(seed=22, n=1500, grid=40, radius=0.1, centre=(0.5, 0.55, 0.5), v0=(0.5, -0.5, 0.0), E=1.5e5)
    c = np.float32([0.5, 0.55, 0.5])
    sc["v"] = (sc["v"] - np.float32(4.0) * (sc["x"] - c)).astype(np.float32)
    return sc


def test_batched_adjoint_equals_each_scene_alone(case):
    """Two scenes (other grid, spacing, modulus and motion) reversed in one
    batched context: each scene's gradients equal its own single-scene
    reverse sweep.  The batch bins and sums in another order, so the
    tolerances are those measured across checkpoint intervals: 1e-5 for
    dL/dx0 and dL/dv0, 1e-3 for dL/dE0."""
    sa, sb = case["sc"], _scene_b()
    na, nb = sa["x"].shape[0], sb["x"].shape[0]
    rng = np.random.default_rng(6)
    ab = rng.normal(size=(nb, 3)).astype(np.float32)
    bb = rng.normal(size=(nb, 3)).astype(np.float32) * np.float32(1e-3)
    gb = _gpu(sb)
    gb.set_particles(sb["x"], sb["v"], sb["m"], sb["E0"])
    gb.init_volumes()
    gb.adjoint_forward(DT, STEPS, EVERY)
    want_b = gb.adjoint_backward(ab, bb)
    gb.close()

    def params(sc):
        return M.default_params(sc["dims"], sc["h"], compose=M.COMPOSE_TEXTBOOK, bc_mode=M.BC_DOMAIN, theta_s=7.5e-3)

    batch = M.Simulation([params(sa), params(sb)], [na, nb])
    batch.scene_set_particles(0, sa["x"], sa["v"], sa["m"], sa["E0"])
    batch.scene_set_particles(1, sb["x"], sb["v"], sb["m"], sb["E0"])
    batch.init_volumes()
    batch.adjoint_forward(DT, STEPS, EVERY)
    got = batch.adjoint_backward(np.concatenate([case["a"], ab]), np.concatenate([case["b"], bb]))
    batch.close()
    want_a = (case["gx0"], case["gv0"], case["gE0"])
    for i, (ref_a, ref_b, tol) in enumerate(zip(want_a, want_b, (1e-5, 1e-5, 1e-3))):
        for part, ref in ((got[i][:na], ref_a), (got[i][na:], ref_b)):
            err = np.linalg.norm(part.astype(np.float64) - ref) / np.linalg.norm(ref.astype(np.float64))
            assert err < tol, (i, err, tol)


def test_torch_autograd_function_matches_adjoint(case):
    torch = pytest.importorskip("torch")
    from mpm_amd.autograd import Rollout

    sc = case["sc"]
    g = _gpu(sc)
    x0 = torch.tensor(sc["x"], requires_grad=True)
    v0 = torch.tensor(sc["v"], requires_grad=True)
    E0 = torch.tensor(sc["E0"], requires_grad=True)
    xT, vT = Rollout.apply(g, DT, STEPS, EVERY, x0, v0, E0)
    loss = (torch.tensor(case["a"]) * xT).sum() + (torch.tensor(case["b"]) * vT).sum()
    loss.backward()
    np.testing.assert_allclose(E0.grad.numpy(), case["gE0"], rtol=1e-3, atol=1e-3 * np.abs(case["gE0"]).max())
    np.testing.assert_allclose(v0.grad.numpy(), case["gv0"], rtol=1e-3, atol=1e-3 * np.abs(case["gv0"]).max())
    g.close()


def test_fit_E0_and_v0_to_a_target_trajectory():
    """Recover a shared Young's modulus and initial velocity from a target
    trajectory (positions after 30 substeps) with torch.optim.Adam through
    the autograd wrapper: the loss drops by 10x and v0 moves most of the way
    to its target."""
    torch = pytest.importorskip("torch")
    from mpm_amd.autograd import Rollout

    sc = _scene()
    g = _gpu(sc)
    c = np.float32([0.5, 0.5, 0.5])
    field = -np.float32(5.0) * (sc["x"] - c)
    steps = 30

    def rollout(logE, v0):
        n = sc["x"].shape[0]
        E0 = torch.exp(logE).expand(n).contiguous()
        v = v0.expand(n, 3) + torch.from_numpy(field)
        return Rollout.apply(g, DT, steps, 10, torch.from_numpy(sc["x"]), v, E0)

    with torch.no_grad():
        xT_target, _ = rollout(torch.tensor(math_log(1.0e5)), torch.tensor([0.3, -1.0, 0.2]))
    logE = torch.tensor(math_log(1.6e5), requires_grad=True)
    v0 = torch.tensor([0.0, -1.0, 0.0], requires_grad=True)
    opt = torch.optim.Adam([logE, v0], lr=0.05)
    losses = []
    for it in range(25):
        opt.zero_grad()
        xT, _ = rollout(logE, v0)
        loss = ((xT - xT_target) ** 2).sum(dim=1).mean() / float(sc["h"]) ** 2
        loss.backward()
        opt.step()
        losses.append(float(loss.detach()))
    dv0 = float(np.linalg.norm(v0.detach().numpy() - np.float32([0.3, -1.0, 0.2])))
    print(f"fit: loss {losses[0]:.4e} -> {losses[-1]:.4e}, |v0 - v0*| {dv0:.4f}, E {float(torch.exp(logE.detach())):.4e}")
    assert losses[-1] < 0.1 * losses[0]
    assert dv0 < 0.2 * float(np.linalg.norm([0.3, 0.0, 0.2]))
    g.close()


def math_log(v):
    return float(np.log(v))
