"""Device scene generators, positions readback and snapshots, through the C
ABI (mpm_scene_generate, mpm_get_positions / mpm_positions_begin / _wait,
mpm_snapshot_save / _load, mpm_create_from_snapshot).

- Generators: every field bit-equal to the oracle's sequential restatement
  (integer Philox draws and single rounded float ops), at small sizes and at
  the full 4,194,304-particle slab; V0 from mpm_init_volumes after a device
  generation against the oracle's V0 (1e-5, summation order differs).
- Readback: positions equal the full readback bit for bit; an asynchronous
  read returns the positions as of its begin while later substeps run.
- Snapshots: the library's file reads back through mpm_amd.snapshot bit for
  bit; a context created from it holds the same state bit for bit and, run
  on, agrees with the original to rounding (grid atomics accumulate in a
  different order: 1e-6); a numpy-written fixture loads and steps as the
  oracle does from the same state (1e-5 after one substep).
"""
import numpy as np
import pytest

import mpm_amd as M
import oracle as O
from helpers import eye9, rel_err
from mpm_amd import snapshot as S

pytestmark = pytest.mark.gpu

DEMO = dict(h=0.05, dims=(20, 20, 20), centre=(0.5, 0.6, 0.5), radius=0.2)


def _cases():
    lat_kept = O.lattice_candidates(DEMO["h"], 3, DEMO["centre"], DEMO["radius"])
    return [
        dict(kind=M.GEN_BALL, n=20000, dims=(64, 64, 64), h=1 / 64, seed=5, centre=(0.5, 0.6, 0.5), radius=0.08,
             E=(1e5, 2e5), velocity=(0.0, -1.0, 0.0), mass=1.0e-4),
        dict(kind=M.GEN_BALL, n=1, dims=(16, 16, 16), h=1 / 16, seed=0, centre=(0.5, 0.5, 0.5), radius=0.1,
             E=(1.4e5, 1.4e5), velocity=(1.0, 2.0, 3.0), mass=2.0),
        dict(kind=M.GEN_CELLS, n=8 * 495 + 37, dims=(64, 64, 64), h=1 / 64, seed=9, cell_lo=(10, 20, 12),
             cells=(9, 5, 11), ppc=(4, 12), E=(1e5, 2e5), velocity=(0.0, 0.0, 0.0), mass=3e-6),
        dict(kind=M.GEN_LATTICE, n=lat_kept - 3, dims=DEMO["dims"], h=DEMO["h"], seed=3, centre=DEMO["centre"],
             radius=DEMO["radius"], E=(1.4e5, 1.4e5), velocity=(0.0, -200.0, 0.0), mass=6e-5),
    ]


def _oracle(c):
    kw = {k: c[k] for k in ("seed", "centre", "radius", "cell_lo", "cells", "ppc", "E") if k in c}
    return O.generate(c["kind"], c["n"], c["h"], **kw)


def _gen_kwargs(c):
    return {k: c[k] for k in ("seed", "centre", "radius", "cell_lo", "cells", "ppc", "E", "velocity", "mass")
            if k in c}


def _sim(c, n=None):
    p = M.default_params(c["dims"], c["h"], bc_mode=M.BC_DOMAIN, theta_s=7.5e-3)
    return M.Simulation(p, c["n"] if n is None else n)


def _check_generated(got, m, E0, c, want_x, want_E0):
    n = c["n"]
    np.testing.assert_array_equal(got["x"], want_x)
    np.testing.assert_array_equal(E0, want_E0)
    np.testing.assert_array_equal(got["v"], np.tile(np.float32(c["velocity"]), (n, 1)))
    np.testing.assert_array_equal(m, np.full(n, np.float32(c["mass"])))
    np.testing.assert_array_equal(got["B"], 0)
    np.testing.assert_array_equal(got["FE"], eye9(n))
    np.testing.assert_array_equal(got["FP"], eye9(n))
    np.testing.assert_array_equal(got["V0"], 0)


@pytest.mark.parametrize("idx", range(4), ids=["ball", "ball_n1", "cells", "lattice_demo"])
def test_generator_bit_equal_to_oracle(idx):
    c = _cases()[idx]
    print(f"generator case {idx}: kind {c['kind']} n {c['n']}", flush=True)
    want_x, want_E0, want_kept = _oracle(c)
    g = _sim(c)
    kept = g.generate(c["kind"], **_gen_kwargs(c))
    assert kept == want_kept
    m, E0 = g.scene_material(0)
    _check_generated(g.particles(), m, E0, c, want_x, want_E0)
    g.close()


@pytest.mark.parametrize("idx", [0, 3], ids=["ball", "lattice_demo"])
def test_volumes_after_device_generation_match_oracle(idx):
    c = _cases()[idx]
    x, E0, _ = _oracle(c)
    g = _sim(c)
    g.generate(c["kind"], **_gen_kwargs(c))
    g.init_volumes()
    V0 = g.particles()["V0"]
    p = O.make_params(c["dims"], c["h"], bc_mode=O.BC_DOMAIN, theta_s=7.5e-3, threads=8)
    o = O.Sim(p, c["n"])
    o.set_particles(x, np.tile(np.float32(c["velocity"]), (c["n"], 1)), np.full(c["n"], np.float32(c["mass"])), E0)
    o.stage(O.STAGE_P2G)
    o.stage(O.STAGE_VOLUMES)
    assert rel_err(V0, o.particles()["V0"]) < 1e-5
    g.close()


def test_batch_scenes_generated_independently():
    cs = _cases()
    scs = [cs[0], cs[2], cs[3]]
    params = [M.default_params(c["dims"], c["h"], bc_mode=M.BC_DOMAIN, theta_s=7.5e-3) for c in scs]
    b = M.Simulation(params, [c["n"] for c in scs])
    for s in (2, 0, 1):  # any order
        b.generate(scs[s]["kind"], scene=s, **_gen_kwargs(scs[s]))
    for s, c in enumerate(scs):
        want_x, want_E0, _ = _oracle(c)
        m, E0 = b.scene_material(s)
        _check_generated(b.scene_particles(s), m, E0, c, want_x, want_E0)
    b.init_volumes()
    b.substeps(1e-4, 3)
    # the CELLS scene's 3e-6 kg particles at 4..12 per cell weigh 3-9 kg/m^3,
    # less at its faces: its elastic wave (over 150 m/s at E0 = 1e5..2e5)
    # crosses one to two cells per substep at dt = 1e-4, near the CFL
    # warning's bound; no error in any scene
    st = b.scene_status()
    assert st[0] == 0 and st[2] == 0 and st[1] in (0, M.MPM_ERR_CFL), st
    # positions readback of a batch: the scenes concatenated, each in caller order
    want = np.concatenate([b.scene_particles(s)["x"] for s in range(3)])
    np.testing.assert_array_equal(b.positions(), want)
    b.positions_begin()
    np.testing.assert_array_equal(b.positions_wait(), want)
    b.close()


def test_generator_errors():
    c = _cases()[2]
    g = _sim(c, n=4 * 495 - 1)  # fewer than ppc_lo per cell can give
    with pytest.raises(M.MpmError) as e:
        g.generate(c["kind"], **_gen_kwargs(c))
    assert e.value.code == M.MPM_ERR_ARG
    g.close()
    d = _cases()[3]
    g = _sim(d, n=d["n"] + 10)  # more particles than the lattice keeps
    with pytest.raises(M.MpmError) as e:
        g.generate(d["kind"], **_gen_kwargs(d))
    assert e.value.code == M.MPM_ERR_ARG and "candidates" in str(e.value)
    g.close()
    b = dict(_cases()[0], centre=(0.03, 0.5, 0.5))  # stencil leaves the grid
    g = _sim(b)
    with pytest.raises(M.MpmError) as e:
        g.generate(b["kind"], **_gen_kwargs(b))
    assert e.value.code == M.MPM_ERR_BOUNDS
    with pytest.raises(M.MpmError) as e:
        g.substeps(1e-4, 1)  # the scene is left unset
    assert e.value.code == M.MPM_ERR_STATE
    g.close()


def test_full_size_slab_generation():
    """The 4,194,304-particle slab (128 x 32 x 128 cells of U{4..12} at
    dx = 1/256), bit-equal to the oracle, then two substeps."""
    c = dict(kind=M.GEN_CELLS, n=4194304, dims=(256, 256, 256), h=1 / 256, seed=2, cell_lo=(64, 128, 64),
             cells=(128, 32, 128), ppc=(4, 12), E=(1e5, 2e5), velocity=(0.0, 0.0, 0.0),
             mass=np.float32(400.0 / 256 ** 3 / 8))
    want_x, want_E0, _ = _oracle(c)
    g = _sim(c)
    g.generate(c["kind"], **_gen_kwargs(c))
    m, E0 = g.scene_material(0)
    _check_generated(g.particles(), m, E0, c, want_x, want_E0)
    g.init_volumes()
    g.substeps(2.5e-5, 2)
    g.sync()
    g.close()


def _running_sim():
    c = _cases()[0]
    g = _sim(c)
    g.generate(c["kind"], **_gen_kwargs(c))
    g.set_colliders([M.collider((0.5, 0.3, 0.5), (0.9238795, 0.0, 0.0, 0.3826834), (0.3, 0.05, 0.3))])
    g.init_volumes()
    g.substeps(1e-4, 7)
    return c, g


def test_positions_readback():
    _, g = _running_sim()
    full = g.particles()["x"]
    np.testing.assert_array_equal(g.positions(), full)
    # an asynchronous read reflects the state at its begin
    g.positions_begin()
    g.substeps(1e-4, 5)
    view = g.positions_wait().copy()
    np.testing.assert_array_equal(view, full)
    after = g.particles()["x"]
    assert not np.array_equal(after, full)
    g.positions_begin()
    np.testing.assert_array_equal(g.positions_wait(), after)
    g.close()


def test_snapshot_save_load_resume(tmp_path):
    _, g = _running_sim()
    path = str(tmp_path / "run.snap")
    g.save(path, step=7)
    state = g.particles()
    m, E0 = g.scene_material(0)
    f = S.read(path)
    assert f["step"] == 7
    sc = f["scenes"][0]
    for k in ("x", "v", "B", "FE", "FP", "V0"):
        np.testing.assert_array_equal(sc[k].reshape(state[k].shape), state[k])
    np.testing.assert_array_equal(sc["m"], m)
    np.testing.assert_array_equal(sc["E0"], E0)
    assert sc["colliders"].shape == (1, 13)
    r, step = M.Simulation.from_snapshot(path)
    assert step == 7
    st2 = r.particles()
    for k in state:
        np.testing.assert_array_equal(st2[k], state[k])
    g.substeps(1e-4, 10)
    r.substeps(1e-4, 10)
    a, b = g.particles(), r.particles()
    for k in ("x", "v", "FE", "FP"):
        assert rel_err(b[k], a[k]) < 1e-6, (k, rel_err(b[k], a[k]))
    # loading rewinds a running context to the snapshot
    assert g.load(path) == 7
    for k, v in g.particles().items():
        np.testing.assert_array_equal(v, state[k])
    g.close()
    r.close()


def test_numpy_fixture_loads_and_steps_like_the_oracle(tmp_path):
    rng = np.random.default_rng(4)
    c = _cases()[0]
    x, E0, _ = _oracle(c)
    n = c["n"]
    v = rng.normal(scale=0.5, size=(n, 3)).astype(np.float32)
    FE = (eye9(n) + rng.normal(scale=1e-3, size=(n, 9))).astype(np.float32)
    FP = (eye9(n) + rng.normal(scale=1e-3, size=(n, 9))).astype(np.float32)
    B = rng.normal(scale=0.1, size=(n, 9)).astype(np.float32)
    mm = np.full(n, np.float32(c["mass"]))
    V0 = np.full(n, np.float32((1 / 64) ** 3 / 8))
    p = M.default_params(c["dims"], c["h"], bc_mode=M.BC_DOMAIN, theta_s=7.5e-3)
    path = str(tmp_path / "fixture.snap")
    S.write(path, [dict(params=p, x=x, v=v, B=B, FE=FE, FP=FP, m=mm, V0=V0, E0=E0)], step=0)
    g, _ = M.Simulation.from_snapshot(path)
    g.substeps(1e-4, 1)
    got = g.particles()
    po = O.make_params(c["dims"], c["h"], bc_mode=O.BC_DOMAIN, theta_s=7.5e-3, threads=8)
    o = O.Sim(po, n)
    o.set_particles(x, v, mm, E0, B=B, FE=FE, FP=FP, V0=V0)
    o.substeps(1e-4, 1)
    ref = o.particles()
    for k in ("x", "v", "FE", "FP"):
        assert rel_err(got[k], ref[k]) < 1e-5, (k, rel_err(got[k], ref[k]))
    g.close()
