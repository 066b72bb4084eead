"""Host-side checks that need no GPU: the C ABI library loads and exports
every function include/mpm_c.h declares; the scene generators meet their
specifications; the CPU binning reference is a valid stable sort with the
documented in-block slot layout; bench.py refuses more than one GPU."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import mpm_amd as M
import oracle as O
from mpm_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*|int32_t)\s+(mpm_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    names = _declared("mpm_c.h")
    assert len(names) >= 15
    lib = ctypes.CDLL(M.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(M.EXPORTS) == sorted(names)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


def test_default_params_are_the_reference_constants():
    p = M.default_params((20, 20, 20), 0.05)
    assert tuple(p.dims) == (20, 20, 20)
    assert p.h == pytest.approx(0.05) and p.E == pytest.approx(1.4e5) and p.nu == pytest.approx(0.2)
    assert p.xi == 10.0 and p.theta_c == 2.5e-2 and p.theta_s == 5e-3
    assert tuple(p.gravity) == pytest.approx((0.0, -9.8, 0.0))
    assert p.compose == M.COMPOSE_LITERAL


def test_snowball_scene():
    sc = scenes.snowball()
    assert sc["x"].shape == (8192, 3)
    r = np.linalg.norm(sc["x"] - np.float32([0.5, 0.6, 0.5]), axis=1)
    assert r.max() <= 0.08 + 1e-6
    assert sc["m"].sum() == pytest.approx(400 * 4 / 3 * np.pi * 0.08 ** 3, rel=1e-5)
    assert np.array_equal(scenes.snowball()["x"], sc["x"])  # seeded


def test_two_snowball_scene():
    sc = scenes.two_snowballs()
    assert sc["x"].shape == (262144, 3)
    assert np.all(sc["v"][:131072, 0] == 3) and np.all(sc["v"][131072:, 0] == -3)


def test_slab_scene_counts():
    sc = scenes.slab(n_total=2 ** 16, cells=(32, 8, 32), grid=128)
    x, h = sc["x"], sc["h"]
    assert x.shape == (2 ** 16, 3)
    cells = np.floor(x / h).astype(np.int64)
    _, counts = np.unique(cells[:, 0] * 10 ** 6 + cells[:, 1] * 10 ** 3 + cells[:, 2], return_counts=True)
    # float32 positions at a cell's upper face may round into the next cell
    assert counts.size == 32 * 8 * 32 and counts.min() >= 3 and counts.max() <= 13
    assert np.mean((counts >= 4) & (counts <= 12)) > 0.99
    assert np.all((sc["E0"] >= 1e5) & (sc["E0"] <= 2e5))
    assert x[:, 1].min() >= 0.5 - 1e-6


def test_binning_reference_is_a_stable_sort_in_key_order():
    for sc in (scenes.two_snowballs(n_total=20000, grid=64), scenes.slab(n_total=2 ** 16, cells=(32, 8, 32), grid=128)):
        keys = O.bin_keys(sc["x"], sc["h"], sc["dims"])
        sk, perm, slot = O.bin_reference(sc["x"], sc["h"], sc["dims"])
        assert np.array_equal(np.sort(perm), np.arange(perm.size))
        assert np.array_equal(keys[perm], sk) and np.all(np.diff(sk.astype(np.int64)) >= 0)
        # stability: equal keys keep index order
        same = np.flatnonzero(np.diff(sk.astype(np.int64)) == 0)
        assert np.all(perm[same] < perm[same + 1])
        # slots in key order: each block's particles contiguous, cell by cell
        assert np.array_equal(slot, perm)
        blk = sk >> 6
        assert np.unique(blk).size == np.flatnonzero(np.r_[True, blk[1:] != blk[:-1]]).size


class _StubSim:
    """Stands in for a device context: the reverse-mode record is the last
    forward's tag; backward returns gradients tagged with it."""

    def __init__(self, n):
        self.n, self.rollout_mass, self.adjoint_generation = n, np.ones(n, np.float32), 0
        self.record = None

    def set_particles(self, x, v, m, E0=None):
        self.x = np.asarray(x)

    def init_volumes(self):
        pass

    def adjoint_forward(self, dt, nsteps, every=0):
        self.record = float(self.x[0, 0])
        self.adjoint_generation += 1

    def particles(self):
        return dict(x=self.x.astype(np.float32), v=np.zeros_like(self.x, np.float32))

    def adjoint_backward(self, gx=None, gv=None):
        g = np.full((self.n, 3), self.record, np.float32)
        return g, g, np.full(self.n, self.record, np.float32)


def test_rollout_backward_refuses_a_replaced_record():
    """One reverse-mode record per context: a second Rollout on the same
    context replaces the first one's record, so the first one's backward
    raises instead of returning the second trajectory's gradients; a rollout
    whose record is current differentiates normally."""
    torch = pytest.importorskip("torch")
    from mpm_amd.autograd import Rollout

    sim = _StubSim(4)
    E0 = torch.ones(4, requires_grad=True)
    v0 = torch.zeros(4, 3, requires_grad=True)
    xa = torch.full((4, 3), 1.0, requires_grad=True)
    xb = torch.full((4, 3), 2.0, requires_grad=True)
    outa = Rollout.apply(sim, 1e-4, 3, 0, xa, v0, E0)
    outb = Rollout.apply(sim, 1e-4, 3, 0, xb, v0, E0)
    with pytest.raises(RuntimeError, match="another adjoint forward"):
        outa[0].sum().backward()
    outb[0].sum().backward()
    assert torch.all(xb.grad == 2.0)
