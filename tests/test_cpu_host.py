"""Host-side checks that need no GPU: the C ABI library loads and exports
every function include/mpm_c.h declares; the scene generators meet their
specifications; the CPU binning reference is a valid stable sort with the
documented round-robin layout."""
import ctypes
import os
import re

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


def test_binning_reference_is_a_stable_sort_with_round_robin_blocks():
    sc = scenes.two_snowballs(n_total=20000, grid=64)
    keys = O.bin_keys(sc["x"], sc["h"], sc["dims"])
    sk, perm, slot = O.bin_reference(sc["x"], sc["h"], sc["dims"])
    assert np.array_equal(np.sort(perm), np.arange(perm.size))
    assert np.array_equal(keys[perm], sk) and np.all(np.diff(sk.astype(np.int64)) >= 0)
    # stability: equal keys keep index order
    same = np.flatnonzero(np.diff(sk.astype(np.int64)) == 0)
    assert np.all(perm[same] < perm[same + 1])
    # slot layout: each block's range is a permutation of its particles and
    # the first min-count rounds list 64 distinct cells in cell order
    blk = sk >> 6
    starts = np.flatnonzero(np.r_[True, blk[1:] != blk[:-1]])
    ends = np.r_[starts[1:], sk.size]
    for s, e in zip(starts[:200], ends[:200]):
        assert sorted(slot[s:e]) == sorted(perm[s:e])
        cells = keys[slot[s:e]] & 63
        cnt = np.bincount(cells, minlength=64)
        full = cnt.min()
        for r in range(full):
            assert np.array_equal(cells[64 * r:64 * r + 64], np.arange(64))
