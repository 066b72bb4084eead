"""The oracle's seeded scene generators (BALL, CELLS, LATTICE), pinned on
the CPU before they are used to check the device generators bit for bit.

- Philox4x64-10 is pinned against numpy.random.Philox (an independent
  implementation of the same published generator), including counter carry
  across words.
- The float draws are pinned by distribution checks that a plausible slip
  fails: uniform in the ball (mass inside r/2 is 1/8, centred), counts in
  range and summing exactly to n with at most |T - n| cells moved, points
  inside their own cell, LATTICE candidates within a quarter cell of their
  quarter-offset site and within the width (material_point_method.cpp:18-63),
  filled from the back in the reference's loop order.
"""
import numpy as np
import pytest

import oracle as O

MASK = (1 << 64) - 1


def _inc(ctr):
    c = [int(v) for v in ctr]
    for i in range(4):
        c[i] = (c[i] + 1) & MASK
        if c[i] != 0:
            break
    return c


@pytest.mark.parametrize("key,ctr", [((0, 0), (0, 0, 0, 0)), ((123, 456), (7, 8, 9, 10)),
                                     ((MASK, MASK), (MASK, MASK, 5, MASK)), ((2**63 + 1, 17), (MASK, 3, 0, 0)),
                                     ((0xDEADBEEF, 0), (MASK, MASK, MASK, MASK))])
def test_philox_matches_numpy(key, ctr):
    # numpy's Philox advances the counter before producing each block
    bg = np.random.Philox(key=np.array(key, np.uint64), counter=np.array(ctr, np.uint64))
    want = bg.random_raw(8)
    c1 = _inc(ctr)
    c2 = _inc(c1)
    got = np.concatenate([O.philox4x64(c1, key), O.philox4x64(c2, key)])
    np.testing.assert_array_equal(got, want)


def test_philox_zero_block():
    # counter 0, key 0 (numpy reaches it by wrapping an all-ones counter)
    z = O.philox4x64((0, 0, 0, 0), (0, 0))
    np.testing.assert_array_equal(z, np.array([0x16554D9ECA36314C, 0xDB20FE9D672D0FDC, 0xD7E772CEE186176B,
                                               0x7E68B68AEC7BA23B], np.uint64))


def test_ball_is_uniform_in_the_ball():
    n, c, r = 40000, np.float32([0.4, 0.5, 0.6]), 0.1
    x, E0, kept = O.generate(O.GEN_BALL, n, 1 / 64, seed=7, centre=c, radius=r, E=(1e5, 2e5))
    assert kept == n
    d = np.linalg.norm(x.astype(np.float64) - c, axis=1)
    assert d.max() <= r * (1 + 1e-5)
    # volume fractions: |d| < r/2 holds 1/8, |d| < r * 0.5^(1/3) holds 1/2
    assert abs(np.mean(d < r / 2) - 0.125) < 0.01
    assert abs(np.mean(d < r * 0.5 ** (1 / 3)) - 0.5) < 0.015
    assert np.all(np.abs(x.mean(0) - c) < 0.01 * r)
    assert E0.min() >= 1e5 and E0.max() < 2e5 and abs(E0.mean() - 1.5e5) < 2e3
    # another seed gives other points; the same seed the same ones
    x2, _, _ = O.generate(O.GEN_BALL, n, 1 / 64, seed=8, centre=c, radius=r)
    x3, _, _ = O.generate(O.GEN_BALL, n, 1 / 64, seed=7, centre=c, radius=r)
    assert not np.array_equal(x, x2)
    np.testing.assert_array_equal(x, x3)


@pytest.mark.parametrize("n_of", ["mean", "low", "high"])
def test_cells_counts_and_placement(n_of):
    h, lo, cells, ppc = 1 / 64, (10, 20, 12), (9, 5, 11), (4, 12)
    nc = int(np.prod(cells))
    n = {"mean": 8 * nc + 37, "low": 4 * nc + 3, "high": 12 * nc - 5}[n_of]
    x, E0, kept = O.generate(O.GEN_CELLS, n, h, seed=11, cell_lo=lo, cells=cells, ppc=ppc, E=(1e5, 2e5))
    assert kept == n
    cell = np.floor(x.astype(np.float64) / h).astype(np.int64) - np.array(lo)
    # points lie in their cell (a draw of exactly 1.0 - ulp may round to the upper face)
    frac = x.astype(np.float64) / h - np.floor(x.astype(np.float64) / h)
    assert np.all((cell >= 0) & (cell <= np.array(cells)))
    q = (np.clip(cell[:, 0], 0, cells[0] - 1) * cells[1] + np.clip(cell[:, 1], 0, cells[1] - 1)) * cells[2] + \
        np.clip(cell[:, 2], 0, cells[2] - 1)
    # particles come cell after cell in q order
    assert np.all(np.diff(q) >= 0)
    cnt = np.bincount(q, minlength=nc)
    assert cnt.sum() == n and cnt.min() >= ppc[0] and cnt.max() <= ppc[1]
    # the draws before adjustment: every cell but |T - n| moved ones keeps it
    raw = np.array([ppc[0] + ((int(O.philox4x64((k, 0, 0, 0), (11, 2))[0]) >> 32) * 9 >> 32) for k in range(nc)])
    moved = np.abs(cnt - raw)
    assert moved.max() <= ppc[1] - ppc[0]
    assert moved.sum() == abs(int(raw.sum()) - n)
    assert np.all(np.sign(cnt - raw) * np.sign(raw.sum() - n) <= 0)
    assert 0.35 < frac.mean() < 0.65


def test_cells_rejects_unreachable_counts():
    with pytest.raises(ValueError):
        O.generate(O.GEN_CELLS, 3 * 8 - 1, 1 / 64, cells=(2, 2, 2), ppc=(3, 5))
    with pytest.raises(ValueError):
        O.generate(O.GEN_CELLS, 5 * 8 + 1, 1 / 64, cells=(2, 2, 2), ppc=(3, 5))


def test_lattice_follows_initialize_particles():
    h, origin, width = 0.05, np.float32([0.5, 0.6, 0.5]), 0.2
    # count the candidates the rule keeps, then ask for exactly that many
    with pytest.raises(ValueError) as e:
        O.generate(O.GEN_LATTICE, 10 ** 6, h, seed=3, centre=origin, radius=width)
    kept = int(str(e.value).split("kept ")[1].split(" ")[0])
    # about (4/3) pi (width/h)^3 cells x 8 candidates
    assert abs(kept / (4 / 3 * np.pi * (width / h) ** 3 * 8) - 1) < 0.1
    n = kept - 5
    x, _, k2 = O.generate(O.GEN_LATTICE, n, h, seed=3, centre=origin, radius=width)
    assert k2 == kept
    d = np.linalg.norm(x - origin, axis=1)
    assert d.max() <= width
    # each point is within a quarter cell of a quarter-offset site
    s = x.astype(np.float64) / h
    site = np.floor(s * 2) / 2 + 0.25
    assert np.all(np.linalg.norm(s - site, axis=1) <= 0.25 + 1e-4)
    # filled from the back in loop order (i, j, k, offset): the last slot
    # holds the first kept candidate, so cell indices descend towards slot 0
    cell = np.floor(s).astype(np.int64)
    lin = (cell[:, 0] * 1000 + cell[:, 1]) * 1000 + cell[:, 2]
    assert np.all(np.diff(lin[::-1]) >= 0)
    w = int(np.float32(width) / np.float32(h))
    c0 = (origin / np.float32(h)).astype(np.int64)
    assert cell.min(0).tolist() >= (c0 - w).tolist() and cell.max(0).tolist() <= (c0 + w - 1).tolist()
