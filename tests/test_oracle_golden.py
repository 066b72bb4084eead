"""The oracle against the reference's own math (CPU).

Fixtures in tests/golden/reference_math.npz were produced by
tests/golden/make_golden.py from the reference's vendored Eigen 3.4.90
JacobiSVD, glm 0.9.7.1 and polar decomposition (compiled in oracle/_ref).
The oracle restates them from the algorithm; FP32 results must be equal bit
for bit (same operation order, no contraction), the polar factor within the
reference's own FP32 accuracy.
"""
import os

import numpy as np
import pytest

import oracle as O

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_math.npz"))


def test_jacobi_svd_bit_exact_against_eigen():
    M = GOLD["svd_M"]
    for i in range(M.shape[0]):
        rc, U, S, V = O.svd3(M[i])
        assert rc == GOLD["svd_rc"][i]
        np.testing.assert_array_equal(U.ravel(), GOLD["svd_U"][i], err_msg=f"case {i}")
        np.testing.assert_array_equal(S, GOLD["svd_S"][i], err_msg=f"case {i}")
        np.testing.assert_array_equal(V.ravel(), GOLD["svd_V"][i], err_msg=f"case {i}")


def test_literal_f_update_bit_exact_against_reference_composition():
    h = float(GOLD["fu_h"])
    for i in range(GOLD["fu_B"].shape[0]):
        rc, FE, FP, sig = O.fupdate(O.LITERAL, float(GOLD["fu_dt"][i]), h, 2.5e-2, 5e-3, GOLD["fu_B"][i],
                                    GOLD["fu_FE"][i], GOLD["fu_FP"][i])
        assert rc == GOLD["fu_rc"][i]
        np.testing.assert_array_equal(sig, GOLD["fu_sig"][i], err_msg=f"case {i}")
        np.testing.assert_array_equal(FE, GOLD["fu_FEo"][i], err_msg=f"case {i}")
        np.testing.assert_array_equal(FP, GOLD["fu_FPo"][i], err_msg=f"case {i}")


def test_polar_matches_reference_higham_noferini():
    F = GOLD["polar_F"]
    worst = 0.0
    for i in range(F.shape[0]):
        R = O.polar(F[i])
        worst = max(worst, float(np.abs(R - GOLD["polar_R"][i]).max()))
    assert worst < 1e-6, worst


def test_glm_mat3_ops_bit_exact():
    for i in range(GOLD["glm_A"].shape[0]):
        p, q, d = O.mat3_ops(GOLD["glm_A"][i], GOLD["glm_B"][i])
        np.testing.assert_array_equal(p, GOLD["glm_prod"][i])
        np.testing.assert_array_equal(q, GOLD["glm_inv"][i])
        assert d == GOLD["glm_det"][i]


def test_collider_pose_inverse_and_sdf_bit_exact():
    c13, pts = GOLD["col_c13"], GOLD["col_pts"]
    for c in range(c13.shape[0]):
        np.testing.assert_array_equal(O.collider_inverse(c13[c]), GOLD["col_inv"][c])
        for j in range(pts.shape[1]):
            assert O.sdf(c13[c], pts[c, j]) == GOLD["col_sdf"][c, j], (c, j)


def test_fresh_reference_vectors_agree():
    """Beyond the committed fixtures: fresh random inputs through the
    compiled reference math, where it is built.  The library is looked up
    here, not at collection, so a run that deselects this test never loads
    it."""
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref needs the reference tree (this container only)")
    rng = np.random.default_rng(99)
    for _ in range(2000):
        M = (np.eye(3) + rng.normal(0, rng.choice([1e-6, 1e-3, 1e-1]), (3, 3))).astype(np.float32).ravel()
        u, s, v = np.zeros(9, np.float32), np.zeros(3, np.float32), np.zeros(9, np.float32)
        R.ref_svd3_f(O.fptr(M), O.fptr(u), O.fptr(s), O.fptr(v))
        rc, U, S, V = O.svd3(M)
        assert np.array_equal(U.ravel(), u) and np.array_equal(S, s) and np.array_equal(V.ravel(), v)
