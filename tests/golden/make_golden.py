"""Write the golden fixtures of tests/golden/ from the reference's own math.

Runs only where the reference tree is present: oracle/_ref/libref_math.so is
built (oracle/Makefile, target `ref`) from the reference's vendored Eigen
3.4.90 JacobiSVD, glm 0.9.7.1 and Higham–Noferini polar decomposition, and
the deformation-gradient update is composed from them in the order of
realtime-deformations/material_point_method.cpp:306-330.  The fixtures hold
inputs and the reference's outputs only; the tests compare the oracle's
restatement against them.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402


def _svd_inputs(rng):
    """Row-major 3x3 inputs covering the Jacobi branches."""
    cases = []
    for sig in (1e-7, 1e-5, 1e-3, 1e-2, 1e-1):  # near identity
        cases += [np.eye(3) + rng.normal(0, sig, (3, 3)) for _ in range(100)]
    cases += [rng.normal(0, 1, (3, 3)) for _ in range(150)]  # general
    for _ in range(100):  # diagonal with ties and permutations
        d = rng.choice([0.97, 1.0, 1.0, 1.004, 1.3], 3)
        P = np.eye(3)[rng.permutation(3)]
        cases.append(P @ np.diag(d) + (rng.normal(0, 1e-7, (3, 3)) if rng.random() < 0.5 else 0))
    for _ in range(50):  # rotations times stretches (F_E-like)
        Q1 = np.linalg.qr(rng.normal(size=(3, 3)))[0]
        Q2 = np.linalg.qr(rng.normal(size=(3, 3)))[0]
        cases.append(Q1 @ np.diag(rng.uniform(0.97, 1.01, 3)) @ Q2.T)
    for _ in range(25):  # rank deficient
        u, v = rng.normal(size=3), rng.normal(size=3)
        cases.append(np.outer(u, v))
    cases += [np.zeros((3, 3)), np.eye(3), -np.eye(3), np.diag([2.0, -1.0, 0.5])]
    return np.array(cases, np.float32).reshape(-1, 9)


def main():
    O.build()
    R = O.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libref_math.so is missing: needs the reference tree")
    rng = np.random.default_rng(20261202)
    out = {}

    M = _svd_inputs(rng)
    U, S, V, ok = np.zeros_like(M), np.zeros((M.shape[0], 3), np.float32), np.zeros_like(M), np.zeros(M.shape[0], np.int32)
    for i in range(M.shape[0]):
        u, s, v = np.zeros(9, np.float32), np.zeros(3, np.float32), np.zeros(9, np.float32)
        ok[i] = R.ref_svd3_f(O.fptr(M[i]), O.fptr(u), O.fptr(s), O.fptr(v))
        U[i], S[i], V[i] = u, s, v
    out.update(svd_M=M, svd_U=U, svd_S=S, svd_V=V, svd_rc=ok)

    n = 600
    h = np.float32(1 / 64)
    dts = rng.choice(np.float32([1e-5, 1e-4, 5e-4]), n).astype(np.float32)
    Bs = (rng.normal(0, 1, (n, 9)) * rng.choice([1e-3, 1e-1, 1.0, 10.0], (n, 1))).astype(np.float32)
    FEs, FPs = np.zeros((n, 9), np.float32), np.zeros((n, 9), np.float32)
    for i in range(n):
        Q1 = np.linalg.qr(rng.normal(size=(3, 3)))[0]
        Q2 = np.linalg.qr(rng.normal(size=(3, 3)))[0]
        fe = Q1 @ np.diag(rng.uniform(0.975, 1.005, 3)) @ Q2.T
        fe *= np.sign(np.linalg.det(fe))
        if i % 5 == 0:
            fe = np.eye(3)
        fp = np.eye(3) + rng.normal(0, 3e-2, (3, 3)) if i % 3 else np.eye(3)
        FEs[i] = fe.T.reshape(9)  # glm column-major
        FPs[i] = fp.T.reshape(9)
    FEo, FPo, sig, rc = np.zeros_like(FEs), np.zeros_like(FPs), np.zeros((n, 3), np.float32), np.zeros(n, np.int32)
    for i in range(n):
        a, b, s = np.zeros(9, np.float32), np.zeros(9, np.float32), np.zeros(3, np.float32)
        rc[i] = R.ref_fupdate_f(float(dts[i]), float(h), O.fptr(Bs[i]), O.fptr(FEs[i]), O.fptr(FPs[i]), O.fptr(a),
                                O.fptr(b), O.fptr(s))
        FEo[i], FPo[i], sig[i] = a, b, s
    out.update(fu_dt=dts, fu_h=np.float32(h), fu_B=Bs, fu_FE=FEs, fu_FP=FPs, fu_FEo=FEo, fu_FPo=FPo, fu_sig=sig,
               fu_rc=rc)

    m = 500
    F = np.zeros((m, 9), np.float32)
    for i in range(m):
        sig_ = [1e-3, 1e-1, 0.3, 0.6][i % 4]
        F[i] = (np.eye(3) + rng.normal(0, sig_, (3, 3))).T.reshape(9)
    Rp, Hp = np.zeros_like(F), np.zeros_like(F)
    for i in range(m):
        r, hh = np.zeros(9, np.float32), np.zeros(9, np.float32)
        R.ref_polar_f(O.fptr(F[i]), O.fptr(r), O.fptr(hh))
        Rp[i], Hp[i] = r, hh
    out.update(polar_F=F, polar_R=Rp, polar_H=Hp)

    k = 300
    A = (rng.normal(0, 1, (k, 9)) + np.eye(3).reshape(9)).astype(np.float32)
    Bm = rng.normal(0, 1, (k, 9)).astype(np.float32)
    prod, inv, det = np.zeros_like(A), np.zeros_like(A), np.zeros(k, np.float32)
    for i in range(k):
        p_, q_, d_ = np.zeros(9, np.float32), np.zeros(9, np.float32), np.zeros(1, np.float32)
        R.ref_mat3_ops_f(O.fptr(A[i]), O.fptr(Bm[i]), O.fptr(p_), O.fptr(q_), O.fptr(d_))
        prod[i], inv[i], det[i] = p_, q_, d_[0]
    out.update(glm_A=A, glm_B=Bm, glm_prod=prod, glm_inv=inv, glm_det=det)

    # box colliders: main.cpp's three plus random poses; SDF at random points
    cols = []
    for t, ang, s in (((0.5, -0.2, 0.5), 0.0, (0.4, 0.4, 0.4)), ((0.0, 0.3, 0.5), 45.0, (0.2, 0.2, 0.3)),
                      ((1.0, 0.3, 0.5), 45.0, (0.2, 0.2, 0.3))):
        a = math.radians(ang) / 2
        cols.append(list(t) + [math.cos(a), 0.0, 0.0, math.sin(a)] + list(s) + [0.0, 0.0, 0.0])
    for _ in range(20):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        cols.append(list(rng.uniform(0.2, 0.8, 3)) + list(q) + list(rng.uniform(0.05, 0.3, 3)) + [0.0, 0.0, 0.0])
    cols = np.array(cols, np.float32)
    pts = rng.uniform(-0.1, 1.1, (cols.shape[0], 40, 3)).astype(np.float32)
    sdf = np.zeros(pts.shape[:2], np.float32)
    cinv = np.zeros((cols.shape[0], 16), np.float32)
    for c in range(cols.shape[0]):
        inv16 = np.zeros(16, np.float32)
        R.ref_collider_inverse_f(O.fptr(cols[c]), O.fptr(inv16))
        cinv[c] = inv16
        for j in range(pts.shape[1]):
            sdf[c, j] = R.ref_sdf_f(O.fptr(cols[c]), O.fptr(np.ascontiguousarray(pts[c, j])))
    out.update(col_c13=cols, col_pts=pts, col_sdf=sdf, col_inv=cinv)

    path = os.path.join(HERE, "reference_math.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
