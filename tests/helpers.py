"""Shared comparison helpers for the parity tests."""
import numpy as np


def rel_err(a, b):
    """Norm-relative error ||a - b||_2 / ||b||_2 (absolute when b == 0)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d


def max_abs(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)))) if np.size(a) else 0.0


def eye9(n):
    e = np.zeros((n, 9), np.float32)
    e[:, 0] = e[:, 4] = e[:, 8] = 1.0
    return e


def to_math(m9):
    """glm column-major 9-vectors -> (n,3,3) math matrices."""
    return np.asarray(m9).reshape(-1, 3, 3).transpose(0, 2, 1)


def from_math(M):
    return np.ascontiguousarray(np.asarray(M).transpose(0, 2, 1).reshape(-1, 9))


def warm_state(rng, n, scale=1e-2, bscale=1.0):
    """A plausible non-trivial particle state: B random, F_E with singular
    values inside the clamp band, F_P near identity."""
    B = rng.normal(0, bscale, (n, 9)).astype(np.float32)
    Q1 = np.linalg.qr(rng.normal(size=(n, 3, 3)))[0]
    Q2 = np.linalg.qr(rng.normal(size=(n, 3, 3)))[0]
    s = rng.uniform(0.976, 1.004, (n, 3))
    FE = np.einsum("nij,nj,nkj->nik", Q1, s, Q2)
    FE *= np.sign(np.linalg.det(FE))[:, None, None]
    FP = np.eye(3) + rng.normal(0, scale, (n, 3, 3))
    return B, from_math(FE).astype(np.float32), from_math(FP).astype(np.float32)
