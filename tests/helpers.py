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


def near_tied_spectra(rng, n):
    """F_E with two singular values within 1e-7..1e-4 (relative) of each
    other, straddling a clamp bound (1 + 7.5e-3 or 1 - 2.5e-2), the third
    inside the band; B = 0 and F_P = I, so F_hat = F_E.  Returns glm
    column-major (B, F_E, F_P)."""
    hi, lo = np.float64(np.float32(1 + 7.5e-3)), np.float64(np.float32(1 - 2.5e-2))
    Q1 = np.linalg.qr(rng.normal(size=(n, 3, 3)))[0]
    Q2 = np.linalg.qr(rng.normal(size=(n, 3, 3)))[0]
    rel = 10 ** rng.uniform(-7, -4, n)
    mid = np.where(rng.uniform(size=n) < 0.5, hi, lo)
    s = np.stack([mid * (1 + rel), mid * (1 - rel), np.full(n, 0.99)], 1)
    FE = np.einsum("nij,nj,nkj->nik", Q1, s, Q2)
    FE *= np.sign(np.linalg.det(FE))[:, None, None]
    return (np.zeros((n, 9), np.float32), from_math(FE).astype(np.float32),
            from_math(np.tile(np.eye(3), (n, 1, 1))).astype(np.float32))


def fp64_textbook_update(B, FE, FP, dt=1e-4, h=1 / 64, theta_s=7.5e-3):
    """The oracle's FP64 build of the TEXTBOOK F update on FP32 inputs."""
    import oracle as O

    n = FE.shape[0]
    x = np.full((n, 3), 0.5, np.float32)
    o = O.Sim(O.make_params((64, 64, 64), h, compose=O.TEXTBOOK, theta_s=theta_s), n, double=True)
    o.set_particles(x, np.zeros_like(x), np.ones(n, np.float32), B=B, FE=FE, FP=FP)
    o.stage(O.STAGE_FUPDATE, dt)
    return o.particles()
