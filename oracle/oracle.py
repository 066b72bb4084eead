"""ORACLE — test infrastructure only.

ctypes view of oracle/liboracle.so (the CPU restatement of the reference's MPM
snow substep, oracle/mpm_oracle.hpp) and, where it was built, of
oracle/_ref/libref_math.so (the reference's own vendored Eigen/glm/polar math).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  Matrices are 9 floats per particle in glm column-major order
(element (row r, col c) at index 3c + r).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_math.so")

STAGE_P2G, STAGE_VOLUMES, STAGE_FORCES, STAGE_GRID, STAGE_COLLIDE, STAGE_FUPDATE, STAGE_G2P, STAGE_ADVECT = range(8)
LITERAL, TEXTBOOK = 0, 1
BC_COLLIDERS, BC_DOMAIN = 1, 2


class Params(C.Structure):
    _fields_ = [
        ("dims", C.c_int32 * 3),
        ("compose", C.c_int32),
        ("bc_mode", C.c_int32),
        ("threads", C.c_int32),
        ("h", C.c_double),
        ("E", C.c_double),
        ("nu", C.c_double),
        ("xi", C.c_double),
        ("theta_c", C.c_double),
        ("theta_s", C.c_double),
        ("g", C.c_double * 3),
    ]


_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_lib = None
_ref = None


def build():
    """Compile liboracle.so (and _ref when the reference tree is present)."""
    import subprocess

    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])
    if os.path.isdir("/root/reference/realtime-deformations"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(Params), C.c_int, C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_particles.argtypes = [C.c_void_p] + [_F] * 8
        L.orc_set_colliders.argtypes = [C.c_void_p, _F, C.c_int]
        L.orc_set_threads.argtypes = [C.c_void_p, C.c_int]
        L.orc_stage.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.orc_substeps.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.orc_get_particles.argtypes = [C.c_void_p] + [_D] * 6
        L.orc_get_grid.argtypes = [C.c_void_p, _D, _D, _D, C.POINTER(C.c_uint8)]
        L.orc_num_active.argtypes = [C.c_void_p]
        L.orc_set_grid.argtypes = [C.c_void_p, _D, _D]
        L.orc_weight_f.restype = C.c_float
        L.orc_weight_f.argtypes = [C.c_float]
        L.orc_weight_d.restype = C.c_double
        L.orc_weight_d.argtypes = [C.c_double]
        L.orc_svd3_f.argtypes = [_F] * 4
        L.orc_svd3_d.argtypes = [_D] * 4
        L.orc_polar_f.argtypes = [_F, _F]
        L.orc_mat3_ops_f.argtypes = [_F] * 5
        L.orc_fupdate_f.argtypes = [C.c_int, C.c_float, C.c_float, C.c_double, C.c_double] + [_F] * 6
        L.orc_sdf_f.restype = C.c_float
        L.orc_sdf_f.argtypes = [_F, _F]
        L.orc_collider_inverse_f.argtypes = [_F, _F]
        L.orc_collide_f.argtypes = [_F, C.c_int, _F, _F, _F]
        L.orc_domain_walls_f.argtypes = [C.c_int] * 6 + [_F, _F]
        _U64 = C.POINTER(C.c_uint64)
        _I = C.POINTER(C.c_int)
        L.orc_philox4x64.argtypes = [_U64, _U64, _U64]
        L.orc_generate.restype = C.c_int64
        L.orc_generate.argtypes = [C.c_int, C.c_uint64, _F, C.c_float, _I, _I, C.c_int, C.c_int, C.c_float,
                                   C.c_float, C.c_float, C.c_int, _F, _F]
        _lib = L
    return _lib


def ref():
    """The reference's vendored math (oracle/_ref), or None where not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = C.CDLL(REF_PATH)
        R.ref_svd3_f.argtypes = [_F] * 4
        R.ref_polar_f.argtypes = [_F] * 3
        R.ref_mat3_ops_f.argtypes = [_F] * 5
        R.ref_fupdate_f.argtypes = [C.c_float, C.c_float] + [_F] * 6
        R.ref_sdf_f.restype = C.c_float
        R.ref_sdf_f.argtypes = [_F, _F]
        R.ref_collider_inverse_f.argtypes = [_F, _F]
        _ref = R
    return _ref


def fptr(a):
    return a.ctypes.data_as(_F) if a is not None else None


def dptr(a):
    return a.ctypes.data_as(_D) if a is not None else None


def f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def make_params(dims, h, compose=LITERAL, bc_mode=0, E=1.4e5, nu=0.2, xi=10.0, theta_c=2.5e-2, theta_s=5e-3,
                g=(0.0, -9.8, 0.0), threads=1):
    p = Params()
    p.dims[:] = list(dims)
    p.compose, p.bc_mode, p.threads = compose, bc_mode, threads
    p.h, p.E, p.nu, p.xi, p.theta_c, p.theta_s = h, E, nu, xi, theta_c, theta_s
    p.g[:] = list(g)
    return p


class Sim:
    """One oracle simulation (FP32 mirror or FP64 build)."""

    def __init__(self, params: Params, n: int, double: bool = False):
        self.L = lib()
        self.n = n
        self.dims = tuple(params.dims)
        self.h = C.c_void_p(self.L.orc_create(C.byref(params), n, 1 if double else 0))

    def __del__(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.L.orc_destroy(self.h)
            self.h = None

    def set_particles(self, x, v, m, E0=None, B=None, FE=None, FP=None, V0=None):
        arrs = [f32(a) for a in (x, v, m, E0, B, FE, FP, V0)]
        self.L.orc_set_particles(self.h, *[fptr(a) for a in arrs])

    def set_colliders(self, c13):
        c = f32(np.asarray(c13).reshape(-1, 13))
        self.L.orc_set_colliders(self.h, fptr(c), c.shape[0])

    def set_threads(self, t):
        self.L.orc_set_threads(self.h, int(t))

    def stage(self, st, dt=0.0):
        rc = self.L.orc_stage(self.h, st, dt)
        if rc:
            raise RuntimeError(f"oracle stage {st} failed: {rc}")

    def substeps(self, dt, n):
        rc = self.L.orc_substeps(self.h, dt, n)
        if rc:
            raise RuntimeError(f"oracle substeps failed: {rc}")

    def particles(self):
        n = self.n
        out = dict(x=np.zeros((n, 3)), v=np.zeros((n, 3)), B=np.zeros((n, 9)), FE=np.zeros((n, 9)),
                   FP=np.zeros((n, 9)), V0=np.zeros(n))
        self.L.orc_get_particles(self.h, *[dptr(out[k]) for k in ("x", "v", "B", "FE", "FP", "V0")])
        return out

    def set_grid(self, m, vel):
        """Dense node masses (cells) and velocities (cells x 3); the used set
        becomes the nodes of non-zero mass."""
        m = np.ascontiguousarray(m, np.float64)
        vel = np.ascontiguousarray(vel, np.float64)
        self.L.orc_set_grid(self.h, dptr(m), dptr(vel))

    def grid(self):
        cells = int(np.prod(self.dims))
        m, vel, force = np.zeros(cells), np.zeros((cells, 3)), np.zeros((cells, 3))
        act = np.zeros(cells, dtype=np.uint8)
        self.L.orc_get_grid(self.h, dptr(m), dptr(vel), dptr(force), act.ctypes.data_as(C.POINTER(C.c_uint8)))
        return dict(m=m, vel=vel, force=force, active=act.astype(bool))


def svd3(M, double=False):
    L = lib()
    if double:
        M = np.ascontiguousarray(M, dtype=np.float64)
        U, S, V = np.zeros(9), np.zeros(3), np.zeros(9)
        rc = L.orc_svd3_d(dptr(M), dptr(U), dptr(S), dptr(V))
    else:
        M = f32(M)
        U, S, V = np.zeros(9, np.float32), np.zeros(3, np.float32), np.zeros(9, np.float32)
        rc = L.orc_svd3_f(fptr(M), fptr(U), fptr(S), fptr(V))
    return rc, U.reshape(3, 3), S, V.reshape(3, 3)


def fupdate(compose, dt, h, theta_c, theta_s, B, FE, FP):
    L = lib()
    B, FE, FP = f32(B), f32(FE), f32(FP)
    FEo, FPo, sig = np.zeros(9, np.float32), np.zeros(9, np.float32), np.zeros(3, np.float32)
    rc = L.orc_fupdate_f(compose, dt, h, theta_c, theta_s, fptr(B), fptr(FE), fptr(FP), fptr(FEo), fptr(FPo),
                         fptr(sig))
    return rc, FEo, FPo, sig


def polar(F):
    F = f32(F)
    R = np.zeros(9, np.float32)
    lib().orc_polar_f(fptr(F), fptr(R))
    return R


def weight(x, double=False):
    return lib().orc_weight_d(x) if double else lib().orc_weight_f(x)


def mat3_ops(A, B):
    """glm A*B, inverse(A), determinant(A) (column-major 9-vectors)."""
    A, B = f32(A), f32(B)
    p, q, d = np.zeros(9, np.float32), np.zeros(9, np.float32), np.zeros(1, np.float32)
    lib().orc_mat3_ops_f(fptr(A), fptr(B), fptr(p), fptr(q), fptr(d))
    return p, q, d[0]


def sdf(c13, pos):
    return lib().orc_sdf_f(fptr(f32(c13)), fptr(f32(pos)))


def collider_inverse(c13):
    out = np.zeros(16, np.float32)
    lib().orc_collider_inverse_f(fptr(f32(c13)), fptr(out))
    return out


def collide(c13s, pos, vel):
    c = f32(np.asarray(c13s).reshape(-1, 13))
    out = np.zeros(3, np.float32)
    lib().orc_collide_f(fptr(c), c.shape[0], fptr(f32(pos)), fptr(f32(vel)), fptr(out))
    return out


def domain_walls(ijk, dims, vel):
    out = np.zeros(3, np.float32)
    lib().orc_domain_walls_f(int(ijk[0]), int(ijk[1]), int(ijk[2]), int(dims[0]), int(dims[1]), int(dims[2]),
                             fptr(f32(vel)), fptr(out))
    return out


# --------------------------------------------------------------- binning
# No reference counterpart (the reference never sorts particles): these pin
# the HIP binning to a CPU stable sort and to the documented slot layout
# inside each 4x4x4-cell block.
def bin_keys(x, h, dims):
    """key = block * 64 + cell-in-block, cell = int(x / h) (FP32 quotient),
    clamped to [1, dims - 3] like the stencil base."""
    dims = np.asarray(dims)
    nb = (dims + 3) // 4
    q = np.asarray(x, np.float32) / np.float32(h)
    c = np.clip(q.astype(np.int32), 1, dims - 3)
    blk = ((c[:, 0] >> 2) * nb[1] + (c[:, 1] >> 2)) * nb[2] + (c[:, 2] >> 2)
    cell = ((c[:, 0] & 3) << 4) | ((c[:, 1] & 3) << 2) | (c[:, 2] & 3)
    return (blk.astype(np.uint32) * np.uint32(64) + cell.astype(np.uint32)).astype(np.uint32)


def bin_reference(x, h, dims):
    """-> (sorted keys, stable permutation, caller index of each slot).  The
    particles take their slots in key order, so the slot's caller index is
    the permutation itself."""
    key = bin_keys(x, h, dims)
    perm = np.argsort(key, kind="stable")
    return key[perm], perm.astype(np.uint32), perm.astype(np.int64)


# ------------------------------------------------------------ generators
# Seeded scene generators (BALL, CELLS, LATTICE); LATTICE follows
# initializeParticles (realtime-deformations/material_point_method.cpp:18-63)
# with a Philox jitter in place of rand().
GEN_BALL, GEN_CELLS, GEN_LATTICE = 0, 1, 2


def philox4x64(ctr, key):
    """Philox4x64-10 block for counter ctr (4 words) and key (2 words)."""
    c = np.ascontiguousarray(ctr, np.uint64)
    k = np.ascontiguousarray(key, np.uint64)
    out = np.zeros(4, np.uint64)
    u = C.POINTER(C.c_uint64)
    lib().orc_philox4x64(c.ctypes.data_as(u), k.ctypes.data_as(u), out.ctypes.data_as(u))
    return out


def generate(kind, n, h, seed=0, centre=(0.5, 0.5, 0.5), radius=0.1, cell_lo=(0, 0, 0), cells=(1, 1, 1),
             ppc=(4, 12), E=(1.4e5, 1.4e5)):
    """-> (x (n,3) f32, E0 (n,) f32, candidates kept).  Raises ValueError when
    the request cannot be met (returned count then says how many fit)."""
    x = np.zeros((n, 3), np.float32)
    E0 = np.zeros(n, np.float32)
    ce = np.ascontiguousarray(centre, np.float32)
    lo = np.ascontiguousarray(cell_lo, np.int32)
    ex = np.ascontiguousarray(cells, np.int32)
    ip = C.POINTER(C.c_int)
    r = lib().orc_generate(int(kind), int(seed), fptr(ce), float(radius), lo.ctypes.data_as(ip),
                           ex.ctypes.data_as(ip), int(ppc[0]), int(ppc[1]), float(E[0]), float(E[1]), float(h), int(n),
                           fptr(x), fptr(E0))
    if r < 0:
        raise ValueError(f"generator request cannot be met (kept {-1 - r} candidates)" if kind == GEN_LATTICE
                         else "generator request cannot be met")
    return x, E0, int(r)


def lattice_candidates(h, seed, centre, radius):
    """Number of LATTICE candidates the rule keeps (an n = 0 request)."""
    return generate(GEN_LATTICE, 0, h, seed=seed, centre=centre, radius=radius)[2]
