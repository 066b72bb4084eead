// Fused, binned substep kernels of libmpm_hip (gfx950).
//
// The reference runs seven passes per substep (main.cpp:192-218).  Because
// the F update consumes the APIC matrix of the PREVIOUS G2P (it runs before
// updateParticleVelocities, cpp:306-330), the G2P of substep n-1 and
// everything particle-side of substep n fuse into one pass per particle:
//
//   G2P from grid n-1 (v, B; cpp:332-342) -> advection and clamp
//   (cpp:344-350,381-388) -> snow stress from F^n (cpp:235-254) -> F update
//   F^n -> F^{n+1} with B (cpp:306-330) -> scatter of mass and of the APIC
//   momentum plus the explicit force impulse,
//   m v + (m B D^-1 - dt V0 D^-1 P F^T)(x_i - x_p), into grid n (cpp:94-129)
//
// so v and B never leave registers inside a run of substeps.  One workgroup
// processes one 4x4x4-cell home block at a time, claiming blocks from the
// binning pass's list through a device counter: the grid velocities around
// the block and the FP64 accumulators it writes live in LDS halos of 9x9x9
// nodes (the block, the cubic stencil, and a one-cell margin for particles
// that drifted since the last sort).  The accumulators are laid out in 4x4
// (y, z) tiles (blk::tile_y, blk::tile_z), so an FP64 atomic's bank pair is
// set by the node's (y mod 4, z mod 4), and a per-lane rotation of the
// stencil traversal (lane_rotation) gives the 16 lanes of each atomic group
// 16 distinct pairs.  A particle outside the margin takes a global-memory
// path.  The grid pass then turns accumulators into velocities on the active
// blocks only.
//
// Every block-list entry carries its scene: one launch steps all scenes of a
// context, each with its own grid spacing, grid, material and colliders read
// from the device scene table.
#include "mpm_block.hpp"

namespace mpm {
namespace {

constexpr int TPB = 256;
using blk::HN;
using blk::SX;
using blk::SY;
using blk::PLANE;
using blk::TPLANE;
using blk::TX;
using blk::tile_y;
using blk::tile_z;
using blk::claim_block;

enum Mode { MODE_P2G = 0, MODE_FUSED = 1, MODE_G2P = 2, MODE_MASS = 3 };
static_assert(ST_NONFINITE == 1 && ST_CFL == 4, "stability_flags returns the status bits");

// Field f of slot p: a wave-uniform row base plus a 32-bit byte offset, which
// maps to one scalar-base global access per field (slots < 2^30).
__device__ __forceinline__ float ld(const KArgs& a, int f, int p) {
  const char* row = reinterpret_cast<const char*>(a.P + (int64_t)f * a.stride);
  return *reinterpret_cast<const float*>(row + (uint32_t)p * 4u);
}
__device__ __forceinline__ void st(const KArgs& a, int f, int p, float v) {
  char* row = reinterpret_cast<char*>(a.P + (int64_t)f * a.stride);
  *reinterpret_cast<float*>(row + (uint32_t)p * 4u) = v;
}

__device__ __forceinline__ void mark_block(const BlockArgs& ba, const Scene& S, int bi, int bj, int bk) {
  if (bi < 0 || bj < 0 || bk < 0 || bi >= S.nb[0] || bj >= S.nb[1] || bk >= S.nb[2]) return;
  const int gb = (int)(S.blk0 + (bi * S.nb[1] + bj) * S.nb[2] + bk);
  const int step = *ba.dstep;
  if (atomicExch(&ba.gflag[gb], step) != step) {
    const int slot = atomicAdd(ba.gcnt + step % 3, 1);
    ba.glist[step & 1][slot] = gb;
  }
}


// The scene record of a workgroup-uniform scene index, its words passed
// through readfirstlane so that the record lives in scalar registers.  Read
// as a vector value, the batched passes kept it in 20-odd vector registers
// and spilled 13 of their own; this way the batched fused pass runs 7%
// faster.
__device__ __forceinline__ Scene uniform_scene(const Scene* scenes, int scene) {
  static_assert(sizeof(Scene) % 4 == 0, "scene record of 32-bit words");
  int w[sizeof(Scene) / 4];
  __builtin_memcpy(w, scenes + scene, sizeof(Scene));
#pragma unroll
  for (int k = 0; k < (int)(sizeof(Scene) / 4); ++k) w[k] = __builtin_amdgcn_readfirstlane(w[k]);
  Scene S;
  __builtin_memcpy(&S, w, sizeof(Scene));
  return S;
}

// Separable G2P accumulation over the 4x4x4 stencil (cpp:332-342):
// v = sum_a wx T_a and the columns of B = sum w v d^T from partial sums
// S (over z) and T (over y).  `fetch(ia, ib, ic)` returns the node velocity.
template <typename Fetch>
__device__ __forceinline__ void g2p_accum(const Axis (&ax)[3], const float (&dd)[3][4], Fetch fetch, F3& v, M3& B) {
  float vx = 0.f, vy = 0.f, vz = 0.f;
  float bxx = 0.f, bxy = 0.f, bxz = 0.f, byx = 0.f, byy = 0.f, byz = 0.f, bzx = 0.f, bzy = 0.f, bzz = 0.f;
#pragma unroll
  for (int ia = 0; ia < 4; ++ia) {
    float tx = 0.f, ty = 0.f, tz = 0.f, tyx = 0.f, tyy = 0.f, tyz = 0.f, tzx = 0.f, tzy = 0.f, tzz = 0.f;
#pragma unroll
    for (int ib = 0; ib < 4; ++ib) {
      float sx = 0.f, sy = 0.f, sz = 0.f, szx = 0.f, szy = 0.f, szz = 0.f;
#pragma unroll
      for (int ic = 0; ic < 4; ++ic) {
        const F3 g = fetch(ia, ib, ic);
        const float wz = ax[2].w[ic], wzd = wz * dd[2][ic];
        sx = fmaf(wz, g.x, sx);
        sy = fmaf(wz, g.y, sy);
        sz = fmaf(wz, g.z, sz);
        szx = fmaf(wzd, g.x, szx);
        szy = fmaf(wzd, g.y, szy);
        szz = fmaf(wzd, g.z, szz);
      }
      const float wy = ax[1].w[ib], wyd = wy * dd[1][ib];
      tx = fmaf(wy, sx, tx);
      ty = fmaf(wy, sy, ty);
      tz = fmaf(wy, sz, tz);
      tyx = fmaf(wyd, sx, tyx);
      tyy = fmaf(wyd, sy, tyy);
      tyz = fmaf(wyd, sz, tyz);
      tzx = fmaf(wy, szx, tzx);
      tzy = fmaf(wy, szy, tzy);
      tzz = fmaf(wy, szz, tzz);
    }
    const float wx = ax[0].w[ia], wxd = wx * dd[0][ia];
    vx = fmaf(wx, tx, vx);
    vy = fmaf(wx, ty, vy);
    vz = fmaf(wx, tz, vz);
    bxx = fmaf(wxd, tx, bxx);
    bxy = fmaf(wxd, ty, bxy);
    bxz = fmaf(wxd, tz, bxz);
    byx = fmaf(wx, tyx, byx);
    byy = fmaf(wx, tyy, byy);
    byz = fmaf(wx, tyz, byz);
    bzx = fmaf(wx, tzx, bzx);
    bzy = fmaf(wx, tzy, bzy);
    bzz = fmaf(wx, tzz, bzz);
  }
  v = f3(vx, vy, vz);
  // B = sum w v d^T : column j holds sum w v d_j (glm B.c[col][row])
  B.c[0][0] = bxx;
  B.c[0][1] = bxy;
  B.c[0][2] = bxz;
  B.c[1][0] = byx;
  B.c[1][1] = byy;
  B.c[1][2] = byz;
  B.c[2][0] = bzx;
  B.c[2][1] = bzy;
  B.c[2][2] = bzz;
}

struct StencilInfo {
  Axis ax[3];
  int lb[3];   // stencil origin inside the halo
  bool fits;   // stencil inside the halo
};
__device__ __forceinline__ StencilInfo stencil_at(const Scene& S, const int o[3], F3 x) {
  StencilInfo s;
  s.ax[0] = axis_weights(x.x, S.h, S.inv_h, 1, S.dims[0] - 3);
  s.ax[1] = axis_weights(x.y, S.h, S.inv_h, 1, S.dims[1] - 3);
  s.ax[2] = axis_weights(x.z, S.h, S.inv_h, 1, S.dims[2] - 3);
  s.fits = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    s.lb[d] = s.ax[d].base - 1 - o[d];
    s.fits = s.fits && s.lb[d] >= 0 && s.lb[d] + 3 < HN;
  }
  return s;
}

// Per-lane traversal order of the 4x4x4 stencil in the LDS halo.  Lane l
// visits the nodes of axis d starting at offset r[d] = (T[d] - lb[d]) mod 4,
// T = (l >> 4, (l >> 2) & 3, l & 3), lb its stencil origin in the halo.  The
// FP64 atomics of a wave issue as four 16-lane groups (FP32 reads as two
// 32-lane groups); lanes of a group that share a cell, or a (cy, cz) column,
// would otherwise hit one address or one bank pair at every step, which is
// common in sparsely filled blocks (surfaces of a body).  Rotated, such lanes
// differ in (y, z) mod 4 at every step and land on distinct addresses in
// distinct bank pairs.  Only the order in which a particle sums its 64 nodes
// changes.
__device__ __forceinline__ void lane_rotation(const StencilInfo& s, int (&r)[3]) {
  const int l = threadIdx.x & 63;
  r[0] = ((l >> 4) - s.lb[0]) & 3;
  r[1] = (((l >> 2) & 3) - s.lb[1]) & 3;
  r[2] = ((l & 3) - s.lb[2]) & 3;
}
// w[k] <- w[(k + r) & 3]
__device__ __forceinline__ void rot4(float (&w)[4], int r) {
  const bool b1 = (r & 1) != 0, b2 = (r & 2) != 0;
  const float a0 = b1 ? w[1] : w[0], a1 = b1 ? w[2] : w[1], a2 = b1 ? w[3] : w[2], a3 = b1 ? w[0] : w[3];
  w[0] = b2 ? a2 : a0;
  w[1] = b2 ? a3 : a1;
  w[2] = b2 ? a0 : a2;
  w[3] = b2 ? a1 : a3;
}

// G2P of one particle from the LDS halo, or from global memory when its
// stencil left the halo, summing the 64 nodes in the plain (i, j, k) order
// of the reference whatever the lane.  Reads, unlike the scatter's atomics,
// gain nothing from the per-lane rotation: lanes of one cell read one
// address (a broadcast), and in the plain order the compiler merges the
// reads of neighbouring z nodes (20% fewer LDS instructions, 2% faster
// passes than rotated).  The order also no longer depends on the slot a
// binning gave the particle, which the reproducible mode needs.
__device__ __forceinline__ void g2p_one(const KArgs& a, const Scene& S, const float* __restrict__ vh, const int o[3],
                                        F3 x, F3& v, M3& B) {
  StencilInfo s = stencil_at(S, o, x);
  float dd[3][4];
  if (s.fits) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float xc = d == 0 ? x.x : (d == 1 ? x.y : x.z);
#pragma unroll
      for (int k = 0; k < 4; ++k) dd[d][k] = (float)(s.ax[d].base - 1 + k) * S.h - xc;
    }
    int qx[4], qy[4], qz[4];  // halo offsets per node and axis
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      qx[k] = (s.lb[0] + k) * SX;
      qy[k] = (s.lb[1] + k) * SY;
      qz[k] = s.lb[2] + k;
    }
    g2p_accum(s.ax, dd, [&](int ia, int ib, int ic) {
      const int q = qx[ia] + qy[ib] + qz[ic];
      return f3(vh[q], vh[PLANE + q], vh[2 * PLANE + q]);
    }, v, B);
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float xc = d == 0 ? x.x : (d == 1 ? x.y : x.z);
#pragma unroll
      for (int k = 0; k < 4; ++k) dd[d][k] = (float)(s.ax[d].base - 1 + k) * S.h - xc;
    }
    g2p_accum(s.ax, dd, [&](int ia, int ib, int ic) {
      const float4 g = a.gvel[node_index(s.ax[0].base - 1 + ia, s.ax[1].base - 1 + ib, s.ax[2].base - 1 + ic, S)];
      return f3(g.x, g.y, g.z);
    }, v, B);
  }
}

// P2G scatter of mass and of m v + Q (x_i - x_p) over the 4x4x4 stencil
// (cpp:99-117 and cpp:249-252 combined); `add(ia, ib, ic, ja, jb, jc, m, px,
// py, pz)` with the traversal steps (ia, ib, ic) and the node offsets
// (ja, jb, jc) from the stencil base.
// `s.ax[d].w` hold the weights in traversal order: step k of axis d is node
// base - 1 + ((k + r[d]) & 3) (r = 0: the plain order).
template <typename Add>
__device__ __forceinline__ void p2g_scatter(const Scene& S, const StencilInfo& s, const int (&r)[3], F3 x, float m,
                                            F3 mv, const M3& Q, Add add) {
#pragma unroll
  for (int ia = 0; ia < 4; ++ia) {
    const int ja = (ia + r[0]) & 3;
    const float ddx = (float)(s.ax[0].base - 1 + ja) * S.h - x.x;
    const float wa = s.ax[0].w[ia];
    const float qax = fmaf(Q.c[0][0], ddx, mv.x), qay = fmaf(Q.c[0][1], ddx, mv.y), qaz = fmaf(Q.c[0][2], ddx, mv.z);
#pragma unroll
    for (int ib = 0; ib < 4; ++ib) {
      const int jb = (ib + r[1]) & 3;
      const float ddy = (float)(s.ax[1].base - 1 + jb) * S.h - x.y;
      const float wxy = wa * s.ax[1].w[ib];
      const float qx = fmaf(Q.c[1][0], ddy, qax);
      const float qy = fmaf(Q.c[1][1], ddy, qay);
      const float qz = fmaf(Q.c[1][2], ddy, qaz);
#pragma unroll
      for (int ic = 0; ic < 4; ++ic) {
        const int jc = (ic + r[2]) & 3;
        const float ddz = (float)(s.ax[2].base - 1 + jc) * S.h - x.z;
        const float w = wxy * s.ax[2].w[ic];
        add(ia, ib, ic, ja, jb, jc, m * w, w * fmaf(Q.c[2][0], ddz, qx), w * fmaf(Q.c[2][1], ddz, qy),
            w * fmaf(Q.c[2][2], ddz, qz));
      }
    }
  }
}

template <int EXACT>
__device__ __forceinline__ void scatter_global(const BlockArgs& ba, const Scene& S, const StencilInfo& s, F3 x, float m,
                                            F3 mv, const M3& Q);

// Scatter into the LDS halo, or with global atomics (and the touched grid
// blocks marked) when the stencil left the halo.  EXACT: each contribution
// rounded to the scene's quanta first (reproducible mode).
template <int EXACT>
__device__ __forceinline__ void scatter_one(const BlockArgs& ba, const Scene& S, double* __restrict__ ah,
                                            const int o[3], F3 x, float m, F3 mv, const M3& Q) {
  StencilInfo s = stencil_at(S, o, x);
  if (s.fits) {
    int r[3];
    lane_rotation(s, r);
#pragma unroll
    for (int d = 0; d < 3; ++d) rot4(s.ax[d].w, r[d]);
    // halo offsets of the rotated traversal, per axis, so a node's address
    // is one add
    int qx[4], qy[4], qz[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      qx[k] = (s.lb[0] + ((k + r[0]) & 3)) * TX;
      qy[k] = tile_y(s.lb[1] + ((k + r[1]) & 3));
      qz[k] = tile_z(s.lb[2] + ((k + r[2]) & 3));
    }
    if (EXACT) {
      const double Mm = quantum_magic(S.qexp), Mp = quantum_magic(S.qexp + 2);
      p2g_scatter(S, s, r, x, m, mv, Q, [&](int ia, int ib, int ic, int, int, int, float vm, float px, float py, float pz) {
        const int q = qx[ia] + qy[ib] + qz[ic];
        atomicAdd(&ah[q], on_quantum((double)vm, Mm));
        atomicAdd(&ah[TPLANE + q], on_quantum((double)px, Mp));
        atomicAdd(&ah[2 * TPLANE + q], on_quantum((double)py, Mp));
        atomicAdd(&ah[3 * TPLANE + q], on_quantum((double)pz, Mp));
      });
    } else {
      p2g_scatter(S, s, r, x, m, mv, Q, [&](int ia, int ib, int ic, int, int, int, float vm, float px, float py, float pz) {
        const int q = qx[ia] + qy[ib] + qz[ic];
        atomicAdd(&ah[q], (double)vm);
        atomicAdd(&ah[TPLANE + q], (double)px);
        atomicAdd(&ah[2 * TPLANE + q], (double)py);
        atomicAdd(&ah[3 * TPLANE + q], (double)pz);
      });
    }
  } else {
    scatter_global<EXACT>(ba, S, s, x, m, mv, Q);
  }
}

// The global-memory scatter of a particle whose stencil left its block's
// halo: atomics into the grid accumulators (FP32, or FP64 on the quanta in
// reproducible mode), touched grid blocks marked.
template <int EXACT>
__device__ __forceinline__ void scatter_global(const BlockArgs& ba, const Scene& S, const StencilInfo& s, F3 x, float m,
                                            F3 mv, const M3& Q) {
  const KArgs& a = ba.a;
  {
    // stencil blocks: base-1 .. base+2 spans at most two blocks per axis
    for (int di = 0; di < 2; ++di)
      for (int dj = 0; dj < 2; ++dj)
        for (int dk = 0; dk < 2; ++dk)
          mark_block(ba, S, (s.ax[0].base - 1 + 3 * di) >> 2, (s.ax[1].base - 1 + 3 * dj) >> 2,
                     (s.ax[2].base - 1 + 3 * dk) >> 2);
    const int r0[3] = {0, 0, 0};
    p2g_scatter(S, s, r0, x, m, mv, Q, [&](int, int, int, int ia, int ib, int ic, float vm, float px, float py, float pz) {
      const int64_t gi = node_index(s.ax[0].base - 1 + ia, s.ax[1].base - 1 + ib, s.ax[2].base - 1 + ic, S);
      if (EXACT) {
        double* g = reinterpret_cast<double*>(a.gacc64 + gi);
        const double Mm = quantum_magic(S.qexp), Mp = quantum_magic(S.qexp + 2);
        atomicAdd(g + 0, on_quantum((double)vm, Mm));
        atomicAdd(g + 1, on_quantum((double)px, Mp));
        atomicAdd(g + 2, on_quantum((double)py, Mp));
        atomicAdd(g + 3, on_quantum((double)pz, Mp));
      } else {
        float* g = reinterpret_cast<float*>(a.gacc + gi);
        atomicAdd(g + 0, vm);
        atomicAdd(g + 1, px);
        atomicAdd(g + 2, py);
        atomicAdd(g + 3, pz);
      }
    });
  }
}

// Stress from F^n and the F update F^n -> F^{n+1} with B of one particle;
// returns its mass and the affine matrix of its P2G momentum,
// Q = m B D^-1 - dt V0 D^-1 P F^T.
template <int COMPOSE, int REC>
__device__ __forceinline__ void particle_update(const BlockArgs& ba, const Scene& S, int scene, int p, F3 x, F3 v,
                                                const M3& B, float& m, M3& Q) {
  const KArgs& a = ba.a;
  {
    M3 FE, FP;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        FE.c[c][r] = ld(a, FFE + 3 * c + r, p);
        FP.c[c][r] = ld(a, FFP + 3 * c + r, p);
      }
    if (REC) {
      float* R = ba.rec;
      const int64_t sd = a.stride;
      R[(int64_t)(FX + 0) * sd + p] = x.x;
      R[(int64_t)(FX + 1) * sd + p] = x.y;
      R[(int64_t)(FX + 2) * sd + p] = x.z;
      R[(int64_t)(FV + 0) * sd + p] = v.x;
      R[(int64_t)(FV + 1) * sd + p] = v.y;
      R[(int64_t)(FV + 2) * sd + p] = v.z;
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          R[(int64_t)(FB + 3 * c + r) * sd + p] = B.c[c][r];
          R[(int64_t)(FFE + 3 * c + r) * sd + p] = FE.c[c][r];
          R[(int64_t)(FFP + 3 * c + r) * sd + p] = FP.c[c][r];
        }
    }
    m = ld(a, FM, p);
    const float V0 = ld(a, FVOL, p);
    const float E = ld(a, FE0, p);
    const M3 Dinv = dinv_matrix(S.dinv);
    float stiff;
    const M3 A = stress_matrix(E, S.lame, V0, Dinv, FE, FP, stiff);
    if (const int fl = stability_flags(v, stiff, V0, m, a.dt, S.h)) atomicOr(a.status + scene, fl);
    const M3 BD = mul_dinv(B, Dinv);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) Q.c[c][r] = fmaf(m, BD.c[c][r], -a.dt * A.c[c][r]);
    float svd[SVD_ROWS];
    if (f_update_t<COMPOSE>(a.dt, Dinv, S.lo, S.hi, S.gap64, B, FE, FP, REC ? svd : nullptr)) {
      if (REC) {
#pragma unroll
        for (int r = 0; r < SVD_ROWS; ++r) ba.rec_svd[(int64_t)r * a.stride + p] = svd[r];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          st(a, FFE + 3 * c + r, p, FE.c[c][r]);
          st(a, FFP + 3 * c + r, p, FP.c[c][r]);
        }
    } else {
      atomicOr(a.status + scene, ST_NONFINITE);
    }
  }
}

// Stress, F update, then the per-particle P2G scatter.
template <int COMPOSE, int REC, int EXACT>
__device__ __forceinline__ void p2g_one(const BlockArgs& ba, const Scene& S, int scene, double* __restrict__ ah,
                                        const int o[3], int p, F3 x, F3 v, const M3& B) {
  float m;
  M3 Q;
  particle_update<COMPOSE, REC>(ba, S, scene, p, x, v, B, m, Q);
  scatter_one<EXACT>(ba, S, ah, o, x, m, v * m, Q);
}

__device__ __forceinline__ F3 advect(const KArgs& a, const Scene& S, F3 x, F3 v) {
  x = x + v * a.dt;
  const float lo = 3.0f * S.h;
  x.x = fminf(fmaxf(x.x, lo), (float)(S.dims[0] - 3) * S.h);
  x.y = fminf(fmaxf(x.y, lo), (float)(S.dims[1] - 3) * S.h);
  x.z = fminf(fmaxf(x.z, lo), (float)(S.dims[2] - 3) * S.h);
  return x;
}

// BATCHED = 0 (one scene): the scene comes from the kernel arguments,
// read-only and loop-invariant; BATCHED = 1: each block's scene from the
// device table, copied once per block into scalar registers.
// Every variant is held to three waves per SIMD (<= 168 registers), the
// occupancy the LDS halo allows.  In the TEXTBOOK variants most spilled
// registers are saved around the out-of-line FP64 re-solve, which runs only
// for near-degenerate spectra: at three waves their passes run 8-9% faster
// than at two (the batched fused pass took 258 registers unbounded, one wave
// per SIMD, and ran 1.5x slower).
// EXACT: reproducible mode, the P2G contributions on the scene's quanta.
template <int MODE, int COMPOSE, int BATCHED, int REC = 0, int EXACT = 0>
__global__ void __launch_bounds__(TPB, 3) k_block_pass(BlockArgs ba) {
  constexpr bool kG2P = MODE == MODE_FUSED || MODE == MODE_G2P;
  constexpr bool kP2G = MODE != MODE_G2P;
  __shared__ float vh[kG2P ? 3 * PLANE : 1];
  __shared__ double ah[kP2G ? 4 * TPLANE : 1];
  __shared__ int claimed;
  const KArgs& a = ba.a;
  const int t = threadIdx.x;
  const int nblocks = *ba.nblk;
  if ((int)blockIdx.x >= nblocks) {
    if (ba.advance && nblocks == 0 && blockIdx.x == 0 && t == 0) {  // nothing to claim
      const int step = *ba.dstep;
      *ba.dstep = step + 1;
      ba.gcnt[(step + 1) % 3] = 0;
    }
    return;
  }
  for (int s; (s = claim_block(ba.bctr, nblocks, &claimed)) >= 0;) {
    const int4 e = ba.blist[s];
    const int scene = e.w, start = e.y, end = e.z;
    const Scene S = BATCHED ? uniform_scene(a.scenes, scene) : a.sc;
    const int blk = (int)(e.x - S.blk0);
    const int bk = blk % S.nb[2];
    const int tt = blk / S.nb[2];
    const int bj = tt % S.nb[1];
    const int bi = tt / S.nb[1];
    const int o[3] = {4 * bi - 2, 4 * bj - 2, 4 * bk - 2};
    if (kG2P) {
      for (int q = t; q < HN * HN * HN; q += TPB) {
        const int hx = q / (HN * HN), hy = (q / HN) % HN, hz = q % HN;
        const int gi = o[0] + hx, gj = o[1] + hy, gk = o[2] + hz;
        float4 g = make_float4(0.f, 0.f, 0.f, 0.f);
        if (gi >= 0 && gj >= 0 && gk >= 0 && gi < S.dims[0] && gj < S.dims[1] && gk < S.dims[2])
          g = a.gvel[node_index(gi, gj, gk, S)];
        const int l = hx * SX + hy * SY + hz;
        vh[l] = g.x;
        vh[PLANE + l] = g.y;
        vh[2 * PLANE + l] = g.z;
      }
    }
    if (kP2G) {
      for (int q = t; q < 4 * TPLANE; q += TPB) ah[q] = 0.0;
    }
    __syncthreads();
    for (int p = start + t; p < end; p += TPB) {
      F3 x = f3(ld(a, FX, p), ld(a, FX + 1, p), ld(a, FX + 2, p));
      if (MODE == MODE_MASS) {
        M3 Z = mdiag(0.0f);
        scatter_one<EXACT>(ba, S, ah, o, x, ld(a, FM, p), f3(0.f, 0.f, 0.f), Z);
        continue;
      }
      F3 v;
      M3 B;
      if (kG2P) {
        g2p_one(a, S, vh, o, x, v, B);
        x = advect(a, S, x, v);
        st(a, FX, p, x.x);
        st(a, FX + 1, p, x.y);
        st(a, FX + 2, p, x.z);
        if (MODE == MODE_G2P) {
          if (REC) {
            ba.rec[(int64_t)(FX + 0) * a.stride + p] = x.x;
            ba.rec[(int64_t)(FX + 1) * a.stride + p] = x.y;
            ba.rec[(int64_t)(FX + 2) * a.stride + p] = x.z;
          }
          st(a, FV, p, v.x);
          st(a, FV + 1, p, v.y);
          st(a, FV + 2, p, v.z);
#pragma unroll
          for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int r = 0; r < 3; ++r) st(a, FB + 3 * c + r, p, B.c[c][r]);
        }
      } else {
        v = f3(ld(a, FV, p), ld(a, FV + 1, p), ld(a, FV + 2, p));
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int r = 0; r < 3; ++r) B.c[c][r] = ld(a, FB + 3 * c + r, p);
      }
      if (kP2G) p2g_one<COMPOSE, REC, EXACT>(ba, S, scene, ah, o, p, x, v, B);
    }
    __syncthreads();
    if (kP2G) {
      for (int q = t; q < HN * HN * HN; q += TPB) {
        const int hx = q / (HN * HN), hy = (q / HN) % HN, hz = q % HN;
        const int l = hx * TX + tile_y(hy) + tile_z(hz);
        if (EXACT) {  // exact FP64 sums of multiples of the quanta
          const double m = ah[l], px = ah[TPLANE + l], py = ah[2 * TPLANE + l], pz = ah[3 * TPLANE + l];
          if (m != 0.0 || px != 0.0 || py != 0.0 || pz != 0.0) {
            const int gi = o[0] + hx, gj = o[1] + hy, gk = o[2] + hz;
            if (gi >= 0 && gj >= 0 && gk >= 0 && gi < S.dims[0] && gj < S.dims[1] && gk < S.dims[2]) {
              double* g = reinterpret_cast<double*>(a.gacc64 + node_index(gi, gj, gk, S));
              atomicAdd(g + 0, m);
              atomicAdd(g + 1, px);
              atomicAdd(g + 2, py);
              atomicAdd(g + 3, pz);
            }
          }
        } else {
          const float m = (float)ah[l], px = (float)ah[TPLANE + l], py = (float)ah[2 * TPLANE + l],
                      pz = (float)ah[3 * TPLANE + l];
          if (m != 0.0f || px != 0.0f || py != 0.0f || pz != 0.0f) {
            const int gi = o[0] + hx, gj = o[1] + hy, gk = o[2] + hz;
            if (gi >= 0 && gj >= 0 && gk >= 0 && gi < S.dims[0] && gj < S.dims[1] && gk < S.dims[2]) {
              float* g = reinterpret_cast<float*>(a.gacc + node_index(gi, gj, gk, S));
              atomicAdd(g + 0, m);
              atomicAdd(g + 1, px);
              atomicAdd(g + 2, py);
              atomicAdd(g + 3, pz);
            }
          }
        }
      }
      if (t < 27) mark_block(ba, S, bi - 1 + t / 9, bj - 1 + (t / 3) % 3, bk - 1 + t % 3);
    }
    __syncthreads();
  }
  // The workgroup whose claim is the last of the launch (each other one has
  // made its final claim after finishing its blocks) advances the epoch for
  // the grid pass and the next substep, and clears the block counter the
  // next substep fills (last used two substeps ago).
  if (ba.advance && t == 0 && claimed == nblocks + ((int)gridDim.x < nblocks ? (int)gridDim.x : nblocks) - 1) {
    const int step = *ba.dstep;
    *ba.dstep = step + 1;
    ba.gcnt[(step + 1) % 3] = 0;
  }
}

// Reverse of advection + G2P for the blocks of the current binning (see
// launch_adj_g2p_block).  Node velocities come from vnode, which the map
// kernel filled from the substep's saved P2G sums (zero outside its listed
// blocks); the advection clamp was active on an axis exactly when the next
// stored position Xn sits on the clamp bound.  The stage order and notation
// are those of mpm_adjoint.hip.
template <int BATCHED>
__global__ void __launch_bounds__(TPB) k_adj_g2p_block(BlockArgs ba, const float* __restrict__ St,
                                                       const float* __restrict__ Xn, const float* __restrict__ Gin,
                                                       float* __restrict__ Gout, float4* __restrict__ gnode,
                                                       const float4* __restrict__ vnode) {
  __shared__ float vh[3 * PLANE];
  __shared__ double gh[3 * TPLANE];  // tiled like the forward accumulators
  __shared__ int claimed;
  const KArgs& a = ba.a;
  const int t = threadIdx.x;
  const int64_t sd = a.stride;
  const int nblocks = *ba.nblk;
  if ((int)blockIdx.x >= nblocks) return;
  for (int s; (s = claim_block(ba.bctr, nblocks, &claimed)) >= 0;) {
    const int4 e = ba.blist[s];
    const int scene = e.w, start = e.y, end = e.z;
    const Scene S = BATCHED ? uniform_scene(a.scenes, scene) : a.sc;
    const int blk = (int)(e.x - S.blk0);
    const int bk = blk % S.nb[2];
    const int tt = blk / S.nb[2];
    const int bj = tt % S.nb[1];
    const int bi = tt / S.nb[1];
    const int o[3] = {4 * bi - 2, 4 * bj - 2, 4 * bk - 2};
    for (int q = t; q < HN * HN * HN; q += TPB) {
      const int hx = q / (HN * HN), hy = (q / HN) % HN, hz = q % HN;
      const int gi = o[0] + hx, gj = o[1] + hy, gk = o[2] + hz;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (gi >= 0 && gj >= 0 && gk >= 0 && gi < S.dims[0] && gj < S.dims[1] && gk < S.dims[2])
        v = vnode[node_index(gi, gj, gk, S)];
      const int l = hx * SX + hy * SY + hz;
      vh[l] = v.x;
      vh[PLANE + l] = v.y;
      vh[2 * PLANE + l] = v.z;
    }
    for (int q = t; q < 3 * TPLANE; q += TPB) gh[q] = 0.0;
    __syncthreads();
    for (int p = start + t; p < end; p += TPB) {
      const F3 x = f3(St[(int64_t)FX * sd + p], St[(int64_t)(FX + 1) * sd + p], St[(int64_t)(FX + 2) * sd + p]);
      AxisD ax[3];
      ax[0] = axis_weights_d(x.x, S.h, S.inv_h, 1, S.dims[0] - 3);
      ax[1] = axis_weights_d(x.y, S.h, S.inv_h, 1, S.dims[1] - 3);
      ax[2] = axis_weights_d(x.z, S.h, S.inv_h, 1, S.dims[2] - 3);
      int lb[3];
      bool fits = true;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        lb[d] = ax[d].base - 1 - o[d];
        fits = fits && lb[d] >= 0 && lb[d] + 3 < HN;
      }
      const int q0 = lb[0] * SX + lb[1] * SY + lb[2];
      float gx[3], gvv[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        // x' = clamp(x + dt v', 3h, (dims - 3) h): the bound when clamped
        const float xn = Xn[(int64_t)(FX + k) * sd + p];
        const float lo = 3.0f * S.h, hi = (float)(S.dims[k] - 3) * S.h;
        const float mask = (xn > lo && xn < hi) ? 1.0f : 0.0f;
        const float gxn = Gin[(int64_t)(FX + k) * sd + p];
        gx[k] = gxn * mask;
        gvv[k] = Gin[(int64_t)(FV + k) * sd + p] + a.dt * gxn * mask;
      }
      M3 gB;
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) gB.c[c][r] = Gin[(int64_t)(FB + 3 * c + r) * sd + p];
      // one node: its velocity vi, weight w and weight gradient dw; returns
      // the node adjoint w (gv' + gB' d) and adds the position terms to gx
      // (ia, ib, ic) index the weight arrays, (ja, jb, jc) the node offsets
      auto node = [&](int ia, int ib, int ic, int ja, int jb, int jc, F3 vi) {
        const int i = ax[0].base - 1 + ja, j = ax[1].base - 1 + jb, k = ax[2].base - 1 + jc;
        const float w = ax[0].w[ia] * ax[1].w[ib] * ax[2].w[ic];
        const F3 dw = f3(ax[0].dw[ia] * ax[1].w[ib] * ax[2].w[ic], ax[0].w[ia] * ax[1].dw[ib] * ax[2].w[ic],
                         ax[0].w[ia] * ax[1].w[ib] * ax[2].dw[ic]);
        const F3 d = f3((float)i * S.h - x.x, (float)j * S.h - x.y, (float)k * S.h - x.z);
        const F3 u = f3(gvv[0] + gB.c[0][0] * d.x + gB.c[1][0] * d.y + gB.c[2][0] * d.z,
                        gvv[1] + gB.c[0][1] * d.x + gB.c[1][1] * d.y + gB.c[2][1] * d.z,
                        gvv[2] + gB.c[0][2] * d.x + gB.c[1][2] * d.y + gB.c[2][2] * d.z);
        const float su = dot(vi, u);
        gx[0] += dw.x * su - w * (gB.c[0][0] * vi.x + gB.c[0][1] * vi.y + gB.c[0][2] * vi.z);
        gx[1] += dw.y * su - w * (gB.c[1][0] * vi.x + gB.c[1][1] * vi.y + gB.c[1][2] * vi.z);
        gx[2] += dw.z * su - w * (gB.c[2][0] * vi.x + gB.c[2][1] * vi.y + gB.c[2][2] * vi.z);
        return u * w;
      };
      if (fits) {
        // the block pass's per-lane traversal order (lane_rotation): step k
        // of axis e visits node (k + r_e) & 3, whose weights rot4 moves to
        // index k.  The position adjoint is separable but for
        // s_i = v_i . u_i: gx = sum dw_i s_i - gB^T sum w_i v_i, summed over
        // z, then y, then x (about 21 vector instructions per node instead
        // of 50, the three scatter atomics included).
        const int l_ = t & 63;
        const int rr[3] = {((l_ >> 4) - lb[0]) & 3, (((l_ >> 2) & 3) - lb[1]) & 3, ((l_ & 3) - lb[2]) & 3};
        float ddr[3][4];
        {
          const float xc[3] = {x.x, x.y, x.z};
#pragma unroll
          for (int e = 0; e < 3; ++e) {
            rot4(ax[e].w, rr[e]);
            rot4(ax[e].dw, rr[e]);
#pragma unroll
            for (int k = 0; k < 4; ++k) ddr[e][k] = (float)(ax[e].base - 1 + ((k + rr[e]) & 3)) * S.h - xc[e];
          }
        }
        // columns of gB: (gB d)_r = sum_c gB.c[c][r] d_c
        const F3 b0 = f3(gB.c[0][0], gB.c[0][1], gB.c[0][2]), b1 = f3(gB.c[1][0], gB.c[1][1], gB.c[1][2]),
                 b2 = f3(gB.c[2][0], gB.c[2][1], gB.c[2][2]);
        const F3 gv3 = f3(gvv[0], gvv[1], gvv[2]);
        // offsets of the tiled accumulator halo per traversal step and axis
        int tqx[4], tqy[4], tqz[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          tqx[k] = (lb[0] + ((k + rr[0]) & 3)) * TX;
          tqy[k] = tile_y(lb[1] + ((k + rr[1]) & 3));
          tqz[k] = tile_z(lb[2] + ((k + rr[2]) & 3));
        }
        float SXs = 0.f, SYs = 0.f, SZs = 0.f;
        F3 V = f3(0.f, 0.f, 0.f);  // sum w_i v_i
#pragma unroll
        for (int ia = 0; ia < 4; ++ia) {
          __asm__ __volatile__("" ::: "memory");  // at most one x slab of nodes in flight
          const int ja = (ia + rr[0]) & 3;
          const F3 ua = gv3 + b0 * ddr[0][ia];
          const float wa = ax[0].w[ia];
          float Q0 = 0.f, Q1 = 0.f, Q2 = 0.f;
          F3 Tv = f3(0.f, 0.f, 0.f);
#pragma unroll
          for (int ib = 0; ib < 4; ++ib) {
            const int jb = (ib + rr[1]) & 3;
            const F3 uab = ua + b1 * ddr[1][ib];
            const float wab = wa * ax[1].w[ib];
            float P0 = 0.f, P1 = 0.f;
            F3 Rv = f3(0.f, 0.f, 0.f);
#pragma unroll
            for (int ic = 0; ic < 4; ++ic) {
              const int jc = (ic + rr[2]) & 3;
              const int q = q0 + ja * SX + jb * SY + jc;
              const F3 vi = f3(vh[q], vh[PLANE + q], vh[2 * PLANE + q]);
              const F3 u = uab + b2 * ddr[2][ic];
              const float su = dot(vi, u);
              const float wz = ax[2].w[ic];
              P0 = fmaf(wz, su, P0);
              P1 = fmaf(ax[2].dw[ic], su, P1);
              Rv = Rv + vi * wz;
              const float w = wab * wz;
              const int qg = tqx[ia] + tqy[ib] + tqz[ic];
              atomicAdd(&gh[qg], (double)(u.x * w));
              atomicAdd(&gh[TPLANE + qg], (double)(u.y * w));
              atomicAdd(&gh[2 * TPLANE + qg], (double)(u.z * w));
            }
            const float wy = ax[1].w[ib];
            Q0 = fmaf(wy, P0, Q0);
            Q1 = fmaf(ax[1].dw[ib], P0, Q1);
            Q2 = fmaf(wy, P1, Q2);
            Tv = Tv + Rv * wy;
          }
          SXs = fmaf(ax[0].dw[ia], Q0, SXs);
          SYs = fmaf(wa, Q1, SYs);
          SZs = fmaf(wa, Q2, SZs);
          V = V + Tv * wa;
        }
        gx[0] += SXs - dot(b0, V);
        gx[1] += SYs - dot(b1, V);
        gx[2] += SZs - dot(b2, V);
      } else {
        for (int ia = 0; ia < 4; ++ia)
          for (int ib = 0; ib < 4; ++ib)
            for (int ic = 0; ic < 4; ++ic) {
              const int i = ax[0].base - 1 + ia, j = ax[1].base - 1 + ib, k = ax[2].base - 1 + ic;
              const int64_t g = node_index(i, j, k, S);
              const int c = ba.grec.map[g >> 6];
              const float4 vn = vnode[g];
              const F3 wu = node(ia, ib, ic, ia, ib, ic, f3(vn.x, vn.y, vn.z));
              // massless nodes (outside the listed blocks) carry no adjoint
              if (c >= 0) {
                float* gn = reinterpret_cast<float*>(gnode + g);
                atomicAdd(gn + 1, wu.x);
                atomicAdd(gn + 2, wu.y);
                atomicAdd(gn + 3, wu.z);
              }
            }
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) Gout[(int64_t)(FX + k) * sd + p] = gx[k];
    }
    __syncthreads();
    for (int q = t; q < HN * HN * HN; q += TPB) {
      const int hx = q / (HN * HN), hy = (q / HN) % HN, hz = q % HN;
      const int l = hx * TX + tile_y(hy) + tile_z(hz);
      const float u0 = (float)gh[l], u1 = (float)gh[TPLANE + l], u2 = (float)gh[2 * TPLANE + l];
      if (u0 != 0.0f || u1 != 0.0f || u2 != 0.0f) {
        const int gi = o[0] + hx, gj = o[1] + hy, gk = o[2] + hz;
        const int64_t g = node_index(gi, gj, gk, S);
        if (gi >= 0 && gj >= 0 && gk >= 0 && gi < S.dims[0] && gj < S.dims[1] && gk < S.dims[2] &&
            ba.grec.map[g >> 6] >= 0) {
          float* gn = reinterpret_cast<float*>(gnode + g);
          atomicAdd(gn + 1, u0);
          atomicAdd(gn + 2, u1);
          atomicAdd(gn + 3, u2);
        }
      }
    }
    __syncthreads();
  }
}

// v = p / m + dt g on the nodes of active blocks, boundary conditions, and
// the accumulator reset; blocks active last substep but not now are zeroed.
template <int BATCHED> __global__ void __launch_bounds__(TPB) k_grid_active(BlockArgs ba) {
  const KArgs& a = ba.a;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int step = *ba.dstep - 1;  // the epoch of the block pass just run
  const int* glist_cur = ba.glist[step & 1];
  const int* glist_prev = ba.glist[(step + 1) & 1];
  const int ncur = ba.gcnt[step % 3];
  const int nprev = ba.gcnt[(step + 2) % 3];
  if (ba.grec.acc && blockIdx.x == 0 && threadIdx.x == 0) ba.grec.list[0] = ncur;
  const int wid = blockIdx.x * 4 + wave, nwaves = gridDim.x * 4;
  for (int s = wid; s < ncur; s += nwaves) {
    {
      const int gb = glist_cur[s];
      const int64_t g = (int64_t)gb * 64 + lane;
      // the sums as FP32, as the grid update uses them; the accumulators
      // are cleared below in their own format (an FP64 sum can round to
      // an FP32 zero)
      bool touched;
      float4 acc;
      if (a.exact) {
        const double4 d = a.gacc64[g];
        acc = make_float4((float)d.x, (float)d.y, (float)d.z, (float)d.w);
        touched = d.x != 0.0 || d.y != 0.0 || d.z != 0.0 || d.w != 0.0;
        if (touched) a.gacc64[g] = make_double4(0.0, 0.0, 0.0, 0.0);
      } else {
        acc = a.gacc[g];
        touched = acc.x != 0.0f || acc.y != 0.0f || acc.z != 0.0f || acc.w != 0.0f;
        if (touched) a.gacc[g] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
      if (ba.grec.acc) {
        ba.grec.acc[(int64_t)s * 64 + lane] = acc;
        if (lane == 0) ba.grec.list[1 + s] = gb;
      }
      float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
      if (acc.x != 0.0f) {
        const Scene S = BATCHED ? a.scenes[a.blk_scene[gb]] : a.sc;
        F3 v = f3(acc.y / acc.x + a.dt * S.g[0], acc.z / acc.x + a.dt * S.g[1], acc.w / acc.x + a.dt * S.g[2]);
        if (S.bc_mode) {
          const int lb = (int)(gb - S.blk0);
          const int bk = lb % S.nb[2];
          const int t2 = lb / S.nb[2];
          const int bj = t2 % S.nb[1];
          const int bi = t2 / S.nb[1];
          const int i = bi * 4 + (lane >> 4), j = bj * 4 + ((lane >> 2) & 3), k = bk * 4 + (lane & 3);
          if ((S.bc_mode & MPM_BC_COLLIDERS) && S.ncoll > 0)
            v = body_collision(S.coll, S.ncoll, f3((float)i * S.h, (float)j * S.h, (float)k * S.h), v);
          if (S.bc_mode & MPM_BC_DOMAIN) v = domain_walls(i, j, k, S.dims[0], S.dims[1], S.dims[2], v);
        }
        out = make_float4(v.x, v.y, v.z, acc.x);
      }
      a.gvel[g] = out;
    }
  }
  // Blocks active in the previous substep but not in this one get zero
  // velocity.  One lane checks one block: most are still active, so a wave
  // per block would mostly idle.
  for (int s = wid * 64 + lane; s < nprev; s += nwaves * 64) {
    const int gb = glist_prev[s];
    if (ba.gflag[gb] != step) {
      float4* gv = a.gvel + (int64_t)gb * 64;
      for (int k = 0; k < 64; ++k) gv[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
}

}  // namespace


template <int BATCHED> void launch_block_pass_t(const BlockArgs& ba, int mode, hipStream_t s) {
  const int grid = ba.grid;
  const bool lit = ba.a.compose == COMPOSE_LITERAL;
  if (ba.rec) {  // reverse-mode recording: TEXTBOOK only (checked by the caller)
    switch (mode) {
      case MODE_P2G: k_block_pass<MODE_P2G, COMPOSE_TEXTBOOK, BATCHED, 1><<<grid, TPB, 0, s>>>(ba); break;
      case MODE_FUSED: k_block_pass<MODE_FUSED, COMPOSE_TEXTBOOK, BATCHED, 1><<<grid, TPB, 0, s>>>(ba); break;
      default: k_block_pass<MODE_G2P, COMPOSE_TEXTBOOK, BATCHED, 1><<<grid, TPB, 0, s>>>(ba); break;
    }
    return;
  }
  if (ba.a.exact) {  // reproducible mode: P2G on the quanta
    switch (mode) {
      case MODE_P2G:
        if (lit)
          k_block_pass<MODE_P2G, COMPOSE_LITERAL, BATCHED, 0, 1><<<grid, TPB, 0, s>>>(ba);
        else
          k_block_pass<MODE_P2G, COMPOSE_TEXTBOOK, BATCHED, 0, 1><<<grid, TPB, 0, s>>>(ba);
        return;
      case MODE_FUSED:
        if (lit)
          k_block_pass<MODE_FUSED, COMPOSE_LITERAL, BATCHED, 0, 1><<<grid, TPB, 0, s>>>(ba);
        else
          k_block_pass<MODE_FUSED, COMPOSE_TEXTBOOK, BATCHED, 0, 1><<<grid, TPB, 0, s>>>(ba);
        return;
      case MODE_MASS: k_block_pass<MODE_MASS, COMPOSE_LITERAL, BATCHED, 0, 1><<<grid, TPB, 0, s>>>(ba); return;
      default: break;  // G2P scatters nothing
    }
  }
  switch (mode) {
    case MODE_P2G:
      if (lit)
        k_block_pass<MODE_P2G, COMPOSE_LITERAL, BATCHED><<<grid, TPB, 0, s>>>(ba);
      else
        k_block_pass<MODE_P2G, COMPOSE_TEXTBOOK, BATCHED><<<grid, TPB, 0, s>>>(ba);
      break;
    case MODE_FUSED:
      if (lit)
        k_block_pass<MODE_FUSED, COMPOSE_LITERAL, BATCHED><<<grid, TPB, 0, s>>>(ba);
      else
        k_block_pass<MODE_FUSED, COMPOSE_TEXTBOOK, BATCHED><<<grid, TPB, 0, s>>>(ba);
      break;
    case MODE_MASS: k_block_pass<MODE_MASS, COMPOSE_LITERAL, BATCHED><<<grid, TPB, 0, s>>>(ba); break;
    default: k_block_pass<MODE_G2P, COMPOSE_LITERAL, BATCHED><<<grid, TPB, 0, s>>>(ba); break;
  }
}

void launch_block_pass(const BlockArgs& ba, int mode, hipStream_t s) {
  if (ba.a.nscenes == 1)
    launch_block_pass_t<0>(ba, mode, s);
  else
    launch_block_pass_t<1>(ba, mode, s);
}

void launch_adj_g2p_block(const BlockArgs& ba, const float* St, const float* Xn, const float* Gin, float* Gout,
                          float4* gnode, const float4* vnode, hipStream_t s) {
  if (ba.a.nscenes == 1)
    k_adj_g2p_block<0><<<ba.grid, TPB, 0, s>>>(ba, St, Xn, Gin, Gout, gnode, vnode);
  else
    k_adj_g2p_block<1><<<ba.grid, TPB, 0, s>>>(ba, St, Xn, Gin, Gout, gnode, vnode);
}

void launch_grid_active(const BlockArgs& ba, hipStream_t s) {
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((ba.nblocks + 3) / 4, 2048));
  if (ba.a.nscenes == 1)
    k_grid_active<0><<<grid, TPB, 0, s>>>(ba);
  else
    k_grid_active<1><<<grid, TPB, 0, s>>>(ba);
}

}  // namespace mpm
