// Reverse-mode kernels of the MPM snow substep (gfx950): the adjoint of one
// substep, stage by stage, for the TEXTBOOK composition with domain walls.
//
// Forward substep n (material_point_method.cpp:94-350 in main.cpp's order,
// with the force impulse folded into the momentum as in the fused pass):
//   P2G   m_i = sum_p w m,  P_i = sum_p w (m v + Q (X_i - x)),
//         Q = m B D^-1 - dt V0 D^-1 P(F_E, F_P, E) F_E^T
//   grid  v_i = walls(P_i / m_i + dt g)
//   F     F^ = (I + dt B D^-1) F_E,  F_tot = F^ F_P,
//         F_E' = V clamp(S) U^T (F^ = V S U^T),  F_P' = F_E'^-1 F_tot
//   G2P   v' = sum_i w v_i,  B' = sum_i w v_i (X_i - x)^T
//   x'    = clamp(x + dt v')
// The reverse runs these backwards from the stored state of substep n:
//   k_adj_map       the block map of the substep's grid record (set before,
//                   reset with the node adjoints cleared after)
//   k_adj_g2p_block advection and G2P (mpm_fused.hip, on the block halos):
//                   node velocities from the substep's recorded P2G sums, node
//                   velocity adjoints and the position adjoint through the
//                   weights
//   k_adj_grid      wall mask and v = P/m + dt g: adjoints of (m_i, P_i),
//                   on the grid blocks the substep's record lists
//   k_adj_p2g       P2G (gathered), stress, hardening and the F update:
//                   adjoints of x, v, B, F_E, F_P of substep n and of E0
// Derivatives of functions of singular values use the divided-difference
// form in the singular basis (stable for repeated singular values); the
// polar rotation is the case f = 1.  V0 is held fixed.
#include "mpm_adjoint.hpp"

namespace mpm {
namespace {

constexpr int TPB = 256;
inline int nblk(int64_t n) { return (int)((n + TPB - 1) / TPB); }

__device__ __forceinline__ float fld(const float* P, int64_t stride, int f, int p) { return P[(int64_t)f * stride + p]; }
__device__ __forceinline__ M3 ld_m3(const float* P, int64_t stride, int f, int p) {
  M3 m;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) m.c[c][r] = P[(int64_t)(f + 3 * c + r) * stride + p];
  return m;
}
__device__ __forceinline__ void st_m3(float* P, int64_t stride, int f, int p, const M3& m) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) P[(int64_t)(f + 3 * c + r) * stride + p] = m.c[c][r];
}

__device__ __forceinline__ AxisD axis_d(float x, const Scene& S, int hi) {
  return axis_weights_d(x, S.h, S.inv_h, 1, hi);
}

__device__ __forceinline__ M3 mzero() { return mdiag(0.0f); }
__device__ __forceinline__ float mdot(const M3& a, const M3& b) {
  float s = 0.0f;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) s = fmaf(a.c[c][r], b.c[c][r], s);
  return s;
}
// row-major math view helpers: A(r, c) = A.c[c][r]
__device__ __forceinline__ M3 from_rows(const float (&a)[3][3]) {
  M3 m;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) m.c[c][r] = a[r][c];
  return m;
}

// Adjoint of F -> Lb diag(f(s)) Rb^T, a function of the singular values of
// F = Lb diag(s) Rb^T (f' = fp): G = Lb^T gOut Rb, then in the singular basis
// H_ii = f'_i G_ii, H_ij = a (G_ij + G_ji)/2 + b (G_ij - G_ji)/2 with the
// divided differences a = (f_i - f_j)/(s_i - s_j), b = (f_i + f_j)/(s_i + s_j);
// returns Lb H Rb^T.
__device__ __forceinline__ M3 spectral_adjoint(const M3& Lb, const M3& Rb, const float s[3], const float f[3],
                                               const float fp[3], const M3& gOut) {
  const M3 G = mmul(mmul(mtrans(Lb), gOut), Rb);  // G(r,c) = G.c[c][r]
  M3 H;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      if (i == j) {
        H.c[i][i] = fp[i] * G.c[i][i];
      } else {
        const float gij = G.c[j][i], gji = G.c[i][j];
        const float ds = s[i] - s[j];
        const float a = fabsf(ds) > 1e-6f * fmaxf(1.0f, fmaxf(s[i], s[j])) ? (f[i] - f[j]) / ds : 0.5f * (fp[i] + fp[j]);
        const float ss = s[i] + s[j];
        const float b = ss > 0.0f ? (f[i] + f[j]) / ss : 0.0f;
        H.c[j][i] = 0.5f * a * (gij + gji) + 0.5f * b * (gij - gji);
      }
    }
  return mmul(mmul(Lb, H), mtrans(Rb));
}

// SVD of F as F = Lb diag(s) Rb^T from the Eigen-convention solver of F^T
// (F^T = U S V^T  =>  F = V S U^T: Lb = V, Rb = U).
__device__ __forceinline__ bool svd_of(const M3& F, M3& Lb, float (&s)[3], M3& Rb) {
  float M[3][3], U[3][3], V[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) M[i][j] = F.c[i][j];  // M(i,j) = F(j,i)
  if (!jacobi_svd3<float>(M, U, s, V)) return false;
  Lb = from_rows(V);
  Rb = from_rows(U);
  return true;
}

// ------------------------------------------------------------------ kernels

// wall mask and v = P/m + dt g: node adjoints (dL/dm, dL/dP) in place, on
// the grid blocks of the substep's record (one wave per listed block; the
// other nodes have no mass and keep a zero adjoint)
__global__ void __launch_bounds__(TPB) k_adj_grid(KArgs a, GridRec r, float4* gnode) {
  const int lane = threadIdx.x & 63;
  const int cnt = r.list[0];
  for (int s = blockIdx.x * 4 + (threadIdx.x >> 6); s < cnt; s += gridDim.x * 4) {
    const int gb = r.list[1 + s];
    const int64_t g = (int64_t)gb * 64 + lane;
    const float4 acc = r.acc[(int64_t)s * 64 + lane];
    const float4 gn = gnode[g];
    if (acc.x == 0.0f) {
      gnode[g] = make_float4(0.f, 0.f, 0.f, 0.f);
      continue;
    }
    const Scene S = a.scenes[a.blk_scene[gb]];
    const F3 vpre = f3(acc.y / acc.x + a.dt * S.g[0], acc.z / acc.x + a.dt * S.g[1], acc.w / acc.x + a.dt * S.g[2]);
    float mk[3] = {1.f, 1.f, 1.f};
    if (S.bc_mode & MPM_BC_DOMAIN) {
      const int64_t lb = gb - S.blk0;
      const int bk = (int)(lb % S.nb[2]);
      const int64_t t2 = lb / S.nb[2];
      const int bj = (int)(t2 % S.nb[1]), bi = (int)(t2 / S.nb[1]);
      const int i = bi * 4 + (lane >> 4), j = bj * 4 + ((lane >> 2) & 3), k = bk * 4 + (lane & 3);
      // domain_walls: floor band sticky, other bands clamp the outward component
      if (j < 3) {
        mk[0] = mk[1] = mk[2] = 0.f;
      } else {
        if ((i < 3 && vpre.x < 0.f) || (i >= S.dims[0] - 3 && vpre.x > 0.f)) mk[0] = 0.f;
        if (j >= S.dims[1] - 3 && vpre.y > 0.f) mk[1] = 0.f;
        if ((k < 3 && vpre.z < 0.f) || (k >= S.dims[2] - 3 && vpre.z > 0.f)) mk[2] = 0.f;
      }
    }
    const F3 gvp = f3(gn.y * mk[0], gn.z * mk[1], gn.w * mk[2]);
    // dL/dP = gv' / m and dL/dm = -(P / m) . dL/dP, by division: a node
    // reached only by stencil weights near the edge can have a denormal m,
    // whose reciprocal overflows (inf, and 0 * inf = NaN) although both
    // quotients are finite
    const F3 gP = f3(gvp.x / acc.x, gvp.y / acc.x, gvp.z / acc.x);
    const float gm = -((acc.y / acc.x) * gP.x + (acc.z / acc.x) * gP.y + (acc.w / acc.x) * gP.z);
    gnode[g] = make_float4(gm, gP.x, gP.y, gP.z);
  }
}

// set the record's block map (set = 1) before the substep is reversed and
// write the node velocities of its blocks, v = walls(P / m + dt g) as the
// grid pass computed them, to vnode (read by the reverse G2P); or after it
// (set = 0) reset the map to -1 and zero the node adjoints and velocities of
// its blocks, so both stay zero outside the blocks being reversed
__global__ void __launch_bounds__(TPB) k_adj_map(KArgs a, GridRec r, int set, float4* gnode, float4* vnode) {
  const int lane = threadIdx.x & 63;
  const int cnt = r.list[0];
  for (int s = blockIdx.x * 4 + (threadIdx.x >> 6); s < cnt; s += gridDim.x * 4) {
    const int gb = r.list[1 + s];
    const int64_t g = (int64_t)gb * 64 + lane;
    if (lane == 0) r.map[gb] = set ? s : -1;
    if (!set) {
      gnode[g] = make_float4(0.f, 0.f, 0.f, 0.f);
      vnode[g] = make_float4(0.f, 0.f, 0.f, 0.f);
      continue;
    }
    const float4 acc = r.acc[(int64_t)s * 64 + lane];
    F3 v = f3(0.f, 0.f, 0.f);
    if (acc.x != 0.0f) {
      const Scene S = a.scenes[a.blk_scene[gb]];
      v = f3(acc.y / acc.x + a.dt * S.g[0], acc.z / acc.x + a.dt * S.g[1], acc.w / acc.x + a.dt * S.g[2]);
      if (S.bc_mode & MPM_BC_DOMAIN) {
        const int64_t lb = gb - S.blk0;
        const int bk = (int)(lb % S.nb[2]);
        const int64_t t2 = lb / S.nb[2];
        const int bj = (int)(t2 % S.nb[1]), bi = (int)(t2 / S.nb[1]);
        v = domain_walls(bi * 4 + (lane >> 4), bj * 4 + ((lane >> 2) & 3), bk * 4 + (lane & 3), S.dims[0], S.dims[1],
                         S.dims[2], v);
      }
    }
    vnode[g] = make_float4(v.x, v.y, v.z, 0.f);
  }
}

// The SVD F = Lb diag(s) Rb^T from the 21 rows f_update_t recorded for F^T
// (U, S, V: Lb = V, Rb = U), singular values clamped to [lo, hi] when the
// rows describe the update that produced F (F_E' = V clamp(S) U^T).
__device__ __forceinline__ void svd_from_rows(const float* Sv, int64_t sd, int p, bool clamp, float lo, float hi,
                                              M3& Lb, float (&s)[3], M3& Rb) {
  float U[3][3], V[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float si = Sv[(int64_t)(9 + i) * sd + p];
    s[i] = clamp ? fminf(fmaxf(si, lo), hi) : si;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      U[i][j] = Sv[(int64_t)(3 * i + j) * sd + p];
      V[i][j] = Sv[(int64_t)(12 + 3 * i + j) * sd + p];
    }
  }
  Lb = from_rows(V);
  Rb = from_rows(U);
}

// P2G, stress, hardening and F-update adjoint for substep n.  St: state of
// substep n; Gin: adjoints of substep n+1 (F_E, F_P); Gout: on entry the
// partial position adjoint, on exit all adjoints of substep n; gE += dL/dE0.
// Sv: the SVD the forward's F update of substep n used; SvPrev: that of
// substep n - 1, whose clamped singular values and vectors are the SVD of
// F_E of substep n (null at a segment's first substep: computed here).
__global__ void __launch_bounds__(TPB) k_adj_p2g(KArgs a, const float* St, const float* Gin, float* Gout, const float4* gnode, float* gE,
                                                 const float* Sv, const float* SvPrev) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  const Scene S = a.scenes[scene_of_slot(a.scenes, a.nscenes, p)];
  const int64_t sd = a.stride;
  const float dinv = S.dinv;
  const M3 Dinv = dinv_matrix(dinv);
  const float nu = S.lame.nu, xi = S.lame.xi;
  // Only the SVD of F_E, Q and the weights stay live across the gather (the
  // other stress pieces are rebuilt from reloaded rows after it), so the
  // gather can keep several node loads in flight.
  M3 Lb, Rb;
  float sv[3];
  M3 Q;
  F3 x, mv;
  float m;
  {
    x = f3(fld(St, sd, FX, p), fld(St, sd, FX + 1, p), fld(St, sd, FX + 2, p));
    const F3 v = f3(fld(St, sd, FV, p), fld(St, sd, FV + 1, p), fld(St, sd, FV + 2, p));
    const M3 B = ld_m3(St, sd, FB, p), FE = ld_m3(St, sd, FFE, p), FP = ld_m3(St, sd, FFP, p);
    m = fld(a.P, sd, FM, p);
    const float V0 = fld(a.P, sd, FVOL, p), E = fld(a.P, sd, FE0, p);
    const float hard = expf(xi * (1.0f - mdet(FP)));
    const float mu = E / S.lame.dmu * hard, lam = (E * nu) / S.lame.dlam * hard;
    // the previous update's SVD is F_E's unless an FP64 re-solve replaced it
    if (SvPrev && S.gap64 <= 0.0f)
      svd_from_rows(SvPrev, sd, p, true, S.lo, S.hi, Lb, sv, Rb);
    else
      svd_of(FE, Lb, sv, Rb);
    const M3 R = mmul(Lb, mtrans(Rb));
    const float J = mdet(FE);
    const M3 K = mtrans(minverse(FE));  // F_E^-T
    const float cJ = (J - 1.0f) * J;
    const M3 Pk = madd(mscale(msub(FE, R), 2.0f * mu), mscale(K, lam * cJ));
    // D^-1 = d I (mul_dinv): the products with it are scalings
    const M3 A = mscale(mmul(Pk, mtrans(FE)), Dinv.c[0][0] * V0);
    const M3 BD = mul_dinv(B, Dinv);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) Q.c[c][r] = m * BD.c[c][r] - a.dt * A.c[c][r];
    mv = v * m;
  }
  // gather the node adjoints.  Per node i (weights w_i, derivatives dw_i,
  // d_i = x_i - x_p) the adjoint of P2G contributes
  //   gx += dw_i s_i - w_i Q^T gP_i,  s_i = m gm_i + gP_i . (m v + Q d_i)
  //   gv += m w_i gP_i,               gQ += w_i gP_i d_i^T
  // The weights are products over the axes, so every sum but s_i's is
  // separable: inner sums over z, then y, then x, as the forward G2P does
  // (about 20 vector instructions per node instead of 55), and
  // sum w_i Q^T gP_i = Q^T sum w_i gP_i.
  AxisD ax[3];
  ax[0] = axis_d(x.x, S, S.dims[0] - 3);
  ax[1] = axis_d(x.y, S, S.dims[1] - 3);
  ax[2] = axis_d(x.z, S, S.dims[2] - 3);
  float dd[3][4];  // d_i per axis, rounded as the forward rounds them
  {
    const float xc[3] = {x.x, x.y, x.z};
#pragma unroll
    for (int e = 0; e < 3; ++e)
#pragma unroll
      for (int k = 0; k < 4; ++k) dd[e][k] = (float)(ax[e].base - 1 + k) * S.h - xc[e];
  }
  // columns of Q: (Q d)_r = sum_c Q.c[c][r] d_c
  const F3 q0 = f3(Q.c[0][0], Q.c[0][1], Q.c[0][2]), q1 = f3(Q.c[1][0], Q.c[1][1], Q.c[1][2]),
           q2 = f3(Q.c[2][0], Q.c[2][1], Q.c[2][2]);
  F3 G0 = f3(0.f, 0.f, 0.f), Gx = G0, Gy = G0, Gz = G0;  // sum w gP, and times d_x, d_y, d_z
  float SX = 0.f, SY = 0.f, SZ = 0.f;                     // sum dw s per axis
#pragma unroll 1
  for (int ia = 0; ia < 4; ++ia) {
    const F3 ua = mv + q0 * dd[0][ia];
    F3 T0 = f3(0.f, 0.f, 0.f), Ty = T0, Tz = T0;
    float Q0 = 0.f, Q1 = 0.f, Q2 = 0.f;
#pragma unroll
    for (int ib = 0; ib < 4; ++ib) {
      __asm__ __volatile__("" ::: "memory");  // one z column of four node loads in flight
      const int i = ax[0].base - 1 + ia, j = ax[1].base - 1 + ib;
      float4 gcol[4];
#pragma unroll
      for (int ic = 0; ic < 4; ++ic) gcol[ic] = gnode[node_index(i, j, ax[2].base - 1 + ic, S)];
      const F3 uab = ua + q1 * dd[1][ib];
      F3 R0 = f3(0.f, 0.f, 0.f), Rz = R0;
      float P0 = 0.f, P1 = 0.f;
#pragma unroll
      for (int ic = 0; ic < 4; ++ic) {
        const float4 gn = gcol[ic];
        const F3 gP = f3(gn.y, gn.z, gn.w);
        const F3 u = uab + q2 * dd[2][ic];
        const float sn = m * gn.x + dot(gP, u);
        const float wz = ax[2].w[ic];
        P0 = fmaf(wz, sn, P0);
        P1 = fmaf(ax[2].dw[ic], sn, P1);
        R0 = R0 + gP * wz;
        Rz = Rz + gP * (wz * dd[2][ic]);
      }
      const float wy = ax[1].w[ib];
      Q0 = fmaf(wy, P0, Q0);
      Q1 = fmaf(ax[1].dw[ib], P0, Q1);
      Q2 = fmaf(wy, P1, Q2);
      T0 = T0 + R0 * wy;
      Ty = Ty + R0 * (wy * dd[1][ib]);
      Tz = Tz + Rz * wy;
    }
    const float wx = ax[0].w[ia];
    SX = fmaf(ax[0].dw[ia], Q0, SX);
    SY = fmaf(wx, Q1, SY);
    SZ = fmaf(wx, Q2, SZ);
    G0 = G0 + T0 * wx;
    Gx = Gx + T0 * (wx * dd[0][ia]);
    Gy = Gy + Ty * wx;
    Gz = Gz + Tz * wx;
  }
  F3 gx = f3(fld(Gout, sd, AX, p), fld(Gout, sd, AX + 1, p), fld(Gout, sd, AX + 2, p));
  // Q^T G0
  const F3 QtG = f3(dot(q0, G0), dot(q1, G0), dot(q2, G0));
  gx = gx + f3(SX, SY, SZ) - QtG;
  F3 gv = G0 * m;
  M3 gQ;
  gQ.c[0][0] = Gx.x;
  gQ.c[0][1] = Gx.y;
  gQ.c[0][2] = Gx.z;
  gQ.c[1][0] = Gy.x;
  gQ.c[1][1] = Gy.y;
  gQ.c[1][2] = Gy.z;
  gQ.c[2][0] = Gz.x;
  gQ.c[2][1] = Gz.y;
  gQ.c[2][2] = Gz.z;
  __asm__ __volatile__("" ::: "memory");  // reload the rows below rather than keep them live
  // forward stress pieces, rebuilt
  const M3 B = ld_m3(St, sd, FB, p), FE = ld_m3(St, sd, FFE, p), FP = ld_m3(St, sd, FFP, p);
  const float V0 = fld(a.P, sd, FVOL, p), E = fld(a.P, sd, FE0, p);
  const float mu0 = E / S.lame.dmu, lam0 = (E * nu) / S.lame.dlam;
  const float JP = mdet(FP);
  const float hard = expf(xi * (1.0f - JP));
  const float mu = mu0 * hard, lam = lam0 * hard;
  const M3 R = mmul(Lb, mtrans(Rb));
  const float J = mdet(FE);
  const M3 K = mtrans(minverse(FE));  // F_E^-T
  const float cJ = (J - 1.0f) * J;
  const M3 Pk = madd(mscale(msub(FE, R), 2.0f * mu), mscale(K, lam * cJ));
  const M3 BD = mul_dinv(B, Dinv);
  // Q = m B D - dt A
  M3 gB = mscale(mul_dinv(gQ, Dinv), m);
  const M3 gA = mscale(gQ, -a.dt);
  // A = V0 D P F_E^T
  const M3 gPk = mscale(mmul(mscale(gA, Dinv.c[0][0]), FE), V0);
  M3 gFE = mmul(mtrans(gA), mscale(mscale(Pk, Dinv.c[0][0]), V0));
  // P = 2 mu (F - R) + lam c(J) K
  const float gmu = 2.0f * mdot(gPk, msub(FE, R));
  const float glam = cJ * mdot(gPk, K);
  gFE = madd(gFE, mscale(gPk, 2.0f * mu));
  {
    const float one[3] = {1.f, 1.f, 1.f}, zero[3] = {0.f, 0.f, 0.f};
    gFE = madd(gFE, spectral_adjoint(Lb, Rb, sv, one, zero, mscale(gPk, -2.0f * mu)));
  }
  const M3 gK = mscale(gPk, lam * cJ);
  const float gJ = lam * mdot(gPk, K) * (2.0f * J - 1.0f);
  gFE = madd(gFE, mscale(K, gJ * J));
  gFE = msub(gFE, mmul(mmul(K, mtrans(gK)), K));
  // mu = E / dmu * hard, lam = E nu / dlam * hard, hard = exp(xi (1 - det F_P))
  gE[p] += hard * (gmu / S.lame.dmu + glam * nu / S.lame.dlam);
  const float ghard = gmu * mu0 + glam * lam0;
  M3 gFP = mscale(mtrans(minverse(FP)), -xi * hard * ghard * JP);
  // F update (TEXTBOOK)
  {
    const M3 gFEn = ld_m3(Gin, sd, AFE, p), gFPn = ld_m3(Gin, sd, AFP, p);
    const M3 Fh = mmul(madd(mdiag(1.0f), mscale(BD, a.dt)), FE);
    const M3 Ftot = mmul(Fh, FP);
    M3 Lh, Rh;
    float sh[3];
    svd_from_rows(Sv, sd, p, false, 0.0f, 0.0f, Lh, sh, Rh);
    float f[3], fp[3];
    for (int i = 0; i < 3; ++i) {
      f[i] = fminf(fmaxf(sh[i], S.lo), S.hi);
      fp[i] = (sh[i] > S.lo && sh[i] < S.hi) ? 1.0f : 0.0f;
    }
    M3 Sc = mdiag(0.0f);
    for (int i = 0; i < 3; ++i) Sc.c[i][i] = f[i];
    const M3 FEn = mmul(mmul(Lh, Sc), mtrans(Rh));
    const M3 Y = minverse(FEn);
    // F_P' = Y F_tot
    const M3 gY = mmul(gFPn, mtrans(Ftot));
    const M3 gFtot = mmul(mtrans(Y), gFPn);
    const M3 gFEtot = msub(gFEn, mmul(mmul(mtrans(Y), gY), mtrans(Y)));
    M3 gFh = spectral_adjoint(Lh, Rh, sh, f, fp, gFEtot);
    gFh = madd(gFh, mmul(gFtot, mtrans(FP)));
    gFP = madd(gFP, mmul(mtrans(Fh), gFtot));
    // F^ = (I + dt B D) F_E
    gFE = madd(gFE, mmul(mtrans(madd(mdiag(1.0f), mscale(BD, a.dt))), gFh));
    gB = madd(gB, mscale(mul_dinv(mmul(gFh, mtrans(FE)), Dinv), a.dt));
  }
  Gout[(int64_t)AX * sd + p] = gx.x;
  Gout[(int64_t)(AX + 1) * sd + p] = gx.y;
  Gout[(int64_t)(AX + 2) * sd + p] = gx.z;
  Gout[(int64_t)AV * sd + p] = gv.x;
  Gout[(int64_t)(AV + 1) * sd + p] = gv.y;
  Gout[(int64_t)(AV + 2) * sd + p] = gv.z;
  st_m3(Gout, sd, AB, p, gB);
  st_m3(Gout, sd, AFE, p, gFE);
  st_m3(Gout, sd, AFP, p, gFP);
}

// element di of `rows` rows of dst <- element si of those of src; eight
// loads in flight before their stores
__device__ __forceinline__ void copy_rows(float* __restrict__ dst, int64_t di, const float* __restrict__ src, int64_t si,
                                          int64_t stride, int rows) {
  int f = 0;
  for (; f + 8 <= rows; f += 8) {
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = src[(int64_t)(f + k) * stride + si];
#pragma unroll
    for (int k = 0; k < 8; ++k) dst[(int64_t)(f + k) * stride + di] = v[k];
  }
  for (; f < rows; ++f) dst[(int64_t)f * stride + di] = src[(int64_t)f * stride + si];
}

// slot <-> caller order of `rows` consecutive rows
__global__ void __launch_bounds__(TPB) k_rows_to_slots(float* dst, const float* src, const int* pid, int64_t stride, int n, int rows) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  copy_rows(dst, p, src, pid[p], stride, rows);
}
// inv = pid^-1: the slot of each caller index
__global__ void __launch_bounds__(TPB) k_invert(const int* pid, int* inv, int n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) inv[pid[p]] = p;
}
__global__ void __launch_bounds__(TPB) k_add_to_caller(float* dst, const float* src, const int* pid, int n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  dst[pid[p]] += src[p];
}
// n x comps interleaved <-> comps rows of `stride`
__global__ void __launch_bounds__(TPB) k_aos_to_rows(float* dst, const float* src, int n, int comps, int64_t stride) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  for (int c = 0; c < comps; ++c) dst[(int64_t)c * stride + p] = src[(int64_t)p * comps + c];
}
__global__ void __launch_bounds__(TPB) k_rows_to_aos(float* dst, const float* src, int n, int comps, int64_t stride) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  for (int c = 0; c < comps; ++c) dst[(int64_t)p * comps + c] = src[(int64_t)c * stride + p];
}

}  // namespace

// one wave per listed grid block, up to 2048 workgroups of 4 waves
inline int rec_grid(const KArgs& a) { return (int)std::max<int64_t>(1, std::min<int64_t>((a.nnodes / 64 + 3) / 4, 2048)); }
void launch_adj_grid(const KArgs& a, const GridRec& r, float4* gnode, hipStream_t s) {
  k_adj_grid<<<rec_grid(a), TPB, 0, s>>>(a, r, gnode);
}
void launch_adj_map(const KArgs& a, const GridRec& r, int set, float4* gnode, float4* vnode, hipStream_t s) {
  k_adj_map<<<rec_grid(a), TPB, 0, s>>>(a, r, set, gnode, vnode);
}
void launch_adj_p2g(const KArgs& a, const float* St, const float* Gin, float* Gout, const float4* gnode, float* gE,
                    const float* Sv, const float* SvPrev, hipStream_t s) {
  if (a.n) k_adj_p2g<<<nblk(a.n), TPB, 0, s>>>(a, St, Gin, Gout, gnode, gE, Sv, SvPrev);
}
void launch_rows_to_slots(float* dst, const float* src, const int* pid, int64_t stride, int n, int rows,
                          hipStream_t s) {
  if (n) k_rows_to_slots<<<nblk(n), TPB, 0, s>>>(dst, src, pid, stride, n, rows);
}
void launch_rows_to_caller(float* dst, const float* src, const int* pid, int* inv, int64_t stride, int n, int rows,
                           hipStream_t s) {
  // written as a gather through the inverse permutation: coalesced stores,
  // where scattering the rows to caller order would store 4 bytes per line
  if (!n) return;
  k_invert<<<nblk(n), TPB, 0, s>>>(pid, inv, n);
  k_rows_to_slots<<<nblk(n), TPB, 0, s>>>(dst, src, inv, stride, n, rows);
}
void launch_add_to_caller(float* dst, const float* src, const int* pid, int n, hipStream_t s) {
  if (n) k_add_to_caller<<<nblk(n), TPB, 0, s>>>(dst, src, pid, n);
}
void launch_aos_to_rows(float* dst, const float* src, int n, int comps, int64_t stride, hipStream_t s) {
  if (n) k_aos_to_rows<<<nblk(n), TPB, 0, s>>>(dst, src, n, comps, stride);
}
void launch_rows_to_aos(float* dst, const float* src, int n, int comps, int64_t stride, hipStream_t s) {
  if (n) k_rows_to_aos<<<nblk(n), TPB, 0, s>>>(dst, src, n, comps, stride);
}

}  // namespace mpm
