// =============================================================================
// ORACLE — test infrastructure only.
//
// Nothing here is part of the product: only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg load it, and only as the checker or as the
// CPU baseline that is timed beside the GPU.  The product (the HIP library)
// never includes, links or calls this file.
//
// A from-scratch CPU restatement of the MPM snow substep of
// brocbyte/realtime-deformations, written from the mathematics and from the
// reference's algorithm, following its structure stage by stage:
//   realtime-deformations/material_point_method.cpp:80-142   stencil, P2G, V0
//   realtime-deformations/material_point_method.cpp:235-350  forces, grid
//       update, collisions, F update, G2P, advection
//   realtime-deformations/material_point_method.cpp:381-388  position clamp
//   realtime-deformations/material_point_method.hpp:18-31,53-58  B-spline
//   realtime-deformations/material_point_method.hpp:79-86  box SDF collider
//   realtime-deformations/mathy.hpp:40-56  finite-difference normal
//   realtime-deformations/utils.h:15-33   glm<->Eigen transposes (F2 quirk)
//   realtime-deformations/main.cpp:192-218 stage order
//
// Scalar T = float mirrors the reference's FP32 path operation for operation:
// glm's column-major matrices and its evaluation order, the double promotion
// inside the weight polynomial, Eigen's two-sided Jacobi SVD conventions.
// T = double is the FP64 build used to state each quantity's error.  Build
// with -ffp-contract=off (no FMA contraction), as the reference is built.
//
// How each function is pinned (tests/test_oracle_*.py):
//   weightN            closed forms N(0)=2/3, N(1)=1/6, N(1/2)=23/48,
//                      N(3/2)=1/48, N(2)=0; partition of unity; linear
//                      reproduction; brute force against exact rationals.
//   glm ops (mul, inverse, det, mat4 inverse)
//                      bit-equal to the reference's vendored glm 0.9.7.1,
//                      compiled in oracle/_ref (golden vectors).
//   jacobi_svd3        bit-equal to the reference's vendored Eigen 3.4.90
//                      JacobiSVD<MatrixXf>, compiled in oracle/_ref, on
//                      near-identity, random, diagonal-tie and degenerate
//                      inputs; U Σ Vᵀ = M and orthogonality in FP64.
//   f_update (LITERAL) bit-equal to the reference's own composition
//                      (glm + Eigen) compiled in oracle/_ref.
//   f_update (TEXTBOOK) J_E·J_P = det F_tot, clamp bounds, idempotence.
//   polar_R            within 1e-6 of the reference's vendored
//                      Higham–Noferini polar decomposition (oracle/_ref);
//                      orthonormality, det +1, RᵀF symmetric.
//   box_sdf / collide  the reference's SDF on glm (oracle/_ref) for SDF
//                      values; hand-computed response cases.
//   p2g                mass and momentum conservation; affine
//                      reproduction: one particle with C = B D^-1 = A
//                      (non-symmetric) gives v_i = v_p + A (x_i - x_p).
//   init_volumes       exact rationals (dyadic x / h): one isolated particle
//                      has V0 = h^3 / prod_axes sum_d N(u_d)^2; overlapping
//                      particles against a brute-force rational sum.
//   grid update        sum m_i v_i = dt M g and the first moment
//                      sum m_i v_i x_i^T = dt (-sum_p V0 P F^T +
//                      g sum_p m_p x_p^T) against an FP64 stress.
//   g2p                uniform velocity gives B = 0, v_p = v_i; an affine
//                      grid field b + A x gives v_p = b + A x_p and
//                      B D^-1 = A (non-symmetric A).
//   advect             exact free fall over 100 substeps (closed form),
//                      clamp bounds.
//   threaded F update  bit-equal to the serial loop, stopping at the first
//                      failing particle as the reference's loop does.
//   forces             F_E = I and F_E = R give zero force; small-strain
//                      linearisation against a finite difference of the
//                      elastic energy.
//   domain walls       no reference counterpart (configs' sticky floor and
//                      separating walls): hand-computed cases.
// =============================================================================
#pragma once
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace mpmo {

// ---------------------------------------------------------------- glm layer
template <typename T> struct V3 { T x, y, z; };
// Column-major like glm: c[col][row].
template <typename T> struct M3 { T c[3][3]; };

template <typename T> inline V3<T> vadd(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> inline V3<T> vsub(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> inline V3<T> vmul(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T> inline V3<T> vdiv(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
// glm dot: tmp = a*b; tmp.x + tmp.y + tmp.z
template <typename T> inline T vdot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

template <typename T> inline M3<T> mdiag(T d) {
  M3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.c[i][j] = (i == j) ? d : T(0);
  return r;
}
template <typename T> inline M3<T> madd(const M3<T>& a, const M3<T>& b) {
  M3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.c[i][j] = a.c[i][j] + b.c[i][j];
  return r;
}
template <typename T> inline M3<T> msub(const M3<T>& a, const M3<T>& b) {
  M3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.c[i][j] = a.c[i][j] - b.c[i][j];
  return r;
}
template <typename T> inline M3<T> mscale(const M3<T>& a, T s) {
  M3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.c[i][j] = a.c[i][j] * s;
  return r;
}
// glm mat3*mat3: R[j][r] = A[0][r]*B[j][0] + A[1][r]*B[j][1] + A[2][r]*B[j][2]
template <typename T> inline M3<T> mmul(const M3<T>& a, const M3<T>& b) {
  M3<T> r;
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i)
      r.c[j][i] = a.c[0][i] * b.c[j][0] + a.c[1][i] * b.c[j][1] + a.c[2][i] * b.c[j][2];
  return r;
}
template <typename T> inline V3<T> mvec(const M3<T>& m, V3<T> v) {
  return {m.c[0][0] * v.x + m.c[1][0] * v.y + m.c[2][0] * v.z,
          m.c[0][1] * v.x + m.c[1][1] * v.y + m.c[2][1] * v.z,
          m.c[0][2] * v.x + m.c[1][2] * v.y + m.c[2][2] * v.z};
}
template <typename T> inline M3<T> mtrans(const M3<T>& a) {
  M3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.c[i][j] = a.c[j][i];
  return r;
}
// glm outerProduct(c, r): column j = c * r[j]
template <typename T> inline M3<T> mouter(V3<T> c, V3<T> r) {
  const T rr[3] = {r.x, r.y, r.z};
  M3<T> m;
  for (int j = 0; j < 3; ++j) {
    m.c[j][0] = c.x * rr[j];
    m.c[j][1] = c.y * rr[j];
    m.c[j][2] = c.z * rr[j];
  }
  return m;
}
// glm func_matrix.inl:231-240
template <typename T> inline T mdet(const M3<T>& m) {
  return +m.c[0][0] * (m.c[1][1] * m.c[2][2] - m.c[2][1] * m.c[1][2])
         - m.c[1][0] * (m.c[0][1] * m.c[2][2] - m.c[2][1] * m.c[0][2])
         + m.c[2][0] * (m.c[0][1] * m.c[1][2] - m.c[1][1] * m.c[0][2]);
}
// glm type_mat3x3.inl:37-56
template <typename T> inline M3<T> minverse(const M3<T>& m) {
  const T ood = static_cast<T>(1) / (+m.c[0][0] * (m.c[1][1] * m.c[2][2] - m.c[2][1] * m.c[1][2])
                                     - m.c[1][0] * (m.c[0][1] * m.c[2][2] - m.c[2][1] * m.c[0][2])
                                     + m.c[2][0] * (m.c[0][1] * m.c[1][2] - m.c[1][1] * m.c[0][2]));
  M3<T> r;
  r.c[0][0] = +(m.c[1][1] * m.c[2][2] - m.c[2][1] * m.c[1][2]) * ood;
  r.c[1][0] = -(m.c[1][0] * m.c[2][2] - m.c[2][0] * m.c[1][2]) * ood;
  r.c[2][0] = +(m.c[1][0] * m.c[2][1] - m.c[2][0] * m.c[1][1]) * ood;
  r.c[0][1] = -(m.c[0][1] * m.c[2][2] - m.c[2][1] * m.c[0][2]) * ood;
  r.c[1][1] = +(m.c[0][0] * m.c[2][2] - m.c[2][0] * m.c[0][2]) * ood;
  r.c[2][1] = -(m.c[0][0] * m.c[2][1] - m.c[2][0] * m.c[0][1]) * ood;
  r.c[0][2] = +(m.c[0][1] * m.c[1][2] - m.c[1][1] * m.c[0][2]) * ood;
  r.c[1][2] = -(m.c[0][0] * m.c[1][2] - m.c[1][0] * m.c[0][2]) * ood;
  r.c[2][2] = +(m.c[0][0] * m.c[1][1] - m.c[1][0] * m.c[0][1]) * ood;
  return r;
}

// 4x4 (collider pose), glm layout c[col][row]
template <typename T> struct M4 { T c[4][4]; };
template <typename T> inline M4<T> m4mul(const M4<T>& a, const M4<T>& b) {
  // Result[j] = A0*B[j][0] + A1*B[j][1] + A2*B[j][2] + A3*B[j][3]
  M4<T> r;
  for (int j = 0; j < 4; ++j)
    for (int i = 0; i < 4; ++i)
      r.c[j][i] = a.c[0][i] * b.c[j][0] + a.c[1][i] * b.c[j][1] + a.c[2][i] * b.c[j][2] + a.c[3][i] * b.c[j][3];
  return r;
}
// glm type_mat4x4.inl:37-100 (cofactor inverse)
template <typename T> inline M4<T> m4inverse(const M4<T>& mm) {
  auto m = [&](int a, int b) { return mm.c[a][b]; };
  const T c00 = m(2, 2) * m(3, 3) - m(3, 2) * m(2, 3), c02 = m(1, 2) * m(3, 3) - m(3, 2) * m(1, 3),
          c03 = m(1, 2) * m(2, 3) - m(2, 2) * m(1, 3);
  const T c04 = m(2, 1) * m(3, 3) - m(3, 1) * m(2, 3), c06 = m(1, 1) * m(3, 3) - m(3, 1) * m(1, 3),
          c07 = m(1, 1) * m(2, 3) - m(2, 1) * m(1, 3);
  const T c08 = m(2, 1) * m(3, 2) - m(3, 1) * m(2, 2), c10 = m(1, 1) * m(3, 2) - m(3, 1) * m(1, 2),
          c11 = m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2);
  const T c12 = m(2, 0) * m(3, 3) - m(3, 0) * m(2, 3), c14 = m(1, 0) * m(3, 3) - m(3, 0) * m(1, 3),
          c15 = m(1, 0) * m(2, 3) - m(2, 0) * m(1, 3);
  const T c16 = m(2, 0) * m(3, 2) - m(3, 0) * m(2, 2), c18 = m(1, 0) * m(3, 2) - m(3, 0) * m(1, 2),
          c19 = m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2);
  const T c20 = m(2, 0) * m(3, 1) - m(3, 0) * m(2, 1), c22 = m(1, 0) * m(3, 1) - m(3, 0) * m(1, 1),
          c23 = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1);
  const T f0[4] = {c00, c00, c02, c03}, f1[4] = {c04, c04, c06, c07}, f2[4] = {c08, c08, c10, c11};
  const T f3[4] = {c12, c12, c14, c15}, f4[4] = {c16, c16, c18, c19}, f5[4] = {c20, c20, c22, c23};
  const T v0[4] = {m(1, 0), m(0, 0), m(0, 0), m(0, 0)}, v1[4] = {m(1, 1), m(0, 1), m(0, 1), m(0, 1)};
  const T v2[4] = {m(1, 2), m(0, 2), m(0, 2), m(0, 2)}, v3[4] = {m(1, 3), m(0, 3), m(0, 3), m(0, 3)};
  T i0[4], i1[4], i2[4], i3[4];
  for (int k = 0; k < 4; ++k) {
    i0[k] = v1[k] * f0[k] - v2[k] * f1[k] + v3[k] * f2[k];
    i1[k] = v0[k] * f0[k] - v2[k] * f3[k] + v3[k] * f4[k];
    i2[k] = v0[k] * f1[k] - v1[k] * f3[k] + v3[k] * f5[k];
    i3[k] = v0[k] * f2[k] - v1[k] * f4[k] + v2[k] * f5[k];
  }
  const T sa[4] = {1, -1, 1, -1}, sb[4] = {-1, 1, -1, 1};
  M4<T> inv;
  for (int k = 0; k < 4; ++k) {
    inv.c[0][k] = i0[k] * sa[k];
    inv.c[1][k] = i1[k] * sb[k];
    inv.c[2][k] = i2[k] * sa[k];
    inv.c[3][k] = i3[k] * sb[k];
  }
  const T row0[4] = {inv.c[0][0], inv.c[1][0], inv.c[2][0], inv.c[3][0]};
  T d0[4];
  for (int k = 0; k < 4; ++k) d0[k] = mm.c[0][k] * row0[k];
  const T dot1 = (d0[0] + d0[1]) + (d0[2] + d0[3]);
  const T ood = static_cast<T>(1) / dot1;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) inv.c[a][b] = inv.c[a][b] * ood;
  return inv;
}
// glm mat4 * vec4: (m0*v0 + m1*v1) + (m2*v2 + m3*v3)
template <typename T> inline void m4vec(const M4<T>& m, const T v[4], T out[4]) {
  for (int r = 0; r < 4; ++r) out[r] = (m.c[0][r] * v[0] + m.c[1][r] * v[1]) + (m.c[2][r] * v[2] + m.c[3][r] * v[3]);
}
// glm mat3_cast(quat) in a mat4; quaternion (w, x, y, z)
template <typename T> inline M4<T> quat_to_m4(const T q[4]) {
  const T w = q[0], x = q[1], y = q[2], z = q[3];
  const T qxx = x * x, qyy = y * y, qzz = z * z, qxz = x * z, qxy = x * y, qyz = y * z;
  const T qwx = w * x, qwy = w * y, qwz = w * z;
  M4<T> r;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) r.c[a][b] = (a == b) ? T(1) : T(0);
  r.c[0][0] = 1 - 2 * (qyy + qzz);
  r.c[0][1] = 2 * (qxy + qwz);
  r.c[0][2] = 2 * (qxz - qwy);
  r.c[1][0] = 2 * (qxy - qwz);
  r.c[1][1] = 1 - 2 * (qxx + qzz);
  r.c[1][2] = 2 * (qyz + qwx);
  r.c[2][0] = 2 * (qxz + qwy);
  r.c[2][1] = 2 * (qyz - qwx);
  r.c[2][2] = 1 - 2 * (qxx + qyy);
  return r;
}
// glm translate(mat4(1), t): column 3 = m0*t0 + m1*t1 + m2*t2 + m3
template <typename T> inline M4<T> translate_m4(const T t[3]) {
  M4<T> r;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) r.c[a][b] = (a == b) ? T(1) : T(0);
  for (int k = 0; k < 4; ++k) {
    const T e0 = (k == 0) ? T(1) : T(0), e1 = (k == 1) ? T(1) : T(0), e2 = (k == 2) ? T(1) : T(0),
            e3 = (k == 3) ? T(1) : T(0);
    r.c[3][k] = e0 * t[0] + e1 * t[1] + e2 * t[2] + e3;
  }
  return r;
}

// ----------------------------------------------------- B-spline weights (a1)
// material_point_method.hpp:20-31.  For float the polynomial is evaluated in
// double, as the reference's double literals promote it, then rounded.
inline float weightN(float x) {
  const float modx = std::fabs(x);
  const float modx2 = modx * modx;
  const float modx3 = modx * modx * modx;
  if (modx < 1.0) return (float)(0.5 * modx3 - modx2 + 2.0 / 3.0);
  if (modx < 2.0) return (float)((1.0 / 6.0) * (2 - modx) * (2 - modx) * (2 - modx));
  return 0.0f;
}
inline double weightN(double x) {
  const double modx = std::fabs(x);
  const double modx2 = modx * modx;
  const double modx3 = modx * modx * modx;
  if (modx < 1.0) return 0.5 * modx3 - modx2 + 2.0 / 3.0;
  if (modx < 2.0) return (1.0 / 6.0) * (2 - modx) * (2 - modx) * (2 - modx);
  return 0.0;
}

// ------------------------------------------- Eigen-convention Jacobi SVD (a7)
// Two-sided Jacobi SVD of a real 3x3 matrix M (math row-major M[r][c]), with
// the conventions of the reference's vendored Eigen 3.4.90 JacobiSVD
// (Eigen/src/SVD/JacobiSVD.h:689-819, misc/RealSvd2x2.h:21-51,
// Jacobi/Jacobi.h:96-127,297-345): scaling by max|M_ij|, sweep order
// (1,0),(2,0),(2,1), threshold max(min_normal, 2*eps*maxDiag), symmetrising
// rotation then makeJacobi, sign fix on U, descending selection sort with the
// first maximum kept.  Returns 0 on success, 1 on non-finite input.
template <typename T> struct Rot { T c, s; };

// x' = c x + s y ; y' = -s x + c y over n entries (skipped for the identity)
template <typename T> inline void rot_apply(T* x, T* y, int stride, Rot<T> j) {
  if (j.c == T(1) && j.s == T(0)) return;
  for (int i = 0; i < 3; ++i) {
    const T xi = x[i * stride], yi = y[i * stride];
    x[i * stride] = j.c * xi + j.s * yi;
    y[i * stride] = -j.s * xi + j.c * yi;
  }
}

template <typename T> inline void real_2x2_jacobi(const T W[3][3], int p, int q, Rot<T>* jl, Rot<T>* jr) {
  T m[2][2] = {{W[p][p], W[p][q]}, {W[q][p], W[q][q]}};
  Rot<T> r1;
  const T t = m[0][0] + m[1][1];
  const T d = m[1][0] - m[0][1];
  if (std::fabs(d) < std::numeric_limits<T>::min()) {
    r1.s = T(0);
    r1.c = T(1);
  } else {
    const T u = t / d;
    const T tmp = std::sqrt(T(1) + u * u);
    r1.s = T(1) / tmp;
    r1.c = u / tmp;
  }
  if (!(r1.c == T(1) && r1.s == T(0))) {
    for (int k = 0; k < 2; ++k) {
      const T xi = m[0][k], yi = m[1][k];
      m[0][k] = r1.c * xi + r1.s * yi;
      m[1][k] = -r1.s * xi + r1.c * yi;
    }
  }
  // makeJacobi(x = m00, y = m01, z = m11)
  const T x = m[0][0], y = m[0][1], z = m[1][1];
  const T deno = T(2) * std::fabs(y);
  if (deno < std::numeric_limits<T>::min()) {
    jr->c = T(1);
    jr->s = T(0);
  } else {
    const T tau = (x - z) / deno;
    const T w = std::sqrt(tau * tau + T(1));
    T tt;
    if (tau > T(0))
      tt = T(1) / (tau + w);
    else
      tt = T(1) / (tau - w);
    const T sign_t = tt > T(0) ? T(1) : T(-1);
    const T n = T(1) / std::sqrt(tt * tt + T(1));
    jr->s = -sign_t * (y / std::fabs(y)) * std::fabs(tt) * n;
    jr->c = n;
  }
  // j_left = rot1 * j_right.transpose()
  const T c2 = jr->c, s2 = -jr->s;
  jl->c = r1.c * c2 - r1.s * s2;
  jl->s = r1.c * s2 + r1.s * c2;
}

template <typename T> inline int jacobi_svd3(const T M[3][3], T U[3][3], T S[3], T V[3][3]) {
  T scale = T(0);
  bool nan = false;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      const T a = std::fabs(M[r][c]);
      if (a != a) nan = true;
      if (a > scale) scale = a;
    }
  if (nan || !std::isfinite(scale)) return 1;
  if (scale == T(0)) scale = T(1);
  T W[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      W[r][c] = M[r][c] / scale;
      U[r][c] = V[r][c] = (r == c) ? T(1) : T(0);
    }
  const T precision = T(2) * std::numeric_limits<T>::epsilon();
  const T zero_lim = std::numeric_limits<T>::min();
  T maxDiag = std::fabs(W[0][0]);
  maxDiag = std::max(maxDiag, std::fabs(W[1][1]));
  maxDiag = std::max(maxDiag, std::fabs(W[2][2]));
  bool finished = false;
  int sweeps = 0;
  while (!finished && sweeps < 64) {
    finished = true;
    ++sweeps;
    for (int p = 1; p < 3; ++p)
      for (int q = 0; q < p; ++q) {
        const T pm = precision * maxDiag;
        const T threshold = (zero_lim < pm) ? pm : zero_lim;
        if (std::fabs(W[p][q]) > threshold || std::fabs(W[q][p]) > threshold) {
          finished = false;
          Rot<T> jl, jr;
          real_2x2_jacobi(W, p, q, &jl, &jr);
          rot_apply(&W[p][0], &W[q][0], 1, jl);                 // rows p,q of W
          rot_apply(&U[0][p], &U[0][q], 3, jl);                 // cols p,q of U
          const Rot<T> jrt{jr.c, -jr.s};
          rot_apply(&W[0][p], &W[0][q], 3, jrt);                // cols p,q of W
          rot_apply(&V[0][p], &V[0][q], 3, jrt);                // cols p,q of V
          const T a = std::fabs(W[p][p]), b = std::fabs(W[q][q]);
          const T mx = (a < b) ? b : a;
          maxDiag = (maxDiag < mx) ? mx : maxDiag;
        }
      }
  }
  for (int i = 0; i < 3; ++i) {
    const T a = W[i][i];
    S[i] = std::fabs(a);
    if (a < T(0))
      for (int r = 0; r < 3; ++r) U[r][i] = -U[r][i];
  }
  for (int i = 0; i < 3; ++i) S[i] *= scale;
  for (int i = 0; i < 3; ++i) {
    int pos = i;
    T mx = S[i];
    for (int k = i + 1; k < 3; ++k)
      if (S[k] > mx) {
        mx = S[k];
        pos = k;
      }
    if (mx == T(0)) break;
    if (pos != i) {
      std::swap(S[i], S[pos]);
      for (int r = 0; r < 3; ++r) {
        std::swap(U[r][i], U[r][pos]);
        std::swap(V[r][i], V[r][pos]);
      }
    }
  }
  return 0;
}

// ------------------------------------------------------- polar (a4, F3)
// Rotation factor R of F = R S.  Any accurate polar decomposition reproduces
// the reference's (realtime-deformations/utils.h:55-75 ->
// include/polar_decomposition_3x3_impl.h:1040): computed here by scaled
// Newton iteration X <- (g X + X^{-T}/g)/2, g = |det X|^{-1/3}, in double,
// then rounded to T.  Like the reference's algorithm, the returned factor is
// always a rotation: for det F < 0 it is minus the orthogonal polar factor
// (and S = R^T F is then negative definite).
template <typename T> inline M3<T> polar_R(const M3<T>& F) {
  M3<double> X;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) X.c[i][j] = (double)F.c[i][j];
  const double sgn = mdet(X) < 0.0 ? -1.0 : 1.0;
  for (int it = 0; it < 100; ++it) {
    const double d = mdet(X);
    if (!(std::fabs(d) > 0.0) || !std::isfinite(d)) break;
    const M3<double> Yt = mtrans(minverse(X));
    const double g = std::pow(std::fabs(d), -1.0 / 3.0);
    M3<double> Xn;
    double diff = 0, nrm = 0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        Xn.c[i][j] = 0.5 * (g * X.c[i][j] + Yt.c[i][j] / g);
        diff += (Xn.c[i][j] - X.c[i][j]) * (Xn.c[i][j] - X.c[i][j]);
        nrm += Xn.c[i][j] * Xn.c[i][j];
      }
    X = Xn;
    if (diff <= 1e-30 * nrm) break;
  }
  M3<T> R;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R.c[i][j] = (T)(sgn * X.c[i][j]);
  return R;
}

// ------------------------------------------- deformation-gradient update (a7)
// material_point_method.cpp:306-330 for one particle.  Bn is the APIC matrix
// of the previous G2P (F1).  LITERAL reproduces the reference's composition
// F_E <- U^T diag(clamp(s)) V of the SVD of M = F_hat^T (utils.h:15-33 each
// transpose); TEXTBOOK uses the correct V diag(clamp(s)) U^T.  Returns 1 if
// the SVD rejects non-finite input (the reference then abandons the loop).
enum Compose { LITERAL = 0, TEXTBOOK = 1 };

template <typename T>
inline int f_update(int compose, T dt, const M3<T>& Dinv, T lo, T hi, const M3<T>& Bn, M3<T>& FE, M3<T>& FP,
                    T sig_out[3] = nullptr) {
  const M3<T> FPn1 = mmul(mmul(madd(mdiag(T(1)), mscale(mmul(Bn, Dinv), dt)), FE), FP);
  const M3<T> FEhat = mmul(FPn1, minverse(FP));
  T M[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) M[i][j] = FEhat.c[i][j];  // Eigen m(i,j) = glm[i][j]
  T U[3][3], S[3], V[3][3];
  if (jacobi_svd3(M, U, S, V)) return 1;
  if (sig_out)
    for (int i = 0; i < 3; ++i) sig_out[i] = S[i];
  for (int i = 0; i < 3; ++i) S[i] = std::clamp(S[i], lo, hi);
  M3<T> Ug, Vg, Sg = mdiag(T(0));
  for (int i = 0; i < 3; ++i) {
    Sg.c[i][i] = S[i];
    for (int j = 0; j < 3; ++j) {
      Ug.c[i][j] = U[i][j];  // eigenToGlm: glm[i][j] = E(i,j)
      Vg.c[i][j] = V[i][j];
    }
  }
  if (compose == LITERAL)
    FE = mmul(mmul(Ug, Sg), mtrans(Vg));
  else
    FE = mmul(mmul(mtrans(Vg), Sg), Ug);
  FP = mmul(minverse(FE), FPn1);
  return 0;
}

// ---------------------------------------------------------- colliders (a6)
// An oriented box given by the pose the reference reads from Object3D
// (translation, rotation quaternion w,x,y,z, scale = half extent) plus a
// velocity.  SDF per material_point_method.hpp:79-86 (L-inf exterior).
template <typename T> struct Collider {
  T t[3], q[4], s[3], vel[3];
  M4<T> inv;  // inverse(translate(t) * toMat4(q))
  void prepare() { inv = m4inverse(m4mul(translate_m4(t), quat_to_m4(q))); }
  T sdf(V3<T> pos) const {
    const T p4in[4] = {pos.x, pos.y, pos.z, T(1)};
    T p[4];
    m4vec(inv, p4in, p);
    const T qx = std::fabs(p[0]) - s[0], qy = std::fabs(p[1]) - s[1], qz = std::fabs(p[2]) - s[2];
    const T mx = std::max({qx, qy, qz, T(0)});
    return std::fabs(mx) + std::min(std::max(qx, std::max(qy, qz)), T(0));
  }
};

// mathy.hpp:40-56: central difference, delta = 1e-3, not normalised
template <typename T> inline V3<T> sdf_gradient(const Collider<T>& c, V3<T> x) {
  const T delta = T(0.001f);
  T g[3];
  for (int i = 0; i < 3; ++i) {
    const T e[3] = {T(i == 0), T(i == 1), T(i == 2)};
    const V3<T> d = {e[0] * delta, e[1] * delta, e[2] * delta};
    const T f1 = c.sdf(vsub(x, d));
    const T f2 = c.sdf(vadd(x, d));
    g[i] = (f2 - f1) / (T(2.0f) * delta);
  }
  return {g[0], g[1], g[2]};
}

// material_point_method.cpp:264-296 (glm vec3::length() is the component
// count 3, so the friction test compares 3 against -mu*vn)
template <typename T>
inline V3<T> body_collision(V3<T> pos, V3<T> vel, const std::vector<Collider<T>>& cs) {
  bool all_out = true;
  for (const auto& c : cs)
    if (!(c.sdf(pos) > T(0))) all_out = false;
  if (all_out) return vel;
  V3<T> out = vel;
  for (const auto& c : cs) {
    if (c.sdf(pos) > T(0)) continue;
    const V3<T> n = sdf_gradient(c, pos);
    const V3<T> vobj = {c.vel[0], c.vel[1], c.vel[2]};
    const V3<T> rel = vsub(out, vobj);
    const T vn = vdot(rel, n);
    if (vn >= T(0)) continue;
    const V3<T> vt = vsub(rel, vmul(n, vn));
    const T mu = T(0.5f);
    V3<T> vrel = {T(0), T(0), T(0)};
    const T len = T(3);
    if (len > -mu * vn && len > T(0)) vrel = vadd(vt, vmul(vt, mu * vn / len));
    out = vadd(vrel, vobj);
  }
  return out;
}

// Domain walls (no reference counterpart; the configs' "sticky floor and
// separating walls"): nodes with index < 3 or >= MAX-3 on an axis.  Floor
// band (j < 3) is sticky (v = 0); every other band removes the velocity
// component that points out of the domain.
template <typename T> inline V3<T> domain_walls(int i, int j, int k, int I, int J, int K, V3<T> v) {
  const int b = 3;
  if (j < b) return {T(0), T(0), T(0)};
  if (i < b && v.x < T(0)) v.x = T(0);
  if (i >= I - b && v.x > T(0)) v.x = T(0);
  if (j >= J - b && v.y > T(0)) v.y = T(0);
  if (k < b && v.z < T(0)) v.z = T(0);
  if (k >= K - b && v.z > T(0)) v.z = T(0);
  return v;
}

// ------------------------------------------------------------- simulation
enum BcMode { BC_COLLIDERS = 1, BC_DOMAIN = 2 };

struct Params {
  int32_t dims[3];
  int32_t compose;
  int32_t bc_mode;
  int32_t threads;
  double h, E, nu, xi, theta_c, theta_s;
  double g[3];
};

template <typename T> struct Sim {
  Params prm{};
  int I = 0, J = 0, K = 0, n = 0;
  T h = 0;
  M3<T> Dinv{};
  T lo = 0, hi = 0;
  std::vector<V3<T>> x, v;
  std::vector<M3<T>> B, FE, FP;
  std::vector<T> m, V0, E0;
  bool per_particle_E = false;
  // dense grid, index i*J*K + j*K + k (material_point_method.hpp:129-131)
  std::vector<T> gm;
  std::vector<V3<T>> gvel, gforce;
  std::vector<int64_t> used;
  std::vector<Collider<T>> colliders;
  int last_error = 0;

  void init(const Params& p, int np) {
    prm = p;
    I = p.dims[0];
    J = p.dims[1];
    K = p.dims[2];
    n = np;
    // Inputs are FP32 values in both builds, so the FP64 build differs from
    // the FP32 one only by arithmetic rounding.
    h = (T)(float)p.h;
    // DpInverse = inverse(mat3(1) * (1/3) * h * h)  (material_point_method.hpp:177)
    Dinv = minverse(mscale(mscale(mscale(mdiag(T(1)), T(1.0f / 3.0f)), h), h));
    // material_point_method.cpp:320: (float)(1 - 2.5f * 1e-2) is evaluated in double
    lo = (T)(float)(1.0 - p.theta_c);
    hi = (T)(float)(1.0 + p.theta_s);
    x.assign(n, {0, 0, 0});
    v.assign(n, {0, 0, 0});
    B.assign(n, mdiag(T(0)));
    FE.assign(n, mdiag(T(1)));
    FP.assign(n, mdiag(T(1)));
    m.assign(n, 0);
    V0.assign(n, 0);
    E0.assign(n, (T)(float)p.E);
    const size_t cells = (size_t)I * J * K;
    gm.assign(cells, 0);
    gvel.assign(cells, {0, 0, 0});
    gforce.assign(cells, {0, 0, 0});
  }
  inline int64_t gidx(int i, int j, int k) const { return ((int64_t)i * J + j) * K + k; }

  // stencil base (material_point_method.cpp:83): ivec3(pos / h)
  inline void base_of(const V3<T>& p, int b[3]) const {
    b[0] = (int)(p.x / h);
    b[1] = (int)(p.y / h);
    b[2] = (int)(p.z / h);
  }
  // 1-D weights of the 5-node stencil base-2..base+2 (wipHost per axis)
  inline void weights1d(const V3<T>& p, const int b[3], T w[3][5]) const {
    const T pc[3] = {p.x, p.y, p.z};
    for (int a = 0; a < 3; ++a)
      for (int d = 0; d < 5; ++d) {
        const int idx = b[a] + d - 2;
        const T comp = pc[a] / h - (T)idx;
        w[a][d] = weightN(comp);
      }
  }
  bool stencil_in_bounds(const int b[3]) const {
    return b[0] - 2 >= 0 && b[1] - 2 >= 0 && b[2] - 2 >= 0 && b[0] + 2 < I && b[1] + 2 < J && b[2] + 2 < K;
  }

  template <typename F> void for_particles(F&& f) {
#ifdef _OPENMP
    if (prm.threads > 1) {
#pragma omp parallel for schedule(static) num_threads(prm.threads)
      for (int p = 0; p < n; ++p) f(p);
      return;
    }
#endif
    for (int p = 0; p < n; ++p) f(p);
  }
  static inline void atomic_add(T& a, T b, bool par) {
#ifdef _OPENMP
    if (par) {
#pragma omp atomic
      a += b;
      return;
    }
#endif
    a += b;
  }

  // a2: material_point_method.cpp:94-129
  int p2g() {
    std::fill(gm.begin(), gm.end(), T(0));
    std::fill(gvel.begin(), gvel.end(), V3<T>{0, 0, 0});
    std::fill(gforce.begin(), gforce.end(), V3<T>{0, 0, 0});
    for (int p = 0; p < n; ++p) {
      int b[3];
      base_of(x[p], b);
      if (!stencil_in_bounds(b)) return last_error = 2;
    }
    const bool par = prm.threads > 1;
    for_particles([&](int p) {
      int b[3];
      base_of(x[p], b);
      T w[3][5];
      weights1d(x[p], b, w);
      for (int dx = 0; dx < 5; ++dx)
        for (int dy = 0; dy < 5; ++dy)
          for (int dz = 0; dz < 5; ++dz) {
            const T wip = w[0][dx] * w[1][dy] * w[2][dz];
            atomic_add(gm[gidx(b[0] + dx - 2, b[1] + dy - 2, b[2] + dz - 2)], m[p] * wip, par);
          }
    });
    used.clear();
    for (int i = 0; i < I; ++i)
      for (int j = 0; j < J; ++j)
        for (int k = 0; k < K; ++k)
          if (gm[gidx(i, j, k)] != T(0)) used.push_back(gidx(i, j, k));
    for_particles([&](int p) {
      int b[3];
      base_of(x[p], b);
      T w[3][5];
      weights1d(x[p], b, w);
      const M3<T> BD = mmul(B[p], Dinv);
      for (int dx = 0; dx < 5; ++dx)
        for (int dy = 0; dy < 5; ++dy)
          for (int dz = 0; dz < 5; ++dz) {
            const int ii = b[0] + dx - 2, jj = b[1] + dy - 2, kk = b[2] + dz - 2;
            const V3<T> Xi = {(T)ii * h, (T)jj * h, (T)kk * h};
            const T wip = w[0][dx] * w[1][dy] * w[2][dz];
            const V3<T> add = vmul(vadd(v[p], mvec(BD, vsub(Xi, x[p]))), wip * m[p]);
            V3<T>& g = gvel[gidx(ii, jj, kk)];
            atomic_add(g.x, add.x, par);
            atomic_add(g.y, add.y, par);
            atomic_add(g.z, add.z, par);
          }
    });
    for (const int64_t c : used) gvel[c] = vdiv(gvel[c], gm[c]);
    return 0;
  }

  // a3: material_point_method.cpp:131-142
  void init_volumes() {
    for_particles([&](int p) {
      int b[3];
      base_of(x[p], b);
      T w[3][5];
      weights1d(x[p], b, w);
      T density = T(0);
      for (int dx = 0; dx < 5; ++dx)
        for (int dy = 0; dy < 5; ++dy)
          for (int dz = 0; dz < 5; ++dz) {
            const T wip = w[0][dx] * w[1][dy] * w[2][dz];
            density += gm[gidx(b[0] + dx - 2, b[1] + dy - 2, b[2] + dz - 2)] * wip;
          }
      density /= (h * h * h);
      V0[p] = density != T(0) ? (m[p] / density) : T(0);
    });
  }

  // per-particle stress matrix V0 * Dinv * P(F_E, J_P) * F_E^T
  // (material_point_method.cpp:236-248,251)
  M3<T> stress_matrix(int p) const {
    const T poisson = (T)(float)prm.nu;
    const T E = E0[p];
    const T xi = (T)(float)prm.xi;
    const T mu0 = E / (T(2.0f) * (1 + poisson));
    const T mu = mu0 * std::exp(xi * (1 - mdet(FP[p])));
    const T lambda0 = (E * poisson) / ((1 + poisson) * (1 - 2 * poisson));
    const T lambda = lambda0 * std::exp(xi * (1 - mdet(FP[p])));
    const M3<T>& F = FE[p];
    const M3<T> R = polar_R(F);
    const T Jd = mdet(F);
    const M3<T> FT = mtrans(minverse(F));
    const M3<T> P = madd(mscale(msub(F, R), T(2) * mu), mscale(FT, lambda * (Jd - 1) * Jd));
    return mmul(mmul(mscale(Dinv, V0[p]), P), mtrans(F));
  }

  // a4: material_point_method.cpp:235-254
  void forces() {
    const bool par = prm.threads > 1;
    for_particles([&](int p) {
      const M3<T> A = stress_matrix(p);
      int b[3];
      base_of(x[p], b);
      T w[3][5];
      weights1d(x[p], b, w);
      for (int dx = 0; dx < 5; ++dx)
        for (int dy = 0; dy < 5; ++dy)
          for (int dz = 0; dz < 5; ++dz) {
            const int ii = b[0] + dx - 2, jj = b[1] + dy - 2, kk = b[2] + dz - 2;
            const V3<T> Xi = {(T)ii * h, (T)jj * h, (T)kk * h};
            const T wip = w[0][dx] * w[1][dy] * w[2][dz];
            const V3<T> f = mvec(mscale(A, wip), vsub(Xi, x[p]));
            V3<T>& g = gforce[gidx(ii, jj, kk)];
            atomic_add(g.x, -f.x, par);
            atomic_add(g.y, -f.y, par);
            atomic_add(g.z, -f.z, par);
          }
    });
  }

  // a5: material_point_method.cpp:256-262
  void grid_update(T dt) {
    const V3<T> g = {(T)(float)prm.g[0], (T)(float)prm.g[1], (T)(float)prm.g[2]};
    for (const int64_t c : used) gvel[c] = vadd(gvel[c], vmul(vadd(vdiv(gforce[c], gm[c]), g), dt));
  }

  // a6: material_point_method.cpp:298-304 (+ the configs' domain walls)
  void collisions(T /*dt*/) {
    for (const int64_t c : used) {
      const int i = (int)(c / ((int64_t)J * K)), j = (int)((c / K) % J), k = (int)(c % K);
      if ((prm.bc_mode & BC_COLLIDERS) && !colliders.empty()) {
        const V3<T> pos = {(T)i * h, (T)j * h, (T)k * h};
        gvel[c] = body_collision(pos, gvel[c], colliders);
      }
      if (prm.bc_mode & BC_DOMAIN) gvel[c] = domain_walls(i, j, k, I, J, K, gvel[c]);
    }
  }

  // a7: material_point_method.cpp:306-330.  The reference leaves the loop
  // at the first particle whose SVD fails (313-316), so particles after it
  // keep their old F.  With threads, every particle is updated into a copy
  // in parallel and only the particles before the first failure commit.
  int deformation_update(T dt) {
    if (prm.threads <= 1) {
      for (int p = 0; p < n; ++p)
        if (f_update(prm.compose, dt, Dinv, lo, hi, B[p], FE[p], FP[p])) return last_error = 3;
      return 0;
    }
    std::vector<M3<T>> fe(FE), fp(FP);
    int first_bad = n;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(prm.threads) reduction(min : first_bad)
#endif
    for (int p = 0; p < n; ++p)
      if (f_update(prm.compose, dt, Dinv, lo, hi, B[p], fe[p], fp[p])) first_bad = std::min(first_bad, p);
    std::copy(fe.begin(), fe.begin() + first_bad, FE.begin());
    std::copy(fp.begin(), fp.begin() + first_bad, FP.begin());
    return first_bad < n ? (last_error = 3) : 0;
  }

  // a8: material_point_method.cpp:332-342
  void g2p() {
    for_particles([&](int p) {
      int b[3];
      base_of(x[p], b);
      T w[3][5];
      weights1d(x[p], b, w);
      V3<T> vp = {0, 0, 0};
      M3<T> Bp = mdiag(T(0));
      for (int dx = 0; dx < 5; ++dx)
        for (int dy = 0; dy < 5; ++dy)
          for (int dz = 0; dz < 5; ++dz) {
            const int ii = b[0] + dx - 2, jj = b[1] + dy - 2, kk = b[2] + dz - 2;
            const T wip = w[0][dx] * w[1][dy] * w[2][dz];
            const V3<T> gv = gvel[gidx(ii, jj, kk)];
            vp = vadd(vp, vmul(gv, wip));
            const V3<T> Xi = {(T)ii * h, (T)jj * h, (T)kk * h};
            Bp = madd(Bp, mscale(mouter(gv, vsub(Xi, x[p])), wip));
          }
      v[p] = vp;
      B[p] = Bp;
    });
  }

  // a9: material_point_method.cpp:344-350,381-388
  void advect(T dt) {
    const T lo_m = (T)(3 * h);
    const T hx = (T)((I - 3) * h), hy = (T)((J - 3) * h), hz = (T)((K - 3) * h);
    for_particles([&](int p) {
      V3<T> q = vadd(x[p], vmul(v[p], dt));
      q.x = std::clamp(q.x, lo_m, hx);
      q.y = std::clamp(q.y, lo_m, hy);
      q.z = std::clamp(q.z, lo_m, hz);
      x[p] = q;
    });
  }

  // One substep in the reference's order (main.cpp:192-218)
  int substep(T dt) {
    if (p2g()) return last_error;
    forces();
    grid_update(dt);
    collisions(dt);
    if (deformation_update(dt)) return last_error;
    g2p();
    advect(dt);
    return 0;
  }
};

}  // namespace mpmo
