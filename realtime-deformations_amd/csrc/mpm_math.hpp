// Device math for the MPM snow substep on gfx950.
//
// Matrices are glm-style column-major (c[col][row]) and every operation keeps
// glm 0.9.7.1's evaluation order, so that with contraction disabled (this
// library is built with -ffp-contract=off; hot loops use explicit fmaf) the
// deformation-gradient update is bit-identical to the reference's FP32 CPU
// path on identical inputs:
//   realtime-deformations/material_point_method.cpp:306-330 (F update)
//   realtime-deformations/utils.h:15-33 (the glm<->Eigen transposes)
//   Eigen/src/SVD/JacobiSVD.h:689-819, Eigen/src/misc/RealSvd2x2.h:21-51,
//   Eigen/src/Jacobi/Jacobi.h:96-127,297-345 (two-sided Jacobi conventions)
// All indices into the small register arrays are compile-time constants
// (templated pair rotations, predicated swaps), so nothing spills to scratch.
//
// The header also compiles as plain host C++ (MPM_HD expands to `inline`),
// so the same source of the weights, the divisions and the F update can be
// checked exhaustively on the CPU against the oracle.
#pragma once
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MPM_HD __host__ __device__ __forceinline__
#define MPM_HDC __host__ __device__
#define MPM_HD_NOINLINE inline __host__ __device__ __noinline__
#else
#include <cmath>
#define MPM_HD inline
#define MPM_HDC
#define MPM_HD_NOINLINE inline
#endif

#include <cstdint>

namespace mpm {

struct F3 {
  float x, y, z;
};
template <typename T> struct M3T {
  T c[3][3];
};
using M3 = M3T<float>;

MPM_HD F3 f3(float a, float b, float c) { return {a, b, c}; }
MPM_HD F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
MPM_HD F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
MPM_HD F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
MPM_HD float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

template <typename T> MPM_HD M3T<T> mdiag(T d) {
  M3T<T> r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.c[i][j] = (i == j) ? d : T(0);
  return r;
}
template <typename T> MPM_HD M3T<T> madd(const M3T<T>& a, const M3T<T>& b) {
  M3T<T> r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.c[i][j] = a.c[i][j] + b.c[i][j];
  return r;
}
template <typename T> MPM_HD M3T<T> msub(const M3T<T>& a, const M3T<T>& b) {
  M3T<T> r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.c[i][j] = a.c[i][j] - b.c[i][j];
  return r;
}
template <typename T> MPM_HD M3T<T> mscale(const M3T<T>& a, T s) {
  M3T<T> r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.c[i][j] = a.c[i][j] * s;
  return r;
}
// glm mat3*mat3: R[j][i] = A[0][i]*B[j][0] + A[1][i]*B[j][1] + A[2][i]*B[j][2]
template <typename T> MPM_HD M3T<T> mmul(const M3T<T>& a, const M3T<T>& b) {
  M3T<T> r;
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i) r.c[j][i] = a.c[0][i] * b.c[j][0] + a.c[1][i] * b.c[j][1] + a.c[2][i] * b.c[j][2];
  return r;
}
MPM_HD F3 mvec(const M3& m, F3 v) {
  return {m.c[0][0] * v.x + m.c[1][0] * v.y + m.c[2][0] * v.z, m.c[0][1] * v.x + m.c[1][1] * v.y + m.c[2][1] * v.z,
          m.c[0][2] * v.x + m.c[1][2] * v.y + m.c[2][2] * v.z};
}
template <typename T> MPM_HD M3T<T> mtrans(const M3T<T>& a) {
  M3T<T> r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.c[i][j] = a.c[j][i];
  return r;
}
template <typename T> MPM_HD T mdet(const M3T<T>& m) {
  return +m.c[0][0] * (m.c[1][1] * m.c[2][2] - m.c[2][1] * m.c[1][2]) -
         m.c[1][0] * (m.c[0][1] * m.c[2][2] - m.c[2][1] * m.c[0][2]) +
         m.c[2][0] * (m.c[0][1] * m.c[1][2] - m.c[1][1] * m.c[0][2]);
}
// glm compute_inverse (cofactors times one division)
template <typename T> MPM_HD M3T<T> minverse(const M3T<T>& m) {
  const T ood = T(1) / (+m.c[0][0] * (m.c[1][1] * m.c[2][2] - m.c[2][1] * m.c[1][2]) -
                        m.c[1][0] * (m.c[0][1] * m.c[2][2] - m.c[2][1] * m.c[0][2]) +
                        m.c[2][0] * (m.c[0][1] * m.c[1][2] - m.c[1][1] * m.c[0][2]));
  M3T<T> r;
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

// ------------------------------------------- fast reciprocal and powers
// For the stress path only, which feeds the P2G sums (compared to the
// oracle at 1e-5) and never the bit-exact F update: the hardware's
// reciprocal (1 ulp) and base-2 log and exponential on the device, the
// plain expressions in the host build.
MPM_HD float fast_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}
// |x|^(p/3) for p = +-1, x > 0 normal: the polar iteration's scale factor
MPM_HD float fast_cbrt_pow(float x, float p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_exp2f(p * (1.0f / 3.0f) * __builtin_amdgcn_logf(x));
#else
  return exp2f(p * (1.0f / 3.0f) * log2f(x));
#endif
}
// glm's cofactor inverse with the hardware reciprocal of the determinant
MPM_HD M3 minverse_fast(const M3& m) {
  const float ood = fast_rcp(+m.c[0][0] * (m.c[1][1] * m.c[2][2] - m.c[2][1] * m.c[1][2]) -
                             m.c[1][0] * (m.c[0][1] * m.c[2][2] - m.c[2][1] * m.c[0][2]) +
                             m.c[2][0] * (m.c[0][1] * m.c[1][2] - m.c[1][1] * m.c[0][2]));
  M3 r;
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

// ------------------------------------------------- division by a constant
// a / b correctly rounded, given rb = 1/b correctly rounded (computed once on
// the host): q = a*rb is within one ulp of a/b, the residual a - b*q is exact
// under a fused multiply-add, and q + r*rb rounds to the correctly rounded
// quotient (Markstein's correction; the quotients here are normal numbers).
MPM_HD float div_rcp(float a, float b, float rb) {
  const float q = a * rb;
  const float r = fmaf(-q, b, a);
  return fmaf(r, rb, q);
}

// glm's inverse of d_in * identity (compute_inverse, cofactors times one
// division) for d_in > 0: the diagonal is d = d_in^2 / d_in^3 and the
// cofactor signs leave these signed zeros off the diagonal.  Built from the
// one scalar, the matrix costs no registers.
MPM_HD M3 dinv_matrix(float d) {
  M3 r;
  r.c[0][0] = d;
  r.c[1][0] = -0.0f;
  r.c[2][0] = 0.0f;
  r.c[0][1] = -0.0f;
  r.c[1][1] = d;
  r.c[2][1] = -0.0f;
  r.c[0][2] = 0.0f;
  r.c[1][2] = -0.0f;
  r.c[2][2] = d;
  return r;
}

// B D^-1 for D^-1 = dinv_matrix(d): glm's product adds B's other columns
// times the signed off-diagonal zeros, which changes no value, only possibly
// the sign of a zero entry, and every use adds the identity or another
// term first; so it is B's columns scaled by d (9 products instead of 45
// operations).  Non-finite B stays non-finite either way.
MPM_HD M3 mul_dinv(const M3& B, const M3& Dinv) { return mscale(B, Dinv.c[0][0]); }

// -------------------------------------------------------- cubic B-spline
// N(u) of material_point_method.hpp:20-31.  The reference's double literals
// promote the polynomial to double before the final rounding; FP64 on gfx950
// runs at the FP32 rate, so the weights are kept bit-exact.
MPM_HD float weight_inner(float modx) {  // |u| < 1
  const float modx2 = modx * modx;
  const float modx3 = modx * modx * modx;
  return (float)(0.5 * (double)modx3 - (double)modx2 + 2.0 / 3.0);
}
MPM_HD float weight_outer(float modx) {  // 1 <= |u| < 2
  const float t = 2.0f - modx;
  return (float)((1.0 / 6.0) * (double)t * (double)t * (double)t);
}
MPM_HD float weightN(float u) {
  const float modx = fabsf(u);
  if (modx < 1.0f) return weight_inner(modx);
  if (modx < 2.0f) return weight_outer(modx);
  return 0.0f;
}

// Stencil base and the four non-zero 1-D weights (nodes base-1 .. base+2) of
// one coordinate; q = x / h is the correctly rounded FP32 quotient.  With
// b = int(q) and q >= 2 every offset u = q - node is exact and its branch of
// N is known: u0 in [1,2), u1 in [0,1), |u2| in (0,1], |u3| in (1,2].  The
// two edge values |u2| = 1 and |u3| = 2 (q an integer) give the same float
// from either branch (1/6 and 0), so the branch-free form equals weightN.
// Positions are kept in [3h, (dims-3)h], where the clamp of b never binds.
struct Axis {
  int base;
  float w[4];
};
MPM_HD Axis axis_weights(float x, float h, float inv_h, int lo, int hi) {
  Axis a;
  const float q = div_rcp(x, h, inv_h);
  int b = (int)q;
  b = b < lo ? lo : (b > hi ? hi : b);  // keeps every access in bounds
  a.base = b;
  const float fb = (float)b;
  a.w[0] = weight_outer(q - (fb - 1.0f));
  a.w[1] = weight_inner(q - fb);
  a.w[2] = weight_inner(fabsf(q - (fb + 1.0f)));
  a.w[3] = weight_outer(fabsf(q - (fb + 2.0f)));
  return a;
}

// dN/du of the cubic B-spline
MPM_HD float dweightN(float u) {
  const float a = fabsf(u);
  if (a < 1.0f) return u * (1.5f * a - 2.0f);
  if (a < 2.0f) {
    const float t = 2.0f - a;
    return (u < 0.0f ? 0.5f : -0.5f) * t * t;
  }
  return 0.0f;
}
// Stencil weights with their derivatives d/dx (reverse mode).
struct AxisD {
  int base;
  float w[4], dw[4];
};
MPM_HD AxisD axis_weights_d(float x, float h, float inv_h, int lo, int hi) {
  const Axis a = axis_weights(x, h, inv_h, lo, hi);
  AxisD r;
  r.base = a.base;
  const float q = div_rcp(x, h, inv_h);
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    r.w[d] = a.w[d];
    r.dw[d] = dweightN(q - (float)(a.base - 1 + d)) * inv_h;
  }
  return r;
}

// ------------------------------------------------ Eigen-convention Jacobi
template <typename T> struct Rot {
  T c, s;
};
template <typename T> struct NumLim;
template <> struct NumLim<float> {
  MPM_HDC static constexpr float min() { return 1.17549435e-38f; }
  MPM_HDC static constexpr float eps() { return 1.1920929e-07f; }
};
template <> struct NumLim<double> {
  MPM_HDC static constexpr double min() { return 2.2250738585072014e-308; }
  MPM_HDC static constexpr double eps() { return 2.220446049250313e-16; }
};
template <typename T> MPM_HD T tabs(T x) { return x < T(0) ? -x : (x == T(0) ? T(0) : x); }
MPM_HD float tsqrt(float x) { return sqrtf(x); }
MPM_HD double tsqrt(double x) { return sqrt(x); }

// real_2x2_jacobi_svd on the (p,q) block of W (row-major math W[r][c]).
template <typename T, int P, int Q>
MPM_HD void jacobi_2x2(const T (&W)[3][3], Rot<T>& jl, Rot<T>& jr) {
  T m00 = W[P][P], m01 = W[P][Q], m10 = W[Q][P], m11 = W[Q][Q];
  Rot<T> r1;
  const T t = m00 + m11;
  const T d = m10 - m01;
  if (tabs(d) < NumLim<T>::min()) {
    r1.s = T(0);
    r1.c = T(1);
  } else {
    const T u = t / d;
    const T tmp = tsqrt(T(1) + u * u);
    r1.s = T(1) / tmp;
    r1.c = u / tmp;
  }
  if (!(r1.c == T(1) && r1.s == T(0))) {
    const T a0 = m00, b0 = m10, a1 = m01, b1 = m11;
    m00 = r1.c * a0 + r1.s * b0;
    m10 = -r1.s * a0 + r1.c * b0;
    m01 = r1.c * a1 + r1.s * b1;
    m11 = -r1.s * a1 + r1.c * b1;
  }
  const T x = m00, y = m01, z = m11;
  const T deno = T(2) * tabs(y);
  if (deno < NumLim<T>::min()) {
    jr.c = T(1);
    jr.s = T(0);
  } else {
    const T tau = (x - z) / deno;
    const T w = tsqrt(tau * tau + T(1));
    const T tt = (tau > T(0)) ? T(1) / (tau + w) : T(1) / (tau - w);
    const T sign_t = tt > T(0) ? T(1) : T(-1);
    const T n = T(1) / tsqrt(tt * tt + T(1));
    const T sgn_y = y < T(0) ? T(-1) : T(1);  // y / |y|, y != 0 here
    jr.s = -sign_t * sgn_y * tabs(tt) * n;
    jr.c = n;
  }
  const T c2 = jr.c, s2 = -jr.s;
  jl.c = r1.c * c2 - r1.s * s2;
  jl.s = r1.c * s2 + r1.s * c2;
}

template <typename T, int P, int Q>
MPM_HD void jacobi_pair(T (&W)[3][3], T (&U)[3][3], T (&V)[3][3], T& maxDiag, bool& finished) {
  const T pm = T(2) * NumLim<T>::eps() * maxDiag;
  const T threshold = (NumLim<T>::min() < pm) ? pm : NumLim<T>::min();
  if (tabs(W[P][Q]) > threshold || tabs(W[Q][P]) > threshold) {
    finished = false;
    Rot<T> jl, jr;
    jacobi_2x2<T, P, Q>(W, jl, jr);
    if (!(jl.c == T(1) && jl.s == T(0))) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {  // rows P,Q of W
        const T xi = W[P][k], yi = W[Q][k];
        W[P][k] = jl.c * xi + jl.s * yi;
        W[Q][k] = -jl.s * xi + jl.c * yi;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {  // cols P,Q of U
        const T xi = U[k][P], yi = U[k][Q];
        U[k][P] = jl.c * xi + jl.s * yi;
        U[k][Q] = -jl.s * xi + jl.c * yi;
      }
    }
    const T rc = jr.c, rs = -jr.s;  // right rotation uses j_right^T
    if (!(rc == T(1) && rs == T(0))) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const T xi = W[k][P], yi = W[k][Q];
        W[k][P] = rc * xi + rs * yi;
        W[k][Q] = -rs * xi + rc * yi;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const T xi = V[k][P], yi = V[k][Q];
        V[k][P] = rc * xi + rs * yi;
        V[k][Q] = -rs * xi + rc * yi;
      }
    }
    const T a = tabs(W[P][P]), b = tabs(W[Q][Q]);
    const T mx = (a < b) ? b : a;
    maxDiag = (maxDiag < mx) ? mx : maxDiag;
  }
}

template <typename T> MPM_HD void swap_cols(T (&A)[3][3], int i, int j, bool doit) {
  // i, j are compile-time after unrolling by the callers
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const T a = A[r][i], b = A[r][j];
    A[r][i] = doit ? b : a;
    A[r][j] = doit ? a : b;
  }
}

// SVD of M (row-major math) = U diag(S) V^T with Eigen's conventions.
// Returns false for non-finite input (Eigen reports InvalidInput).
template <typename T>
MPM_HD bool jacobi_svd3(const T (&M)[3][3], T (&U)[3][3], T (&S)[3], T (&V)[3][3]) {
  T scale = T(0);
  bool bad = false;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T a = tabs(M[r][c]);
      bad |= !(a == a) || (a > T(1e308) && sizeof(T) == 8) || (sizeof(T) == 4 && a > T(3.4028235e38f));
      scale = (a > scale) ? a : scale;
    }
  if (bad) return false;
  if (scale == T(0)) scale = T(1);
  T W[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      W[r][c] = M[r][c] / scale;
      U[r][c] = V[r][c] = (r == c) ? T(1) : T(0);
    }
  T maxDiag = tabs(W[0][0]);
  {
    const T b = tabs(W[1][1]);
    maxDiag = (maxDiag < b) ? b : maxDiag;
    const T c = tabs(W[2][2]);
    maxDiag = (maxDiag < c) ? c : maxDiag;
  }
  bool finished = false;
  for (int sweep = 0; sweep < 48 && !finished; ++sweep) {
    finished = true;
    jacobi_pair<T, 1, 0>(W, U, V, maxDiag, finished);
    jacobi_pair<T, 2, 0>(W, U, V, maxDiag, finished);
    jacobi_pair<T, 2, 1>(W, U, V, maxDiag, finished);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const T a = W[i][i];
    S[i] = tabs(a);
    if (a < T(0)) {
#pragma unroll
      for (int r = 0; r < 3; ++r) U[r][i] = -U[r][i];
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) S[i] *= scale;
  // selection sort, descending, first maximum kept
  {
    int pos = 0;
    T mx = S[0];
    if (S[1] > mx) {
      mx = S[1];
      pos = 1;
    }
    if (S[2] > mx) {
      mx = S[2];
      pos = 2;
    }
    if (mx == T(0)) return true;
    const bool s1 = pos == 1, s2 = pos == 2;
    T t0 = S[0];
    S[0] = s1 ? S[1] : (s2 ? S[2] : S[0]);
    if (s1) S[1] = t0;
    if (s2) S[2] = t0;
    swap_cols(U, 0, 1, s1);
    swap_cols(V, 0, 1, s1);
    swap_cols(U, 0, 2, s2);
    swap_cols(V, 0, 2, s2);
  }
  {
    const T mx = (S[2] > S[1]) ? S[2] : S[1];
    if (mx == T(0)) return true;
    const bool s2 = S[2] > S[1];
    if (s2) {
      const T t = S[1];
      S[1] = S[2];
      S[2] = t;
    }
    swap_cols(U, 1, 2, s2);
    swap_cols(V, 1, 2, s2);
  }
  return true;
}

// ------------------------------------------------------------- polar factor
// Rotation factor R of F = R S by scaled Newton iteration
// X <- (g X + X^{-T} / g) / 2, g = |det X|^{-1/3}.  Like the reference's
// Higham–Noferini routine, the result is always a rotation (minus the
// orthogonal factor when det F < 0).  The scale g only speeds convergence
// (the iteration's fixed point is R for any g > 0), so it and its inverse
// come from the hardware's log2 / exp2, and X^{-T} from the hardware
// reciprocal of det X.
MPM_HD M3 polar_R(const M3& F) {
  M3 X = F;
  const float d0 = mdet(F);
  for (int it = 0; it < 12; ++it) {
    const float d = mdet(X);
    if (!(fabsf(d) >= 1.17549435e-38f)) break;  // zero, denormal (the log would be -inf) or NaN
    const M3 Yt = mtrans(minverse_fast(X));
    const float g = fast_cbrt_pow(fabsf(d), -1.0f);
    const float ig = fast_cbrt_pow(fabsf(d), 1.0f);
    float diff = 0.0f;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const float xn = 0.5f * (g * X.c[i][j] + Yt.c[i][j] * ig);
        diff = fmaxf(diff, fabsf(xn - X.c[i][j]));
        X.c[i][j] = xn;
      }
    if (diff <= 2.0e-7f) break;
  }
  if (d0 < 0.0f) X = mscale(X, -1.0f);
  return X;
}

// --------------------------------------------------- deformation update
enum { COMPOSE_LITERAL = 0, COMPOSE_TEXTBOOK = 1 };

// FP64 re-solve of the clamp for TEXTBOOK mode: the singular vectors of a
// nearly degenerate spectrum are ill-conditioned in FP32, so when the
// relative gap falls below `gap` the SVD and the recomposition run in FP64.
MPM_HD_NOINLINE void textbook_fp64(const M3& FEhat, double lo, double hi, M3& FE) {
  double M[3][3], U[3][3], S[3], V[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) M[i][j] = (double)FEhat.c[i][j];
  jacobi_svd3<double>(M, U, S, V);
#pragma unroll
  for (int i = 0; i < 3; ++i) S[i] = S[i] < lo ? lo : (S[i] > hi ? hi : S[i]);
  // M = FEhat^T = U S V^T  ->  FE = V S U^T ; glm FE.c[col][row] = FE(row, col)
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) acc += V[r][k] * S[k] * U[c][k];
      FE.c[c][r] = (float)acc;
    }
}

// material_point_method.cpp:306-330 for one particle; Bn is the APIC matrix
// of the previous G2P.  Returns false when the SVD rejects its input; the
// particle's F is then left unchanged and the caller raises a status flag.
// svd (optional): the 21 values of the SVD of F^T = U S V^T the update used
// (U row-major, S unclamped, V row-major), which the reverse mode reads back.
template <int COMPOSE>
MPM_HD bool f_update_t(float dt, const M3& Dinv, float lo, float hi, float gap64, const M3& Bn,
                                           M3& FE, M3& FP, float* svd = nullptr) {
  const int compose = COMPOSE;
  const M3 FPn1 = mmul(mmul(madd(mdiag(1.0f), mscale(mul_dinv(Bn, Dinv), dt)), FE), FP);
  const M3 FEhat = mmul(FPn1, minverse(FP));
  float M[3][3], U[3][3], S[3], V[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) M[i][j] = FEhat.c[i][j];
  if (!jacobi_svd3<float>(M, U, S, V)) return false;
  if (svd) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      svd[9 + i] = S[i];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        svd[3 * i + j] = U[i][j];
        svd[12 + 3 * i + j] = V[i][j];
      }
    }
  }
  M3 FEn;
  const bool fp64 = COMPOSE == COMPOSE_TEXTBOOK && gap64 > 0.0f &&
                    fminf(S[0] - S[1], S[1] - S[2]) < gap64 * S[0];
  if (COMPOSE == COMPOSE_TEXTBOOK && fp64) {
    textbook_fp64(FEhat, (double)lo, (double)hi, FEn);
  } else {
#pragma unroll
    for (int i = 0; i < 3; ++i) S[i] = fminf(fmaxf(S[i], lo), hi);  // std::clamp for finite values
    M3 Ug, Vg, Sg = mdiag(0.0f);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      Sg.c[i][i] = S[i];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        Ug.c[i][j] = U[i][j];
        Vg.c[i][j] = V[i][j];
      }
    }
    FEn = (compose == COMPOSE_LITERAL) ? mmul(mmul(Ug, Sg), mtrans(Vg)) : mmul(mmul(mtrans(Vg), Sg), Ug);
  }
  FE = FEn;
  FP = mmul(minverse(FEn), FPn1);
  return true;
}
MPM_HD bool f_update(int compose, float dt, const M3& Dinv, float lo, float hi, float gap64,
                                         const M3& Bn, M3& FE, M3& FP) {
  return compose == COMPOSE_LITERAL ? f_update_t<COMPOSE_LITERAL>(dt, Dinv, lo, hi, gap64, Bn, FE, FP)
                                    : f_update_t<COMPOSE_TEXTBOOK>(dt, Dinv, lo, hi, gap64, Bn, FE, FP);
}

// ------------------------------------------------------------ snow stress
// V0 * Dinv * P(F_E, J_P) * F_E^T with P = 2 mu (F - R) + lambda (J-1) J F^-T,
// mu, lambda hardened by exp(xi (1 - det F_P))
// (material_point_method.cpp:236-251).
// Lame denominators 2(1+nu) and (1+nu)(1-2nu) with their correctly rounded
// reciprocals, computed once on the host.
struct Lame {
  float nu, xi, dmu, rmu, dlam, rlam;
};
// `stiff` receives the hardened P-wave modulus lambda + 2 mu, which sets
// the elastic wave speed c^2 = (lambda + 2 mu) / rho of the stability check.
MPM_HD M3 stress_matrix(float E, const Lame& L, float V0, const M3& Dinv, const M3& FE, const M3& FP,
                        float& stiff) {
  const float nu = L.nu, xi = L.xi;
  const float mu0 = div_rcp(E, L.dmu, L.rmu);  // E / (2 (1 + nu))
  const float jp = mdet(FP);
  const float hard = expf(xi * (1.0f - jp));
  const float mu = mu0 * hard;
  const float lambda0 = div_rcp(E * nu, L.dlam, L.rlam);  // E nu / ((1 + nu)(1 - 2 nu))
  const float lambda = lambda0 * hard;
  stiff = fmaf(2.0f, mu, lambda);
  const M3 R = polar_R(FE);
  const float J = mdet(FE);
  const M3 FT = mtrans(minverse_fast(FE));
  const M3 P = madd(mscale(msub(FE, R), 2.0f * mu), mscale(FT, lambda * (J - 1.0f) * J));
  // (D^-1 V0) P: the diagonal scaling of glm's product (see mul_dinv)
  return mmul(mscale(P, Dinv.c[0][0] * V0), mtrans(FE));
}
MPM_HD M3 stress_matrix(float E, const Lame& L, float V0, const M3& Dinv, const M3& FE, const M3& FP) {
  float stiff;
  return stress_matrix(E, L, V0, Dinv, FE, FP, stiff);
}

// Explicit-integration stability of one particle for one substep:
// |v| dt <= h (it moves at most one cell) and c dt <= 2h with
// c^2 = (lambda + 2 mu) V0 / m (the elastic wave crosses at most two cells;
// the margin covers surface particles, whose initial volumes come from a
// grid density that their missing neighbours make low, so m / V0
// underestimates their density by up to a few times).
// Returns 0 when both hold, ST-style bits otherwise: 1 when a quantity is not
// finite, 4 when a bound is exceeded by finite values.
MPM_HD int stability_flags(F3 v, float stiff, float V0, float m, float dt, float h) {
  const float dt2 = dt * dt, h2 = h * h;
  const float vv = v.x * v.x + v.y * v.y + v.z * v.z;
  const float cc = stiff * V0;
  if (vv * dt2 <= h2 && cc * dt2 <= 4.0f * m * h2) return 0;
  // x - x is 0 exactly for finite x, NaN for NaN and infinities
  const float z = (v.x - v.x) + (v.y - v.y) + (v.z - v.z) + (stiff - stiff) + (V0 - V0) + (m - m);
  return z == 0.0f ? 4 : 1;
}

// ---------------------------------------------------------------- colliders
// Oriented box with the inverse pose matrix precomputed on the host in glm's
// operation order; SDF per material_point_method.hpp:79-86.
struct DevCollider {
  float inv[4][4];  // glm column-major
  float s[3];       // half extents
  float vel[3];
};
MPM_HD float box_sdf(const DevCollider& c, F3 p) {
  float q[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const float pr = (c.inv[0][r] * p.x + c.inv[1][r] * p.y) + (c.inv[2][r] * p.z + c.inv[3][r] * 1.0f);
    q[r] = fabsf(pr) - c.s[r];
  }
  float mx = q[0];
  mx = (mx < q[1]) ? q[1] : mx;
  mx = (mx < q[2]) ? q[2] : mx;
  float mx0 = (mx < 0.0f) ? 0.0f : mx;
  const float inner = fmaxf(q[0], fmaxf(q[1], q[2]));
  return fabsf(mx0) + fminf(inner, 0.0f);
}
// mathy.hpp:40-56 and material_point_method.cpp:264-296
MPM_HD F3 body_collision(const DevCollider* cs, int k, F3 pos, F3 vel) {
  bool all_out = true;
  for (int i = 0; i < k; ++i)
    if (!(box_sdf(cs[i], pos) > 0.0f)) all_out = false;
  if (all_out) return vel;
  F3 out = vel;
  const float delta = 0.001f;
  for (int i = 0; i < k; ++i) {
    const DevCollider& c = cs[i];
    if (box_sdf(c, pos) > 0.0f) continue;
    float g[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const F3 d = f3(a == 0 ? delta : 0.0f, a == 1 ? delta : 0.0f, a == 2 ? delta : 0.0f);
      const float f1 = box_sdf(c, pos - d);
      const float f2 = box_sdf(c, pos + d);
      g[a] = (f2 - f1) / (2.0f * delta);
    }
    const F3 n = f3(g[0], g[1], g[2]);
    const F3 vobj = f3(c.vel[0], c.vel[1], c.vel[2]);
    const F3 rel = out - vobj;
    const float vn = dot(rel, n);
    if (vn >= 0.0f) continue;
    const F3 vt = rel - n * vn;
    const float mu = 0.5f;
    F3 vrel = f3(0.0f, 0.0f, 0.0f);
    if (3.0f > -mu * vn) vrel = vt + vt * (mu * vn / 3.0f);  // glm vec3::length() == 3
    out = vrel + vobj;
  }
  return out;
}
// Configs' domain walls: floor band sticky, other bands separating.
MPM_HD F3 domain_walls(int i, int j, int k, int I, int J, int K, F3 v) {
  const int b = 3;
  if (j < b) return f3(0.0f, 0.0f, 0.0f);
  if (i < b && v.x < 0.0f) v.x = 0.0f;
  if (i >= I - b && v.x > 0.0f) v.x = 0.0f;
  if (j >= J - b && v.y > 0.0f) v.y = 0.0f;
  if (k < b && v.z < 0.0f) v.z = 0.0f;
  if (k >= K - b && v.z > 0.0f) v.z = 0.0f;
  return v;
}

}  // namespace mpm
