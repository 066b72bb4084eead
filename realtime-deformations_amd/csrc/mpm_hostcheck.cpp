// Host build of the device math in mpm_math.hpp (plain C++, same source), so
// that CPU tests can hold the weights, the constant divisions, the Jacobi SVD
// and the deformation-gradient update to the oracle bit for bit.  Built as
// lib/libmpm_hostcheck.so; not part of libmpm_hip.
#include <algorithm>
#include <cstring>

#include "mpm_hostutil.hpp"
#include "mpm_math.hpp"

using namespace mpm;

namespace {
M3 load9(const float* p) {
  M3 m;
  for (int i = 0; i < 9; ++i) m.c[i / 3][i % 3] = p[i];
  return m;
}
void store9(const M3& m, float* p) {
  for (int i = 0; i < 9; ++i) p[i] = m.c[i / 3][i % 3];
}
// inverse(mat3(1) * (1/3) * h * h) in glm's operation order
M3 dp_inverse(float h) { return minverse(mscale(mscale(mscale(mdiag(1.0f), 1.0f / 3.0f), h), h)); }
}  // namespace

extern "C" {

// 1 when glm's inverse of diag(h^2/3) (the operation order of the context's
// set-up) equals dinv_matrix of its diagonal bit for bit, zero signs included
int hc_dinv_pattern(float h) {
  const M3 full = dp_inverse(h);
  const M3 fast = dinv_matrix(full.c[0][0]);
  return std::memcmp(&full, &fast, sizeof(M3)) == 0 ? 1 : 0;
}

// base and the four 1-D weights of every coordinate x[i]
void hc_axis_weights(const float* x, int n, float h, int lo, int hi, int* base, float* w) {
  const float inv_h = 1.0f / h;
  for (int i = 0; i < n; ++i) {
    const Axis a = axis_weights(x[i], h, inv_h, lo, hi);
    base[i] = a.base;
    for (int d = 0; d < 4; ++d) w[4 * i + d] = a.w[d];
  }
}

void hc_weightN(const float* u, int n, float* out) {
  for (int i = 0; i < n; ++i) out[i] = weightN(u[i]);
}

// a[i] / b through div_rcp with the correctly rounded reciprocal of b
void hc_div_rcp(const float* a, int n, float b, float* out) {
  const float rb = 1.0f / b;
  for (int i = 0; i < n; ++i) out[i] = div_rcp(a[i], b, rb);
}

// M, U, V math row-major (coefficient (r, c) at 3r + c), as the oracle's KAT
int hc_svd3(const float* M, float* U, float* S, float* V) {
  float m[3][3], u[3][3], v[3][3], s[3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) m[r][c] = M[3 * r + c];
  const bool ok = jacobi_svd3<float>(m, u, s, v);
  for (int i = 0; i < 3; ++i) S[i] = s[i];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      U[3 * r + c] = u[r][c];
      V[3 * r + c] = v[r][c];
    }
  return ok ? 0 : 1;
}

// glm column-major matrices; returns 0, or 1 when the SVD rejects its input
int hc_fupdate(int compose, float dt, float h, double theta_c, double theta_s, const float* B, const float* FE,
               const float* FP, float* FEo, float* FPo) {
  M3 fe = load9(FE), fp = load9(FP);
  const bool ok = f_update(compose, dt, dp_inverse(h), (float)(1.0 - theta_c), (float)(1.0 + theta_s), 0.0f,
                           load9(B), fe, fp);
  store9(fe, FEo);
  store9(fp, FPo);
  return ok ? 0 : 1;
}

// hc_fupdate with the TEXTBOOK FP64 re-solve threshold (svd_gap_fp64)
int hc_fupdate_gap(int compose, float dt, float h, double theta_c, double theta_s, float gap64, const float* B,
                   const float* FE, const float* FP, float* FEo, float* FPo) {
  M3 fe = load9(FE), fp = load9(FP);
  const bool ok = f_update(compose, dt, dp_inverse(h), (float)(1.0 - theta_c), (float)(1.0 + theta_s), gap64,
                           load9(B), fe, fp);
  store9(fe, FEo);
  store9(fp, FPo);
  return ok ? 0 : 1;
}

// The fused pass's per-particle stability check (0, 1 non-finite, 4 CFL).
int hc_stability_flags(const float* v, float stiff, float V0, float m, float dt, float h) {
  return stability_flags(f3(v[0], v[1], v[2]), stiff, V0, m, dt, h);
}

// The byte ranges mpm_get_positions hands to nt copy threads for a read of
// `bytes` bytes: begin[t], end[t] for t < nt.
void hc_readback_parts(size_t bytes, unsigned nt, size_t* begin, size_t* end) {
  const size_t part = readback_part_bytes(bytes, nt);
  for (unsigned t = 0; t < nt; ++t) {
    begin[t] = std::min(bytes, part * t);
    end[t] = t + 1 == nt ? bytes : std::min(bytes, part * (t + 1));
  }
}

}  // extern "C"
