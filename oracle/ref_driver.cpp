// ORACLE pinning driver — test infrastructure only.
//
// Built only in a container that holds the reference tree, by oracle/Makefile,
// into oracle/_ref/libref_math.so (never committed, never part of the
// product).  It compiles the reference's OWN vendored, header-only math —
// Eigen 3.4.90 JacobiSVD, glm 0.9.7.1 and the Higham–Noferini polar
// decomposition under realtime-deformations/include — and exposes them to
// tests/golden/make_golden.py, which records their outputs as fixtures.
//
// The reference's material_point_method.cpp itself cannot be built here: it
// includes cuda_runtime_api.h and links cudaMalloc/cudaMemcpyToSymbol and the
// nvcc-compiled cuP2G (cudaCalc.cu), none of which the image provides.  So the
// F-update composition below calls glm and Eigen in the order of
// realtime-deformations/material_point_method.cpp:308-327, with the two
// transposing conversions of realtime-deformations/utils.h:15-33.
#include <Eigen/Dense>
#include <algorithm>
#include <glm/glm.hpp>
#include <glm/gtc/matrix_transform.hpp>
#include <glm/gtx/quaternion.hpp>

#include "polar_decomposition_3x3.h"

namespace {
glm::mat3 load_glm(const float* a) {  // column-major: a[3c + r] = m[c][r]
  glm::mat3 m;
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) m[c][r] = a[3 * c + r];
  return m;
}
void store_glm(const glm::mat3& m, float* a) {
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) a[3 * c + r] = m[c][r];
}
}  // namespace

extern "C" {

// Eigen JacobiSVD<MatrixXf, ComputeFullU|ComputeFullV>, as called at
// material_point_method.cpp:312.  M, U, V are row-major math matrices.
int ref_svd3_f(const float* M, float* U, float* S, float* V) {
  Eigen::MatrixXf m(3, 3);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) m(r, c) = M[3 * r + c];
  Eigen::JacobiSVD<Eigen::MatrixXf, Eigen::ComputeFullU | Eigen::ComputeFullV> svd(m);
  if (svd.info() != Eigen::ComputationInfo::Success) return 1;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      U[3 * r + c] = svd.matrixU()(r, c);
      V[3 * r + c] = svd.matrixV()(r, c);
    }
  for (int i = 0; i < 3; ++i) S[i] = svd.singularValues()(i);
  return 0;
}

// polar::polar_decomposition<float> on a column-major matrix, as utils.h:71-73
void ref_polar_f(const float* A, float* R, float* H) { polar::polar_decomposition<float>(R, H, A); }

void ref_mat3_ops_f(const float* A, const float* B, float* prod, float* inv, float* det) {
  const glm::mat3 a = load_glm(A), b = load_glm(B);
  store_glm(a * b, prod);
  store_glm(glm::inverse(a), inv);
  *det = glm::determinant(a);
}

// The deformation-gradient update with the reference's constants
// (theta_c = 2.5e-2, theta_s = 5e-3), material_point_method.cpp:306-330.
int ref_fupdate_f(float dt, float h, const float* Bp, const float* FEp, const float* FPp, float* FEo, float* FPo,
                  float* sig) {
  const glm::mat3 Dinv = glm::inverse(glm::mat3(1.0) * (1.0f / 3.0f) * h * h);
  const glm::mat3 B = load_glm(Bp), FE = load_glm(FEp), FP = load_glm(FPp);
  const glm::mat3 total = (glm::mat3(1.0) + B * Dinv * dt) * FE * FP;
  const glm::mat3 hat = total * glm::inverse(FP);
  const glm::dmat3 hat_d = hat;
  Eigen::MatrixXf m(3, 3);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) m(i, j) = hat_d[i][j];  // utils.h:15-23
  Eigen::JacobiSVD<Eigen::MatrixXf, Eigen::ComputeFullU | Eigen::ComputeFullV> svd(m);
  if (svd.info() != Eigen::ComputationInfo::Success) return 1;
  Eigen::VectorXf s = svd.singularValues();
  for (int i = 0; i < 3; ++i) sig[i] = s(i);
  for (int i = 0; i < 3; ++i) s(i) = std::clamp(s(i), (float)(1 - 2.5f * 1e-2), (float)(1 + 5.0f * 1e-3));
  Eigen::MatrixXf Sm{{s(0), 0, 0}, {0, s(1), 0}, {0, 0, s(2)}};
  glm::mat3 U, V, S;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {  // utils.h:25-33
      U[i][j] = svd.matrixU()(i, j);
      V[i][j] = svd.matrixV()(i, j);
      S[i][j] = Sm(i, j);
    }
  const glm::mat3 feNew = U * S * glm::transpose(V);
  const glm::mat3 fpNew = glm::inverse(feNew) * total;
  store_glm(feNew, FEo);
  store_glm(fpNew, FPo);
  return 0;
}

// The box SDF of material_point_method.hpp:79-86 on glm, for a pose given as
// translation t, quaternion q (w,x,y,z) and scale s.
float ref_sdf_f(const float* c13, const float* pos) {
  const glm::vec3 t(c13[0], c13[1], c13[2]);
  const glm::quat q(c13[3], c13[4], c13[5], c13[6]);
  const glm::vec3 s(c13[7], c13[8], c13[9]);
  const glm::vec4 b4 = glm::scale(glm::mat4(), s) * glm::vec4{1, 1, 1, 1};
  const glm::vec3 b = {b4.x, b4.y, b4.z};
  const glm::vec4 p4 = glm::inverse(glm::translate(glm::mat4(), t) * glm::toMat4(q)) * glm::vec4{pos[0], pos[1], pos[2], 1};
  const glm::vec3 p = {p4.x, p4.y, p4.z};
  const glm::vec3 qq = glm::abs(p) - b;
  return glm::length(std::max({qq.x, qq.y, qq.z, 0.0f})) + std::min({std::max(qq.x, std::max(qq.y, qq.z)), 0.0f});
}

void ref_collider_inverse_f(const float* c13, float* inv16) {
  const glm::vec3 t(c13[0], c13[1], c13[2]);
  const glm::quat q(c13[3], c13[4], c13[5], c13[6]);
  const glm::mat4 inv = glm::inverse(glm::translate(glm::mat4(), t) * glm::toMat4(q));
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) inv16[4 * c + r] = inv[c][r];
}

}  // extern "C"
