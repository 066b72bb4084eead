// material_point_method.hpp — drop-in C++ API of the MI355X MPM snow library.
//
// The same namespace, types and methods as the reference's simulation API
// (realtime-deformations/material_point_method.hpp:14-235 and
// realtime-deformations/logger.hpp:7-46), implemented over the C ABI of
// libmpm_hip.so (include/mpm_c.h).  A caller written against the reference —
// realtime-deformations/main.cpp:48-54,119-218,255-271 — compiles against
// this header and runs every stage on the GPU.
//
// Differences a caller can observe, all deliberate:
//  - No OpenGL: MeshCollider keeps its pose (translation, rotation, scale)
//    and velocity but draws nothing; the reference constructor's GL
//    arguments are accepted and ignored.
//  - initializeParticles places particles with the reference's algorithm but
//    draws its jitter from a seeded counter-based generator instead of the
//    platform's rand(), so placements are reproducible across platforms.
//  - getParticles() returns a host mirror refreshed from the device on each
//    call; edits made through the pointer reach the simulation after
//    commitParticles().
//  - The seven per-substep calls in main.cpp's order are queued and run as
//    one fused substep on the GPU; any other order, or a read, runs the
//    queued stages one by one with the reference's per-stage semantics.
//  - Math types are glm's when <glm/glm.hpp> is on the include path (as in
//    the reference), otherwise the small value types below with the same
//    layout and element access (m[col][row]).
//
// Defining MPM_FACADE_IN_TREE builds the header inside the reference's own
// tree, where main.cpp also includes mesh.hpp and logger.hpp: ftype, v3t,
// m3t, Logger, Loggable, Object3D and Mesh then come from the reference's
// constants.hpp, logger.hpp and mesh.hpp, and MeshCollider holds a drawable
// Mesh built from the reference constructor's GL arguments (only its pose is
// sent to the GPU).
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <iostream>
#include <string>
#include <vector>

#include "../mpm_c.h"

#if defined(MPM_FACADE_IN_TREE)
#include <glm/glm.hpp>
#include <constants.hpp>
#include <logger.hpp>
#include <mesh.hpp>
namespace mpmf {
using vec3 = glm::vec3;
using ivec3 = glm::ivec3;
using mat3 = glm::mat3;
using mat4 = glm::mat4;
using quat = glm::quat;
}  // namespace mpmf
#define MPM_FACADE_GLM 1
#elif __has_include(<glm/glm.hpp>) && !defined(MPM_FACADE_NO_GLM)
#include <glm/glm.hpp>
#include <glm/gtc/quaternion.hpp>
namespace mpmf {
using vec3 = glm::vec3;
using ivec3 = glm::ivec3;
using mat3 = glm::mat3;
using mat4 = glm::mat4;
using quat = glm::quat;
}  // namespace mpmf
#define MPM_FACADE_GLM 1
#else
namespace mpmf {
struct vec3 {
  float x = 0.f, y = 0.f, z = 0.f;
  vec3() = default;
  explicit vec3(float s) : x(s), y(s), z(s) {}
  vec3(float a, float b, float c) : x(a), y(b), z(c) {}
  float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  vec3 operator+(const vec3& o) const { return {x + o.x, y + o.y, z + o.z}; }
  vec3 operator-(const vec3& o) const { return {x - o.x, y - o.y, z - o.z}; }
  vec3 operator*(float s) const { return {x * s, y * s, z * s}; }
  vec3& operator+=(const vec3& o) { return *this = *this + o; }
  vec3& operator-=(const vec3& o) { return *this = *this - o; }
  vec3& operator*=(float s) { return *this = *this * s; }
};
inline vec3 operator*(float s, const vec3& v) { return v * s; }
struct ivec3 {
  int x = 0, y = 0, z = 0;
};
// column-major like glm: m[col][row]
struct mat3 {
  vec3 c[3];
  mat3() : mat3(1.0f) {}
  explicit mat3(float d) {
    for (int i = 0; i < 3; ++i) c[i][i] = d;
  }
  vec3& operator[](int i) { return c[i]; }
  const vec3& operator[](int i) const { return c[i]; }
};
struct mat4 {
  float c[4][4];
  mat4() : mat4(1.0f) {}
  explicit mat4(float d) {
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) c[i][j] = i == j ? d : 0.f;
  }
  float* operator[](int i) { return c[i]; }
  const float* operator[](int i) const { return c[i]; }
};
struct quat {
  float w = 1.f, x = 0.f, y = 0.f, z = 0.f;
};
}  // namespace mpmf
#endif

#ifndef MPM_FACADE_IN_TREE
typedef float ftype;
typedef mpmf::vec3 v3t;
typedef mpmf::mat3 m3t;

// realtime-deformations/logger.hpp:7-46
class Logger {
 public:
  enum class LogLevel { INFO, WARNING, ERROR };
  void log(LogLevel level, const std::string& msg) {
    if (level >= level_) std::cout << msg << "\n";
  }
  void log(LogLevel level, const std::string& name, ftype v) {
    if (level >= level_) log(level, name + " is: " + std::to_string(v));
  }
  void log(LogLevel level, const std::string& name, int v) {
    if (level >= level_) log(level, name + " is: " + std::to_string(v));
  }
  void setLevel(LogLevel level) { level_ = level; }

 private:
  LogLevel level_ = LogLevel::INFO;
};

class Loggable {
 public:
  void setLevel(Logger::LogLevel level) { logger.setLevel(level); }

 protected:
  Logger logger;
};

// Pose of a rigid object (realtime-deformations/mesh.hpp:11-24), without GL.
class Object3D {
 public:
  mpmf::mat4 matrixWorld{1.0f};
  v3t scale{1.0f, 1.0f, 1.0f};
  mpmf::quat rotation{};
  v3t translation{0.0f, 0.0f, 0.0f};
  // matrixWorld = m * matrixWorld, then translation / rotation / scale are
  // taken apart again (translate * rotate * scale matrices, no shear).
  void applyMatrix4(const mpmf::mat4& m);
};
#endif  // !MPM_FACADE_IN_TREE

namespace MaterialPointMethod {
const auto DEFAULT_LOG_LEVEL_OPTIMIZER = Logger::LogLevel::WARNING;
const auto DEFAULT_LOG_LEVEL_MPM = Logger::LogLevel::WARNING;

// realtime-deformations/material_point_method.hpp:18-72
namespace WeightCalculator {
extern ftype h;
inline ftype weightNx(ftype x) {
  const ftype modx = std::fabs(x);
  const ftype modx2 = modx * modx;
  const ftype modx3 = modx * modx * modx;
  if (modx < 1.0) return 0.5 * modx3 - modx2 + 2.0 / 3.0;
  if (modx < 2.0) return (1.0 / 6.0) * (2 - modx) * (2 - modx) * (2 - modx);
  return 0.0;
}
inline ftype weightNxDerivative(ftype x) {
  const ftype modx = std::fabs(x);
  if (modx < 1.0f) return x >= 0 ? +3.0 / 2.0 * x * x - 2 * x : -3.0 / 2.0 * x * x - 2 * x;
  if (modx < 2.0) return x >= 0 ? -0.5 * (2 - x) * (2 - x) : 0.5 * (2 + x) * (2 + x);
  return 0.0;
}
inline ftype wipHost(mpmf::ivec3 idx, v3t pos) {
  return weightNx(pos.x / h - idx.x) * weightNx(pos.y / h - idx.y) * weightNx(pos.z / h - idx.z);
}
inline v3t wipGrad(mpmf::ivec3 idx, v3t pos) {
  const ftype xc = (pos.x - idx.x * h) / h, yc = (pos.y - idx.y * h) / h, zc = (pos.z - idx.z * h) / h;
  const ftype wx = weightNx(xc), wy = weightNx(yc), wz = weightNx(zc);
  return v3t((1.0f / h) * weightNxDerivative(xc) * wy * wz, (1.0f / h) * wx * weightNxDerivative(yc) * wz,
             (1.0f / h) * wx * wy * weightNxDerivative(zc));
}
}  // namespace WeightCalculator

// realtime-deformations/material_point_method.hpp:74-93: an oriented box
// collider; `mesh` holds its pose, scale is the half extent.
#ifdef MPM_FACADE_IN_TREE
// The reference's collider (material_point_method.hpp:74-93): a drawable
// Mesh, its velocity and a box SDF of the mesh pose.
class MeshCollider {
 public:
  Mesh mesh;
  MeshCollider(GLuint programID, glm::mat4& VP, const std::vector<GLfloat>& vertices,
               const std::vector<GLfloat>& colors, const glm::vec3& vel);
  MeshCollider(const MeshCollider& o);
  std::function<ftype(const v3t&)> sdf;
  glm::vec3 velocity;
  void move(ftype timeDelta);
};
#else
class MeshCollider {
 public:
  Object3D mesh;
  v3t velocity;
  std::function<ftype(const v3t&)> sdf;

  explicit MeshCollider(const v3t& vel = v3t(0.0f, 0.0f, 0.0f));
  // The reference's (programID, VP, vertices, colors, velocity) signature;
  // the GL arguments are ignored.
  template <typename ProgramId, typename ViewProj, typename Vertices, typename Colors>
  MeshCollider(ProgramId, ViewProj&, const Vertices&, const Colors&, const v3t& vel) : MeshCollider(vel) {}
  MeshCollider(const MeshCollider& o);
  MeshCollider& operator=(const MeshCollider& o);
  void move(ftype timeDelta);
};
#endif

// realtime-deformations/material_point_method.hpp:96-110
struct Particle {
  ftype mass;
  v3t velocity;
  ftype volume;
  v3t pos;
  m3t FElastic{1.0f};
  m3t FPlastic{1.0f};
  m3t B{0.0f};
  m3t D{0.0f};
  std::vector<mpmf::ivec3> neighs;
  unsigned char r, g, b, a;
  float size;
};

// realtime-deformations/material_point_method.hpp:112-117
struct Cell {
  ftype mass;
  v3t force{0.0f, 0.0f, 0.0f};
  v3t velocity{0.0f, 0.0f, 0.0f};
  int nParticles{0};
};

// realtime-deformations/material_point_method.hpp:174-233
struct LagrangeEulerView : Loggable {
 public:
  bool useCuda = false;
  const m3t DpInverse;
  LagrangeEulerView(int max_i, int max_j, int max_k, int particlesNum);
  ~LagrangeEulerView();
  LagrangeEulerView(const LagrangeEulerView&) = delete;
  LagrangeEulerView& operator=(const LagrangeEulerView&) = delete;

  void initializeParticles(const v3t& particlesOrigin, const v3t& velocity);
  Particle* getParticles();
  int getNumParticles() { return nParticles; }
  void precalculateWeights();
  void rasterizeParticlesToGrid();
  void computeParticleVolumesAndDensities();
  void computeExplicitGridForces();
  void gridVelocitiesUpdate(ftype timeDelta);
  void timeIntegration(ftype timeDelta);
  void gridBasedCollisions(ftype timeDelta, const std::vector<MeshCollider>& objects);
  void updateDeformationGradient(ftype timeDelta);
  void updateParticleVelocities();
  void updateParticlePositions(ftype timeDelta);

  const int MAX_I, MAX_J, MAX_K;

  // ---- additions (no reference counterpart)
  // Upload edits made through getParticles() (positions, velocities, masses,
  // B, F_E, F_P, volumes).
  void commitParticles();
  // n fused substeps at once (the canonical seven calls n times).
  void substeps(ftype timeDelta, int n, const std::vector<MeshCollider>& objects);
  // Positions only, x y z per particle (what a viewer reads), from the
  // context's pinned host mirror: valid until the next getPositions().
  const float* getPositions();
  // Device reads of the whole particle state made by getParticles() so far.
  // The host mirror is re-read only after something changed the device state
  // (a stage, a substep, the initial volumes), so a viewer that calls
  // getParticles()[i] for every i each frame, as the reference's does
  // (realtime-deformations/main.cpp:257-271), costs one read per frame.
  int64_t particleReads() const { return reads; }
  mpm_ctx* context() { return ctx; }
  // Seed of the placement generator used by initializeParticles.
  uint64_t seed = 0x5eedULL;

 private:
  enum Pending : int { P_P2G, P_FORCES, P_GRID, P_COLLIDE, P_FUPDATE, P_G2P, P_ADVECT };
  struct Call {
    int stage;
    ftype dt;
  };
  mpm_ctx* ctx = nullptr;
  int nParticles;
  std::vector<Particle> particles;
  std::vector<float> positions;
  std::vector<Call> queue;
  std::vector<mpm_collider> colliders;
  bool mirrorStale = true;  // device state changed since the mirror was read
  int64_t reads = 0;
  void check(int rc, const char* what);
  void enqueue(int stage, ftype dt);
  void flush();
  void run_stage(int stage, ftype dt);
  void set_colliders(const std::vector<MeshCollider>& objects);
  void download();
};

}  // namespace MaterialPointMethod
