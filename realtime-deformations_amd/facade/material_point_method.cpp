// The reference's MaterialPointMethod API (include/mpm/material_point_method.hpp)
// over libmpm_hip's C ABI.  Compile this file with the caller's program, as
// the reference compiles its own material_point_method.cpp, and link
// libmpm_hip.so; it touches the math types only through element access, so
// it builds with glm or with the header's own value types.
#include "mpm/material_point_method.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace {

// ---- 4x4 helpers on element access (m[col][row]), valid for glm and mpmf
mpmf::mat4 mul4(const mpmf::mat4& a, const mpmf::mat4& b) {
  mpmf::mat4 r(1.0f);
  for (int j = 0; j < 4; ++j)
    for (int i = 0; i < 4; ++i) {
      float s = 0.f;
      for (int k = 0; k < 4; ++k) s += a[k][i] * b[j][k];
      r[j][i] = s;
    }
  return r;
}

// Rotation matrix (columns) to quaternion (Shepperd's method).
mpmf::quat rot_to_quat(const float R[3][3]) {  // R[col][row]
  const float m00 = R[0][0], m11 = R[1][1], m22 = R[2][2];
  const float tr = m00 + m11 + m22;
  float w, x, y, z;
  if (tr > 0.f) {
    const float s = std::sqrt(tr + 1.f) * 2.f;
    w = 0.25f * s;
    x = (R[1][2] - R[2][1]) / s;
    y = (R[2][0] - R[0][2]) / s;
    z = (R[0][1] - R[1][0]) / s;
  } else if (m00 > m11 && m00 > m22) {
    const float s = std::sqrt(1.f + m00 - m11 - m22) * 2.f;
    w = (R[1][2] - R[2][1]) / s;
    x = 0.25f * s;
    y = (R[1][0] + R[0][1]) / s;
    z = (R[2][0] + R[0][2]) / s;
  } else if (m11 > m22) {
    const float s = std::sqrt(1.f + m11 - m00 - m22) * 2.f;
    w = (R[2][0] - R[0][2]) / s;
    x = (R[1][0] + R[0][1]) / s;
    y = 0.25f * s;
    z = (R[2][1] + R[1][2]) / s;
  } else {
    const float s = std::sqrt(1.f + m22 - m00 - m11) * 2.f;
    w = (R[0][1] - R[1][0]) / s;
    x = (R[2][0] + R[0][2]) / s;
    y = (R[2][1] + R[1][2]) / s;
    z = 0.25f * s;
  }
  mpmf::quat q;
  q.w = w;
  q.x = x;
  q.y = y;
  q.z = z;
  return q;
}

mpm_collider to_c(const MaterialPointMethod::MeshCollider& c) {
  mpm_collider out;
  for (int a = 0; a < 3; ++a) {
    out.translation[a] = c.mesh.translation[a];
    out.scale[a] = c.mesh.scale[a];
    out.velocity[a] = c.velocity[a];
  }
  out.rotation[0] = c.mesh.rotation.w;
  out.rotation[1] = c.mesh.rotation.x;
  out.rotation[2] = c.mesh.rotation.y;
  out.rotation[3] = c.mesh.rotation.z;
  return out;
}

// splitmix64: the counter-based source of initializeParticles' jitter
struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  // utils.h:110-112: low + (r % 2000) / 2000 * (high - low)
  float rand_float(float low, float high) { return low + (float)(next() % 2000) / 2000.0f * (high - low); }
};

// utils.h:114-127: uniform point in a ball of radius R
v3t random_in_ball(SplitMix& g, ftype R) {
  const float phi = g.rand_float(0.0f, 2.0f * 3.1415f);
  const float costheta = g.rand_float(0.0f, 2.0f) - 1.0f;
  const float u = g.rand_float(0.0f, 1.0f);
  const float theta = std::acos(costheta);
  const float r = R * std::cbrt(u);
  return v3t(r * std::sin(theta) * std::cos(phi), r * std::sin(theta) * std::sin(phi), r * std::cos(theta));
}

}  // namespace

#ifndef MPM_FACADE_IN_TREE
void Object3D::applyMatrix4(const mpmf::mat4& m) {
  matrixWorld = mul4(m, matrixWorld);
  for (int a = 0; a < 3; ++a) translation[a] = matrixWorld[3][a];
  float R[3][3];
  for (int c = 0; c < 3; ++c) {
    const float len = std::sqrt(matrixWorld[c][0] * matrixWorld[c][0] + matrixWorld[c][1] * matrixWorld[c][1] +
                                matrixWorld[c][2] * matrixWorld[c][2]);
    scale[c] = len;
    for (int r = 0; r < 3; ++r) R[c][r] = len > 0.f ? matrixWorld[c][r] / len : (c == r ? 1.f : 0.f);
  }
  rotation = rot_to_quat(R);
}
#endif

namespace MaterialPointMethod {

ftype WeightCalculator::h = 0.05f;

namespace {
// the box SDF of material_point_method.hpp:79-86 for a host-side sdf()
ftype box_sdf(const Object3D& o, const v3t& pos) {
  const float w = o.rotation.w, x = o.rotation.x, y = o.rotation.y, z = o.rotation.z;
  // R^T (pos - t), R from the unit quaternion
  const float R[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)},
                         {2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)},
                         {2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)}};
  float q[3];
  for (int c = 0; c < 3; ++c) {
    const float p = R[c][0] * (pos[0] - o.translation[0]) + R[c][1] * (pos[1] - o.translation[1]) +
                    R[c][2] * (pos[2] - o.translation[2]);
    q[c] = std::fabs(p) - o.scale[c];
  }
  const float mx = std::max(std::max(q[0], q[1]), q[2]);
  return std::fabs(std::max(mx, 0.0f)) + std::min(mx, 0.0f);
}
}  // namespace

// sdf reads this collider's own pose, also in copies (the reference's lambda
// keeps pointing at the object it was built in).
#ifdef MPM_FACADE_IN_TREE
MeshCollider::MeshCollider(GLuint programID, glm::mat4& VP, const std::vector<GLfloat>& vertices,
                           const std::vector<GLfloat>& colors, const glm::vec3& vel)
    : mesh{programID, VP, vertices, colors}, velocity(vel) {
  sdf = [this](const v3t& p) -> ftype { return box_sdf(mesh, p); };
}
MeshCollider::MeshCollider(const MeshCollider& o) : mesh(o.mesh), velocity(o.velocity) {
  sdf = [this](const v3t& p) -> ftype { return box_sdf(mesh, p); };
}
#else
MeshCollider::MeshCollider(const v3t& vel) : velocity(vel) {
  sdf = [this](const v3t& p) -> ftype { return box_sdf(mesh, p); };
}
MeshCollider::MeshCollider(const MeshCollider& o) : mesh(o.mesh), velocity(o.velocity) {
  sdf = [this](const v3t& p) -> ftype { return box_sdf(mesh, p); };
}
MeshCollider& MeshCollider::operator=(const MeshCollider& o) {
  mesh = o.mesh;
  velocity = o.velocity;
  sdf = [this](const v3t& p) -> ftype { return box_sdf(mesh, p); };
  return *this;
}
#endif
void MeshCollider::move(ftype timeDelta) {
  mpmf::mat4 T(1.0f);
  for (int a = 0; a < 3; ++a) T[3][a] = timeDelta * velocity[a];
  mesh.applyMatrix4(T);
}

static m3t dp_inverse() {
  // inverse(mat3(1) * (1/3) * h * h) (material_point_method.hpp:177)
  const float d = (1.0f * (1.0f / 3.0f)) * WeightCalculator::h * WeightCalculator::h;
  m3t r(0.0f);
  for (int i = 0; i < 3; ++i) r[i][i] = (d * d) * (1.0f / (d * (d * d)));
  return r;
}

LagrangeEulerView::LagrangeEulerView(int max_i, int max_j, int max_k, int particlesNum)
    : DpInverse(dp_inverse()), MAX_I(max_i), MAX_J(max_j), MAX_K(max_k), nParticles(particlesNum) {
  mpm_params p;
  mpm_default_params(&p, max_i, max_j, max_k, WeightCalculator::h);
  p.bc_mode = MPM_BC_COLLIDERS;
  check(mpm_create(&p, particlesNum, &ctx), "mpm_create");
  particles.resize(particlesNum);
}

LagrangeEulerView::~LagrangeEulerView() {
  if (ctx) mpm_destroy(ctx);
}

void LagrangeEulerView::check(int rc, const char* what) {
  if (rc != MPM_OK) logger.log(Logger::LogLevel::ERROR, std::string(what) + ": " + mpm_last_error(ctx));
}

// material_point_method.cpp:18-63 with a seeded generator in place of rand()
void LagrangeEulerView::initializeParticles(const v3t& particlesOrigin, const v3t& velocity) {
  SplitMix g{seed};
  const ftype width = 0.2f;
  const int cx = (int)(particlesOrigin.x / WeightCalculator::h);
  const int cy = (int)(particlesOrigin.y / WeightCalculator::h);
  const int cz = (int)(particlesOrigin.z / WeightCalculator::h);
  const int widthInCells = (int)(width / WeightCalculator::h);
  int left = nParticles, needMore = 0;
  static const float ds[8][3] = {{1, 1, 1}, {1, 1, 3}, {1, 3, 1}, {1, 3, 3}, {3, 1, 1}, {3, 1, 3}, {3, 3, 1}, {3, 3, 3}};
  for (int i = cx - widthInCells; i < cx + widthInCells; ++i)
    for (int j = cy - widthInCells; j < cy + widthInCells; ++j)
      for (int k = cz - widthInCells; k < cz + widthInCells; ++k)
        for (int d = 0; d < 8; ++d) {
          const v3t jit = random_in_ball(g, 0.25f);
          const v3t pos((i + ds[d][0] * 0.25f + jit.x) * WeightCalculator::h,
                        (j + ds[d][1] * 0.25f + jit.y) * WeightCalculator::h,
                        (k + ds[d][2] * 0.25f + jit.z) * WeightCalculator::h);
          const v3t rel = pos - particlesOrigin;
          if (std::sqrt(rel.x * rel.x + rel.y * rel.y + rel.z * rel.z) > width) continue;
          if (!left) {
            ++needMore;
            continue;
          }
          Particle& p = particles[--left];
          p.pos = pos;
          p.velocity = velocity;
          p.r = p.g = p.b = p.a = 255;
          p.size = 0.02f;
          p.mass = 0.00006f;
          p.volume = 0.0f;
          p.FElastic = m3t(1.0f);
          p.FPlastic = m3t(1.0f);
          p.B = m3t(0.0f);
        }
  if (needMore) std::cout << needMore << " more!!!\n";
  if (left) {
    // the reference leaves these at the origin, where their stencil leaves the grid
    logger.log(Logger::LogLevel::WARNING, std::to_string(left) + " particles not placed: put at the origin point");
    for (int i = 0; i < left; ++i) {
      particles[i].pos = particlesOrigin;
      particles[i].velocity = velocity;
      particles[i].mass = 0.00006f;
      particles[i].size = 0.02f;
    }
  }
  queue.clear();
  std::vector<float> x(3 * (size_t)nParticles), v(3 * (size_t)nParticles), m(nParticles);
  for (int i = 0; i < nParticles; ++i) {
    for (int a = 0; a < 3; ++a) {
      x[3 * i + a] = particles[i].pos[a];
      v[3 * i + a] = particles[i].velocity[a];
    }
    m[i] = particles[i].mass;
  }
  check(mpm_set_particles(ctx, x.data(), v.data(), m.data(), nullptr), "initializeParticles");
  mirrorStale = false;  // device state = the mirror just placed
}

void LagrangeEulerView::commitParticles() {
  flush();
  const size_t n = (size_t)nParticles;
  std::vector<float> x(3 * n), v(3 * n), m(n), B(9 * n), FE(9 * n), FP(9 * n), V0(n);
  for (size_t i = 0; i < n; ++i) {
    const Particle& p = particles[i];
    for (int a = 0; a < 3; ++a) {
      x[3 * i + a] = p.pos[a];
      v[3 * i + a] = p.velocity[a];
    }
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r) {
        B[9 * i + 3 * c + r] = p.B[c][r];
        FE[9 * i + 3 * c + r] = p.FElastic[c][r];
        FP[9 * i + 3 * c + r] = p.FPlastic[c][r];
      }
    m[i] = p.mass;
    V0[i] = p.volume;
  }
  check(mpm_set_particles(ctx, x.data(), v.data(), m.data(), nullptr), "commitParticles");
  check(mpm_set_state(ctx, B.data(), FE.data(), FP.data(), V0.data()), "commitParticles");
  mirrorStale = false;  // device state = the mirror just uploaded
}

void LagrangeEulerView::download() {
  const size_t n = (size_t)nParticles;
  std::vector<float> x(3 * n), v(3 * n), B(9 * n), FE(9 * n), FP(9 * n), V0(n);
  check(mpm_get_particles(ctx, x.data(), v.data(), B.data(), FE.data(), FP.data(), V0.data()), "getParticles");
  for (size_t i = 0; i < n; ++i) {
    Particle& p = particles[i];
    for (int a = 0; a < 3; ++a) {
      p.pos[a] = x[3 * i + a];
      p.velocity[a] = v[3 * i + a];
    }
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r) {
        p.B[c][r] = B[9 * i + 3 * c + r];
        p.FElastic[c][r] = FE[9 * i + 3 * c + r];
        p.FPlastic[c][r] = FP[9 * i + 3 * c + r];
      }
    p.volume = V0[i];
  }
}

// The reference returns a pointer into its live particle array
// (material_point_method.hpp:181-183) and its viewer calls this once per
// particle (main.cpp:257-271); the mirror is re-read only when the device
// state changed since the last read.
Particle* LagrangeEulerView::getParticles() {
  flush();
  if (mirrorStale) {
    download();
    ++reads;
    mirrorStale = false;
  }
  return particles.data();
}

const float* LagrangeEulerView::getPositions() {
  flush();
  // gather + copy into the context's pinned mirror; no host-side copy
  const float* view = nullptr;
  int rc = mpm_positions_begin(ctx);
  if (rc == MPM_OK) rc = mpm_positions_wait(ctx, &view);
  if (rc == MPM_OK) rc = mpm_sync(ctx);
  check(rc, "getPositions");
  if (rc != MPM_OK || !view) {
    positions.assign(3 * (size_t)nParticles, 0.0f);
    return positions.data();
  }
  return view;
}

void LagrangeEulerView::precalculateWeights() {
  // The reference's dense cells x particles weight tensor is not used by any
  // stage (material_point_method.cpp:65-76); nothing to do.
}

void LagrangeEulerView::timeIntegration(ftype) {
  logger.log(Logger::LogLevel::ERROR, "timeIntegration: the implicit LBFGS path is not provided (explicit stages only)");
}

void LagrangeEulerView::run_stage(int stage, ftype dt) {
  static const int map[] = {MPM_STAGE_P2G, MPM_STAGE_FORCES, MPM_STAGE_GRID, MPM_STAGE_COLLIDE,
                            MPM_STAGE_FUPDATE, MPM_STAGE_G2P, MPM_STAGE_ADVECT};
  check(mpm_stage(ctx, map[stage], dt), "stage");
  mirrorStale = true;
}

// The canonical order of main.cpp:192-218; a completed run with one dt
// becomes one fused substep, anything else runs stage by stage.
void LagrangeEulerView::enqueue(int stage, ftype dt) {
  if ((int)queue.size() != stage) {
    flush();
    if (stage != P_P2G) {
      run_stage(stage, dt);
      return;
    }
  }
  queue.push_back({stage, dt});
  if (stage == P_ADVECT) {
    const ftype d = queue[P_GRID].dt;
    if (queue[P_COLLIDE].dt == d && queue[P_FUPDATE].dt == d && queue[P_ADVECT].dt == d) {
      check(mpm_substeps(ctx, d, 1), "substep");
      mirrorStale = true;
      queue.clear();
    } else {
      flush();
    }
  }
}

void LagrangeEulerView::flush() {
  for (const Call& c : queue) run_stage(c.stage, c.dt);
  queue.clear();
}

void LagrangeEulerView::set_colliders(const std::vector<MeshCollider>& objects) {
  std::vector<mpm_collider> cs;
  for (const MeshCollider& o : objects) cs.push_back(to_c(o));
  const bool same = cs.size() == colliders.size() &&
                    (cs.empty() || std::memcmp(cs.data(), colliders.data(), cs.size() * sizeof(mpm_collider)) == 0);
  if (same) return;
  colliders = cs;  // queued stages before the collision stage do not read colliders
  check(mpm_set_colliders(ctx, colliders.data(), (int32_t)colliders.size()), "gridBasedCollisions");
}

void LagrangeEulerView::rasterizeParticlesToGrid() { enqueue(P_P2G, 0.0f); }
void LagrangeEulerView::computeParticleVolumesAndDensities() {
  flush();
  check(mpm_stage(ctx, MPM_STAGE_VOLUMES, 0.0f), "computeParticleVolumesAndDensities");
  mirrorStale = true;
}
void LagrangeEulerView::computeExplicitGridForces() { enqueue(P_FORCES, 0.0f); }
void LagrangeEulerView::gridVelocitiesUpdate(ftype dt) { enqueue(P_GRID, dt); }
void LagrangeEulerView::gridBasedCollisions(ftype dt, const std::vector<MeshCollider>& objects) {
  set_colliders(objects);
  enqueue(P_COLLIDE, dt);
}
void LagrangeEulerView::updateDeformationGradient(ftype dt) { enqueue(P_FUPDATE, dt); }
void LagrangeEulerView::updateParticleVelocities() { enqueue(P_G2P, 0.0f); }
void LagrangeEulerView::updateParticlePositions(ftype dt) { enqueue(P_ADVECT, dt); }

void LagrangeEulerView::substeps(ftype dt, int n, const std::vector<MeshCollider>& objects) {
  flush();
  set_colliders(objects);
  check(mpm_substeps(ctx, dt, n), "substeps");
  mirrorStale = true;
}

}  // namespace MaterialPointMethod
