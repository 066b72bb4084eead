// The reference demo's simulation loop (realtime-deformations/main.cpp:48-54,
// 119-218) without the viewer, written against the drop-in API.
//
//   facade_demo FRAMES OUT_PREFIX
//
// writes OUT_PREFIX.init.f32 (x, v, m of every particle after
// initializeParticles), OUT_PREFIX.colliders.f32 (13 floats per collider:
// translation, quaternion w x y z, scale, velocity), OUT_PREFIX.final.f32
// (x, v, F_E per particle after FRAMES substeps) and OUT_PREFIX.drawn.f32
// (x, y, z, size per particle as the viewer read them in the last frame), and
// prints the time per frame and the number of device reads of the particle
// state that the per-frame getParticles() loop made.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "mpm/material_point_method.hpp"

namespace {
mpmf::mat4 translate(float x, float y, float z) {
  mpmf::mat4 m(1.0f);
  m[3][0] = x;
  m[3][1] = y;
  m[3][2] = z;
  return m;
}
mpmf::mat4 rotate_z(float degrees) {
  const float a = degrees * 3.14159265358979f / 180.0f, c = std::cos(a), s = std::sin(a);
  mpmf::mat4 m(1.0f);
  m[0][0] = c;
  m[0][1] = s;
  m[1][0] = -s;
  m[1][1] = c;
  return m;
}
mpmf::mat4 scale(float x, float y, float z) {
  mpmf::mat4 m(1.0f);
  m[0][0] = x;
  m[1][1] = y;
  m[2][2] = z;
  return m;
}
mpmf::mat4 mul(const mpmf::mat4& a, const mpmf::mat4& b) {
  mpmf::mat4 r(1.0f);
  for (int j = 0; j < 4; ++j)
    for (int i = 0; i < 4; ++i) {
      float s = 0.f;
      for (int k = 0; k < 4; ++k) s += a[k][i] * b[j][k];
      r[j][i] = s;
    }
  return r;
}
void dump(const std::string& path, const std::vector<float>& v) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return;
  std::fwrite(v.data(), sizeof(float), v.size(), f);
  std::fclose(f);
}
}  // namespace

int main(int argc, char** argv) {
  using namespace MaterialPointMethod;
  const int frames = argc > 1 ? std::atoi(argv[1]) : 100;
  const std::string prefix = argc > 2 ? argv[2] : "facade_demo";
  const int numParticles = 270 + 1877;
  const v3t origin(0.5f, 0.6f, 0.5f);
  LagrangeEulerView MPM{20, 20, 20, numParticles};
  MPM.setLevel(DEFAULT_LOG_LEVEL_MPM);
  MPM.initializeParticles(origin, v3t(0.0f, -200.0f, 0.0f));
  {
    const Particle* p = MPM.getParticles();
    std::vector<float> init;
    for (int i = 0; i < numParticles; ++i) {
      for (int a = 0; a < 3; ++a) init.push_back(p[i].pos[a]);
      for (int a = 0; a < 3; ++a) init.push_back(p[i].velocity[a]);
      init.push_back(p[i].mass);
    }
    dump(prefix + ".init.f32", init);
  }
  MPM.rasterizeParticlesToGrid();
  MPM.computeParticleVolumesAndDensities();

  std::vector<MeshCollider> solidObjects;
  MeshCollider box1(v3t(0, 0, 0)), box3(v3t(0, 0, 0)), box4(v3t(0, 0, 0));
  box1.mesh.applyMatrix4(mul(mul(translate(0.5f, -0.2f, 0.5f), rotate_z(0.0f)), scale(0.4f, 0.4f, 0.4f)));
  box3.mesh.applyMatrix4(mul(mul(translate(0.0f, 0.3f, 0.5f), rotate_z(45.0f)), scale(0.2f, 0.2f, 0.3f)));
  box4.mesh.applyMatrix4(mul(mul(translate(1.0f, 0.3f, 0.5f), rotate_z(45.0f)), scale(0.2f, 0.2f, 0.3f)));
  solidObjects.push_back(box1);
  solidObjects.push_back(box3);
  solidObjects.push_back(box4);
  {
    std::vector<float> cols;
    for (const MeshCollider& c : solidObjects) {
      for (int a = 0; a < 3; ++a) cols.push_back(c.mesh.translation[a]);
      cols.push_back(c.mesh.rotation.w);
      cols.push_back(c.mesh.rotation.x);
      cols.push_back(c.mesh.rotation.y);
      cols.push_back(c.mesh.rotation.z);
      for (int a = 0; a < 3; ++a) cols.push_back(c.mesh.scale[a]);
      for (int a = 0; a < 3; ++a) cols.push_back(c.velocity[a]);
    }
    dump(prefix + ".colliders.f32", cols);
  }

  const double delta = 1e-5;
  const int64_t reads0 = MPM.particleReads();
  std::vector<float> drawn(4 * (size_t)numParticles);
  const auto t0 = std::chrono::steady_clock::now();
  for (int f = 0; f < frames; ++f) {
    for (auto& box : solidObjects) box.move((ftype)delta);
    MPM.rasterizeParticlesToGrid();
    MPM.computeExplicitGridForces();
    MPM.gridVelocitiesUpdate((ftype)delta);
    MPM.gridBasedCollisions((ftype)delta, solidObjects);
    MPM.updateDeformationGradient((ftype)delta);
    MPM.updateParticleVelocities();
    MPM.updateParticlePositions((ftype)delta);
    // the viewer's read (main.cpp:257-271): getParticles() once per particle
    for (int i = 0; i < MPM.getNumParticles(); i++) {
      const auto q = MPM.getParticles()[i];
      drawn[4 * i + 0] = q.pos.x;
      drawn[4 * i + 1] = q.pos.y;
      drawn[4 * i + 2] = q.pos.z;
      drawn[4 * i + 3] = q.size;
    }
  }
  const int64_t frame_reads = MPM.particleReads() - reads0;
  const Particle* p = MPM.getParticles();
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  std::vector<float> fin;
  double cx = 0, cy = 0, cz = 0;
  for (int i = 0; i < numParticles; ++i) {
    for (int a = 0; a < 3; ++a) fin.push_back(p[i].pos[a]);
    for (int a = 0; a < 3; ++a) fin.push_back(p[i].velocity[a]);
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r) fin.push_back(p[i].FElastic[c][r]);
    cx += p[i].pos.x;
    cy += p[i].pos.y;
    cz += p[i].pos.z;
  }
  dump(prefix + ".final.f32", fin);
  // the viewer's positions-only read (pinned mirror)
  const float* pos = MPM.getPositions();
  dump(prefix + ".positions.f32", std::vector<float>(pos, pos + 3 * (size_t)numParticles));
  dump(prefix + ".drawn.f32", drawn);
  std::printf("frames %d  ms/frame %.4f  centre %.6f %.6f %.6f\n", frames, ms / frames, cx / numParticles,
              cy / numParticles, cz / numParticles);
  std::printf("particle reads %lld in %d frames of %d getParticles() calls each\n", (long long)frame_reads, frames,
              numParticles);
  return 0;
}
