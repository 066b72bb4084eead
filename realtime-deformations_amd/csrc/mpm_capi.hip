// C ABI of libmpm_hip: context lifetime, particle/grid transfer, the legacy
// stage calls and the fused substep loop.  Declarations: include/mpm_c.h.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>

#include "mpm_adjoint.hpp"
#include "mpm_bin.hpp"
#include "mpm_fused.hpp"
#include "mpm_internal.hpp"

using namespace mpm;

#define MPM_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);                     \
      return MPM_ERR_HIP;                                                               \
    }                                                                                   \
  } while (0)

namespace {

// glm inverse(translate(mat4(1), t) * toMat4(q)) in glm 0.9.7.1's operation
// order (gtc/quaternion.inl mat3_cast, gtc/matrix_transform.inl translate,
// detail/type_mat4x4.inl operator* and compute_inverse), so the collider SDF
// sees the same transform the reference's MeshCollider::sdf builds
// (realtime-deformations/material_point_method.hpp:79-86).
struct M4h {
  float c[4][4];
};
M4h m4_ident() {
  M4h r;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) r.c[a][b] = (a == b) ? 1.0f : 0.0f;
  return r;
}
M4h m4_mul(const M4h& a, const M4h& b) {
  M4h r;
  for (int j = 0; j < 4; ++j)
    for (int i = 0; i < 4; ++i)
      r.c[j][i] = a.c[0][i] * b.c[j][0] + a.c[1][i] * b.c[j][1] + a.c[2][i] * b.c[j][2] + a.c[3][i] * b.c[j][3];
  return r;
}
M4h m4_inverse(const M4h& M) {
  auto m = [&](int a, int b) { return M.c[a][b]; };
  const float c00 = m(2, 2) * m(3, 3) - m(3, 2) * m(2, 3), c02 = m(1, 2) * m(3, 3) - m(3, 2) * m(1, 3),
              c03 = m(1, 2) * m(2, 3) - m(2, 2) * m(1, 3), c04 = m(2, 1) * m(3, 3) - m(3, 1) * m(2, 3),
              c06 = m(1, 1) * m(3, 3) - m(3, 1) * m(1, 3), c07 = m(1, 1) * m(2, 3) - m(2, 1) * m(1, 3),
              c08 = m(2, 1) * m(3, 2) - m(3, 1) * m(2, 2), c10 = m(1, 1) * m(3, 2) - m(3, 1) * m(1, 2),
              c11 = m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2), c12 = m(2, 0) * m(3, 3) - m(3, 0) * m(2, 3),
              c14 = m(1, 0) * m(3, 3) - m(3, 0) * m(1, 3), c15 = m(1, 0) * m(2, 3) - m(2, 0) * m(1, 3),
              c16 = m(2, 0) * m(3, 2) - m(3, 0) * m(2, 2), c18 = m(1, 0) * m(3, 2) - m(3, 0) * m(1, 2),
              c19 = m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2), c20 = m(2, 0) * m(3, 1) - m(3, 0) * m(2, 1),
              c22 = m(1, 0) * m(3, 1) - m(3, 0) * m(1, 1), c23 = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1);
  const float f0[4] = {c00, c00, c02, c03}, f1[4] = {c04, c04, c06, c07}, f2[4] = {c08, c08, c10, c11},
              f3_[4] = {c12, c12, c14, c15}, f4[4] = {c16, c16, c18, c19}, f5[4] = {c20, c20, c22, c23};
  const float v0[4] = {m(1, 0), m(0, 0), m(0, 0), m(0, 0)}, v1[4] = {m(1, 1), m(0, 1), m(0, 1), m(0, 1)},
              v2[4] = {m(1, 2), m(0, 2), m(0, 2), m(0, 2)}, v3[4] = {m(1, 3), m(0, 3), m(0, 3), m(0, 3)};
  const float sa[4] = {1, -1, 1, -1}, sb[4] = {-1, 1, -1, 1};
  M4h inv;
  for (int k = 0; k < 4; ++k) {
    inv.c[0][k] = (v1[k] * f0[k] - v2[k] * f1[k] + v3[k] * f2[k]) * sa[k];
    inv.c[1][k] = (v0[k] * f0[k] - v2[k] * f3_[k] + v3[k] * f4[k]) * sb[k];
    inv.c[2][k] = (v0[k] * f1[k] - v1[k] * f3_[k] + v3[k] * f5[k]) * sa[k];
    inv.c[3][k] = (v0[k] * f2[k] - v1[k] * f4[k] + v2[k] * f5[k]) * sb[k];
  }
  float d0[4];
  const float row0[4] = {inv.c[0][0], inv.c[1][0], inv.c[2][0], inv.c[3][0]};
  for (int k = 0; k < 4; ++k) d0[k] = M.c[0][k] * row0[k];
  const float ood = 1.0f / ((d0[0] + d0[1]) + (d0[2] + d0[3]));
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) inv.c[a][b] = inv.c[a][b] * ood;
  return inv;
}

int check_ctx(mpm_ctx* ctx) { return ctx ? MPM_OK : MPM_ERR_ARG; }

// Run one launch; while profiling, bracket it with a pair of HIP events on
// the context stream.
template <typename F> int timed(mpm_ctx* ctx, int kernel, F&& launch) {
  if (!ctx->profiling || !ctx->per_kernel) {
    launch();
    return MPM_OK;
  }
  while (ctx->events.size() < ctx->events_used + 2) {
    hipEvent_t e;
    MPM_HIP(hipEventCreate(&e));
    ctx->events.push_back(e);
  }
  MPM_HIP(hipEventRecord(ctx->events[ctx->events_used], ctx->stream));
  launch();
  MPM_HIP(hipEventRecord(ctx->events[ctx->events_used + 1], ctx->stream));
  ctx->events_used += 2;
  ctx->event_kernel.push_back(kernel);
  return MPM_OK;
}

// Every particle's 5^3 stencil (base-2 .. base+2, base = int(x/h)) must lie
// inside its scene's grid; the reference indexes out of bounds otherwise.
int check_positions(mpm_ctx* ctx, int scene, const float* x) {
  const Scene& S = ctx->scenes[scene];
  for (int p = 0; p < S.np; ++p)
    for (int a = 0; a < 3; ++a) {
      const float q = x[3 * p + a] / S.h;
      if (!(q >= 2.0f) || !(q < (float)(S.dims[a] - 2))) {
        char buf[200];
        snprintf(buf, sizeof buf, "scene %d particle %d coordinate %d = %g puts its stencil outside the grid", scene,
                 p, a, (double)x[3 * p + a]);
        ctx->err = buf;
        return MPM_ERR_BOUNDS;
      }
    }
  return MPM_OK;
}

int status_code(int st) {
  if (st & ST_NONFINITE) return MPM_ERR_NONFINITE;
  if (st & ST_BOUNDS) return MPM_ERR_BOUNDS;
  return (st & ST_CFL) ? MPM_ERR_CFL : MPM_OK;
}

// Per-scene device status words.  The error bits (non-finite, bounds) of
// every scene are returned as the first failing scene's code, with the
// context error set, and cleared; the CFL warning bit is left pending for
// mpm_scene_status, unless `all` (that call) reads and clears every bit.
int read_status(mpm_ctx* ctx, std::vector<int>* per_scene = nullptr, bool all = false) {
  std::vector<int> st(ctx->nscenes, 0);
  MPM_HIP(hipMemcpyAsync(st.data(), ctx->status, sizeof(int) * ctx->nscenes, hipMemcpyDeviceToHost, ctx->stream));
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  if (per_scene) *per_scene = st;
  const int errbits = ST_NONFINITE | ST_BOUNDS;
  int rc = MPM_OK;
  bool any = false;
  for (int s = 0; s < ctx->nscenes; ++s) {
    any = any || (all ? st[s] != 0 : (st[s] & errbits) != 0);
    if ((st[s] & errbits) && rc == MPM_OK) {
      rc = status_code(st[s]);
      ctx->err = (st[s] & ST_NONFINITE)
                     ? "non-finite state: the SVD met a non-finite F (particle F left unchanged) or a "
                       "particle's velocity or stress input is not finite"
                     : "device status " + std::to_string(st[s]);
      if (ctx->nscenes > 1) ctx->err = "scene " + std::to_string(s) + ": " + ctx->err;
    }
  }
  if (any) {
    if (all) {
      MPM_HIP(hipMemsetAsync(ctx->status, 0, sizeof(int) * ctx->nscenes, ctx->stream));
    } else {
      // clear the error bits, keep pending warnings
      for (int s = 0; s < ctx->nscenes; ++s) st[s] &= ~errbits;
      MPM_HIP(hipMemcpyAsync(ctx->status, st.data(), sizeof(int) * ctx->nscenes, hipMemcpyHostToDevice,
                             ctx->stream));
    }
    MPM_HIP(hipStreamSynchronize(ctx->stream));
  }
  return rc;
}

// Scene table entry for one mpm_params; grid placement filled by the caller.
Scene make_scene(const mpm_params& p) {
  Scene S{};
  for (int a = 0; a < 3; ++a) {
    S.dims[a] = p.dims[a];
    S.nb[a] = (p.dims[a] + 3) / 4;
    S.g[a] = p.gravity[a];
  }
  S.nblocks = (int64_t)S.nb[0] * S.nb[1] * S.nb[2];
  S.h = p.h;
  S.inv_h = 1.0f / p.h;
  S.E = p.E;
  S.lo = (float)(1.0 - p.theta_c);
  S.hi = (float)(1.0 + p.theta_s);
  S.gap64 = p.svd_gap_fp64;
  S.lame.nu = p.nu;
  S.lame.xi = p.xi;
  S.lame.dmu = 2.0f * (1.0f + p.nu);
  S.lame.rmu = 1.0f / S.lame.dmu;
  S.lame.dlam = (1.0f + p.nu) * (1.0f - 2.0f * p.nu);
  S.lame.rlam = 1.0f / S.lame.dlam;
  S.bc_mode = p.bc_mode;
  // DpInverse = inverse(mat3(1) * (1/3) * h * h) (material_point_method.hpp:177),
  // glm compute_inverse on the host
  float d = 1.0f * (1.0f / 3.0f);
  d = d * p.h;
  d = d * p.h;
  M3 D;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) D.c[i][j] = (i == j) ? d : 0.0f;
  const float ood = 1.0f / (+D.c[0][0] * (D.c[1][1] * D.c[2][2] - D.c[2][1] * D.c[1][2]) -
                            D.c[1][0] * (D.c[0][1] * D.c[2][2] - D.c[2][1] * D.c[0][2]) +
                            D.c[2][0] * (D.c[0][1] * D.c[1][2] - D.c[1][1] * D.c[0][2]));
  M3 r;
  r.c[0][0] = +(D.c[1][1] * D.c[2][2] - D.c[2][1] * D.c[1][2]) * ood;
  r.c[1][0] = -(D.c[1][0] * D.c[2][2] - D.c[2][0] * D.c[1][2]) * ood;
  r.c[2][0] = +(D.c[1][0] * D.c[2][1] - D.c[2][0] * D.c[1][1]) * ood;
  r.c[0][1] = -(D.c[0][1] * D.c[2][2] - D.c[2][1] * D.c[0][2]) * ood;
  r.c[1][1] = +(D.c[0][0] * D.c[2][2] - D.c[2][0] * D.c[0][2]) * ood;
  r.c[2][1] = -(D.c[0][0] * D.c[2][1] - D.c[2][0] * D.c[0][1]) * ood;
  r.c[0][2] = +(D.c[0][1] * D.c[1][2] - D.c[1][1] * D.c[0][2]) * ood;
  r.c[1][2] = -(D.c[0][0] * D.c[1][2] - D.c[1][0] * D.c[0][2]) * ood;
  r.c[2][2] = +(D.c[0][0] * D.c[1][1] - D.c[1][0] * D.c[0][1]) * ood;
  S.dinv = r.c[0][0];  // the off-diagonal signed zeros are those of dinv_matrix
  return S;
}

int upload_scene_table(mpm_ctx* ctx) {
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  MPM_HIP(hipMemcpy(ctx->d_scenes, ctx->scenes.data(), sizeof(Scene) * ctx->nscenes, hipMemcpyHostToDevice));
  return MPM_OK;
}

int valid_scene(mpm_ctx* ctx, int scene) {
  if (scene < 0 || scene >= ctx->nscenes) {
    ctx->err = "scene index out of range";
    return MPM_ERR_ARG;
  }
  return MPM_OK;
}

}  // namespace

namespace mpm {
KArgs make_args(const mpm_ctx* c, float dt) {
  KArgs a{};
  a.P = c->P;
  a.stride = c->stride;
  a.n = c->n;
  a.gacc = reinterpret_cast<float4*>(c->gacc);
  a.gacc64 = c->gacc;
  a.gvel = c->gvel;
  a.gforce = c->gforce;
  a.nnodes = c->nnodes;
  a.dt = dt;
  a.compose = c->prm.compose;
  a.nscenes = c->nscenes;
  a.scenes = c->d_scenes;
  a.blk_scene = c->blk_scene;
  a.sc = c->scenes[0];
  a.status = c->status;
  a.exact = c->exact ? 1 : 0;
  return a;
}
void host_collider_inverse(const float t[3], const float q[4], float inv[4][4]) {
  M4h T = m4_ident();
  for (int k = 0; k < 4; ++k) {
    const float e0 = k == 0 ? 1.f : 0.f, e1 = k == 1 ? 1.f : 0.f, e2 = k == 2 ? 1.f : 0.f, e3 = k == 3 ? 1.f : 0.f;
    T.c[3][k] = e0 * t[0] + e1 * t[1] + e2 * t[2] + e3;
  }
  const float w = q[0], x = q[1], y = q[2], z = q[3];
  const float qxx = x * x, qyy = y * y, qzz = z * z, qxz = x * z, qxy = x * y, qyz = y * z, qwx = w * x, qwy = w * y,
              qwz = w * z;
  M4h R = m4_ident();
  R.c[0][0] = 1 - 2 * (qyy + qzz);
  R.c[0][1] = 2 * (qxy + qwz);
  R.c[0][2] = 2 * (qxz - qwy);
  R.c[1][0] = 2 * (qxy - qwz);
  R.c[1][1] = 1 - 2 * (qxx + qzz);
  R.c[1][2] = 2 * (qyz + qwx);
  R.c[2][0] = 2 * (qxz + qwy);
  R.c[2][1] = 2 * (qyz - qwx);
  R.c[2][2] = 1 - 2 * (qxx + qyy);
  const M4h I = m4_inverse(m4_mul(T, R));
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) inv[a][b] = I.c[a][b];
}
}  // namespace mpm

namespace mpm {
int upload_field(mpm_ctx* ctx, int scene, int field, const float* host, int comps) {
  const Scene& S = ctx->scenes[scene];
  if (!host || S.np == 0) return MPM_OK;
  DevBuf<float> tmp;
  const size_t bytes = sizeof(float) * (size_t)S.np * comps;
  MPM_HIP(tmp.alloc(bytes));
  MPM_HIP(hipMemcpyAsync(tmp.p, host, bytes, hipMemcpyHostToDevice, ctx->stream));
  launch_scatter_field(ctx->P + (int64_t)field * ctx->stride, ctx->stride, tmp.p, ctx->pid, S.p0, S.np, comps,
                       ctx->stream);
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  return MPM_OK;
}
int download_field(mpm_ctx* ctx, int scene, int field, float* host, int comps) {
  const Scene& S = ctx->scenes[scene];
  if (!host || S.np == 0) return MPM_OK;
  DevBuf<float> tmp;
  const size_t bytes = sizeof(float) * (size_t)S.np * comps;
  MPM_HIP(tmp.alloc(bytes));
  launch_gather_field(tmp.p, ctx->P + (int64_t)field * ctx->stride, ctx->stride, ctx->pid, S.p0, S.np, comps,
                      ctx->stream);
  MPM_HIP(hipMemcpyAsync(host, tmp.p, bytes, hipMemcpyDeviceToHost, ctx->stream));
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  return MPM_OK;
}
int check_scene_positions(mpm_ctx* ctx, int scene, const float* x) { return check_positions(ctx, scene, x); }
int reset_scene_order(mpm_ctx* ctx, int scene) {
  const Scene& S = ctx->scenes[scene];
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<int> ids((size_t)S.np);
  for (int i = 0; i < S.np; ++i) ids[i] = S.p0 + i;
  if (S.np) MPM_HIP(hipMemcpy(ctx->pid + S.p0, ids.data(), sizeof(int) * S.np, hipMemcpyHostToDevice));
  ctx->since_sort = -1;
  return MPM_OK;
}
}  // namespace mpm

extern "C" {

void mpm_default_params(mpm_params* p, int32_t ni, int32_t nj, int32_t nk, float h) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  p->dims[0] = ni;
  p->dims[1] = nj;
  p->dims[2] = nk;
  p->compose = MPM_COMPOSE_LITERAL;
  p->bc_mode = 0;
  p->h = h;
  p->E = 1.4e5f;
  p->nu = 0.2f;
  p->xi = 10.0f;
  p->theta_c = 2.5e-2;
  p->theta_s = 5e-3;
  p->gravity[0] = 0.0f;
  p->gravity[1] = -9.8f;
  p->gravity[2] = 0.0f;
  p->svd_gap_fp64 = 0.0f;
}

int mpm_create_batch(const mpm_params* params, const int32_t* n_particles, int32_t nscenes, mpm_ctx** out) {
  if (!params || !n_particles || !out || nscenes < 1) return MPM_ERR_ARG;
  int64_t ntot = 0;
  for (int s = 0; s < nscenes; ++s) {
    const mpm_params& p = params[s];
    if (n_particles[s] < 0) return MPM_ERR_ARG;
    for (int a = 0; a < 3; ++a)
      if (p.dims[a] < 6) return MPM_ERR_ARG;
    if (!(p.h > 0.0f)) return MPM_ERR_ARG;
    if (p.compose != MPM_COMPOSE_LITERAL && p.compose != MPM_COMPOSE_TEXTBOOK) return MPM_ERR_ARG;
    if (p.compose != params[0].compose) return MPM_ERR_ARG;  // one fused kernel per launch
    ntot += n_particles[s];
  }
  if (ntot >= (int64_t)1 << 30) return MPM_ERR_ARG;  // slot byte offsets are 32-bit
  auto* ctx = new (std::nothrow) mpm_ctx();
  if (!ctx) return MPM_ERR_NOMEM;
  *out = ctx;
  ctx->prm = params[0];
  ctx->sprm.assign(params, params + nscenes);
  ctx->hcoll.assign(nscenes, {});
  ctx->nscenes = nscenes;
  ctx->scenes.resize(nscenes);
  ctx->scoll.assign(nscenes, nullptr);
  ctx->scene_set.assign(nscenes, 0);
  int64_t blk0 = 0;
  int p0 = 0;
  for (int s = 0; s < nscenes; ++s) {
    Scene& S = ctx->scenes[s];
    S = make_scene(params[s]);
    S.blk0 = blk0;
    S.p0 = p0;
    S.np = n_particles[s];
    blk0 += S.nblocks;
    p0 += S.np;
  }
  ctx->nblocks = blk0;
  ctx->nnodes = blk0 * 64;
  if (ctx->nnodes >= ((int64_t)1 << 32)) {  // sort keys are block * 64 + cell in 32 bits
    ctx->err = "grids too large for 32-bit sort keys";
    return MPM_ERR_ARG;
  }
  ctx->n = (int)ntot;
  ctx->stride = ((int64_t)ctx->n + 63) / 64 * 64;
  if (ctx->stride == 0) ctx->stride = 64;
  ctx->resort_every = params[0].resort_every > 0 ? params[0].resort_every : 16;
  ctx->block_grid = (int)std::max<int64_t>(1, std::min<int64_t>(ctx->nblocks, 2048));
  MPM_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  const size_t pbytes = sizeof(float) * (size_t)NFIELDS * (size_t)ctx->stride;
  const size_t gbytes = sizeof(float4) * (size_t)ctx->nnodes;
  const int64_t ntiles = ((int64_t)ctx->n + 4095) / 4096;
  const size_t nn = (size_t)std::max(1, ctx->n);
  const size_t nb = (size_t)ctx->nblocks;
  bool ok = true;
  auto alloc = [&](auto** ptr, size_t bytes) {
    if (ok && hipMalloc(reinterpret_cast<void**>(ptr), bytes) != hipSuccess) ok = false;
  };
  alloc(&ctx->P, pbytes);
  alloc(&ctx->P2, pbytes);
  alloc(&ctx->pid, sizeof(int) * nn);
  alloc(&ctx->pid2, sizeof(int) * nn);
  alloc(&ctx->keys, sizeof(uint32_t) * nn);
  alloc(&ctx->vals, sizeof(uint32_t) * nn);
  alloc(&ctx->keys2, sizeof(uint32_t) * nn);
  alloc(&ctx->vals2, sizeof(uint32_t) * nn);
  alloc(&ctx->hist, sizeof(uint32_t) * 256 * (size_t)(std::max<int64_t>(1, ntiles) + 1));
  alloc(&ctx->blk_end, sizeof(int) * nb);
  alloc(&ctx->blist, sizeof(int4) * nb);
  alloc(&ctx->nblk, sizeof(int));
  alloc(&ctx->gflag, sizeof(int) * nb);
  alloc(&ctx->glist[0], sizeof(int) * nb);
  alloc(&ctx->glist[1], sizeof(int) * nb);
  alloc(&ctx->gcnt, sizeof(int) * 4);
  alloc(&ctx->dstep, sizeof(int));
  alloc(&ctx->bctr, sizeof(int));
  alloc(&ctx->gacc, 2 * gbytes);  // a double4 per node (reproducible mode)
  alloc(&ctx->gvel, gbytes);
  alloc(&ctx->gforce, gbytes);
  alloc(&ctx->status, sizeof(int) * nscenes);
  alloc(&ctx->dmref, sizeof(uint32_t) * nscenes);
  alloc(&ctx->d_scenes, sizeof(Scene) * nscenes);
  alloc(&ctx->blk_scene, sizeof(int) * nb);
  if (!ok) {
    ctx->err = "device allocation failed";
    return MPM_ERR_NOMEM;
  }
  MPM_HIP(hipMemsetAsync(ctx->P, 0, pbytes, ctx->stream));
  MPM_HIP(hipMemsetAsync(ctx->P2, 0, pbytes, ctx->stream));
  launch_iota(ctx->pid, ctx->n, ctx->stream);
  for (int s = 0; s < nscenes; ++s)
    launch_fill_int(ctx->blk_scene + ctx->scenes[s].blk0, ctx->scenes[s].nblocks, s, ctx->stream);
  launch_fill_int(ctx->gflag, ctx->nblocks, 0, ctx->stream);
  launch_fill_int(ctx->gcnt, 4, 0, ctx->stream);
  launch_fill_int(ctx->dstep, 1, 1, ctx->stream);
  launch_fill_int(ctx->bctr, 1, 0, ctx->stream);
  launch_fill_int(ctx->nblk, 1, 0, ctx->stream);
  launch_clear(reinterpret_cast<float4*>(ctx->gacc), 2 * ctx->nnodes, ctx->stream);
  launch_clear(ctx->gvel, ctx->nnodes, ctx->stream);
  launch_clear(ctx->gforce, ctx->nnodes, ctx->stream);
  MPM_HIP(hipMemsetAsync(ctx->status, 0, sizeof(int) * nscenes, ctx->stream));
  if (int rc = upload_scene_table(ctx)) return rc;
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  ctx->gacc_clean = true;
  ctx->P_first = ctx->P;
  return MPM_OK;
}

int mpm_create(const mpm_params* p, int32_t n, mpm_ctx** out) {
  if (!p || !out || n < 0) return MPM_ERR_ARG;
  return mpm_create_batch(p, &n, 1, out);
}

int mpm_destroy(mpm_ctx* ctx) {
  if (!ctx) return MPM_ERR_ARG;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (hipGraphExec_t& g : ctx->gexec)
    if (g) (void)hipGraphExecDestroy(g);
  void* bufs[] = {ctx->P,       ctx->P2,      ctx->pid,      ctx->pid2,  ctx->keys,     ctx->vals,     ctx->keys2,
                  ctx->vals2,   ctx->hist,    ctx->blk_end, ctx->blist,   ctx->nblk,
                  ctx->gflag,   ctx->glist[0], ctx->glist[1], ctx->gcnt,  ctx->dstep,    ctx->gacc,     ctx->gvel,
                  ctx->gforce,  ctx->status,  ctx->dcount,  ctx->d_scenes, ctx->blk_scene, ctx->bctr,
                  ctx->dmref};
  for (void* b : bufs) (void)hipFree(b);
  for (DevCollider* c : ctx->scoll) (void)hipFree(c);
  for (auto& k : ctx->adj_ckpt) {
    void* b[] = {k.rows, k.pid, k.blist, k.nblk};
    for (void* q : b) (void)hipFree(q);
  }
  (void)hipFree(ctx->adj_const);
  (void)hipFree(ctx->adj_inv);
  for (auto& g : ctx->adj_segs) {
    void* b[] = {g.st, g.gs, g.gl, g.sv, g.pid, g.blist, g.nblk};
    for (void* q : b) (void)hipFree(q);
  }
  {
    const mpm_ctx::AdjWork& w = ctx->adj_ws;
    void* b[] = {w.Gc, w.Gs, w.Gs2, w.gEs, w.gEc, w.tmp, w.gnode, w.vnode, w.gmap, w.seg, w.gsave, w.glist, w.svd};
    for (void* q : b) (void)hipFree(q);
  }
  for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
  for (hipEvent_t e : ctx->ev_region)
    if (e) (void)hipEventDestroy(e);
  if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
  for (hipEvent_t e : {ctx->ev_gathered, ctx->ev_copied})
    if (e) (void)hipEventDestroy(e);
  (void)hipFree(ctx->pos_stage);
  if (ctx->pos_host) (void)hipHostFree(ctx->pos_host);
  if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return MPM_OK;
}

const char* mpm_last_error(const mpm_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }
int32_t mpm_num_particles(const mpm_ctx* ctx) { return ctx ? ctx->n : 0; }
int32_t mpm_num_scenes(const mpm_ctx* ctx) { return ctx ? ctx->nscenes : 0; }


int mpm_scene_set_particles(mpm_ctx* ctx, int32_t scene, const float* x, const float* v, const float* m,
                            const float* E0) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  if (int rc = valid_scene(ctx, scene)) return rc;
  const Scene& S = ctx->scenes[scene];
  if (S.np && (!x || !v || !m)) {
    ctx->err = "x, v and m are required";
    return MPM_ERR_ARG;
  }
  if (int rc = check_positions(ctx, scene, x)) return rc;
  // caller order again inside this scene's slots; every scene re-bins next
  int rc;
  if ((rc = reset_scene_order(ctx, scene))) return rc;
  if ((rc = upload_field(ctx, scene, FX, x, 3)) || (rc = upload_field(ctx, scene, FV, v, 3)) ||
      (rc = upload_field(ctx, scene, FM, m, 1)))
    return rc;
  std::vector<float> ident((size_t)S.np * 9, 0.0f), zeros((size_t)S.np * 9, 0.0f), e((size_t)S.np);
  for (int p = 0; p < S.np; ++p) {
    ident[9 * p + 0] = ident[9 * p + 4] = ident[9 * p + 8] = 1.0f;
    e[p] = E0 ? E0[p] : S.E;
  }
  if ((rc = upload_field(ctx, scene, FB, zeros.data(), 9)) || (rc = upload_field(ctx, scene, FFE, ident.data(), 9)) ||
      (rc = upload_field(ctx, scene, FFP, ident.data(), 9)) || (rc = upload_field(ctx, scene, FVOL, zeros.data(), 1)) ||
      (rc = upload_field(ctx, scene, FE0, e.data(), 1)))
    return rc;
  ctx->scene_set[scene] = 1;
  ctx->quanta_stale = true;
  ctx->particles_set = true;
  for (char c : ctx->scene_set) ctx->particles_set = ctx->particles_set && c;
  return MPM_OK;
}

int mpm_set_particles(mpm_ctx* ctx, const float* x, const float* v, const float* m, const float* E0) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  for (int s = 0; s < ctx->nscenes; ++s) {
    const int p0 = ctx->scenes[s].p0;
    if (int rc = mpm_scene_set_particles(ctx, s, x ? x + 3 * (int64_t)p0 : nullptr, v ? v + 3 * (int64_t)p0 : nullptr,
                                         m ? m + p0 : nullptr, E0 ? E0 + p0 : nullptr))
      return rc;
  }
  return MPM_OK;
}

int mpm_scene_set_state(mpm_ctx* ctx, int32_t scene, const float* B, const float* FE, const float* FP,
                        const float* V0) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  if (int rc = valid_scene(ctx, scene)) return rc;
  int rc;
  if ((rc = upload_field(ctx, scene, FB, B, 9)) || (rc = upload_field(ctx, scene, FFE, FE, 9)) ||
      (rc = upload_field(ctx, scene, FFP, FP, 9)) || (rc = upload_field(ctx, scene, FVOL, V0, 1)))
    return rc;
  return MPM_OK;
}

int mpm_set_state(mpm_ctx* ctx, const float* B, const float* FE, const float* FP, const float* V0) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  for (int s = 0; s < ctx->nscenes; ++s) {
    const int64_t p0 = ctx->scenes[s].p0;
    if (int rc = mpm_scene_set_state(ctx, s, B ? B + 9 * p0 : nullptr, FE ? FE + 9 * p0 : nullptr,
                                     FP ? FP + 9 * p0 : nullptr, V0 ? V0 + p0 : nullptr))
      return rc;
  }
  return MPM_OK;
}

int mpm_scene_set_colliders(mpm_ctx* ctx, int32_t scene, const mpm_collider* c, int32_t k) {
  if (check_ctx(ctx) || k < 0 || (k > 0 && !c)) return MPM_ERR_ARG;
  if (int rc = valid_scene(ctx, scene)) return rc;
  std::vector<DevCollider> host((size_t)k);
  for (int i = 0; i < k; ++i) {
    host_collider_inverse(c[i].translation, c[i].rotation, host[i].inv);
    for (int a = 0; a < 3; ++a) {
      host[i].s[a] = c[i].scale[a];
      host[i].vel[a] = c[i].velocity[a];
    }
  }
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  DevCollider*& dc = ctx->scoll[scene];
  (void)hipFree(dc);
  dc = nullptr;
  if (k) {
    MPM_HIP(hipMalloc(&dc, sizeof(DevCollider) * k));
    MPM_HIP(hipMemcpy(dc, host.data(), sizeof(DevCollider) * k, hipMemcpyHostToDevice));
  }
  ctx->scenes[scene].coll = dc;
  ctx->scenes[scene].ncoll = k;
  ctx->hcoll[scene].assign(c, c + k);
  return upload_scene_table(ctx);
}

int mpm_set_colliders(mpm_ctx* ctx, const mpm_collider* c, int32_t k) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  for (int s = 0; s < ctx->nscenes; ++s)
    if (int rc = mpm_scene_set_colliders(ctx, s, c, k)) return rc;
  return MPM_OK;
}

int mpm_stage(mpm_ctx* ctx, int32_t stage, float dt) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  if (!ctx->particles_set) {
    ctx->err = "no particles set";
    return MPM_ERR_STATE;
  }
  if (ctx->nscenes != 1) {
    ctx->err = "the per-stage calls run single-scene contexts only";
    return MPM_ERR_STATE;
  }
  const KArgs a = make_args(ctx, dt);
  hipStream_t s = ctx->stream;
  switch (stage) {
    case MPM_STAGE_P2G:
      launch_clear(reinterpret_cast<float4*>(ctx->gacc), 2 * ctx->nnodes, s);
      launch_clear(ctx->gforce, ctx->nnodes, s);
      launch_stage_p2g(a, s);
      launch_stage_normalise(a, s);
      ctx->gacc_clean = false;
      break;
    case MPM_STAGE_VOLUMES: launch_stage_volumes(a, s); break;
    case MPM_STAGE_FORCES: launch_stage_forces(a, s); break;
    case MPM_STAGE_GRID: launch_stage_grid(a, s); break;
    case MPM_STAGE_COLLIDE: launch_stage_collide(a, s); break;
    case MPM_STAGE_FUPDATE: launch_stage_fupdate(a, s); break;
    case MPM_STAGE_G2P: launch_stage_g2p(a, s); break;
    case MPM_STAGE_ADVECT: launch_stage_advect(a, s); break;
    default: ctx->err = "unknown stage"; return MPM_ERR_ARG;
  }
  MPM_HIP(hipGetLastError());
  return read_status(ctx);
}

}  // extern "C"

namespace {

BlockArgs block_args(mpm_ctx* ctx, float dt) {
  BlockArgs ba{};
  ba.a = make_args(ctx, dt);
  ba.blist = ctx->blist;
  ba.nblk = ctx->nblk;
  ba.gflag = ctx->gflag;
  ba.glist[0] = ctx->glist[0];
  ba.glist[1] = ctx->glist[1];
  ba.gcnt = ctx->gcnt;
  ba.dstep = ctx->dstep;
  ba.advance = 0;
  ba.nblocks = ctx->nblocks;
  ba.grid = ctx->block_grid;
  ba.rec = nullptr;
  ba.rec_svd = nullptr;
  ba.grec = GridRec{nullptr, nullptr, nullptr};
  ba.bctr = ctx->bctr;
  return ba;
}

BinArgs bin_args(mpm_ctx* ctx) {
  BinArgs b{};
  b.P = ctx->P;
  b.Pout = ctx->P2;
  b.pid = ctx->pid;
  b.pid_out = ctx->pid2;
  b.stride = ctx->stride;
  b.n = ctx->n;
  b.nscenes = ctx->nscenes;
  b.scenes = ctx->d_scenes;
  b.blk_scene = ctx->blk_scene;
  b.nnodes = ctx->nnodes;
  b.keys = ctx->keys;
  b.vals = ctx->vals;
  b.keys2 = ctx->keys2;
  b.vals2 = ctx->vals2;
  b.hist = ctx->hist;
  b.blk_end = ctx->blk_end;
  b.blist = ctx->blist;
  b.nblk = ctx->nblk;
  return b;
}

// Sort, reorder into ctx->P2, swap.
int rebin(mpm_ctx* ctx) {
  const BinArgs b = bin_args(ctx);
  int rc = timed(ctx, MPM_KERNEL_BIN, [&] { launch_bin(b, ctx->stream); });
  if (rc) return rc;
  std::swap(ctx->P, ctx->P2);
  std::swap(ctx->pid, ctx->pid2);
  ctx->since_sort = 0;
  return MPM_OK;
}

// Grid accumulators, velocities and block lists back to their clean state
// (after the legacy stages or the initial-volume pass used them).
void reset_grid(mpm_ctx* ctx) {
  hipStream_t s = ctx->stream;
  launch_clear(reinterpret_cast<float4*>(ctx->gacc), 2 * ctx->nnodes, s);
  launch_clear(ctx->gvel, ctx->nnodes, s);
  launch_fill_int(ctx->gflag, ctx->nblocks, 0, s);
  launch_fill_int(ctx->gcnt, 4, 0, s);
  launch_fill_int(ctx->dstep, 1, 1, s);
  ctx->gacc_clean = true;
}

// One substep: re-bin when due, then the block pass (PASS_P2G on the first
// substep of a call, PASS_FUSED after it), which advances the epoch at its
// end, and the grid pass.
int substep(mpm_ctx* ctx, float dt, bool first) {
  int rc;
  if (ctx->since_sort < 0 || ctx->since_sort >= ctx->resort_every)
    if ((rc = rebin(ctx))) return rc;
  BlockArgs ba = block_args(ctx, dt);
  ba.advance = 1;
  hipStream_t s = ctx->stream;
  if ((rc = timed(ctx, first ? MPM_KERNEL_P2G : MPM_KERNEL_FUSED,
                  [&] { launch_block_pass(ba, first ? PASS_P2G : PASS_FUSED, s); })) ||
      (rc = timed(ctx, MPM_KERNEL_GRID, [&] { launch_grid_active(ba, s); })))
    return rc;
  ++ctx->since_sort;
  return MPM_OK;
}

// Drops the captured segment graphs (they bake in the kernel arguments).
int drop_graphs(mpm_ctx* ctx) {
  for (hipGraphExec_t& g : ctx->gexec)
    if (g) {
      MPM_HIP(hipGraphExecDestroy(g));
      g = nullptr;
    }
  return MPM_OK;
}

// Reproducible mode: each scene's quanta from its largest particle mass
// m_ref, after masses changed.  2^e is the power of two just above m_ref,
// and the mass quantum is 2^(e - 40): node sums stay exact up to 2^13 m_ref
// of mass and 2^15 m_ref m/s of momentum (quantum 2^(e - 38)), and a
// particle's largest contribution keeps 40 bits.
int ensure_quanta(mpm_ctx* ctx) {
  if (!ctx->exact || !ctx->quanta_stale) return MPM_OK;
  int max_np = 0;
  for (const Scene& S : ctx->scenes) max_np = std::max(max_np, S.np);
  launch_scene_mass_max(ctx->P, ctx->stride, ctx->d_scenes, ctx->nscenes, max_np, ctx->dmref, ctx->stream);
  MPM_HIP(hipGetLastError());
  std::vector<uint32_t> bits((size_t)ctx->nscenes);
  MPM_HIP(hipMemcpyAsync(bits.data(), ctx->dmref, sizeof(uint32_t) * bits.size(), hipMemcpyDeviceToHost, ctx->stream));
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  for (int s = 0; s < ctx->nscenes; ++s) {
    float m;
    std::memcpy(&m, &bits[s], sizeof m);
    // no mass (nothing scatters) or a non-finite one (flagged by the pass):
    // any quantum will do
    const int e = (m > 0.0f && std::isfinite(m)) ? std::ilogb(m) + 1 : 0;
    ctx->scenes[s].qexp = e - 40;
  }
  if (int rc = upload_scene_table(ctx)) return rc;
  ctx->quanta_stale = false;
  return drop_graphs(ctx);
}

// resort_every fused substeps that start with a re-bin, replayed from a HIP
// graph.  The launches bake in the particle buffers (which alternate with
// every re-bin, hence one graph per buffer parity), dt, the segment length
// and the block-pass width; scene data (colliders included), the substep
// epoch and the block lists are device state, so one graph serves every
// segment.
int graph_segment(mpm_ctx* ctx, float dt) {
  const int par = ctx->P == ctx->P_first ? 0 : 1;
  const int key = ctx->resort_every * 65536 + ctx->block_grid;
  hipGraphExec_t& g = ctx->gexec[par];
  if (g && (ctx->graph_dt[par] != dt || ctx->graph_key[par] != key)) {
    MPM_HIP(hipGraphExecDestroy(g));
    g = nullptr;
  }
  if (!g) {
    float *const P0 = ctx->P, *const P20 = ctx->P2;
    int *const pid0 = ctx->pid, *const pid20 = ctx->pid2;
    const int since0 = ctx->since_sort;
    hipGraph_t graph = nullptr;
    MPM_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    int rc = MPM_OK;
    for (int i = 0; i < ctx->resort_every && !rc; ++i) rc = substep(ctx, dt, false);
    const hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
    // capture executed nothing: restore the host view of the state
    ctx->P = P0;
    ctx->P2 = P20;
    ctx->pid = pid0;
    ctx->pid2 = pid20;
    ctx->since_sort = since0;
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    MPM_HIP(ce);
    const hipError_t ie = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    MPM_HIP(ie);
    ctx->graph_dt[par] = dt;
    ctx->graph_key[par] = key;
  }
  MPM_HIP(hipGraphLaunch(g, ctx->stream));
  std::swap(ctx->P, ctx->P2);
  std::swap(ctx->pid, ctx->pid2);
  ctx->since_sort = ctx->resort_every;
  return MPM_OK;
}

}  // namespace

extern "C" {

// rasterizeParticlesToGrid + computeParticleVolumesAndDensities for every
// scene on the device: bin, scatter the masses through the block pass,
// node mass, then each particle's density gathered with its weights.
int mpm_init_volumes(mpm_ctx* ctx) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  if (!ctx->particles_set) {
    ctx->err = "no particles set";
    return MPM_ERR_STATE;
  }
  if (ctx->n == 0) return MPM_OK;
  int rc;
  if ((rc = ensure_quanta(ctx))) return rc;
  if (!ctx->gacc_clean) reset_grid(ctx);
  if ((rc = rebin(ctx))) return rc;
  const BlockArgs ba = block_args(ctx, 0.0f);
  launch_block_pass(ba, PASS_MASS, ctx->stream);
  launch_stage_normalise(ba.a, ctx->stream);
  launch_volumes(ba.a, ctx->stream);
  ctx->gacc_clean = false;
  MPM_HIP(hipGetLastError());
  return read_status(ctx);
}

int mpm_substeps(mpm_ctx* ctx, float dt, int32_t nsteps) {
  if (check_ctx(ctx) || nsteps < 0) return MPM_ERR_ARG;
  if (!ctx->particles_set) {
    ctx->err = "no particles set";
    return MPM_ERR_STATE;
  }
  if (nsteps == 0 || ctx->n == 0) return MPM_OK;
  hipStream_t s = ctx->stream;
  int rc;
  if ((rc = ensure_quanta(ctx))) return rc;
  if (!ctx->gacc_clean) reset_grid(ctx);
  const bool graphs = ctx->use_graphs && !(ctx->profiling && ctx->per_kernel);
  for (int i = 0; i < nsteps;) {
    if (graphs && i > 0 && ctx->since_sort >= ctx->resort_every && nsteps - i >= ctx->resort_every) {
      if ((rc = graph_segment(ctx, dt))) return rc;
      i += ctx->resort_every;
    } else {
      if ((rc = substep(ctx, dt, i == 0))) return rc;
      ++i;
    }
  }
  const BlockArgs ba = block_args(ctx, dt);
  if ((rc = timed(ctx, MPM_KERNEL_G2P, [&] { launch_block_pass(ba, PASS_G2P, s); }))) return rc;
  MPM_HIP(hipGetLastError());
  return MPM_OK;
}

int mpm_use_graphs(mpm_ctx* ctx, int32_t enable) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  ctx->use_graphs = enable != 0;
  return MPM_OK;
}

int mpm_set_reproducible(mpm_ctx* ctx, int32_t enable) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  if ((enable != 0) == ctx->exact) return MPM_OK;
  ctx->exact = enable != 0;
  ctx->quanta_stale = true;
  return drop_graphs(ctx);
}

int mpm_get_reproducible(mpm_ctx* ctx, int32_t* enabled) {
  if (check_ctx(ctx) || !enabled) return MPM_ERR_ARG;
  *enabled = ctx->exact ? 1 : 0;
  return MPM_OK;
}

int mpm_set_tuning(mpm_ctx* ctx, int32_t resort_every, int32_t block_grid) {
  if (check_ctx(ctx) || resort_every < 1 || block_grid < 1) return MPM_ERR_ARG;
  ctx->resort_every = resort_every;
  ctx->block_grid = block_grid;
  ctx->since_sort = -1;  // a new re-bin cycle starts at the next substep
  return MPM_OK;
}

int mpm_get_tuning(mpm_ctx* ctx, int32_t* resort_every, int32_t* block_grid) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  if (resort_every) *resort_every = ctx->resort_every;
  if (block_grid) *block_grid = ctx->block_grid;
  return MPM_OK;
}

// Times `steps` substeps for every candidate (re-bin interval x block-pass
// width) from the same starting state, keeps the fastest, and puts the
// particle state back as it was.  Status flags raised by the candidate runs
// (which run ahead of the caller's state) are read and cleared here, so they
// never surface later as errors of the restored state; a flagged run makes
// the tuning fail with that status and keeps the caller's settings.
int mpm_autotune(mpm_ctx* ctx, float dt, int32_t steps, double* best_ms) {
  if (check_ctx(ctx) || steps < 1) return MPM_ERR_ARG;
  if (!ctx->particles_set) {
    ctx->err = "no particles set";
    return MPM_ERR_STATE;
  }
  if (int rc = read_status(ctx)) return rc;  // the caller's own pending flags first
  const size_t pbytes = sizeof(float) * (size_t)NFIELDS * (size_t)ctx->stride;
  DevBuf<float> snapP;
  DevBuf<int> snapId;
  MPM_HIP(snapP.alloc(pbytes));
  MPM_HIP(snapId.alloc(sizeof(int) * (size_t)std::max(1, ctx->n)));
  MPM_HIP(hipMemcpyAsync(snapP.p, ctx->P, pbytes, hipMemcpyDeviceToDevice, ctx->stream));
  MPM_HIP(hipMemcpyAsync(snapId.p, ctx->pid, sizeof(int) * ctx->n, hipMemcpyDeviceToDevice, ctx->stream));
  float* const snapBuf = ctx->P;  // the buffer the snapshot belongs in
  // puts the snapshot back into the buffer it came from
  auto restore = [&]() -> int {
    if (ctx->P != snapBuf) {
      std::swap(ctx->P, ctx->P2);
      std::swap(ctx->pid, ctx->pid2);
    }
    MPM_HIP(hipMemcpyAsync(ctx->P, snapP.p, pbytes, hipMemcpyDeviceToDevice, ctx->stream));
    MPM_HIP(hipMemcpyAsync(ctx->pid, snapId.p, sizeof(int) * ctx->n, hipMemcpyDeviceToDevice, ctx->stream));
    ctx->since_sort = -1;
    ctx->gacc_clean = false;
    return MPM_OK;
  };
  EventGuard e0, e1;
  MPM_HIP(hipEventCreate(&e0.e));
  MPM_HIP(hipEventCreate(&e1.e));
  const int resorts[] = {8, 16, 32, 64};
  const int64_t wmax = std::max<int64_t>(1, ctx->nblocks);
  const int widths[] = {(int)std::min<int64_t>(wmax, 1024), (int)std::min<int64_t>(wmax, 2048),
                        (int)std::min<int64_t>(wmax, 4096)};
  const int keep_r = ctx->resort_every, keep_w = ctx->block_grid;
  // every candidate is timed over whole re-bin cycles of every interval, and
  // over two of the longest, so the re-bin cost and the drift out of the
  // halo margins between re-bins are both counted
  const int cycle = 64;
  const int window = (std::max(steps, 2 * cycle) + cycle - 1) / cycle * cycle;
  double best = 1e300;
  int best_r = keep_r, best_w = keep_w;
  auto candidates = [&]() -> int {
    for (int r : resorts)
      for (int w : widths) {
        if (int rc = restore()) return rc;
        ctx->resort_every = r;
        ctx->block_grid = w;
        MPM_HIP(hipEventRecord(e0.e, ctx->stream));
        if (int rc = mpm_substeps(ctx, dt, window)) return rc;
        MPM_HIP(hipEventRecord(e1.e, ctx->stream));
        MPM_HIP(hipEventSynchronize(e1.e));
        float ms = 0.0f;
        MPM_HIP(hipEventElapsedTime(&ms, e0.e, e1.e));
        if (ms < best) {
          best = ms;
          best_r = r;
          best_w = w;
        }
      }
    return MPM_OK;
  };
  int rc = candidates();
  // the caller's state back, whatever happened above
  const int rrc = restore();
  if (!rc) rc = rrc;
  if (!rc) MPM_HIP(hipStreamSynchronize(ctx->stream));
  const int src = read_status(ctx, nullptr, true);  // flags of the candidate runs, cleared
  if (!rc && src) {
    rc = src;
    ctx->err = "mpm_autotune: a candidate run raised a status flag (" + ctx->err +
               "); tuning left unchanged, particle state restored";
  }
  if (rc) {
    ctx->resort_every = keep_r;
    ctx->block_grid = keep_w;
    return rc;
  }
  ctx->resort_every = best_r;
  ctx->block_grid = best_w;
  if (best_ms) *best_ms = best / window;
  return MPM_OK;
}

int mpm_profile_begin(mpm_ctx* ctx, int32_t per_kernel) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  for (hipEvent_t& e : ctx->ev_region)
    if (!e) MPM_HIP(hipEventCreate(&e));
  ctx->profiling = true;
  ctx->per_kernel = per_kernel != 0;
  ctx->events_used = 0;
  ctx->event_kernel.clear();
  MPM_HIP(hipEventRecord(ctx->ev_region[0], ctx->stream));
  return MPM_OK;
}

int mpm_profile_end(mpm_ctx* ctx, double* kernel_ms, int32_t* launches, double* total_ms) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  if (!ctx->profiling) {
    ctx->err = "mpm_profile_end without mpm_profile_begin";
    return MPM_ERR_STATE;
  }
  MPM_HIP(hipEventRecord(ctx->ev_region[1], ctx->stream));
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  ctx->profiling = false;
  for (int k = 0; k < MPM_NKERNELS; ++k) {
    if (kernel_ms) kernel_ms[k] = 0.0;
    if (launches) launches[k] = 0;
  }
  for (size_t i = 0; i < ctx->event_kernel.size(); ++i) {
    float ms = 0.0f;
    MPM_HIP(hipEventElapsedTime(&ms, ctx->events[2 * i], ctx->events[2 * i + 1]));
    const int k = ctx->event_kernel[i];
    if (kernel_ms) kernel_ms[k] += ms;
    if (launches) launches[k] += 1;
  }
  if (total_ms) {
    float ms = 0.0f;
    MPM_HIP(hipEventElapsedTime(&ms, ctx->ev_region[0], ctx->ev_region[1]));
    *total_ms = ms;
  }
  return MPM_OK;
}

int mpm_count_active(mpm_ctx* ctx, int64_t* active_nodes) {
  if (check_ctx(ctx) || !active_nodes) return MPM_ERR_ARG;
  if (!ctx->dcount) MPM_HIP(hipMalloc(&ctx->dcount, sizeof(int64_t)));
  MPM_HIP(hipMemsetAsync(ctx->dcount, 0, sizeof(int64_t), ctx->stream));
  launch_count_active(make_args(ctx, 0.0f), ctx->dcount, ctx->stream);
  MPM_HIP(hipMemcpyAsync(active_nodes, ctx->dcount, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  return MPM_OK;
}

int mpm_sync(mpm_ctx* ctx) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  return read_status(ctx);
}

int mpm_scene_status(mpm_ctx* ctx, int32_t* status) {
  if (check_ctx(ctx) || !status) return MPM_ERR_ARG;
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<int> st;
  const int rc = read_status(ctx, &st, true);
  for (int s = 0; s < ctx->nscenes; ++s) status[s] = status_code(st[s]);
  return rc == MPM_ERR_HIP ? rc : MPM_OK;
}

int mpm_scene_get_particles(mpm_ctx* ctx, int32_t scene, float* x, float* v, float* B, float* FE, float* FP,
                            float* V0) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  if (int rc = valid_scene(ctx, scene)) return rc;
  int rc;
  if ((rc = download_field(ctx, scene, FX, x, 3)) || (rc = download_field(ctx, scene, FV, v, 3)) ||
      (rc = download_field(ctx, scene, FB, B, 9)) || (rc = download_field(ctx, scene, FFE, FE, 9)) ||
      (rc = download_field(ctx, scene, FFP, FP, 9)) || (rc = download_field(ctx, scene, FVOL, V0, 1)))
    return rc;
  return MPM_OK;
}

int mpm_get_particles(mpm_ctx* ctx, float* x, float* v, float* B, float* FE, float* FP, float* V0) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  for (int s = 0; s < ctx->nscenes; ++s) {
    const int64_t p0 = ctx->scenes[s].p0;
    if (int rc = mpm_scene_get_particles(ctx, s, x ? x + 3 * p0 : nullptr, v ? v + 3 * p0 : nullptr,
                                         B ? B + 9 * p0 : nullptr, FE ? FE + 9 * p0 : nullptr,
                                         FP ? FP + 9 * p0 : nullptr, V0 ? V0 + p0 : nullptr))
      return rc;
  }
  return read_status(ctx);
}

int mpm_scene_get_grid(mpm_ctx* ctx, int32_t scene, float* m, float* vel, float* force, uint8_t* active) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  if (int rc = valid_scene(ctx, scene)) return rc;
  const Scene& S = ctx->scenes[scene];
  const int64_t cells = (int64_t)S.dims[0] * S.dims[1] * S.dims[2];
  DevBuf<float> dm, dv, df;
  DevBuf<uint8_t> da;
  if (m) MPM_HIP(dm.alloc(sizeof(float) * cells));
  if (vel) MPM_HIP(dv.alloc(sizeof(float) * 3 * cells));
  if (force) MPM_HIP(df.alloc(sizeof(float) * 3 * cells));
  if (active) MPM_HIP(da.alloc(cells));
  launch_grid_to_dense(make_args(ctx, 0.0f), S, dm.p, dv.p, df.p, da.p, ctx->stream);
  if (m) MPM_HIP(hipMemcpyAsync(m, dm.p, sizeof(float) * cells, hipMemcpyDeviceToHost, ctx->stream));
  if (vel) MPM_HIP(hipMemcpyAsync(vel, dv.p, sizeof(float) * 3 * cells, hipMemcpyDeviceToHost, ctx->stream));
  if (force) MPM_HIP(hipMemcpyAsync(force, df.p, sizeof(float) * 3 * cells, hipMemcpyDeviceToHost, ctx->stream));
  if (active) MPM_HIP(hipMemcpyAsync(active, da.p, cells, hipMemcpyDeviceToHost, ctx->stream));
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  return MPM_OK;
}

int mpm_get_grid(mpm_ctx* ctx, float* m, float* vel, float* force, uint8_t* active) {
  return mpm_scene_get_grid(ctx, 0, m, vel, force, active);
}

}  // extern "C"


// ---------------------------------------------------------------- reverse mode
namespace {

int adjoint_supported(mpm_ctx* ctx) {
  if (!ctx->particles_set) {
    ctx->err = "no particles set";
    return MPM_ERR_STATE;
  }
  if (ctx->prm.compose != MPM_COMPOSE_TEXTBOOK) {
    ctx->err = "reverse mode needs the TEXTBOOK composition (LITERAL's U^T S V has no derivative at repeated "
               "singular values)";
    return MPM_ERR_STATE;
  }
  for (const Scene& S : ctx->scenes)
    if ((S.bc_mode & MPM_BC_COLLIDERS) && S.ncoll > 0) {
      ctx->err = "reverse mode supports the domain walls, not box colliders";
      return MPM_ERR_STATE;
    }
  return MPM_OK;
}

// A checkpoint: the particle state with the binning it is in.  Saving bins
// first, so the state is in a fresh binning's slot order; restoring puts
// state and binning back, and the substeps that follow start without a
// re-bin.
int save_checkpoint(mpm_ctx* ctx, mpm_ctx::AdjCheckpoint& k) {
  if (!k.rows) {
    bool ok = hipMalloc(&k.rows, sizeof(float) * NFIELDS * ctx->stride) == hipSuccess &&
              hipMalloc(&k.pid, sizeof(int) * (size_t)std::max(1, ctx->n)) == hipSuccess &&
              hipMalloc(&k.blist, sizeof(int4) * (size_t)ctx->nblocks) == hipSuccess &&
              hipMalloc(&k.nblk, sizeof(int)) == hipSuccess;
    if (!ok) {  // all or nothing: a later call allocates again
      void* b[] = {k.rows, k.pid, k.blist, k.nblk};
      for (void* q : b) (void)hipFree(q);
      k = mpm_ctx::AdjCheckpoint{};
      ctx->err = "device memory for the reverse-mode checkpoints is exhausted";
      return MPM_ERR_NOMEM;
    }
  }
  if (ctx->since_sort != 0)
    if (int rc = rebin(ctx)) return rc;
  hipStream_t s = ctx->stream;
  MPM_HIP(hipMemcpyAsync(k.rows, ctx->P, sizeof(float) * NFIELDS * ctx->stride, hipMemcpyDeviceToDevice, s));
  MPM_HIP(hipMemcpyAsync(k.pid, ctx->pid, sizeof(int) * ctx->n, hipMemcpyDeviceToDevice, s));
  MPM_HIP(hipMemcpyAsync(k.blist, ctx->blist, sizeof(int4) * ctx->nblocks, hipMemcpyDeviceToDevice, s));
  MPM_HIP(hipMemcpyAsync(k.nblk, ctx->nblk, sizeof(int), hipMemcpyDeviceToDevice, s));
  return MPM_OK;
}

void free_checkpoints(mpm_ctx* ctx, size_t keep) {
  while (ctx->adj_ckpt.size() > keep) {
    const mpm_ctx::AdjCheckpoint& k = ctx->adj_ckpt.back();
    void* b[] = {k.rows, k.pid, k.blist, k.nblk};
    for (void* q : b) (void)hipFree(q);
    ctx->adj_ckpt.pop_back();
  }
}

int restore_checkpoint(mpm_ctx* ctx, const mpm_ctx::AdjCheckpoint& k) {
  hipStream_t s = ctx->stream;
  MPM_HIP(hipMemcpyAsync(ctx->P, k.rows, sizeof(float) * NFIELDS * ctx->stride, hipMemcpyDeviceToDevice, s));
  MPM_HIP(hipMemcpyAsync(ctx->pid, k.pid, sizeof(int) * ctx->n, hipMemcpyDeviceToDevice, s));
  MPM_HIP(hipMemcpyAsync(ctx->blist, k.blist, sizeof(int4) * ctx->nblocks, hipMemcpyDeviceToDevice, s));
  MPM_HIP(hipMemcpyAsync(ctx->nblk, k.nblk, sizeof(int), hipMemcpyDeviceToDevice, s));
  ctx->since_sort = 0;
  return MPM_OK;
}

// Grid record k of a segment: sums in gs (nnodes float4 per substep, room for
// every grid block), block lists in gl (nblocks + 1 ints per substep).
GridRec grid_record(const mpm_ctx* ctx, float4* gs, int* gl, int* map, int k) {
  return GridRec{gs + (size_t)k * ctx->nnodes, gl + (size_t)k * (ctx->nblocks + 1), map};
}

// Run ns substeps from the current binning on the fused path (the first one
// a P2G pass, then fused passes, then a final G2P pass: mpm_substeps(dt, ns)
// without re-binning), the passes writing each substep's starting state
// (rows 0..32 in slot order, seg + k * 33 rows), the SVD its F update used
// (sv + k * SVD_ROWS rows) and the positions after the last one (seg + ns *
// 33 rows), and the grid pass recording the P2G sums of each substep's
// active grid blocks (grid_record(k)).
int record_segment(mpm_ctx* ctx, float dt, int ns, float* seg, float4* gs, int* gl, float* sv) {
  const int64_t sd = ctx->stride;
  hipStream_t s = ctx->stream;
  if (!ctx->gacc_clean) reset_grid(ctx);
  for (int k = 0; k < ns; ++k) {
    BlockArgs ba = block_args(ctx, dt);
    // the recording passes sum in the default (FP32) format, so the grid
    // pass must read that format in reproducible mode too
    ba.a.exact = 0;
    ba.rec = seg + (size_t)k * NA * sd;
    ba.rec_svd = sv + (size_t)k * SVD_ROWS * sd;
    ba.grec = grid_record(ctx, gs, gl, nullptr, k);
    ba.advance = 1;
    launch_block_pass(ba, k == 0 ? PASS_P2G : PASS_FUSED, s);
    launch_grid_active(ba, s);
    ++ctx->since_sort;
  }
  BlockArgs ba = block_args(ctx, dt);
  ba.a.exact = 0;
  ba.rec = seg + (size_t)ns * NA * sd;
  launch_block_pass(ba, PASS_G2P, s);
  MPM_HIP(hipGetLastError());
  return MPM_OK;
}

void free_segments(mpm_ctx* ctx) {
  for (auto& g : ctx->adj_segs) {
    void* b[] = {g.st, g.gs, g.gl, g.sv, g.pid, g.blist, g.nblk};
    for (void* q : b) (void)hipFree(q);
  }
  ctx->adj_segs.clear();
}

// Adjoint kernels over the ns recorded substeps of one segment, last to
// first; Gs/Gs2 (slot order) swap per substep, Gs holds the result.
// The reverse sweep's workspaces in ctx->adj_ws, allocated on first use and
// kept for later calls; the segment buffers grow to `every` substeps.
int adjoint_workspace(mpm_ctx* ctx, int every) {
  mpm_ctx::AdjWork& w = ctx->adj_ws;
  const int64_t sd = ctx->stride;
  const size_t rows = sizeof(float) * NA * sd;
  bool ok = true;
  auto alloc = [&](auto** ptr, size_t bytes) {
    if (ok && !*ptr && hipMalloc(reinterpret_cast<void**>(ptr), bytes) != hipSuccess) {
      *ptr = nullptr;
      ok = false;
    }
  };
  alloc(&w.Gc, rows);
  alloc(&w.Gs, rows);
  alloc(&w.Gs2, rows);
  alloc(&w.gEs, sizeof(float) * sd);
  alloc(&w.gEc, sizeof(float) * sd);
  alloc(&w.tmp, sizeof(float) * 3 * (size_t)std::max(1, ctx->n));
  alloc(&w.gnode, sizeof(float4) * (size_t)ctx->nnodes);
  alloc(&w.vnode, sizeof(float4) * (size_t)ctx->nnodes);
  alloc(&w.gmap, sizeof(int) * (size_t)ctx->nblocks);
  if (every > 1 && every > w.seg_every) {
    // states of a segment's substeps plus the positions after it; the grid
    // records of each substep
    void* b[] = {w.seg, w.gsave, w.glist, w.svd};
    for (void* q : b) (void)hipFree(q);
    w.seg = nullptr;
    w.gsave = nullptr;
    w.glist = nullptr;
    w.svd = nullptr;
    w.seg_every = 0;
    alloc(&w.seg, rows * (size_t)(every + 1));
    alloc(&w.gsave, sizeof(float4) * (size_t)ctx->nnodes * every);
    alloc(&w.glist, sizeof(int) * (size_t)(ctx->nblocks + 1) * every);
    alloc(&w.svd, sizeof(float) * SVD_ROWS * sd * (size_t)every);
    if (ok) w.seg_every = every;
  }
  if (!ok) {
    ctx->err = "device memory for the reverse sweep is exhausted";
    return MPM_ERR_NOMEM;
  }
  return MPM_OK;
}

// gnode (zero on entry and exit outside the blocks being reversed) holds the
// node adjoints; map: one int per grid block, -1 on entry and exit.
void adjoint_segment(mpm_ctx* ctx, float dt, int ns, const float* seg, float4* gs, int* gl, const float* sv, int* map,
                     float*& Gs, float*& Gs2, float4* gnode, float4* vnode, float* gEs) {
  const int64_t sd = ctx->stride;
  hipStream_t s = ctx->stream;
  for (int k = ns - 1; k >= 0; --k) {
    const float* St = seg + (size_t)k * NA * sd;
    const float* Xn = seg + (size_t)(k + 1) * NA * sd;
    BlockArgs ba = block_args(ctx, dt);
    ba.grec = grid_record(ctx, gs, gl, map, k);
    launch_adj_map(ba.a, ba.grec, 1, gnode, vnode, s);
    launch_adj_g2p_block(ba, St, Xn, Gs, Gs2, gnode, vnode, s);
    launch_adj_grid(ba.a, ba.grec, gnode, s);
    const float* Sv = sv + (size_t)k * SVD_ROWS * sd;
    launch_adj_p2g(ba.a, St, Gs, Gs2, gnode, gEs, Sv, k > 0 ? Sv - (size_t)SVD_ROWS * sd : nullptr, s);
    launch_adj_map(ba.a, ba.grec, 0, gnode, vnode, s);
    std::swap(Gs, Gs2);
  }
}

}  // namespace

extern "C" {

int mpm_adjoint_forward(mpm_ctx* ctx, float dt, int32_t nsteps, int32_t every) {
  if (check_ctx(ctx) || nsteps < 1 || every < 0) return MPM_ERR_ARG;
  if (int rc = adjoint_supported(ctx)) return rc;
  const int64_t sd = ctx->stride;
  ctx->adj_steps = 0;
  // checkpoint buffers and per-substep records are kept from the previous
  // call and reused when the shape matches
  const int L = std::max(1, ctx->resort_every);
  std::vector<int> seg_ns;
  for (int done = 0; done < nsteps; done += L) seg_ns.push_back(std::min(L, nsteps - done));
  bool same = seg_ns.size() == ctx->adj_segs.size();
  for (size_t i = 0; same && i < seg_ns.size(); ++i) same = ctx->adj_segs[i].ns == seg_ns[i];
  if (every == 0) {
    // automatic: keep every substep when its records fit in half of the free
    // device memory (or are already held for this shape), else checkpoints
    // every 20 substeps
    size_t free_b = 0, total_b = 0;
    const double per_step = (double)sizeof(float) * (NA + SVD_ROWS) * sd + (double)sizeof(float4) * ctx->nnodes +
                            (double)sizeof(int) * (ctx->nblocks + 1);
    const double need = per_step * (nsteps + (double)seg_ns.size()) + (double)sizeof(float) * NFIELDS * sd +
                        (double)sizeof(int) * ctx->n + (double)sizeof(int4) * ctx->nblocks;  // + the checkpoint
    every = (same && !ctx->adj_segs.empty()) ||
                    (hipMemGetInfo(&free_b, &total_b) == hipSuccess && need < 0.5 * (double)free_b)
                ? 1
                : 20;
  }
  if (every != 1) seg_ns.clear();
  same = same && every == 1;
  if (every == 1 && !same) free_segments(ctx);  // other calls leave them for the next keep-all call
  if (!ctx->adj_const) MPM_HIP(hipMalloc(&ctx->adj_const, sizeof(float) * 3 * sd));
  if (!ctx->adj_inv) MPM_HIP(hipMalloc(&ctx->adj_inv, sizeof(int) * (size_t)std::max(1, ctx->n)));
  launch_rows_to_caller(ctx->adj_const, ctx->P + (int64_t)FM * sd, ctx->pid, ctx->adj_inv, sd, ctx->n, 3, ctx->stream);
  const int nck = every == 1 ? 1 : (nsteps + every - 1) / every;
  free_checkpoints(ctx, (size_t)nck);
  ctx->adj_ckpt.resize((size_t)nck);
  {
    // with checkpoints, each segment runs from its checkpoint's binning
    // without re-binning, as the reverse sweep re-runs it
    const int keep_resort = ctx->resort_every;
    if (every > 1) ctx->resort_every = std::max(keep_resort, every);
    auto restore_resort = on_scope_exit([ctx, keep_resort] { ctx->resort_every = keep_resort; });
    for (int c = 0; c < nck; ++c) {
      if (int rc = save_checkpoint(ctx, ctx->adj_ckpt[c])) return rc;
      if (every > 1)
        if (int rc = mpm_substeps(ctx, dt, std::min(every, nsteps - c * every))) return rc;
    }
  }
  if (every == 1) {
    // keep every substep: segments of resort_every substeps, one binning each
    const size_t rows = sizeof(float) * NA * sd;
    for (size_t i = 0; i < seg_ns.size(); ++i) {
      if (i == ctx->adj_segs.size()) {
        mpm_ctx::AdjSegment g;
        g.ns = seg_ns[i];
        bool ok = hipMalloc(&g.st, rows * (size_t)(g.ns + 1)) == hipSuccess &&
                  hipMalloc(&g.gs, sizeof(float4) * (size_t)ctx->nnodes * g.ns) == hipSuccess &&
                  hipMalloc(&g.gl, sizeof(int) * (size_t)(ctx->nblocks + 1) * g.ns) == hipSuccess &&
                  hipMalloc(&g.sv, sizeof(float) * SVD_ROWS * sd * (size_t)g.ns) == hipSuccess &&
                  hipMalloc(&g.pid, sizeof(int) * (size_t)std::max(1, ctx->n)) == hipSuccess &&
                  hipMalloc(&g.blist, sizeof(int4) * (size_t)ctx->nblocks) == hipSuccess &&
                  hipMalloc(&g.nblk, sizeof(int)) == hipSuccess;
        ctx->adj_segs.push_back(g);
        if (!ok) {
          free_segments(ctx);
          ctx->err = "device memory for the per-substep records (checkpoint_every = 1) is exhausted";
          return MPM_ERR_NOMEM;
        }
      }
      const mpm_ctx::AdjSegment& g = ctx->adj_segs[i];
      if (ctx->since_sort != 0)  // the first segment starts in the checkpoint's binning
        if (int rc = rebin(ctx)) return rc;
      MPM_HIP(hipMemcpyAsync(g.pid, ctx->pid, sizeof(int) * ctx->n, hipMemcpyDeviceToDevice, ctx->stream));
      MPM_HIP(hipMemcpyAsync(g.blist, ctx->blist, sizeof(int4) * ctx->nblocks, hipMemcpyDeviceToDevice, ctx->stream));
      MPM_HIP(hipMemcpyAsync(g.nblk, ctx->nblk, sizeof(int), hipMemcpyDeviceToDevice, ctx->stream));
      if (int rc = record_segment(ctx, dt, g.ns, g.st, g.gs, g.gl, g.sv)) return rc;
    }
    ctx->gacc_clean = false;
  }
  ctx->adj_dt = dt;
  ctx->adj_steps = nsteps;
  ctx->adj_every = every;
  MPM_HIP(hipGetLastError());
  return MPM_OK;
}

int mpm_adjoint_interval(mpm_ctx* ctx, int32_t* every) {
  if (check_ctx(ctx) || !every) return MPM_ERR_ARG;
  *every = ctx->adj_steps > 0 ? ctx->adj_every : 0;
  return MPM_OK;
}

int mpm_adjoint_backward(mpm_ctx* ctx, const float* gx, const float* gv, float* gx0, float* gv0, float* gE0) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  if (ctx->adj_steps < 1) {
    ctx->err = "mpm_adjoint_backward without mpm_adjoint_forward";
    return MPM_ERR_STATE;
  }
  const int n = ctx->n;
  const int64_t sd = ctx->stride;
  const float dt = ctx->adj_dt;
  const int every = ctx->adj_every, steps = ctx->adj_steps;
  hipStream_t s = ctx->stream;
  const size_t rows = sizeof(float) * NA * sd;
  if (int rc = adjoint_workspace(ctx, every)) return rc;
  mpm_ctx::AdjWork& w = ctx->adj_ws;
  float *const Gc = w.Gc, *const gEs = w.gEs, *const gEc = w.gEc, *const seg = w.seg, *const tmp = w.tmp;
  float*& Gs = w.Gs;  // Gs and Gs2 swap per substep
  float*& Gs2 = w.Gs2;
  float4 *const gnode = w.gnode, *const vnode = w.vnode, *const gsave = w.gsave;
  int *const glist = w.glist, *const gmap = w.gmap;
  int rc = MPM_OK;
  MPM_HIP(hipMemsetAsync(Gc, 0, rows, s));
  MPM_HIP(hipMemsetAsync(gEc, 0, sizeof(float) * sd, s));
  launch_clear(gnode, ctx->nnodes, s);
  launch_clear(vnode, ctx->nnodes, s);
  launch_fill_int(gmap, ctx->nblocks, -1, s);
  if (gx) {
    MPM_HIP(hipMemcpyAsync(tmp, gx, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s));
    launch_aos_to_rows(Gc + (int64_t)AX * sd, tmp, n, 3, sd, s);
  }
  if (gv) {
    MPM_HIP(hipStreamSynchronize(s));
    MPM_HIP(hipMemcpyAsync(tmp, gv, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s));
    launch_aos_to_rows(Gc + (int64_t)AV * sd, tmp, n, 3, sd, s);
  }
  // one binning per segment: slots stay fixed inside it; the caller's
  // interval comes back on every exit, early error returns included
  const int keep_resort = ctx->resort_every;
  ctx->resort_every = 1 << 30;
  auto restore_resort = on_scope_exit([ctx, keep_resort] { ctx->resort_every = keep_resort; });
  if (every == 1) {
    // recorded segments: restore each one's slot order and block list
    for (int c = (int)ctx->adj_segs.size() - 1; c >= 0 && !rc; --c) {
      mpm_ctx::AdjSegment& g = ctx->adj_segs[c];
      MPM_HIP(hipMemcpyAsync(ctx->pid, g.pid, sizeof(int) * n, hipMemcpyDeviceToDevice, s));
      MPM_HIP(hipMemcpyAsync(ctx->blist, g.blist, sizeof(int4) * ctx->nblocks, hipMemcpyDeviceToDevice, s));
      MPM_HIP(hipMemcpyAsync(ctx->nblk, g.nblk, sizeof(int), hipMemcpyDeviceToDevice, s));
      // m, V0, E0 in this segment's slot order (k_adj_p2g reads them from P)
      launch_rows_to_slots(ctx->P + (int64_t)FM * sd, ctx->adj_const, ctx->pid, sd, n, 3, s);
      launch_rows_to_slots(Gs, Gc, ctx->pid, sd, n, NA, s);
      MPM_HIP(hipMemsetAsync(gEs, 0, sizeof(float) * sd, s));
      adjoint_segment(ctx, dt, g.ns, g.st, g.gs, g.gl, g.sv, gmap, Gs, Gs2, gnode, vnode, gEs);
      launch_rows_to_caller(Gc, Gs, ctx->pid, ctx->adj_inv, sd, n, NA, s);
      launch_add_to_caller(gEc, gEs, ctx->pid, n, s);
    }
  }
  const int nck = every == 1 ? 0 : (steps + every - 1) / every;
  for (int c = nck - 1; c >= 0 && !rc; --c) {
    const int ns = std::min(every, steps - c * every);
    if ((rc = restore_checkpoint(ctx, ctx->adj_ckpt[c]))) break;
    // re-run the segment, keeping each substep's starting state and P2G sums
    if ((rc = record_segment(ctx, dt, ns, seg, gsave, glist, w.svd))) break;
    launch_rows_to_slots(Gs, Gc, ctx->pid, sd, n, NA, s);
    MPM_HIP(hipMemsetAsync(gEs, 0, sizeof(float) * sd, s));
    adjoint_segment(ctx, dt, ns, seg, gsave, glist, w.svd, gmap, Gs, Gs2, gnode, vnode, gEs);
    launch_rows_to_caller(Gc, Gs, ctx->pid, ctx->adj_inv, sd, n, NA, s);
    launch_add_to_caller(gEc, gEs, ctx->pid, n, s);
    ctx->gacc_clean = false;
  }
  ctx->resort_every = keep_resort;
  if (!rc) rc = restore_checkpoint(ctx, ctx->adj_ckpt[0]);  // leave the initial state
  ctx->gacc_clean = false;
  if (!rc) {
    MPM_HIP(hipGetLastError());
    if (gx0) {
      launch_rows_to_aos(tmp, Gc + (int64_t)AX * sd, n, 3, sd, s);
      MPM_HIP(hipMemcpyAsync(gx0, tmp, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s));
      MPM_HIP(hipStreamSynchronize(s));
    }
    if (gv0) {
      launch_rows_to_aos(tmp, Gc + (int64_t)AV * sd, n, 3, sd, s);
      MPM_HIP(hipMemcpyAsync(gv0, tmp, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s));
      MPM_HIP(hipStreamSynchronize(s));
    }
    if (gE0) MPM_HIP(hipMemcpyAsync(gE0, gEc, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    MPM_HIP(hipStreamSynchronize(s));
    rc = read_status(ctx);
  }
  return rc;
}

}  // extern "C"

extern "C" int mpm_debug_bin(mpm_ctx* ctx, uint32_t* sorted_keys, uint32_t* perm, int32_t* slot_caller_index) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  if (ctx->n == 0) return MPM_OK;
  BinArgs b = bin_args(ctx);
  uint32_t *k = nullptr, *v = nullptr;
  launch_sort_only(b, ctx->stream, &k, &v);
  if (sorted_keys) MPM_HIP(hipMemcpyAsync(sorted_keys, k, sizeof(uint32_t) * ctx->n, hipMemcpyDeviceToHost, ctx->stream));
  if (perm) MPM_HIP(hipMemcpyAsync(perm, v, sizeof(uint32_t) * ctx->n, hipMemcpyDeviceToHost, ctx->stream));
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  if (int rc = rebin(ctx)) return rc;
  if (slot_caller_index)
    MPM_HIP(hipMemcpyAsync(slot_caller_index, ctx->pid, sizeof(int) * ctx->n, hipMemcpyDeviceToHost, ctx->stream));
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  return MPM_OK;
}

extern "C" int mpm_debug_blocks(mpm_ctx* ctx, int32_t* blocks, int32_t max_blocks, int32_t* count) {
  if (check_ctx(ctx) || !count || max_blocks < 0) return MPM_ERR_ARG;
  int nb = 0;
  if (ctx->n > 0 && ctx->since_sort >= 0) {
    MPM_HIP(hipMemcpyAsync(&nb, ctx->nblk, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    MPM_HIP(hipStreamSynchronize(ctx->stream));
  }
  *count = nb;
  if (blocks && nb > 0)
    MPM_HIP(hipMemcpy(blocks, ctx->blist, sizeof(int4) * (size_t)std::min(nb, max_blocks), hipMemcpyDeviceToHost));
  return MPM_OK;
}

namespace {
__global__ void __launch_bounds__(256) k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst,
                                                    int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
}  // namespace

extern "C" int mpm_copy_probe(mpm_ctx* ctx, int64_t bytes) {
  if (check_ctx(ctx) || bytes <= 0 || bytes % 16) return MPM_ERR_ARG;
  float4 *a = nullptr, *b = nullptr;
  MPM_HIP(hipMalloc(&a, bytes));
  MPM_HIP(hipMalloc(&b, bytes));
  MPM_HIP(hipMemsetAsync(a, 0, bytes, ctx->stream));
  const int64_t n = bytes / 16;
  k_copy_probe<<<2048, 256, 0, ctx->stream>>>(a, b, n);
  MPM_HIP(hipGetLastError());
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  MPM_HIP(hipFree(a));
  MPM_HIP(hipFree(b));
  return MPM_OK;
}
