// C ABI of libmpm_hip: context lifetime, particle/grid transfer, the legacy
// stage calls and the fused substep loop.  Declarations: include/mpm_c.h.
#include <cstdio>
#include <cstring>
#include <new>

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
// inside the grid; the reference indexes out of bounds otherwise.
int check_positions(mpm_ctx* ctx, const float* x) {
  for (int p = 0; p < ctx->n; ++p)
    for (int a = 0; a < 3; ++a) {
      const float q = x[3 * p + a] / ctx->h;
      if (!(q >= 2.0f) || !(q < (float)(ctx->dims[a] - 2))) {
        char buf[160];
        snprintf(buf, sizeof buf, "particle %d coordinate %d = %g puts its stencil outside the grid", p, a,
                 (double)x[3 * p + a]);
        ctx->err = buf;
        return MPM_ERR_BOUNDS;
      }
    }
  return MPM_OK;
}

int read_status(mpm_ctx* ctx) {
  int st = 0;
  MPM_HIP(hipMemcpyAsync(&st, ctx->status, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  if (st) {
    MPM_HIP(hipMemsetAsync(ctx->status, 0, sizeof(int), ctx->stream));
    MPM_HIP(hipStreamSynchronize(ctx->stream));
    if (st & ST_NONFINITE) {
      ctx->err = "SVD Error: non-finite deformation gradient (particle F left unchanged)";
      return MPM_ERR_NONFINITE;
    }
    ctx->err = "device status " + std::to_string(st);
    return MPM_ERR_BOUNDS;
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
  a.gacc = c->gacc;
  a.gvel = c->gvel;
  a.gforce = c->gforce;
  for (int i = 0; i < 3; ++i) {
    a.dims[i] = c->dims[i];
    a.nb[i] = c->nb[i];
    a.g[i] = c->prm.gravity[i];
  }
  a.nnodes = c->nnodes;
  a.h = c->h;
  a.inv_h = 1.0f / c->h;
  a.dt = dt;
  a.lo = c->lo;
  a.hi = c->hi;
  a.gap64 = c->prm.svd_gap_fp64;
  a.Dinv = c->Dinv;
  const float nu = c->prm.nu;
  a.lame.nu = nu;
  a.lame.xi = c->prm.xi;
  a.lame.dmu = 2.0f * (1.0f + nu);
  a.lame.rmu = 1.0f / a.lame.dmu;
  a.lame.dlam = (1.0f + nu) * (1.0f - 2.0f * nu);
  a.lame.rlam = 1.0f / a.lame.dlam;
  a.compose = c->prm.compose;
  a.bc_mode = c->prm.bc_mode;
  a.ncoll = c->ncoll;
  a.coll = c->coll;
  a.status = c->status;
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

int mpm_create(const mpm_params* p, int32_t n, mpm_ctx** out) {
  if (!p || !out || n < 0) return MPM_ERR_ARG;
  for (int a = 0; a < 3; ++a)
    if (p->dims[a] < 6) return MPM_ERR_ARG;
  if (!(p->h > 0.0f)) return MPM_ERR_ARG;
  if (p->compose != MPM_COMPOSE_LITERAL && p->compose != MPM_COMPOSE_TEXTBOOK) return MPM_ERR_ARG;
  auto* ctx = new (std::nothrow) mpm_ctx();
  if (!ctx) return MPM_ERR_NOMEM;
  *out = ctx;
  ctx->prm = *p;
  ctx->n = n;
  ctx->stride = ((int64_t)n + 63) / 64 * 64;
  if (ctx->stride == 0) ctx->stride = 64;
  for (int a = 0; a < 3; ++a) {
    ctx->dims[a] = p->dims[a];
    ctx->nb[a] = (p->dims[a] + 3) / 4;
  }
  ctx->nnodes = (int64_t)ctx->nb[0] * ctx->nb[1] * ctx->nb[2] * 64;
  ctx->h = p->h;
  ctx->lo = (float)(1.0 - p->theta_c);
  ctx->hi = (float)(1.0 + p->theta_s);
  // DpInverse = inverse(mat3(1) * (1/3) * h * h) (material_point_method.hpp:177)
  {
    float d = 1.0f * (1.0f / 3.0f);
    d = d * p->h;
    d = d * p->h;
    M3 D;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) D.c[i][j] = (i == j) ? d : 0.0f;
    // glm compute_inverse on the host
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
    ctx->Dinv = r;
  }
  ctx->resort_every = p->resort_every > 0 ? p->resort_every : 16;
  MPM_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  const size_t pbytes = sizeof(float) * (size_t)NFIELDS * (size_t)ctx->stride;
  const size_t gbytes = sizeof(float4) * (size_t)ctx->nnodes;
  const int64_t nblocks = (int64_t)ctx->nb[0] * ctx->nb[1] * ctx->nb[2];
  const int64_t ntiles = ((int64_t)n + 4095) / 4096;
  const size_t nn = (size_t)std::max(1, n);
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
  alloc(&ctx->blk_start, sizeof(int) * nblocks);
  alloc(&ctx->blk_end, sizeof(int) * nblocks);
  alloc(&ctx->blk_ids, sizeof(int) * nblocks);
  alloc(&ctx->blist, sizeof(int4) * nblocks);
  alloc(&ctx->nblk, sizeof(int));
  alloc(&ctx->gflag, sizeof(int) * nblocks);
  alloc(&ctx->glist[0], sizeof(int) * nblocks);
  alloc(&ctx->glist[1], sizeof(int) * nblocks);
  alloc(&ctx->gcnt, sizeof(int) * 4);
  alloc(&ctx->dstep, sizeof(int));
  alloc(&ctx->gacc, gbytes);
  alloc(&ctx->gvel, gbytes);
  alloc(&ctx->gforce, gbytes);
  alloc(&ctx->status, sizeof(int));
  if (!ok) {
    ctx->err = "device allocation failed";
    return MPM_ERR_NOMEM;
  }
  MPM_HIP(hipMemsetAsync(ctx->P, 0, pbytes, ctx->stream));
  MPM_HIP(hipMemsetAsync(ctx->P2, 0, pbytes, ctx->stream));
  launch_iota(ctx->pid, n, ctx->stream);
  launch_fill_int(ctx->gflag, nblocks, 0, ctx->stream);
  launch_fill_int(ctx->gcnt, 4, 0, ctx->stream);
  launch_fill_int(ctx->dstep, 1, 1, ctx->stream);
  launch_fill_int(ctx->nblk, 1, 0, ctx->stream);
  launch_clear(ctx->gacc, ctx->nnodes, ctx->stream);
  launch_clear(ctx->gvel, ctx->nnodes, ctx->stream);
  launch_clear(ctx->gforce, ctx->nnodes, ctx->stream);
  MPM_HIP(hipMemsetAsync(ctx->status, 0, sizeof(int), ctx->stream));
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  ctx->gacc_clean = true;
  ctx->P_first = ctx->P;
  return MPM_OK;
}

int mpm_destroy(mpm_ctx* ctx) {
  if (!ctx) return MPM_ERR_ARG;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (hipGraphExec_t& g : ctx->gexec)
    if (g) (void)hipGraphExecDestroy(g);
  void* bufs[] = {ctx->P,       ctx->P2,      ctx->pid,      ctx->pid2,  ctx->keys,     ctx->vals,     ctx->keys2,
                  ctx->vals2,   ctx->hist,    ctx->blk_start, ctx->blk_end, ctx->blk_ids, ctx->blist,   ctx->nblk,
                  ctx->gflag,   ctx->glist[0], ctx->glist[1], ctx->gcnt,  ctx->dstep,    ctx->gacc,     ctx->gvel,
                  ctx->gforce,  ctx->coll,    ctx->status,  ctx->dcount};
  for (void* b : bufs) (void)hipFree(b);
  for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
  for (hipEvent_t e : ctx->ev_region)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return MPM_OK;
}

const char* mpm_last_error(const mpm_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }
int32_t mpm_num_particles(const mpm_ctx* ctx) { return ctx ? ctx->n : 0; }

static int upload_fields(mpm_ctx* ctx, int field, const float* host, int comps) {
  if (!host || ctx->n == 0) return MPM_OK;
  float* tmp = nullptr;
  const size_t bytes = sizeof(float) * (size_t)ctx->n * comps;
  MPM_HIP(hipMalloc(&tmp, bytes));
  MPM_HIP(hipMemcpyAsync(tmp, host, bytes, hipMemcpyHostToDevice, ctx->stream));
  launch_scatter_field(ctx->P + (int64_t)field * ctx->stride, ctx->stride, tmp, ctx->pid, ctx->n, comps, ctx->stream);
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  MPM_HIP(hipFree(tmp));
  return MPM_OK;
}
static int download_fields(mpm_ctx* ctx, int field, float* host, int comps) {
  if (!host || ctx->n == 0) return MPM_OK;
  float* tmp = nullptr;
  const size_t bytes = sizeof(float) * (size_t)ctx->n * comps;
  MPM_HIP(hipMalloc(&tmp, bytes));
  launch_gather_field(tmp, ctx->P + (int64_t)field * ctx->stride, ctx->stride, ctx->pid, ctx->n, comps, ctx->stream);
  MPM_HIP(hipMemcpyAsync(host, tmp, bytes, hipMemcpyDeviceToHost, ctx->stream));
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  MPM_HIP(hipFree(tmp));
  return MPM_OK;
}

int mpm_set_particles(mpm_ctx* ctx, const float* x, const float* v, const float* m, const float* E0) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  if (ctx->n && (!x || !v || !m)) {
    ctx->err = "x, v and m are required";
    return MPM_ERR_ARG;
  }
  if (int rc = check_positions(ctx, x)) return rc;
  launch_iota(ctx->pid, ctx->n, ctx->stream);  // caller order again
  ctx->since_sort = -1;
  int rc;
  if ((rc = upload_fields(ctx, FX, x, 3)) || (rc = upload_fields(ctx, FV, v, 3)) || (rc = upload_fields(ctx, FM, m, 1)))
    return rc;
  std::vector<float> ident((size_t)ctx->n * 9, 0.0f), zeros((size_t)ctx->n * 9, 0.0f), e((size_t)ctx->n);
  for (int p = 0; p < ctx->n; ++p) {
    ident[9 * p + 0] = ident[9 * p + 4] = ident[9 * p + 8] = 1.0f;
    e[p] = E0 ? E0[p] : ctx->prm.E;
  }
  if ((rc = upload_fields(ctx, FB, zeros.data(), 9)) || (rc = upload_fields(ctx, FFE, ident.data(), 9)) ||
      (rc = upload_fields(ctx, FFP, ident.data(), 9)) || (rc = upload_fields(ctx, FVOL, zeros.data(), 1)) ||
      (rc = upload_fields(ctx, FE0, e.data(), 1)))
    return rc;
  ctx->particles_set = true;
  return MPM_OK;
}

int mpm_set_state(mpm_ctx* ctx, const float* B, const float* FE, const float* FP, const float* V0) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  int rc;
  if ((rc = upload_fields(ctx, FB, B, 9)) || (rc = upload_fields(ctx, FFE, FE, 9)) ||
      (rc = upload_fields(ctx, FFP, FP, 9)) || (rc = upload_fields(ctx, FVOL, V0, 1)))
    return rc;
  return MPM_OK;
}

int mpm_set_colliders(mpm_ctx* ctx, const mpm_collider* c, int32_t k) {
  if (check_ctx(ctx) || k < 0 || (k > 0 && !c)) return MPM_ERR_ARG;
  std::vector<DevCollider> host((size_t)k);
  for (int i = 0; i < k; ++i) {
    host_collider_inverse(c[i].translation, c[i].rotation, host[i].inv);
    for (int a = 0; a < 3; ++a) {
      host[i].s[a] = c[i].scale[a];
      host[i].vel[a] = c[i].velocity[a];
    }
  }
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  (void)hipFree(ctx->coll);
  ctx->coll = nullptr;
  ctx->ncoll = k;
  if (k) {
    MPM_HIP(hipMalloc(&ctx->coll, sizeof(DevCollider) * k));
    MPM_HIP(hipMemcpy(ctx->coll, host.data(), sizeof(DevCollider) * k, hipMemcpyHostToDevice));
  }
  return MPM_OK;
}

int mpm_stage(mpm_ctx* ctx, int32_t stage, float dt) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  if (!ctx->particles_set) {
    ctx->err = "no particles set";
    return MPM_ERR_STATE;
  }
  const KArgs a = make_args(ctx, dt);
  hipStream_t s = ctx->stream;
  switch (stage) {
    case MPM_STAGE_P2G:
      launch_clear(ctx->gacc, ctx->nnodes, s);
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

int mpm_init_volumes(mpm_ctx* ctx) {
  int rc = mpm_stage(ctx, MPM_STAGE_P2G, 0.0f);
  if (rc) return rc;
  return mpm_stage(ctx, MPM_STAGE_VOLUMES, 0.0f);
}

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
  for (int i = 0; i < 3; ++i) {
    b.dims[i] = ctx->dims[i];
    b.nb[i] = ctx->nb[i];
  }
  b.nnodes = ctx->nnodes;
  b.h = ctx->h;
  b.keys = ctx->keys;
  b.vals = ctx->vals;
  b.keys2 = ctx->keys2;
  b.vals2 = ctx->vals2;
  b.hist = ctx->hist;
  b.blk_start = ctx->blk_start;
  b.blk_end = ctx->blk_end;
  b.blk_ids = ctx->blk_ids;
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

// One substep: re-bin when due, then the block pass (PASS_P2G on the first
// substep of a call, PASS_FUSED after it), the grid pass and the epoch advance.
int substep(mpm_ctx* ctx, float dt, bool first) {
  int rc;
  if (ctx->since_sort < 0 || ctx->since_sort >= ctx->resort_every)
    if ((rc = rebin(ctx))) return rc;
  const BlockArgs ba = block_args(ctx, dt);
  hipStream_t s = ctx->stream;
  if ((rc = timed(ctx, first ? MPM_KERNEL_P2G : MPM_KERNEL_FUSED,
                  [&] { launch_block_pass(ba, first ? PASS_P2G : PASS_FUSED, s); })) ||
      (rc = timed(ctx, MPM_KERNEL_GRID, [&] { launch_grid_active(ba, s); })))
    return rc;
  launch_advance_step(ctx->dstep, ctx->gcnt, s);
  ++ctx->since_sort;
  return MPM_OK;
}

// resort_every fused substeps that start with a re-bin, replayed from a HIP
// graph.  The launches bake in the particle buffers (which alternate with
// every re-bin, hence one graph per buffer parity), dt and the collider
// table; the substep epoch and block lists are device state, so one graph
// serves every segment.
int graph_segment(mpm_ctx* ctx, float dt) {
  const int par = ctx->P == ctx->P_first ? 0 : 1;
  hipGraphExec_t& g = ctx->gexec[par];
  if (g && (ctx->graph_dt[par] != dt || ctx->graph_coll[par] != (const void*)ctx->coll ||
            ctx->graph_ncoll[par] != ctx->ncoll)) {
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
    ctx->graph_coll[par] = ctx->coll;
    ctx->graph_ncoll[par] = ctx->ncoll;
  }
  MPM_HIP(hipGraphLaunch(g, ctx->stream));
  std::swap(ctx->P, ctx->P2);
  std::swap(ctx->pid, ctx->pid2);
  ctx->since_sort = ctx->resort_every;
  return MPM_OK;
}

}  // namespace

int mpm_substeps(mpm_ctx* ctx, float dt, int32_t nsteps) {
  if (check_ctx(ctx) || nsteps < 0) return MPM_ERR_ARG;
  if (!ctx->particles_set) {
    ctx->err = "no particles set";
    return MPM_ERR_STATE;
  }
  if (nsteps == 0 || ctx->n == 0) return MPM_OK;
  hipStream_t s = ctx->stream;
  if (!ctx->gacc_clean) {
    // entering the fused path after legacy stages: clean grid and block lists
    const int64_t nblocks = (int64_t)ctx->nb[0] * ctx->nb[1] * ctx->nb[2];
    launch_clear(ctx->gacc, ctx->nnodes, s);
    launch_clear(ctx->gvel, ctx->nnodes, s);
    launch_fill_int(ctx->gflag, nblocks, 0, s);
    launch_fill_int(ctx->gcnt, 4, 0, s);
    launch_fill_int(ctx->dstep, 1, 1, s);
    ctx->gacc_clean = true;
  }
  const bool graphs = ctx->use_graphs && !(ctx->profiling && ctx->per_kernel);
  int rc;
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

int mpm_get_particles(mpm_ctx* ctx, float* x, float* v, float* B, float* FE, float* FP, float* V0) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  int rc;
  if ((rc = download_fields(ctx, FX, x, 3)) || (rc = download_fields(ctx, FV, v, 3)) ||
      (rc = download_fields(ctx, FB, B, 9)) || (rc = download_fields(ctx, FFE, FE, 9)) ||
      (rc = download_fields(ctx, FFP, FP, 9)) || (rc = download_fields(ctx, FVOL, V0, 1)))
    return rc;
  return read_status(ctx);
}

int mpm_get_grid(mpm_ctx* ctx, float* m, float* vel, float* force, uint8_t* active) {
  if (check_ctx(ctx)) return MPM_ERR_ARG;
  const int64_t cells = (int64_t)ctx->dims[0] * ctx->dims[1] * ctx->dims[2];
  float *dm = nullptr, *dv = nullptr, *df = nullptr;
  uint8_t* da = nullptr;
  if (m) MPM_HIP(hipMalloc(&dm, sizeof(float) * cells));
  if (vel) MPM_HIP(hipMalloc(&dv, sizeof(float) * 3 * cells));
  if (force) MPM_HIP(hipMalloc(&df, sizeof(float) * 3 * cells));
  if (active) MPM_HIP(hipMalloc(&da, cells));
  launch_grid_to_dense(make_args(ctx, 0.0f), dm, dv, df, da, ctx->stream);
  if (m) MPM_HIP(hipMemcpyAsync(m, dm, sizeof(float) * cells, hipMemcpyDeviceToHost, ctx->stream));
  if (vel) MPM_HIP(hipMemcpyAsync(vel, dv, sizeof(float) * 3 * cells, hipMemcpyDeviceToHost, ctx->stream));
  if (force) MPM_HIP(hipMemcpyAsync(force, df, sizeof(float) * 3 * cells, hipMemcpyDeviceToHost, ctx->stream));
  if (active) MPM_HIP(hipMemcpyAsync(active, da, cells, hipMemcpyDeviceToHost, ctx->stream));
  MPM_HIP(hipStreamSynchronize(ctx->stream));
  (void)hipFree(dm);
  (void)hipFree(dv);
  (void)hipFree(df);
  (void)hipFree(da);
  return MPM_OK;
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
