// Readback and snapshots.
//
// Positions readback (the viewer's per-frame read of getParticles(),
// realtime-deformations/main.cpp:257-271): one gather kernel into a
// context-owned device staging buffer, then a device-to-host copy into a
// context-owned pinned mirror on a second stream, so the copy of frame k
// overlaps the substeps of frame k + 1 on the compute stream.
//
// Snapshot file (little-endian), shared by resume, checkpoints on disk and
// test fixtures:
//   header, 64 bytes: "MPMSNAP1", u32 version = 1, u32 nscenes, i64 step,
//     u32 fields per particle = 36, u32 scene record bytes = 88,
//     u64 total particles, 24 zero bytes;
//   nscenes scene records of 88 bytes: the mpm_params fields in declaration
//     order (dims[3], compose, bc_mode, flags: i32; h, E, nu, xi: f32;
//     theta_c, theta_s: f64; gravity[3], svd_gap_fp64: f32; resort_every: i32),
//     then i32 particles, i32 colliders, 4 zero bytes;
//   per scene, its colliders: 13 f32 each (translation, rotation w x y z,
//     scale, velocity);
//   per scene, its particles in caller order, field after field:
//     x (n x 3), v (n x 3), B (n x 9), F_E (n x 9), F_P (n x 9), m, V0, E0
//     (n each), all f32, matrices glm column-major.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "mpm_hostutil.hpp"
#include "mpm_internal.hpp"

using namespace mpm;

#define IO_HIP(call)                                                    \
  do {                                                                  \
    hipError_t e_ = (call);                                             \
    if (e_ != hipSuccess) {                                             \
      ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);    \
      return MPM_ERR_HIP;                                               \
    }                                                                   \
  } while (0)

namespace {

constexpr char MAGIC[8] = {'M', 'P', 'M', 'S', 'N', 'A', 'P', '1'};
constexpr uint32_t VERSION = 1, RECORD = 88;

struct Header {
  char magic[8];
  uint32_t version, nscenes;
  int64_t step;
  uint32_t nfields, record;
  uint64_t total;
  uint8_t zero[24];
};
static_assert(sizeof(Header) == 64, "snapshot header is 64 bytes");

// the particle blocks of a scene, in file order
struct Block {
  int field, comps;
};
constexpr Block BLOCKS[8] = {{FX, 3}, {FV, 3}, {FB, 9}, {FFE, 9}, {FFP, 9}, {FM, 1}, {FVOL, 1}, {FE0, 1}};

void put_record(const mpm_params& p, int32_t n, int32_t ncoll, uint8_t out[RECORD]) {
  std::memset(out, 0, RECORD);
  size_t o = 0;
  auto put = [&](const void* v, size_t b) {
    std::memcpy(out + o, v, b);
    o += b;
  };
  put(p.dims, 12);
  put(&p.compose, 4);
  put(&p.bc_mode, 4);
  put(&p.flags, 4);
  put(&p.h, 4);
  put(&p.E, 4);
  put(&p.nu, 4);
  put(&p.xi, 4);
  put(&p.theta_c, 8);
  put(&p.theta_s, 8);
  put(p.gravity, 12);
  put(&p.svd_gap_fp64, 4);
  put(&p.resort_every, 4);
  put(&n, 4);
  put(&ncoll, 4);
}

void get_record(const uint8_t in[RECORD], mpm_params* p, int32_t* n, int32_t* ncoll) {
  size_t o = 0;
  auto get = [&](void* v, size_t b) {
    std::memcpy(v, in + o, b);
    o += b;
  };
  *p = mpm_params{};
  get(p->dims, 12);
  get(&p->compose, 4);
  get(&p->bc_mode, 4);
  get(&p->flags, 4);
  get(&p->h, 4);
  get(&p->E, 4);
  get(&p->nu, 4);
  get(&p->xi, 4);
  get(&p->theta_c, 8);
  get(&p->theta_s, 8);
  get(p->gravity, 12);
  get(&p->svd_gap_fp64, 4);
  get(&p->resort_every, 4);
  get(n, 4);
  get(ncoll, 4);
}

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) fclose(f);
  }
};

// header + scene records + colliders of a snapshot
struct Meta {
  Header h{};
  std::vector<mpm_params> prm;
  std::vector<int32_t> n;
  std::vector<std::vector<mpm_collider>> coll;
};

int read_meta(FILE* f, Meta* m, std::string* err) {
  if (fread(&m->h, sizeof(Header), 1, f) != 1 || std::memcmp(m->h.magic, MAGIC, 8) != 0) {
    *err = "not a snapshot file (bad magic)";
    return MPM_ERR_ARG;
  }
  if (m->h.version != VERSION || m->h.nfields != NFIELDS || m->h.record != RECORD || m->h.nscenes < 1 ||
      m->h.nscenes > (1u << 20)) {
    *err = "unsupported snapshot version or layout";
    return MPM_ERR_ARG;
  }
  const int ns = (int)m->h.nscenes;
  m->prm.resize(ns);
  m->n.resize(ns);
  m->coll.resize(ns);
  std::vector<int32_t> nc(ns);
  uint64_t total = 0;
  for (int s = 0; s < ns; ++s) {
    uint8_t rec[RECORD];
    if (fread(rec, RECORD, 1, f) != 1) {
      *err = "truncated snapshot (scene records)";
      return MPM_ERR_ARG;
    }
    get_record(rec, &m->prm[s], &m->n[s], &nc[s]);
    if (m->n[s] < 0 || nc[s] < 0 || nc[s] > (1 << 20)) {
      *err = "corrupt snapshot scene record";
      return MPM_ERR_ARG;
    }
    total += (uint64_t)m->n[s];
  }
  if (total != m->h.total) {
    *err = "corrupt snapshot: particle counts do not add up";
    return MPM_ERR_ARG;
  }
  for (int s = 0; s < ns; ++s) {
    m->coll[s].resize((size_t)nc[s]);
    for (auto& c : m->coll[s]) {
      float v[13];
      if (fread(v, sizeof v, 1, f) != 1) {
        *err = "truncated snapshot (colliders)";
        return MPM_ERR_ARG;
      }
      std::memcpy(c.translation, v, 12);
      std::memcpy(c.rotation, v + 3, 16);
      std::memcpy(c.scale, v + 7, 12);
      std::memcpy(c.velocity, v + 10, 12);
    }
  }
  return MPM_OK;
}

bool same_params(const mpm_params& a, const mpm_params& b) {
  uint8_t ra[RECORD], rb[RECORD];
  put_record(a, 0, 0, ra);
  put_record(b, 0, 0, rb);
  return std::memcmp(ra, rb, RECORD) == 0;
}

int ensure_readback(mpm_ctx* ctx) {
  const size_t bytes = sizeof(float) * 3 * (size_t)std::max(1, ctx->n);
  if (!ctx->pos_stage) IO_HIP(hipMalloc(&ctx->pos_stage, bytes));
  if (!ctx->pos_host) IO_HIP(hipHostMalloc(&ctx->pos_host, bytes, hipHostMallocDefault));
  if (!ctx->copy_stream) IO_HIP(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
  if (!ctx->ev_gathered) IO_HIP(hipEventCreateWithFlags(&ctx->ev_gathered, hipEventDisableTiming));
  if (!ctx->ev_copied) IO_HIP(hipEventCreateWithFlags(&ctx->ev_copied, hipEventDisableTiming));
  return MPM_OK;
}

// ---- readback that allocates nothing: a file-scope staging array.
// Each field of one scene is gathered into caller order in `scratch` and
// copied to the caller's host array.  In host code `scratch` names a
// host-side stand-in, not the device array: its device address comes from
// hipGetSymbolAddress, once per device, and only that address is used.
constexpr int64_t NOALLOC_SCRATCH_FLOATS = 9 * (int64_t)MPM_NOALLOC_MAX_PARTICLES;
static __device__ __align__(256) float scratch[NOALLOC_SCRATCH_FLOATS];
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex scratch_mu;
float* scratch_dev[NOALLOC_MAX_DEVICES] = {};  // device address of `scratch`, per device

int scratch_address(mpm_ctx* ctx, float** out) {
  int dev = -1;
  IO_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    ctx->err = "mpm_scene_get_particles_noalloc: device index out of the supported range";
    return MPM_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void* p = nullptr;
    IO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    scratch_dev[dev] = static_cast<float*>(p);
  }
  *out = scratch_dev[dev];
  return MPM_OK;
}

}  // namespace

extern "C" {

int mpm_scene_get_particles_noalloc(mpm_ctx* ctx, int32_t scene, float* x, float* v, float* B, float* FE,
                                    float* FP, float* V0) {
  if (!ctx) return MPM_ERR_ARG;
  if (scene < 0 || scene >= ctx->nscenes) {
    ctx->err = "scene index out of range";
    return MPM_ERR_ARG;
  }
  const Scene& S = ctx->scenes[scene];
  if (S.np > MPM_NOALLOC_MAX_PARTICLES) {
    ctx->err = "mpm_scene_get_particles_noalloc: the scene has " + std::to_string(S.np) +
               " particles, more than the staging array holds (MPM_NOALLOC_MAX_PARTICLES = " +
               std::to_string(MPM_NOALLOC_MAX_PARTICLES) + ")";
    return MPM_ERR_ARG;
  }
  float* stage = nullptr;
  if (int rc = scratch_address(ctx, &stage)) return rc;
  if (S.np == 0) return MPM_OK;
  const Block fields[6] = {{FX, 3}, {FV, 3}, {FB, 9}, {FFE, 9}, {FFP, 9}, {FVOL, 1}};
  float* const outs[6] = {x, v, B, FE, FP, V0};
  for (int f = 0; f < 6; ++f) {
    if (!outs[f]) continue;
    // every element [0, np * comps) of the stage is written: caller indices
    // of a scene's slots are a permutation of [p0, p0 + np)
    launch_gather_field(stage, ctx->P + (int64_t)fields[f].field * ctx->stride, ctx->stride, ctx->pid, S.p0, S.np,
                        fields[f].comps, ctx->stream);
    IO_HIP(hipGetLastError());
    IO_HIP(hipMemcpyAsync(outs[f], stage, sizeof(float) * (size_t)S.np * fields[f].comps, hipMemcpyDeviceToHost,
                          ctx->stream));
    IO_HIP(hipStreamSynchronize(ctx->stream));
  }
  return MPM_OK;
}

int mpm_scene_get_material(mpm_ctx* ctx, int32_t scene, float* m, float* E0) {
  if (!ctx) return MPM_ERR_ARG;
  if (scene < 0 || scene >= ctx->nscenes) {
    ctx->err = "scene index out of range";
    return MPM_ERR_ARG;
  }
  int rc;
  if ((rc = download_field(ctx, scene, FM, m, 1)) || (rc = download_field(ctx, scene, FE0, E0, 1))) return rc;
  return MPM_OK;
}

int mpm_positions_begin(mpm_ctx* ctx) {
  if (!ctx) return MPM_ERR_ARG;
  if (int rc = ensure_readback(ctx)) return rc;
  // the staging buffer is reused: the gather waits for the previous copy
  if (ctx->pos_pending) IO_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_copied, 0));
  for (int s = 0; s < ctx->nscenes; ++s) {
    const Scene& S = ctx->scenes[s];
    if (S.np)
      launch_gather_field(ctx->pos_stage + 3 * (int64_t)S.p0, ctx->P + (int64_t)FX * ctx->stride, ctx->stride,
                          ctx->pid, S.p0, S.np, 3, ctx->stream);
  }
  IO_HIP(hipGetLastError());
  IO_HIP(hipEventRecord(ctx->ev_gathered, ctx->stream));
  IO_HIP(hipStreamWaitEvent(ctx->copy_stream, ctx->ev_gathered, 0));
  IO_HIP(hipMemcpyAsync(ctx->pos_host, ctx->pos_stage, sizeof(float) * 3 * (size_t)ctx->n, hipMemcpyDeviceToHost,
                        ctx->copy_stream));
  IO_HIP(hipEventRecord(ctx->ev_copied, ctx->copy_stream));
  ctx->pos_pending = true;
  return MPM_OK;
}

int mpm_positions_wait(mpm_ctx* ctx, const float** x) {
  if (!ctx || !x) return MPM_ERR_ARG;
  if (!ctx->pos_pending) {
    ctx->err = "mpm_positions_wait without mpm_positions_begin";
    return MPM_ERR_STATE;
  }
  IO_HIP(hipEventSynchronize(ctx->ev_copied));
  *x = ctx->pos_host;
  return MPM_OK;
}

int mpm_get_positions(mpm_ctx* ctx, float* x) {
  if (!ctx || !x) return MPM_ERR_ARG;
  const float* view = nullptr;
  int rc;
  if ((rc = mpm_positions_begin(ctx)) || (rc = mpm_positions_wait(ctx, &view))) return rc;
  // from the pinned mirror into the caller's array: one host thread copies
  // (and first-touches) about 9 GB/s, so large reads are split over threads
  const size_t bytes = sizeof(float) * 3 * (size_t)ctx->n;
  const size_t chunk = size_t(4) << 20;
  const unsigned nt = (unsigned)std::min<size_t>(8, std::max<size_t>(1, bytes / chunk));
  if (nt == 1) {
    std::memcpy(x, view, bytes);
  } else {
    // parts of ceil(bytes / nt) rounded up to 64 bytes: nt of them cover
    // every byte, and the last one ends at `bytes` whatever the rounding
    const size_t part = readback_part_bytes(bytes, nt);
    auto copy_part = [=](unsigned t) {
      const size_t o = std::min(bytes, part * t);
      const size_t e = t + 1 == nt ? bytes : std::min(bytes, part * (t + 1));
      std::memcpy(reinterpret_cast<char*>(x) + o, reinterpret_cast<const char*>(view) + o, e - o);
    };
    std::vector<std::thread> pool;
    unsigned handed = 1;
    try {
      pool.reserve(nt - 1);
      for (; handed < nt; ++handed) pool.emplace_back(copy_part, handed);
    } catch (...) {
      // fewer threads than asked for: the parts not handed out are copied here
    }
    copy_part(0);
    for (unsigned t = handed; t < nt; ++t) copy_part(t);
    for (std::thread& th : pool) th.join();
  }
  return mpm_sync(ctx);
}

int mpm_snapshot_save(mpm_ctx* ctx, const char* path, int64_t step) {
  if (!ctx || !path) return MPM_ERR_ARG;
  if (int rc = mpm_sync(ctx)) return rc;
  File out;
  out.f = fopen(path, "wb");
  if (!out.f) {
    ctx->err = std::string("cannot open ") + path + " for writing";
    return MPM_ERR_ARG;
  }
  Header h{};
  std::memcpy(h.magic, MAGIC, 8);
  h.version = VERSION;
  h.nscenes = (uint32_t)ctx->nscenes;
  h.step = step;
  h.nfields = NFIELDS;
  h.record = RECORD;
  h.total = (uint64_t)ctx->n;
  bool ok = fwrite(&h, sizeof h, 1, out.f) == 1;
  for (int s = 0; s < ctx->nscenes && ok; ++s) {
    uint8_t rec[RECORD];
    put_record(ctx->sprm[s], ctx->scenes[s].np, (int32_t)ctx->hcoll[s].size(), rec);
    ok = fwrite(rec, RECORD, 1, out.f) == 1;
  }
  for (int s = 0; s < ctx->nscenes && ok; ++s)
    for (const mpm_collider& c : ctx->hcoll[s]) {
      float v[13];
      std::memcpy(v, c.translation, 12);
      std::memcpy(v + 3, c.rotation, 16);
      std::memcpy(v + 7, c.scale, 12);
      std::memcpy(v + 10, c.velocity, 12);
      ok = ok && fwrite(v, sizeof v, 1, out.f) == 1;
    }
  std::vector<float> buf;
  for (int s = 0; s < ctx->nscenes && ok; ++s) {
    const int np = ctx->scenes[s].np;
    for (const Block& b : BLOCKS) {
      buf.resize((size_t)np * b.comps);
      if (int rc = download_field(ctx, s, b.field, buf.data(), b.comps)) return rc;
      ok = ok && (np == 0 || fwrite(buf.data(), sizeof(float) * buf.size(), 1, out.f) == 1);
    }
  }
  if (!ok) {
    ctx->err = std::string("write to ") + path + " failed";
    return MPM_ERR_ARG;
  }
  return MPM_OK;
}

int mpm_snapshot_load(mpm_ctx* ctx, const char* path, int64_t* step) {
  if (!ctx || !path) return MPM_ERR_ARG;
  File in;
  in.f = fopen(path, "rb");
  if (!in.f) {
    ctx->err = std::string("cannot open ") + path;
    return MPM_ERR_ARG;
  }
  Meta m;
  if (int rc = read_meta(in.f, &m, &ctx->err)) return rc;
  if ((int)m.h.nscenes != ctx->nscenes) {
    ctx->err = "snapshot has a different number of scenes";
    return MPM_ERR_ARG;
  }
  for (int s = 0; s < ctx->nscenes; ++s)
    if (m.n[s] != ctx->scenes[s].np || !same_params(m.prm[s], ctx->sprm[s])) {
      ctx->err = "snapshot scene " + std::to_string(s) + " differs from the context's (particles or parameters)";
      return MPM_ERR_ARG;
    }
  // read everything first, so a truncated file leaves the context untouched
  std::vector<std::vector<float>> data(ctx->nscenes * 8);
  for (int s = 0; s < ctx->nscenes; ++s)
    for (int b = 0; b < 8; ++b) {
      auto& v = data[s * 8 + b];
      v.resize((size_t)m.n[s] * BLOCKS[b].comps);
      if (!v.empty() && fread(v.data(), sizeof(float) * v.size(), 1, in.f) != 1) {
        ctx->err = "truncated snapshot (particles)";
        return MPM_ERR_ARG;
      }
    }
  for (int s = 0; s < ctx->nscenes; ++s)
    if (int rc = check_scene_positions(ctx, s, data[s * 8].data())) return rc;
  // a load replaces whatever state raised the status flags: clear them
  IO_HIP(hipStreamSynchronize(ctx->stream));
  IO_HIP(hipMemsetAsync(ctx->status, 0, sizeof(int) * ctx->nscenes, ctx->stream));
  for (int s = 0; s < ctx->nscenes; ++s) {
    int rc;
    if ((rc = reset_scene_order(ctx, s))) return rc;
    for (int b = 0; b < 8; ++b)
      if ((rc = upload_field(ctx, s, BLOCKS[b].field, data[s * 8 + b].data(), BLOCKS[b].comps))) return rc;
    if ((rc = mpm_scene_set_colliders(ctx, s, m.coll[s].data(), (int32_t)m.coll[s].size()))) return rc;
    ctx->scene_set[s] = 1;
    ctx->quanta_stale = true;
  }
  ctx->particles_set = true;
  if (step) *step = m.h.step;
  return MPM_OK;
}

int mpm_create_from_snapshot(const char* path, mpm_ctx** out, int64_t* step) {
  if (!path || !out) return MPM_ERR_ARG;
  *out = nullptr;
  Meta m;
  std::string err;
  {
    File in;
    in.f = fopen(path, "rb");
    if (!in.f) return MPM_ERR_ARG;
    if (int rc = read_meta(in.f, &m, &err)) return rc;
  }
  if (int rc = mpm_create_batch(m.prm.data(), m.n.data(), (int32_t)m.h.nscenes, out)) return rc;
  return mpm_snapshot_load(*out, path, step);
}

}  // extern "C"
