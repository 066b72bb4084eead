// Context, device layout and kernel launchers of libmpm_hip.
//
// Particle state lives in one device buffer of NFIELDS rows of `stride`
// floats (structure of arrays): every field read by a wave is one coalesced
// 256-byte line per row.  The grid is dense over 4x4x4 node blocks stored
// block-major (node (i,j,k) -> block (i>>2, j>>2, k>>2) * 64 + local), so a
// block's 64 nodes are one contiguous kilobyte of float4.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "mpm_c.h"
#include "mpm_math.hpp"

namespace mpm {

enum Field : int {
  FX = 0,    // x, y, z
  FV = 3,    // velocity
  FB = 6,    // APIC B (9, glm column-major)
  FFE = 15,  // F_E (9)
  FFP = 24,  // F_P (9)
  FM = 33,   // mass
  FVOL = 34, // initial volume V0
  FE0 = 35,  // Young's modulus
  NFIELDS = 36
};

// Device status bits per scene.  ST_CFL (a particle moved more than a cell,
// or the elastic wave crosses more than a cell, in one substep) is a
// warning: only mpm_scene_status reports it.
enum StatusBits : int { ST_NONFINITE = 1, ST_BOUNDS = 2, ST_CFL = 4 };

// A device allocation freed when the scope ends, error returns included.
template <typename T> struct DevBuf {
  T* p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes); }
};
// A HIP event destroyed when the scope ends.
struct EventGuard {
  hipEvent_t e = nullptr;
  EventGuard() = default;
  EventGuard(const EventGuard&) = delete;
  EventGuard& operator=(const EventGuard&) = delete;
  ~EventGuard() {
    if (e) (void)hipEventDestroy(e);
  }
};
// Runs f when the scope ends (restores context settings on early returns).
template <typename F> struct ScopeExit {
  F f;
  ~ScopeExit() { f(); }
};
template <typename F> ScopeExit<F> on_scope_exit(F f) { return ScopeExit<F>{f}; }

// One scene: its grid (dims, spacing, place in the shared grid arrays), its
// particle slots, material and boundary conditions.  A context steps any
// number of scenes together; kernels read a scene's entry per block or per
// particle from a device table.
struct Scene {
  int dims[3];
  int nb[3];        // 4x4x4 node blocks per axis
  int64_t blk0;     // first grid block of this scene in the shared grid arrays
  int64_t nblocks;  // nb0 * nb1 * nb2
  int p0, np;       // particle slots [p0, p0 + np) (fixed: sorting stays inside a scene)
  float h, inv_h;   // inv_h: 1/h correctly rounded
  float E;          // Young's modulus of particles uploaded without E0
  float lo, hi, gap64;
  float dinv;       // D^-1 = dinv_matrix(dinv), glm inverse(mat3(1) h^2 / 3)
  Lame lame;
  float g[3];
  int bc_mode, ncoll;
  const DevCollider* coll;
  // reproducible mode: P2G contributions are rounded to multiples of
  // 2^qexp (mass) and 2^(qexp + 2) (momentum), 2^qexp = 2^-40 of the
  // scene's largest particle mass rounded up to a power of two
  int qexp;
};

// Everything a kernel needs, passed by value.
struct KArgs {
  float* P;
  int64_t stride;
  int n;           // particles over all scenes
  float4* gacc;    // (m, p) accumulators, FP32 (default mode)
  double4* gacc64; // the same memory as FP64 accumulators (reproducible mode)
  float4* gvel;    // (v, m) after the grid update
  float4* gforce;  // (f, 0), legacy stage path only
  int64_t nnodes;  // nodes over all scenes
  float dt;
  int compose;
  int nscenes;
  const Scene* scenes;   // device table
  const int* blk_scene;  // grid block -> scene
  Scene sc;              // scene 0 by value (legacy stage kernels: one scene)
  int* status;           // per scene
  int exact;             // reproducible mode (P2G sums on the scenes' quanta)
};

// Reproducible P2G sums.  M = quantum_magic(k) = 1.5 * 2^(52 + k); for
// |c| < 2^(51 + k), c + M lies in [2^(52 + k), 2^(53 + k)), where doubles are
// spaced by 2^k, so on_quantum(c, M) is c rounded to the nearest multiple of
// 2^k (subtracting M is exact).  A sum of such multiples is exact, and so
// the same in any order, while it stays below 2^(53 + k) in magnitude;
// beyond that it is rounded like any FP64 sum.  Non-finite c stays
// non-finite.  (Built without fast-math: the add and subtract are kept.)
__device__ __forceinline__ double quantum_magic(int k) { return ldexp(1.5, 52 + k); }
__device__ __forceinline__ double on_quantum(double c, double M) { return (c + M) - M; }

// P2G sums of node g as FP32, from the accumulators of the mode
__device__ __forceinline__ float4 load_acc(const KArgs& a, int64_t g) {
  if (a.exact) {
    const double4 d = a.gacc64[g];
    return make_float4((float)d.x, (float)d.y, (float)d.z, (float)d.w);
  }
  return a.gacc[g];
}

__device__ __forceinline__ int64_t node_index(int i, int j, int k, const Scene& S) {
  return (S.blk0 + ((int64_t)((i >> 2) * S.nb[1] + (j >> 2)) * S.nb[2] + (k >> 2))) * 64 + ((i & 3) << 4) +
         ((j & 3) << 2) + (k & 3);
}

// scene owning particle slot p (slots are grouped by scene)
__device__ __forceinline__ int scene_of_slot(const Scene* scenes, int nscenes, int p) {
  int lo = 0, hi = nscenes - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (scenes[mid].p0 <= p) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

}  // namespace mpm

struct mpm_ctx {
  mpm_params prm{};            // scene 0's parameters (compose, resort_every for all)
  int nscenes = 1;
  std::vector<mpm_params> sprm;                    // each scene's parameters as created
  std::vector<std::vector<mpm_collider>> hcoll;    // each scene's colliders as set
  std::vector<mpm::Scene> scenes;          // host copy of the scene table
  std::vector<mpm::DevCollider*> scoll;    // per-scene collider arrays (device)
  mpm::Scene* d_scenes = nullptr;          // device scene table
  int* blk_scene = nullptr;                // grid block -> scene
  int n = 0;
  int64_t stride = 0;
  int64_t nnodes = 0;       // over all scenes
  int64_t nblocks = 0;      // grid blocks over all scenes
  float* P = nullptr;       // particle fields (binned order once sorted)
  float* P2 = nullptr;      // reorder target, swapped with P
  int* pid = nullptr;       // caller index of each particle slot
  int* pid2 = nullptr;
  // binning
  uint32_t *keys = nullptr, *vals = nullptr, *keys2 = nullptr, *vals2 = nullptr, *hist = nullptr;
  int *blk_end = nullptr, *nblk = nullptr;
  int4* blist = nullptr;
  // active grid blocks of the fused path
  int* gflag = nullptr;
  int* glist[2] = {nullptr, nullptr};
  int* gcnt = nullptr;      // three rotating counters
  int* dstep = nullptr;     // device substep epoch (> 0)
  int* bctr = nullptr;      // block passes' home-block claim counter
  int since_sort = -1;      // substeps since the last binning (-1: never binned)
  int resort_every = 16;
  int block_grid = 2048;    // workgroups of the persistent block pass
  double4* gacc = nullptr;  // (m, p) accumulators: FP64 in reproducible mode,
                            // else FP32 in the first half
  float4* gvel = nullptr;
  float4* gforce = nullptr;
  int* status = nullptr;    // device status word per scene
  hipStream_t stream = nullptr;
  std::vector<char> scene_set;  // particles uploaded, per scene
  bool particles_set = false;   // every scene has particles
  bool gacc_clean = false;      // fused path needs zeroed accumulators
  bool exact = false;           // reproducible mode
  bool quanta_stale = true;     // masses changed since the scenes' quanta were set
  uint32_t* dmref = nullptr;    // per scene: bits of its largest particle mass
  std::string err;
  // HIP graphs of [rebin + resort_every substeps], one per particle-buffer
  // parity, valid for one dt (scene data, colliders included, is read from
  // the device table at run time)
  float* P_first = nullptr;  // P as allocated: parity 0 when P == P_first
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  float graph_dt[2] = {0.f, 0.f};
  int graph_key[2] = {0, 0};  // resort_every and block_grid baked into the graph
  bool use_graphs = true;
  // event profiling of fused-path launches
  bool profiling = false;
  bool per_kernel = true;
  hipEvent_t ev_region[2] = {nullptr, nullptr};  // profiled region begin, end
  std::vector<hipEvent_t> events;   // pool, reused across profiles
  size_t events_used = 0;
  std::vector<int> event_kernel;    // kernel id per (start, stop) pair
  int64_t* dcount = nullptr;        // device counter scratch
  // reverse mode: checkpoints every adj_every substeps of the last
  // mpm_adjoint_forward (each the whole particle state in the slot order of a
  // binning made at the checkpoint, with that binning), and the constant
  // fields (m, V0, E0) in caller order
  float adj_dt = 0.f;
  int adj_steps = 0, adj_every = 0;
  struct AdjCheckpoint {
    float* rows = nullptr;  // NFIELDS rows, slot order
    int* pid = nullptr;     // n: caller index of each slot
    int4* blist = nullptr;  // nblocks: home-block list
    int* nblk = nullptr;    // 1
  };
  std::vector<AdjCheckpoint> adj_ckpt;
  float* adj_const = nullptr;
  int* adj_inv = nullptr;  // slot of each caller index (row transfers to caller order)
  // checkpoint_every == 1: every substep's record kept, segment by segment
  // (one binning each): starting states (+ the positions after the segment),
  // P2G sums, and the segment's slot order and block list
  struct AdjSegment {
    int ns = 0;
    float* st = nullptr;     // (ns + 1) x 33 rows
    float4* gs = nullptr;    // ns x nnodes: P2G sums of the active grid blocks
    int* gl = nullptr;       // ns x (nblocks + 1): their count and list
    float* sv = nullptr;     // ns x 21 rows: the F updates' SVDs
    int* pid = nullptr;      // n
    int4* blist = nullptr;   // nblocks
    int* nblk = nullptr;     // 1
  };
  std::vector<AdjSegment> adj_segs;  // kept between forward calls of the same shape
  // reverse-sweep workspaces, kept between calls (freed by mpm_destroy)
  struct AdjWork {
    float *Gc = nullptr, *Gs = nullptr, *Gs2 = nullptr, *gEs = nullptr, *gEc = nullptr, *tmp = nullptr;
    float4* gnode = nullptr;
    float4* vnode = nullptr;  // node velocities of the substep being reversed
    int* gmap = nullptr;
    float* seg = nullptr;     // (every + 1) x 33 rows
    float4* gsave = nullptr;  // every x nnodes
    int* glist = nullptr;     // every x (nblocks + 1)
    float* svd = nullptr;     // every x 21 rows: the F updates' SVDs
    int seg_every = 0;        // substeps seg / gsave / glist / svd hold
  } adj_ws;
  // positions readback: device staging, pinned host mirror, copy stream
  float* pos_stage = nullptr;
  float* pos_host = nullptr;     // pinned
  hipStream_t copy_stream = nullptr;
  hipEvent_t ev_gathered = nullptr, ev_copied = nullptr;
  bool pos_pending = false;
};

namespace mpm {
KArgs make_args(const mpm_ctx* c, float dt);
// legacy stages
void launch_clear(float4* p, int64_t count, hipStream_t s);
void launch_stage_p2g(const KArgs& a, hipStream_t s);
void launch_stage_normalise(const KArgs& a, hipStream_t s);
void launch_stage_volumes(const KArgs& a, hipStream_t s);
void launch_stage_forces(const KArgs& a, hipStream_t s);
void launch_stage_grid(const KArgs& a, hipStream_t s);
void launch_stage_collide(const KArgs& a, hipStream_t s);
void launch_stage_fupdate(const KArgs& a, hipStream_t s);
void launch_stage_g2p(const KArgs& a, hipStream_t s);
void launch_stage_advect(const KArgs& a, hipStream_t s);
// layout conversion (pid: caller index of each slot)
// slots [p0, p0 + np) of one scene; pid holds caller indices, src/dst are
// the scene's caller-ordered arrays (index pid - p0)
void launch_scatter_field(float* dst, int64_t stride, const float* src_aos, const int* pid, int p0, int np,
                          int comps, hipStream_t s);
void launch_gather_field(float* dst_aos, const float* src, int64_t stride, const int* pid, int p0, int np, int comps,
                         hipStream_t s);
void launch_iota(int* p, int n, hipStream_t s);
void launch_fill_int(int* p, int64_t n, int v, hipStream_t s);
void launch_grid_to_dense(const KArgs& a, const Scene& S, float* m, float* vel, float* force, uint8_t* active,
                         hipStream_t s);
void host_collider_inverse(const float t[3], const float q[4], float inv[4][4]);
void launch_count_active(const KArgs& a, int64_t* out, hipStream_t s);
// batched initial volumes: density gathered from the node masses in gvel.w
void launch_volumes(const KArgs& a, hipStream_t s);
void launch_scene_mass_max(const float* P, int64_t stride, const Scene* scenes, int nscenes, int max_np, uint32_t* out,
                           hipStream_t s);
// one field of one scene between the caller-ordered host array and the
// particle rows (synchronous); NULL host skips
int upload_field(mpm_ctx* ctx, int scene, int field, const float* host, int comps);
int download_field(mpm_ctx* ctx, int scene, int field, float* host, int comps);
// every particle's stencil inside its scene's grid (MPM_ERR_BOUNDS otherwise)
int check_scene_positions(mpm_ctx* ctx, int scene, const float* x);
// caller order again inside a scene's slots (slot p0 + i holds particle i)
int reset_scene_order(mpm_ctx* ctx, int scene);
}  // namespace mpm
