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

enum StatusBits : int { ST_NONFINITE = 1, ST_BOUNDS = 2 };

// Everything a kernel needs, passed by value.
struct KArgs {
  float* P;
  int64_t stride;
  int n;
  float4* gacc;    // (m, p) accumulators
  float4* gvel;    // (v, m) after the grid update
  float4* gforce;  // (f, 0), legacy stage path only
  int dims[3];
  int nb[3];
  int64_t nnodes;  // nb0*nb1*nb2*64
  float h, inv_h, dt, lo, hi, gap64;  // inv_h: 1/h correctly rounded
  M3 Dinv;
  Lame lame;
  float g[3];
  int compose, bc_mode, ncoll;
  const DevCollider* coll;
  int* status;
};

__device__ __forceinline__ int64_t node_index(int i, int j, int k, const int nb[3]) {
  return ((int64_t)((i >> 2) * nb[1] + (j >> 2)) * nb[2] + (k >> 2)) * 64 + ((i & 3) << 4) + ((j & 3) << 2) +
         (k & 3);
}

}  // namespace mpm

struct mpm_ctx {
  mpm_params prm{};
  int n = 0;
  int64_t stride = 0;
  int dims[3] = {0, 0, 0};
  int nb[3] = {0, 0, 0};
  int64_t nnodes = 0;
  float h = 0, lo = 0, hi = 0;
  mpm::M3 Dinv{};
  float* P = nullptr;       // particle fields (binned order once sorted)
  float* P2 = nullptr;      // reorder target, swapped with P
  int* pid = nullptr;       // caller index of each particle slot
  int* pid2 = nullptr;
  // binning
  uint32_t *keys = nullptr, *vals = nullptr, *keys2 = nullptr, *vals2 = nullptr, *hist = nullptr;
  int *blk_start = nullptr, *blk_end = nullptr, *blk_ids = nullptr, *nblk = nullptr;
  int4* blist = nullptr;
  // active grid blocks of the fused path
  int* gflag = nullptr;
  int* glist[2] = {nullptr, nullptr};
  int* gcnt = nullptr;      // three rotating counters
  int* dstep = nullptr;     // device substep epoch (> 0)
  int since_sort = -1;      // substeps since the last binning (-1: never binned)
  int resort_every = 16;
  float4* gacc = nullptr;
  float4* gvel = nullptr;
  float4* gforce = nullptr;
  mpm::DevCollider* coll = nullptr;
  int ncoll = 0;
  int* status = nullptr;  // device status word
  hipStream_t stream = nullptr;
  bool particles_set = false;
  bool gacc_clean = false;  // fused path needs zeroed accumulators
  std::string err;
  // HIP graphs of [rebin + resort_every substeps], one per particle-buffer
  // parity, valid for one dt and collider table
  float* P_first = nullptr;  // P as allocated: parity 0 when P == P_first
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  float graph_dt[2] = {0.f, 0.f};
  const void* graph_coll[2] = {nullptr, nullptr};
  int graph_ncoll[2] = {-1, -1};
  bool use_graphs = true;
  // event profiling of fused-path launches
  bool profiling = false;
  bool per_kernel = true;
  hipEvent_t ev_region[2] = {nullptr, nullptr};  // profiled region begin, end
  std::vector<hipEvent_t> events;   // pool, reused across profiles
  size_t events_used = 0;
  std::vector<int> event_kernel;    // kernel id per (start, stop) pair
  int64_t* dcount = nullptr;        // device counter scratch
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
void launch_scatter_field(float* dst, int64_t stride, const float* src_aos, const int* pid, int n, int comps,
                          hipStream_t s);
void launch_gather_field(float* dst_aos, const float* src, int64_t stride, const int* pid, int n, int comps,
                         hipStream_t s);
void launch_iota(int* p, int n, hipStream_t s);
void launch_fill_int(int* p, int64_t n, int v, hipStream_t s);
void launch_grid_to_dense(const KArgs& a, float* m, float* vel, float* force, uint8_t* active, hipStream_t s);
void host_collider_inverse(const float t[3], const float q[4], float inv[4][4]);
void launch_count_active(const KArgs& a, int64_t* out, hipStream_t s);
}  // namespace mpm
