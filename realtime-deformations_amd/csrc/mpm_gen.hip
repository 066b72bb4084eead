// Device scene generators: seeded particle placement written straight into
// a scene's particle slots, so large scenes are set up without host loops or
// a host-to-device copy of the particle state.
//
// Randomness is counter-based (Philox4x64-10, Salmon et al., "Parallel
// random numbers: as easy as 1, 2, 3", SC'11 — the generator behind
// numpy.random.Philox): every draw is a pure function of (seed, stream tag,
// particle or candidate index, attempt), so results do not depend on launch
// shape or thread timing.  Floats are built only from exact conversions and
// single rounded + - * / sqrt, so a CPU restatement reproduces them bit for
// bit.
//
// Kinds (mpm_c.h, MPM_GEN_*):
//   BALL     n points uniform in a ball (rejection from the cube [-1, 1)^3);
//   CELLS    a box of cells with U{lo..hi} particles per cell, adjusted by +-1
//            steps in a seeded cell order to exactly n, uniform in each cell;
//   LATTICE  initializeParticles (realtime-deformations/material_point_method.cpp:18-63):
//            eight quarter-offset candidates per cell of the cube of half
//            width int(width / h) cells around the origin cell, each jittered
//            by a point uniform in a ball of radius 1/4 cell, kept when within
//            `width` of the origin, placed from the back of the particle array.
//            The reference draws the jitter from rand() (utils.h:110-127); here
//            it is a Philox draw, so the candidate set follows the reference's
//            rule but not its platform-dependent random sequence.
#include <cstdio>
#include <string>

#include "mpm_internal.hpp"

namespace mpm {
namespace {

constexpr uint64_t TAG_BALL = 1, TAG_CELL_COUNT = 2, TAG_CELL_POS = 3, TAG_LATTICE = 4, TAG_E0 = 5;
constexpr int MAX_TRIES = 64;  // P(no accept in 64 tries) = (1 - pi/6)^64 ~ 2e-21; then the centre

struct U4 {
  uint64_t w[4];
};

// Philox4x64 with 10 rounds: multipliers 0xD2E7470EE14C6C93 / 0xCA5A826395121157,
// Weyl key increments 0x9E3779B97F4A7C15 / 0xBB67AE8584CAA73B.  Counter
// (c0, c1, 0, 0), key (k0, k1).
__device__ U4 philox(uint64_t c0, uint64_t c1, uint64_t k0, uint64_t k1) {
  uint64_t x0 = c0, x1 = c1, x2 = 0, x3 = 0;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t lo0 = 0xD2E7470EE14C6C93ull * x0, hi0 = __umul64hi(0xD2E7470EE14C6C93ull, x0);
    const uint64_t lo1 = 0xCA5A826395121157ull * x2, hi1 = __umul64hi(0xCA5A826395121157ull, x2);
    const uint64_t y0 = hi1 ^ x1 ^ k0, y2 = hi0 ^ x3 ^ k1;
    x0 = y0;
    x1 = lo1;
    x2 = y2;
    x3 = lo0;
    k0 += 0x9E3779B97F4A7C15ull;
    k1 += 0xBB67AE8584CAA73Bull;
  }
  return U4{{x0, x1, x2, x3}};
}

// top 24 bits as a float in [0, 1), exact
__device__ __forceinline__ float u24(uint64_t w) { return (float)(uint32_t)(w >> 40) * 0x1p-24f; }

// a point uniform in the unit ball: first accepted candidate of the cube
// [-1, 1)^3 over attempts 0, 1, ... of counter (idx, attempt)
__device__ F3 unit_ball(uint64_t idx, uint64_t seed, uint64_t tag) {
  for (int t = 0; t < MAX_TRIES; ++t) {
    const U4 r = philox(idx, (uint64_t)t, seed, tag);
    const float a = 2.0f * u24(r.w[0]) - 1.0f, b = 2.0f * u24(r.w[1]) - 1.0f, c = 2.0f * u24(r.w[2]) - 1.0f;
    const float s = (a * a + b * b) + c * c;
    if (s <= 1.0f) return f3(a, b, c);
  }
  return f3(0.0f, 0.0f, 0.0f);
}

struct GenArgs {
  float* P;
  int64_t stride;
  int* pid;
  int p0, np;
  int dims[3];
  float h;
  uint64_t seed;
  mpm_generator g;
  int* bad;  // set when a generated particle's stencil leaves the grid
};

// one particle into slot p0 + i: x, v, m, E0; B = 0, F_E = F_P = I, V0 = 0;
// its caller index is the slot itself
__device__ void put_particle(const GenArgs& a, int i, F3 x, float E0) {
  const int64_t s = a.p0 + i, st = a.stride;
  float* P = a.P;
  P[(FX + 0) * st + s] = x.x;
  P[(FX + 1) * st + s] = x.y;
  P[(FX + 2) * st + s] = x.z;
#pragma unroll
  for (int k = 0; k < 3; ++k) P[(FV + k) * st + s] = a.g.velocity[k];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const float d = (k % 4 == 0) ? 1.0f : 0.0f;
    P[(FB + k) * st + s] = 0.0f;
    P[(FFE + k) * st + s] = d;
    P[(FFP + k) * st + s] = d;
  }
  P[FM * st + s] = a.g.mass;
  P[FVOL * st + s] = 0.0f;
  P[FE0 * st + s] = E0;
  a.pid[s] = (int)s;
  // the stencil base-2 .. base+2 (base = int(x / h)) must lie in the grid
  const float q[3] = {x.x / a.h, x.y / a.h, x.z / a.h};
#pragma unroll
  for (int k = 0; k < 3; ++k)
    if (!(q[k] >= 2.0f) || !(q[k] < (float)(a.dims[k] - 2))) atomicOr(a.bad, 1);
}

__device__ __forceinline__ float draw_E0(const GenArgs& a, float u) { return a.g.E_lo + (a.g.E_hi - a.g.E_lo) * u; }

__global__ void k_gen_ball(GenArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
  const F3 c = unit_ball((uint64_t)i, a.seed, TAG_BALL);
  const F3 x = f3(a.g.centre[0] + a.g.radius * c.x, a.g.centre[1] + a.g.radius * c.y,
                  a.g.centre[2] + a.g.radius * c.z);
  put_particle(a, i, x, draw_E0(a, u24(philox((uint64_t)i, 0, a.seed, TAG_E0).w[0])));
}

// CELLS: cell q = (a * cy + b) * cz + c of the box
__global__ void k_gen_cell_counts(GenArgs a, int nc, int* cnt) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nc) return;
  const uint64_t span = (uint64_t)(a.g.ppc_hi - a.g.ppc_lo + 1);
  const uint64_t w = philox((uint64_t)q, 0, a.seed, TAG_CELL_COUNT).w[0];
  cnt[q] = a.g.ppc_lo + (int)(((w >> 32) * span) >> 32);
}

// cells eligible for one +-1 step, in the seeded order pi(k) = (k A + B) mod nc
__global__ void k_gen_adjust_flags(GenArgs a, int nc, uint64_t A, uint64_t B, const int* cnt, const int* total,
                                   int* flag) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nc) return;
  const int d = *total - a.np;
  const int q = (int)(((uint64_t)k * A + B) % (uint64_t)nc);
  flag[k] = d > 0 ? (cnt[q] > a.g.ppc_lo) : (d < 0 ? (cnt[q] < a.g.ppc_hi) : 0);
}

__global__ void k_gen_adjust_apply(GenArgs a, int nc, uint64_t A, uint64_t B, int* cnt, const int* flag,
                                   const int* rank, const int* total) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nc) return;
  const int d = *total - a.np;
  const int need = d > 0 ? d : -d;
  if (!flag[k] || rank[k] >= need) return;
  const int q = (int)(((uint64_t)k * A + B) % (uint64_t)nc);
  cnt[q] += d > 0 ? -1 : 1;
}

// particle i: its cell from the exclusive offsets (last q with off[q] <= i),
// its rank r inside the cell, position uniform in the cell
__global__ void k_gen_cells(GenArgs a, int nc, const int* off) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
  int lo = 0, hi = nc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  const int q = lo, r = i - off[q];
  const int cz = a.g.cells[2], cy = a.g.cells[1];
  const int ci = q / (cy * cz), cj = (q / cz) % cy, ck = q % cz;
  const U4 w = philox((uint64_t)q, (uint64_t)r, a.seed, TAG_CELL_POS);
  const F3 x = f3(((float)(a.g.cell_lo[0] + ci) + u24(w.w[0])) * a.h, ((float)(a.g.cell_lo[1] + cj) + u24(w.w[1])) * a.h,
                  ((float)(a.g.cell_lo[2] + ck) + u24(w.w[2])) * a.h);
  put_particle(a, i, x, draw_E0(a, u24(w.w[3])));
}

// LATTICE candidate q = ((a * L + b) * L + c) * 8 + d, cell (c0 + a, c0 + b, c0 + c)
struct Lattice {
  int c0[3];
  int L;
};

__device__ F3 lattice_pos(const GenArgs& a, const Lattice& lt, int q) {
  const int d = q & 7, cell = q >> 3;
  const int ck = cell % lt.L, cj = (cell / lt.L) % lt.L, ci = cell / (lt.L * lt.L);
  // quarter offsets ds = {1 or 3} per axis in the order (1,1,1), (1,1,3), ... (3,3,3)
  const float ds0 = (d & 4) ? 3.0f : 1.0f, ds1 = (d & 2) ? 3.0f : 1.0f, ds2 = (d & 1) ? 3.0f : 1.0f;
  const F3 u = unit_ball((uint64_t)q, a.seed, TAG_LATTICE);
  const F3 jit = f3(0.25f * u.x, 0.25f * u.y, 0.25f * u.z);
  return f3((((float)(lt.c0[0] + ci) + ds0 * 0.25f) + jit.x) * a.h, (((float)(lt.c0[1] + cj) + ds1 * 0.25f) + jit.y) * a.h,
            (((float)(lt.c0[2] + ck) + ds2 * 0.25f) + jit.z) * a.h);
}

__global__ void k_gen_lattice_flags(GenArgs a, Lattice lt, int ncand, int* flag) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ncand) return;
  const F3 p = lattice_pos(a, lt, q);
  const float rx = p.x - a.g.centre[0], ry = p.y - a.g.centre[1], rz = p.z - a.g.centre[2];
  const float len = sqrtf((rx * rx + ry * ry) + rz * rz);
  flag[q] = !(len > a.g.radius);
}

__global__ void k_gen_lattice_place(GenArgs a, Lattice lt, int ncand, const int* flag, const int* rank) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ncand || !flag[q] || rank[q] >= a.np) return;
  put_particle(a, a.np - 1 - rank[q], lattice_pos(a, lt, q), draw_E0(a, u24(philox((uint64_t)q, 0, a.seed, TAG_E0).w[0])));
}

// Exclusive prefix sum of n ints (in place allowed) by one workgroup walking
// tiles of SCAN_T * SCAN_ITEMS; *total = the sum.  Setup-time only.
constexpr int SCAN_T = 1024, SCAN_ITEMS = 8, SCAN_WAVES = SCAN_T / 64;
__global__ __launch_bounds__(SCAN_T) void k_scan_single(const int* in, int* out, int n, int* total) {
  __shared__ int wsum[SCAN_WAVES];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  int carry = 0;
  for (int base = 0; base < n; base += SCAN_T * SCAN_ITEMS) {
    const int b = base + t * SCAN_ITEMS;
    int v[SCAN_ITEMS], s = 0;
#pragma unroll
    for (int u = 0; u < SCAN_ITEMS; ++u) {
      v[u] = (b + u < n) ? in[b + u] : 0;
      s += v[u];
    }
    int incl = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(incl, o);
      if (lane >= o) incl += y;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    if (wv == 0) {
      int x = lane < SCAN_WAVES ? wsum[lane] : 0;
#pragma unroll
      for (int o = 1; o < SCAN_WAVES; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
      }
      if (lane < SCAN_WAVES) wsum[lane] = x;
    }
    __syncthreads();
    int run = carry + (wv ? wsum[wv - 1] : 0) + incl - s;
#pragma unroll
    for (int u = 0; u < SCAN_ITEMS; ++u)
      if (b + u < n) {
        out[b + u] = run;
        run += v[u];
      }
    carry += wsum[SCAN_WAVES - 1];
    __syncthreads();
  }
  if (t == 0) *total = carry;
}

// Exclusive prefix sum of n ints (in place allowed) over many workgroups:
// the sums of tiles of PS_T * PS_ITEMS ints, k_scan_single over those sums,
// then each tile's local scan offset by its sum.  *total = the sum.
constexpr int PS_T = 256, PS_ITEMS = 8, PS_TILE = PS_T * PS_ITEMS;
inline int scan_tiles(int n) { return (n + PS_TILE - 1) / PS_TILE; }
__global__ __launch_bounds__(PS_T) void k_scan_tile_sums(const int* in, int n, int* wsum) {
  __shared__ int ws[PS_T / 64];
  const int base = blockIdx.x * PS_TILE + threadIdx.x * PS_ITEMS;
  int s = 0;
#pragma unroll
  for (int u = 0; u < PS_ITEMS; ++u) s += (base + u < n) ? in[base + u] : 0;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
#pragma unroll
    for (int w = 0; w < PS_T / 64; ++w) t += ws[w];
    wsum[blockIdx.x] = t;
  }
}
__global__ __launch_bounds__(PS_T) void k_scan_tiles(const int* in, int* out, int n, const int* woff) {
  __shared__ int ws[PS_T / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int base = blockIdx.x * PS_TILE + threadIdx.x * PS_ITEMS;
  int v[PS_ITEMS], s = 0;
#pragma unroll
  for (int u = 0; u < PS_ITEMS; ++u) {
    v[u] = (base + u < n) ? in[base + u] : 0;
    s += v[u];
  }
  int incl = s;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o);
    if (lane >= o) incl += y;
  }
  if (lane == 63) ws[wv] = incl;
  __syncthreads();
  int run = woff[blockIdx.x] + incl - s;
  for (int w = 0; w < wv; ++w) run += ws[w];
#pragma unroll
  for (int u = 0; u < PS_ITEMS; ++u)
    if (base + u < n) {
      out[base + u] = run;
      run += v[u];
    }
}
// wsum: scan_tiles(n) ints of scratch
void launch_scan(const int* in, int* out, int n, int* total, int* wsum, hipStream_t st) {
  const int nt = scan_tiles(n);
  if (nt == 0) {
    k_scan_single<<<1, SCAN_T, 0, st>>>(in, out, 0, total);
    return;
  }
  k_scan_tile_sums<<<nt, PS_T, 0, st>>>(in, n, wsum);
  k_scan_single<<<1, SCAN_T, 0, st>>>(wsum, wsum, nt, total);
  k_scan_tiles<<<nt, PS_T, 0, st>>>(in, out, n, wsum);
}

inline unsigned grid_for(int64_t n) { return (unsigned)((n + 255) / 256); }

uint64_t gcd64(uint64_t a, uint64_t b) {
  while (b) {
    const uint64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

}  // namespace

// the seeded cell order's multiplier: the first odd A >= (2654435761 mod nc)
// coprime with nc, so k -> (k A + B) mod nc is a permutation
uint64_t gen_cell_order_multiplier(uint64_t nc) {
  uint64_t A = (2654435761ull % nc) | 1ull;
  while (gcd64(A, nc) != 1) A += 2;
  return A;
}

}  // namespace mpm

using namespace mpm;

#define GEN_HIP(call)                                                       \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) {                                                 \
      ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);        \
      rc = MPM_ERR_HIP;                                                     \
      goto done;                                                            \
    }                                                                       \
  } while (0)

#define GEN_FAIL(code, msg) \
  do {                      \
    ctx->err = (msg);       \
    rc = (code);            \
    goto done;              \
  } while (0)

extern "C" {

int mpm_scene_generate(mpm_ctx* ctx, int32_t scene, const mpm_generator* g, int64_t* candidates) {
  if (!ctx) return MPM_ERR_ARG;
  if (scene < 0 || scene >= ctx->nscenes) {
    ctx->err = "scene index out of range";
    return MPM_ERR_ARG;
  }
  if (!g) {
    ctx->err = "null generator";
    return MPM_ERR_ARG;
  }
  const Scene& S = ctx->scenes[scene];
  int rc = MPM_OK;
  int *bad = nullptr, *tot = nullptr, *buf = nullptr;
  int bad_h = 0;
  GenArgs a{};
  a.P = ctx->P;
  a.stride = ctx->stride;
  a.pid = ctx->pid;
  a.p0 = S.p0;
  a.np = S.np;
  for (int k = 0; k < 3; ++k) a.dims[k] = S.dims[k];
  a.h = S.h;
  a.seed = g->seed;
  a.g = *g;
  if (candidates) *candidates = 0;
  if (!(g->E_hi >= g->E_lo) || !(g->mass >= 0.0f)) {
    ctx->err = "generator needs E_hi >= E_lo and mass >= 0";
    return MPM_ERR_ARG;
  }
  if (g->kind != MPM_GEN_BALL && g->kind != MPM_GEN_CELLS && g->kind != MPM_GEN_LATTICE) {
    ctx->err = "unknown generator kind";
    return MPM_ERR_ARG;
  }
  GEN_HIP(hipStreamSynchronize(ctx->stream));
  GEN_HIP(hipMalloc(&bad, 4 * sizeof(int)));
  tot = bad + 1;
  a.bad = bad;
  GEN_HIP(hipMemsetAsync(bad, 0, 4 * sizeof(int), ctx->stream));
  if (g->kind == MPM_GEN_BALL) {
    if (S.np) k_gen_ball<<<grid_for(S.np), 256, 0, ctx->stream>>>(a);
    if (candidates) *candidates = S.np;
  } else if (g->kind == MPM_GEN_CELLS) {
    const int64_t nc64 = (int64_t)g->cells[0] * g->cells[1] * g->cells[2];
    if (g->cells[0] <= 0 || g->cells[1] <= 0 || g->cells[2] <= 0 || g->ppc_lo < 0 || g->ppc_hi < g->ppc_lo ||
        nc64 * (int64_t)g->ppc_hi >= ((int64_t)1 << 31))
      GEN_FAIL(MPM_ERR_ARG, "CELLS needs a non-empty box, 0 <= ppc_lo <= ppc_hi and fewer than 2^31 particles");
    if ((int64_t)S.np < nc64 * g->ppc_lo || (int64_t)S.np > nc64 * g->ppc_hi)
      GEN_FAIL(MPM_ERR_ARG, "CELLS: the scene's particle count is outside [ppc_lo, ppc_hi] x cells");
    const int nc = (int)nc64;
    const uint64_t A = gen_cell_order_multiplier((uint64_t)nc), B = g->seed % (uint64_t)nc;
    GEN_HIP(hipMalloc(&buf, sizeof(int) * (3 * (size_t)nc + scan_tiles(nc))));
    int *cnt = buf, *off = buf + nc, *flag = buf + 2 * nc, *wsum = buf + 3 * (size_t)nc;
    k_gen_cell_counts<<<grid_for(nc), 256, 0, ctx->stream>>>(a, nc, cnt);
    // +-1 passes in the seeded cell order until the total is exactly n (one
    // pass unless |total - n| exceeds the cells that can still move)
    for (int pass = 0;; ++pass) {
      int total = 0;
      launch_scan(cnt, off, nc, tot, wsum, ctx->stream);
      GEN_HIP(hipMemcpyAsync(&total, tot, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
      GEN_HIP(hipStreamSynchronize(ctx->stream));
      if (total == S.np) break;
      if (pass > g->ppc_hi - g->ppc_lo) GEN_FAIL(MPM_ERR_HIP, "CELLS: count adjustment did not converge");
      k_gen_adjust_flags<<<grid_for(nc), 256, 0, ctx->stream>>>(a, nc, A, B, cnt, tot, flag);
      launch_scan(flag, off, nc, tot + 1, wsum, ctx->stream);
      k_gen_adjust_apply<<<grid_for(nc), 256, 0, ctx->stream>>>(a, nc, A, B, cnt, flag, off, tot);
    }
    if (S.np) k_gen_cells<<<grid_for(S.np), 256, 0, ctx->stream>>>(a, nc, off);
    if (candidates) *candidates = S.np;
  } else {
    // initializeParticles: origin cell ivec3(origin / h), half width int(width / h)
    Lattice lt{};
    const int w = (int)(g->radius / S.h);
    for (int k = 0; k < 3; ++k) lt.c0[k] = (int)(g->centre[k] / S.h) - w;
    lt.L = 2 * (w > 0 ? w : 0);
    const int64_t nq = (int64_t)lt.L * lt.L * lt.L * 8;
    if (nq >= ((int64_t)1 << 31)) GEN_FAIL(MPM_ERR_ARG, "LATTICE: too many candidates (width / h too large)");
    const int ncand = (int)nq;
    int kept = 0;
    if (ncand) {
      GEN_HIP(hipMalloc(&buf, sizeof(int) * (2 * (size_t)ncand + scan_tiles(ncand))));
      int *flag = buf, *rank = buf + ncand, *wsum = buf + 2 * (size_t)ncand;
      k_gen_lattice_flags<<<grid_for(ncand), 256, 0, ctx->stream>>>(a, lt, ncand, flag);
      launch_scan(flag, rank, ncand, tot, wsum, ctx->stream);
      GEN_HIP(hipMemcpyAsync(&kept, tot, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
      GEN_HIP(hipStreamSynchronize(ctx->stream));
    }
    if (candidates) *candidates = kept;
    if (kept < S.np) {
      char m[160];
      snprintf(m, sizeof m, "LATTICE: %d candidates lie within the width, fewer than the scene's %d particles", kept,
               S.np);
      GEN_FAIL(MPM_ERR_ARG, m);
    }
    if (ncand) k_gen_lattice_place<<<grid_for(ncand), 256, 0, ctx->stream>>>(a, lt, ncand, buf, buf + ncand);
  }
  GEN_HIP(hipGetLastError());
  GEN_HIP(hipMemcpyAsync(&bad_h, bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  GEN_HIP(hipStreamSynchronize(ctx->stream));
  ctx->since_sort = -1;
  if (bad_h) {
    ctx->scene_set[scene] = 0;
    ctx->particles_set = false;
    GEN_FAIL(MPM_ERR_BOUNDS, "a generated particle puts its stencil outside the scene's grid");
  }
  ctx->scene_set[scene] = 1;
  ctx->quanta_stale = true;
  ctx->particles_set = true;
  for (char c : ctx->scene_set) ctx->particles_set = ctx->particles_set && c;
done:
  if (buf) (void)hipFree(buf);
  if (bad) (void)hipFree(bad);
  return rc;
}

int mpm_generate(mpm_ctx* ctx, const mpm_generator* g, int64_t* candidates) {
  return mpm_scene_generate(ctx, 0, g, candidates);
}

}  // extern "C"
