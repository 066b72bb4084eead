// Particle binning for the fused substep (gfx950).
//
// key(p) = block(p) * 64 + cell-in-block(p), with cell = int(x / h) per axis
// (correctly rounded FP32 quotient, as material_point_method.cpp:83) and
// 4x4x4-cell blocks numbered blk0 + (bi * nb1 + bj) * nb2 + bk, blk0 being the
// first block of the particle's scene: one sort bins every scene.  A stable LSD radix
// sort (8-bit digits) orders (key, index) pairs exactly like a CPU stable
// sort: ties keep index order.  The non-empty blocks are listed from
// per-workgroup counts and a scan (no atomics), in a spread claim order
// (claim_slot).  Each block's particles are then moved into one contiguous
// range in column groups (see the layout note above k_reorder), so the lanes
// of each 16-lane group start their stencils in different (y, z) columns and
// the FP64 LDS atomics of P2G spread over the bank pairs.
#include "mpm_bin.hpp"

namespace mpm {
namespace {

constexpr int RS_T = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_T * RS_ITEMS;

__global__ void __launch_bounds__(256) k_keys(BinArgs b) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= b.n) return;
  const Scene S = b.scenes[scene_of_slot(b.scenes, b.nscenes, p)];
  int c[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float q = b.P[(int64_t)(FX + a) * b.stride + p] / S.h;
    int v = (int)q;
    v = v < 1 ? 1 : (v > S.dims[a] - 3 ? S.dims[a] - 3 : v);
    c[a] = v;
  }
  const uint32_t blk = (uint32_t)(S.blk0 + ((c[0] >> 2) * S.nb[1] + (c[1] >> 2)) * S.nb[2] + (c[2] >> 2));
  b.keys[p] = blk * 64u + (uint32_t)(((c[0] & 3) << 4) | ((c[1] & 3) << 2) | (c[2] & 3));
  b.vals[p] = (uint32_t)p;
}

__global__ void __launch_bounds__(RS_T) k_rs_hist(const uint32_t* keys, int n, int shift, uint32_t* hist, int ntiles) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll
  for (int i = 0; i < RS_ITEMS; ++i) {
    const int64_t idx = base + i * RS_T + threadIdx.x;
    if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Inclusive scan of one value per thread across a 1024-thread workgroup.
__device__ __forceinline__ uint32_t block_incl_scan_1024(uint32_t v, uint32_t* wsum) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  if (lane == 63) wsum[wave] = v;
  __syncthreads();
  if (wave == 0) {
    uint32_t s = lane < 16 ? wsum[lane] : 0u;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
      const uint32_t u = __shfl_up(s, off, 64);
      if (lane >= off) s += u;
    }
    if (lane < 16) wsum[lane] = s;
  }
  __syncthreads();
  const uint32_t r = v + (wave > 0 ? wsum[wave - 1] : 0u);
  __syncthreads();
  return r;
}

// Per digit (one workgroup each): exclusive scan over the tiles' counts in
// place, and the digit's total in dtot[digit].
__global__ void __launch_bounds__(1024) k_rs_scan_digit(uint32_t* hist, int ntiles, uint32_t* dtot) {
  __shared__ uint32_t wsum[16];
  uint32_t* h = hist + (int64_t)blockIdx.x * ntiles;
  uint32_t carry = 0;
  for (int base = 0; base < ntiles; base += 1024) {
    const int i = base + threadIdx.x;
    const uint32_t v = i < ntiles ? h[i] : 0u;
    const uint32_t incl = block_incl_scan_1024(v, wsum);
    if (i < ntiles) h[i] = carry + incl - v;
    // total of this segment = inclusive value of the last thread
    __shared__ uint32_t seg_total;
    if (threadIdx.x == 1023) seg_total = incl;
    __syncthreads();
    carry += seg_total;
    __syncthreads();
  }
  if (threadIdx.x == 0) dtot[blockIdx.x] = carry;
}

__global__ void __launch_bounds__(RS_T) k_rs_scatter(const uint32_t* keys, const uint32_t* vals, uint32_t* okeys,
                                                     uint32_t* ovals, int n, int shift, const uint32_t* hist,
                                                     int ntiles, const uint32_t* dtot) {
  __shared__ uint32_t skeys[RS_TILE];
  __shared__ uint32_t svals[RS_TILE];
  __shared__ uint32_t wcnt[4][256];
  __shared__ uint32_t running[256];
  __shared__ uint32_t dstart[256];
  __shared__ uint32_t gbase[256];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  running[t] = 0;
  // digit bases: exclusive scan of the 256 digit totals (wave 0, 4 per lane)
  if (t < 64) {
    const uint32_t c0 = dtot[4 * t], c1 = dtot[4 * t + 1], c2 = dtot[4 * t + 2], c3 = dtot[4 * t + 3];
    const uint32_t s4 = c0 + c1 + c2 + c3;
    uint32_t incl = s4;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t u = __shfl_up(incl, off, 64);
      if (t >= off) incl += u;
    }
    const uint32_t ex = incl - s4;
    gbase[4 * t] = ex;
    gbase[4 * t + 1] = ex + c0;
    gbase[4 * t + 2] = ex + c0 + c1;
    gbase[4 * t + 3] = ex + c0 + c1 + c2;
  }
  __syncthreads();
  gbase[t] += hist[(int64_t)t * ntiles + blockIdx.x];
#pragma unroll
  for (int w = 0; w < 4; ++w) wcnt[w][t] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  uint32_t rk[RS_ITEMS], rv[RS_ITEMS], rd[RS_ITEMS], rp[RS_ITEMS];
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    const int64_t idx = base + r * RS_T + t;
    const bool valid = idx < n;
    const uint32_t k = valid ? keys[idx] : 0u;
    const uint32_t d = (k >> shift) & 255u;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const bool set = (d >> bit) & 1u;
      const uint64_t bb = __ballot(valid && set);
      m &= set ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t)__popcll(m & lt);
    if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t pre = running[d];
    for (int w = 0; w < wave; ++w) pre += wcnt[w][d];
    rk[r] = k;
    rv[r] = valid ? vals[idx] : 0u;
    rd[r] = valid ? d : 256u;
    rp[r] = pre + rank;
    __syncthreads();
    running[t] += wcnt[0][t] + wcnt[1][t] + wcnt[2][t] + wcnt[3][t];
#pragma unroll
    for (int w = 0; w < 4; ++w) wcnt[w][t] = 0;
    __syncthreads();
  }
  // digit starts inside the tile
  if (t < 64) {
    // wave 0 scans the 256 counts, 4 per lane
    uint32_t c0 = running[4 * t], c1 = running[4 * t + 1], c2 = running[4 * t + 2], c3 = running[4 * t + 3];
    uint32_t s = c0 + c1 + c2 + c3;
    uint32_t incl = s;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t v = __shfl_up(incl, off, 64);
      if (t >= off) incl += v;
    }
    const uint32_t ex = incl - s;
    dstart[4 * t] = ex;
    dstart[4 * t + 1] = ex + c0;
    dstart[4 * t + 2] = ex + c0 + c1;
    dstart[4 * t + 3] = ex + c0 + c1 + c2;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    if (rd[r] < 256u) {
      const uint32_t pos = dstart[rd[r]] + rp[r];
      skeys[pos] = rk[r];
      svals[pos] = rv[r];
    }
  }
  __syncthreads();
  const int tile_n = (int)min((int64_t)RS_TILE, (int64_t)n - base);
  for (int i = t; i < tile_n; i += RS_T) {
    const uint32_t k = skeys[i];
    const uint32_t d = (k >> shift) & 255u;
    const uint32_t o = gbase[d] + ((uint32_t)i - dstart[d]);
    okeys[o] = k;
    ovals[o] = svals[i];
  }
}

// First and last sorted index of every non-empty block; compact block list.
// Block boundaries of the sorted keys: blk_start / blk_end per grid block,
// and the number of blocks that start in each workgroup (wgcnt).
__global__ void __launch_bounds__(256) k_block_detect(BinArgs b, uint32_t* wgcnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool start = false;
  if (i < b.n) {
    const uint32_t blk = b.keys[i] >> 6;
    start = i == 0 || (b.keys[i - 1] >> 6) != blk;
    if (start) b.blk_start[blk] = i;
    if (i == b.n - 1 || (b.keys[i + 1] >> 6) != blk) b.blk_end[blk] = i + 1;
  }
  const int c = __syncthreads_count(start);
  if (threadIdx.x == 0) wgcnt[blockIdx.x] = (uint32_t)c;
}

// Exclusive scan of the per-workgroup block counts in place (one workgroup,
// BS_ITEMS consecutive counts per thread) and their total into *nblk.
constexpr int BS_T = 1024, BS_ITEMS = 8;
__global__ void __launch_bounds__(BS_T) k_block_scan(uint32_t* wgcnt, int nwg, int* nblk) {
  __shared__ uint32_t wsum[BS_T / 64];
  __shared__ uint32_t carry;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  if (t == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nwg; base += BS_T * BS_ITEMS) {
    uint32_t v[BS_ITEMS], sum = 0;
#pragma unroll
    for (int k = 0; k < BS_ITEMS; ++k) {
      const int i = base + t * BS_ITEMS + k;
      v[k] = i < nwg ? wgcnt[i] : 0u;
      sum += v[k];
    }
    uint32_t x = sum;  // inclusive scan over the wave
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(x, off, 64);
      if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    if (t < 64) {
      uint32_t w = t < BS_T / 64 ? wsum[t] : 0u;
#pragma unroll
      for (int off = 1; off < BS_T / 64; off <<= 1) {
        const uint32_t y = __shfl_up(w, off, 64);
        if (lane >= off) w += y;
      }
      if (t < BS_T / 64) wsum[t] = w;
    }
    __syncthreads();
    uint32_t run = carry + (wave ? wsum[wave - 1] : 0u) + x - sum;
#pragma unroll
    for (int k = 0; k < BS_ITEMS; ++k) {
      const int i = base + t * BS_ITEMS + k;
      if (i < nwg) wgcnt[i] = run;
      run += v[k];
    }
    __syncthreads();
    if (t == 0) carry += wsum[BS_T / 64 - 1];
    __syncthreads();
  }
  if (t == 0) *nblk = (int)carry;
}

// The block passes claim home blocks in list order.  Rank r in key order
// goes to list slot (r * stride) mod nblk, so that blocks claimed at the same
// time lie apart and do not flush into the same grid nodes at once.
// The stride is the first of three primes that does not divide nblk; nblk
// < 2^31 cannot be a multiple of all three (their product exceeds 2^38), so
// the map is always a permutation.
constexpr uint32_t CLAIM_STRIDES[3] = {7919, 7907, 7901};
__device__ __forceinline__ uint32_t claim_slot(uint32_t r, uint32_t nblk) {
  const uint64_t m = nblk % CLAIM_STRIDES[0] ? CLAIM_STRIDES[0]
                     : nblk % CLAIM_STRIDES[1] ? CLAIM_STRIDES[1] : CLAIM_STRIDES[2];
  return (uint32_t)(((uint64_t)r * m) % nblk);
}

// Compact list of the non-empty blocks: the block starting in workgroup w
// has rank wgoff[w] + its rank among that workgroup's starts in key order.
__global__ void __launch_bounds__(256) k_block_ids(BinArgs b, const uint32_t* wgoff) {
  __shared__ uint32_t wcnt[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool start = i < b.n && (i == 0 || (b.keys[i - 1] >> 6) != (b.keys[i] >> 6));
  const unsigned long long m = __ballot(start);
  if (lane == 0) wcnt[wave] = (uint32_t)__popcll(m);
  __syncthreads();
  if (start) {
    uint32_t r = wgoff[blockIdx.x];
    for (int w = 0; w < wave; ++w) r += wcnt[w];
    r += (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    b.blk_ids[claim_slot(r, (uint32_t)*b.nblk)] = (int)(b.keys[i] >> 6);
  }
}

// Slot layout inside a block, chosen for the LDS banks of the block pass.
// A particle's stencil origin in the halo sets the FP64 bank pair of all its
// atomics by its cell's column (cy, cz); a wave's 64 lanes issue an atomic as
// four 16-lane groups, each free of conflicts when its 16 lanes come from 16
// different columns.  The slots are therefore cut into groups of 16, and
// each group takes one particle from every column that still has particles;
// when fewer than 16 columns remain, the group's spare slots go one by one
// to the column with the most particles left (lowest column on ties).
// Inside a group the columns come in order.  A column's k-th taken particle
// is entry k of its list, which is cell-major: the particles of cell cx = 0
// in sorted order, then cx = 1, 2, 3.  So the same lane of consecutive
// groups mostly holds particles of one cell, and G2P reads of the two 16-lane
// groups that issue together hit one address (a broadcast).
//
// Tables (one wave builds them): per group and column the first slot in the
// group and the list position at the group's start.  Threads then walk
// destination slots (coalesced writes) and find their sources by table
// lookups.  Blocks with a cell of more than RR_MAX particles or more than
// GMAX groups take the plain round-robin of the per-source formula instead
// (slot of the r-th particle of cell c: sum_c' min(cnt[c'], r) +
// #{c' < c : cnt[c'] > r}).
constexpr int RR_MAX = 64;
constexpr int GMAX = 256;
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}
// Particle fields and caller index from slot src (current order) to slot
// dst (binned order), in two halves whose loads are all issued before their
// stores (half the fields at a time keeps the register count, and with it
// the number of blocks being planned at once, up).
__device__ __forceinline__ void move_particle(const BinArgs& b, int src, int dst) {
  const float* __restrict__ P = b.P;
  float* __restrict__ Pout = b.Pout;
  constexpr int HALF = NFIELDS / 2;
#pragma unroll 1
  for (int f0 = 0; f0 < NFIELDS; f0 += HALF) {
    float v[HALF];
#pragma unroll
    for (int f = 0; f < HALF; ++f) v[f] = P[(int64_t)(f0 + f) * b.stride + src];
#pragma unroll
    for (int f = 0; f < HALF; ++f) Pout[(int64_t)(f0 + f) * b.stride + dst] = v[f];
  }
  b.pid_out[dst] = b.pid[src];
}

constexpr int RO_T = 256;
__global__ void __launch_bounds__(RO_T) k_reorder(BinArgs b) {
  __shared__ int cnt[64];
  __shared__ int cstart[64];
  __shared__ unsigned char gofs[GMAX][16];
  __shared__ unsigned short gptr[GMAX][16];
  __shared__ int table_s;
  const int t = threadIdx.x;
  const int nblk = *b.nblk;
  for (int r = blockIdx.x; r < nblk; r += gridDim.x) {  // blocks in key order
    const int s = (int)claim_slot((uint32_t)r, (uint32_t)nblk);
    const int blk = b.blk_ids[s];
    const int start = b.blk_start[blk], end = b.blk_end[blk];
    const int nb = end - start;
    if (t < 64) cnt[t] = 0;
    __syncthreads();
    for (int i = start + t; i < end; i += RO_T) atomicAdd(&cnt[b.keys[i] & 63u], 1);
    __syncthreads();
    if (t < 64) {
      const int c = cnt[t];
      int incl = c;
      for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(incl, off, 64);
        if (t >= off) incl += v;
      }
      cstart[t] = incl - c;
      const int maxc = wave_max(c);
      const bool table = maxc <= RR_MAX && (nb + 15) / 16 <= GMAX;
      if (t == 0) table_s = table ? 1 : 0;
      if (table) {
        // groups of 16 slots (the whole wave runs the reductions; lanes 0..15
        // hold column t, lanes >= 16 nothing)
        int rem = t < 16 ? cnt[t] + cnt[16 | t] + cnt[32 | t] + cnt[48 | t] : 0, ptr = 0, left = nb;
        for (int gi = 0; left > 0; ++gi) {
          const int size = left < 16 ? left : 16;
          int take = rem > 0 ? 1 : 0;
          const int k = __popcll(__ballot(take != 0));
          for (int e = size - k; e > 0; --e) {
            const int key = (t < 16 && rem - take > 0) ? (((rem - take) << 6) | (63 - t)) : -1;
            const int best = wave_max(key);
            if (key >= 0 && key == best) ++take;
          }
          int pre = take;
#pragma unroll
          for (int off = 1; off < 16; off <<= 1) {
            const int v = __shfl_up(pre, off, 64);
            if ((t & 15) >= off) pre += v;
          }
          if (t < 16) {
            gofs[gi][t] = (unsigned char)(pre - take);
            gptr[gi][t] = (unsigned short)ptr;
          }
          ptr += take;
          rem -= take;
          left -= size;
        }
      }
    }
    __syncthreads();
    if (table_s) {
      for (int dst = start + t; dst < end; dst += RO_T) {
        const int j = dst - start, gi = j >> 4, r = j & 15;
        int g = 15;
        while (gofs[gi][g] > r) --g;
        const int p = gptr[gi][g] + (r - gofs[gi][g]);
        int cx = 0, d = p;  // entry p of the cell-major list: particle d of cell cx
        while (cx < 3 && d >= cnt[(cx << 4) | g]) d -= cnt[(cx++ << 4) | g];
        const int c = (cx << 4) | g;
        move_particle(b, (int)b.vals[start + cstart[c] + d], dst);
      }
    } else {
      for (int i = start + t; i < end; i += RO_T) {
        const int c = (int)(b.keys[i] & 63u);
        const int r = i - start - cstart[c];
        int pos = 0;
        for (int cc = 0; cc < 64; ++cc) {
          const int k = cnt[cc];
          pos += (k < r ? k : r) + ((cc < c && k > r) ? 1 : 0);
        }
        move_particle(b, (int)b.vals[i], start + pos);
      }
    }
    if (t == 0) b.blist[s] = make_int4(blk, start, end, b.blk_scene[blk]);
    __syncthreads();
  }
}

}  // namespace

int sort_passes(int64_t nnodes) {
  int bits = 0;
  while (bits < 32 && ((int64_t)1 << bits) < nnodes) ++bits;
  return (bits + 7) / 8;
}

void launch_bin(BinArgs b, hipStream_t s) {
  if (b.n == 0) return;
  const int nb = (b.n + 255) / 256;
  k_keys<<<nb, 256, 0, s>>>(b);
  const int ntiles = (b.n + RS_TILE - 1) / RS_TILE;
  const int passes = sort_passes(b.nnodes);
  uint32_t *k0 = b.keys, *v0 = b.vals, *k1 = b.keys2, *v1 = b.vals2;
  for (int pass = 0; pass < passes; ++pass) {
    k_rs_hist<<<ntiles, RS_T, 0, s>>>(k0, b.n, 8 * pass, b.hist, ntiles);
    k_rs_scan_digit<<<256, 1024, 0, s>>>(b.hist, ntiles, b.hist + 256 * (int64_t)ntiles);
    k_rs_scatter<<<ntiles, RS_T, 0, s>>>(k0, v0, k1, v1, b.n, 8 * pass, b.hist, ntiles,
                                         b.hist + 256 * (int64_t)ntiles);
    std::swap(k0, k1);
    std::swap(v0, v1);
  }
  BinArgs c = b;
  c.keys = k0;
  c.vals = v0;
  // per-workgroup block counts in the sort's histogram space (free after the
  // sort; 256 * ntiles >= nb)
  k_block_detect<<<nb, 256, 0, s>>>(c, b.hist);
  k_block_scan<<<1, BS_T, 0, s>>>(b.hist, nb, b.nblk);
  k_block_ids<<<nb, 256, 0, s>>>(c, b.hist);
  const int grid = (int)std::min<int64_t>(std::max<int64_t>(1, (int64_t)b.n / 64), 4096);
  k_reorder<<<grid, RO_T, 0, s>>>(c);
}

// Keys and permutation only (the test hook for bit-exact sort parity).
void launch_sort_only(BinArgs b, hipStream_t s, uint32_t** keys_out, uint32_t** vals_out) {
  const int nb = (b.n + 255) / 256;
  k_keys<<<nb, 256, 0, s>>>(b);
  const int ntiles = (b.n + RS_TILE - 1) / RS_TILE;
  const int passes = sort_passes(b.nnodes);
  uint32_t *k0 = b.keys, *v0 = b.vals, *k1 = b.keys2, *v1 = b.vals2;
  for (int pass = 0; pass < passes; ++pass) {
    k_rs_hist<<<ntiles, RS_T, 0, s>>>(k0, b.n, 8 * pass, b.hist, ntiles);
    k_rs_scan_digit<<<256, 1024, 0, s>>>(b.hist, ntiles, b.hist + 256 * (int64_t)ntiles);
    k_rs_scatter<<<ntiles, RS_T, 0, s>>>(k0, v0, k1, v1, b.n, 8 * pass, b.hist, ntiles,
                                         b.hist + 256 * (int64_t)ntiles);
    std::swap(k0, k1);
    std::swap(v0, v1);
  }
  *keys_out = k0;
  *vals_out = v0;
}

}  // namespace mpm
