// Particle binning for the fused substep (gfx950).
//
// key(p) = block(p) * 64 + cell-in-block(p), with cell = int(x / h) per axis
// (correctly rounded FP32 quotient, as material_point_method.cpp:83) and
// 4x4x4-cell blocks numbered blk0 + (bi * nb1 + bj) * nb2 + bk, blk0 being the
// first block of the particle's scene: one sort bins every scene.  A stable LSD radix
// sort (8-bit digits) orders (key, index) pairs exactly like a CPU stable
// sort: ties keep index order.  The non-empty blocks are listed from
// per-workgroup counts and a scan (no atomics), in a spread claim order
// (claim_slot), and the particles are moved into key order, so each block's
// particles take one contiguous slot range, cell by cell.
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

// Block boundaries of the sorted keys: the end of every non-empty block's
// slot range (blk_end, per grid block) and the number of blocks that start
// in each workgroup (wgcnt).
__global__ void __launch_bounds__(256) k_block_detect(BinArgs b, uint32_t* wgcnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool start = false;
  if (i < b.n) {
    const uint32_t blk = b.keys[i] >> 6;
    start = i == 0 || (b.keys[i - 1] >> 6) != blk;
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

// List of the non-empty blocks, (block, first slot, end slot, scene) each:
// the block starting in workgroup w has rank wgoff[w] + its rank among that
// workgroup's starts in key order, and list slot claim_slot(rank).
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
    const int blk = (int)(b.keys[i] >> 6);
    b.blist[claim_slot(r, (uint32_t)*b.nblk)] = make_int4(blk, i, b.blk_end[blk], b.blk_scene[blk]);
  }
}

// Particle fields and caller index of every slot moved into key order: slot
// i takes the particle the sort put at i (vals[i], its slot in the current
// order).  Particles move little between binnings, so the gathered loads
// stay close to in order.  The loads of half the fields are all issued
// before their stores.
constexpr int RO_T = 256;
__global__ void __launch_bounds__(RO_T) k_reorder(BinArgs b) {
  const float* __restrict__ P = b.P;
  float* __restrict__ Pout = b.Pout;
  constexpr int HALF = NFIELDS / 2;
  for (int dst = blockIdx.x * RO_T + threadIdx.x; dst < b.n; dst += gridDim.x * RO_T) {
    const int src = (int)b.vals[dst];
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
  k_reorder<<<(b.n + RO_T - 1) / RO_T, RO_T, 0, s>>>(c);
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
