// Block-pass building blocks shared by the forward passes (mpm_fused.hip)
// and the reverse-mode block kernels (mpm_adjoint.hip): the LDS halo of a
// 4x4x4-cell home block and the claiming of home blocks.
#pragma once
#include "mpm_fused.hpp"

namespace mpm {
namespace blk {

// Halo of a home block: 9 nodes per axis (the block, the cubic stencil and
// one cell of drift), strides (SX, SY, 1) chosen so a 4x4x4 cube of lanes
// spreads over the LDS banks.
constexpr int HN = 9;
constexpr int SX = 112, SY = 12;
constexpr int PLANE = 1008;  // >= 8*SX + 8*SY + 8 + 1

// The FP64 accumulator halo of the forward passes is tiled instead: node
// (x, y, z) sits at x TX + ty(y) + tz(z) with ty(y) = (y >> 2) 48 + (y & 3) 4
// and tz(z) = (z >> 2) 16 + (z & 3), that is in 4x4 (y, z) tiles of 16
// doubles, 3 x 3 tiles per x plane.  A double's bank pair (its index mod 32)
// is then 16 (tile parity) + 4 (y & 3) + (z & 3): the 16 lanes of an FP64
// atomic's lane group, whose rotated traversals (lane_rotation) visit 16
// distinct (y & 3, z & 3) at every step, never share a bank pair.
constexpr int TX = 144;
constexpr int TPLANE = HN * TX;  // 1296 doubles per accumulated field
__device__ __forceinline__ int tile_y(int y) { return ((y >> 2) * 48) + ((y & 3) << 2); }
__device__ __forceinline__ int tile_z(int z) { return ((z >> 2) << 4) + (z & 3); }

// The next home block for this workgroup, claimed from the pass's counter
// (home blocks differ in cost, so claiming balances the workgroups where a
// fixed share would leave a tail); -1 when none is left.  Only workgroups
// blockIdx.x < nblocks take part (the others return at once).  Each ends with
// exactly one failed claim, and the one that makes the last of them (value
// nblocks + participants - 1) re-zeroes the counter for the next launch.
__device__ __forceinline__ int claim_block(int* ctr, int nblocks, int* slot) {
  if (threadIdx.x == 0) {
    const int participants = (int)gridDim.x < nblocks ? (int)gridDim.x : nblocks;
    const int s = atomicAdd(ctr, 1);
    if (s == nblocks + participants - 1) *ctr = 0;
    *slot = s;
  }
  __syncthreads();
  // the same value in every lane: say so, so that the block list entry and
  // the scene record derived from it are loaded once into scalar registers
  const int s = __builtin_amdgcn_readfirstlane(*slot);
  return s < nblocks ? s : -1;
}

}  // namespace blk
}  // namespace mpm
