// Binning interface (keys, stable radix sort, block ranges, reorder).
#pragma once
#include "mpm_internal.hpp"

namespace mpm {

struct BinArgs {
  const float* P;    // particle fields, current order
  float* Pout;       // particle fields, binned order
  const int* pid;    // caller index of each slot
  int* pid_out;
  int64_t stride;
  int n;
  int nscenes;
  const Scene* scenes;    // device scene table (grids, particle slot ranges)
  const int* blk_scene;   // grid block -> scene
  int64_t nnodes;         // nodes over all scenes
  uint32_t *keys, *vals, *keys2, *vals2;  // 4 x n
  uint32_t* hist;                         // 256 x ceil(n / 4096)
  int* blk_end;                           // per grid block: end of its slot range
  int4* blist;                            // (block, start, end, scene) per listed block
  int* nblk;                              // number of listed blocks
};

int sort_passes(int64_t nnodes);
void launch_bin(BinArgs b, hipStream_t s);
void launch_sort_only(BinArgs b, hipStream_t s, uint32_t** keys_out, uint32_t** vals_out);

}  // namespace mpm
