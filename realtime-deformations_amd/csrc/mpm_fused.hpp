// Fused binned substep: launch interface.
#pragma once
#include "mpm_internal.hpp"

namespace mpm {

struct BlockArgs {
  KArgs a;
  const int4* blist;  // (home block, start, end, 0)
  const int* nblk;    // number of home blocks
  int* gflag;         // per grid block: substep epoch of its last activation
  int* glist_cur;     // grid blocks activated this substep
  int* gcnt_cur;
  const int* glist_prev;  // grid blocks activated last substep
  const int* gcnt_prev;
  int* gcnt_next;     // reset for the next substep
  int step;           // substep epoch (> 0)
};

enum { PASS_P2G = 0, PASS_FUSED = 1, PASS_G2P = 2 };
void launch_block_pass(const BlockArgs& ba, int mode, hipStream_t s);
void launch_grid_active(const BlockArgs& ba, hipStream_t s);

}  // namespace mpm
