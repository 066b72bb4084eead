// Fused binned substep: launch interface.
#pragma once
#include "mpm_internal.hpp"

namespace mpm {

// The substep epoch lives on the device (*dstep, advanced at the end of each
// substep's block pass; the grid pass that follows works on epoch *dstep - 1),
// so one captured substep can be replayed as a HIP graph: grid-block lists
// alternate with the epoch's parity, their counters rotate with it modulo 3.
// Reverse-mode record of one substep's grid: the P2G sums of its active
// grid blocks, 64 nodes each, in the order of list (list[0]: the count,
// list[1 + s]: the grid block of record slot s).  While the substep is
// reversed, map (one entry per grid block, -1 for blocks not listed) holds
// each listed block's slot.
struct GridRec {
  float4* acc;
  int* list;
  int* map;
};

struct BlockArgs {
  KArgs a;
  const int4* blist;  // (home block, start, end, 0)
  const int* nblk;    // number of home blocks
  int* gflag;         // per grid block: epoch of its last activation
  int* glist[2];      // grid blocks activated, by epoch parity
  int* gcnt;          // [3] counters, by epoch modulo 3
  int* dstep;         // substep epoch (> 0)
  int advance;        // block pass: advance the epoch once every workgroup is done
  int64_t nblocks;    // grid blocks over all scenes
  int grid;           // workgroups of the persistent block pass
  int* bctr;          // home-block claim counter (0 between launches)
  // reverse-mode recording (TEXTBOOK passes only): P2G and fused passes
  // write each particle's substep-start state (33 rows: x, v, B, F_E, F_P)
  // to rec, the G2P pass its new positions to rec's x rows; with grec.acc
  // set the grid pass records the substep's P2G sums there
  float* rec;
  // with rec: the SVD the F update used (SVD_ROWS rows of f_update_t's svd
  // output, slot order), for the reverse mode to read back
  float* rec_svd;
  GridRec grec;
};

enum { PASS_P2G = 0, PASS_FUSED = 1, PASS_G2P = 2, PASS_MASS = 3 };
constexpr int SVD_ROWS = 21;

// Reverse mode's advection + G2P adjoint on the block structure: per block,
// node velocities (vnode, filled from the substep's grid record by the map
// kernel; ba.grec's map set) in an
// LDS halo and the node-velocity adjoints (w (gv' + gB' d)) accumulated in an
// FP64 LDS halo, flushed with one set of global atomics per block into the
// listed grid blocks.  St: stored state of the substep (slot order); Xn: rows
// holding the next state's positions; Gin: the adjoints of the next state;
// Gout.x receives the partial position adjoint; gnode.yzw the node adjoints.
void launch_adj_g2p_block(const BlockArgs& ba, const float* St, const float* Xn, const float* Gin, float* Gout,
                          float4* gnode, const float4* vnode, hipStream_t s);
void launch_block_pass(const BlockArgs& ba, int mode, hipStream_t s);
void launch_grid_active(const BlockArgs& ba, hipStream_t s);

}  // namespace mpm
