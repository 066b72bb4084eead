// Reverse-mode (adjoint) kernels: launch interface.
#pragma once
#include "mpm_fused.hpp"

namespace mpm {

// Adjoint and stored-state rows: the first 33 particle fields (x, v, B, F_E,
// F_P), same layout and stride as the particle buffer.
enum AdjField : int { AX = FX, AV = FV, AB = FB, AFE = FFE, AFP = FFP, NA = FM };

void launch_adj_grid(const KArgs& a, const GridRec& r, float4* gnode, hipStream_t s);
// r.map from r.list (set = 1); or back to -1 with the listed blocks' node
// adjoints zeroed (set = 0)
void launch_adj_map(const KArgs& a, const GridRec& r, int set, float4* gnode, float4* vnode, hipStream_t s);
void launch_adj_p2g(const KArgs& a, const float* St, const float* Gin, float* Gout, const float4* gnode, float* gE,
                    const float* Sv, const float* SvPrev, hipStream_t s);
// row transfers: slot order (pid: caller index per slot) <-> caller order
void launch_rows_to_slots(float* dst, const float* src, const int* pid, int64_t stride, int n, int rows,
                          hipStream_t s);
void launch_rows_to_caller(float* dst, const float* src, const int* pid, int* inv, int64_t stride, int n, int rows,
                           hipStream_t s);  // inv: n ints of scratch
void launch_add_to_caller(float* dst, const float* src, const int* pid, int n, hipStream_t s);
// interleaved host layout (n x comps) <-> rows
void launch_aos_to_rows(float* dst, const float* src, int n, int comps, int64_t stride, hipStream_t s);
void launch_rows_to_aos(float* dst, const float* src, int n, int comps, int64_t stride, hipStream_t s);

}  // namespace mpm
