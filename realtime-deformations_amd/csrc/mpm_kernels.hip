// Legacy-stage and layout kernels of libmpm_hip (gfx950).
//
// The legacy stages keep the reference's seven separate passes
// (realtime-deformations/main.cpp:192-218), one thread per particle or per
// node, with global float atomics into the dense block-major grid.  They back
// the 1:1 stage calls of the C ABI and the facade; the throughput path is in
// mpm_fused.hip.
#include "mpm_internal.hpp"

namespace mpm {

namespace {

__device__ __forceinline__ void load_m3(const float* P, int64_t stride, int f, int p, M3& m) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) m.c[c][r] = P[(int64_t)(f + 3 * c + r) * stride + p];
}
__device__ __forceinline__ void store_m3(float* P, int64_t stride, int f, int p, const M3& m) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) P[(int64_t)(f + 3 * c + r) * stride + p] = m.c[c][r];
}
__device__ __forceinline__ F3 load_f3(const float* P, int64_t stride, int f, int p) {
  return f3(P[(int64_t)f * stride + p], P[(int64_t)(f + 1) * stride + p], P[(int64_t)(f + 2) * stride + p]);
}
__device__ __forceinline__ void store_f3(float* P, int64_t stride, int f, int p, F3 v) {
  P[(int64_t)f * stride + p] = v.x;
  P[(int64_t)(f + 1) * stride + p] = v.y;
  P[(int64_t)(f + 2) * stride + p] = v.z;
}

struct Stencil {
  Axis ax[3];
};
__device__ __forceinline__ Stencil stencil_of(const Scene& S, F3 x) {
  Stencil s;
  s.ax[0] = axis_weights(x.x, S.h, S.inv_h, 1, S.dims[0] - 3);
  s.ax[1] = axis_weights(x.y, S.h, S.inv_h, 1, S.dims[1] - 3);
  s.ax[2] = axis_weights(x.z, S.h, S.inv_h, 1, S.dims[2] - 3);
  return s;
}

constexpr int TPB = 256;
inline int nblk(int64_t n) { return (int)((n + TPB - 1) / TPB); }

__global__ void k_clear(float4* p, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// rasterizeParticlesToGrid scatter: m_i += m_p w_ip,
// p_i += w_ip m_p (v_p + B_p D^-1 (x_i - x_p))   (cpp:99-117)
__global__ void k_stage_p2g(KArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  const F3 x = load_f3(a.P, a.stride, FX, p);
  const F3 v = load_f3(a.P, a.stride, FV, p);
  M3 B;
  load_m3(a.P, a.stride, FB, p, B);
  const float m = a.P[(int64_t)FM * a.stride + p];
  const M3 BD = mmul(B, dinv_matrix(a.sc.dinv));
  const Stencil s = stencil_of(a.sc, x);
  for (int dx = 0; dx < 4; ++dx)
    for (int dy = 0; dy < 4; ++dy)
#pragma unroll
      for (int dz = 0; dz < 4; ++dz) {
        const int i = s.ax[0].base - 1 + dx, j = s.ax[1].base - 1 + dy, k = s.ax[2].base - 1 + dz;
        const float w = s.ax[0].w[dx] * s.ax[1].w[dy] * s.ax[2].w[dz];
        const F3 Xi = f3((float)i * a.sc.h, (float)j * a.sc.h, (float)k * a.sc.h);
        const F3 mom = (v + mvec(BD, Xi - x)) * (w * m);
        const int64_t gi = node_index(i, j, k, a.sc);
        if (a.exact) {  // reproducible mode: on the scene's quanta, FP64
          double* g = reinterpret_cast<double*>(a.gacc64 + gi);
          const double Mm = quantum_magic(a.sc.qexp), Mp = quantum_magic(a.sc.qexp + 2);
          atomicAdd(g + 0, on_quantum((double)(m * w), Mm));
          atomicAdd(g + 1, on_quantum((double)mom.x, Mp));
          atomicAdd(g + 2, on_quantum((double)mom.y, Mp));
          atomicAdd(g + 3, on_quantum((double)mom.z, Mp));
        } else {
          float* g = reinterpret_cast<float*>(a.gacc + gi);
          atomicAdd(g + 0, m * w);
          atomicAdd(g + 1, mom.x);
          atomicAdd(g + 2, mom.y);
          atomicAdd(g + 3, mom.z);
        }
      }
}

// coordinates of node g (scene-local index) of scene S
__device__ __forceinline__ bool node_coords(const Scene& S, int64_t g, int& i, int& j, int& k) {
  const int64_t blk = g >> 6;
  const int l = (int)(g & 63);
  const int bk = (int)(blk % S.nb[2]);
  const int64_t t = blk / S.nb[2];
  const int bj = (int)(t % S.nb[1]);
  const int bi = (int)(t / S.nb[1]);
  i = bi * 4 + (l >> 4);
  j = bj * 4 + ((l >> 2) & 3);
  k = bk * 4 + (l & 3);
  return i < S.dims[0] && j < S.dims[1] && k < S.dims[2];
}

// velocity = momentum / mass on nodes with mass (cpp:118-121)
__global__ void k_stage_normalise(KArgs a) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.nnodes) return;
  const float4 acc = load_acc(a, g);
  float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
  if (acc.x != 0.0f) out = make_float4(acc.y / acc.x, acc.z / acc.x, acc.w / acc.x, acc.x);
  a.gvel[g] = out;
}

// computeParticleVolumesAndDensities (cpp:131-142)
__global__ void k_stage_volumes(KArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  const F3 x = load_f3(a.P, a.stride, FX, p);
  const Stencil s = stencil_of(a.sc, x);
  float density = 0.0f;
  for (int dx = 0; dx < 4; ++dx)
    for (int dy = 0; dy < 4; ++dy)
#pragma unroll
      for (int dz = 0; dz < 4; ++dz) {
        const int i = s.ax[0].base - 1 + dx, j = s.ax[1].base - 1 + dy, k = s.ax[2].base - 1 + dz;
        const float w = s.ax[0].w[dx] * s.ax[1].w[dy] * s.ax[2].w[dz];
        density += a.gvel[node_index(i, j, k, a.sc)].w * w;
      }
  density /= (a.sc.h * a.sc.h * a.sc.h);
  const float m = a.P[(int64_t)FM * a.stride + p];
  a.P[(int64_t)FVOL * a.stride + p] = density != 0.0f ? (m / density) : 0.0f;
}

// computeExplicitGridForces (cpp:235-254)
__global__ void k_stage_forces(KArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  const F3 x = load_f3(a.P, a.stride, FX, p);
  M3 FE, FP;
  load_m3(a.P, a.stride, FFE, p, FE);
  load_m3(a.P, a.stride, FFP, p, FP);
  const float V0 = a.P[(int64_t)FVOL * a.stride + p];
  const float E = a.P[(int64_t)FE0 * a.stride + p];
  const M3 A = stress_matrix(E, a.sc.lame, V0, dinv_matrix(a.sc.dinv), FE, FP);
  const Stencil s = stencil_of(a.sc, x);
  for (int dx = 0; dx < 4; ++dx)
    for (int dy = 0; dy < 4; ++dy)
#pragma unroll
      for (int dz = 0; dz < 4; ++dz) {
        const int i = s.ax[0].base - 1 + dx, j = s.ax[1].base - 1 + dy, k = s.ax[2].base - 1 + dz;
        const float w = s.ax[0].w[dx] * s.ax[1].w[dy] * s.ax[2].w[dz];
        const F3 Xi = f3((float)i * a.sc.h, (float)j * a.sc.h, (float)k * a.sc.h);
        const F3 f = mvec(mscale(A, w), Xi - x);
        float* g = reinterpret_cast<float*>(a.gforce + node_index(i, j, k, a.sc));
        atomicAdd(g + 0, -f.x);
        atomicAdd(g + 1, -f.y);
        atomicAdd(g + 2, -f.z);
      }
}

// gridVelocitiesUpdate (cpp:256-262)
__global__ void k_stage_grid(KArgs a) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.nnodes) return;
  float4 v = a.gvel[g];
  if (v.w == 0.0f) return;
  const float4 f = a.gforce[g];
  const F3 acc = f3(f.x / v.w + a.sc.g[0], f.y / v.w + a.sc.g[1], f.z / v.w + a.sc.g[2]);
  v.x = v.x + a.dt * acc.x;
  v.y = v.y + a.dt * acc.y;
  v.z = v.z + a.dt * acc.z;
  a.gvel[g] = v;
}

__device__ __forceinline__ F3 apply_bc(const KArgs& a, int i, int j, int k, F3 v) {
  if ((a.sc.bc_mode & MPM_BC_COLLIDERS) && a.sc.ncoll > 0)
    v = body_collision(a.sc.coll, a.sc.ncoll, f3((float)i * a.sc.h, (float)j * a.sc.h, (float)k * a.sc.h), v);
  if (a.sc.bc_mode & MPM_BC_DOMAIN) v = domain_walls(i, j, k, a.sc.dims[0], a.sc.dims[1], a.sc.dims[2], v);
  return v;
}

// gridBasedCollisions (cpp:298-304)
__global__ void k_stage_collide(KArgs a) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.nnodes) return;
  float4 v = a.gvel[g];
  if (v.w == 0.0f) return;
  int i, j, k;
  if (!node_coords(a.sc, g, i, j, k)) return;
  const F3 r = apply_bc(a, i, j, k, f3(v.x, v.y, v.z));
  a.gvel[g] = make_float4(r.x, r.y, r.z, v.w);
}

// updateDeformationGradient (cpp:306-330)
__global__ void k_stage_fupdate(KArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  M3 B, FE, FP;
  load_m3(a.P, a.stride, FB, p, B);
  load_m3(a.P, a.stride, FFE, p, FE);
  load_m3(a.P, a.stride, FFP, p, FP);
  if (!f_update(a.compose, a.dt, dinv_matrix(a.sc.dinv), a.sc.lo, a.sc.hi, a.sc.gap64, B, FE, FP)) {
    atomicOr(a.status, ST_NONFINITE);
    return;
  }
  store_m3(a.P, a.stride, FFE, p, FE);
  store_m3(a.P, a.stride, FFP, p, FP);
}

// updateParticleVelocities (cpp:332-342)
__global__ void k_stage_g2p(KArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  const F3 x = load_f3(a.P, a.stride, FX, p);
  const Stencil s = stencil_of(a.sc, x);
  F3 v = f3(0.f, 0.f, 0.f);
  M3 B = mdiag(0.0f);
  for (int dx = 0; dx < 4; ++dx)
    for (int dy = 0; dy < 4; ++dy)
#pragma unroll
      for (int dz = 0; dz < 4; ++dz) {
        const int i = s.ax[0].base - 1 + dx, j = s.ax[1].base - 1 + dy, k = s.ax[2].base - 1 + dz;
        const float w = s.ax[0].w[dx] * s.ax[1].w[dy] * s.ax[2].w[dz];
        const float4 gv4 = a.gvel[node_index(i, j, k, a.sc)];
        const F3 gv = f3(gv4.x, gv4.y, gv4.z);
        v = v + gv * w;
        const F3 d = f3((float)i * a.sc.h, (float)j * a.sc.h, (float)k * a.sc.h) - x;
        const float dd[3] = {d.x, d.y, d.z};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          B.c[c][0] = B.c[c][0] + (gv.x * dd[c]) * w;
          B.c[c][1] = B.c[c][1] + (gv.y * dd[c]) * w;
          B.c[c][2] = B.c[c][2] + (gv.z * dd[c]) * w;
        }
      }
  store_f3(a.P, a.stride, FV, p, v);
  store_m3(a.P, a.stride, FB, p, B);
}

// updateParticlePositions + clampPosition (cpp:344-350,381-388)
__global__ void k_stage_advect(KArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  F3 x = load_f3(a.P, a.stride, FX, p);
  const F3 v = load_f3(a.P, a.stride, FV, p);
  x = x + v * a.dt;
  const float lo = 3.0f * a.sc.h;
  x.x = fminf(fmaxf(x.x, lo), (float)(a.sc.dims[0] - 3) * a.sc.h);
  x.y = fminf(fmaxf(x.y, lo), (float)(a.sc.dims[1] - 3) * a.sc.h);
  x.z = fminf(fmaxf(x.z, lo), (float)(a.sc.dims[2] - 3) * a.sc.h);
  store_f3(a.P, a.stride, FX, p, x);
}

__global__ void k_scatter_field(float* dst, int64_t stride, const float* src, const int* pid, int p0, int np,
                                int comps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int p = p0 + i;
  const int q = pid[p] - p0;
  for (int c = 0; c < comps; ++c) dst[(int64_t)c * stride + p] = src[(int64_t)q * comps + c];
}
__global__ void k_gather_field(float* dst, const float* src, int64_t stride, const int* pid, int p0, int np,
                               int comps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int p = p0 + i;
  const int q = pid[p] - p0;
  for (int c = 0; c < comps; ++c) dst[(int64_t)q * comps + c] = src[(int64_t)c * stride + p];
}
__global__ void k_iota(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}
__global__ void k_fill_int(int* p, int64_t n, int v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_grid_to_dense(KArgs a, Scene S, float* m, float* vel, float* force, uint8_t* active) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= S.nblocks * 64) return;
  int i, j, k;
  if (!node_coords(S, g, i, j, k)) return;
  const int64_t d = ((int64_t)i * S.dims[1] + j) * S.dims[2] + k;
  const int64_t gg = S.blk0 * 64 + g;
  const float4 v = a.gvel[gg];
  if (m) m[d] = v.w;
  if (vel) {
    vel[3 * d] = v.x;
    vel[3 * d + 1] = v.y;
    vel[3 * d + 2] = v.z;
  }
  if (force) {
    const float4 f = a.gforce[gg];
    force[3 * d] = f.x;
    force[3 * d + 1] = f.y;
    force[3 * d + 2] = f.z;
  }
  if (active) active[d] = v.w != 0.0f ? 1 : 0;
}

// computeParticleVolumesAndDensities (cpp:131-142) for every scene: node
// masses (gvel.w) gathered with the particle's weights, V0 = m / density
__global__ void k_volumes(KArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  const Scene S = a.scenes[scene_of_slot(a.scenes, a.nscenes, p)];
  const F3 x = load_f3(a.P, a.stride, FX, p);
  const Stencil s = stencil_of(S, x);
  float density = 0.0f;
  for (int dx = 0; dx < 4; ++dx)
    for (int dy = 0; dy < 4; ++dy)
#pragma unroll
      for (int dz = 0; dz < 4; ++dz) {
        const int i = s.ax[0].base - 1 + dx, j = s.ax[1].base - 1 + dy, k = s.ax[2].base - 1 + dz;
        const float w = s.ax[0].w[dx] * s.ax[1].w[dy] * s.ax[2].w[dz];
        density += a.gvel[node_index(i, j, k, S)].w * w;
      }
  density /= (S.h * S.h * S.h);
  const float m = a.P[(int64_t)FM * a.stride + p];
  a.P[(int64_t)FVOL * a.stride + p] = density != 0.0f ? (m / density) : 0.0f;
}

__global__ void k_count_active(KArgs a, int64_t* out) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = g < a.nnodes && a.gvel[g].w != 0.0f;
  const unsigned long long ball = __ballot(act);
  if ((threadIdx.x & 63) == 0 && ball) atomicAdd(reinterpret_cast<unsigned long long*>(out), (unsigned long long)__popcll(ball));
}

}  // namespace

void launch_count_active(const KArgs& a, int64_t* out, hipStream_t s) {
  k_count_active<<<nblk(a.nnodes), TPB, 0, s>>>(a, out);
}
void launch_clear(float4* p, int64_t count, hipStream_t s) {
  if (count > 0) k_clear<<<nblk(count), TPB, 0, s>>>(p, count);
}
void launch_stage_p2g(const KArgs& a, hipStream_t s) {
  if (a.n) k_stage_p2g<<<nblk(a.n), TPB, 0, s>>>(a);
}
void launch_stage_normalise(const KArgs& a, hipStream_t s) { k_stage_normalise<<<nblk(a.nnodes), TPB, 0, s>>>(a); }
void launch_stage_volumes(const KArgs& a, hipStream_t s) {
  if (a.n) k_stage_volumes<<<nblk(a.n), TPB, 0, s>>>(a);
}
void launch_stage_forces(const KArgs& a, hipStream_t s) {
  if (a.n) k_stage_forces<<<nblk(a.n), TPB, 0, s>>>(a);
}
void launch_stage_grid(const KArgs& a, hipStream_t s) { k_stage_grid<<<nblk(a.nnodes), TPB, 0, s>>>(a); }
void launch_stage_collide(const KArgs& a, hipStream_t s) { k_stage_collide<<<nblk(a.nnodes), TPB, 0, s>>>(a); }
void launch_stage_fupdate(const KArgs& a, hipStream_t s) {
  if (a.n) k_stage_fupdate<<<nblk(a.n), TPB, 0, s>>>(a);
}
void launch_stage_g2p(const KArgs& a, hipStream_t s) {
  if (a.n) k_stage_g2p<<<nblk(a.n), TPB, 0, s>>>(a);
}
void launch_stage_advect(const KArgs& a, hipStream_t s) {
  if (a.n) k_stage_advect<<<nblk(a.n), TPB, 0, s>>>(a);
}
void launch_scatter_field(float* dst, int64_t stride, const float* src, const int* pid, int p0, int np, int comps,
                          hipStream_t s) {
  if (np) k_scatter_field<<<nblk(np), TPB, 0, s>>>(dst, stride, src, pid, p0, np, comps);
}
void launch_gather_field(float* dst, const float* src, int64_t stride, const int* pid, int p0, int np, int comps,
                         hipStream_t s) {
  if (np) k_gather_field<<<nblk(np), TPB, 0, s>>>(dst, src, stride, pid, p0, np, comps);
}
void launch_iota(int* p, int n, hipStream_t s) {
  if (n) k_iota<<<nblk(n), TPB, 0, s>>>(p, n);
}
void launch_fill_int(int* p, int64_t n, int v, hipStream_t s) {
  if (n) k_fill_int<<<nblk(n), TPB, 0, s>>>(p, n, v);
}
void launch_grid_to_dense(const KArgs& a, const Scene& S, float* m, float* vel, float* force, uint8_t* active,
                         hipStream_t s) {
  k_grid_to_dense<<<nblk(S.nblocks * 64), TPB, 0, s>>>(a, S, m, vel, force, active);
}
void launch_volumes(const KArgs& a, hipStream_t s) {
  if (a.n) k_volumes<<<nblk(a.n), TPB, 0, s>>>(a);
}

// out[scene] = bits of the largest |m| over the scene's particle slots
// (blockIdx.y: scene; out cleared by the caller).  The bits of non-negative
// floats order like the floats; a NaN mass orders above every finite one.
__global__ void k_scene_mass_max(const float* P, int64_t stride, const Scene* scenes, uint32_t* out) {
  const Scene& S = scenes[blockIdx.y];
  uint32_t best = 0;
  for (int p = blockIdx.x * TPB + threadIdx.x; p < S.np; p += gridDim.x * TPB)
    best = max(best, __float_as_uint(fabsf(P[(int64_t)FM * stride + S.p0 + p])));
  for (int off = 32; off > 0; off >>= 1) best = max(best, (uint32_t)__shfl_xor((int)best, off));
  if ((threadIdx.x & 63) == 0 && best) atomicMax(out + blockIdx.y, best);
}
void launch_scene_mass_max(const float* P, int64_t stride, const Scene* scenes, int nscenes, int max_np, uint32_t* out,
                           hipStream_t s) {
  launch_fill_int(reinterpret_cast<int*>(out), nscenes, 0, s);
  const int gx = std::max(1, std::min(nblk(max_np), 1024));
  k_scene_mass_max<<<dim3(gx, nscenes), TPB, 0, s>>>(P, stride, scenes, out);
}

}  // namespace mpm
