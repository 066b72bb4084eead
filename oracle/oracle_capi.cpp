// ORACLE — test infrastructure only (see mpm_oracle.hpp).  A C interface over
// the CPU restatement for the Python tests (ctypes) and for the CPU baseline
// timed by bench.py.  Never linked into the product.
#include <type_traits>
#include <chrono>
#include <cstdio>
#include <memory>

#include "mpm_oracle.hpp"

using namespace mpmo;

namespace {

struct Handle {
  bool dbl = false;
  Sim<float> f;
  Sim<double> d;
};

template <typename T> void load_m3(M3<T>& m, const float* src) {
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) m.c[c][r] = (T)src[3 * c + r];
}
template <typename T> void store_m3(const M3<T>& m, double* dst) {
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) dst[3 * c + r] = (double)m.c[c][r];
}

template <typename T>
void set_particles(Sim<T>& s, const float* x, const float* v, const float* m, const float* E0, const float* B,
                   const float* FE, const float* FP, const float* V0) {
  for (int p = 0; p < s.n; ++p) {
    s.x[p] = {(T)x[3 * p], (T)x[3 * p + 1], (T)x[3 * p + 2]};
    s.v[p] = {(T)v[3 * p], (T)v[3 * p + 1], (T)v[3 * p + 2]};
    s.m[p] = (T)m[p];
    if (E0) s.E0[p] = (T)E0[p];
    if (B) load_m3(s.B[p], B + 9 * p);
    if (FE) load_m3(s.FE[p], FE + 9 * p);
    if (FP) load_m3(s.FP[p], FP + 9 * p);
    if (V0) s.V0[p] = (T)V0[p];
  }
}

template <typename T>
void get_particles(const Sim<T>& s, double* x, double* v, double* B, double* FE, double* FP, double* V0) {
  for (int p = 0; p < s.n; ++p) {
    if (x) {
      x[3 * p] = s.x[p].x;
      x[3 * p + 1] = s.x[p].y;
      x[3 * p + 2] = s.x[p].z;
    }
    if (v) {
      v[3 * p] = s.v[p].x;
      v[3 * p + 1] = s.v[p].y;
      v[3 * p + 2] = s.v[p].z;
    }
    if (B) store_m3(s.B[p], B + 9 * p);
    if (FE) store_m3(s.FE[p], FE + 9 * p);
    if (FP) store_m3(s.FP[p], FP + 9 * p);
    if (V0) V0[p] = s.V0[p];
  }
}

template <typename T> void get_grid(const Sim<T>& s, double* m, double* vel, double* force, unsigned char* active) {
  const size_t cells = (size_t)s.I * s.J * s.K;
  for (size_t c = 0; c < cells; ++c) {
    if (m) m[c] = s.gm[c];
    if (vel) {
      vel[3 * c] = s.gvel[c].x;
      vel[3 * c + 1] = s.gvel[c].y;
      vel[3 * c + 2] = s.gvel[c].z;
    }
    if (force) {
      force[3 * c] = s.gforce[c].x;
      force[3 * c + 1] = s.gforce[c].y;
      force[3 * c + 2] = s.gforce[c].z;
    }
    if (active) active[c] = 0;
  }
  if (active)
    for (int64_t c : s.used) active[c] = 1;
}

template <typename T> void set_colliders(Sim<T>& s, const float* c13, int k) {
  s.colliders.clear();
  for (int i = 0; i < k; ++i) {
    Collider<T> c;
    const float* d = c13 + 13 * i;
    for (int a = 0; a < 3; ++a) c.t[a] = (T)d[a];
    for (int a = 0; a < 4; ++a) c.q[a] = (T)d[3 + a];
    for (int a = 0; a < 3; ++a) c.s[a] = (T)d[7 + a];
    for (int a = 0; a < 3; ++a) c.vel[a] = (T)d[10 + a];
    c.prepare();
    s.colliders.push_back(c);
  }
}

template <typename T> int stage(Sim<T>& s, int st, double dt) {
  const T t = (T)(float)dt;
  switch (st) {
    case 0: return s.p2g();
    case 1: s.init_volumes(); return 0;
    case 2: s.forces(); return 0;
    case 3: s.grid_update(t); return 0;
    case 4: s.collisions(t); return 0;
    case 5: return s.deformation_update(t);
    case 6: s.g2p(); return 0;
    case 7: s.advect(t); return 0;
    default: return -1;
  }
}

}  // namespace

extern "C" {

void* orc_create(const Params* p, int n, int use_double) {
  auto* h = new Handle();
  h->dbl = use_double != 0;
  if (h->dbl)
    h->d.init(*p, n);
  else
    h->f.init(*p, n);
  return h;
}
void orc_destroy(void* hp) { delete static_cast<Handle*>(hp); }

int orc_set_particles(void* hp, const float* x, const float* v, const float* m, const float* E0, const float* B,
                      const float* FE, const float* FP, const float* V0) {
  auto* h = static_cast<Handle*>(hp);
  if (h->dbl)
    set_particles(h->d, x, v, m, E0, B, FE, FP, V0);
  else
    set_particles(h->f, x, v, m, E0, B, FE, FP, V0);
  return 0;
}
int orc_set_colliders(void* hp, const float* c13, int k) {
  auto* h = static_cast<Handle*>(hp);
  if (h->dbl)
    set_colliders(h->d, c13, k);
  else
    set_colliders(h->f, c13, k);
  return 0;
}
int orc_set_threads(void* hp, int threads) {
  auto* h = static_cast<Handle*>(hp);
  h->f.prm.threads = threads;
  h->d.prm.threads = threads;
  return 0;
}
// stage: 0 p2g, 1 init volumes, 2 forces, 3 grid update, 4 collisions,
//        5 F update, 6 g2p, 7 advect
int orc_stage(void* hp, int st, double dt) {
  auto* h = static_cast<Handle*>(hp);
  return h->dbl ? stage(h->d, st, dt) : stage(h->f, st, dt);
}
int orc_substeps(void* hp, double dt, int nsteps) {
  auto* h = static_cast<Handle*>(hp);
  for (int i = 0; i < nsteps; ++i) {
    const int rc = h->dbl ? h->d.substep((double)(float)dt) : h->f.substep((float)dt);
    if (rc) return rc;
  }
  return 0;
}
int orc_get_particles(void* hp, double* x, double* v, double* B, double* FE, double* FP, double* V0) {
  auto* h = static_cast<Handle*>(hp);
  if (h->dbl)
    get_particles(h->d, x, v, B, FE, FP, V0);
  else
    get_particles(h->f, x, v, B, FE, FP, V0);
  return 0;
}
int orc_get_grid(void* hp, double* m, double* vel, double* force, unsigned char* active) {
  auto* h = static_cast<Handle*>(hp);
  if (h->dbl)
    get_grid(h->d, m, vel, force, active);
  else
    get_grid(h->f, m, vel, force, active);
  return 0;
}
// Dense grid masses and velocities from the caller (index i*J*K + j*K + k);
// the used set becomes the nodes of non-zero mass, as after p2g.  Lets a
// known field (e.g. an affine velocity field) drive g2p directly.
int orc_set_grid(void* hp, const double* m, const double* vel) {
  auto* h = static_cast<Handle*>(hp);
  auto put = [&](auto& s) {
    using T = std::decay_t<decltype(s.gm[0])>;
    s.used.clear();
    for (size_t c = 0; c < s.gm.size(); ++c) {
      s.gm[c] = (T)m[c];
      s.gvel[c] = {(T)vel[3 * c], (T)vel[3 * c + 1], (T)vel[3 * c + 2]};
      s.gforce[c] = {T(0), T(0), T(0)};
      if (s.gm[c] != T(0)) s.used.push_back((int64_t)c);
    }
  };
  if (h->dbl)
    put(h->d);
  else
    put(h->f);
  return 0;
}
int orc_num_active(void* hp) {
  auto* h = static_cast<Handle*>(hp);
  return (int)(h->dbl ? h->d.used.size() : h->f.used.size());
}

// ----------------------------------------------------------- single-op KATs
float orc_weight_f(float x) { return weightN(x); }
double orc_weight_d(double x) { return weightN(x); }

// M, U, V are math row-major 3x3 (Eigen coefficient (r,c) at 3r+c)
int orc_svd3_f(const float* M, float* U, float* S, float* V) {
  float m[3][3], u[3][3], v[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) m[r][c] = M[3 * r + c];
  const int rc = jacobi_svd3(m, u, S, v);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      U[3 * r + c] = u[r][c];
      V[3 * r + c] = v[r][c];
    }
  return rc;
}
int orc_svd3_d(const double* M, double* U, double* S, double* V) {
  double m[3][3], u[3][3], v[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) m[r][c] = M[3 * r + c];
  const int rc = jacobi_svd3(m, u, S, v);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      U[3 * r + c] = u[r][c];
      V[3 * r + c] = v[r][c];
    }
  return rc;
}

// matrices in glm column-major (9 floats, element (row r, col c) at 3c + r)
void orc_polar_f(const float* F, float* R) {
  M3<float> f;
  for (int i = 0; i < 9; ++i) f.c[i / 3][i % 3] = F[i];
  const M3<float> r = polar_R(f);
  for (int i = 0; i < 9; ++i) R[i] = r.c[i / 3][i % 3];
}
void orc_mat3_ops_f(const float* A, const float* Bm, float* prod, float* inv, float* det) {
  M3<float> a, b;
  for (int i = 0; i < 9; ++i) {
    a.c[i / 3][i % 3] = A[i];
    b.c[i / 3][i % 3] = Bm[i];
  }
  const M3<float> p = mmul(a, b), q = minverse(a);
  for (int i = 0; i < 9; ++i) {
    prod[i] = p.c[i / 3][i % 3];
    inv[i] = q.c[i / 3][i % 3];
  }
  *det = mdet(a);
}
int orc_fupdate_f(int compose, float dt, float h, double theta_c, double theta_s, const float* B, const float* FE,
                  const float* FP, float* FEo, float* FPo, float* sig) {
  M3<float> b, fe, fp;
  for (int i = 0; i < 9; ++i) {
    b.c[i / 3][i % 3] = B[i];
    fe.c[i / 3][i % 3] = FE[i];
    fp.c[i / 3][i % 3] = FP[i];
  }
  const M3<float> Dinv = minverse(mscale(mscale(mscale(mdiag(1.0f), 1.0f / 3.0f), h), h));
  const int rc = f_update(compose, dt, Dinv, (float)(1.0 - theta_c), (float)(1.0 + theta_s), b, fe, fp, sig);
  for (int i = 0; i < 9; ++i) {
    FEo[i] = fe.c[i / 3][i % 3];
    FPo[i] = fp.c[i / 3][i % 3];
  }
  return rc;
}
// collider: 13 floats (t[3], q[4] as w,x,y,z, half extent s[3], vel[3])
float orc_sdf_f(const float* c13, const float* pos) {
  Collider<float> c;
  for (int a = 0; a < 3; ++a) c.t[a] = c13[a];
  for (int a = 0; a < 4; ++a) c.q[a] = c13[3 + a];
  for (int a = 0; a < 3; ++a) c.s[a] = c13[7 + a];
  for (int a = 0; a < 3; ++a) c.vel[a] = c13[10 + a];
  c.prepare();
  return c.sdf({pos[0], pos[1], pos[2]});
}
void orc_collider_inverse_f(const float* c13, float* inv16) {
  Collider<float> c;
  for (int a = 0; a < 3; ++a) c.t[a] = c13[a];
  for (int a = 0; a < 4; ++a) c.q[a] = c13[3 + a];
  for (int a = 0; a < 3; ++a) c.s[a] = c13[7 + a];
  c.prepare();
  for (int i = 0; i < 16; ++i) inv16[i] = c.inv.c[i / 4][i % 4];
}
void orc_collide_f(const float* c13, int k, const float* pos, const float* vel, float* out) {
  std::vector<Collider<float>> cs;
  for (int i = 0; i < k; ++i) {
    Collider<float> c;
    const float* d = c13 + 13 * i;
    for (int a = 0; a < 3; ++a) c.t[a] = d[a];
    for (int a = 0; a < 4; ++a) c.q[a] = d[3 + a];
    for (int a = 0; a < 3; ++a) c.s[a] = d[7 + a];
    for (int a = 0; a < 3; ++a) c.vel[a] = d[10 + a];
    c.prepare();
    cs.push_back(c);
  }
  const V3<float> r = body_collision<float>({pos[0], pos[1], pos[2]}, {vel[0], vel[1], vel[2]}, cs);
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}
void orc_domain_walls_f(int i, int j, int k, int I, int J, int K, const float* vel, float* out) {
  const V3<float> r = domain_walls<float>(i, j, k, I, J, K, {vel[0], vel[1], vel[2]});
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}

int orc_openmp_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
