// ORACLE — test infrastructure only.
//
// Sequential CPU restatement of the seeded scene generators (BALL, CELLS,
// LATTICE), against which the library's device generators are checked bit
// for bit.  LATTICE follows initializeParticles
// (realtime-deformations/material_point_method.cpp:18-63): the candidate
// loop i, j, k over [c - w, c + w) with c = ivec3(origin / h) and
// w = int(width / h) (cpp:19-24), eight quarter offsets per cell (cpp:28-29),
// position (ijk + ds / 4 + jitter) * h (cpp:34), kept unless
// length(pos - origin) > width (cpp:35-36), written from the back of the
// array (cpp:41).  The reference's jitter comes from rand()
// (utils.h:110-127); here it is a point of the unit ball scaled by 1/4, drawn
// by rejection from Philox4x64-10 (Salmon et al., SC'11), whose output is
// pinned against numpy.random.Philox in the tests.
//
// Every float is an exact conversion or one rounded + - * / sqrt, compiled
// without contraction, so the device generators must match exactly.
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

struct W4 {
  uint64_t v[4];
};

uint64_t mulhi64(uint64_t a, uint64_t b) { return (uint64_t)(((unsigned __int128)a * b) >> 64); }

// Philox4x64-10: counter (c0, c1, c2, c3), key (k0, k1)
W4 philox4x64(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3, uint64_t k0, uint64_t k1) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
  const uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
  W4 x{{c0, c1, c2, c3}};
  for (int round = 0; round < 10; ++round) {
    if (round > 0) {
      k0 += W0;
      k1 += W1;
    }
    const uint64_t h0 = mulhi64(M0, x.v[0]), l0 = M0 * x.v[0];
    const uint64_t h1 = mulhi64(M1, x.v[2]), l1 = M1 * x.v[2];
    x = W4{{h1 ^ x.v[1] ^ k0, l1, h0 ^ x.v[3] ^ k1, l0}};
  }
  return x;
}

W4 draw(uint64_t idx, uint64_t attempt, uint64_t seed, uint64_t tag) { return philox4x64(idx, attempt, 0, 0, seed, tag); }

float unit24(uint64_t w) { return (float)(uint32_t)(w >> 40) * (1.0f / 16777216.0f); }

enum : uint64_t { T_BALL = 1, T_COUNT = 2, T_POS = 3, T_LATTICE = 4, T_E0 = 5 };

void ball_point(uint64_t idx, uint64_t seed, uint64_t tag, float out[3]) {
  out[0] = out[1] = out[2] = 0.0f;
  for (uint64_t t = 0; t < 64; ++t) {
    const W4 r = draw(idx, t, seed, tag);
    float c[3];
    for (int a = 0; a < 3; ++a) c[a] = 2.0f * unit24(r.v[a]) - 1.0f;
    const float s = (c[0] * c[0] + c[1] * c[1]) + c[2] * c[2];
    if (s <= 1.0f) {
      for (int a = 0; a < 3; ++a) out[a] = c[a];
      return;
    }
  }
}

uint64_t gcd(uint64_t a, uint64_t b) {
  while (b) {
    uint64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

}  // namespace

extern "C" {

void orc_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) {
  const W4 r = philox4x64(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}

// kind 0 BALL, 1 CELLS, 2 LATTICE.  Writes n particles (x: n x 3, E0: n);
// returns the number of candidates kept (LATTICE) or n, or -1 when the
// request cannot be met (CELLS count outside the box's range, LATTICE with
// fewer kept candidates than n: x is then untouched).
int64_t orc_generate(int kind, uint64_t seed, const float centre[3], float radius, const int cell_lo[3],
                     const int cells[3], int ppc_lo, int ppc_hi, float E_lo, float E_hi, float h, int n, float* x,
                     float* E0) {
  auto E_of = [&](float u) { return E_lo + (E_hi - E_lo) * u; };
  if (kind == 0) {
    for (int i = 0; i < n; ++i) {
      float c[3];
      ball_point((uint64_t)i, seed, T_BALL, c);
      for (int a = 0; a < 3; ++a) x[3 * i + a] = centre[a] + radius * c[a];
      E0[i] = E_of(unit24(draw((uint64_t)i, 0, seed, T_E0).v[0]));
    }
    return n;
  }
  if (kind == 1) {
    const int64_t nc = (int64_t)cells[0] * cells[1] * cells[2];
    if (nc <= 0 || n < nc * ppc_lo || n > nc * ppc_hi) return -1;
    std::vector<int> cnt((size_t)nc);
    int64_t total = 0;
    for (int64_t q = 0; q < nc; ++q) {
      const uint64_t w = draw((uint64_t)q, 0, seed, T_COUNT).v[0];
      cnt[q] = ppc_lo + (int)(((w >> 32) * (uint64_t)(ppc_hi - ppc_lo + 1)) >> 32);
      total += cnt[q];
    }
    // +-1 steps in the cell order k -> (k A + B) mod nc, A the first odd
    // number >= 2654435761 mod nc coprime with nc, B = seed mod nc; each
    // pass moves every eligible cell at most once
    uint64_t A = (2654435761ull % (uint64_t)nc) | 1ull;
    while (gcd(A, (uint64_t)nc) != 1) A += 2;
    const uint64_t B = seed % (uint64_t)nc;
    while (total != n) {
      const int64_t d = total - n;
      int64_t moved = 0;
      for (int64_t k = 0; k < nc && moved < (d > 0 ? d : -d); ++k) {
        const int64_t q = (int64_t)(((uint64_t)k * A + B) % (uint64_t)nc);
        if (d > 0 && cnt[q] > ppc_lo) {
          --cnt[q];
          ++moved;
        } else if (d < 0 && cnt[q] < ppc_hi) {
          ++cnt[q];
          ++moved;
        }
      }
      total += d > 0 ? -moved : moved;
    }
    int i = 0;
    for (int64_t q = 0; q < nc; ++q) {
      const int ck = (int)(q % cells[2]), cj = (int)((q / cells[2]) % cells[1]), ci = (int)(q / ((int64_t)cells[1] * cells[2]));
      const int cidx[3] = {cell_lo[0] + ci, cell_lo[1] + cj, cell_lo[2] + ck};
      for (int r = 0; r < cnt[q]; ++r, ++i) {
        const W4 w = draw((uint64_t)q, (uint64_t)r, seed, T_POS);
        for (int a = 0; a < 3; ++a) x[3 * i + a] = ((float)cidx[a] + unit24(w.v[a])) * h;
        E0[i] = E_of(unit24(w.v[3]));
      }
    }
    return n;
  }
  // LATTICE (cpp:18-63), origin = centre, width = radius
  const int w = (int)(radius / h);
  int c[3];
  for (int a = 0; a < 3; ++a) c[a] = (int)(centre[a] / h);
  static const float ds[8][3] = {{1, 1, 1}, {1, 1, 3}, {1, 3, 1}, {1, 3, 3}, {3, 1, 1}, {3, 1, 3}, {3, 3, 1}, {3, 3, 3}};
  std::vector<float> kept;
  std::vector<uint64_t> kept_q;
  uint64_t q = 0;
  for (int i = c[0] - w; i < c[0] + w; ++i)
    for (int j = c[1] - w; j < c[1] + w; ++j)
      for (int k = c[2] - w; k < c[2] + w; ++k)
        for (int d = 0; d < 8; ++d, ++q) {
          float u[3];
          ball_point(q, seed, T_LATTICE, u);
          const int ijk[3] = {i, j, k};
          float p[3];
          for (int a = 0; a < 3; ++a) p[a] = (((float)ijk[a] + ds[d][a] * 0.25f) + 0.25f * u[a]) * h;
          const float r0 = p[0] - centre[0], r1 = p[1] - centre[1], r2 = p[2] - centre[2];
          if (std::sqrt((r0 * r0 + r1 * r1) + r2 * r2) > radius) continue;
          kept.insert(kept.end(), p, p + 3);
          kept_q.push_back(q);
        }
  const int64_t nk = (int64_t)kept_q.size();
  if (nk < n) return -1 - nk;
  for (int64_t r = 0; r < n; ++r) {
    const int64_t slot = n - 1 - r;
    for (int a = 0; a < 3; ++a) x[3 * slot + a] = kept[3 * r + a];
    E0[slot] = E_of(unit24(draw(kept_q[r], 0, seed, T_E0).v[0]));
  }
  return nk;
}

}  // extern "C"
