// Gaussian control forces on the MAC faces and their parameter adjoint
// (gfx950): apply_gausion_force_kernel (2D/vortex/lfm_midpoint_simulator.py:309-326,
// 3D/lfm_midpoint_simulator.py:316-344) and calculate_adj_gausion_force_kernel
// (2D/shape/lfm_midpoint_diff_simulator.py:75-103,
// 3D/lfm_midpoint_diff_simulator.py:70-110).
//
// Force I has centre c_I, inverse squared radius r_I and one strength per
// axis s_I.  At the node p of a face of axis a, with m = |p - c_I| (the
// correctly rounded square root of the FP32 sum of squares), the force
// contributes when (m*m) r_I < 5, with e = exp(-(m*m) r_I):
//     f_a(p) = sum over I, in increasing I, of e * s_I[a]
// and the adjoint of a face field lam (one component per axis) is, per force,
//     adj_strength[a] = sum over faces of axis a of e * lam
//     adj_center      = sum over all faces of s_I[a] e 2 (p - c_I) r_I lam
//     pen_strength[a] = sum over faces of axis a of 2 (e s_I[a]) e
//     pen_center      = sum over all faces of 2 (e s_I[a]) s_I[a] e 2 (p - c_I) r_I.
// The membership test is evaluated in FP32 exactly as written (no
// multiply-add contraction), so a face is inside for exactly the forces the
// reference's test admits.  The forward sums e = exp(-q) of the FP32 q, as the
// reference does; the adjoint forms e and p - c in FP64 (its sums cancel
// strongly around the centre, where the FP32 e's 1e-6 relative error would
// dominate them).
//
// Forward: the reference visits every force at every face.  Here each
// workgroup owns a tile of faces, streams the forces through in chunks of 256,
// keeps (by an ordered prefix-sum compaction in LDS) those whose cut-off ball
// reaches the tile's bounding box, and each face thread sums the kept forces
// in increasing I.  The sum is that of the reference's loop, without the
// forces that cannot reach the tile.
//
// Adjoint: the reference scatters with atomics per (face, force).  Here one
// workgroup per force walks the faces inside the force's cut-off box, sums in
// FP64 per thread and reduces across the workgroup in a fixed order: no
// atomics, deterministic, FP32 results.
#include <math.h>

#include "dfm_common.hpp"
#include "dfm_internal.hpp"
#include "../../include/dfm.h"

using namespace dfm;

namespace {

constexpr int kFB = 256;          // threads per workgroup (four waves)
constexpr double kCullCut = 5.0005;  // conservative cut-off for tile/box culling (exact test is 5, FP32)

// face tile of one workgroup (index directions i, j, k; k fastest)
template <int DIM> struct Tile;
template <> struct Tile<2> { static constexpr int I = 16, J = 16, K = 1; };
template <> struct Tile<3> { static constexpr int I = 4, J = 8, K = 8; };

struct Forces {
  const float* c;  // (count, dim) centres
  const float* r;  // (count) inverse squared radii
  const float* s;  // (count, dim) strengths
  int count;
};

struct TileCounts {
  int per_axis[3];  // tiles of face grid a
  int t[3][3];      // tiles along each index direction of face grid a
};

// FP32 membership test and weight of force (cb, r) at node p, as the
// reference evaluates them; returns false outside the cut-off.
template <int DIM>
DFM_HD bool gauss_weight(const float p[3], const float* cb, float r, float& e) {
#pragma clang fp contract(off)
  float n2 = 0.f;
  for (int b = 0; b < DIM; ++b) {
    const float d = p[b] - cb[b];
    n2 = n2 + d * d;
  }
  const float m = sqrtf(n2);
  const float q = (m * m) * r;
  if (!(q < 5.f)) return false;
  e = expf(-q);
  return true;
}

template <int DIM>
__global__ void __launch_bounds__(kFB) k_gauss_force(Geo g, Forces F, TileCounts tc, dfm_mac out) {
#pragma clang fp contract(off)
  __shared__ float lc[DIM][kFB];
  __shared__ float lr[kFB];
  __shared__ float ls[kFB];
  __shared__ int wave_tot[kFB / 64];
  // workgroup -> (axis, tile)
  int a = 0, tile = (int)blockIdx.x;
  while (a < DIM - 1 && tile >= tc.per_axis[a]) tile -= tc.per_axis[a++];
  int S[3];
  shape_of(g, a, S);
  constexpr int T[3] = {Tile<DIM>::I, Tile<DIM>::J, Tile<DIM>::K};
  const int t2 = tile % tc.t[a][2], t01 = tile / tc.t[a][2];
  const int t1 = t01 % tc.t[a][1], t0 = t01 / tc.t[a][1];
  const int org[3] = {t0 * T[0], t1 * T[1], t2 * T[2]};
  const int tid = threadIdx.x;
  const int li = tid / (T[1] * T[2]), lj = (tid / T[2]) % T[1], lk = tid % T[2];
  const int I[3] = {org[0] + li, org[1] + lj, org[2] + lk};
  const bool valid = I[0] < S[0] && I[1] < S[1] && I[2] < S[2];
  float p[3];
  node_pos(g, a, I[0], I[1], I[2], p);
  // bounding box of the tile's nodes (FP32 node positions, widened to FP64)
  double lo[3], hi[3];
  for (int b = 0; b < DIM; ++b) {
    const int e0 = org[b], e1 = min(org[b] + T[b], S[b]) - 1;
    const float off = b == a ? 0.f : 0.5f;
    lo[b] = (double)(((float)e0 + off) * g.dx);
    hi[b] = (double)(((float)e1 + off) * g.dx);
  }
  const int lane = tid & 63, wave = tid >> 6;
  float acc = 0.f;
  for (int base = 0; base < F.count; base += kFB) {
    const int q = base + tid;
    bool keep = false;
    float cq[DIM], rq = 0.f, sq = 0.f;
    if (q < F.count) {
      rq = F.r[q];
      sq = F.s[(size_t)q * DIM + a];
      double d2 = 0.0;
      for (int b = 0; b < DIM; ++b) {
        cq[b] = F.c[(size_t)q * DIM + b];
        const double c = (double)cq[b];
        const double d = c < lo[b] ? lo[b] - c : (c > hi[b] ? c - hi[b] : 0.0);
        d2 += d * d;
      }
      keep = !(d2 * (double)rq >= kCullCut);  // r <= 0 or NaN: kept
    }
    const unsigned long long bal = __ballot(keep);
    if (lane == 0) wave_tot[wave] = __popcll(bal);
    __syncthreads();
    int off = 0, tot = 0;
    for (int w = 0; w < kFB / 64; ++w) {
      off += w < wave ? wave_tot[w] : 0;
      tot += wave_tot[w];
    }
    if (keep) {
      const int pos = off + __popcll(bal & ((1ull << lane) - 1ull));
      for (int b = 0; b < DIM; ++b) lc[b][pos] = cq[b];
      lr[pos] = rq;
      ls[pos] = sq;
    }
    __syncthreads();
    if (valid) {
      for (int m = 0; m < tot; ++m) {
        float cb[DIM];
        for (int b = 0; b < DIM; ++b) cb[b] = lc[b][m];
        float e;
        if (gauss_weight<DIM>(p, cb, lr[m], e)) acc = acc + e * ls[m];
      }
    }
    __syncthreads();
  }
  if (valid) out.c[a][lin(S, I[0], I[1], I[2])] = acc;
}

// Face-index box of grid a that can lie within the cut-off of force (c, r);
// the whole grid when r is not a positive finite number.  Returns the number
// of nodes (0 when empty).
template <int DIM>
DFM_HD int64_t force_box(const Geo& g, int a, const int S[3], const float* c, float r, int lo[3], int ext[3]) {
  const bool bounded = r > 0.f && isfinite(r);
  const double R = bounded ? sqrt(kCullCut / (double)r) : 0.0;
  int64_t cnt = 1;
  for (int b = 0; b < 3; ++b) {
    int l = 0, h = S[b] - 1;
    if (b < DIM && !isfinite(c[b])) {
      h = -1;  // a NaN or infinite centre admits no face
    } else if (b < DIM && bounded) {
      const double off = b == a ? 0.0 : 0.5;
      const double fl = floor(((double)c[b] - R) / (double)g.dx - off) - 1.0;
      const double fh = ceil(((double)c[b] + R) / (double)g.dx - off) + 1.0;
      l = fl < 0.0 ? 0 : (fl > (double)(S[b] - 1) ? S[b] : (int)fl);
      h = fh > (double)(S[b] - 1) ? S[b] - 1 : (fh < 0.0 ? -1 : (int)fh);
    }
    lo[b] = l;
    ext[b] = h >= l ? h - l + 1 : 0;
    cnt *= ext[b];
  }
  return cnt;
}

DFM_HD double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// One workgroup per force: FP64 sums over the faces of every axis inside the
// force's box, then a fixed-order reduction over the workgroup.
template <int DIM>
__global__ void __launch_bounds__(kFB) k_gauss_force_adj(Geo g, Forces F, dfm_mac lam, float* __restrict__ adj_c,
                                                         float* __restrict__ adj_s, float* __restrict__ pen_c,
                                                         float* __restrict__ pen_s) {
#pragma clang fp contract(off)
  constexpr int NA = 4 * DIM;  // adj_s[DIM], adj_c[DIM], pen_s[DIM], pen_c[DIM]
  __shared__ double red[kFB / 64][NA];
  const int I = (int)blockIdx.x;
  float c[DIM], s[DIM];
  for (int b = 0; b < DIM; ++b) {
    c[b] = F.c[(size_t)I * DIM + b];
    s[b] = F.s[(size_t)I * DIM + b];
  }
  const float r = F.r[I];
  double acc[NA];
  for (int v = 0; v < NA; ++v) acc[v] = 0.0;
  for (int a = 0; a < DIM; ++a) {
    int S[3], lo[3], ext[3];
    shape_of(g, a, S);
    const int64_t cnt = force_box<DIM>(g, a, S, c, r, lo, ext);
    for (int64_t t = threadIdx.x; t < cnt; t += kFB) {
      const uint32_t u = (uint32_t)t;
      const uint32_t q = u / (uint32_t)ext[2];
      const int k = lo[2] + (int)(u - q * (uint32_t)ext[2]);
      const uint32_t w = q / (uint32_t)ext[1];
      const int j = lo[1] + (int)(q - w * (uint32_t)ext[1]);
      const int i = lo[0] + (int)w;
      float p[3];
      node_pos(g, a, i, j, k, p);
      float e32;
      if (!gauss_weight<DIM>(p, c, r, e32)) continue;
      double d[DIM], d2 = 0.0;
      for (int b = 0; b < DIM; ++b) {
        d[b] = (double)p[b] - (double)c[b];
        d2 += d[b] * d[b];
      }
      const double rd = (double)r;
      const double ed = exp(-d2 * rd);
      const double l = (double)lam.c[a][lin(S, i, j, k)];
      const double sa = (double)s[a];
      const double f = ed * sa;
      acc[a] += ed * l;
      acc[2 * DIM + a] += 2.0 * f * ed;
      for (int b = 0; b < DIM; ++b) {
        const double gb = ed * 2.0 * d[b] * rd;  // d e / d c_b
        acc[DIM + b] += sa * gb * l;
        acc[3 * DIM + b] += 2.0 * f * sa * gb;
      }
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int v = 0; v < NA; ++v) {
    const double t = wave_sum(acc[v]);
    if (lane == 0) red[wave][v] = t;
  }
  __syncthreads();
  if (threadIdx.x < NA) {
    const int v = threadIdx.x;
    double t = 0.0;
    for (int w = 0; w < kFB / 64; ++w) t += red[w][v];
    const int b = v % DIM;
    float* dst = v < DIM ? adj_s : (v < 2 * DIM ? adj_c : (v < 3 * DIM ? pen_s : pen_c));
    dst[(size_t)I * DIM + b] = (float)t;
  }
}

int check_force_args(const dfm_grid* g, const float* center, const float* inv_radius, const float* strength,
                     int count, Geo& G) {
  if (!g) return set_error(DFM_E_NULL, "gaussian force: NULL grid");
  if (g->dim != 2 && g->dim != 3) return set_error(DFM_E_ARG, "dim must be 2 or 3");
  for (int a = 0; a < 3; ++a)
    if (g->n[a] < 1) return set_error(DFM_E_ARG, "grid extents must be >= 1");
  if (g->dim == 2 && g->n[2] != 1) return set_error(DFM_E_ARG, "2D grid must have n[2] == 1");
  if ((int64_t)(g->n[0] + 1) * (g->n[1] + 1) * (g->n[2] + 1) * 3 >= (int64_t(1) << 31))
    return set_error(DFM_E_ARG, "grid too large");
  if (!(g->dx > 0.f)) return set_error(DFM_E_ARG, "dx must be > 0");
  if (count < 0) return set_error(DFM_E_ARG, "force count must be >= 0");
  if (count > 0 && (!center || !inv_radius || !strength))
    return set_error(DFM_E_NULL, "gaussian force: NULL force parameters");
  G.dim = g->dim;
  for (int a = 0; a < 3; ++a) G.n[a] = g->n[a];
  G.dx = g->dx;
  G.inv_dx = 1.0f / g->dx;
  return DFM_OK;
}

int check_mac(const dfm_mac* m, int dim, const char* what) {
  if (!m) return set_error(DFM_E_NULL, what);
  for (int a = 0; a < dim; ++a)
    if (!m->c[a]) return set_error(DFM_E_NULL, what);
  return DFM_OK;
}

template <int DIM>
TileCounts tiles_of(const Geo& G, int& total) {
  TileCounts tc{};
  total = 0;
  for (int a = 0; a < DIM; ++a) {
    int S[3];
    shape_of_host(G, a, S);
    int n = 1;
    const int T[3] = {Tile<DIM>::I, Tile<DIM>::J, Tile<DIM>::K};
    for (int b = 0; b < 3; ++b) {
      tc.t[a][b] = (S[b] + T[b] - 1) / T[b];
      n *= tc.t[a][b];
    }
    tc.per_axis[a] = n;
    total += n;
  }
  return tc;
}

}  // namespace

extern "C" int dfm_gaussian_force(const dfm_grid* g, const float* center, const float* inv_radius,
                                  const float* strength, int count, const dfm_mac* f, void* stream) {
  Geo G;
  int rc = check_force_args(g, center, inv_radius, strength, count, G);
  if (rc != DFM_OK) return rc;
  if ((rc = check_mac(f, g->dim, "dfm_gaussian_force: NULL force field")) != DFM_OK) return rc;
  const Forces F{center, inv_radius, strength, count};
  hipStream_t st = static_cast<hipStream_t>(stream);
  int blocks = 0;
  if (g->dim == 2) {
    const TileCounts tc = tiles_of<2>(G, blocks);
    hipLaunchKernelGGL(k_gauss_force<2>, dim3(blocks), dim3(kFB), 0, st, G, F, tc, *f);
  } else {
    const TileCounts tc = tiles_of<3>(G, blocks);
    hipLaunchKernelGGL(k_gauss_force<3>, dim3(blocks), dim3(kFB), 0, st, G, F, tc, *f);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? DFM_OK : set_error(DFM_E_HIP, "dfm_gaussian_force: launch failed", hipGetErrorString(e));
}

extern "C" int dfm_gaussian_force_adjoint(const dfm_grid* g, const float* center, const float* inv_radius,
                                          const float* strength, int count, const dfm_mac* adj_u,
                                          float* adj_center, float* adj_strength, float* penalty_center,
                                          float* penalty_strength, void* stream) {
  Geo G;
  int rc = check_force_args(g, center, inv_radius, strength, count, G);
  if (rc != DFM_OK) return rc;
  if ((rc = check_mac(adj_u, g->dim, "dfm_gaussian_force_adjoint: NULL adjoint velocity")) != DFM_OK) return rc;
  if (count == 0) return DFM_OK;
  if (!adj_center || !adj_strength || !penalty_center || !penalty_strength)
    return set_error(DFM_E_NULL, "dfm_gaussian_force_adjoint: NULL output");
  const Forces F{center, inv_radius, strength, count};
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (g->dim == 2)
    hipLaunchKernelGGL(k_gauss_force_adj<2>, dim3(count), dim3(kFB), 0, st, G, F, *adj_u, adj_center, adj_strength,
                       penalty_center, penalty_strength);
  else
    hipLaunchKernelGGL(k_gauss_force_adj<3>, dim3(count), dim3(kFB), 0, st, G, F, *adj_u, adj_center, adj_strength,
                       penalty_center, penalty_strength);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? DFM_OK
                         : set_error(DFM_E_HIP, "dfm_gaussian_force_adjoint: launch failed", hipGetErrorString(e));
}
