// MGPCG pressure projection for the all-Neumann box (MI355X / gfx950).
//
// Algorithm (restated from MGPCG, 2D/vortex/mgpcg.py:8-282 and
// 3D/mgpcg.py:13-291): PCG on A x = b, A = (#neighbours) I - adjacency on
// cells, preconditioned by one multigrid V-cycle with red-black Gauss-Seidel
// (colour = (i+j+k) & 1), residual restriction by summing the 2^d children
// and scaling by 1/(dim-1), piecewise-constant prolongation, `bottom` RBGS
// sweeps on the coarsest level; the residual is recentred to mean zero before
// every preconditioner application.
//
// MI355X mapping:
//  * PCG scalars (alpha, beta, r.r, z.r, mean) live on the device in FP64 and
//    are produced by deterministic two-stage reductions (per-block partials +
//    one finalize block), so no per-iteration host round trip is needed to
//    compute them; a device "done" flag turns every later kernel of the solve
//    into a no-op, so the host only polls the flag.
//  * Fusions: divergence + apply_bc + partial sums; the first pre-smoothing
//    sweep from z = 0 is one kernel (black cells read neighbour r/nn directly);
//    prolongation is fused into the first post-smoothing half-sweep (the
//    colour whose value is overwritten next is never materialised); the
//    recentring of r is applied on the fly by the level-0 V-cycle kernels and
//    written back by the z.r reduction; the coarsest level is solved by one
//    workgroup in LDS.
#include "dfm_common.hpp"
#include "dfm_internal.hpp"
#include "../../include/dfm.h"

#include <cmath>
#include <cstring>
#include <new>

using namespace dfm;

namespace {

constexpr int kB = 256;          // threads per block
constexpr int kRedBlocks = 1024; // blocks of the reduction kernels
constexpr int kMaxLevels = 12;
constexpr int kLdsCells = 8192;  // bottom solve in one workgroup up to this size
constexpr double kEps = 1e-12;   // denominator guard (2D/vortex/mgpcg.py:204)

enum Slot { S_ZR = 0, S_ALPHA, S_BETA, S_RR, S_BB, S_THRESH, S_MEAN, S_PAP, S_NSLOT };
enum Flag { F_DONE = 0, F_ITERS, F_STATUS, F_NFLAG };

struct Lvl {
  int n[3];
  int64_t size;
  float* r;
  float* z;
};

struct Dims {
  int n[3];
};

DFM_HD int64_t cidx(const Dims& d, int i, int j, int k) { return ((int64_t)i * d.n[1] + j) * d.n[2] + k; }

// neighbour sum and count (num_fluid_neighbors / neighbor_sum, 2D/vortex/mgpcg.py:92-115)
template <int DIM>
DFM_HD float nbr_sum(const float* __restrict__ f, const Dims& d, int i, int j, int k, float& nn) {
  float s = 0.f;
  nn = 0.f;
  if (i > 0) { s += f[cidx(d, i - 1, j, k)]; nn += 1.f; }
  if (i < d.n[0] - 1) { s += f[cidx(d, i + 1, j, k)]; nn += 1.f; }
  if (j > 0) { s += f[cidx(d, i, j - 1, k)]; nn += 1.f; }
  if (j < d.n[1] - 1) { s += f[cidx(d, i, j + 1, k)]; nn += 1.f; }
  if (DIM == 3) {
    if (k > 0) { s += f[cidx(d, i, j, k - 1)]; nn += 1.f; }
    if (k < d.n[2] - 1) { s += f[cidx(d, i, j, k + 1)]; nn += 1.f; }
  }
  return s;
}

template <int DIM>
DFM_HD float nbr_count(const Dims& d, int i, int j, int k) {
  float nn = 0.f;
  nn += (i > 0) + (i < d.n[0] - 1) + (j > 0) + (j < d.n[1] - 1);
  if (DIM == 3) nn += (k > 0) + (k < d.n[2] - 1);
  return nn;
}

// Cell index -> (i, j, k).  Grids here stay below 2^32 cells, so 32-bit
// unsigned division is used (64-bit division costs ~5x more on the VALU).
DFM_HD void cell_of(const Dims& d, int64_t id, int& i, int& j, int& k) {
  uint32_t u = (uint32_t)id;
  uint32_t n2 = (uint32_t)d.n[2], n1 = (uint32_t)d.n[1];
  uint32_t t = u / n2;
  k = (int)(u - t * n2);
  uint32_t q = t / n1;
  j = (int)(t - q * n1);
  i = (int)q;
}

// 1/nn for the neighbour counts that occur (1..6): one multiply instead of an
// IEEE divide per update.
DFM_HD float inv_count(float nn) {
  const int c = (int)nn;
  return c == 6 ? (1.0f / 6.0f)
       : c == 5 ? 0.2f
       : c == 4 ? 0.25f
       : c == 3 ? (1.0f / 3.0f)
       : c == 2 ? 0.5f
                : 1.0f;
}

DFM_HD bool is_done(const int* flags) { return flags[F_DONE] != 0; }

// ---------------------------------------------------------- reductions ----
DFM_HD double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Block sum of up to two values; thread 0 gets the totals.
DFM_HD void block_sum2(double& a, double& b) {
  __shared__ double sa[kB / 64], sb[kB / 64];
  a = wave_sum(a);
  b = wave_sum(b);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) { sa[w] = a; sb[w] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = 0; b = 0;
    for (int t = 0; t < kB / 64; ++t) { a += sa[t]; b += sb[t]; }
  }
}

// Finalize: sum partials and run the scalar step of the PCG recurrence.
enum FinOp { FIN_MEAN = 0, FIN_BNORM, FIN_ZR_INIT, FIN_ALPHA, FIN_RR, FIN_BETA, FIN_TRUE };

__global__ void __launch_bounds__(kB) k_finalize(int op, const double* __restrict__ part, int nparts, double* sc,
                                                 int* flags, int64_t ncells, double tol, int relative) {
  if (op != FIN_MEAN && op != FIN_BNORM && op != FIN_TRUE && is_done(flags)) return;
  double a = 0, b = 0;
  for (int t = threadIdx.x; t < nparts; t += kB) {
    a += part[t];
    b += part[kRedBlocks + t];
  }
  block_sum2(a, b);
  if (threadIdx.x != 0) return;
  switch (op) {
    case FIN_MEAN:  // a = sum r
      sc[S_MEAN] = a / (double)ncells;
      break;
    case FIN_BNORM:  // a = b.b (after recentring)
      sc[S_BB] = a;
      sc[S_RR] = a;
      sc[S_THRESH] = relative ? tol * tol * a : tol;
      flags[F_ITERS] = 0;
      flags[F_STATUS] = 0;
      flags[F_DONE] = (a == 0.0) ? 1 : 0;
      break;
    case FIN_ZR_INIT:  // a = z.r
      sc[S_ZR] = a;
      break;
    case FIN_ALPHA:  // a = p.Ap
      sc[S_PAP] = a;
      sc[S_ALPHA] = sc[S_ZR] / (a + kEps);
      break;
    case FIN_RR: {  // a = r.r, b = sum r
      sc[S_RR] = a;
      sc[S_MEAN] = b / (double)ncells;
      flags[F_ITERS] += 1;
      if (!isfinite(a)) {
        flags[F_STATUS] = 2;
        flags[F_DONE] = 1;
      } else if (a < sc[S_THRESH]) {
        flags[F_STATUS] = 1;
        flags[F_DONE] = 1;
      }
      break;
    }
    case FIN_BETA: {  // a = z.r (new)
      sc[S_BETA] = a / (sc[S_ZR] + kEps);
      sc[S_ZR] = a;
      break;
    }
    case FIN_TRUE: {  // a = |b - A x|^2 in FP64, b = sum(b - A x): refinement check
      sc[S_RR] = a;
      sc[S_MEAN] = b / (double)ncells;
      if (!isfinite(a)) {
        flags[F_STATUS] = 2;
        flags[F_DONE] = 1;
      } else if (a < sc[S_THRESH]) {
        flags[F_STATUS] = 1;
        flags[F_DONE] = 1;
      } else {
        flags[F_STATUS] = 0;
        flags[F_DONE] = 0;
      }
      break;
    }
  }
}

DFM_HD void store_partials(double* part, double a, double b) {
  block_sum2(a, b);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[kRedBlocks + blockIdx.x] = b;
  }
}

// ------------------------------------------------------ projection ends ----
struct Faces {
  float* c[3];
};

// apply_bc + divergence (no 1/dx) + b = -div + partial sum of b
// (apply_bc/divergence, 2D/vortex/mgpcg.py:294-336; 3D/mgpcg.py:380-411).
template <int DIM>
__global__ void __launch_bounds__(kB) k_div(Dims d, Faces u, float* __restrict__ b, double* part) {
  double acc = 0;
  int64_t N = (int64_t)d.n[0] * d.n[1] * d.n[2];
  for (int64_t id = (int64_t)blockIdx.x * kB + threadIdx.x; id < N; id += (int64_t)gridDim.x * kB) {
    int i, j, k;
    cell_of(d, id, i, j, k);
    // face grids: x (n0+1, n1, n2), y (n0, n1+1, n2), z (n0, n1, n2+1)
    Dims fx{{d.n[0] + 1, d.n[1], d.n[2]}}, fy{{d.n[0], d.n[1] + 1, d.n[2]}}, fz{{d.n[0], d.n[1], d.n[2] + 1}};
    float div = 0.f;
    float xl = (i == 0) ? 0.f : u.c[0][cidx(fx, i, j, k)];
    float xr = (i == d.n[0] - 1) ? 0.f : u.c[0][cidx(fx, i + 1, j, k)];
    float yl = (j == 0) ? 0.f : u.c[1][cidx(fy, i, j, k)];
    float yr = (j == d.n[1] - 1) ? 0.f : u.c[1][cidx(fy, i, j + 1, k)];
    div = (xr - xl) + (yr - yl);
    if (i == 0) u.c[0][cidx(fx, 0, j, k)] = 0.f;
    if (i == d.n[0] - 1) u.c[0][cidx(fx, d.n[0], j, k)] = 0.f;
    if (j == 0) u.c[1][cidx(fy, i, 0, k)] = 0.f;
    if (j == d.n[1] - 1) u.c[1][cidx(fy, i, d.n[1], k)] = 0.f;
    if (DIM == 3) {
      float zl = (k == 0) ? 0.f : u.c[2][cidx(fz, i, j, k)];
      float zr = (k == d.n[2] - 1) ? 0.f : u.c[2][cidx(fz, i, j, k + 1)];
      div += (zr - zl);
      if (k == 0) u.c[2][cidx(fz, i, j, 0)] = 0.f;
      if (k == d.n[2] - 1) u.c[2][cidx(fz, i, j, d.n[2])] = 0.f;
    }
    float bv = -div;
    b[id] = bv;
    acc += (double)bv;
  }
  store_partials(part, acc, 0.0);
}

// copy b into r (for dfm_mgpcg_solve) with partial sum
__global__ void __launch_bounds__(kB) k_load_b(int64_t N, const float* __restrict__ b, float* __restrict__ r,
                                               double* part) {
  double acc = 0;
  for (int64_t id = (int64_t)blockIdx.x * kB + threadIdx.x; id < N; id += (int64_t)gridDim.x * kB) {
    float v = b[id];
    r[id] = v;
    acc += (double)v;
  }
  store_partials(part, acc, 0.0);
}

// r -= mean (kept as b for the refinement residual); partial r.r; x = 0
__global__ void __launch_bounds__(kB) k_recenter_norm(int64_t N, float* __restrict__ r, float* __restrict__ bkeep,
                                                      double* __restrict__ x, const double* sc, double* part) {
  float m = (float)sc[S_MEAN];
  double acc = 0;
  for (int64_t id = (int64_t)blockIdx.x * kB + threadIdx.x; id < N; id += (int64_t)gridDim.x * kB) {
    float v = r[id] - m;
    r[id] = v;
    bkeep[id] = v;
    x[id] = 0.0;
    acc += (double)v * (double)v;
  }
  store_partials(part, acc, 0.0);
}

// subtract_grad_p + apply_bc (2D/vortex/mgpcg.py:338-356, 3D/mgpcg.py:413-439):
// interior faces u_a -= p[I] - p[I - e_a]; boundary faces = 0.
template <int DIM>
__global__ void __launch_bounds__(kB) k_subgrad(Dims d, Faces u, const double* __restrict__ p) {
  int64_t nf[3];
  Dims fd[3];
  for (int a = 0; a < 3; ++a) {
    fd[a] = d;
    if (a < DIM) fd[a].n[a] += 1;
    nf[a] = (a < DIM) ? (int64_t)fd[a].n[0] * fd[a].n[1] * fd[a].n[2] : 0;
  }
  int64_t total = nf[0] + nf[1] + nf[2];
  for (int64_t g = (int64_t)blockIdx.x * kB + threadIdx.x; g < total; g += (int64_t)gridDim.x * kB) {
    int a = 0;
    int64_t id = g;
    if (id >= nf[0]) { id -= nf[0]; a = 1; if (id >= nf[1]) { id -= nf[1]; a = 2; } }
    int i, j, k;
    cell_of(fd[a], id, i, j, k);
    int I[3] = {i, j, k};
    float* c = u.c[a];
    if (I[a] == 0 || I[a] == d.n[a]) {
      c[id] = 0.f;
    } else {
      int lo[3] = {i, j, k};
      lo[a] -= 1;
      // the pressure difference is formed in FP64: p grows like n |u| in
      // these units, so an FP32 p would cost ~n ulps of velocity
      c[id] = (float)((double)c[id] - (p[cidx(d, i, j, k)] - p[cidx(d, lo[0], lo[1], lo[2])]));
    }
  }
}

// Iterative-refinement residual r = b - A x evaluated in FP64 (x is FP64;
// b is the recentred FP32 right-hand side); partial |r|^2 and sum r.
template <int DIM>
__global__ void __launch_bounds__(kB) k_true_resid(Dims d, const double* __restrict__ x, const float* __restrict__ b,
                                                   float* __restrict__ r, double* part) {
  double rr = 0, rs = 0;
  int64_t N = (int64_t)d.n[0] * d.n[1] * d.n[2];
  for (int64_t id = (int64_t)blockIdx.x * kB + threadIdx.x; id < N; id += (int64_t)gridDim.x * kB) {
    int i, j, k;
    cell_of(d, id, i, j, k);
    double s = 0.0, nn = 0.0;
    if (i > 0) { s += x[cidx(d, i - 1, j, k)]; nn += 1.0; }
    if (i < d.n[0] - 1) { s += x[cidx(d, i + 1, j, k)]; nn += 1.0; }
    if (j > 0) { s += x[cidx(d, i, j - 1, k)]; nn += 1.0; }
    if (j < d.n[1] - 1) { s += x[cidx(d, i, j + 1, k)]; nn += 1.0; }
    if (DIM == 3) {
      if (k > 0) { s += x[cidx(d, i, j, k - 1)]; nn += 1.0; }
      if (k < d.n[2] - 1) { s += x[cidx(d, i, j, k + 1)]; nn += 1.0; }
    }
    double v = (double)b[id] - (nn * x[id] - s);
    r[id] = (float)v;
    rr += v * v;
    rs += v;
  }
  store_partials(part, rr, rs);
}

// FP64 solution -> FP32 output
__global__ void __launch_bounds__(kB) k_to_float(int64_t N, const double* __restrict__ x, float* __restrict__ y) {
  for (int64_t id = (int64_t)blockIdx.x * kB + threadIdx.x; id < N; id += (int64_t)gridDim.x * kB) y[id] = (float)x[id];
}

// ------------------------------------------------------------ PCG steps ----
// Ap = A p ; partial p.Ap
template <int DIM>
__global__ void __launch_bounds__(kB) k_Ap_dot(Dims d, const float* __restrict__ p, float* __restrict__ Ap,
                                               double* part, const int* flags) {
  if (is_done(flags)) return;
  double acc = 0;
  int64_t N = (int64_t)d.n[0] * d.n[1] * d.n[2];
  for (int64_t id = (int64_t)blockIdx.x * kB + threadIdx.x; id < N; id += (int64_t)gridDim.x * kB) {
    int i, j, k;
    cell_of(d, id, i, j, k);
    float nn;
    float s = nbr_sum<DIM>(p, d, i, j, k, nn);
    float pv = p[id];
    float v = nn * pv - s;
    Ap[id] = v;
    acc += (double)pv * (double)v;
  }
  store_partials(part, acc, 0.0);
}

// x += alpha p (x in FP64) ; r -= alpha Ap ; partial r.r and sum r
__global__ void __launch_bounds__(kB) k_update_xr(int64_t N, double* __restrict__ x, float* __restrict__ r,
                                                  const float* __restrict__ p, const float* __restrict__ Ap,
                                                  const double* sc, double* part, const int* flags) {
  if (is_done(flags)) return;
  const double alpha_d = sc[S_ALPHA];
  float alpha = (float)alpha_d;
  double rr = 0, rs = 0;
  for (int64_t id = (int64_t)blockIdx.x * kB + threadIdx.x; id < N; id += (int64_t)gridDim.x * kB) {
    x[id] = fma(alpha_d, (double)p[id], x[id]);
    float v = fmaf(-alpha, Ap[id], r[id]);
    r[id] = v;
    rr += (double)v * (double)v;
    rs += (double)v;
  }
  store_partials(part, rr, rs);
}

// r -= shift (written back) ; partial z.r   (shift = 0 for the first pass)
__global__ void __launch_bounds__(kB) k_zr(int64_t N, const float* __restrict__ z, float* __restrict__ r,
                                           const double* shift, double* part, const int* flags, int init_p,
                                           float* __restrict__ p) {
  if (is_done(flags)) return;
  float m = shift ? (float)*shift : 0.f;
  double acc = 0;
  for (int64_t id = (int64_t)blockIdx.x * kB + threadIdx.x; id < N; id += (int64_t)gridDim.x * kB) {
    float rv = r[id] - m;
    if (shift) r[id] = rv;
    float zv = z[id];
    if (init_p) p[id] = zv;
    acc += (double)zv * (double)rv;
  }
  store_partials(part, acc, 0.0);
}

// p = z + beta p
__global__ void __launch_bounds__(kB) k_update_p(int64_t N, const float* __restrict__ z, float* __restrict__ p,
                                                 const double* sc, const int* flags) {
  if (is_done(flags)) return;
  float beta = (float)sc[S_BETA];
  for (int64_t id = (int64_t)blockIdx.x * kB + threadIdx.x; id < N; id += (int64_t)gridDim.x * kB)
    p[id] = fmaf(beta, p[id], z[id]);
}

// -------------------------------------------------------------- V-cycle ----
// First RBGS sweep from z = 0: colour 0 cells z = r/nn; colour 1 cells use the
// colour-0 neighbours' r/nn.  `shift` (level 0) recentres r on the fly.
template <int DIM>
__global__ void __launch_bounds__(kB) k_first_sweep(Dims d, const float* __restrict__ r, float* __restrict__ z,
                                                    const double* shift, const int* flags) {
  if (is_done(flags)) return;
  float m = shift ? (float)*shift : 0.f;
  int64_t N = (int64_t)d.n[0] * d.n[1] * d.n[2];
  int64_t id = xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
  if (id >= N) return;
  int i, j, k;
  cell_of(d, id, i, j, k);
  float nn = nbr_count<DIM>(d, i, j, k);
  float rv = r[id] - m;
  if (((i + j + k) & 1) == 0) {
    z[id] = rv * inv_count(nn);
    return;
  }
  float s = 0.f;
  if (i > 0) s += (r[cidx(d, i - 1, j, k)] - m) * inv_count(nbr_count<DIM>(d, i - 1, j, k));
  if (i < d.n[0] - 1) s += (r[cidx(d, i + 1, j, k)] - m) * inv_count(nbr_count<DIM>(d, i + 1, j, k));
  if (j > 0) s += (r[cidx(d, i, j - 1, k)] - m) * inv_count(nbr_count<DIM>(d, i, j - 1, k));
  if (j < d.n[1] - 1) s += (r[cidx(d, i, j + 1, k)] - m) * inv_count(nbr_count<DIM>(d, i, j + 1, k));
  if (DIM == 3) {
    if (k > 0) s += (r[cidx(d, i, j, k - 1)] - m) * inv_count(nbr_count<DIM>(d, i, j, k - 1));
    if (k < d.n[2] - 1) s += (r[cidx(d, i, j, k + 1)] - m) * inv_count(nbr_count<DIM>(d, i, j, k + 1));
  }
  z[id] = (rv + s) * inv_count(nn);
}

// Colour-`phase` cell of pair index t (pairs run along the innermost axis);
// false when the pair's cell of that colour falls outside an odd-sized row.
template <int DIM>
DFM_HD bool colour_cell(const Dims& d, uint32_t t, int phase, int& i, int& j, int& k) {
  const uint32_t inner = DIM == 3 ? d.n[2] : d.n[1];
  const uint32_t half = (inner + 1) >> 1;
  uint32_t o = t / half;
  int hm = (int)(t - o * half);
  if (DIM == 3) {
    uint32_t q = o / (uint32_t)d.n[1];
    i = (int)q;
    j = (int)(o - q * (uint32_t)d.n[1]);
    k = 2 * hm + ((phase - i - j) & 1);
    return k < d.n[2];
  }
  i = (int)o;
  k = 0;
  j = 2 * hm + ((phase - i) & 1);
  return j < d.n[1];
}

// One half-sweep of colour `phase`; threads cover only that colour (the
// innermost axis is walked in steps of two).  If zc is given, the
// neighbours' values are read as z + zc[coarse parent] (fused prolongation).
template <int DIM>
__global__ void __launch_bounds__(kB) k_smooth(Dims d, const float* __restrict__ r, float* __restrict__ z, int phase,
                                               const double* shift, const float* __restrict__ zc, Dims dc,
                                               const int* flags) {
  if (is_done(flags)) return;
  float m = shift ? (float)*shift : 0.f;
  const int inner = DIM == 3 ? d.n[2] : d.n[1];
  const int half = (inner + 1) >> 1;
  const int64_t outer = DIM == 3 ? (int64_t)d.n[0] * d.n[1] : d.n[0];
  int64_t t = xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
  if (t >= outer * half) return;
  int i, j, k;
  if (!colour_cell<DIM>(d, (uint32_t)t, phase, i, j, k)) return;
  int64_t id = cidx(d, i, j, k);
  float nn, s;
  if (zc) {
    auto zf = [&](int a, int b, int c) {
      return z[cidx(d, a, b, c)] + zc[cidx(dc, a >> 1, b >> 1, DIM == 3 ? (c >> 1) : 0)];
    };
    s = 0.f;
    nn = 0.f;
    if (i > 0) { s += zf(i - 1, j, k); nn += 1.f; }
    if (i < d.n[0] - 1) { s += zf(i + 1, j, k); nn += 1.f; }
    if (j > 0) { s += zf(i, j - 1, k); nn += 1.f; }
    if (j < d.n[1] - 1) { s += zf(i, j + 1, k); nn += 1.f; }
    if (DIM == 3) {
      if (k > 0) { s += zf(i, j, k - 1); nn += 1.f; }
      if (k < d.n[2] - 1) { s += zf(i, j, k + 1); nn += 1.f; }
    }
  } else {
    s = nbr_sum<DIM>(z, d, i, j, k, nn);
  }
  z[id] = (r[id] - m + s) * inv_count(nn);
}

// restrict (2D/vortex/mgpcg.py:152-158): coarse r = sum over the 2^d children
// of the fine residual r - A z, times 1/(dim-1).  One thread per coarse cell,
// so the sum order is fixed.
template <int DIM>
__global__ void __launch_bounds__(kB) k_restrict(Dims d, const float* __restrict__ r, const float* __restrict__ z,
                                                 Dims dc, float* __restrict__ rc, const double* shift,
                                                 const int* flags) {
  if (is_done(flags)) return;
  float m = shift ? (float)*shift : 0.f;
  int64_t Nc = (int64_t)dc.n[0] * dc.n[1] * dc.n[2];
  int64_t cid = xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
  if (cid >= Nc) return;
  int ci, cj, ck;
  cell_of(dc, cid, ci, cj, ck);
  float acc = 0.f;
  const int kz = DIM == 3 ? 2 : 1;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < kz; ++c) {
        int i = 2 * ci + a, j = 2 * cj + b, k = DIM == 3 ? 2 * ck + c : 0;
        float nn;
        float s = nbr_sum<DIM>(z, d, i, j, k, nn);
        int64_t id = cidx(d, i, j, k);
        acc += (r[id] - m) - (nn * z[id] - s);
      }
  rc[cid] = acc * (1.0f / (float)(DIM - 1));
}

// --------------------------------------------------- coarse tail in LDS ----
// Levels Lt..L-1 of the V-cycle run inside one workgroup with every level's
// r and z resident in LDS (FP32, or BF16 storage with FP32 arithmetic in the
// mixed-precision mode): pre-smoothing from zero, restriction, the `bottom`
// sweeps on the coarsest level, prolongation and post-smoothing -- the launch
// sequence the reference issues per level collapsed into one kernel.
DFM_HD float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
DFM_HD uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (uint16_t)(u >> 16);
}
template <typename T> struct Lds;
template <> struct Lds<float> {
  DFM_HD static float ld(const float* p, int i) { return p[i]; }
  DFM_HD static void st(float* p, int i, float v) { p[i] = v; }
};
template <> struct Lds<uint16_t> {
  DFM_HD static float ld(const uint16_t* p, int i) { return bf16_to_f32(p[i]); }
  DFM_HD static void st(uint16_t* p, int i, float v) { p[i] = f32_to_bf16(v); }
};

constexpr int kTailMax = 8;
struct TailArgs {
  int nl;             // levels in the tail
  int n[kTailMax][3]; // their dims
  int off[kTailMax];  // element offset of level l's r in LDS (z follows r)
  int pre, bottom;
  int bottom_exact;   // 1: solve the coarsest level exactly (DCT-II) instead of RBGS
  float inv_dm1;      // restriction scale 1/(dim-1)
};

// Orthonormal DCT-II along one axis of an LDS array (or its inverse, DCT-III):
// out[.., k, ..] = sum_m in[.., m, ..] * c_k cos(pi k (2m+1) / 2N) (forward).
// The all-Neumann Laplacian factorises over axes and each factor is
// diagonalised by this basis with eigenvalues 2 - 2 cos(pi k / N).
template <typename T>
DFM_HD void lds_dct_axis(const T* in, T* out, const int* n, int axis, bool inverse) {
  const int N = n[axis];
  const int stride = axis == 0 ? n[1] * n[2] : (axis == 1 ? n[2] : 1);
  const int total = n[0] * n[1] * n[2];
  const float pi_2n = 3.14159265358979f / (2.0f * (float)N);
  const float c0 = sqrtf(1.0f / (float)N), c1 = sqrtf(2.0f / (float)N);
  for (int id = threadIdx.x; id < total; id += blockDim.x) {
    int coord = (id / stride) % N;
    int base = id - coord * stride;
    float acc = 0.f;
    for (int m = 0; m < N; ++m) {
      // forward: k = coord (output frequency), m = sample; inverse: swapped
      int kf = inverse ? m : coord, xs = inverse ? coord : m;
      float c = (kf == 0 ? c0 : c1) * __cosf(pi_2n * (float)(kf * (2 * xs + 1)));
      acc = fmaf(Lds<T>::ld(in, base + m * stride), c, acc);
    }
    Lds<T>::st(out, id, acc);
  }
  __syncthreads();
}

// z = A^+ r on one LDS level (mean-zero solution of the Neumann problem):
// forward DCT on every axis, divide by the eigenvalue sum (zero mode -> 0),
// inverse DCT.  r is used as scratch; the result lands in z.
template <int DIM, typename T>
DFM_HD void lds_exact_solve(T* r, T* z, const int* n) {
  T* a = r;
  T* b = z;
  for (int ax = 0; ax < DIM; ++ax) {
    lds_dct_axis<T>(a, b, n, ax, false);
    T* t = a; a = b; b = t;
  }
  const int total = n[0] * n[1] * n[2];
  for (int id = threadIdx.x; id < total; id += blockDim.x) {
    int k = id % n[2], t = id / n[2];
    int j = t % n[1], i = t / n[1];
    float lam = 2.0f - 2.0f * __cosf(3.14159265358979f * (float)i / (float)n[0]) +
                2.0f - 2.0f * __cosf(3.14159265358979f * (float)j / (float)n[1]);
    if (DIM == 3) lam += 2.0f - 2.0f * __cosf(3.14159265358979f * (float)k / (float)n[2]);
    float v = Lds<T>::ld(a, id);
    Lds<T>::st(a, id, (i == 0 && j == 0 && k == 0) ? 0.f : v / lam);
  }
  __syncthreads();
  for (int ax = 0; ax < DIM; ++ax) {
    lds_dct_axis<T>(a, b, n, ax, true);
    T* t = a; a = b; b = t;
  }
  if (a != z) {
    for (int id = threadIdx.x; id < total; id += blockDim.x) Lds<T>::st(z, id, Lds<T>::ld(a, id));
    __syncthreads();
  }
}

// One half-sweep of colour `colour` on an LDS-resident level: every thread
// owns one cell of that colour per pass (pairs along the innermost axis).
template <int DIM, typename T>
DFM_HD void lds_half_sweep(T* r, T* z, const int* n, int colour) {
  const Dims d{{n[0], n[1], n[2]}};
  const uint32_t inner = DIM == 3 ? n[2] : n[1];
  const uint32_t pairs = (uint32_t)(DIM == 3 ? n[0] * n[1] : n[0]) * ((inner + 1) >> 1);
  const int sx = n[1] * n[2], sy = n[2];
  for (uint32_t t = threadIdx.x; t < pairs; t += blockDim.x) {
    int i, j, k;
    if (!colour_cell<DIM>(d, t, colour, i, j, k)) continue;
    int id = (i * n[1] + j) * n[2] + k;
    float s = 0.f, nn = 0.f;
    if (i > 0) { s += Lds<T>::ld(z, id - sx); nn += 1.f; }
    if (i < n[0] - 1) { s += Lds<T>::ld(z, id + sx); nn += 1.f; }
    if (j > 0) { s += Lds<T>::ld(z, id - sy); nn += 1.f; }
    if (j < n[1] - 1) { s += Lds<T>::ld(z, id + sy); nn += 1.f; }
    if (DIM == 3) {
      if (k > 0) { s += Lds<T>::ld(z, id - 1); nn += 1.f; }
      if (k < n[2] - 1) { s += Lds<T>::ld(z, id + 1); nn += 1.f; }
    }
    Lds<T>::st(z, id, (Lds<T>::ld(r, id) + s) * inv_count(nn));
  }
  __syncthreads();
}

template <int DIM, typename T>
__global__ void __launch_bounds__(1024) k_tail(TailArgs A, const float* __restrict__ rin, float* __restrict__ zout,
                                               const double* shift, const int* flags) {
  if (is_done(flags)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* smem = reinterpret_cast<T*>(smem_raw);
  const float m = shift ? (float)*shift : 0.f;
  auto size_of = [&](int l) { return A.n[l][0] * A.n[l][1] * A.n[l][2]; };
  {
    T* r0 = smem + A.off[0];
    T* z0 = r0 + size_of(0);
    for (int id = threadIdx.x; id < size_of(0); id += blockDim.x) {
      Lds<T>::st(r0, id, rin[id] - m);
      Lds<T>::st(z0, id, 0.f);
    }
  }
  __syncthreads();
  // down: smooth from zero, restrict
  for (int l = 0; l < A.nl - 1; ++l) {
    T* r = smem + A.off[l];
    T* z = r + size_of(l);
    const int* n = A.n[l];
    for (int s = 0; s < A.pre; ++s) {
      lds_half_sweep<DIM, T>(r, z, n, 0);
      lds_half_sweep<DIM, T>(r, z, n, 1);
    }
    T* rc = smem + A.off[l + 1];
    T* zc = rc + size_of(l + 1);
    const int* nc = A.n[l + 1];
    for (int cid = threadIdx.x; cid < size_of(l + 1); cid += blockDim.x) {
      int ck = cid % nc[2], ct = cid / nc[2];
      int cj = ct % nc[1], ci = ct / nc[1];
      float acc = 0.f;
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
          for (int c = 0; c < (DIM == 3 ? 2 : 1); ++c) {
            int i = 2 * ci + a, j = 2 * cj + b, k = DIM == 3 ? 2 * ck + c : 0;
            int id = (i * n[1] + j) * n[2] + k;
            float s = 0.f, nn = 0.f;
            if (i > 0) { s += Lds<T>::ld(z, id - n[1] * n[2]); nn += 1.f; }
            if (i < n[0] - 1) { s += Lds<T>::ld(z, id + n[1] * n[2]); nn += 1.f; }
            if (j > 0) { s += Lds<T>::ld(z, id - n[2]); nn += 1.f; }
            if (j < n[1] - 1) { s += Lds<T>::ld(z, id + n[2]); nn += 1.f; }
            if (DIM == 3) {
              if (k > 0) { s += Lds<T>::ld(z, id - 1); nn += 1.f; }
              if (k < n[2] - 1) { s += Lds<T>::ld(z, id + 1); nn += 1.f; }
            }
            acc += Lds<T>::ld(r, id) - (nn * Lds<T>::ld(z, id) - s);
          }
      Lds<T>::st(rc, cid, acc * A.inv_dm1);
      Lds<T>::st(zc, cid, 0.f);
    }
    __syncthreads();
  }
  // bottom
  {
    int l = A.nl - 1;
    T* r = smem + A.off[l];
    T* z = r + size_of(l);
    if (A.bottom_exact) {
      lds_exact_solve<DIM, T>(r, z, A.n[l]);
    } else {
      for (int s = 0; s < A.bottom; ++s) {
        lds_half_sweep<DIM, T>(r, z, A.n[l], 0);
        lds_half_sweep<DIM, T>(r, z, A.n[l], 1);
      }
    }
  }
  // up: prolongate, post-smooth (colour 1 first)
  for (int l = A.nl - 2; l >= 0; --l) {
    T* r = smem + A.off[l];
    T* z = r + size_of(l);
    T* zc = smem + A.off[l + 1] + size_of(l + 1);
    const int* n = A.n[l];
    const int* nc = A.n[l + 1];
    for (int id = threadIdx.x; id < size_of(l); id += blockDim.x) {
      int k = id % n[2], t = id / n[2];
      int j = t % n[1], i = t / n[1];
      int cid = ((i >> 1) * nc[1] + (j >> 1)) * nc[2] + (DIM == 3 ? (k >> 1) : 0);
      Lds<T>::st(z, id, Lds<T>::ld(z, id) + Lds<T>::ld(zc, cid));
    }
    __syncthreads();
    for (int s = 0; s < A.pre; ++s) {
      lds_half_sweep<DIM, T>(r, z, n, 1);
      lds_half_sweep<DIM, T>(r, z, n, 0);
    }
  }
  T* z0 = smem + A.off[0] + size_of(0);
  for (int id = threadIdx.x; id < size_of(0); id += blockDim.x) zout[id] = Lds<T>::ld(z0, id);
}

// ------------------------------------------------------------- handle -----
struct Solver {
  Geo G;
  int L, pre, bottom;
  Lvl lv[kMaxLevels];
  double* x;  // PCG iterate (pressure), FP64
  float* b;   // recentred right-hand side, kept for the FP64 refinement residual
  int max_refine;
  float *p, *Ap;
  double* part;
  double* sc;
  int* flags;
  int* h_flags;
  double* h_sc;
  int last_iters;
  double last_res;
  int check_every;
  // coarse tail: levels tail_lt..L-1 run in one workgroup with LDS storage
  int bf16_coarse;   // option: store levels >= 2 in BF16 inside the tail
  int bottom_exact;  // option: exact (DCT) coarsest-level solve in the tail
  int tail_bf16;     // the tail in use is BF16
  int tail_lt;       // first tail level (L = no tail)
  size_t tail_bytes; // dynamic LDS of the tail kernel
  TailArgs tail;
  float* z0_final;   // level-0 buffer holding M^-1 r after the last V-cycle
  // captured PCG iterations
  int use_graph;
  hipStream_t cap;
  hipGraphExec_t gexec;
  int gexec_iters;
};

// Largest LDS a workgroup may request on this device.
size_t max_lds_bytes() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 65536;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) return 65536;
  return (size_t)v;
}

// Pick the first tail level: the largest suffix of levels whose r and z fit
// in LDS (BF16 storage allowed only for levels >= 2).
void configure_tail(Solver& s, size_t cap) {
  int best = s.L;
  size_t best_bytes = 0;
  for (int l = s.L - 1; l >= 0 && s.L - l <= kTailMax; --l) {
    const size_t es = (s.bf16_coarse && l >= 2) ? 2 : 4;
    size_t tot = 0;
    for (int k = l; k < s.L; ++k) tot += 2 * (size_t)s.lv[k].size * es;
    if (tot > cap) break;
    best = l;
    best_bytes = tot;
  }
  s.tail_lt = best;
  s.tail_bytes = best_bytes;
  s.tail_bf16 = (s.bf16_coarse && best >= 2) ? 1 : 0;
  if (best >= s.L) return;
  s.tail.nl = s.L - best;
  int off = 0;
  for (int k = 0; k < s.tail.nl; ++k) {
    for (int a = 0; a < 3; ++a) s.tail.n[k][a] = s.lv[best + k].n[a];
    s.tail.off[k] = off;
    off += 2 * (int)s.lv[best + k].size;
  }
  s.tail.pre = s.pre;
  s.tail.bottom = s.bottom;
  s.tail.bottom_exact = s.bottom_exact;
  s.tail.inv_dm1 = 1.0f / (float)(s.G.dim - 1);
}

// Dynamic-LDS opt-in for the tail instantiation this handle launches.
bool enable_tail_lds(const Solver& s) {
  if (s.tail_lt >= s.L) return true;
  const void* fn = s.G.dim == 2 ? (s.tail_bf16 ? (const void*)k_tail<2, uint16_t> : (const void*)k_tail<2, float>)
                                : (s.tail_bf16 ? (const void*)k_tail<3, uint16_t> : (const void*)k_tail<3, float>);
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)s.tail_bytes) == hipSuccess;
}

// Choose the tail for the current options; falls back to the 64 KB static
// limit (then to no tail) if the runtime refuses the larger allocation.
void setup_tail(Solver& s) {
  configure_tail(s, max_lds_bytes());
  if (!enable_tail_lds(s)) {
    (void)hipGetLastError();
    configure_tail(s, 65536);
    if (!enable_tail_lds(s)) {
      (void)hipGetLastError();
      s.tail_lt = s.L;
    }
  }
}

inline Dims dims_of(const Lvl& l) { return Dims{{l.n[0], l.n[1], l.n[2]}}; }

inline unsigned grid1(int64_t n) { return (unsigned)((n + kB - 1) / kB); }
inline unsigned gridr(int64_t n) {
  int64_t b = (n + kB - 1) / kB;
  return (unsigned)(b < kRedBlocks ? b : kRedBlocks);
}

template <int DIM>
void fin(Solver& s, hipStream_t st, int op, unsigned nparts, double tol = 0, int rel = 0) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kB), 0, st, op, s.part, (int)nparts, s.sc, s.flags, s.lv[0].size, tol,
                     rel);
}

// `sweeps` RBGS sweeps (colour c1 first, then the other) in place on lv[l].z,
// starting from z = 0 when from_zero (first sweep fused into one kernel).
// zc_prolong fuses the prolongation into the first half-sweep.  Returns lv[l].z.
template <int DIM>
float* sweeps_global(Solver& s, hipStream_t st, int l, int sweeps, int c1, bool from_zero,
                     const double* shift, const float* zc_prolong, Dims dc) {
  Lvl& L = s.lv[l];
  Dims d = dims_of(L);
  Dims none{{1, 1, 1}};
  int k0 = 0;
  if (from_zero) {
    hipLaunchKernelGGL((k_first_sweep<DIM>), dim3(grid1(L.size)), dim3(kB), 0, st, d, L.r, L.z, shift, s.flags);
    k0 = 1;
  }
  int inner = DIM == 3 ? d.n[2] : d.n[1];
  int64_t outer = DIM == 3 ? (int64_t)d.n[0] * d.n[1] : d.n[0];
  unsigned gs = grid1(outer * ((inner + 1) / 2));
  for (int it = k0; it < sweeps; ++it) {
    const float* zc = (it == k0 && zc_prolong) ? zc_prolong : nullptr;
    hipLaunchKernelGGL((k_smooth<DIM>), dim3(gs), dim3(kB), 0, st, d, L.r, L.z, c1, shift, zc, zc ? dc : none,
                       s.flags);
    hipLaunchKernelGGL((k_smooth<DIM>), dim3(gs), dim3(kB), 0, st, d, L.r, L.z, 1 - c1, shift, nullptr, none,
                       s.flags);
  }
  return L.z;
}

// apply_preconditioner (2D/vortex/mgpcg.py:182-200): z0 = M^-1 r0.  Levels
// below tail_lt use global kernels; levels tail_lt.. run in one LDS kernel.
template <int DIM>
void vcycle(Solver& s, hipStream_t st, const double* shift0) {
  Dims none{{1, 1, 1}};
  float* cur[kMaxLevels] = {nullptr};
  const bool tail = s.tail_lt < s.L;
  const int Lg = tail ? s.tail_lt : s.L - 1;  // levels smoothed with global kernels before the bottom
  for (int l = 0; l < Lg; ++l) {
    const double* sh = l == 0 ? shift0 : nullptr;
    cur[l] = sweeps_global<DIM>(s, st, l, s.pre, 0, true, sh, nullptr, none);
    Lvl& F = s.lv[l];
    Lvl& C = s.lv[l + 1];
    hipLaunchKernelGGL((k_restrict<DIM>), dim3(grid1(C.size)), dim3(kB), 0, st, dims_of(F), F.r, cur[l], dims_of(C),
                       C.r, sh, s.flags);
  }
  if (tail) {
    Lvl& T = s.lv[s.tail_lt];
    const double* sh = s.tail_lt == 0 ? shift0 : nullptr;
    if (s.tail_bf16)
      hipLaunchKernelGGL((k_tail<DIM, uint16_t>), dim3(1), dim3(1024), s.tail_bytes, st, s.tail, T.r, T.z, sh, s.flags);
    else
      hipLaunchKernelGGL((k_tail<DIM, float>), dim3(1), dim3(1024), s.tail_bytes, st, s.tail, T.r, T.z, sh, s.flags);
    cur[s.tail_lt] = T.z;
  } else {
    const double* sh = s.L == 1 ? shift0 : nullptr;
    cur[s.L - 1] = sweeps_global<DIM>(s, st, s.L - 1, s.bottom, 0, true, sh, nullptr, none);
  }
  for (int l = Lg - 1; l >= 0; --l) {
    const double* sh = l == 0 ? shift0 : nullptr;
    cur[l] = sweeps_global<DIM>(s, st, l, s.pre, 1, false, sh, cur[l + 1], dims_of(s.lv[l + 1]));
  }
  s.z0_final = cur[0];
}

// One PCG iteration once z, p and z.r are set (solve(), 2D/vortex/mgpcg.py:235-278):
// Ap and p.Ap, alpha, x/r update and r.r (convergence flag), recentring
// V-cycle, z.r and beta, p update.  Every kernel is a no-op once `done` is set,
// so iterations may be queued past convergence.
template <int DIM>
void pcg_iteration(Solver& s, hipStream_t st) {
  Lvl& L0 = s.lv[0];
  int64_t N = L0.size;
  unsigned gr = gridr(N);
  Dims d0 = dims_of(L0);
  hipLaunchKernelGGL((k_Ap_dot<DIM>), dim3(gr), dim3(kB), 0, st, d0, s.p, s.Ap, s.part, s.flags);
  fin<DIM>(s, st, FIN_ALPHA, gr);
  hipLaunchKernelGGL(k_update_xr, dim3(gr), dim3(kB), 0, st, N, s.x, L0.r, s.p, s.Ap, s.sc, s.part, s.flags);
  fin<DIM>(s, st, FIN_RR, gr);
  vcycle<DIM>(s, st, s.sc + S_MEAN);
  hipLaunchKernelGGL(k_zr, dim3(gr), dim3(kB), 0, st, N, s.z0_final, L0.r, (const double*)(s.sc + S_MEAN), s.part,
                     s.flags, 0, (float*)nullptr);
  fin<DIM>(s, st, FIN_BETA, gr);
  hipLaunchKernelGGL(k_update_p, dim3(gr), dim3(kB), 0, st, N, s.z0_final, s.p, s.sc, s.flags);
}

void drop_graph(Solver& s) {
  if (s.gexec) (void)hipGraphExecDestroy(s.gexec);
  s.gexec = nullptr;
  s.gexec_iters = 0;
}

// Capture check_every PCG iterations into one executable graph (on a private
// capture stream, default priority); replayed on the caller's stream.  The
// graph holds kernel launches only -- every pointer it uses is owned by the
// handle and stable -- so one capture serves every later solve.
template <int DIM>
bool build_graph(Solver& s) {
  drop_graph(s);
  if (!s.cap && hipStreamCreateWithFlags(&s.cap, hipStreamNonBlocking) != hipSuccess) {
    s.cap = nullptr;
    return false;
  }
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(s.cap, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
  for (int q = 0; q < s.check_every; ++q) pcg_iteration<DIM>(s, s.cap);
  if (hipStreamEndCapture(s.cap, &g) != hipSuccess || !g) {
    (void)hipGetLastError();
    return false;
  }
  bool ok = hipGraphInstantiate(&s.gexec, g, nullptr, nullptr, 0) == hipSuccess;
  (void)hipGraphDestroy(g);
  if (!ok) {
    s.gexec = nullptr;
    (void)hipGetLastError();
    return false;
  }
  s.gexec_iters = s.check_every;
  return true;
}

// PCG from r0 (already in lv[0].r, mean sum in partials): solve(),
// 2D/vortex/mgpcg.py:202-282.
template <int DIM>
int pcg(Solver& s, hipStream_t st, double tol, int relative, int max_iters) {
  Lvl& L0 = s.lv[0];
  int64_t N = L0.size;
  unsigned gr = gridr(N);
  fin<DIM>(s, st, FIN_MEAN, gr);
  hipLaunchKernelGGL(k_recenter_norm, dim3(gr), dim3(kB), 0, st, N, L0.r, s.b, s.x, s.sc, s.part);
  fin<DIM>(s, st, FIN_BNORM, gr, tol, relative);
  vcycle<DIM>(s, st, nullptr);
  hipLaunchKernelGGL(k_zr, dim3(gr), dim3(kB), 0, st, N, s.z0_final, L0.r, (const double*)nullptr, s.part, s.flags, 1, s.p);
  fin<DIM>(s, st, FIN_ZR_INIT, gr);
  Dims d0 = dims_of(L0);
  int it = 0;
  int refinements = 0;
  bool done = false;
  auto read_flags = [&]() {
    return hipMemcpyAsync(s.h_flags, s.flags, sizeof(int) * F_NFLAG, hipMemcpyDeviceToHost, st) == hipSuccess &&
           hipStreamSynchronize(st) == hipSuccess;
  };
  bool use_graph = s.use_graph;
  if (use_graph && (!s.gexec || s.gexec_iters != s.check_every)) use_graph = build_graph<DIM>(s);
  while (it < max_iters) {
    int k = s.check_every;
    if (it + k > max_iters) k = max_iters - it;
    if (use_graph && k == s.gexec_iters) {
      if (hipGraphLaunch(s.gexec, st) != hipSuccess) return set_error(DFM_E_HIP, "mgpcg: graph launch failed");
    } else {
      for (int q = 0; q < k; ++q) pcg_iteration<DIM>(s, st);
    }
    it += k;
    {
      if (!read_flags()) return set_error(DFM_E_HIP, "mgpcg: flag readback failed");
      if (s.h_flags[F_DONE] && s.h_flags[F_STATUS] == 1 && refinements < s.max_refine) {
        // Converged on the recursive FP32 residual: measure the true residual
        // b - A x in FP64; if it misses the tolerance, restart PCG from it.
        hipLaunchKernelGGL((k_true_resid<DIM>), dim3(gr), dim3(kB), 0, st, d0, s.x, s.b, L0.r, s.part);
        fin<DIM>(s, st, FIN_TRUE, gr);
        if (!read_flags()) return set_error(DFM_E_HIP, "mgpcg: flag readback failed");
        if (!s.h_flags[F_DONE]) {
          ++refinements;
          vcycle<DIM>(s, st, s.sc + S_MEAN);
          hipLaunchKernelGGL(k_zr, dim3(gr), dim3(kB), 0, st, N, s.z0_final, L0.r, (const double*)(s.sc + S_MEAN), s.part,
                             s.flags, 1, s.p);
          fin<DIM>(s, st, FIN_ZR_INIT, gr);
          continue;
        }
      }
      if (s.h_flags[F_DONE]) { done = true; break; }
    }
  }
  bool rb = hipMemcpyAsync(s.h_sc, s.sc, sizeof(double) * S_NSLOT, hipMemcpyDeviceToHost, st) == hipSuccess;
  if (!done) rb = rb && hipMemcpyAsync(s.h_flags, s.flags, sizeof(int) * F_NFLAG, hipMemcpyDeviceToHost, st) == hipSuccess;
  if (!rb || hipStreamSynchronize(st) != hipSuccess) return set_error(DFM_E_HIP, "mgpcg: status readback failed");
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(DFM_E_HIP, "mgpcg launch", hipGetErrorString(e));
  s.last_iters = s.h_flags[F_ITERS];
  double bb = s.h_sc[S_BB];
  s.last_res = bb > 0 ? std::sqrt(s.h_sc[S_RR] / bb) : 0.0;
  if (s.h_flags[F_STATUS] == 2) return set_error(DFM_E_NONFINITE, "mgpcg: residual is not finite");
  if (!s.h_flags[F_DONE]) return set_error(DFM_E_NOCONV, "mgpcg: max_iters reached without convergence");
  return DFM_OK;
}

void destroy(Solver* s) {
  if (!s) return;
  for (int l = 0; l < s->L; ++l) {
    if (s->lv[l].r) (void)hipFree(s->lv[l].r);
    if (s->lv[l].z) (void)hipFree(s->lv[l].z);
  }
  if (s->x) (void)hipFree(s->x);
  if (s->b) (void)hipFree(s->b);
  if (s->p) (void)hipFree(s->p);
  if (s->Ap) (void)hipFree(s->Ap);
  if (s->part) (void)hipFree(s->part);
  if (s->sc) (void)hipFree(s->sc);
  if (s->flags) (void)hipFree(s->flags);
  if (s->h_flags) (void)hipHostFree(s->h_flags);
  if (s->h_sc) (void)hipHostFree(s->h_sc);
  drop_graph(*s);
  if (s->cap) (void)hipStreamDestroy(s->cap);
  delete s;
}

}  // namespace

extern "C" int dfm_mgpcg_create(const dfm_grid* g, int levels, int pre_post, int bottom, void** handle) {
  if (!g || !handle) return set_error(DFM_E_NULL, "dfm_mgpcg_create: NULL argument");
  *handle = nullptr;
  if (g->dim != 2 && g->dim != 3) return set_error(DFM_E_ARG, "dim must be 2 or 3");
  if (pre_post < 1 || bottom < 1) return set_error(DFM_E_ARG, "pre_post and bottom must be >= 1");
  int nmin = g->n[0];
  for (int a = 1; a < g->dim; ++a) nmin = g->n[a] < nmin ? g->n[a] : nmin;
  if (nmin < 2) return set_error(DFM_E_ARG, "grid too small");
  if (levels <= 0) {
    int lg = 0;
    while ((1 << (lg + 1)) <= nmin) ++lg;
    levels = lg - 3 + 1;
    if (levels < 1) levels = 1;
  }
  if (levels > kMaxLevels) return set_error(DFM_E_ARG, "too many levels");
  for (int a = 0; a < g->dim; ++a)
    if (g->n[a] % (1 << (levels - 1)) != 0) return set_error(DFM_E_ARG, "n must be divisible by 2^(levels-1)");
  if (g->dim == 2 && g->n[2] != 1) return set_error(DFM_E_ARG, "2D grid must have n[2] == 1");
  Solver* s = new (std::nothrow) Solver();
  if (!s) return set_error(DFM_E_HIP, "out of host memory");
  std::memset(s, 0, sizeof(Solver));
  s->G.dim = g->dim;
  for (int a = 0; a < 3; ++a) s->G.n[a] = g->n[a];
  s->G.dx = g->dx;
  s->G.inv_dx = 1.0f / g->dx;
  s->L = levels;
  s->pre = pre_post;
  s->bottom = bottom;
  s->check_every = 1;
  s->max_refine = 3;
  s->use_graph = 1;
  bool ok = true;
  for (int l = 0; l < levels && ok; ++l) {
    for (int a = 0; a < 3; ++a) s->lv[l].n[a] = a < g->dim ? (g->n[a] >> l) : 1;
    s->lv[l].size = (int64_t)s->lv[l].n[0] * s->lv[l].n[1] * s->lv[l].n[2];
    ok = hipMalloc(&s->lv[l].r, sizeof(float) * s->lv[l].size) == hipSuccess &&
         hipMalloc(&s->lv[l].z, sizeof(float) * s->lv[l].size) == hipSuccess;
  }
  int64_t N = s->lv[0].size;
  ok = ok && hipMalloc(&s->x, sizeof(double) * N) == hipSuccess && hipMalloc(&s->b, sizeof(float) * N) == hipSuccess &&
       hipMalloc(&s->p, sizeof(float) * N) == hipSuccess &&
       hipMalloc(&s->Ap, sizeof(float) * N) == hipSuccess &&
       hipMalloc(&s->part, sizeof(double) * 2 * kRedBlocks) == hipSuccess &&
       hipMalloc(&s->sc, sizeof(double) * S_NSLOT) == hipSuccess &&
       hipMalloc(&s->flags, sizeof(int) * F_NFLAG) == hipSuccess &&
       hipHostMalloc(&s->h_flags, sizeof(int) * F_NFLAG, hipHostMallocDefault) == hipSuccess &&
       hipHostMalloc(&s->h_sc, sizeof(double) * S_NSLOT, hipHostMallocDefault) == hipSuccess;
  if (ok) ok = hipMemset(s->part, 0, sizeof(double) * 2 * kRedBlocks) == hipSuccess &&
               hipMemset(s->sc, 0, sizeof(double) * S_NSLOT) == hipSuccess &&
               hipMemset(s->flags, 0, sizeof(int) * F_NFLAG) == hipSuccess &&
               hipMemset(s->x, 0, sizeof(double) * N) == hipSuccess && hipDeviceSynchronize() == hipSuccess;
  if (!ok) {
    destroy(s);
    return set_error(DFM_E_HIP, "dfm_mgpcg_create: device allocation failed");
  }
  setup_tail(*s);
  *handle = s;
  return DFM_OK;
}

extern "C" int dfm_mgpcg_destroy(void* handle) {
  destroy(static_cast<Solver*>(handle));
  return DFM_OK;
}

extern "C" int dfm_mgpcg_set_option(void* handle, int option, int value) {
  Solver* s = static_cast<Solver*>(handle);
  if (!s) return set_error(DFM_E_NULL, "mgpcg handle is NULL");
  switch (option) {
    case DFM_MG_BF16_COARSE:
      s->bf16_coarse = value ? 1 : 0;
      setup_tail(*s);
      drop_graph(*s);
      return DFM_OK;
    case DFM_MG_TAIL:
      if (value) setup_tail(*s);
      else s->tail_lt = s->L;
      drop_graph(*s);
      return DFM_OK;
    case DFM_MG_GRAPH:
      s->use_graph = value ? 1 : 0;
      drop_graph(*s);
      return DFM_OK;
    case DFM_MG_BOTTOM_EXACT:
      s->bottom_exact = value ? 1 : 0;
      setup_tail(*s);
      drop_graph(*s);
      return DFM_OK;
    case DFM_MG_MAX_REFINE:
      if (value < 0) return set_error(DFM_E_ARG, "max_refine must be >= 0");
      s->max_refine = value;
      return DFM_OK;
    case DFM_MG_CHECK_EVERY:
      if (value < 1) return set_error(DFM_E_ARG, "check_every must be >= 1");
      s->check_every = value;
      drop_graph(*s);
      return DFM_OK;
    default:
      return set_error(DFM_E_ARG, "unknown MGPCG option");
  }
}

extern "C" int dfm_mgpcg_get_option(void* handle, int option, int* value) {
  Solver* s = static_cast<Solver*>(handle);
  if (!s || !value) return set_error(DFM_E_NULL, "dfm_mgpcg_get_option: NULL argument");
  switch (option) {
    case DFM_MG_BF16_COARSE: *value = s->tail_bf16; return DFM_OK;
    case DFM_MG_TAIL: *value = s->tail_lt; return DFM_OK;
    case DFM_MG_MAX_REFINE: *value = s->max_refine; return DFM_OK;
    case DFM_MG_CHECK_EVERY: *value = s->check_every; return DFM_OK;
    case DFM_MG_GRAPH: *value = s->gexec ? 2 : s->use_graph; return DFM_OK;
    case DFM_MG_BOTTOM_EXACT: *value = (s->tail_lt < s->L) ? s->bottom_exact : 0; return DFM_OK;
    default: return set_error(DFM_E_ARG, "unknown MGPCG option");
  }
}

extern "C" int dfm_mgpcg_project(void* handle, const dfm_mac* u, double tol, int relative, int max_iters,
                                 void* stream) {
  Solver* s = static_cast<Solver*>(handle);
  if (!s) return set_error(DFM_E_NULL, "mgpcg handle is NULL");
  if (!u) return set_error(DFM_E_NULL, "velocity is NULL");
  for (int a = 0; a < s->G.dim; ++a)
    if (!u->c[a]) return set_error(DFM_E_NULL, "velocity component is NULL");
  if (max_iters < 1) return set_error(DFM_E_ARG, "max_iters must be >= 1");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  Lvl& L0 = s->lv[0];
  Dims d0 = dims_of(L0);
  Faces F{{u->c[0], u->c[1], u->c[2]}};
  unsigned gr = gridr(L0.size);
  int rc;
  if (s->G.dim == 2) {
    hipLaunchKernelGGL((k_div<2>), dim3(gr), dim3(kB), 0, st, d0, F, L0.r, s->part);
    rc = pcg<2>(*s, st, tol, relative, max_iters);
  } else {
    hipLaunchKernelGGL((k_div<3>), dim3(gr), dim3(kB), 0, st, d0, F, L0.r, s->part);
    rc = pcg<3>(*s, st, tol, relative, max_iters);
  }
  if (rc != DFM_OK && rc != DFM_E_NOCONV) return rc;
  // subtract the pressure gradient even when max_iters was hit (the reference
  // prints and continues, 2D/vortex/mgpcg.py:280-282)
  int64_t nf = 0;
  for (int a = 0; a < s->G.dim; ++a) nf += (L0.size / L0.n[a]) * (L0.n[a] + 1);
  unsigned gf = (unsigned)((nf + kB - 1) / kB);
  if (gf > 65535u * 8u) gf = 65535u * 8u;
  if (s->G.dim == 2) hipLaunchKernelGGL((k_subgrad<2>), dim3(gf), dim3(kB), 0, st, d0, F, s->x);
  else hipLaunchKernelGGL((k_subgrad<3>), dim3(gf), dim3(kB), 0, st, d0, F, s->x);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(DFM_E_HIP, "mgpcg subtract_grad_p", hipGetErrorString(e));
  return rc;
}

extern "C" int dfm_mgpcg_solve(void* handle, const float* b, float* x, double tol, int relative, int max_iters,
                               void* stream) {
  Solver* s = static_cast<Solver*>(handle);
  if (!s) return set_error(DFM_E_NULL, "mgpcg handle is NULL");
  if (!b || !x) return set_error(DFM_E_NULL, "b/x is NULL");
  if (max_iters < 1) return set_error(DFM_E_ARG, "max_iters must be >= 1");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  Lvl& L0 = s->lv[0];
  unsigned gr = gridr(L0.size);
  hipLaunchKernelGGL(k_load_b, dim3(gr), dim3(kB), 0, st, L0.size, b, L0.r, s->part);
  int rc = s->G.dim == 2 ? pcg<2>(*s, st, tol, relative, max_iters) : pcg<3>(*s, st, tol, relative, max_iters);
  if (rc != DFM_OK && rc != DFM_E_NOCONV) return rc;
  hipLaunchKernelGGL(k_to_float, dim3(gr), dim3(kB), 0, st, L0.size, s->x, x);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(DFM_E_HIP, "mgpcg: solution copy", hipGetErrorString(e));
  return rc;
}

extern "C" int dfm_mgpcg_status(void* handle, int* iters, double* rel_residual) {
  Solver* s = static_cast<Solver*>(handle);
  if (!s) return set_error(DFM_E_NULL, "mgpcg handle is NULL");
  if (iters) *iters = s->last_iters;
  if (rel_residual) *rel_residual = s->last_res;
  return DFM_OK;
}

extern "C" int dfm_mgpcg_pressure(void* handle, double** p) {
  Solver* s = static_cast<Solver*>(handle);
  if (!s || !p) return set_error(DFM_E_NULL, "dfm_mgpcg_pressure: NULL argument");
  *p = s->x;
  return DFM_OK;
}
