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

DFM_HD void cell_of(const Dims& d, int64_t id, int& i, int& j, int& k) {
  k = (int)(id % d.n[2]);
  int64_t t = id / d.n[2];
  j = (int)(t % d.n[1]);
  i = (int)(t / d.n[1]);
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
    z[id] = rv / nn;
    return;
  }
  float s = 0.f;
  if (i > 0) s += (r[cidx(d, i - 1, j, k)] - m) / nbr_count<DIM>(d, i - 1, j, k);
  if (i < d.n[0] - 1) s += (r[cidx(d, i + 1, j, k)] - m) / nbr_count<DIM>(d, i + 1, j, k);
  if (j > 0) s += (r[cidx(d, i, j - 1, k)] - m) / nbr_count<DIM>(d, i, j - 1, k);
  if (j < d.n[1] - 1) s += (r[cidx(d, i, j + 1, k)] - m) / nbr_count<DIM>(d, i, j + 1, k);
  if (DIM == 3) {
    if (k > 0) s += (r[cidx(d, i, j, k - 1)] - m) / nbr_count<DIM>(d, i, j, k - 1);
    if (k < d.n[2] - 1) s += (r[cidx(d, i, j, k + 1)] - m) / nbr_count<DIM>(d, i, j, k + 1);
  }
  z[id] = (rv + s) / nn;
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
  int hm = (int)(t % half);
  int64_t o = t / half;
  int i, j, k;
  if (DIM == 3) {
    i = (int)(o / d.n[1]);
    j = (int)(o % d.n[1]);
    k = 2 * hm + ((phase - i - j) & 1);
    if (k >= d.n[2]) return;
  } else {
    i = (int)o;
    k = 0;
    j = 2 * hm + ((phase - i) & 1);
    if (j >= d.n[1]) return;
  }
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
  z[id] = (r[id] - m + s) / nn;
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

// Coarsest level in one workgroup: `iters` RBGS sweeps from z = 0 in LDS.
template <int DIM>
__global__ void __launch_bounds__(1024) k_bottom(Dims d, const float* __restrict__ r, float* __restrict__ z, int iters,
                                                 const double* shift, const int* flags) {
  if (is_done(flags)) return;
  __shared__ float sr[kLdsCells];
  __shared__ float sz[kLdsCells];
  float m = shift ? (float)*shift : 0.f;
  int N = d.n[0] * d.n[1] * d.n[2];
  for (int id = threadIdx.x; id < N; id += blockDim.x) {
    sr[id] = r[id] - m;
    sz[id] = 0.f;
  }
  __syncthreads();
  for (int it = 0; it < iters; ++it) {
    for (int phase = 0; phase < 2; ++phase) {
      for (int id = threadIdx.x; id < N; id += blockDim.x) {
        int i, j, k;
        cell_of(d, id, i, j, k);
        if (((i + j + k) & 1) != phase) continue;
        float nn;
        float s = nbr_sum<DIM>(sz, d, i, j, k, nn);
        sz[id] = (sr[id] + s) / nn;
      }
      __syncthreads();
    }
  }
  for (int id = threadIdx.x; id < N; id += blockDim.x) z[id] = sz[id];
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
};

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

template <int DIM>
void smooth_levels_global(Solver& s, hipStream_t st, int l, int sweeps, const double* shift) {
  Lvl& L = s.lv[l];
  Dims d = dims_of(L);
  Dims none{{1, 1, 1}};
  hipLaunchKernelGGL((k_first_sweep<DIM>), dim3(grid1(L.size)), dim3(kB), 0, st, d, L.r, L.z, shift, s.flags);
  int inner = DIM == 3 ? d.n[2] : d.n[1];
  int64_t outer = DIM == 3 ? (int64_t)d.n[0] * d.n[1] : d.n[0];
  unsigned gs = grid1(outer * ((inner + 1) / 2));
  for (int it = 1; it < sweeps; ++it) {
    hipLaunchKernelGGL((k_smooth<DIM>), dim3(gs), dim3(kB), 0, st, d, L.r, L.z, 0, shift, nullptr, none, s.flags);
    hipLaunchKernelGGL((k_smooth<DIM>), dim3(gs), dim3(kB), 0, st, d, L.r, L.z, 1, shift, nullptr, none, s.flags);
  }
}

// apply_preconditioner (2D/vortex/mgpcg.py:182-200): z0 = M^-1 r0
template <int DIM>
void vcycle(Solver& s, hipStream_t st, const double* shift0) {
  Dims none{{1, 1, 1}};
  for (int l = 0; l < s.L - 1; ++l) {
    const double* sh = l == 0 ? shift0 : nullptr;
    smooth_levels_global<DIM>(s, st, l, s.pre, sh);
    Lvl& F = s.lv[l];
    Lvl& C = s.lv[l + 1];
    hipLaunchKernelGGL((k_restrict<DIM>), dim3(grid1(C.size)), dim3(kB), 0, st, dims_of(F), F.r, F.z, dims_of(C), C.r,
                       sh, s.flags);
  }
  Lvl& B = s.lv[s.L - 1];
  const double* shb = s.L == 1 ? shift0 : nullptr;
  if (B.size <= kLdsCells) {
    hipLaunchKernelGGL((k_bottom<DIM>), dim3(1), dim3(1024), 0, st, dims_of(B), B.r, B.z, s.bottom, shb, s.flags);
  } else {
    smooth_levels_global<DIM>(s, st, s.L - 1, s.bottom, shb);
  }
  for (int l = s.L - 2; l >= 0; --l) {
    const double* sh = l == 0 ? shift0 : nullptr;
    Lvl& F = s.lv[l];
    Lvl& C = s.lv[l + 1];
    Dims d = dims_of(F);
    int inner = DIM == 3 ? d.n[2] : d.n[1];
    int64_t outer = DIM == 3 ? (int64_t)d.n[0] * d.n[1] : d.n[0];
    unsigned gs = grid1(outer * ((inner + 1) / 2));
    // prolongate + smooth(l, 1), then smooth(l, 0); repeat
    hipLaunchKernelGGL((k_smooth<DIM>), dim3(gs), dim3(kB), 0, st, d, F.r, F.z, 1, sh, C.z, dims_of(C), s.flags);
    hipLaunchKernelGGL((k_smooth<DIM>), dim3(gs), dim3(kB), 0, st, d, F.r, F.z, 0, sh, nullptr, none, s.flags);
    for (int it = 1; it < s.pre; ++it) {
      hipLaunchKernelGGL((k_smooth<DIM>), dim3(gs), dim3(kB), 0, st, d, F.r, F.z, 1, sh, nullptr, none, s.flags);
      hipLaunchKernelGGL((k_smooth<DIM>), dim3(gs), dim3(kB), 0, st, d, F.r, F.z, 0, sh, nullptr, none, s.flags);
    }
  }
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
  hipLaunchKernelGGL(k_zr, dim3(gr), dim3(kB), 0, st, N, L0.z, L0.r, (const double*)nullptr, s.part, s.flags, 1, s.p);
  fin<DIM>(s, st, FIN_ZR_INIT, gr);
  Dims d0 = dims_of(L0);
  int it = 0;
  int refinements = 0;
  bool done = false;
  auto read_flags = [&]() {
    return hipMemcpyAsync(s.h_flags, s.flags, sizeof(int) * F_NFLAG, hipMemcpyDeviceToHost, st) == hipSuccess &&
           hipStreamSynchronize(st) == hipSuccess;
  };
  while (it < max_iters) {
    hipLaunchKernelGGL((k_Ap_dot<DIM>), dim3(gr), dim3(kB), 0, st, d0, s.p, s.Ap, s.part, s.flags);
    fin<DIM>(s, st, FIN_ALPHA, gr);
    hipLaunchKernelGGL(k_update_xr, dim3(gr), dim3(kB), 0, st, N, s.x, L0.r, s.p, s.Ap, s.sc, s.part, s.flags);
    fin<DIM>(s, st, FIN_RR, gr);
    ++it;
    bool poll = (it % s.check_every) == 0 || it == max_iters;
    if (poll) {
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
          hipLaunchKernelGGL(k_zr, dim3(gr), dim3(kB), 0, st, N, L0.z, L0.r, (const double*)(s.sc + S_MEAN), s.part,
                             s.flags, 1, s.p);
          fin<DIM>(s, st, FIN_ZR_INIT, gr);
          continue;
        }
      }
      if (s.h_flags[F_DONE]) { done = true; break; }
    }
    if (it == max_iters) break;
    vcycle<DIM>(s, st, s.sc + S_MEAN);
    hipLaunchKernelGGL(k_zr, dim3(gr), dim3(kB), 0, st, N, L0.z, L0.r, (const double*)(s.sc + S_MEAN), s.part,
                       s.flags, 0, (float*)nullptr);
    fin<DIM>(s, st, FIN_BETA, gr);
    hipLaunchKernelGGL(k_update_p, dim3(gr), dim3(kB), 0, st, N, L0.z, s.p, s.sc, s.flags);
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
  *handle = s;
  return DFM_OK;
}

extern "C" int dfm_mgpcg_destroy(void* handle) {
  destroy(static_cast<Solver*>(handle));
  return DFM_OK;
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
