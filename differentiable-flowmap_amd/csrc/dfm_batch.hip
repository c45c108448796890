// Batched forward + adjoint of many independent 2D flow-map problems of
// uneven size (MI355X / gfx950).
//
// Every instance keeps the single-grid simulator's step order
// (LFM_Simulator.step_flow_map / LFM_Diff_Simulator.backtrack_flow_map,
// 2D/vortex/lfm_midpoint_simulator.py:517-661,
// 2D/vortex/lfm_midpoint_diff_simulator.py:260-503, with RK2, the quartic
// convective source, plain BFECC and the record-everything cache), but each
// stage is ONE launch over all instances:
//  * an instance table on the device holds each instance's grid, dt and the
//    offsets of its fields in shared pools (one allocation per field kind);
//  * every launch domain (face points, cells, a multigrid level, ...) gives
//    each instance a whole number of blocks, so a block belongs to one
//    instance and finds it by a binary search of the block prefix table;
//  * the MGPCG is batched the same way: per-instance FP64 scalars and
//    convergence flags, deterministic per-instance reductions (the last
//    block of an instance sums the instance's block partials in block
//    order), a converged instance's blocks return at once, the global
//    multigrid levels are batched launches and each instance's coarse levels
//    run in one LDS-resident workgroup of a single "tail" launch;
//  * the launches between two convergence checks are captured once into a
//    hipGraph (per step and segment) and replayed; a check reads the
//    per-instance flags back (one small copy) and replays a chunk of PCG
//    iterations until every instance has converged.
#include "dfm_common.hpp"
#include "dfm_points.hpp"
#include "dfm_internal.hpp"
#include "../../include/dfm.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

using namespace dfm;

namespace {

constexpr int BS = 256;             // threads per block of the point kernels
constexpr int kMaxLv = 10;          // multigrid levels per instance
constexpr int kTailThreads = 1024;  // threads of an instance's LDS tail workgroup
constexpr int kCtrStride = 32;      // counters on separate 128-byte lines
constexpr int kLdsBytes = 160 * 1024;

// Per-instance scalars (FP64) and flags.
enum { S_MEAN = 0, S_BB, S_RR, S_THRESH, S_ZR, S_ALPHA, S_BETA, S_LOSS, S_COUNT };
enum { F_DONE = 0, F_ITERS, F_STATUS, F_COUNT };  // status: 1 converged, 2 non-finite, 3 max_iters

struct Inst {
  int n0, n1;
  float dx, inv_dx, dt;
  int L, T;            // multigrid levels; first level held by the LDS tail
  int64_t oc, of;      // offsets into the cell pools (n0 n1) and MAC pools ((n0+1) n1 + n0 (n1+1))
  int64_t om[kMaxLv];  // offsets of levels 0..T in the multigrid level pools
  double inv_n;
  int max_iters;
};

__host__ __device__ inline int n_of(const Inst& I, int l, int a) { return (a == 0 ? I.n0 : I.n1) >> l; }
__host__ __device__ inline int64_t faces0(const Inst& I) { return (int64_t)(I.n0 + 1) * I.n1; }
__host__ __device__ inline int64_t faces1(const Inst& I) { return (int64_t)I.n0 * (I.n1 + 1); }
__host__ __device__ inline int64_t cells(const Inst& I) { return (int64_t)I.n0 * I.n1; }

DFM_HD Geo geo_of(const Inst& I) {
  Geo g;
  g.dim = 2;
  g.n[0] = I.n0;
  g.n[1] = I.n1;
  g.n[2] = 1;
  g.dx = I.dx;
  g.inv_dx = I.inv_dx;
  return g;
}
DFM_HD Mac mac_of(const float* base, const Inst& I) {
  Mac m;
  m.c[0] = base + I.of;
  m.c[1] = base + I.of + faces0(I);
  m.c[2] = nullptr;
  return m;
}
DFM_HD float* mac_c(float* base, const Inst& I, int a) { return base + I.of + (a ? faces0(I) : 0); }

// Flow map of an instance in the map pool: psi (2 C), psi_f[0], psi_f[1]
// (2 F_a each), T_f[0], T_f[1] (2 F_a each), SoA per set.
struct MapV {
  float* psi;
  float* pf[2];
  float* tf[2];
};
DFM_HD MapV map_of(float* base, const Inst& I) {
  MapV m;
  float* p = base + 2 * I.oc + 4 * I.of;
  const int64_t C = cells(I), F0 = faces0(I), F1 = faces1(I);
  m.psi = p;
  m.pf[0] = p + 2 * C;
  m.pf[1] = m.pf[0] + 2 * F0;
  m.tf[0] = m.pf[1] + 2 * F1;
  m.tf[1] = m.tf[0] + 2 * F0;
  return m;
}

// Launch domain: blk[i] .. blk[i+1] are instance i's blocks.
struct Dom {
  const int* blk;
  int count;
};

// Largest instance whose first block is <= b (instances with no blocks are
// skipped, since they share their first block with the next instance).
DFM_HD int find_inst(const Dom& D, int b) {
  int lo = 0, hi = D.count;
  while (hi - lo > 1) {
    const int m = (lo + hi) >> 1;
    if (D.blk[m] <= b) lo = m;
    else hi = m;
  }
  return lo;
}

struct Ctx {
  const Inst* I;
  double* sc;   // S_COUNT per instance
  int* fl;      // F_COUNT per instance
  double* part; // 2 per block of the largest reducing launch
  unsigned* ctr;
  double tol;
};

DFM_HD bool inst_done(const Ctx& C, int i) { return C.fl[i * F_COUNT + F_DONE] != 0; }

// Point index within its instance of this thread (block-aligned domains).
struct Where {
  int i;
  int64_t t;
};
DFM_HD Where where(const Dom& D) {
  Where w;
  w.i = find_inst(D, blockIdx.x);
  w.t = (int64_t)(blockIdx.x - D.blk[w.i]) * BS + threadIdx.x;
  return w;
}

// ------------------------------------------------------------ reductions ----
DFM_HD double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
template <int NT>
DFM_HD void block_sum2(double& a, double& b) {
  __shared__ double sa[NT / 64], sb[NT / 64];
  a = wave_sum(a);
  b = wave_sum(b);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();  // the arrays may be reused by a previous call
  if (l == 0) {
    sa[w] = a;
    sb[w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = 0;
    b = 0;
    for (int t = 0; t < NT / 64; ++t) {
      a += sa[t];
      b += sb[t];
    }
  }
}
DFM_HD void wait_stores() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

enum Fin { FIN_MEAN = 0, FIN_BNORM, FIN_ZR_INIT, FIN_ALPHA, FIN_RR, FIN_BETA, FIN_LOSS };

DFM_HD double safe_ratio(double a, double b) { return b != 0.0 ? a / b : 0.0; }

DFM_HD void fin_op(const Ctx& C, int i, int op, double a, double b) {
  double* sc = C.sc + (int64_t)i * S_COUNT;
  int* fl = C.fl + i * F_COUNT;
  switch (op) {
    case FIN_MEAN:
      sc[S_MEAN] = a * C.I[i].inv_n;
      break;
    case FIN_BNORM:
      sc[S_MEAN] = 0.0;
      sc[S_BB] = a;
      sc[S_RR] = a;
      sc[S_THRESH] = C.tol * C.tol * a;
      fl[F_ITERS] = 0;
      fl[F_STATUS] = (a == 0.0) ? 1 : 0;
      fl[F_DONE] = (a == 0.0) ? 1 : 0;
      break;
    case FIN_ZR_INIT:
      sc[S_ZR] = a;
      break;
    case FIN_ALPHA:
      sc[S_ALPHA] = safe_ratio(sc[S_ZR], a);
      break;
    case FIN_RR:
      sc[S_RR] = a;
      sc[S_MEAN] = b * C.I[i].inv_n;
      fl[F_ITERS] += 1;
      if (!isfinite(a)) {
        fl[F_STATUS] = 2;
        fl[F_DONE] = 1;
      } else if (a < sc[S_THRESH]) {
        fl[F_STATUS] = 1;
        fl[F_DONE] = 1;
      } else if (fl[F_ITERS] >= C.I[i].max_iters) {
        fl[F_STATUS] = 3;
        fl[F_DONE] = 1;
      }
      break;
    case FIN_BETA:
      sc[S_BETA] = safe_ratio(a, sc[S_ZR]);
      sc[S_ZR] = a;
      break;
    case FIN_LOSS:
      sc[S_LOSS] = a;
      break;
  }
}

// Per-instance deterministic reduction of (a, b) then fin_op; every thread
// of every block of the instance calls it.  Partials are published as
// device-scope atomic stores and read back as device-scope atomic loads (no
// device-wide fence: the L2s of the XCDs are not coherent with each other).
template <int NT>
DFM_HD void inst_reduce(const Ctx& C, const Dom& D, int i, int op, double a, double b) {
  __shared__ int s_last;
  block_sum2<NT>(a, b);
  const int first = D.blk[i], nb = D.blk[i + 1] - first;
  if (threadIdx.x == 0) {
    __hip_atomic_store(&C.part[2 * blockIdx.x], a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&C.part[2 * blockIdx.x + 1], b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    wait_stores();
    unsigned* c = C.ctr + i * kCtrStride;
    const int last = __hip_atomic_fetch_add(c, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == (unsigned)(nb - 1);
    if (last) __hip_atomic_store(c, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_last = last;
  }
  __syncthreads();
  if (!s_last) return;
  double ta = 0, tb = 0;
  for (int t = threadIdx.x; t < nb; t += NT) {
    ta += __hip_atomic_load(&C.part[2 * (first + t)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    tb += __hip_atomic_load(&C.part[2 * (first + t) + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  block_sum2<NT>(ta, tb);
  if (threadIdx.x == 0) fin_op(C, i, op, ta, tb);
}

// ------------------------------------------------------- flow-map stages ----
// Face points [0, F0) x faces, [F0, F0 + F1) y faces, then (ALL domains)
// cells [F0 + F1, F0 + F1 + C).
DFM_HD int point_grid(const Inst& I, int64_t t, int64_t& k) {
  const int64_t F0 = faces0(I), F1 = faces1(I);
  if (t < F0) {
    k = t;
    return 0;
  }
  if (t < F0 + F1) {
    k = t - F0;
    return 1;
  }
  k = t - F0 - F1;
  return 3;
}
DFM_HD int64_t npoints(const Inst& I, bool with_cells) { return faces0(I) + faces1(I) + (with_cells ? cells(I) : 0); }

struct SLB {
  const float* vel;      // MAC pool of the advecting velocity
  float dt_scale;        // dt = dt_scale * instance dt
  int nvec, nscal;
  const float* vsrc[2];  // MAC pools
  float* vdst[2];
  const float* ssrc[2];  // cell pools
  float* sdst[2];
};

__global__ void __launch_bounds__(BS) k_b_sl(Ctx C, Dom D, SLB A) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  if (w.t >= npoints(I, A.nscal > 0)) return;
  int64_t k;
  const int grid = point_grid(I, w.t, k);
  if (grid < 3 && A.nvec == 0) return;
  const Geo g = geo_of(I);
  int sh[3];
  shape_of(g, grid, sh);
  int a, b, c;
  unlin(sh, k, a, b, c);
  float x[3] = {0.f, 0.f, 0.f}, p[3] = {0.f, 0.f, 0.f};
  node_pos(g, grid, a, b, c, x);
  backtrace<2, 2>(g, mac_of(A.vel, I), A.dt_scale * I.dt, x, p);
  float bl[3] = {0.5f, 0.5f, 0.5f};
  if (grid < 3) {
    bl[grid] = 0.0f;
    for (int m = 0; m < A.nvec; ++m)
      mac_c(A.vdst[m], I, grid)[k] = interp<2, false, 2>(mac_of(A.vsrc[m], I).c[grid], sh, bl, p, g.inv_dx, nullptr);
  } else {
    for (int m = 0; m < A.nscal; ++m)
      A.sdst[m][I.oc + k] = interp<2, false, 2>(A.ssrc[m] + I.oc, sh, bl, p, g.inv_dx, nullptr);
  }
}

// RK2 march of every point set of a map (faces with T, cells psi only).
__global__ void __launch_bounds__(BS) k_b_march(Ctx C, Dom D, const float* vel, float* map, float dt_scale) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  if (w.t >= npoints(I, true)) return;
  int64_t k;
  const int grid = point_grid(I, w.t, k);
  const Geo g = geo_of(I);
  const MapV M = map_of(map, I);
  const Mac u = mac_of(vel, I);
  const float dt = dt_scale * I.dt;
  float p[3] = {0.f, 0.f, 0.f}, T[3] = {0.f, 0.f, 0.f};
  if (grid < 3) {
    const int64_t cnt = grid ? faces1(I) : faces0(I);
    float* ps = M.pf[grid];
    float* ts = M.tf[grid];
    p[0] = ps[k];
    p[1] = ps[cnt + k];
    T[0] = ts[k];
    T[1] = ts[cnt + k];
    rk_point<2, 2, true>(g, u, dt, p, T);
    ps[k] = p[0];
    ps[cnt + k] = p[1];
    ts[k] = T[0];
    ts[cnt + k] = T[1];
  } else {
    const int64_t cnt = cells(I);
    p[0] = M.psi[k];
    p[1] = M.psi[cnt + k];
    rk_point<2, 2, false>(g, u, dt, p, T);
    M.psi[k] = p[0];
    M.psi[cnt + k] = p[1];
  }
}

__global__ void __launch_bounds__(BS) k_b_reset(Ctx C, Dom D, float* map) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  if (w.t >= npoints(I, true)) return;
  int64_t k;
  const int grid = point_grid(I, w.t, k);
  const Geo g = geo_of(I);
  const MapV M = map_of(map, I);
  int sh[3];
  shape_of(g, grid, sh);
  int a, b, c;
  unlin(sh, k, a, b, c);
  float x[3] = {0.f, 0.f, 0.f};
  node_pos(g, grid, a, b, c, x);
  if (grid < 3) {
    const int64_t cnt = grid ? faces1(I) : faces0(I);
    M.pf[grid][k] = x[0];
    M.pf[grid][cnt + k] = x[1];
    M.tf[grid][k] = grid == 0 ? 1.f : 0.f;
    M.tf[grid][cnt + k] = grid == 1 ? 1.f : 0.f;
  } else {
    M.psi[k] = x[0];
    M.psi[cells(I) + k] = x[1];
  }
}

// out = alpha T . src(psi) + beta add on the face points (cells: src(psi)).
struct PBB {
  const float* src;
  const float* map;
  float alpha, beta;
  const float* add;
  float* out;
};
__global__ void __launch_bounds__(BS) k_b_pullback_u(Ctx C, Dom D, PBB A) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  if (w.t >= npoints(I, false)) return;
  int64_t k;
  const int grid = point_grid(I, w.t, k);
  const Geo g = geo_of(I);
  const MapV M = map_of(const_cast<float*>(A.map), I);
  const int64_t cnt = grid ? faces1(I) : faces0(I);
  float p[3] = {M.pf[grid][k], M.pf[grid][cnt + k], 0.f}, v[3];
  mac_eval<false, 2>(g, mac_of(A.src, I), p, v, nullptr);
  const float d = fmaf(M.tf[grid][cnt + k], v[1], M.tf[grid][k] * v[0]);
  float r = A.alpha * d;
  if (A.beta != 0.f) r = fmaf(A.beta, mac_of(A.add, I).c[grid][k], r);
  mac_c(A.out, I, grid)[k] = r;
}
__global__ void __launch_bounds__(BS) k_b_pullback_s(Ctx C, Dom D, PBB A) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  const int64_t cnt = cells(I);
  if (w.t >= cnt) return;
  const int64_t k = w.t;
  const Geo g = geo_of(I);
  const MapV M = map_of(const_cast<float*>(A.map), I);
  const float p[3] = {M.psi[k], M.psi[cnt + k], 0.f};
  int S[3];
  shape_of(g, 3, S);
  const float bl[3] = {0.5f, 0.5f, 0.5f};
  float r = A.alpha * interp<2, false, 2>(A.src + I.oc, S, bl, p, g.inv_dx, nullptr);
  if (A.beta != 0.f) r = fmaf(A.beta, A.add[I.oc + k], r);
  A.out[I.oc + k] = r;
}

// accumulate_init_new (AB3) of f through the map, plus f2 at dt2.
struct AccB {
  const float* f;
  const float* f2;
  const float* map;
  float* init;
  float* h1;
  float* h2;
  int ind;
  float dt_scale, dt2_scale;
};
__global__ void __launch_bounds__(BS) k_b_accumulate(Ctx C, Dom D, AccB A) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  if (w.t >= npoints(I, false)) return;
  int64_t k;
  const int grid = point_grid(I, w.t, k);
  const Geo g = geo_of(I);
  const MapV M = map_of(const_cast<float*>(A.map), I);
  const int64_t cnt = grid ? faces1(I) : faces0(I);
  const float p[3] = {M.pf[grid][k], M.pf[grid][cnt + k], 0.f};
  const float T0 = M.tf[grid][k], T1 = M.tf[grid][cnt + k];
  float v[3];
  mac_eval<false, 2>(g, mac_of(A.f, I), p, v, nullptr);
  const float gv = fmaf(T1, v[1], T0 * v[0]);
  float* init = mac_c(A.init, I, grid) + k;
  float* h1 = mac_c(A.h1, I, grid) + k;
  float* h2 = mac_c(A.h2, I, grid) + k;
  const float dt = A.dt_scale * I.dt, dt2 = A.dt2_scale * I.dt;
  float acc = *init;
  const float hp = *h1;
  if (A.ind == 0) acc += gv * dt;
  else if (A.ind == 1) acc += (23.0f * gv - 16.0f * gv + 5.0f * hp) / 12.0f * dt;
  else acc += (23.0f * gv - 16.0f * hp + 5.0f * *h2) / 12.0f * dt;
  mac_eval<false, 2>(g, mac_of(A.f2, I), p, v, nullptr);
  acc += fmaf(T1, v[1], T0 * v[0]) * dt2;
  *init = acc;
  *h2 = hp;
  *h1 = gv;
}

// Both adjoint sources at the face points: r_u = ((grad u)^T lam_u) with the
// quartic spline, r_s = grad s . lam_s with the cubic one.
__global__ void __launch_bounds__(BS) k_b_sources(Ctx C, Dom D, const float* u, const float* lam, const float* s,
                                                  const float* lam_s, float* r_u, float* r_s) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  if (w.t >= npoints(I, false)) return;
  int64_t k;
  const int grid = point_grid(I, w.t, k);
  const Geo g = geo_of(I);
  int sh[3];
  shape_of(g, grid, sh);
  int a, b, c;
  unlin(sh, k, a, b, c);
  float x[3] = {0.f, 0.f, 0.f};
  node_pos(g, grid, a, b, c, x);
  const Mac U = mac_of(u, I), W = mac_of(lam, I);
  float r = 0.f;
#pragma unroll
  for (int cc = 0; cc < 2; ++cc) {
    int Sc[3];
    shape_of(g, cc, Sc);
    float bl[3] = {0.5f, 0.5f, 0.5f};
    bl[cc] = 0.0f;
    float d, v;
    dual_interp<4, 2>(U.c[cc], W.c[cc], Sc, bl, x, g.inv_dx, grid, d, v);
    r = fmaf(d, v, r);
  }
  mac_c(r_u, I, grid)[k] = r;
  int Sc[3];
  shape_of(g, 3, Sc);
  const float bl[3] = {0.5f, 0.5f, 0.5f};
  float d, v;
  dual_interp<3, 2>(s + I.oc, lam_s + I.oc, Sc, bl, x, g.inv_dx, grid, d, v);
  mac_c(r_s, I, grid)[k] = d * v;
}

// out = x + (alpha dt) y + (beta dt) z on MAC pools (z may be null).
__global__ void __launch_bounds__(BS) k_b_axpy(Ctx C, Dom D, const float* x, const float* y, const float* z, float* out,
                                               float alpha_dt, float beta_dt) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  if (w.t >= npoints(I, false)) return;
  const int64_t o = I.of + w.t;
  float r = fmaf(alpha_dt * I.dt, y[o], x[o]);
  if (z) r = fmaf(beta_dt * I.dt, z[o], r);
  out[o] = r;
}

// Copies: up to 3 MAC pools (face points) and 3 cell pools in one launch.
struct CopyB {
  int nmac, ncell;
  const float* msrc[3];
  float* mdst[3];
  const float* csrc[3];
  float* cdst[3];
  float czero;  // cells: dst = src * (1 - czero) (czero = 1 zeroes)
};
__global__ void __launch_bounds__(BS) k_b_copy(Ctx C, Dom D, CopyB A) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  const int64_t F = faces0(I) + faces1(I);
  if (w.t >= npoints(I, A.ncell > 0)) return;
  if (w.t < F) {
    for (int m = 0; m < A.nmac; ++m) A.mdst[m][I.of + w.t] = A.msrc[m] ? A.msrc[m][I.of + w.t] : 0.f;
  } else {
    const int64_t k = w.t - F;
    for (int m = 0; m < A.ncell; ++m) A.cdst[m][I.oc + k] = A.csrc[m] ? A.csrc[m][I.oc + k] : 0.f;
  }
}

// Terminal adjoint: lam_s = 2 (s - target), loss = sum (s - target)^2 (FP64).
__global__ void __launch_bounds__(BS) k_b_terminal(Ctx C, Dom D, const float* s, const float* tgt, float* lam_s) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  double l = 0;
  if (w.t < cells(I)) {
    const int64_t o = I.oc + w.t;
    const float d = s[o] - tgt[o];
    lam_s[o] = 2.0f * d;
    l = (double)d * (double)d;
  }
  inst_reduce<BS>(C, D, w.i, FIN_LOSS, l, 0.0);
}

// ------------------------------------------------------------------ MGPCG ----
// Level-l arrays of an instance: natural row-major (i n1 + j), n >> l.
struct MG {
  float* z;  // level pools (levels 0..T): z, r
  float* r;
  float* p;
  float* Ap;
  double* x;
};

// Neumann 5-point operator pieces of cell (i, j) on an a x b grid.
DFM_HD int nnbr(int i, int j, int a, int b) { return (i > 0) + (i < a - 1) + (j > 0) + (j < b - 1); }
DFM_HD float nsum(const float* f, int i, int j, int a, int b) {
  const int64_t id = (int64_t)i * b + j;
  float s = 0.f;
  if (i > 0) s += f[id - b];
  if (i < a - 1) s += f[id + b];
  if (j > 0) s += f[id - 1];
  if (j < b - 1) s += f[id + 1];
  return s;
}
DFM_HD float inv_nn(int nn) { return nn == 4 ? 0.25f : (nn == 3 ? (1.0f / 3.0f) : (nn == 2 ? 0.5f : (nn == 1 ? 1.f : 0.f))); }

// b = -div(u) (wall faces read as 0), partial sum -> mean; level-0 r = b.
__global__ void __launch_bounds__(BS) k_b_div(Ctx C, Dom D, const float* u, MG M) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  double s = 0;
  if (w.t < cells(I)) {
    const int i = (int)(w.t / I.n1), j = (int)(w.t % I.n1);
    const Mac U = mac_of(u, I);
    const int64_t fx = (int64_t)i * I.n1 + j, fy = (int64_t)i * (I.n1 + 1) + j;
    const float uxl = i > 0 ? U.c[0][fx] : 0.f, uxh = i < I.n0 - 1 ? U.c[0][fx + I.n1] : 0.f;
    const float uyl = j > 0 ? U.c[1][fy] : 0.f, uyh = j < I.n1 - 1 ? U.c[1][fy + 1] : 0.f;
    const float b = -((uxh - uxl) + (uyh - uyl));
    M.r[I.om[0] + w.t] = b;
    s = b;
  }
  inst_reduce<BS>(C, D, w.i, FIN_MEAN, s, 0.0);
}

// r -= mean, x = 0, partial r.r -> |b|^2, threshold, flags reset.
__global__ void __launch_bounds__(BS) k_b_recenter(Ctx C, Dom D, MG M) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  double rr = 0;
  if (w.t < cells(I)) {
    const float m = (float)C.sc[(int64_t)w.i * S_COUNT + S_MEAN];
    float* r = M.r + I.om[0] + w.t;
    const float v = *r - m;
    *r = v;
    M.x[I.oc + w.t] = 0.0;
    rr = (double)v * v;
  }
  inst_reduce<BS>(C, D, w.i, FIN_BNORM, rr, 0.0);
}

// The level-l residual right-hand side of a cell: level 0 carries the
// deferred mean of the PCG residual.
DFM_HD float rhs(const MG& M, const Inst& I, const Ctx& C, int inst, int l, int64_t id) {
  const float v = M.r[I.om[l] + id];
  return l == 0 ? v - (float)C.sc[(int64_t)inst * S_COUNT + S_MEAN] : v;
}

// First pre-smoothing sweep from z = 0, both colours: red z = r/nn, black z
// = (r + sum of the red neighbours' r/nn)/nn.
__global__ void __launch_bounds__(BS) k_b_pre1(Ctx C, Dom D, MG M, int l) {
  const Where w = where(D);
  if (inst_done(C, w.i)) return;
  const Inst I = C.I[w.i];
  const int a = n_of(I, l, 0), b = n_of(I, l, 1);
  if (w.t >= (int64_t)a * b) return;
  const int i = (int)(w.t / b), j = (int)(w.t % b);
  const float m = l == 0 ? (float)C.sc[(int64_t)w.i * S_COUNT + S_MEAN] : 0.f;
  const float* r = M.r + I.om[l];
  float z;
  if (((i + j) & 1) == 0) {
    z = (r[w.t] - m) * inv_nn(nnbr(i, j, a, b));
  } else {
    float s = 0.f;
    if (i > 0) s += (r[w.t - b] - m) * inv_nn(nnbr(i - 1, j, a, b));
    if (i < a - 1) s += (r[w.t + b] - m) * inv_nn(nnbr(i + 1, j, a, b));
    if (j > 0) s += (r[w.t - 1] - m) * inv_nn(nnbr(i, j - 1, a, b));
    if (j < b - 1) s += (r[w.t + 1] - m) * inv_nn(nnbr(i, j + 1, a, b));
    z = (r[w.t] - m + s) * inv_nn(nnbr(i, j, a, b));
  }
  M.z[I.om[l] + w.t] = z;
}

// Half-sweep of colour `phase` on level l (one thread per cell of that
// colour; rows have an even length, so each row holds b/2 of them).  With
// prolong, the neighbours (of the other colour) are read with the coarse
// correction z_{l+1}[I/2] added (the prolongation of those cells is never
// written: they are overwritten by the next half-sweep).
__global__ void __launch_bounds__(BS) k_b_half(Ctx C, Dom D, MG M, int l, int phase, int prolong) {
  const Where w = where(D);
  if (inst_done(C, w.i)) return;
  const Inst I = C.I[w.i];
  const int a = n_of(I, l, 0), b = n_of(I, l, 1), hb = b >> 1;
  if (w.t >= (int64_t)a * hb) return;
  const int i = (int)(w.t / hb);
  const int j = 2 * (int)(w.t % hb) + ((i + phase) & 1);
  const int64_t id = (int64_t)i * b + j;
  float* z = M.z + I.om[l];
  float s;
  if (!prolong) {
    s = nsum(z, i, j, a, b);
  } else {
    const float* zc = M.z + I.om[l + 1];
    const int bc = b >> 1;
    s = 0.f;
    if (i > 0) s += z[id - b] + zc[(int64_t)((i - 1) >> 1) * bc + (j >> 1)];
    if (i < a - 1) s += z[id + b] + zc[(int64_t)((i + 1) >> 1) * bc + (j >> 1)];
    if (j > 0) s += z[id - 1] + zc[(int64_t)(i >> 1) * bc + ((j - 1) >> 1)];
    if (j < b - 1) s += z[id + 1] + zc[(int64_t)(i >> 1) * bc + ((j + 1) >> 1)];
  }
  z[id] = (rhs(M, I, C, w.i, l, id) + s) * inv_nn(nnbr(i, j, a, b));
}

// r_{l+1}[I] = sum over the 4 children of (r - A z) (the 2D restriction
// scale 1/(dim - 1) is 1); one thread per coarse cell.
__global__ void __launch_bounds__(BS) k_b_restrict(Ctx C, Dom D, MG M, int l) {
  const Where w = where(D);
  if (inst_done(C, w.i)) return;
  const Inst I = C.I[w.i];
  const int a = n_of(I, l, 0), b = n_of(I, l, 1), bc = b >> 1;
  if (w.t >= (int64_t)(a >> 1) * bc) return;
  const int ic = (int)(w.t / bc), jc = (int)(w.t % bc);
  const float* z = M.z + I.om[l];
  float s = 0.f;
#pragma unroll
  for (int di = 0; di < 2; ++di)
#pragma unroll
    for (int dj = 0; dj < 2; ++dj) {
      const int i = 2 * ic + di, j = 2 * jc + dj;
      const int64_t id = (int64_t)i * b + j;
      s += rhs(M, I, C, w.i, l, id) - ((float)nnbr(i, j, a, b) * z[id] - nsum(z, i, j, a, b));
    }
  M.r[I.om[l + 1] + w.t] = s;
}

// One workgroup per instance runs levels T..L-1 in LDS: from r_T (global)
// the V-cycle's remaining legs (2 pre sweeps, restriction, `bottom` sweeps
// on the coarsest level, prolongation, 2 post sweeps), z_T written back.
__device__ void lds_half(float* z, const float* r, int a, int b, int phase, float m) {
  const int hb = b >> 1;
  for (int t = threadIdx.x; t < a * hb; t += kTailThreads) {
    const int i = t / hb, j = 2 * (t % hb) + ((i + phase) & 1);
    const int id = i * b + j;
    z[id] = (r[id] - m + nsum(z, i, j, a, b)) * inv_nn(nnbr(i, j, a, b));
  }
}
__global__ void __launch_bounds__(kTailThreads) k_b_tail(Ctx C, MG M, int bottom) {
  extern __shared__ float lds[];
  const int inst = blockIdx.x;
  if (inst_done(C, inst)) return;
  const Inst I = C.I[inst];
  const int T = I.T, L = I.L;
  // LDS: z_T..z_{L-1}, then r_{T+1}..r_{L-1}; r_T stays in global memory
  float* zl[kMaxLv];
  float* rl[kMaxLv];
  int off = 0;
  for (int l = T; l < L; ++l) {
    zl[l] = lds + off;
    off += n_of(I, l, 0) * n_of(I, l, 1);
  }
  for (int l = T + 1; l < L; ++l) {
    rl[l] = lds + off;
    off += n_of(I, l, 0) * n_of(I, l, 1);
  }
  const float* rT = M.r + I.om[T];
  const float mT = T == 0 ? (float)C.sc[(int64_t)inst * S_COUNT + S_MEAN] : 0.f;
  for (int l = T; l < L; ++l) {
    const int a = n_of(I, l, 0), b = n_of(I, l, 1);
    const float* r = l == T ? rT : rl[l];
    const float m = l == T ? mT : 0.f;
    for (int t = threadIdx.x; t < a * b; t += kTailThreads) zl[l][t] = 0.f;
    __syncthreads();
    const int sweeps = l == L - 1 ? bottom : 2;
    for (int s = 0; s < sweeps; ++s) {
      lds_half(zl[l], r, a, b, 0, m);
      __syncthreads();
      lds_half(zl[l], r, a, b, 1, m);
      __syncthreads();
    }
    if (l == L - 1) break;
    const int bc = b >> 1;
    for (int t = threadIdx.x; t < (a >> 1) * bc; t += kTailThreads) {
      const int ic = t / bc, jc = t % bc;
      float s = 0.f;
      for (int di = 0; di < 2; ++di)
        for (int dj = 0; dj < 2; ++dj) {
          const int i = 2 * ic + di, j = 2 * jc + dj, id = i * b + j;
          s += (r[id] - m) - ((float)nnbr(i, j, a, b) * zl[l][id] - nsum(zl[l], i, j, a, b));
        }
      rl[l + 1][t] = s;
    }
    __syncthreads();
  }
  for (int l = L - 2; l >= T; --l) {
    const int a = n_of(I, l, 0), b = n_of(I, l, 1), bc = b >> 1;
    const float* r = l == T ? rT : rl[l];
    const float m = l == T ? mT : 0.f;
    for (int t = threadIdx.x; t < a * b; t += kTailThreads) {
      const int i = t / b, j = t % b;
      zl[l][t] += zl[l + 1][(i >> 1) * bc + (j >> 1)];
    }
    __syncthreads();
    for (int s = 0; s < 2; ++s) {
      lds_half(zl[l], r, a, b, 1, m);
      __syncthreads();
      lds_half(zl[l], r, a, b, 0, m);
      __syncthreads();
    }
  }
  const int nT = n_of(I, T, 0) * n_of(I, T, 1);
  for (int t = threadIdx.x; t < nT; t += kTailThreads) M.z[I.om[T] + t] = zl[T][t];
}

// z.r (r with its deferred mean) -> FIN_ZR_INIT (p = z) or FIN_BETA.
__global__ void __launch_bounds__(BS) k_b_zr(Ctx C, Dom D, MG M, int init) {
  const Where w = where(D);
  if (inst_done(C, w.i)) return;
  const Inst I = C.I[w.i];
  double zr = 0;
  if (w.t < cells(I)) {
    const float z = M.z[I.om[0] + w.t];
    zr = (double)z * (double)rhs(M, I, C, w.i, 0, w.t);
    if (init) M.p[I.oc + w.t] = z;
  }
  inst_reduce<BS>(C, D, w.i, init ? FIN_ZR_INIT : FIN_BETA, zr, 0.0);
}

// p = z + beta p
__global__ void __launch_bounds__(BS) k_b_p(Ctx C, Dom D, MG M) {
  const Where w = where(D);
  if (inst_done(C, w.i)) return;
  const Inst I = C.I[w.i];
  if (w.t >= cells(I)) return;
  const float beta = (float)C.sc[(int64_t)w.i * S_COUNT + S_BETA];
  float* p = M.p + I.oc + w.t;
  *p = fmaf(beta, *p, M.z[I.om[0] + w.t]);
}

// Ap = A p, partial p.Ap -> alpha.
__global__ void __launch_bounds__(BS) k_b_Ap(Ctx C, Dom D, MG M) {
  const Where w = where(D);
  if (inst_done(C, w.i)) return;
  const Inst I = C.I[w.i];
  double pap = 0;
  if (w.t < cells(I)) {
    const int a = I.n0, b = I.n1;
    const int i = (int)(w.t / b), j = (int)(w.t % b);
    const float* p = M.p + I.oc;
    const float ap = (float)nnbr(i, j, a, b) * p[w.t] - nsum(p, i, j, a, b);
    M.Ap[I.oc + w.t] = ap;
    pap = (double)p[w.t] * ap;
  }
  inst_reduce<BS>(C, D, w.i, FIN_ALPHA, pap, 0.0);
}

// x += alpha p, r = (r - mean) - alpha Ap; partial r.r and sum r.
__global__ void __launch_bounds__(BS) k_b_update(Ctx C, Dom D, MG M) {
  const Where w = where(D);
  if (inst_done(C, w.i)) return;
  const Inst I = C.I[w.i];
  double rr = 0, rs = 0;
  if (w.t < cells(I)) {
    const double alpha = C.sc[(int64_t)w.i * S_COUNT + S_ALPHA];
    const int64_t o = I.oc + w.t;
    M.x[o] += alpha * (double)M.p[o];
    float* r = M.r + I.om[0] + w.t;
    const float v = fmaf(-(float)alpha, M.Ap[o], rhs(M, I, C, w.i, 0, w.t));
    *r = v;
    rr = (double)v * v;
    rs = v;
  }
  inst_reduce<BS>(C, D, w.i, FIN_RR, rr, rs);
}

// u_a -= x[I] - x[I - e_a] on interior faces, wall faces 0 (one thread per face).
__global__ void __launch_bounds__(BS) k_b_subgrad(Ctx C, Dom D, float* u, MG M) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  if (w.t >= npoints(I, false)) return;
  int64_t k;
  const int grid = point_grid(I, w.t, k);
  const double* x = M.x + I.oc;
  float* c = mac_c(u, I, grid) + k;
  if (grid == 0) {
    const int i = (int)(k / I.n1), j = (int)(k % I.n1);
    *c = (i > 0 && i < I.n0) ? (float)((double)*c - (x[(int64_t)i * I.n1 + j] - x[(int64_t)(i - 1) * I.n1 + j])) : 0.f;
  } else {
    const int i = (int)(k / (I.n1 + 1)), j = (int)(k % (I.n1 + 1));
    *c = (j > 0 && j < I.n1) ? (float)((double)*c - (x[(int64_t)i * I.n1 + j] - x[(int64_t)i * I.n1 + j - 1])) : 0.f;
  }
}

// ----------------------------------------------------------------- engine ----
enum DomKind { D_FACES = 0, D_CELLS, D_ALL, D_TAILGRID, D_LV_CELLS, D_LV_HALF = D_LV_CELLS + kMaxLv,
               D_LV_COARSE = D_LV_HALF + kMaxLv, D_COUNT = D_LV_COARSE + kMaxLv };

struct Batch {
  int count = 0, R = 0, steps = 0, bottom = 50, chunk = 4, first_iters = 12, graphs = 1, tail_cells = 16384;
  double tol = 1e-6;
  int max_iters = 400;
  std::vector<Inst> hI;
  Inst* dI = nullptr;
  std::vector<std::vector<int>> dom_h;  // per kind: count+1 block prefix
  int* dom_d = nullptr;                 // all tables, D_COUNT x (count + 1)
  int64_t C = 0, F = 0, MGcells = 0;    // pool sizes
  int max_blocks = 0;
  size_t tail_lds = 0;
  // pools
  float *u = nullptr, *u_new = nullptr, *init_u = nullptr, *err = nullptr, *mid = nullptr;
  float *adj = nullptr, *adj_new = nullptr, *init_adj = nullptr, *r_u = nullptr, *r_s = nullptr, *h1 = nullptr,
        *h2 = nullptr;
  float *s = nullptr, *s_new = nullptr, *init_s = nullptr, *err_s = nullptr, *tgt = nullptr, *s_start = nullptr;
  float *ls = nullptr, *ls_new = nullptr, *init_ls = nullptr;
  float *phi = nullptr, *psi = nullptr, *phis = nullptr;  // maps; phis: one per interval
  float *mz = nullptr, *mr = nullptr, *mp = nullptr, *mAp = nullptr;
  double* mx = nullptr;
  double* sc = nullptr;
  int* fl = nullptr;
  double* part = nullptr;
  unsigned* ctr = nullptr;
  int* h_fl = nullptr;      // pinned mirror of the flags
  double* h_sc = nullptr;   // pinned mirror of the scalars
  hipStream_t cap = nullptr;
  std::map<std::tuple<int, int, int>, hipGraphExec_t> graphs_c;
  int step = -1;            // forward: last completed step; adjoint: next step to undo + 1
  int projections = 0, max_seen_iters = 0, noconv = 0;
  std::vector<void*> allocs;

  Ctx ctx() const { return Ctx{dI, sc, fl, part, ctr, tol}; }
  Dom dom(int kind) const { return Dom{dom_d + (size_t)kind * (count + 1), count}; }
  int blocks(int kind) const { return dom_h[kind][count]; }
  MG mg() const { return MG{mz, mr, mp, mAp, mx}; }
  int64_t map_size() const { return 2 * C + 4 * F; }
};

template <class K, class... A>
void launch(hipStream_t st, K kern, int nblocks, A... args) {
  if (nblocks > 0) hipLaunchKernelGGL(kern, dim3(nblocks), dim3(BS), 0, st, args...);
}

int levels_of(int n0, int n1) {  // int(log2(min(N))) - base_level + 1, base_level 3 (2D/vortex/mgpcg.py:28)
  const int m = std::min(n0, n1);
  int lg = 0;
  while ((2 << lg) <= m) ++lg;
  return std::max(1, lg - 3 + 1);
}

int build_domains(Batch& B) {
  B.dom_h.assign(D_COUNT, std::vector<int>(B.count + 1, 0));
  auto add = [&](int kind, int i, int64_t pts) {
    B.dom_h[kind][i + 1] = B.dom_h[kind][i] + (int)((pts + BS - 1) / BS);
  };
  for (int i = 0; i < B.count; ++i) {
    const Inst& I = B.hI[i];
    add(D_FACES, i, faces0(I) + faces1(I));
    add(D_CELLS, i, cells(I));
    add(D_ALL, i, faces0(I) + faces1(I) + cells(I));
    B.dom_h[D_TAILGRID][i + 1] = i + 1;
    for (int l = 0; l < kMaxLv; ++l) {
      const bool global = l < I.T;  // levels 0..T-1 run as batched kernels
      const int64_t nl = (int64_t)n_of(I, l, 0) * n_of(I, l, 1);
      add(D_LV_CELLS + l, i, global ? nl : 0);
      add(D_LV_HALF + l, i, global ? nl / 2 : 0);
      add(D_LV_COARSE + l, i, global ? nl / 4 : 0);
    }
  }
  B.max_blocks = 0;
  for (auto& v : B.dom_h) B.max_blocks = std::max(B.max_blocks, v[B.count]);
  return DFM_OK;
}

template <class T>
int dalloc(Batch& B, T** p, size_t n) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) return set_error(DFM_E_HIP, "dfm_batch_create: hipMalloc", hipGetErrorString(e));
  B.allocs.push_back(q);
  *p = static_cast<T*>(q);
  return DFM_OK;
}

// ----------------------------------------------------------------- solves ----
void enq_pcg_iters(Batch& B, hipStream_t st, int iters) {
  const Ctx c = B.ctx();
  const MG M = B.mg();
  for (int it = 0; it < iters; ++it) {
    launch(st, k_b_Ap, B.blocks(D_CELLS), c, B.dom(D_CELLS), M);
    launch(st, k_b_update, B.blocks(D_CELLS), c, B.dom(D_CELLS), M);
    // V-cycle on the recentred residual
    int G = 0;
    for (auto& I : B.hI) G = std::max(G, I.T);
    for (int l = 0; l < G; ++l) {
      launch(st, k_b_pre1, B.blocks(D_LV_CELLS + l), c, B.dom(D_LV_CELLS + l), M, l);
      launch(st, k_b_half, B.blocks(D_LV_HALF + l), c, B.dom(D_LV_HALF + l), M, l, 0, 0);
      launch(st, k_b_half, B.blocks(D_LV_HALF + l), c, B.dom(D_LV_HALF + l), M, l, 1, 0);
      launch(st, k_b_restrict, B.blocks(D_LV_COARSE + l), c, B.dom(D_LV_COARSE + l), M, l);
    }
    hipLaunchKernelGGL(k_b_tail, dim3(B.count), dim3(kTailThreads), B.tail_lds, st, c, M, B.bottom);
    for (int l = G - 1; l >= 0; --l) {
      launch(st, k_b_half, B.blocks(D_LV_HALF + l), c, B.dom(D_LV_HALF + l), M, l, 1, 1);
      launch(st, k_b_half, B.blocks(D_LV_HALF + l), c, B.dom(D_LV_HALF + l), M, l, 0, 0);
      launch(st, k_b_half, B.blocks(D_LV_HALF + l), c, B.dom(D_LV_HALF + l), M, l, 1, 0);
      launch(st, k_b_half, B.blocks(D_LV_HALF + l), c, B.dom(D_LV_HALF + l), M, l, 0, 0);
    }
    launch(st, k_b_zr, B.blocks(D_CELLS), c, B.dom(D_CELLS), M, 0);
    launch(st, k_b_p, B.blocks(D_CELLS), c, B.dom(D_CELLS), M);
  }
}

void enq_vcycle_init(Batch& B, hipStream_t st) {
  const Ctx c = B.ctx();
  const MG M = B.mg();
  int G = 0;
  for (auto& I : B.hI) G = std::max(G, I.T);
  for (int l = 0; l < G; ++l) {
    launch(st, k_b_pre1, B.blocks(D_LV_CELLS + l), c, B.dom(D_LV_CELLS + l), M, l);
    launch(st, k_b_half, B.blocks(D_LV_HALF + l), c, B.dom(D_LV_HALF + l), M, l, 0, 0);
    launch(st, k_b_half, B.blocks(D_LV_HALF + l), c, B.dom(D_LV_HALF + l), M, l, 1, 0);
    launch(st, k_b_restrict, B.blocks(D_LV_COARSE + l), c, B.dom(D_LV_COARSE + l), M, l);
  }
  hipLaunchKernelGGL(k_b_tail, dim3(B.count), dim3(kTailThreads), B.tail_lds, st, c, M, B.bottom);
  for (int l = G - 1; l >= 0; --l) {
    launch(st, k_b_half, B.blocks(D_LV_HALF + l), c, B.dom(D_LV_HALF + l), M, l, 1, 1);
    launch(st, k_b_half, B.blocks(D_LV_HALF + l), c, B.dom(D_LV_HALF + l), M, l, 0, 0);
    launch(st, k_b_half, B.blocks(D_LV_HALF + l), c, B.dom(D_LV_HALF + l), M, l, 1, 0);
    launch(st, k_b_half, B.blocks(D_LV_HALF + l), c, B.dom(D_LV_HALF + l), M, l, 0, 0);
  }
}

// Projection start: b = -div u, recentre, first V-cycle, z.r, p = z, then
// `iters` PCG iterations.
void enq_pcg_begin(Batch& B, hipStream_t st, const float* u, int iters) {
  const Ctx c = B.ctx();
  const MG M = B.mg();
  launch(st, k_b_div, B.blocks(D_CELLS), c, B.dom(D_CELLS), u, M);
  launch(st, k_b_recenter, B.blocks(D_CELLS), c, B.dom(D_CELLS), M);
  enq_vcycle_init(B, st);
  launch(st, k_b_zr, B.blocks(D_CELLS), c, B.dom(D_CELLS), M, 1);
  enq_pcg_iters(B, st, iters);
}

// Runs enqueue() on st: captured once per key into a graph and replayed.
template <class F>
int run_seg(Batch& B, hipStream_t st, std::tuple<int, int, int> key, F&& enqueue) {
  if (!B.graphs) {
    enqueue(st);
    return DFM_OK;
  }
  auto it = B.graphs_c.find(key);
  if (it == B.graphs_c.end()) {
    // captured on the engine's own stream (the caller's may be the null
    // stream, which cannot be captured); nothing runs during the capture,
    // and the graph is then launched on the caller's stream, in its order
    hipGraph_t g = nullptr;
    hipGraphExec_t ex = nullptr;
    hipError_t e = hipStreamBeginCapture(B.cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return set_error(DFM_E_HIP, "dfm_batch: hipStreamBeginCapture", hipGetErrorString(e));
    enqueue(B.cap);
    e = hipStreamEndCapture(B.cap, &g);
    if (e != hipSuccess) return set_error(DFM_E_HIP, "dfm_batch: hipStreamEndCapture", hipGetErrorString(e));
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return set_error(DFM_E_HIP, "dfm_batch: hipGraphInstantiate", hipGetErrorString(e));
    it = B.graphs_c.emplace(key, ex).first;
  }
  hipError_t e = hipGraphLaunch(it->second, st);
  if (e != hipSuccess) return set_error(DFM_E_HIP, "dfm_batch: hipGraphLaunch", hipGetErrorString(e));
  return DFM_OK;
}

// After a segment that ended inside a projection: read the flags back and
// run PCG chunks until every instance is done.
int finish_pcg(Batch& B, hipStream_t st) {
  for (;;) {
    hipError_t e = hipMemcpyAsync(B.h_fl, B.fl, sizeof(int) * F_COUNT * B.count, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return set_error(DFM_E_HIP, "dfm_batch: convergence check", hipGetErrorString(e));
    bool all = true;
    for (int i = 0; i < B.count; ++i) all = all && B.h_fl[i * F_COUNT + F_DONE];
    if (all) break;
    int rc = run_seg(B, st, std::make_tuple(9, 0, 0), [&](hipStream_t s) { enq_pcg_iters(B, s, B.chunk); });
    if (rc) return rc;
  }
  B.projections += 1;
  for (int i = 0; i < B.count; ++i) {
    B.max_seen_iters = std::max(B.max_seen_iters, B.h_fl[i * F_COUNT + F_ITERS]);
    if (B.h_fl[i * F_COUNT + F_STATUS] != 1) B.noconv += 1;
  }
  return DFM_OK;
}

// ------------------------------------------------------------ step pieces ----
void enq_copy(Batch& B, hipStream_t st, std::vector<std::pair<const float*, float*>> macs,
              std::vector<std::pair<const float*, float*>> cellsv) {
  CopyB A{};
  A.nmac = (int)macs.size();
  A.ncell = (int)cellsv.size();
  for (int m = 0; m < A.nmac; ++m) {
    A.msrc[m] = macs[m].first;
    A.mdst[m] = macs[m].second;
  }
  for (int m = 0; m < A.ncell; ++m) {
    A.csrc[m] = cellsv[m].first;
    A.cdst[m] = cellsv[m].second;
  }
  const int kind = A.ncell ? D_ALL : D_FACES;
  launch(st, k_b_copy, B.blocks(kind), B.ctx(), B.dom(kind), A);
}

float* mid_at(Batch& B, int step) { return B.mid + (size_t)step * B.F; }
float* phi_at(Batch& B, int m) { return B.phis + (size_t)m * B.map_size(); }
float* sstart_at(Batch& B, int m) { return B.s_start + (size_t)m * B.C; }

void enq_bfecc_u(Batch& B, hipStream_t st, const float* init, const float* fwd, const float* bwd, float* err,
                 float* out) {
  const Ctx c = B.ctx();
  launch(st, k_b_pullback_u, B.blocks(D_FACES), c, B.dom(D_FACES), PBB{init, fwd, 1.f, 0.f, nullptr, out});
  launch(st, k_b_pullback_u, B.blocks(D_FACES), c, B.dom(D_FACES), PBB{out, bwd, 0.5f, -0.5f, init, err});
  launch(st, k_b_pullback_u, B.blocks(D_FACES), c, B.dom(D_FACES), PBB{err, fwd, -1.f, 1.f, out, out});
}
void enq_bfecc_s(Batch& B, hipStream_t st, const float* init, const float* fwd, const float* bwd, float* err,
                 float* out) {
  const Ctx c = B.ctx();
  launch(st, k_b_pullback_s, B.blocks(D_CELLS), c, B.dom(D_CELLS), PBB{init, fwd, 1.f, 0.f, nullptr, out});
  launch(st, k_b_pullback_s, B.blocks(D_CELLS), c, B.dom(D_CELLS), PBB{out, bwd, 0.5f, -0.5f, init, err});
  launch(st, k_b_pullback_s, B.blocks(D_CELLS), c, B.dom(D_CELLS), PBB{err, fwd, -1.f, 1.f, out, out});
}

// Forward step (step_flow_map); u, s hold the step's start and end values.
int forward_step(Batch& B, hipStream_t st) {
  const int step = ++B.step;
  const int j = step % B.R;
  const Ctx c = B.ctx();
  float* mid = mid_at(B, step);
  int rc;
  // seg 0: interval checkpoint, half-step self-advection, projection of mid
  rc = run_seg(B, st, std::make_tuple(0, step, 0), [&](hipStream_t s) {
    if (j == 0) enq_copy(B, s, {}, {{B.s, sstart_at(B, step / B.R)}});
    SLB A{};
    A.vel = B.u;
    A.dt_scale = 0.5f;
    A.nvec = 1;
    A.vsrc[0] = B.u;
    A.vdst[0] = mid;
    launch(s, k_b_sl, B.blocks(D_FACES), c, B.dom(D_FACES), A);
    enq_pcg_begin(B, s, mid, B.first_iters);
  });
  if (rc || (rc = finish_pcg(B, st))) return rc;
  // seg 1: finish mid, march phi, advect u and s by mid, project u
  rc = run_seg(B, st, std::make_tuple(0, step, 1), [&](hipStream_t s) {
    launch(s, k_b_subgrad, B.blocks(D_FACES), c, B.dom(D_FACES), mid, B.mg());
    launch(s, k_b_march, B.blocks(D_ALL), c, B.dom(D_ALL), (const float*)mid, B.phi, -1.f);
    SLB A{};
    A.vel = mid;
    A.dt_scale = 1.f;
    A.nvec = 1;
    A.vsrc[0] = B.u;
    A.vdst[0] = B.u_new;
    A.nscal = 1;
    A.ssrc[0] = B.s;
    A.sdst[0] = B.s_new;
    launch(s, k_b_sl, B.blocks(D_ALL), c, B.dom(D_ALL), A);
    enq_copy(B, s, {}, {{B.s_new, B.s}});
    enq_pcg_begin(B, s, B.u_new, B.first_iters);
  });
  if (rc || (rc = finish_pcg(B, st))) return rc;
  if (j != B.R - 1) {
    return run_seg(B, st, std::make_tuple(0, step, 2), [&](hipStream_t s) {
      launch(s, k_b_subgrad, B.blocks(D_FACES), c, B.dom(D_FACES), B.u_new, B.mg());
      enq_copy(B, s, {{B.u_new, B.u}}, {});
    });
  }
  // reinitialisation: psi re-marched through the end-of-step u and the
  // interval's midpoint velocities, BFECC of init_u, projection
  rc = run_seg(B, st, std::make_tuple(0, step, 2), [&](hipStream_t s) {
    launch(s, k_b_subgrad, B.blocks(D_FACES), c, B.dom(D_FACES), B.u_new, B.mg());
    enq_copy(B, s, {{B.u_new, B.u}}, {});
    launch(s, k_b_reset, B.blocks(D_ALL), c, B.dom(D_ALL), B.psi);
    launch(s, k_b_march, B.blocks(D_ALL), c, B.dom(D_ALL), (const float*)B.u, B.psi, 1.f);
    for (int t = step - 1; t >= step - j; --t)
      launch(s, k_b_march, B.blocks(D_ALL), c, B.dom(D_ALL), (const float*)mid_at(B, t), B.psi, 1.f);
    enq_bfecc_u(B, s, B.init_u, B.psi, B.phi, B.err, B.u);
    enq_pcg_begin(B, s, B.u, B.first_iters);
  });
  if (rc || (rc = finish_pcg(B, st))) return rc;
  return run_seg(B, st, std::make_tuple(0, step, 3), [&](hipStream_t s) {
    launch(s, k_b_subgrad, B.blocks(D_FACES), c, B.dom(D_FACES), B.u, B.mg());
    enq_bfecc_s(B, s, B.init_s, B.psi, B.phi, B.err_s, B.s);
    // keep (phi, F) of the interval, reset phi, init_u = u, init_s = s
    enq_copy(B, s, {{B.u, B.init_u}}, {{B.s, B.init_s}});
    (void)hipMemcpyAsync(phi_at(B, step / B.R), B.phi, sizeof(float) * B.map_size(), hipMemcpyDeviceToDevice, s);
    launch(s, k_b_reset, B.blocks(D_ALL), c, B.dom(D_ALL), B.phi);
  });
}

// Adjoint step (backtrack_flow_map); on entry B.step is the step to undo + 1.
int backward_step(Batch& B, hipStream_t st) {
  const int step = --B.step;
  const int j = step % B.R;
  const Ctx c = B.ctx();
  float* mid = mid_at(B, step);
  int rc = run_seg(B, st, std::make_tuple(1, step, 0), [&](hipStream_t s) {
    launch(s, k_b_sources, B.blocks(D_FACES), c, B.dom(D_FACES), (const float*)B.u, (const float*)B.adj,
           (const float*)B.s, (const float*)B.ls, B.r_u, B.r_s);
    SLB A{};
    A.vel = mid;
    A.dt_scale = -1.f;
    A.nvec = 2;
    A.vsrc[0] = B.adj;
    A.vdst[0] = B.adj_new;
    A.vsrc[1] = B.u;
    A.vdst[1] = B.u_new;
    A.nscal = 2;
    A.ssrc[0] = B.s;
    A.sdst[0] = B.s_new;
    A.ssrc[1] = B.ls;
    A.sdst[1] = B.ls_new;
    launch(s, k_b_sl, B.blocks(D_ALL), c, B.dom(D_ALL), A);
    enq_copy(B, s, {}, {{B.s_new, B.s}, {B.ls_new, B.ls}});
    // lam_u += -dt r_u - dt r_s
    launch(s, k_b_axpy, B.blocks(D_FACES), c, B.dom(D_FACES), (const float*)B.adj_new, (const float*)B.r_u,
           (const float*)B.r_s, B.adj_new, -1.f, -1.f);
    launch(s, k_b_march, B.blocks(D_ALL), c, B.dom(D_ALL), (const float*)mid, B.psi, 1.f);
    launch(s, k_b_accumulate, B.blocks(D_FACES), c, B.dom(D_FACES),
           AccB{B.r_u, B.r_s, B.psi, B.init_adj, B.h1, B.h2, B.R - 1 - j, -2.f, -1.f});
    enq_pcg_begin(B, s, B.adj_new, B.first_iters);
  });
  if (rc || (rc = finish_pcg(B, st))) return rc;
  rc = run_seg(B, st, std::make_tuple(1, step, 1), [&](hipStream_t s) {
    launch(s, k_b_subgrad, B.blocks(D_FACES), c, B.dom(D_FACES), B.adj_new, B.mg());
    enq_copy(B, s, {{B.adj_new, B.adj}}, {});
    enq_pcg_begin(B, s, B.u_new, B.first_iters);
  });
  if (rc || (rc = finish_pcg(B, st))) return rc;
  if (j != 0) {
    return run_seg(B, st, std::make_tuple(1, step, 2), [&](hipStream_t s) {
      launch(s, k_b_subgrad, B.blocks(D_FACES), c, B.dom(D_FACES), B.u_new, B.mg());
      enq_copy(B, s, {{B.u_new, B.u}}, {});
    });
  }
  const float* fwd = phi_at(B, step / B.R);
  rc = run_seg(B, st, std::make_tuple(1, step, 2), [&](hipStream_t s) {
    launch(s, k_b_subgrad, B.blocks(D_FACES), c, B.dom(D_FACES), B.u_new, B.mg());
    enq_bfecc_u(B, s, B.init_adj, fwd, B.psi, B.err, B.adj);
    enq_pcg_begin(B, s, B.adj, B.first_iters);
  });
  if (rc || (rc = finish_pcg(B, st))) return rc;
  rc = run_seg(B, st, std::make_tuple(1, step, 3), [&](hipStream_t s) {
    launch(s, k_b_subgrad, B.blocks(D_FACES), c, B.dom(D_FACES), B.adj, B.mg());
    enq_bfecc_u(B, s, B.init_u, fwd, B.psi, B.err, B.u);
    enq_pcg_begin(B, s, B.u, B.first_iters);
  });
  if (rc || (rc = finish_pcg(B, st))) return rc;
  return run_seg(B, st, std::make_tuple(1, step, 4), [&](hipStream_t s) {
    launch(s, k_b_subgrad, B.blocks(D_FACES), c, B.dom(D_FACES), B.u, B.mg());
    // init_adj = lam_u, init_u = u; s from the interval-start checkpoint;
    // lam_s = BFECC of init_lam_s; init_lam_s = lam_s; psi reset
    enq_copy(B, s, {{B.adj, B.init_adj}, {B.u, B.init_u}}, {{sstart_at(B, step / B.R), B.s}});
    enq_bfecc_s(B, s, B.init_ls, fwd, B.psi, B.err_s, B.ls);
    enq_copy(B, s, {}, {{B.ls, B.init_ls}, {B.s, B.init_s}});
    launch(s, k_b_reset, B.blocks(D_ALL), c, B.dom(D_ALL), B.psi);
  });
}

// Multigrid plan: each instance's first tail level T is the finest level of
// at most tail_cells cells whose levels fit in LDS (z of every tail level, r
// of all but the first, whose r stays in global memory); levels 0..T get
// global z / r arrays.  (Re)allocates the level pools and the partials,
// rebuilds and uploads the instance table and the launch domains.
int plan_mg(Batch& B, int tail_cells) {
  int64_t om = 0;
  size_t lds_max = 0;
  for (auto& I : B.hI) {
    const int n0 = I.n0, n1 = I.n1;
    int T = I.L - 1;
    for (int l = I.L - 1; l >= 0; --l) {
      size_t bytes = 0;
      for (int k = l; k < I.L; ++k) bytes += sizeof(float) * (size_t)(n0 >> k) * (n1 >> k) * (k == l ? 1 : 2);
      if ((int64_t)(n0 >> l) * (n1 >> l) > tail_cells || bytes > (size_t)kLdsBytes) break;
      T = l;
    }
    I.T = T;
    size_t bytes = 0;
    for (int k = T; k < I.L; ++k) bytes += sizeof(float) * (size_t)(n0 >> k) * (n1 >> k) * (k == T ? 1 : 2);
    lds_max = std::max(lds_max, bytes);
    for (int l = 0; l < kMaxLv; ++l) I.om[l] = 0;
    for (int l = 0; l <= T; ++l) {
      I.om[l] = om;
      om += (int64_t)(n0 >> l) * (n1 >> l);
    }
  }
  B.tail_cells = tail_cells;
  B.MGcells = om;
  B.tail_lds = lds_max;
  build_domains(B);
  for (float** p : {&B.mz, &B.mr})
    if (*p) {
      (void)hipFree(*p);
      *p = nullptr;
    }
  if (B.part) {
    (void)hipFree(B.part);
    B.part = nullptr;
  }
  hipError_t e = hipMalloc((void**)&B.mz, sizeof(float) * std::max<int64_t>(om, 1));
  if (e == hipSuccess) e = hipMalloc((void**)&B.mr, sizeof(float) * std::max<int64_t>(om, 1));
  if (e == hipSuccess) e = hipMalloc((void**)&B.part, sizeof(double) * 2 * std::max(B.max_blocks, 1));
  std::vector<int> flat;
  for (auto& v : B.dom_h) flat.insert(flat.end(), v.begin(), v.end());
  if (e == hipSuccess) e = hipMemcpy(B.dom_d, flat.data(), sizeof(int) * flat.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(B.dI, B.hI.data(), sizeof(Inst) * B.count, hipMemcpyHostToDevice);
  if (e == hipSuccess && B.tail_lds > 64 * 1024)
    e = hipFuncSetAttribute((const void*)k_b_tail, hipFuncAttributeMaxDynamicSharedMemorySize, (int)B.tail_lds);
  if (e != hipSuccess) return set_error(DFM_E_HIP, "dfm_batch: multigrid plan", hipGetErrorString(e));
  return DFM_OK;
}

}  // namespace

// ------------------------------------------------------------------ C-ABI ----
struct dfm_batch {
  Batch b;
};

extern "C" int dfm_batch_create(const dfm_batch_instance* inst, int count, int reinit_every, int total_steps,
                                double tol, int max_iters, dfm_batch** out) {
  if (!inst || !out) return set_error(DFM_E_NULL, "dfm_batch_create: NULL pointer");
  *out = nullptr;
  if (count < 1 || count > 4096) return set_error(DFM_E_ARG, "dfm_batch_create: count must be 1..4096");
  if (reinit_every < 1 || total_steps < 1 || total_steps % reinit_every)
    return set_error(DFM_E_ARG, "dfm_batch_create: total_steps must be a positive multiple of reinit_every");
  if (!(tol > 0) || max_iters < 1) return set_error(DFM_E_ARG, "dfm_batch_create: tol > 0 and max_iters >= 1");
  auto* h = new dfm_batch();
  Batch& B = h->b;
  B.count = count;
  B.R = reinit_every;
  B.steps = total_steps;
  B.tol = tol;
  B.max_iters = max_iters;
  B.hI.resize(count);
  int64_t oc = 0, of = 0;
  for (int i = 0; i < count; ++i) {
    const int n0 = inst[i].n[0], n1 = inst[i].n[1];
    if (n0 < 8 || n1 < 8 || n0 > 8192 || n1 > 8192) {
      delete h;
      return set_error(DFM_E_ARG, "dfm_batch_create: each extent must be 8..8192");
    }
    if (!(inst[i].dx > 0) || !(inst[i].dt > 0)) {
      delete h;
      return set_error(DFM_E_ARG, "dfm_batch_create: dx and dt must be positive");
    }
    Inst& I = B.hI[i];
    std::memset(&I, 0, sizeof(I));
    I.n0 = n0;
    I.n1 = n1;
    I.dx = inst[i].dx;
    I.inv_dx = 1.0f / inst[i].dx;
    I.dt = inst[i].dt;
    I.L = std::min(kMaxLv, levels_of(n0, n1));
    if ((n0 % (2 << (I.L - 1))) || (n1 % (2 << (I.L - 1)))) {
      // every level's extents must stay even (the colour half-sweeps)
      delete h;
      return set_error(DFM_E_ARG, "dfm_batch_create: extents must be divisible by 2^levels");
    }
    I.oc = oc;
    I.of = of;
    I.inv_n = 1.0 / ((double)n0 * n1);
    I.max_iters = max_iters;
    oc += (int64_t)n0 * n1;
    of += faces0(I) + faces1(I);
  }
  B.C = oc;
  B.F = of;
  int rc = 0;
  const int nint = total_steps / reinit_every;
  auto A = [&](float** p, size_t n) { return rc ? rc : (rc = dalloc(B, p, n)); };
  A(&B.u, B.F); A(&B.u_new, B.F); A(&B.init_u, B.F); A(&B.err, B.F); A(&B.mid, (size_t)B.F * total_steps);
  A(&B.adj, B.F); A(&B.adj_new, B.F); A(&B.init_adj, B.F); A(&B.r_u, B.F); A(&B.r_s, B.F); A(&B.h1, B.F);
  A(&B.h2, B.F);
  A(&B.s, B.C); A(&B.s_new, B.C); A(&B.init_s, B.C); A(&B.err_s, B.C); A(&B.tgt, B.C);
  A(&B.s_start, (size_t)B.C * nint); A(&B.ls, B.C); A(&B.ls_new, B.C); A(&B.init_ls, B.C);
  A(&B.phi, B.map_size()); A(&B.psi, B.map_size()); A(&B.phis, (size_t)B.map_size() * nint);
  A(&B.mp, B.C); A(&B.mAp, B.C);
  if (!rc) rc = dalloc(B, &B.mx, B.C);
  if (!rc) rc = dalloc(B, &B.sc, (size_t)S_COUNT * count);
  if (!rc) rc = dalloc(B, &B.fl, (size_t)F_COUNT * count);
  if (!rc) rc = dalloc(B, &B.ctr, (size_t)kCtrStride * count);
  if (!rc) rc = dalloc(B, &B.dI, count);
  if (!rc) rc = dalloc(B, &B.dom_d, (size_t)D_COUNT * (count + 1));
  if (!rc) rc = plan_mg(B, B.tail_cells);
  hipError_t e = hipSuccess;
  if (!rc) {
    e = hipMemset(B.ctr, 0, sizeof(unsigned) * kCtrStride * count);
    if (e == hipSuccess) e = hipMemset(B.fl, 0, sizeof(int) * F_COUNT * count);
    if (e == hipSuccess) e = hipHostMalloc((void**)&B.h_fl, sizeof(int) * F_COUNT * count, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&B.h_sc, sizeof(double) * S_COUNT * count, hipHostMallocDefault);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&B.cap, hipStreamNonBlocking);
    if (e != hipSuccess) rc = set_error(DFM_E_HIP, "dfm_batch_create", hipGetErrorString(e));
  }
  if (rc) {
    dfm_batch_destroy(h);
    return rc;
  }
  *out = h;
  return DFM_OK;
}

extern "C" int dfm_batch_destroy(dfm_batch* h) {
  if (!h) return DFM_OK;
  Batch& B = h->b;
  (void)hipDeviceSynchronize();
  for (auto& kv : B.graphs_c) (void)hipGraphExecDestroy(kv.second);
  for (void* p : B.allocs) (void)hipFree(p);
  for (void* p : {(void*)B.mz, (void*)B.mr, (void*)B.part})
    if (p) (void)hipFree(p);
  if (B.h_fl) (void)hipHostFree(B.h_fl);
  if (B.h_sc) (void)hipHostFree(B.h_sc);
  if (B.cap) (void)hipStreamDestroy(B.cap);
  delete h;
  return DFM_OK;
}

extern "C" int dfm_batch_field(dfm_batch* h, int field, int inst, float** c0, float** c1) {
  if (!h || !c0) return set_error(DFM_E_NULL, "dfm_batch_field: NULL pointer");
  Batch& B = h->b;
  if (inst < 0 || inst >= B.count) return set_error(DFM_E_ARG, "dfm_batch_field: instance out of range");
  const Inst& I = B.hI[inst];
  float* mac = nullptr;
  float* cell = nullptr;
  switch (field) {
    case DFM_BATCH_U: mac = B.u; break;
    case DFM_BATCH_ADJ_U: mac = B.adj; break;
    case DFM_BATCH_S: cell = B.s; break;
    case DFM_BATCH_ADJ_S: cell = B.ls; break;
    case DFM_BATCH_TARGET: cell = B.tgt; break;
    default: return set_error(DFM_E_ARG, "dfm_batch_field: unknown field");
  }
  if (mac) {
    *c0 = mac + I.of;
    if (c1) *c1 = mac + I.of + faces0(I);
  } else {
    *c0 = cell + I.oc;
    if (c1) *c1 = nullptr;
  }
  return DFM_OK;
}

extern "C" int dfm_batch_copy(dfm_batch* h, int field, int inst, float* c0, float* c1, int to_engine, void* stream) {
  float *e0 = nullptr, *e1 = nullptr;
  int rc = dfm_batch_field(h, field, inst, &e0, &e1);
  if (rc) return rc;
  if (!c0 || (e1 && !c1)) return set_error(DFM_E_NULL, "dfm_batch_copy: NULL buffer");
  const Inst& I = h->b.hI[inst];
  hipStream_t st = (hipStream_t)stream;
  const size_t n0 = sizeof(float) * (size_t)(e1 ? faces0(I) : cells(I)), n1 = sizeof(float) * (size_t)faces1(I);
  hipError_t e = to_engine ? hipMemcpyAsync(e0, c0, n0, hipMemcpyDeviceToDevice, st)
                           : hipMemcpyAsync(c0, e0, n0, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess && e1)
    e = to_engine ? hipMemcpyAsync(e1, c1, n1, hipMemcpyDeviceToDevice, st)
                  : hipMemcpyAsync(c1, e1, n1, hipMemcpyDeviceToDevice, st);
  return e == hipSuccess ? DFM_OK : set_error(DFM_E_HIP, "dfm_batch_copy", hipGetErrorString(e));
}

extern "C" int dfm_batch_set_option(dfm_batch* h, int option, int value) {
  if (!h) return set_error(DFM_E_NULL, "dfm_batch_set_option: NULL handle");
  Batch& B = h->b;
  switch (option) {
    case DFM_BATCH_GRAPH: B.graphs = value ? 1 : 0; break;
    case DFM_BATCH_FIRST_ITERS:
      if (value < 0 || value > 200) return set_error(DFM_E_ARG, "dfm_batch_set_option: first iterations 0..200");
      B.first_iters = value;
      break;
    case DFM_BATCH_CHUNK:
      if (value < 1 || value > 200) return set_error(DFM_E_ARG, "dfm_batch_set_option: chunk 1..200");
      B.chunk = value;
      break;
    case DFM_BATCH_TAIL_CELLS:
      if (value < 1) return set_error(DFM_E_ARG, "dfm_batch_set_option: tail cells >= 1");
      break;
    default: return set_error(DFM_E_ARG, "dfm_batch_set_option: unknown option");
  }
  // captured graphs hold the old schedule
  (void)hipDeviceSynchronize();
  for (auto& kv : B.graphs_c) (void)hipGraphExecDestroy(kv.second);
  B.graphs_c.clear();
  if (option == DFM_BATCH_TAIL_CELLS) return plan_mg(B, value);
  return DFM_OK;
}

extern "C" int dfm_batch_init(dfm_batch* h, void* stream) {
  if (!h) return set_error(DFM_E_NULL, "dfm_batch_init: NULL handle");
  Batch& B = h->b;
  hipStream_t st = (hipStream_t)stream;
  int rc = run_seg(B, st, std::make_tuple(2, 0, 0), [&](hipStream_t s) { enq_pcg_begin(B, s, B.u, B.first_iters); });
  if (rc || (rc = finish_pcg(B, st))) return rc;
  rc = run_seg(B, st, std::make_tuple(2, 0, 1), [&](hipStream_t s) {
    const Ctx c = B.ctx();
    launch(s, k_b_subgrad, B.blocks(D_FACES), c, B.dom(D_FACES), B.u, B.mg());
    enq_copy(B, s, {{B.u, B.init_u}}, {{B.s, B.init_s}});
    launch(s, k_b_reset, B.blocks(D_ALL), c, B.dom(D_ALL), B.phi);
  });
  B.step = -1;
  return rc ? rc : DFM_OK;
}

extern "C" int dfm_batch_forward(dfm_batch* h, int steps, void* stream) {
  if (!h) return set_error(DFM_E_NULL, "dfm_batch_forward: NULL handle");
  Batch& B = h->b;
  if (steps < 0 || B.step + steps >= B.steps) return set_error(DFM_E_ARG, "dfm_batch_forward: beyond total_steps");
  for (int k = 0; k < steps; ++k) {
    int rc = forward_step(B, (hipStream_t)stream);
    if (rc) return rc;
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? DFM_OK : set_error(DFM_E_HIP, "dfm_batch_forward", hipGetErrorString(e));
}

extern "C" int dfm_batch_init_gradient(dfm_batch* h, double* loss, void* stream) {
  if (!h) return set_error(DFM_E_NULL, "dfm_batch_init_gradient: NULL handle");
  Batch& B = h->b;
  if (B.step + 1 != B.steps) return set_error(DFM_E_ARG, "dfm_batch_init_gradient: the forward run is not complete");
  hipStream_t st = (hipStream_t)stream;
  const Ctx c = B.ctx();
  launch(st, k_b_terminal, B.blocks(D_CELLS), c, B.dom(D_CELLS), (const float*)B.s, (const float*)B.tgt, B.ls);
  enq_copy(B, st, {{nullptr, B.adj}, {nullptr, B.h1}, {nullptr, B.h2}}, {{B.ls, B.init_ls}, {B.s, B.init_s}});
  enq_copy(B, st, {{nullptr, B.init_adj}, {B.u, B.init_u}}, {});
  launch(st, k_b_reset, B.blocks(D_ALL), c, B.dom(D_ALL), B.psi);
  B.step = B.steps;
  hipError_t e = hipMemcpyAsync(B.h_sc, B.sc, sizeof(double) * S_COUNT * B.count, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return set_error(DFM_E_HIP, "dfm_batch_init_gradient", hipGetErrorString(e));
  if (loss)
    for (int i = 0; i < B.count; ++i) loss[i] = B.h_sc[i * S_COUNT + S_LOSS];
  return DFM_OK;
}

extern "C" int dfm_batch_backward(dfm_batch* h, int steps, void* stream) {
  if (!h) return set_error(DFM_E_NULL, "dfm_batch_backward: NULL handle");
  Batch& B = h->b;
  if (steps < 0 || B.step - steps < 0) return set_error(DFM_E_ARG, "dfm_batch_backward: beyond step 0");
  for (int k = 0; k < steps; ++k) {
    int rc = backward_step(B, (hipStream_t)stream);
    if (rc) return rc;
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? DFM_OK : set_error(DFM_E_HIP, "dfm_batch_backward", hipGetErrorString(e));
}

extern "C" int dfm_batch_stats(dfm_batch* h, int* projections, int* max_iters, int* noconv) {
  if (!h) return set_error(DFM_E_NULL, "dfm_batch_stats: NULL handle");
  Batch& B = h->b;
  if (projections) *projections = B.projections;
  if (max_iters) *max_iters = B.max_seen_iters;
  if (noconv) *noconv = B.noconv;
  return DFM_OK;
}

extern "C" int dfm_batch_levels(dfm_batch* h, int inst, int* levels, int* tail_level) {
  if (!h) return set_error(DFM_E_NULL, "dfm_batch_levels: NULL handle");
  Batch& B = h->b;
  if (inst < 0 || inst >= B.count) return set_error(DFM_E_ARG, "dfm_batch_levels: instance out of range");
  if (levels) *levels = B.hI[inst].L;
  if (tail_level) *tail_level = B.hI[inst].T;
  return DFM_OK;
}
