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

// Launch domain: blk[i] .. blk[i+1] are instance i's blocks; bt[b] =
// (instance, first block of that instance) for every block b, so a block
// finds its instance with one load.
struct Dom {
  const int* blk;
  const int2* bt;
  int count;
};

struct Ctx {
  const Inst* I;
  double* sc;   // S_COUNT per instance
  int* fl;      // F_COUNT per instance
  double* part; // 2 per block of the largest reducing launch
  unsigned* ctr;
  double tol;
};

DFM_HD bool inst_done(const Ctx& C, int i) { return C.fl[i * F_COUNT + F_DONE] != 0; }

// Each thread handles PPT points of its instance, BS apart: the block's
// set-up latency (instance lookup, instance record) is paid once per
// PPT * BS points, and the points' loads are in flight together.
constexpr int PPT = 4;
struct Where {
  int i;
  int t0;
};
DFM_HD Where where(const Dom& D) {
  const int2 e = D.bt[blockIdx.x];
  return Where{e.x, (int)(blockIdx.x - e.y) * (BS * PPT) + (int)threadIdx.x};
}

// q = t / d, r = t % d for 0 <= t < 2^24, d >= 1: a float quotient estimate
// (off by at most a few units) corrected exactly in integers, instead of
// the long integer-division sequence.
DFM_HD void divmod(int t, int d, int& q, int& r) {
  q = (int)((float)t * __frcp_rn((float)d));
  r = t - q * d;
  while (r < 0) {
    --q;
    r += d;
  }
  while (r >= d) {
    ++q;
    r -= d;
  }
}
// (i, j) of row-major index t in a grid whose rows hold `cols` entries.
DFM_HD void rc_of(int t, int cols, int& i, int& j) { divmod(t, cols, i, j); }

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
DFM_HD int point_grid(const Inst& I, int t, int& k) {
  const int F0 = (int)faces0(I), F1 = (int)faces1(I);
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
DFM_HD int npoints(const Inst& I, bool with_cells) {
  return (int)(faces0(I) + faces1(I) + (with_cells ? cells(I) : 0));
}
DFM_HD void node_of(const Geo& g, int grid, int k, float x[3]) {
  int sh[3];
  shape_of(g, grid, sh);
  int a, b;
  rc_of(k, sh[1], a, b);
  x[2] = 0.f;
  node_pos(g, grid, a, b, 0, x);
}

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
  const int N = npoints(I, A.nscal > 0);
  const Geo g = geo_of(I);
  const Mac vel = mac_of(A.vel, I);
  const float dt = A.dt_scale * I.dt;
#pragma unroll 1
  for (int u = 0, t = w.t0; u < PPT && t < N; ++u, t += BS) {
    int k;
    const int grid = point_grid(I, t, k);
    if (grid < 3 && A.nvec == 0) continue;
    int sh[3];
    shape_of(g, grid, sh);
    float x[3], p[3] = {0.f, 0.f, 0.f};
    node_of(g, grid, k, x);
    backtrace<2, 2>(g, vel, dt, x, p);
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
}

// RK2 march of every point set of a map (faces with T, cells psi only).
__global__ void __launch_bounds__(BS) k_b_march(Ctx C, Dom D, const float* vel, float* map, float dt_scale) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  const int N = npoints(I, true);
  const Geo g = geo_of(I);
  const MapV M = map_of(map, I);
  const Mac u = mac_of(vel, I);
  const float dt = dt_scale * I.dt;
#pragma unroll 1
  for (int uu = 0, t = w.t0; uu < PPT && t < N; ++uu, t += BS) {
    int k;
    const int grid = point_grid(I, t, k);
    float p[3] = {0.f, 0.f, 0.f}, T[3] = {0.f, 0.f, 0.f};
    if (grid < 3) {
      const int cnt = (int)(grid ? faces1(I) : faces0(I));
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
      const int cnt = (int)cells(I);
      p[0] = M.psi[k];
      p[1] = M.psi[cnt + k];
      rk_point<2, 2, false>(g, u, dt, p, T);
      M.psi[k] = p[0];
      M.psi[cnt + k] = p[1];
    }
  }
}

__global__ void __launch_bounds__(BS) k_b_reset(Ctx C, Dom D, float* map) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  const int N = npoints(I, true);
  const Geo g = geo_of(I);
  const MapV M = map_of(map, I);
  for (int u = 0, t = w.t0; u < PPT && t < N; ++u, t += BS) {
    int k;
    const int grid = point_grid(I, t, k);
    float x[3];
    node_of(g, grid, k, x);
    if (grid < 3) {
      const int cnt = (int)(grid ? faces1(I) : faces0(I));
      M.pf[grid][k] = x[0];
      M.pf[grid][cnt + k] = x[1];
      M.tf[grid][k] = grid == 0 ? 1.f : 0.f;
      M.tf[grid][cnt + k] = grid == 1 ? 1.f : 0.f;
    } else {
      M.psi[k] = x[0];
      M.psi[cells(I) + k] = x[1];
    }
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
  const int N = npoints(I, false);
  const Geo g = geo_of(I);
  const MapV M = map_of(const_cast<float*>(A.map), I);
  const Mac src = mac_of(A.src, I);
#pragma unroll 1
  for (int u = 0, t = w.t0; u < PPT && t < N; ++u, t += BS) {
    int k;
    const int grid = point_grid(I, t, k);
    const int cnt = (int)(grid ? faces1(I) : faces0(I));
    float p[3] = {M.pf[grid][k], M.pf[grid][cnt + k], 0.f}, v[3];
    mac_eval<false, 2>(g, src, p, v, nullptr);
    const float d = fmaf(M.tf[grid][cnt + k], v[1], M.tf[grid][k] * v[0]);
    float r = A.alpha * d;
    if (A.beta != 0.f) r = fmaf(A.beta, mac_of(A.add, I).c[grid][k], r);
    mac_c(A.out, I, grid)[k] = r;
  }
}
__global__ void __launch_bounds__(BS) k_b_pullback_s(Ctx C, Dom D, PBB A) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  const int cnt = (int)cells(I);
  const Geo g = geo_of(I);
  const MapV M = map_of(const_cast<float*>(A.map), I);
  int S[3];
  shape_of(g, 3, S);
  const float bl[3] = {0.5f, 0.5f, 0.5f};
#pragma unroll 1
  for (int u = 0, k = w.t0; u < PPT && k < cnt; ++u, k += BS) {
    const float p[3] = {M.psi[k], M.psi[cnt + k], 0.f};
    float r = A.alpha * interp<2, false, 2>(A.src + I.oc, S, bl, p, g.inv_dx, nullptr);
    if (A.beta != 0.f) r = fmaf(A.beta, A.add[I.oc + k], r);
    A.out[I.oc + k] = r;
  }
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
  const int N = npoints(I, false);
  const Geo g = geo_of(I);
  const MapV M = map_of(const_cast<float*>(A.map), I);
  const float dt = A.dt_scale * I.dt, dt2 = A.dt2_scale * I.dt;
#pragma unroll 1
  for (int u = 0, t = w.t0; u < PPT && t < N; ++u, t += BS) {
    int k;
    const int grid = point_grid(I, t, k);
    const int cnt = (int)(grid ? faces1(I) : faces0(I));
    const float p[3] = {M.pf[grid][k], M.pf[grid][cnt + k], 0.f};
    const float T0 = M.tf[grid][k], T1 = M.tf[grid][cnt + k];
    float v[3];
    mac_eval<false, 2>(g, mac_of(A.f, I), p, v, nullptr);
    const float gv = fmaf(T1, v[1], T0 * v[0]);
    float* init = mac_c(A.init, I, grid) + k;
    float* h1 = mac_c(A.h1, I, grid) + k;
    float* h2 = mac_c(A.h2, I, grid) + k;
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
}

// Both adjoint sources at the face points: r_u = ((grad u)^T lam_u) with the
// quartic spline, r_s = grad s . lam_s with the cubic one.
__global__ void __launch_bounds__(BS) k_b_sources(Ctx C, Dom D, const float* u, const float* lam, const float* s,
                                                  const float* lam_s, float* r_u, float* r_s) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  const int N = npoints(I, false);
  const Geo g = geo_of(I);
  const Mac U = mac_of(u, I), W = mac_of(lam, I);
#pragma unroll 1
  for (int uu = 0, t = w.t0; uu < PPT && t < N; ++uu, t += BS) {
    int k;
    const int grid = point_grid(I, t, k);
    float x[3];
    node_of(g, grid, k, x);
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
}

// out = x + (alpha dt) y + (beta dt) z on MAC pools (z may be null).
__global__ void __launch_bounds__(BS) k_b_axpy(Ctx C, Dom D, const float* x, const float* y, const float* z, float* out,
                                               float alpha_dt, float beta_dt) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  const int N = npoints(I, false);
  const float a = alpha_dt * I.dt, b = beta_dt * I.dt;
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    const int t = w.t0 + u * BS;
    if (t < N) {
      const int64_t o = I.of + t;
      float r = fmaf(a, y[o], x[o]);
      if (z) r = fmaf(b, z[o], r);
      out[o] = r;
    }
  }
}

// Copies: up to 3 MAC pools (face points) and 3 cell pools in one launch; a
// null source writes zeros.
struct CopyB {
  int nmac, ncell;
  const float* msrc[3];
  float* mdst[3];
  const float* csrc[3];
  float* cdst[3];
};
__global__ void __launch_bounds__(BS) k_b_copy(Ctx C, Dom D, CopyB A) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  const int F = (int)(faces0(I) + faces1(I));
  const int N = npoints(I, A.ncell > 0);
  for (int u = 0, t = w.t0; u < PPT && t < N; ++u, t += BS) {
    if (t < F) {
      for (int m = 0; m < A.nmac; ++m) A.mdst[m][I.of + t] = A.msrc[m] ? A.msrc[m][I.of + t] : 0.f;
    } else {
      const int k = t - F;
      for (int m = 0; m < A.ncell; ++m) A.cdst[m][I.oc + k] = A.csrc[m] ? A.csrc[m][I.oc + k] : 0.f;
    }
  }
}

// Terminal adjoint: lam_s = 2 (s - target), loss = sum (s - target)^2 (FP64).
__global__ void __launch_bounds__(BS) k_b_terminal(Ctx C, Dom D, const float* s, const float* tgt, float* lam_s) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  const int N = (int)cells(I);
  double l = 0;
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    const int t = w.t0 + u * BS;
    if (t < N) {
      const int64_t o = I.oc + t;
      const float d = s[o] - tgt[o];
      lam_s[o] = 2.0f * d;
      l += (double)d * (double)d;
    }
  }
  inst_reduce<BS>(C, D, w.i, FIN_LOSS, l, 0.0);
}

// ------------------------------------------------------------------ MGPCG ----
// Level-l arrays of an instance: natural row-major (i n1 + j), n >> l.
struct MG {
  float* z;   // level pools (levels 0..T): z (down legs and tail), r
  float* zo;  // z after the up leg of a global level (levels 0..T-1)
  float* r;
  float* p;
  float* Ap;
  double* x;
};

// Neumann 5-point operator pieces of cell (i, j) on an a x b grid.
DFM_HD int nnbr(int i, int j, int a, int b) { return (i > 0) + (i < a - 1) + (j > 0) + (j < b - 1); }
DFM_HD float nsum(const float* f, int i, int j, int a, int b) {
  const int id = i * b + j;
  float s = 0.f;
  if (i > 0) s += f[id - b];
  if (i < a - 1) s += f[id + b];
  if (j > 0) s += f[id - 1];
  if (j < b - 1) s += f[id + 1];
  return s;
}
DFM_HD float inv_nn(int nn) { return nn == 4 ? 0.25f : (nn == 3 ? (1.0f / 3.0f) : (nn == 2 ? 0.5f : (nn == 1 ? 1.f : 0.f))); }
DFM_HD float mean_of(const Ctx& C, int inst) { return (float)C.sc[(int64_t)inst * S_COUNT + S_MEAN]; }

// b = -div(u) (wall faces read as 0), partial sum -> mean; level-0 r = b.
__global__ void __launch_bounds__(BS) k_b_div(Ctx C, Dom D, const float* u, MG M) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  const int N = (int)cells(I);
  const Mac U = mac_of(u, I);
  double s = 0;
#pragma unroll
  for (int uu = 0; uu < PPT; ++uu) {
    const int t = w.t0 + uu * BS;
    if (t < N) {
      int i, j;
      rc_of(t, I.n1, i, j);
      const int fx = i * I.n1 + j, fy = i * (I.n1 + 1) + j;
      const float uxl = i > 0 ? U.c[0][fx] : 0.f, uxh = i < I.n0 - 1 ? U.c[0][fx + I.n1] : 0.f;
      const float uyl = j > 0 ? U.c[1][fy] : 0.f, uyh = j < I.n1 - 1 ? U.c[1][fy + 1] : 0.f;
      const float b = -((uxh - uxl) + (uyh - uyl));
      M.r[I.om[0] + t] = b;
      s += b;
    }
  }
  inst_reduce<BS>(C, D, w.i, FIN_MEAN, s, 0.0);
}

// r -= mean, x = 0, partial r.r -> |b|^2, threshold, flags reset.
__global__ void __launch_bounds__(BS) k_b_recenter(Ctx C, Dom D, MG M) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  const int N = (int)cells(I);
  const float m = mean_of(C, w.i);
  double rr = 0;
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    const int t = w.t0 + u * BS;
    if (t < N) {
      float* r = M.r + I.om[0] + t;
      const float v = *r - m;
      *r = v;
      M.x[I.oc + t] = 0.0;
      rr += (double)v * v;
    }
  }
  inst_reduce<BS>(C, D, w.i, FIN_BNORM, rr, 0.0);
}

// Global multigrid levels: LDS-tiled legs.  A block owns a TS x TS tile of
// an instance's level and loads the tile plus a kTH-cell halo (cells of the
// grid only) into LDS; every half-sweep updates the whole region, reading a
// neighbour outside the region as 0.  The down leg starts from z = 0 and
// the up leg from the exact prolonged z, so after the leg's four
// half-sweeps only cells within 2 (down) or 3 (up) of the region's edge
// can differ from the global red-black sweep; with kTH = 4 the tile's
// cells, and for the restriction their neighbours, are the global sweep's
// values, and each leg is one launch instead of four.
constexpr int kTH = 4;

template <int R>
__device__ void tile_half(float* zt, const float* rt, int i0, int j0, int a, int b, int c) {
  constexpr int HR = R / 2;
  for (int t = threadIdx.x; t < R * HR; t += BS) {
    const int li = t / HR;
    const int i = i0 + li;
    const int lj = 2 * (t - li * HR) + ((c + i + j0) & 1);
    const int j = j0 + lj;
    if (i < 0 || i >= a || j < 0 || j >= b) continue;
    const int id = li * R + lj;
    float s = 0.f;
    if (i > 0 && li > 0) s += zt[id - R];
    if (i < a - 1 && li < R - 1) s += zt[id + R];
    if (j > 0 && lj > 0) s += zt[id - 1];
    if (j < b - 1 && lj < R - 1) s += zt[id + 1];
    zt[id] = (rt[id] + s) * inv_nn(nnbr(i, j, a, b));
  }
}

// Tile of block b of a tiled domain: instance, tile row / column.
DFM_HD void tile_of(const Dom& D, int TS, const Inst& I, int l, int& ti, int& tj) {
  const int2 e = D.bt[blockIdx.x];
  const int tiles_j = (n_of(I, l, 1) + TS - 1) / TS;
  rc_of((int)blockIdx.x - e.y, tiles_j, ti, tj);
}

// Down leg on level l: two red-black sweeps from z = 0, then r_{l+1} = the
// restriction (sum over 4 children) of r - A z; z written back.
template <int TS>
__global__ void __launch_bounds__(BS) k_b_down(Ctx C, Dom D, MG M, int l) {
  constexpr int R = TS + 2 * kTH;
  __shared__ float zt[R * R], rt[R * R];
  const int inst = D.bt[blockIdx.x].x;
  if (inst_done(C, inst)) return;
  const Inst I = C.I[inst];
  const int a = n_of(I, l, 0), b = n_of(I, l, 1);
  int ti, tj;
  tile_of(D, TS, I, l, ti, tj);
  const int i0 = ti * TS - kTH, j0 = tj * TS - kTH;
  const float m = l == 0 ? mean_of(C, inst) : 0.f;
  const float* r = M.r + I.om[l];
  for (int t = threadIdx.x; t < R * R; t += BS) {
    const int li = t / R, lj = t - li * R, i = i0 + li, j = j0 + lj;
    rt[t] = (i >= 0 && i < a && j >= 0 && j < b) ? r[i * b + j] - m : 0.f;
    zt[t] = 0.f;
  }
  __syncthreads();
#pragma unroll 1
  for (int h = 0; h < 4; ++h) {
    tile_half<R>(zt, rt, i0, j0, a, b, h & 1);
    __syncthreads();
  }
  // restriction of the tile's residual (the tile origin is even, so its
  // coarse cells are whole)
  constexpr int CT = TS / 2;
  const int cb = b >> 1;
  float* rc = M.r + I.om[l + 1];
  for (int t = threadIdx.x; t < CT * CT; t += BS) {
    const int ci = t / CT, cj = t - ci * CT;
    const int ic = ti * CT + ci, jc = tj * CT + cj;
    if (2 * ic >= a || 2 * jc >= b) continue;
    float sres = 0.f;
#pragma unroll
    for (int di = 0; di < 2; ++di)
#pragma unroll
      for (int dj = 0; dj < 2; ++dj) {
        const int li = kTH + 2 * ci + di, lj = kTH + 2 * cj + dj;
        const int i = i0 + li, j = j0 + lj, id = li * R + lj;
        float ns = 0.f;
        if (i > 0) ns += zt[id - R];
        if (i < a - 1) ns += zt[id + R];
        if (j > 0) ns += zt[id - 1];
        if (j < b - 1) ns += zt[id + 1];
        sres += rt[id] - ((float)nnbr(i, j, a, b) * zt[id] - ns);
      }
    rc[ic * cb + jc] = sres;
  }
  float* z = M.z + I.om[l];
  for (int t = threadIdx.x; t < TS * TS; t += BS) {
    const int li = t / TS, lj = t - li * TS, i = ti * TS + li, j = tj * TS + lj;
    if (i < a && j < b) z[i * b + j] = zt[(li + kTH) * R + lj + kTH];
  }
}

// Up leg on level l: z += z_{l+1}[I/2] (piecewise-constant prolongation),
// then two black-red sweeps.  The leg reads the down leg's z (halo included)
// and writes zo: writing z in place would race with the neighbouring tiles'
// reads of their halos.  The coarse correction is the tail's z at l + 1 = T,
// else the up leg's zo.
template <int TS>
__global__ void __launch_bounds__(BS) k_b_up(Ctx C, Dom D, MG M, int l) {
  constexpr int R = TS + 2 * kTH;
  __shared__ float zt[R * R], rt[R * R];
  const int inst = D.bt[blockIdx.x].x;
  if (inst_done(C, inst)) return;
  const Inst I = C.I[inst];
  const int a = n_of(I, l, 0), b = n_of(I, l, 1), cb = b >> 1;
  int ti, tj;
  tile_of(D, TS, I, l, ti, tj);
  const int i0 = ti * TS - kTH, j0 = tj * TS - kTH;
  const float m = l == 0 ? mean_of(C, inst) : 0.f;
  const float* r = M.r + I.om[l];
  const float* z = M.z + I.om[l];
  const float* zc = (l + 1 == I.T ? M.z : M.zo) + I.om[l + 1];
  for (int t = threadIdx.x; t < R * R; t += BS) {
    const int li = t / R, lj = t - li * R, i = i0 + li, j = j0 + lj;
    const bool in = i >= 0 && i < a && j >= 0 && j < b;
    rt[t] = in ? r[i * b + j] - m : 0.f;
    zt[t] = in ? z[i * b + j] + zc[(i >> 1) * cb + (j >> 1)] : 0.f;
  }
  __syncthreads();
#pragma unroll 1
  for (int h = 0; h < 4; ++h) {
    tile_half<R>(zt, rt, i0, j0, a, b, (h & 1) ^ 1);
    __syncthreads();
  }
  float* zo = M.zo + I.om[l];
  for (int t = threadIdx.x; t < TS * TS; t += BS) {
    const int li = t / TS, lj = t - li * TS, i = ti * TS + li, j = tj * TS + lj;
    if (i < a && j < b) zo[i * b + j] = zt[(li + kTH) * R + lj + kTH];
  }
}

// The V-cycle's level-0 result: the up leg's output, or the tail's when the
// tail holds level 0.
DFM_HD const float* z_final(const MG& M, const Inst& I) { return (I.T > 0 ? M.zo : M.z) + I.om[0]; }

// One workgroup per instance runs levels T..L-1 in LDS: from r_T (global)
// the V-cycle's remaining legs (2 pre sweeps, restriction, `bottom` sweeps
// on the coarsest level, prolongation, 2 post sweeps), z_T written back.
// Level T holds at most kTailMax cells (16 per thread): each thread keeps
// the r values of its level-T cells in registers for the whole V-cycle (r_T
// does not change during it), so the level-T sweeps read only LDS.  Every
// half-sweep loop is unrolled (a colour's cells are independent), so a
// thread's LDS loads for its cells are in flight together.
constexpr int kTailMax = 16384;
constexpr int kTailQ = kTailMax / 2 / kTailThreads;  // cells of one colour per thread at level T

// Half-sweep of colour `phase` over an a x b LDS level, r from LDS.
template <int Q>
__device__ void lds_half(float* z, const float* r, int a, int b, int phase) {
  const int hb = b >> 1, N = a * hb;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int t = threadIdx.x + q * kTailThreads;
    if (t < N) {
      int i, jh;
      rc_of(t, hb, i, jh);
      const int j = 2 * jh + ((i + phase) & 1);
      const int id = i * b + j;
      z[id] = (r[id] + nsum(z, i, j, a, b)) * inv_nn(nnbr(i, j, a, b));
    }
  }
}
// Level-T half-sweep with the thread's r values from registers.
__device__ void reg_half(float* z, const float (&rr)[kTailQ], int a, int b, int phase) {
  const int hb = b >> 1, N = a * hb;
#pragma unroll
  for (int q = 0; q < kTailQ; ++q) {
    const int t = threadIdx.x + q * kTailThreads;
    if (t < N) {
      int i, jh;
      rc_of(t, hb, i, jh);
      const int j = 2 * jh + ((i + phase) & 1);
      const int id = i * b + j;
      z[id] = (rr[q] + nsum(z, i, j, a, b)) * inv_nn(nnbr(i, j, a, b));
    }
  }
}
// The same half-sweep by the 64 lanes of wave 0 alone: a wavefront's LDS
// operations are performed in order, so consecutive half-sweeps of one wave
// need only a wavefront-scope fence between them (no workgroup barrier);
// used for the `bottom` sweeps of the coarsest level (at most 1024 cells).
__device__ void wave_half(float* z, const float* r, int a, int b, int phase) {
  const int hb = b >> 1;
  for (int t = threadIdx.x; t < a * hb; t += 64) {
    int i, jh;
    rc_of(t, hb, i, jh);
    const int j = 2 * jh + ((i + phase) & 1);
    const int id = i * b + j;
    z[id] = (r[id] + nsum(z, i, j, a, b)) * inv_nn(nnbr(i, j, a, b));
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// z_l of level l (>= T) restricted: r_{l+1} = sum over 4 children of (r - A z).
template <int Q, class RF>
__device__ void lds_restrict(const float* z, RF r_at, float* rc, int a, int b) {
  const int bc = b >> 1, N = (a >> 1) * bc;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int t = threadIdx.x + q * kTailThreads;
    if (t < N) {
      int ic, jc;
      rc_of(t, bc, ic, jc);
      float s = 0.f;
#pragma unroll
      for (int di = 0; di < 2; ++di)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj) {
          const int i = 2 * ic + di, j = 2 * jc + dj, id = i * b + j;
          s += r_at(id) - ((float)nnbr(i, j, a, b) * z[id] - nsum(z, i, j, a, b));
        }
      rc[t] = s;
    }
  }
}
// z_l += z_{l+1}[I/2] (piecewise-constant prolongation).
template <int Q>
__device__ void lds_prolong(float* z, const float* zc, int a, int b) {
  const int bc = b >> 1, N = a * b;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int t = threadIdx.x + q * kTailThreads;
    if (t < N) {
      int i, j;
      rc_of(t, b, i, j);
      z[t] += zc[(i >> 1) * bc + (j >> 1)];
    }
  }
}

// Exact solve of the coarsest level (a, b <= 16): the all-Neumann operator
// (#neighbours) I - adjacency is diagonalised by the orthonormal DCT-II
// (eigenvalues (2 - 2 cos(pi k / a)) + (2 - 2 cos(pi m / b))), so z =
// C_a^T [(C_a r C_b^T) / lambda] C_b with the constant mode set to 0.
// w holds 4 x 256 floats of LDS scratch.  Called by every thread.
constexpr int kBottomMax = 16;
__device__ void dct_bottom(float* z, const float* r, int a, int b, float* w) {
  float* Ca = w;            // Ca[k * a + i]
  float* Cb = w + 256;      // Cb[m * b + j]
  float* t1 = w + 512;      // a x b
  float* t2 = w + 768;      // a x b
  const float pi = 3.14159265358979f;
  for (int t = threadIdx.x; t < a * a; t += kTailThreads) {
    const int k = t / a, i = t % a;
    Ca[t] = sqrtf((k ? 2.0f : 1.0f) / a) * cosf(pi * k * (i + 0.5f) / a);
  }
  for (int t = threadIdx.x; t < b * b; t += kTailThreads) {
    const int m = t / b, j = t % b;
    Cb[t] = sqrtf((m ? 2.0f : 1.0f) / b) * cosf(pi * m * (j + 0.5f) / b);
  }
  __syncthreads();
  const int n = a * b;
  if (threadIdx.x < n) {  // t1 = Ca r
    const int k = threadIdx.x / b, j = threadIdx.x % b;
    float v = 0.f;
    for (int i = 0; i < a; ++i) v = fmaf(Ca[k * a + i], r[i * b + j], v);
    t1[threadIdx.x] = v;
  }
  __syncthreads();
  if (threadIdx.x < n) {  // t2 = (t1 Cb^T) / lambda
    const int k = threadIdx.x / b, m = threadIdx.x % b;
    float v = 0.f;
    for (int j = 0; j < b; ++j) v = fmaf(t1[k * b + j], Cb[m * b + j], v);
    const float lam = (2.0f - 2.0f * cosf(pi * k / a)) + (2.0f - 2.0f * cosf(pi * m / b));
    t2[threadIdx.x] = (k == 0 && m == 0) ? 0.f : v / lam;
  }
  __syncthreads();
  if (threadIdx.x < n) {  // t1 = Ca^T t2
    const int i = threadIdx.x / b, m = threadIdx.x % b;
    float v = 0.f;
    for (int k = 0; k < a; ++k) v = fmaf(Ca[k * a + i], t2[k * b + m], v);
    t1[threadIdx.x] = v;
  }
  __syncthreads();
  if (threadIdx.x < n) {  // z = t1 Cb
    const int i = threadIdx.x / b, j = threadIdx.x % b;
    float v = 0.f;
    for (int m = 0; m < b; ++m) v = fmaf(t1[i * b + m], Cb[m * b + j], v);
    z[threadIdx.x] = v;
  }
  __syncthreads();
}

__global__ void __launch_bounds__(kTailThreads) k_b_tail(Ctx C, MG M, int bottom) {
  extern __shared__ float lds[];
  const int inst = blockIdx.x;
  if (inst_done(C, inst)) return;
  const Inst I = C.I[inst];
  const int T = I.T, L = I.L;
  // LDS: z_T..z_{L-1}, then r_{T+1}..r_{L-1}; r_T stays in global memory
  int ztot = 0;
  for (int l = T; l < L; ++l) ztot += n_of(I, l, 0) * n_of(I, l, 1);
  auto zl = [&](int l) {
    int o = 0;
    for (int k = T; k < l; ++k) o += n_of(I, k, 0) * n_of(I, k, 1);
    return lds + o;
  };
  auto rl = [&](int l) {
    int o = ztot;
    for (int k = T + 1; k < l; ++k) o += n_of(I, k, 0) * n_of(I, k, 1);
    return lds + o;
  };
  float* dct_w = rl(L);  // 4 x 256 floats after the r levels
  const float* rT = M.r + I.om[T];
  const float mT = T == 0 ? mean_of(C, inst) : 0.f;
  const int aT = n_of(I, T, 0), bT = n_of(I, T, 1);
  float rr[2][kTailQ];
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int q = 0; q < kTailQ; ++q) {
      const int t = threadIdx.x + q * kTailThreads, hb = bT >> 1;
      rr[c][q] = 0.f;
      if (t < aT * hb) {
        int i, jh;
        rc_of(t, hb, i, jh);
        rr[c][q] = rT[i * bT + 2 * jh + ((i + c) & 1)] - mT;
      }
    }
  for (int t = threadIdx.x; t < ztot; t += kTailThreads) lds[t] = 0.f;
  __syncthreads();
  // down: level T from registers, the others from LDS
  for (int l = T; l < L; ++l) {
    const int a = n_of(I, l, 0), b = n_of(I, l, 1);
    float* z = zl(l);
    if (l == L - 1 && bottom < 0 && l > T && a <= kBottomMax && b <= kBottomMax) {
      dct_bottom(z, rl(l), a, b, dct_w);
      break;
    }
    if (l == L - 1 && a * b <= 1024) {
      const float* r = l == T ? nullptr : rl(l);
      if (l == T) {
        // a single-level tail: the bottom sweeps over the register values
        for (int s = 0; s < (bottom < 0 ? 50 : bottom); ++s) {
          reg_half(z, rr[0], a, b, 0);
          __syncthreads();
          reg_half(z, rr[1], a, b, 1);
          __syncthreads();
        }
      } else {
        if (threadIdx.x < 64)
          for (int s = 0; s < (bottom < 0 ? 50 : bottom); ++s) {
            wave_half(z, r, a, b, 0);
            wave_half(z, r, a, b, 1);
          }
        __syncthreads();
      }
      break;
    }
    const int sweeps = l == L - 1 ? (bottom < 0 ? 50 : bottom) : 2;
    for (int s = 0; s < sweeps; ++s) {
      if (l == T) reg_half(z, rr[0], a, b, 0);
      else lds_half<2>(z, rl(l), a, b, 0);
      __syncthreads();
      if (l == T) reg_half(z, rr[1], a, b, 1);
      else lds_half<2>(z, rl(l), a, b, 1);
      __syncthreads();
    }
    if (l == L - 1) break;
    if (l == T)
      lds_restrict<kTailQ / 2>(z, [&](int id) { return rT[id] - mT; }, rl(l + 1), a, b);
    else {
      const float* r = rl(l);
      lds_restrict<1>(z, [&](int id) { return r[id]; }, rl(l + 1), a, b);
    }
    __syncthreads();
  }
  // up
  for (int l = L - 2; l >= T; --l) {
    const int a = n_of(I, l, 0), b = n_of(I, l, 1);
    float* z = zl(l);
    if (l == T) lds_prolong<2 * kTailQ>(z, zl(l + 1), a, b);
    else lds_prolong<4>(z, zl(l + 1), a, b);
    __syncthreads();
    for (int s = 0; s < 2; ++s) {
      if (l == T) reg_half(z, rr[1], a, b, 1);
      else lds_half<2>(z, rl(l), a, b, 1);
      __syncthreads();
      if (l == T) reg_half(z, rr[0], a, b, 0);
      else lds_half<2>(z, rl(l), a, b, 0);
      __syncthreads();
    }
  }
  const float* zT = zl(T);
  for (int t = threadIdx.x; t < aT * bT; t += kTailThreads) M.z[I.om[T] + t] = zT[t];
}

// z.r (r with its deferred mean) -> FIN_ZR_INIT (p = z) or FIN_BETA.
__global__ void __launch_bounds__(BS) k_b_zr(Ctx C, Dom D, MG M, int init) {
  const Where w = where(D);
  if (inst_done(C, w.i)) return;
  const Inst I = C.I[w.i];
  const int N = (int)cells(I);
  const float m = mean_of(C, w.i);
  double zr = 0;
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    const int t = w.t0 + u * BS;
    if (t < N) {
      const float z = z_final(M, I)[t];
      zr += (double)z * (double)(M.r[I.om[0] + t] - m);
      if (init) M.p[I.oc + t] = z;
    }
  }
  inst_reduce<BS>(C, D, w.i, init ? FIN_ZR_INIT : FIN_BETA, zr, 0.0);
}

// p = z + beta p
__global__ void __launch_bounds__(BS) k_b_p(Ctx C, Dom D, MG M) {
  const Where w = where(D);
  if (inst_done(C, w.i)) return;
  const Inst I = C.I[w.i];
  const int N = (int)cells(I);
  const float beta = (float)C.sc[(int64_t)w.i * S_COUNT + S_BETA];
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    const int t = w.t0 + u * BS;
    if (t < N) {
      float* p = M.p + I.oc + t;
      *p = fmaf(beta, *p, z_final(M, I)[t]);
    }
  }
}

// Ap = A p, partial p.Ap -> alpha.
__global__ void __launch_bounds__(BS) k_b_Ap(Ctx C, Dom D, MG M) {
  const Where w = where(D);
  if (inst_done(C, w.i)) return;
  const Inst I = C.I[w.i];
  const int a = I.n0, b = I.n1, N = a * b;
  const float* p = M.p + I.oc;
  double pap = 0;
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    const int t = w.t0 + u * BS;
    if (t < N) {
      int i, j;
      rc_of(t, b, i, j);
      const float ap = (float)nnbr(i, j, a, b) * p[t] - nsum(p, i, j, a, b);
      M.Ap[I.oc + t] = ap;
      pap += (double)p[t] * ap;
    }
  }
  inst_reduce<BS>(C, D, w.i, FIN_ALPHA, pap, 0.0);
}

// x += alpha p, r = (r - mean) - alpha Ap; partial r.r and sum r.
__global__ void __launch_bounds__(BS) k_b_update(Ctx C, Dom D, MG M) {
  const Where w = where(D);
  if (inst_done(C, w.i)) return;
  const Inst I = C.I[w.i];
  const int N = (int)cells(I);
  const double alpha = C.sc[(int64_t)w.i * S_COUNT + S_ALPHA];
  const float m = mean_of(C, w.i);
  double rr = 0, rs = 0;
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    const int t = w.t0 + u * BS;
    if (t < N) {
      const int64_t o = I.oc + t;
      M.x[o] += alpha * (double)M.p[o];
      float* r = M.r + I.om[0] + t;
      const float v = fmaf(-(float)alpha, M.Ap[o], *r - m);
      *r = v;
      rr += (double)v * v;
      rs += v;
    }
  }
  inst_reduce<BS>(C, D, w.i, FIN_RR, rr, rs);
}

// u_a -= x[I] - x[I - e_a] on interior faces, wall faces 0 (one thread per face).
__global__ void __launch_bounds__(BS) k_b_subgrad(Ctx C, Dom D, float* u, MG M) {
  const Where w = where(D);
  const Inst I = C.I[w.i];
  const int N = npoints(I, false);
  const double* x = M.x + I.oc;
#pragma unroll
  for (int uu = 0; uu < PPT; ++uu) {
    const int t = w.t0 + uu * BS;
    if (t >= N) break;
    int k;
    const int grid = point_grid(I, t, k);
    float* c = mac_c(u, I, grid) + k;
    if (grid == 0) {
      int i, j;
      rc_of(k, I.n1, i, j);
      *c = (i > 0 && i < I.n0) ? (float)((double)*c - (x[i * I.n1 + j] - x[(i - 1) * I.n1 + j])) : 0.f;
    } else {
      int i, j;
      rc_of(k, I.n1 + 1, i, j);
      *c = (j > 0 && j < I.n1) ? (float)((double)*c - (x[i * I.n1 + j] - x[i * I.n1 + j - 1])) : 0.f;
    }
  }
}

// ----------------------------------------------------------------- engine ----
enum DomKind { D_FACES = 0, D_CELLS, D_ALL, D_TAILGRID, D_LV_TILES, D_COUNT = D_LV_TILES + kMaxLv };

struct Batch {
  int count = 0, R = 0, steps = 0, bottom = 50, bottom_exact = 1, chunk = 4, first_iters = 8, graphs = 1,
      tail_cells = 16384, tile = 32;
  double tol = 1e-6;
  int max_iters = 400;
  std::vector<Inst> hI;
  Inst* dI = nullptr;
  std::vector<std::vector<int>> dom_h;  // per kind: count+1 block prefix
  std::vector<int> bt_off;              // per kind: offset of its block table in bt_d
  int* dom_d = nullptr;                 // all prefix tables, D_COUNT x (count + 1)
  int2* bt_d = nullptr;                 // all block tables (instance, first block)
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
  float *mz = nullptr, *mzo = nullptr, *mr = nullptr, *mp = nullptr, *mAp = nullptr;
  double* mx = nullptr;
  double* sc = nullptr;
  int* fl = nullptr;
  double* part = nullptr;
  unsigned* ctr = nullptr;
  int* h_fl = nullptr;      // pinned mirror of the flags
  double* h_sc = nullptr;   // pinned mirror of the scalars
  hipStream_t cap = nullptr;
  hipEvent_t rdy = nullptr;  // polled for the flag readbacks (spin_sync)
  std::map<std::tuple<int, int, int>, hipGraphExec_t> graphs_c;
  int step = -1;            // forward: last completed step; adjoint: next step to undo + 1
  int projections = 0, max_seen_iters = 0, noconv = 0;
  std::vector<void*> allocs;

  Ctx ctx() const { return Ctx{dI, sc, fl, part, ctr, tol}; }
  Dom dom(int kind) const { return Dom{dom_d + (size_t)kind * (count + 1), bt_d + bt_off[kind], count}; }
  int blocks(int kind) const { return dom_h[kind][count]; }
  MG mg() const { return MG{mz, mzo, mr, mp, mAp, mx}; }
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
    B.dom_h[kind][i + 1] = B.dom_h[kind][i] + (int)((pts + BS * PPT - 1) / (BS * PPT));
  };
  for (int i = 0; i < B.count; ++i) {
    const Inst& I = B.hI[i];
    add(D_FACES, i, faces0(I) + faces1(I));
    add(D_CELLS, i, cells(I));
    add(D_ALL, i, faces0(I) + faces1(I) + cells(I));
    B.dom_h[D_TAILGRID][i + 1] = i + 1;
    for (int l = 0; l < kMaxLv; ++l) {
      // levels 0..T-1 run as batched tiled kernels: one block per tile
      const int TS = B.tile;
      const int tiles = l < I.T ? ((n_of(I, l, 0) + TS - 1) / TS) * ((n_of(I, l, 1) + TS - 1) / TS) : 0;
      B.dom_h[D_LV_TILES + l][i + 1] = B.dom_h[D_LV_TILES + l][i] + tiles;
    }
  }
  B.max_blocks = 0;
  for (auto& v : B.dom_h) B.max_blocks = std::max(B.max_blocks, v[B.count]);
  B.bt_off.assign(D_COUNT, 0);
  int tot = 0;
  for (int k = 0; k < D_COUNT; ++k) {
    B.bt_off[k] = tot;
    tot += B.dom_h[k][B.count];
  }
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
// One V-cycle on the recentred level-0 residual: tiled down legs on the
// global levels, the per-instance LDS tail, tiled up legs.
void enq_vcycle(Batch& B, hipStream_t st) {
  const Ctx c = B.ctx();
  const MG M = B.mg();
  int G = 0;
  for (auto& I : B.hI) G = std::max(G, I.T);
  for (int l = 0; l < G; ++l) {
    const int nb = B.blocks(D_LV_TILES + l);
    if (B.tile == 16) launch(st, k_b_down<16>, nb, c, B.dom(D_LV_TILES + l), M, l);
    else launch(st, k_b_down<32>, nb, c, B.dom(D_LV_TILES + l), M, l);
  }
  hipLaunchKernelGGL(k_b_tail, dim3(B.count), dim3(kTailThreads), B.tail_lds, st, c, M,
                     B.bottom_exact ? -1 : B.bottom);
  for (int l = G - 1; l >= 0; --l) {
    const int nb = B.blocks(D_LV_TILES + l);
    if (B.tile == 16) launch(st, k_b_up<16>, nb, c, B.dom(D_LV_TILES + l), M, l);
    else launch(st, k_b_up<32>, nb, c, B.dom(D_LV_TILES + l), M, l);
  }
}

void enq_pcg_iters(Batch& B, hipStream_t st, int iters) {
  const Ctx c = B.ctx();
  const MG M = B.mg();
  for (int it = 0; it < iters; ++it) {
    launch(st, k_b_Ap, B.blocks(D_CELLS), c, B.dom(D_CELLS), M);
    launch(st, k_b_update, B.blocks(D_CELLS), c, B.dom(D_CELLS), M);
    enq_vcycle(B, st);
    launch(st, k_b_zr, B.blocks(D_CELLS), c, B.dom(D_CELLS), M, 0);
    launch(st, k_b_p, B.blocks(D_CELLS), c, B.dom(D_CELLS), M);
  }
}

// Projection start: b = -div u, recentre, first V-cycle, z.r, p = z, then
// `iters` PCG iterations.
void enq_pcg_begin(Batch& B, hipStream_t st, const float* u, int iters) {
  const Ctx c = B.ctx();
  const MG M = B.mg();
  launch(st, k_b_div, B.blocks(D_CELLS), c, B.dom(D_CELLS), u, M);
  launch(st, k_b_recenter, B.blocks(D_CELLS), c, B.dom(D_CELLS), M);
  enq_vcycle(B, st);
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
    if (e == hipSuccess) e = spin_sync(st, B.rdy);
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
  tail_cells = std::min(tail_cells, kTailMax);
  int64_t om = 0;
  size_t lds_max = 0;
  for (auto& I : B.hI) {
    const int n0 = I.n0, n1 = I.n1;
    int T = I.L - 1;
    for (int l = I.L - 1; l >= 0; --l) {
      size_t bytes = 4 * 256 * sizeof(float);
      for (int k = l; k < I.L; ++k) bytes += sizeof(float) * (size_t)(n0 >> k) * (n1 >> k) * (k == l ? 1 : 2);
      if ((int64_t)(n0 >> l) * (n1 >> l) > tail_cells || bytes > (size_t)kLdsBytes) break;
      T = l;
    }
    I.T = T;
    size_t bytes = 4 * 256 * sizeof(float);  // exact-bottom scratch
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
  for (float** p : {&B.mz, &B.mzo, &B.mr})
    if (*p) {
      (void)hipFree(*p);
      *p = nullptr;
    }
  if (B.part) {
    (void)hipFree(B.part);
    B.part = nullptr;
  }
  hipError_t e = hipMalloc((void**)&B.mz, sizeof(float) * std::max<int64_t>(om, 1));
  if (e == hipSuccess) e = hipMalloc((void**)&B.mzo, sizeof(float) * std::max<int64_t>(om, 1));
  if (e == hipSuccess) e = hipMalloc((void**)&B.mr, sizeof(float) * std::max<int64_t>(om, 1));
  if (e == hipSuccess) e = hipMalloc((void**)&B.part, sizeof(double) * 2 * std::max(B.max_blocks, 1));
  std::vector<int> flat;
  for (auto& v : B.dom_h) flat.insert(flat.end(), v.begin(), v.end());
  if (e == hipSuccess) e = hipMemcpy(B.dom_d, flat.data(), sizeof(int) * flat.size(), hipMemcpyHostToDevice);
  std::vector<int2> bt;
  for (int k = 0; k < D_COUNT; ++k)
    for (int i = 0; i < B.count; ++i)
      for (int b = B.dom_h[k][i]; b < B.dom_h[k][i + 1]; ++b) bt.push_back(make_int2(i, B.dom_h[k][i]));
  if (B.bt_d) {
    (void)hipFree(B.bt_d);
    B.bt_d = nullptr;
  }
  if (e == hipSuccess) e = hipMalloc((void**)&B.bt_d, sizeof(int2) * std::max<size_t>(bt.size(), 1));
  if (e == hipSuccess) e = hipMemcpy(B.bt_d, bt.data(), sizeof(int2) * bt.size(), hipMemcpyHostToDevice);
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
    if (n0 < 8 || n1 < 8 || n0 > 2048 || n1 > 2048) {
      delete h;
      return set_error(DFM_E_ARG, "dfm_batch_create: each extent must be 8..2048");
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
    if (e == hipSuccess) e = hipEventCreateWithFlags(&B.rdy, hipEventDisableTiming);
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
  for (void* p : {(void*)B.mz, (void*)B.mzo, (void*)B.mr, (void*)B.part, (void*)B.bt_d})
    if (p) (void)hipFree(p);
  if (B.h_fl) (void)hipHostFree(B.h_fl);
  if (B.h_sc) (void)hipHostFree(B.h_sc);
  if (B.cap) (void)hipStreamDestroy(B.cap);
  if (B.rdy) (void)hipEventDestroy(B.rdy);
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
    case DFM_BATCH_BOTTOM_EXACT: B.bottom_exact = value ? 1 : 0; break;
    case DFM_BATCH_TILE:
      if (value != 16 && value != 32) return set_error(DFM_E_ARG, "dfm_batch_set_option: tile must be 16 or 32");
      B.tile = value;
      break;
    default: return set_error(DFM_E_ARG, "dfm_batch_set_option: unknown option");
  }
  // captured graphs hold the old schedule
  (void)hipDeviceSynchronize();
  for (auto& kv : B.graphs_c) (void)hipGraphExecDestroy(kv.second);
  B.graphs_c.clear();
  if (option == DFM_BATCH_TAIL_CELLS) return plan_mg(B, value);
  if (option == DFM_BATCH_TILE) return plan_mg(B, B.tail_cells);
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
  if (e == hipSuccess) e = spin_sync(st, B.rdy);
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
