// Flow-map kernels: march, pullback, BFECC, semi-Lagrangian advection,
// accumulation and the adjoint sources, plus their C-ABI entry points.
#include "dfm_common.hpp"
#include "dfm_points.hpp"
#include "dfm_tiled_sources.hpp"
#include "dfm_internal.hpp"
#include "../../include/dfm.h"

#include <atomic>
#include <mutex>
#include <type_traits>

using namespace dfm;

// BFECC error field of dfm_bfecc_u_noalloc: a file-scope device array, one
// copy per GPU, held until the process ends (DFM_NOALLOC_MAX_VALUES floats,
// 32 MiB).  Host code never passes this name to a kernel: it is a host-side
// stand-in, and the device address of each GPU's copy comes from
// hipGetSymbolAddress (noalloc_scratch below).
static __device__ __align__(256) float scratch[DFM_NOALLOC_MAX_VALUES];

namespace {

constexpr int kBlock = 256;
constexpr int kMarchCapMax = 16;  // largest workgroups-per-CU cap of dfm_march_capped

// Up to four point sets (3 face grids + cells) addressed by one flat index.
struct Sets {
  int nset;
  int grid[4];     // face axis 0..2 or 3 = cells
  int64_t cnt[4];  // points per set
  int64_t off[5];  // prefix offsets
  DFM_HD int find(int64_t gid) const {
    int s = 0;
#pragma unroll
    for (int t = 1; t < 4; ++t)
      if (t < nset && gid >= off[t]) s = t;
    return s;
  }
};

inline void sets_add(Sets& S, const Geo& g, int grid) {
  int k = S.nset++;
  S.grid[k] = grid;
  S.cnt[k] = count_of_host(g, grid);
  S.off[k + 1] = S.off[k] + S.cnt[k];
}

inline Sets sets_init() {
  Sets S;
  S.nset = 0;
  S.off[0] = 0;
  for (int k = 0; k < 4; ++k) {
    S.grid[k] = 0;
    S.cnt[k] = 0;
    S.off[k + 1] = 0;
  }
  return S;
}

inline int64_t blocks_for(int64_t n) { return (n + kBlock - 1) / kBlock; }

// ------------------------------------------------------------- identity ----
struct MapPtrs {
  float* psi[4];  // per set
  float* T[4];
};

__global__ void __launch_bounds__(kBlock) k_reset(Geo g, Sets S, MapPtrs M) {
  int64_t gid = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (gid >= S.off[S.nset]) return;
  int s = S.find(gid);
  int64_t i = gid - S.off[s];
  int grid = S.grid[s];
  int sh[3];
  shape_of(g, grid, sh);
  int I, J, K;
  unlin(sh, i, I, J, K);
  float p[3];
  node_pos(g, grid, I, J, K, p);
  int64_t cnt = S.cnt[s];
  for (int c = 0; c < g.dim; ++c) {
    M.psi[s][c * cnt + i] = p[c];
    if (M.T[s]) M.T[s][c * cnt + i] = (c == grid) ? 1.0f : 0.0f;
  }
}

// ----------------------------------------------------------------- march ----
template <int DIM, int ORDER>
__global__ void __launch_bounds__(kBlock) k_march(Geo g, Mac u, float dt, Sets S, MapPtrs M) {
  int64_t gid = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (gid >= S.off[S.nset]) return;
  int s = S.find(gid);
  int64_t i = gid - S.off[s];
  int64_t cnt = S.cnt[s];
  float* psi = M.psi[s];
  float* T = M.T[s];
  float p[3] = {0.f, 0.f, 0.f}, t[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < DIM; ++c) p[c] = psi[c * cnt + i];
  if (T) {
#pragma unroll
    for (int c = 0; c < DIM; ++c) t[c] = T[c * cnt + i];
    rk_point<DIM, ORDER, true>(g, u, dt, p, t);
#pragma unroll
    for (int c = 0; c < DIM; ++c) T[c * cnt + i] = t[c];
  } else {
    rk_point<DIM, ORDER, false>(g, u, dt, p, t);
  }
#pragma unroll
  for (int c = 0; c < DIM; ++c) psi[c * cnt + i] = p[c];
}

// -------------------------------------------------------------- pullback ----
struct Epi {
  float alpha, beta;
  const float* add[4];
  float* out[4];
};

template <int DIM, int K>
__global__ void __launch_bounds__(kBlock) k_pullback_u(Geo g, Mac src, Sets S, MapPtrs M, Epi E) {
  int64_t gid = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (gid >= S.off[S.nset]) return;
  int s = S.find(gid);
  int64_t i = gid - S.off[s];
  int64_t cnt = S.cnt[s];
  float p[3] = {0.f, 0.f, 0.f}, v[3];
#pragma unroll
  for (int c = 0; c < DIM; ++c) p[c] = M.psi[s][c * cnt + i];
  mac_eval<false, DIM, K>(g, src, p, v, nullptr);
  float d = 0.f;
#pragma unroll
  for (int c = 0; c < DIM; ++c) d = fmaf(M.T[s][c * cnt + i], v[c], d);
  float r = E.alpha * d;
  if (E.beta != 0.f) r = fmaf(E.beta, E.add[s][i], r);
  E.out[s][i] = r;
}

template <int DIM, int K>
__global__ void __launch_bounds__(kBlock) k_pullback_s(Geo g, const float* __restrict__ src,
                                                       const float* __restrict__ psi, float alpha, float beta,
                                                       const float* add, float* out, int64_t cnt) {
  int64_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= cnt) return;
  float p[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < DIM; ++c) p[c] = psi[c * cnt + i];
  int S[3];
  shape_of(g, 3, S);
  const float bl[3] = {0.5f, 0.5f, 0.5f};
  float v = interp<K, false, DIM>(src, S, bl, p, g.inv_dx, nullptr);
  float r = alpha * v;
  if (beta != 0.f) r = fmaf(beta, add[i], r);
  out[i] = r;
}

// ---------------------------------------------------------- accumulate -----
struct AccPtrs {
  float* init[4];
  float* h1[4];
  float* h2[4];
};

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_accumulate(Geo g, Mac f, Mac f2, bool has_f2, Sets S, MapPtrs M,
                                                       AccPtrs A, int ind, float dt, float dt2) {
  int64_t gid = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (gid >= S.off[S.nset]) return;
  int s = S.find(gid);
  int64_t i = gid - S.off[s];
  int64_t cnt = S.cnt[s];
  float p[3] = {0.f, 0.f, 0.f}, T[3] = {0.f, 0.f, 0.f}, v[3];
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
    p[c] = M.psi[s][c * cnt + i];
    T[c] = M.T[s][c * cnt + i];
  }
  mac_eval<false, DIM>(g, f, p, v, nullptr);
  float gv = 0.f;
#pragma unroll
  for (int c = 0; c < DIM; ++c) gv = fmaf(T[c], v[c], gv);
  float acc = A.init[s][i];
  float hp = A.h1[s][i];
  if (ind == 0) {
    acc += gv * dt;
  } else if (ind == 1) {
    acc += (23.0f * gv - 16.0f * gv + 5.0f * hp) / 12.0f * dt;
  } else {
    acc += (23.0f * gv - 16.0f * hp + 5.0f * A.h2[s][i]) / 12.0f * dt;
  }
  if (has_f2) {
    mac_eval<false, DIM>(g, f2, p, v, nullptr);
    float g2 = 0.f;
#pragma unroll
    for (int c = 0; c < DIM; ++c) g2 = fmaf(T[c], v[c], g2);
    acc += g2 * dt2;
  }
  A.init[s][i] = acc;
  A.h2[s][i] = hp;
  A.h1[s][i] = gv;
}

// ------------------------------------------------ semi-Lagrangian advect ----
constexpr int kMaxVec = 3;
constexpr int kMaxScal = 8;

struct SLArgs {
  int nvec, nscal;
  Mac src[kMaxVec];
  float* dst[kMaxVec][3];
  const float* ssrc[kMaxScal];
  float* sdst[kMaxScal];
};

// The back-trace samples the velocity quadratically; the sources are sampled
// with the spline of order KS.
template <int DIM, int ORDER, int KS>
__global__ void __launch_bounds__(kBlock) k_sl_advect(Geo g, Mac vel, float dt, Sets S, SLArgs A) {
  int64_t gid = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (gid >= S.off[S.nset]) return;
  int s = S.find(gid);
  int64_t i = gid - S.off[s];
  int grid = S.grid[s];
  int sh[3];
  shape_of(g, grid, sh);
  int I, J, K;
  unlin(sh, i, I, J, K);
  float x[3] = {0.f, 0.f, 0.f}, p[3] = {0.f, 0.f, 0.f};
  node_pos(g, grid, I, J, K, x);
  backtrace<DIM, ORDER>(g, vel, dt, x, p);
  float bl[3] = {0.5f, 0.5f, 0.5f};
  if (grid < 3) {
    bl[grid] = 0.0f;
    for (int m = 0; m < A.nvec; ++m) A.dst[m][grid][i] = interp<KS, false, DIM>(A.src[m].c[grid], sh, bl, p, g.inv_dx, nullptr);
  } else {
    for (int m = 0; m < A.nscal; ++m) A.sdst[m][i] = interp<KS, false, DIM>(A.ssrc[m], sh, bl, p, g.inv_dx, nullptr);
  }
}

// ------------------------------------------------------- adjoint sources ----
template <int DIM, int K>
__global__ void __launch_bounds__(kBlock) k_adj_conv(Geo g, Mac u, Mac w, Sets S, MapPtrs R) {
  int64_t gid = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (gid >= S.off[S.nset]) return;
  int s = S.find(gid);
  int64_t i = gid - S.off[s];
  int grid = S.grid[s];
  int sh[3];
  shape_of(g, grid, sh);
  int I, J, Kk;
  unlin(sh, i, I, J, Kk);
  float x[3] = {0.f, 0.f, 0.f};
  node_pos(g, grid, I, J, Kk, x);
  float r = 0.f;
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
    int Sc[3];
    shape_of(g, c, Sc);
    float bl[3] = {0.5f, 0.5f, 0.5f};
    bl[c] = 0.0f;
    float d, v;
    dual_interp<K, DIM>(u.c[c], w.c[c], Sc, bl, x, g.inv_dx, grid, d, v);
    r = fmaf(d, v, r);
  }
  R.psi[s][i] = r;
}

struct ScalPairs {
  int n;
  const float* s[kMaxScal];
  const float* lam[kMaxScal];
};

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_adj_scalar(Geo g, ScalPairs P, Sets S, MapPtrs R) {
  int64_t gid = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (gid >= S.off[S.nset]) return;
  int s = S.find(gid);
  int64_t i = gid - S.off[s];
  int grid = S.grid[s];
  int sh[3];
  shape_of(g, grid, sh);
  int I, J, K;
  unlin(sh, i, I, J, K);
  float x[3] = {0.f, 0.f, 0.f};
  node_pos(g, grid, I, J, K, x);
  int Sc[3];
  shape_of(g, 3, Sc);
  const float bl[3] = {0.5f, 0.5f, 0.5f};
  float r = 0.f;
  for (int m = 0; m < P.n; ++m) {
    float d, v;
    dual_interp<3, DIM>(P.s[m], P.lam[m], Sc, bl, x, g.inv_dx, grid, d, v);
    r = fmaf(d, v, r);
  }
  R.psi[s][i] = r;
}

// ---------------------------------------------------------------- axpy -----
struct Axpy {
  const float* x[4];
  const float* y[4];
  const float* z[4];
  float* out[4];
  float alpha, beta;
};

__global__ void __launch_bounds__(kBlock) k_axpy(Sets S, Axpy A) {
  int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (gid >= S.off[S.nset]) return;
  int s = S.find(gid);
  int64_t i = gid - S.off[s];
  float r = fmaf(A.alpha, A.y[s][i], A.x[s][i]);
  if (A.z[s]) r = fmaf(A.beta, A.z[s][i], r);
  A.out[s][i] = r;
}

// ------------------------------------------------ vector-Jacobian products ----
// Transposes of the stages that are linear in a field (the pullbacks and the
// semi-Lagrangian sample), with the gradients w.r.t. the map or the
// advecting velocity, and the reverse of the march.  The transpose of a
// B-spline sample scatters the incoming value with the sample's own tap
// weights to the sample's own (clamped) taps by FP32 atomic adds, so the sums
// are order-nondeterministic at rounding level.  A sample's derivative
// w.r.t. its point is zero along an axis whose sample coordinate is clamped
// at the wall band (the forward's value is constant there).

// g[taps of the sample at p] += coef * weight (transpose of interp<K, false>).
template <int K, int DIM>
DFM_HD void scatter_sample(float* __restrict__ g, const int S[3], const float bl[3], const float p[3], float inv_dx,
                           float coef) {
  constexpr int NT = Spline<K>::NT;
  Axis<K, false> ax, ay, az;
  ax.setup(p[0] * inv_dx, bl[0], S[0]);
  ay.setup(p[1] * inv_dx, bl[1], S[1]);
  if (DIM == 3) az.setup(p[2] * inv_dx, bl[2], S[2]);
#pragma unroll
  for (int a = 0; a < NT; ++a) {
    const float ca = coef * ax.w[a];
#pragma unroll
    for (int b = 0; b < NT; ++b) {
      const float cb = ca * ay.w[b];
      const int64_t row = (int64_t)ax.idx[a] * S[1] + ay.idx[b];
      if (DIM == 2) {
        atomicAdd(g + row, cb);
      } else {
#pragma unroll
        for (int c = 0; c < NT; ++c) atomicAdd(g + row * S[2] + az.idx[c], cb * az.w[c]);
      }
    }
  }
}

// Gradient w.r.t. the point of the order-K sample of f at p, zero along
// clamped axes; returns the value.
template <int K, int DIM>
DFM_HD float sample_dp(const float* __restrict__ f, const int S[3], const float bl[3], const float p[3], float inv_dx,
                       float gr[3]) {
  const float v = interp<K, true, DIM>(f, S, bl, p, inv_dx, gr);
#pragma unroll
  for (int c = 0; c < DIM; ++c) gr[c] *= clamp_live<K>(p[c] * inv_dx, bl[c], S[c]);
  return v;
}

// Reverse of the march (k_march) for one point.  eval_vjp takes the incoming
// gradients of one velocity sample u(q) (gv) and of k = J(q) Tq (gk, so
// that the gradient of J[a][b] = du_a/dx_b is gk[a] Tq[b]), scatters them
// into the velocity's gradient gu, and adds their derivative w.r.t. q to gq
// (through the value's gradient and the quadratic spline's second
// derivative, weights {1, -2, 1}; zero along an axis whose sample coordinate
// is clamped at the wall band) and J^T gk to gTq.
template <int DIM>
DFM_HD void eval_vjp(const Geo& g, const Mac& u, const float q[3], const float gv[3], const float gk[3],
                     const float Tq[3], float* const gu[3], float gq[3], float gTq[3]) {
  constexpr float d2[3] = {1.0f, -2.0f, 1.0f};
#pragma unroll 1
  for (int a = 0; a < DIM; ++a) {
    int S[3];
    shape_of(g, a, S);
    float bl[3] = {0.5f, 0.5f, 0.5f};
    bl[a] = 0.0f;
    Axis<2, true> ax[3];
    float live[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      ax[c].setup(q[c] * g.inv_dx, bl[c], S[c]);
      live[c] = clamp_live<2>(q[c] * g.inv_dx, bl[c], S[c]);
    }
    float gJ[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < DIM; ++b) gJ[b] = gk[a] * Tq[b];
    float G[3] = {0.f, 0.f, 0.f}, H[3] = {0.f, 0.f, 0.f};
    const int NZ = DIM == 3 ? 3 : 1;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
          const float w0 = ax[0].w[i], w1 = ax[1].w[j], w2 = DIM == 3 ? ax[2].w[k] : 1.0f;
          const float e0 = ax[0].dw[i], e1 = ax[1].dw[j], e2 = DIM == 3 ? ax[2].dw[k] : 0.0f;
          const float h0 = d2[i], h1 = d2[j], h2 = DIM == 3 ? d2[k] : 0.0f;
          // first and second derivative weights of the tap (grid units)
          const float D[3] = {e0 * w1 * w2, w0 * e1 * w2, w0 * w1 * e2};
          const float Hm[3][3] = {{h0 * w1 * w2, e0 * e1 * w2, e0 * w1 * e2},
                                  {e0 * e1 * w2, w0 * h1 * w2, w0 * e1 * e2},
                                  {e0 * w1 * e2, w0 * e1 * e2, w0 * w1 * h2}};
          int64_t id = (int64_t)ax[0].idx[i] * S[1] + ax[1].idx[j];
          if (DIM == 3) id = id * S[2] + ax[2].idx[k];
          float coef = gv[a] * (w0 * w1 * w2);
#pragma unroll
          for (int b = 0; b < DIM; ++b) coef = fmaf(gJ[b] * g.inv_dx, D[b], coef);
          atomicAdd(gu[a] + id, coef);
          const float f = u.c[a][id];
#pragma unroll
          for (int c = 0; c < DIM; ++c) {
            G[c] = fmaf(f, D[c], G[c]);
            float hc = 0.f;
#pragma unroll
            for (int b = 0; b < DIM; ++b) hc = fmaf(gJ[b], Hm[c][b], hc);
            H[c] = fmaf(f, hc, H[c]);
          }
        }
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      const float J = G[c] * g.inv_dx;  // du_a / dx_c
      gq[c] += live[c] * (gv[a] * J + H[c] * g.inv_dx * g.inv_dx);
      gTq[c] = fmaf(J, gk[a], gTq[c]);
    }
  }
}

struct GuPtrs {
  float* c[3];
};

// k_pullback_u is out_a = alpha T_a . src(psi_a) (+ beta add_a).  Per face
// point: g_src_b += alpha T_a[b] g_out_a at psi_a; with GM.psi set, also
// g_T_a[b] += alpha g_out_a src_b(psi_a) and g_psi_a += alpha g_out_a
// sum_b T_a[b] grad src_b(psi_a).
template <int DIM, int K>
__global__ void __launch_bounds__(kBlock) k_pullback_u_vjp(Geo g, Mac gout, Mac src, Sets S, MapPtrs M, float alpha,
                                                          Epi G, MapPtrs GM) {
  int64_t gid = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (gid >= S.off[S.nset]) return;
  const int s = S.find(gid);
  const int64_t i = gid - S.off[s], cnt = S.cnt[s];
  const int a = S.grid[s];
  const bool want_map = GM.psi[s] != nullptr;
  float p[3] = {0.f, 0.f, 0.f}, gp[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < DIM; ++c) p[c] = M.psi[s][c * cnt + i];
  const float go = alpha * gout.c[a][i];
#pragma unroll 1
  for (int b = 0; b < DIM; ++b) {
    int Sb[3];
    shape_of(g, b, Sb);
    float bl[3] = {0.5f, 0.5f, 0.5f};
    bl[b] = 0.0f;
    const float Tb = M.T[s][b * cnt + i];
    scatter_sample<K, DIM>(G.out[b], Sb, bl, p, g.inv_dx, go * Tb);
    if (want_map) {
      float gr[3] = {0.f, 0.f, 0.f};
      const float v = sample_dp<K, DIM>(src.c[b], Sb, bl, p, g.inv_dx, gr);
      GM.T[s][b * cnt + i] += go * v;
#pragma unroll
      for (int c = 0; c < DIM; ++c) gp[c] = fmaf(go * Tb, gr[c], gp[c]);
    }
  }
  if (want_map) {
#pragma unroll
    for (int c = 0; c < DIM; ++c) GM.psi[s][c * cnt + i] += gp[c];
  }
}

// k_pullback_s is out = alpha src(psi) (+ beta add): g_src += alpha g_out at
// psi; with gpsi set, also gpsi += alpha g_out grad src(psi).
template <int DIM, int K>
__global__ void __launch_bounds__(kBlock) k_pullback_s_vjp(Geo g, const float* __restrict__ gout,
                                                          const float* __restrict__ src,
                                                          const float* __restrict__ psi, float alpha,
                                                          float* __restrict__ gsrc, float* __restrict__ gpsi,
                                                          int64_t cnt) {
  int64_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i >= cnt) return;
  float p[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < DIM; ++c) p[c] = psi[c * cnt + i];
  int Sc[3];
  shape_of(g, 3, Sc);
  const float bl[3] = {0.5f, 0.5f, 0.5f};
  const float go = alpha * gout[i];
  scatter_sample<K, DIM>(gsrc, Sc, bl, p, g.inv_dx, go);
  if (gpsi) {
    float gr[3] = {0.f, 0.f, 0.f};
    sample_dp<K, DIM>(src, Sc, bl, p, g.inv_dx, gr);
#pragma unroll
    for (int c = 0; c < DIM; ++c) gpsi[c * cnt + i] += go * gr[c];
  }
}

struct SLVjpArgs {
  int nvec, nscal;
  Mac fsrc[kMaxVec];           // forward sources (read when the velocity gradient is wanted)
  Mac gdst[kMaxVec];           // incoming gradients of the destinations
  float* gsrc[kMaxVec][3];     // accumulated source gradients
  const float* fs[kMaxScal];
  const float* gds[kMaxScal];
  float* gss[kMaxScal];
  float* gvel[3];              // velocity gradient, or null
};

// Transpose of k_sl_advect (dst(x) = src(p(x)), p the back-trace of node x
// through vel): the same back-trace, each destination's incoming value
// scattered to its source at p; with the velocity gradient wanted, the
// gradient at p is run back through the RK stages of the back-trace.
template <int DIM, int ORDER, int KS>
__global__ void __launch_bounds__(kBlock) k_sl_advect_vjp(Geo g, Mac vel, float dt, Sets S, SLVjpArgs A) {
  int64_t gid = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (gid >= S.off[S.nset]) return;
  int s = S.find(gid);
  int64_t i = gid - S.off[s];
  int grid = S.grid[s];
  int sh[3];
  shape_of(g, grid, sh);
  int I, J, K;
  unlin(sh, i, I, J, K);
  float x[3] = {0.f, 0.f, 0.f}, p[3] = {0.f, 0.f, 0.f};
  node_pos(g, grid, I, J, K, x);
  backtrace<DIM, ORDER>(g, vel, dt, x, p);
  float bl[3] = {0.5f, 0.5f, 0.5f};
  if (grid < 3) bl[grid] = 0.0f;
  const bool want_vel = A.gvel[0] != nullptr;
  float gp[3] = {0.f, 0.f, 0.f};
  const int nm = grid < 3 ? A.nvec : A.nscal;
  for (int m = 0; m < nm; ++m) {
    const float gi = grid < 3 ? A.gdst[m].c[grid][i] : A.gds[m][i];
    float* dst = grid < 3 ? A.gsrc[m][grid] : A.gss[m];
    scatter_sample<KS, DIM>(dst, sh, bl, p, g.inv_dx, gi);
    if (want_vel) {
      float gr[3] = {0.f, 0.f, 0.f};
      sample_dp<KS, DIM>(grid < 3 ? A.fsrc[m].c[grid] : A.fs[m], sh, bl, p, g.inv_dx, gr);
#pragma unroll
      for (int c = 0; c < DIM; ++c) gp[c] = fmaf(gi, gr[c], gp[c]);
    }
  }
  if (!want_vel) return;
  // back-trace stage points (as backtrace) and their reverse
  constexpr int NS = ORDER == 2 ? 2 : 4;
  float q[NS][3], v[3];
  const float step[3] = {-0.5f * dt, -0.5f * dt, -dt};
#pragma unroll
  for (int c = 0; c < 3; ++c) q[0][c] = x[c];
#pragma unroll
  for (int st = 0; st + 1 < NS; ++st) {
    mac_eval<false, DIM>(g, vel, q[st], v, nullptr);
#pragma unroll
    for (int c = 0; c < 3; ++c) q[st + 1][c] = c < DIM ? x[c] + step[st] * v[c] : 0.f;
  }
  float gv[NS][3];
  const float wt4[4] = {1.f / 6.f, 2.f / 6.f, 2.f / 6.f, 1.f / 6.f};
#pragma unroll
  for (int st = 0; st < NS; ++st)
#pragma unroll
    for (int c = 0; c < 3; ++c) gv[st][c] = -dt * (ORDER == 2 ? (st == NS - 1 ? 1.0f : 0.0f) : wt4[st]) * gp[c];
  const float zero[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int st = NS - 1; st >= 0; --st) {
    float gq[3] = {0.f, 0.f, 0.f}, unused[3] = {0.f, 0.f, 0.f};
    eval_vjp<DIM>(g, vel, q[st], gv[st], zero, zero, A.gvel, gq, unused);
    if (st > 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) gv[st - 1][c] += step[st - 1] * gq[c];
    }
  }
}

// gu += the velocity part of the march's vector-Jacobian product; the
// incoming gradients (GM, w.r.t. the marched map) are replaced in place by
// those w.r.t. the map before the march (M0).  Each point recomputes its RK
// stage points from M0 and runs the stages backwards.
template <int DIM, int ORDER>
__global__ void __launch_bounds__(kBlock) k_march_vjp(Geo g, Mac u, float dt, Sets S, MapPtrs M0, MapPtrs GM,
                                                     GuPtrs gu) {
  int64_t gid = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (gid >= S.off[S.nset]) return;
  const int s = S.find(gid);
  const int64_t i = gid - S.off[s], cnt = S.cnt[s];
  const bool has_t = M0.T[s] != nullptr;
  constexpr int NS = ORDER == 2 ? 2 : 4;
  float q[NS][3], Tq[NS][3];
  float gp[3] = {0.f, 0.f, 0.f}, gT[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int st = 0; st < NS; ++st)
#pragma unroll
    for (int c = 0; c < 3; ++c) q[st][c] = Tq[st][c] = 0.f;
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
    q[0][c] = M0.psi[s][c * cnt + i];
    gp[c] = GM.psi[s][c * cnt + i];
    if (has_t) {
      Tq[0][c] = M0.T[s][c * cnt + i];
      gT[c] = GM.T[s][c * cnt + i];
    }
  }
  const float neg = -dt;
  // forward: the stage points and transported columns of rk_point
  float v[3], J[3][3];
  const float step[3] = {0.5f * neg, 0.5f * neg, neg};
#pragma unroll
  for (int st = 0; st + 1 < NS; ++st) {
    mac_eval<true, DIM>(g, u, q[st], v, J);
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      float k = 0.f;
#pragma unroll
      for (int b = 0; b < DIM; ++b) k = fmaf(J[a][b], Tq[st][b], k);
      q[st + 1][a] = q[0][a] + step[st] * v[a];
      Tq[st + 1][a] = Tq[0][a] + step[st] * k;
    }
  }
  // backward
  float gu_st[NS][3], gk_st[NS][3];
  const float wt4[4] = {1.f / 6.f, 2.f / 6.f, 2.f / 6.f, 1.f / 6.f};
#pragma unroll
  for (int st = 0; st < NS; ++st)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float wgt = ORDER == 2 ? (st == NS - 1 ? 1.0f : 0.0f) : wt4[st];
      gu_st[st][c] = neg * wgt * gp[c];
      gk_st[st][c] = has_t ? neg * wgt * gT[c] : 0.f;
    }
#pragma unroll
  for (int st = NS - 1; st >= 0; --st) {
    float gq[3] = {0.f, 0.f, 0.f}, gTq[3] = {0.f, 0.f, 0.f};
    eval_vjp<DIM>(g, u, q[st], gu_st[st], gk_st[st], Tq[st], gu.c, gq, gTq);
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      gp[c] += gq[c];
      if (has_t) gT[c] += gTq[c];
      if (st > 0) {
        gu_st[st - 1][c] += step[st - 1] * gq[c];
        if (has_t) gk_st[st - 1][c] += step[st - 1] * gTq[c];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
    GM.psi[s][c * cnt + i] = gp[c];
    if (has_t) GM.T[s][c * cnt + i] = gT[c];
  }
}

// ------------------------------------------------------------ host glue ----
int check_grid(const dfm_grid* g, Geo& G) {
  if (!g) return set_error(DFM_E_NULL, "grid is NULL");
  if (g->dim != 2 && g->dim != 3) return set_error(DFM_E_ARG, "dim must be 2 or 3");
  for (int a = 0; a < 3; ++a) {
    int want_min = a < g->dim ? 4 : 1;
    if (g->n[a] < want_min) return set_error(DFM_E_ARG, "grid too small (need >= 4 cells per axis)");
    if (a >= g->dim && g->n[a] != 1) return set_error(DFM_E_ARG, "2D grid must have n[2] == 1");
  }
  if (!(g->dx > 0.f)) return set_error(DFM_E_ARG, "dx must be positive");
  // The samplers address a field with 32-bit offsets and 24-bit multiplies:
  // every face or cell grid holds fewer than 2^30 values; the multiplied
  // operands (a 3D (x, y) plane, or in 2D each extent) stay below 2^24.
  const int64_t plane = (int64_t)(g->n[0] + 1) * (g->n[1] + 1);
  const int64_t vol = g->dim == 3 ? plane * (g->n[2] + 1) : plane;
  const int64_t lim24 = int64_t(1) << 24;
  if (vol >= (int64_t(1) << 30) || (g->dim == 3 && plane >= lim24) ||
      (g->dim == 2 && (g->n[0] + 1 >= lim24 || g->n[1] + 1 >= lim24)))
    return set_error(DFM_E_ARG,
                     "grid too large (a face grid must hold fewer than 2^30 values, a 3D plane and each 2D extent "
                     "fewer than 2^24)");
  G.dim = g->dim;
  for (int a = 0; a < 3; ++a) G.n[a] = g->n[a];
  G.dx = g->dx;
  G.inv_dx = 1.0f / g->dx;
  return DFM_OK;
}

int check_mac(const Geo& G, const dfm_mac* m, const char* what) {
  if (!m) return set_error(DFM_E_NULL, what);
  for (int a = 0; a < G.dim; ++a)
    if (!m->c[a]) return set_error(DFM_E_NULL, what);
  return DFM_OK;
}

Mac to_mac(const dfm_mac* m) {
  Mac r;
  for (int a = 0; a < 3; ++a) r.c[a] = m ? m->c[a] : nullptr;
  return r;
}

inline hipStream_t S_(void* s) { return reinterpret_cast<hipStream_t>(s); }

int launch_check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(DFM_E_HIP, what, hipGetErrorString(e));
  return DFM_OK;
}

// face sets of a MAC field, in axis order
Sets face_sets(const Geo& G) {
  Sets S = sets_init();
  for (int a = 0; a < G.dim; ++a) sets_add(S, G, a);
  return S;
}

}  // namespace

#define DFM_DISPATCH_DIM(G, BODY2, BODY3) \
  do {                                    \
    if ((G).dim == 2) {                   \
      BODY2;                              \
    } else {                              \
      BODY3;                              \
    }                                     \
  } while (0)

// Calls f(std::integral_constant<int, K>) for the sampling spline of order
// kern: 1 (linear, N_1), 2 (quadratic, N_2, the simulator's) or 3 (cubic, N_3).
int check_kern(int kern) {
  if (kern < 1 || kern > 3)
    return set_error(DFM_E_ARG, "interpolation kernel must be 1 (linear), 2 (quadratic) or 3 (cubic)");
  return DFM_OK;
}
template <class F> int with_kern(int kern, F&& f) {
  switch (kern) {
    case 1: f(std::integral_constant<int, 1>{}); return DFM_OK;
    case 2: f(std::integral_constant<int, 2>{}); return DFM_OK;
    case 3: f(std::integral_constant<int, 3>{}); return DFM_OK;
    default: return check_kern(kern);
  }
}

extern "C" int dfm_version(void) { return 1; }

extern "C" int dfm_reset_identity(const dfm_grid* g, const dfm_flowmap* m, void* stream) {
  Geo G;
  int rc = check_grid(g, G);
  if (rc) return rc;
  if (!m) return set_error(DFM_E_NULL, "flowmap is NULL");
  Sets S = sets_init();
  MapPtrs M{};
  for (int a = 0; a < G.dim; ++a)
    if (m->psi_f[a]) {
      M.psi[S.nset] = m->psi_f[a];
      M.T[S.nset] = m->T_f[a];
      sets_add(S, G, a);
    }
  if (m->psi) {
    M.psi[S.nset] = m->psi;
    M.T[S.nset] = nullptr;
    sets_add(S, G, 3);
  }
  if (S.nset == 0) return DFM_OK;
  int64_t total = S.off[S.nset];
  hipLaunchKernelGGL(k_reset, dim3(blocks_for(total)), dim3(kBlock), 0, S_(stream), G, S, M);
  return launch_check("dfm_reset_identity");
}

// The march of every point set of m, with `lds` bytes of dynamic LDS
// requested per workgroup (the kernel uses none; a request caps how many of
// its workgroups a CU holds at once).
static int march_launch(const dfm_grid* g, const dfm_mac* u, const dfm_flowmap* m, float dt, int order,
                        size_t lds, void* stream) {
  Geo G;
  int rc = check_grid(g, G);
  if (rc) return rc;
  if ((rc = check_mac(G, u, "velocity is NULL"))) return rc;
  if (!m) return set_error(DFM_E_NULL, "flowmap is NULL");
  if (order != 2 && order != 4) return set_error(DFM_E_ARG, "order must be 2 or 4");
  Sets S = sets_init();
  MapPtrs M{};
  for (int a = 0; a < G.dim; ++a)
    if (m->psi_f[a]) {
      M.psi[S.nset] = m->psi_f[a];
      M.T[S.nset] = m->T_f[a];
      sets_add(S, G, a);
    }
  if (m->psi) {
    M.psi[S.nset] = m->psi;
    M.T[S.nset] = nullptr;
    sets_add(S, G, 3);
  }
  if (S.nset == 0) return DFM_OK;
  int64_t total = S.off[S.nset];
  Mac U = to_mac(u);
  dim3 grid(blocks_for(total)), blk(kBlock);
  hipStream_t st = S_(stream);
  if (G.dim == 2) {
    if (order == 2) hipLaunchKernelGGL((k_march<2, 2>), grid, blk, lds, st, G, U, dt, S, M);
    else hipLaunchKernelGGL((k_march<2, 4>), grid, blk, lds, st, G, U, dt, S, M);
  } else {
    if (order == 2) hipLaunchKernelGGL((k_march<3, 2>), grid, blk, lds, st, G, U, dt, S, M);
    else hipLaunchKernelGGL((k_march<3, 4>), grid, blk, lds, st, G, U, dt, S, M);
  }
  return launch_check("dfm_march");
}

// Dynamic LDS that lets at most n march workgroups share a CU: the least
// request of which n + 1 no longer fit in the CU's LDS (1 KB steps), so the
// rest of the CU's LDS stays free for other kernels.  Instantiations that
// need more than the default per-workgroup limit are opted in once per
// device.
static int march_cap_lds(int n, size_t* lds) {
  int dev = 0, per_cu = 0, per_block = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess ||
      hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || per_cu <= 0)
    return set_error(DFM_E_HIP, "march: LDS size query failed");
  const size_t kb = 1024, req = ((size_t)per_cu / (size_t)(n + 1) / kb + 1) * kb;
  if ((size_t)n * req > (size_t)per_cu) return set_error(DFM_E_ARG, "march: blocks_per_cu too large for the CU's LDS");
  if (req > (size_t)per_block) {
    static std::atomic<uint64_t> opted{0};  // bit per device
    if (dev < 64 && !(opted.load() >> dev & 1u)) {
      const void* fns[4] = {(const void*)k_march<2, 2>, (const void*)k_march<2, 4>, (const void*)k_march<3, 2>,
                            (const void*)k_march<3, 4>};
      for (const void* f : fns)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)req) != hipSuccess)
          return set_error(DFM_E_HIP, "march: LDS opt-in refused");
      opted.fetch_or(uint64_t(1) << dev);
    }
  }
  *lds = req;
  return DFM_OK;
}

extern "C" int dfm_march(const dfm_grid* g, const dfm_mac* u, const dfm_flowmap* m, float dt, int order,
                         void* stream) {
  return march_launch(g, u, m, dt, order, 0, stream);
}

extern "C" int dfm_march_capped(const dfm_grid* g, const dfm_mac* u, const dfm_flowmap* m, float dt, int order,
                                int blocks_per_cu, void* stream) {
  if (blocks_per_cu < 0 || blocks_per_cu > kMarchCapMax)
    return set_error(DFM_E_ARG, "blocks_per_cu must be 0 (no cap) .. 16");
  size_t lds = 0;
  if (blocks_per_cu > 0) {
    const int rc = march_cap_lds(blocks_per_cu, &lds);
    if (rc) return rc;
  }
  return march_launch(g, u, m, dt, order, lds, stream);
}

extern "C" int dfm_pullback_u_k(const dfm_grid* g, int kern, const dfm_mac* src, const dfm_flowmap* m, float alpha,
                                float beta, const dfm_mac* add, const dfm_mac* out, void* stream) {
  Geo G;
  int rc = check_grid(g, G);
  if (rc) return rc;
  if ((rc = check_kern(kern))) return rc;
  if ((rc = check_mac(G, src, "pullback source is NULL"))) return rc;
  if ((rc = check_mac(G, out, "pullback output is NULL"))) return rc;
  if (beta != 0.f && (rc = check_mac(G, add, "pullback add field is NULL"))) return rc;
  if (!m) return set_error(DFM_E_NULL, "flowmap is NULL");
  for (int a = 0; a < G.dim; ++a)
    if (!m->psi_f[a] || !m->T_f[a]) return set_error(DFM_E_NULL, "pullback needs psi_f and T_f on every face axis");
  Sets S = face_sets(G);
  MapPtrs M{};
  Epi E{};
  E.alpha = alpha;
  E.beta = beta;
  for (int a = 0; a < G.dim; ++a) {
    M.psi[a] = m->psi_f[a];
    M.T[a] = m->T_f[a];
    E.add[a] = add ? add->c[a] : nullptr;
    E.out[a] = out->c[a];
  }
  int64_t total = S.off[S.nset];
  dim3 grid(blocks_for(total)), blk(kBlock);
  Mac U = to_mac(src);
  rc = with_kern(kern, [&](auto kc) {
    constexpr int K = decltype(kc)::value;
    DFM_DISPATCH_DIM(G, hipLaunchKernelGGL((k_pullback_u<2, K>), grid, blk, 0, S_(stream), G, U, S, M, E),
                     hipLaunchKernelGGL((k_pullback_u<3, K>), grid, blk, 0, S_(stream), G, U, S, M, E));
  });
  return rc ? rc : launch_check("dfm_pullback_u");
}

extern "C" int dfm_pullback_u(const dfm_grid* g, const dfm_mac* src, const dfm_flowmap* m, float alpha, float beta,
                              const dfm_mac* add, const dfm_mac* out, void* stream) {
  return dfm_pullback_u_k(g, 2, src, m, alpha, beta, add, out, stream);
}

extern "C" int dfm_pullback_s_k(const dfm_grid* g, int kern, const float* src, const float* psi, float alpha,
                                float beta, const float* add, float* out, void* stream) {
  Geo G;
  int rc = check_grid(g, G);
  if (rc) return rc;
  if ((rc = check_kern(kern))) return rc;
  if (!src || !psi || !out || (beta != 0.f && !add)) return set_error(DFM_E_NULL, "dfm_pullback_s: NULL pointer");
  int64_t cnt = count_of_host(G, 3);
  dim3 grid(blocks_for(cnt)), blk(kBlock);
  rc = with_kern(kern, [&](auto kc) {
    constexpr int K = decltype(kc)::value;
    DFM_DISPATCH_DIM(G,
                     hipLaunchKernelGGL((k_pullback_s<2, K>), grid, blk, 0, S_(stream), G, src, psi, alpha, beta, add,
                                        out, cnt),
                     hipLaunchKernelGGL((k_pullback_s<3, K>), grid, blk, 0, S_(stream), G, src, psi, alpha, beta, add,
                                        out, cnt));
  });
  return rc ? rc : launch_check("dfm_pullback_s");
}

extern "C" int dfm_pullback_s(const dfm_grid* g, const float* src, const float* psi, float alpha, float beta,
                              const float* add, float* out, void* stream) {
  return dfm_pullback_s_k(g, 2, src, psi, alpha, beta, add, out, stream);
}

extern "C" int dfm_bfecc_u_k(const dfm_grid* g, int kern, const dfm_mac* init, const dfm_flowmap* fwd,
                             const dfm_flowmap* bwd, const dfm_mac* err, const dfm_mac* out, void* stream) {
  int rc;
  if ((rc = check_kern(kern))) return rc;
  // out = T.init(psi)
  if ((rc = dfm_pullback_u_k(g, kern, init, fwd, 1.0f, 0.0f, nullptr, out, stream))) return rc;
  // err = 0.5 F.out(phi) - 0.5 init
  if ((rc = dfm_pullback_u_k(g, kern, out, bwd, 0.5f, -0.5f, init, err, stream))) return rc;
  // out = out - T.err(psi)
  return dfm_pullback_u_k(g, kern, err, fwd, -1.0f, 1.0f, out, out, stream);
}

extern "C" int dfm_bfecc_u(const dfm_grid* g, const dfm_mac* init, const dfm_flowmap* fwd, const dfm_flowmap* bwd,
                           const dfm_mac* err, const dfm_mac* out, void* stream) {
  return dfm_bfecc_u_k(g, 2, init, fwd, bwd, err, out, stream);
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_dev[kMaxDevices] = {};

// Device address of the current GPU's copy of `scratch`, looked up once per
// device (with that device current) and kept.
int noalloc_scratch(float** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_error(DFM_E_HIP, "dfm_bfecc_u_noalloc: hipGetDevice", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return set_error(DFM_E_ARG, "dfm_bfecc_u_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(DFM_E_HIP, "dfm_bfecc_u_noalloc: hipGetSymbolAddress(scratch)", hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_dev[dev];
  return DFM_OK;
}
}  // namespace

extern "C" int dfm_bfecc_u_noalloc(const dfm_grid* g, const dfm_mac* init, const dfm_flowmap* fwd,
                                   const dfm_flowmap* bwd, const dfm_mac* out, void* stream) {
  Geo G;
  int rc = check_grid(g, G);
  if (rc) return rc;
  int64_t off[4] = {0, 0, 0, 0};
  for (int a = 0; a < G.dim; ++a) off[a + 1] = off[a] + count_of_host(G, a);
  if (off[G.dim] > (int64_t)DFM_NOALLOC_MAX_VALUES)
    return set_error(DFM_E_ARG, "dfm_bfecc_u_noalloc: the grid's face values exceed DFM_NOALLOC_MAX_VALUES "
                                "(use dfm_bfecc_u with a caller-provided error field)");
  float* base = nullptr;
  if ((rc = noalloc_scratch(&base))) return rc;
  // the error field's components lie one after another in the scratch; every
  // value is written by the second pullback before the third reads it, so
  // nothing needs clearing between calls
  dfm_mac err{};
  for (int a = 0; a < G.dim; ++a) err.c[a] = base + off[a];
  return dfm_bfecc_u(g, init, fwd, bwd, &err, out, stream);
}

extern "C" int dfm_bfecc_s_k(const dfm_grid* g, int kern, const float* init, const float* psi, const float* phi,
                             float* err, float* out, void* stream) {
  int rc;
  if ((rc = check_kern(kern))) return rc;
  if ((rc = dfm_pullback_s_k(g, kern, init, psi, 1.0f, 0.0f, nullptr, out, stream))) return rc;
  if ((rc = dfm_pullback_s_k(g, kern, out, phi, 0.5f, -0.5f, init, err, stream))) return rc;
  return dfm_pullback_s_k(g, kern, err, psi, -1.0f, 1.0f, out, out, stream);
}

extern "C" int dfm_bfecc_s(const dfm_grid* g, const float* init, const float* psi, const float* phi, float* err,
                           float* out, void* stream) {
  return dfm_bfecc_s_k(g, 2, init, psi, phi, err, out, stream);
}

extern "C" int dfm_accumulate(const dfm_grid* g, const dfm_mac* f, const dfm_flowmap* m, const dfm_mac* init,
                              const dfm_mac* h1, const dfm_mac* h2, int ind, float dt, const dfm_mac* f2, float dt2,
                              void* stream) {
  Geo G;
  int rc = check_grid(g, G);
  if (rc) return rc;
  if ((rc = check_mac(G, f, "accumulate source is NULL"))) return rc;
  if ((rc = check_mac(G, init, "accumulate target is NULL"))) return rc;
  if ((rc = check_mac(G, h1, "accumulate history h1 is NULL"))) return rc;
  if ((rc = check_mac(G, h2, "accumulate history h2 is NULL"))) return rc;
  if (f2 && (rc = check_mac(G, f2, "accumulate second source incomplete"))) return rc;
  if (!m) return set_error(DFM_E_NULL, "flowmap is NULL");
  for (int a = 0; a < G.dim; ++a)
    if (!m->psi_f[a] || !m->T_f[a]) return set_error(DFM_E_NULL, "accumulate needs psi_f and T_f");
  Sets S = face_sets(G);
  MapPtrs M{};
  AccPtrs A{};
  for (int a = 0; a < G.dim; ++a) {
    M.psi[a] = m->psi_f[a];
    M.T[a] = m->T_f[a];
    A.init[a] = init->c[a];
    A.h1[a] = h1->c[a];
    A.h2[a] = h2->c[a];
  }
  int64_t total = S.off[S.nset];
  dim3 grid(blocks_for(total)), blk(kBlock);
  Mac F = to_mac(f), F2 = to_mac(f2);
  bool has2 = f2 != nullptr;
  DFM_DISPATCH_DIM(G, hipLaunchKernelGGL((k_accumulate<2>), grid, blk, 0, S_(stream), G, F, F2, has2, S, M, A, ind, dt, dt2),
                   hipLaunchKernelGGL((k_accumulate<3>), grid, blk, 0, S_(stream), G, F, F2, has2, S, M, A, ind, dt, dt2));
  return launch_check("dfm_accumulate");
}

extern "C" int dfm_sl_advect_k(const dfm_grid* g, int kern, const dfm_mac* vel, float dt, int order, int nvec,
                               const dfm_mac* src_vec, const dfm_mac* dst_vec, int nscal, const float* const* src_s,
                               float* const* dst_s, void* stream) {
  Geo G;
  int rc = check_grid(g, G);
  if (rc) return rc;
  if ((rc = check_kern(kern))) return rc;
  if ((rc = check_mac(G, vel, "advection velocity is NULL"))) return rc;
  if (order != 2 && order != 4) return set_error(DFM_E_ARG, "order must be 2 or 4");
  if (nvec < 0 || nvec > kMaxVec || nscal < 0 || nscal > kMaxScal) return set_error(DFM_E_ARG, "too many sources");
  SLArgs A{};
  A.nvec = nvec;
  A.nscal = nscal;
  for (int m = 0; m < nvec; ++m) {
    if ((rc = check_mac(G, &src_vec[m], "advected MAC source is NULL"))) return rc;
    if ((rc = check_mac(G, &dst_vec[m], "advected MAC destination is NULL"))) return rc;
    A.src[m] = to_mac(&src_vec[m]);
    for (int a = 0; a < 3; ++a) A.dst[m][a] = dst_vec[m].c[a];
  }
  for (int m = 0; m < nscal; ++m) {
    if (!src_s || !dst_s || !src_s[m] || !dst_s[m]) return set_error(DFM_E_NULL, "advected scalar is NULL");
    A.ssrc[m] = src_s[m];
    A.sdst[m] = dst_s[m];
  }
  Sets S = sets_init();
  if (nvec > 0)
    for (int a = 0; a < G.dim; ++a) sets_add(S, G, a);
  if (nscal > 0) sets_add(S, G, 3);
  if (S.nset == 0) return DFM_OK;
  int64_t total = S.off[S.nset];
  dim3 grid(blocks_for(total)), blk(kBlock);
  Mac V = to_mac(vel);
  hipStream_t st = S_(stream);
  rc = with_kern(kern, [&](auto kc) {
    constexpr int K = decltype(kc)::value;
    if (G.dim == 2) {
      if (order == 2) hipLaunchKernelGGL((k_sl_advect<2, 2, K>), grid, blk, 0, st, G, V, dt, S, A);
      else hipLaunchKernelGGL((k_sl_advect<2, 4, K>), grid, blk, 0, st, G, V, dt, S, A);
    } else {
      if (order == 2) hipLaunchKernelGGL((k_sl_advect<3, 2, K>), grid, blk, 0, st, G, V, dt, S, A);
      else hipLaunchKernelGGL((k_sl_advect<3, 4, K>), grid, blk, 0, st, G, V, dt, S, A);
    }
  });
  return rc ? rc : launch_check("dfm_sl_advect");
}

extern "C" int dfm_sl_advect(const dfm_grid* g, const dfm_mac* vel, float dt, int order, int nvec,
                             const dfm_mac* src_vec, const dfm_mac* dst_vec, int nscal, const float* const* src_s,
                             float* const* dst_s, void* stream) {
  return dfm_sl_advect_k(g, 2, vel, dt, order, nvec, src_vec, dst_vec, nscal, src_s, dst_s, stream);
}

namespace {

// Library-wide kernel choices (dfm_flowmap_set_option).
std::atomic<int> g_tiled_sources{1};

// The 3D adjoint sources by the LDS-tiled kernel: one workgroup per output
// tile of each face grid, grids in axis order.
int launch_src_tiled(const Geo& G, int K, const tiled::Pairs& P, const dfm_mac* r, hipStream_t st) {
  tiled::Tiles T{};
  const int te[3] = {tiled::TX, tiled::TY, tiled::TZ};
  T.off[0] = 0;
  for (int a = 0; a < 3; ++a) {
    int So[3];
    shape_of_host(G, a, So);
    int64_t nb = 1;
    for (int b = 0; b < 3; ++b) {
      T.nt[a][b] = (So[b] + te[b] - 1) / te[b];
      nb *= T.nt[a][b];
    }
    T.off[a + 1] = (int)(T.off[a] + nb);
    T.out[a] = r->c[a];
  }
  dim3 grid(T.off[3]), blk(tiled::THREADS);
  if (K == 3) hipLaunchKernelGGL((tiled::k_src_tiled<3>), grid, blk, 0, st, G, P, T);
  else hipLaunchKernelGGL((tiled::k_src_tiled<4>), grid, blk, 0, st, G, P, T);
  return DFM_OK;
}

}  // namespace

extern "C" int dfm_flowmap_set_option(int option, int value) {
  if (option != DFM_FM_TILED_SOURCES) return set_error(DFM_E_ARG, "unknown flow-map option");
  if (value != 0 && value != 1) return set_error(DFM_E_ARG, "DFM_FM_TILED_SOURCES takes 0 or 1");
  g_tiled_sources.store(value);
  return DFM_OK;
}

extern "C" int dfm_flowmap_get_option(int option, int* value) {
  if (!value) return set_error(DFM_E_NULL, "value is NULL");
  if (option != DFM_FM_TILED_SOURCES) return set_error(DFM_E_ARG, "unknown flow-map option");
  *value = g_tiled_sources.load();
  return DFM_OK;
}

extern "C" int dfm_adj_conv_source(const dfm_grid* g, const dfm_mac* u, const dfm_mac* w, int kern,
                                   const dfm_mac* r, void* stream) {
  Geo G;
  int rc = check_grid(g, G);
  if (rc) return rc;
  if ((rc = check_mac(G, u, "u is NULL")) || (rc = check_mac(G, w, "w is NULL")) ||
      (rc = check_mac(G, r, "r is NULL")))
    return rc;
  if (kern != 3 && kern != 4) return set_error(DFM_E_ARG, "kern must be 3 (cubic) or 4 (quartic)");
  hipStream_t st = S_(stream);
  if (G.dim == 3 && g_tiled_sources.load()) {
    tiled::Pairs P{};
    P.n = 3;
    P.nslot = 3;
    for (int c = 0; c < 3; ++c) {
      P.slot[c] = c;
      P.slot_lattice[c] = c;
      P.fd[c] = u->c[c];
      P.fv[c] = w->c[c];
    }
    launch_src_tiled(G, kern, P, r, st);
    return launch_check("dfm_adj_conv_source");
  }
  Sets S = face_sets(G);
  MapPtrs R{};
  for (int a = 0; a < G.dim; ++a) R.psi[a] = r->c[a];
  int64_t total = S.off[S.nset];
  dim3 grid(blocks_for(total)), blk(kBlock);
  Mac U = to_mac(u), W = to_mac(w);
  if (G.dim == 2) {
    if (kern == 3) hipLaunchKernelGGL((k_adj_conv<2, 3>), grid, blk, 0, st, G, U, W, S, R);
    else hipLaunchKernelGGL((k_adj_conv<2, 4>), grid, blk, 0, st, G, U, W, S, R);
  } else {
    if (kern == 3) hipLaunchKernelGGL((k_adj_conv<3, 3>), grid, blk, 0, st, G, U, W, S, R);
    else hipLaunchKernelGGL((k_adj_conv<3, 4>), grid, blk, 0, st, G, U, W, S, R);
  }
  return launch_check("dfm_adj_conv_source");
}

extern "C" int dfm_adj_scalar_source(const dfm_grid* g, int nscal, const float* const* s, const float* const* lam,
                                     const dfm_mac* r, void* stream) {
  Geo G;
  int rc = check_grid(g, G);
  if (rc) return rc;
  if ((rc = check_mac(G, r, "r is NULL"))) return rc;
  if (nscal < 0 || nscal > kMaxScal) return set_error(DFM_E_ARG, "nscal must be 0..8");
  ScalPairs P{};
  P.n = nscal;
  for (int m = 0; m < nscal; ++m) {
    if (!s || !lam || !s[m] || !lam[m]) return set_error(DFM_E_NULL, "scalar pair is NULL");
    P.s[m] = s[m];
    P.lam[m] = lam[m];
  }
  if (G.dim == 3 && g_tiled_sources.load()) {
    tiled::Pairs T{};
    T.n = nscal;
    T.nslot = 1;
    T.slot_lattice[0] = 3;
    for (int m = 0; m < nscal; ++m) {
      T.slot[m] = 0;
      T.fd[m] = P.s[m];
      T.fv[m] = P.lam[m];
    }
    launch_src_tiled(G, 3, T, r, S_(stream));
    return launch_check("dfm_adj_scalar_source");
  }
  Sets S = face_sets(G);
  MapPtrs R{};
  for (int a = 0; a < G.dim; ++a) R.psi[a] = r->c[a];
  int64_t total = S.off[S.nset];
  dim3 grid(blocks_for(total)), blk(kBlock);
  DFM_DISPATCH_DIM(G, hipLaunchKernelGGL((k_adj_scalar<2>), grid, blk, 0, S_(stream), G, P, S, R),
                   hipLaunchKernelGGL((k_adj_scalar<3>), grid, blk, 0, S_(stream), G, P, S, R));
  return launch_check("dfm_adj_scalar_source");
}

extern "C" int dfm_axpy_mac(const dfm_grid* g, const dfm_mac* x, float alpha, const dfm_mac* y, float beta,
                            const dfm_mac* z, const dfm_mac* out, void* stream) {
  Geo G;
  int rc = check_grid(g, G);
  if (rc) return rc;
  if ((rc = check_mac(G, x, "x is NULL")) || (rc = check_mac(G, y, "y is NULL")) ||
      (rc = check_mac(G, out, "out is NULL")))
    return rc;
  if (z && (rc = check_mac(G, z, "z incomplete"))) return rc;
  Sets S = face_sets(G);
  Axpy A{};
  A.alpha = alpha;
  A.beta = beta;
  for (int a = 0; a < G.dim; ++a) {
    A.x[a] = x->c[a];
    A.y[a] = y->c[a];
    A.z[a] = z ? z->c[a] : nullptr;
    A.out[a] = out->c[a];
  }
  int64_t total = S.off[S.nset];
  hipLaunchKernelGGL(k_axpy, dim3(blocks_for(total)), dim3(kBlock), 0, S_(stream), S, A);
  return launch_check("dfm_axpy_mac");
}

// ------------------------------------------------ vector-Jacobian products ----
extern "C" int dfm_pullback_u_vjp(const dfm_grid* g, int kern, const dfm_mac* g_out, const dfm_mac* src,
                                  const dfm_flowmap* m, float alpha, const dfm_mac* g_src, const dfm_flowmap* g_m,
                                  void* stream) {
  Geo G;
  int rc = check_grid(g, G);
  if (rc) return rc;
  if ((rc = check_kern(kern))) return rc;
  if ((rc = check_mac(G, g_out, "pullback output gradient is NULL"))) return rc;
  if ((rc = check_mac(G, g_src, "pullback source gradient is NULL"))) return rc;
  if (!m) return set_error(DFM_E_NULL, "flowmap is NULL");
  for (int a = 0; a < G.dim; ++a)
    if (!m->psi_f[a] || !m->T_f[a]) return set_error(DFM_E_NULL, "pullback needs psi_f and T_f on every face axis");
  if (g_m) {
    if ((rc = check_mac(G, src, "the map gradient needs the pullback source"))) return rc;
    for (int a = 0; a < G.dim; ++a)
      if (!g_m->psi_f[a] || !g_m->T_f[a]) return set_error(DFM_E_NULL, "map gradient needs psi_f and T_f");
  }
  Sets S = face_sets(G);
  MapPtrs M{}, GM{};
  Epi E{};
  for (int a = 0; a < G.dim; ++a) {
    M.psi[a] = m->psi_f[a];
    M.T[a] = m->T_f[a];
    E.out[a] = g_src->c[a];
    if (g_m) {
      GM.psi[a] = g_m->psi_f[a];
      GM.T[a] = g_m->T_f[a];
    }
  }
  dim3 grid(blocks_for(S.off[S.nset])), blk(kBlock);
  Mac GO = to_mac(g_out);
  Mac SR = g_m ? to_mac(src) : Mac{};
  rc = with_kern(kern, [&](auto kc) {
    constexpr int K = decltype(kc)::value;
    DFM_DISPATCH_DIM(
        G, hipLaunchKernelGGL((k_pullback_u_vjp<2, K>), grid, blk, 0, S_(stream), G, GO, SR, S, M, alpha, E, GM),
        hipLaunchKernelGGL((k_pullback_u_vjp<3, K>), grid, blk, 0, S_(stream), G, GO, SR, S, M, alpha, E, GM));
  });
  return rc ? rc : launch_check("dfm_pullback_u_vjp");
}

extern "C" int dfm_pullback_s_vjp(const dfm_grid* g, int kern, const float* g_out, const float* src, const float* psi,
                                  float alpha, float* g_src, float* g_psi, void* stream) {
  Geo G;
  int rc = check_grid(g, G);
  if (rc) return rc;
  if ((rc = check_kern(kern))) return rc;
  if (!g_out || !psi || !g_src) return set_error(DFM_E_NULL, "dfm_pullback_s_vjp: NULL pointer");
  if (g_psi && !src) return set_error(DFM_E_NULL, "the map gradient needs the pullback source");
  const int64_t cnt = count_of_host(G, 3);
  dim3 grid(blocks_for(cnt)), blk(kBlock);
  rc = with_kern(kern, [&](auto kc) {
    constexpr int K = decltype(kc)::value;
    DFM_DISPATCH_DIM(G,
                     hipLaunchKernelGGL((k_pullback_s_vjp<2, K>), grid, blk, 0, S_(stream), G, g_out, src, psi, alpha,
                                        g_src, g_psi, cnt),
                     hipLaunchKernelGGL((k_pullback_s_vjp<3, K>), grid, blk, 0, S_(stream), G, g_out, src, psi, alpha,
                                        g_src, g_psi, cnt));
  });
  return rc ? rc : launch_check("dfm_pullback_s_vjp");
}

extern "C" int dfm_sl_advect_vjp(const dfm_grid* g, int kern, const dfm_mac* vel, float dt, int order, int nvec,
                                 const dfm_mac* src_vec, const dfm_mac* g_dst_vec, const dfm_mac* g_src_vec, int nscal,
                                 const float* const* src_s, const float* const* g_dst_s, float* const* g_src_s,
                                 const dfm_mac* g_vel, void* stream) {
  Geo G;
  int rc = check_grid(g, G);
  if (rc) return rc;
  if ((rc = check_kern(kern))) return rc;
  if ((rc = check_mac(G, vel, "advection velocity is NULL"))) return rc;
  if (order != 2 && order != 4) return set_error(DFM_E_ARG, "order must be 2 or 4");
  if (nvec < 0 || nvec > kMaxVec || nscal < 0 || nscal > kMaxScal) return set_error(DFM_E_ARG, "too many sources");
  SLVjpArgs A{};
  A.nvec = nvec;
  A.nscal = nscal;
  if (g_vel) {
    if ((rc = check_mac(G, g_vel, "velocity gradient is NULL"))) return rc;
    for (int a = 0; a < 3; ++a) A.gvel[a] = a < G.dim ? g_vel->c[a] : nullptr;
  }
  for (int m = 0; m < nvec; ++m) {
    if ((rc = check_mac(G, &g_dst_vec[m], "advected MAC gradient is NULL"))) return rc;
    if ((rc = check_mac(G, &g_src_vec[m], "MAC source gradient is NULL"))) return rc;
    A.gdst[m] = to_mac(&g_dst_vec[m]);
    for (int a = 0; a < 3; ++a) A.gsrc[m][a] = g_src_vec[m].c[a];
    if (g_vel) {
      if (!src_vec) return set_error(DFM_E_NULL, "the velocity gradient needs the forward sources");
      if ((rc = check_mac(G, &src_vec[m], "MAC source is NULL"))) return rc;
      A.fsrc[m] = to_mac(&src_vec[m]);
    }
  }
  for (int m = 0; m < nscal; ++m) {
    if (!g_dst_s || !g_src_s || !g_dst_s[m] || !g_src_s[m]) return set_error(DFM_E_NULL, "scalar gradient is NULL");
    A.gds[m] = g_dst_s[m];
    A.gss[m] = g_src_s[m];
    if (g_vel) {
      if (!src_s || !src_s[m]) return set_error(DFM_E_NULL, "the velocity gradient needs the forward sources");
      A.fs[m] = src_s[m];
    }
  }
  Sets S = sets_init();
  if (nvec > 0)
    for (int a = 0; a < G.dim; ++a) sets_add(S, G, a);
  if (nscal > 0) sets_add(S, G, 3);
  if (S.nset == 0) return DFM_OK;
  dim3 grid(blocks_for(S.off[S.nset])), blk(kBlock);
  Mac V = to_mac(vel);
  hipStream_t st = S_(stream);
  rc = with_kern(kern, [&](auto kc) {
    constexpr int K = decltype(kc)::value;
    if (G.dim == 2) {
      if (order == 2) hipLaunchKernelGGL((k_sl_advect_vjp<2, 2, K>), grid, blk, 0, st, G, V, dt, S, A);
      else hipLaunchKernelGGL((k_sl_advect_vjp<2, 4, K>), grid, blk, 0, st, G, V, dt, S, A);
    } else {
      if (order == 2) hipLaunchKernelGGL((k_sl_advect_vjp<3, 2, K>), grid, blk, 0, st, G, V, dt, S, A);
      else hipLaunchKernelGGL((k_sl_advect_vjp<3, 4, K>), grid, blk, 0, st, G, V, dt, S, A);
    }
  });
  return rc ? rc : launch_check("dfm_sl_advect_vjp");
}

extern "C" int dfm_march_vjp(const dfm_grid* g, const dfm_mac* u, const dfm_flowmap* m0, float dt, int order,
                             const dfm_flowmap* gm, const dfm_mac* g_u, void* stream) {
  Geo G;
  int rc = check_grid(g, G);
  if (rc) return rc;
  if ((rc = check_mac(G, u, "velocity is NULL"))) return rc;
  if ((rc = check_mac(G, g_u, "velocity gradient is NULL"))) return rc;
  if (!m0 || !gm) return set_error(DFM_E_NULL, "flowmap is NULL");
  if (order != 2 && order != 4) return set_error(DFM_E_ARG, "order must be 2 or 4");
  Sets S = sets_init();
  MapPtrs M{}, GM{};
  for (int a = 0; a < G.dim; ++a)
    if (m0->psi_f[a]) {
      if (!gm->psi_f[a] || (m0->T_f[a] && !gm->T_f[a]))
        return set_error(DFM_E_NULL, "map gradient lacks a set the map has");
      M.psi[S.nset] = m0->psi_f[a];
      M.T[S.nset] = m0->T_f[a];
      GM.psi[S.nset] = gm->psi_f[a];
      GM.T[S.nset] = gm->T_f[a];
      sets_add(S, G, a);
    }
  if (m0->psi) {
    if (!gm->psi) return set_error(DFM_E_NULL, "map gradient lacks a set the map has");
    M.psi[S.nset] = m0->psi;
    GM.psi[S.nset] = gm->psi;
    sets_add(S, G, 3);
  }
  if (S.nset == 0) return DFM_OK;
  Mac U = to_mac(u);
  GuPtrs GU{{g_u->c[0], g_u->c[1], g_u->c[2]}};
  dim3 grid(blocks_for(S.off[S.nset])), blk(kBlock);
  hipStream_t st = S_(stream);
  if (G.dim == 2) {
    if (order == 2) hipLaunchKernelGGL((k_march_vjp<2, 2>), grid, blk, 0, st, G, U, dt, S, M, GM, GU);
    else hipLaunchKernelGGL((k_march_vjp<2, 4>), grid, blk, 0, st, G, U, dt, S, M, GM, GU);
  } else {
    if (order == 2) hipLaunchKernelGGL((k_march_vjp<3, 2>), grid, blk, 0, st, G, U, dt, S, M, GM, GU);
    else hipLaunchKernelGGL((k_march_vjp<3, 4>), grid, blk, 0, st, G, U, dt, S, M, GM, GU);
  }
  return launch_check("dfm_march_vjp");
}
