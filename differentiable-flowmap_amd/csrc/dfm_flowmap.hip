// Flow-map kernels: march, pullback, BFECC, semi-Lagrangian advection,
// accumulation and the adjoint sources, plus their C-ABI entry points.
#include "dfm_common.hpp"
#include "dfm_internal.hpp"
#include "../../include/dfm.h"

#include <mutex>

using namespace dfm;

// BFECC error field of dfm_bfecc_u_noalloc: a file-scope device array, one
// copy per GPU, held until the process ends (DFM_NOALLOC_MAX_VALUES floats,
// 32 MiB).  Host code never passes this name to a kernel: it is a host-side
// stand-in, and the device address of each GPU's copy comes from
// hipGetSymbolAddress (noalloc_scratch below).
static __device__ __align__(256) float scratch[DFM_NOALLOC_MAX_VALUES];

namespace {

constexpr int kBlock = 256;

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
template <int DIM, int ORDER, bool HAS_T>
DFM_HD void rk_point(const Geo& g, const Mac& u, float dt, float p[3], float T[3]) {
  const float neg = -dt;
  float u1[3], u2[3], u3[3], u4[3], J[3][3];
  float k1[3], k2[3], k3[3], k4[3], q[3], Tq[3];
  auto JT = [&](const float* t, float* out) {
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      float s = 0.f;
#pragma unroll
      for (int b = 0; b < DIM; ++b) s = fmaf(J[a][b], t[b], s);
      out[a] = s;
    }
  };
  mac_eval<HAS_T, DIM>(g, u, p, u1, J);
  if (HAS_T) JT(T, k1);
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    q[a] = p[a] + 0.5f * neg * u1[a];
    if (HAS_T) Tq[a] = T[a] + 0.5f * neg * k1[a];
  }
  mac_eval<HAS_T, DIM>(g, u, q, u2, J);
  if (HAS_T) JT(Tq, k2);
  if (ORDER == 2) {
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      p[a] = p[a] + neg * u2[a];
      if (HAS_T) T[a] = T[a] + neg * k2[a];
    }
    return;
  }
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    q[a] = p[a] + 0.5f * neg * u2[a];
    if (HAS_T) Tq[a] = T[a] + 0.5f * neg * k2[a];
  }
  mac_eval<HAS_T, DIM>(g, u, q, u3, J);
  if (HAS_T) JT(Tq, k3);
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    q[a] = p[a] + neg * u3[a];
    if (HAS_T) Tq[a] = T[a] + neg * k3[a];
  }
  mac_eval<HAS_T, DIM>(g, u, q, u4, J);
  if (HAS_T) JT(Tq, k4);
  const float sixth = 1.0f / 6.0f;
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    p[a] = p[a] + neg * sixth * (u1[a] + 2.0f * u2[a] + 2.0f * u3[a] + u4[a]);
    if (HAS_T) T[a] = T[a] + neg * sixth * (k1[a] + 2.0f * k2[a] + 2.0f * k3[a] + k4[a]);
  }
}

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

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_pullback_u(Geo g, Mac src, Sets S, MapPtrs M, Epi E) {
  int64_t gid = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (gid >= S.off[S.nset]) return;
  int s = S.find(gid);
  int64_t i = gid - S.off[s];
  int64_t cnt = S.cnt[s];
  float p[3] = {0.f, 0.f, 0.f}, v[3];
#pragma unroll
  for (int c = 0; c < DIM; ++c) p[c] = M.psi[s][c * cnt + i];
  mac_eval<false, DIM>(g, src, p, v, nullptr);
  float d = 0.f;
#pragma unroll
  for (int c = 0; c < DIM; ++c) d = fmaf(M.T[s][c * cnt + i], v[c], d);
  float r = E.alpha * d;
  if (E.beta != 0.f) r = fmaf(E.beta, E.add[s][i], r);
  E.out[s][i] = r;
}

template <int DIM>
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
  float v = interp<2, false, DIM>(src, S, bl, p, g.inv_dx, nullptr);
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

template <int DIM, int ORDER>
DFM_HD void backtrace(const Geo& g, const Mac& vel, float dt, const float x[3], float p[3]) {
  float v1[3], v2[3], v3[3], v4[3], q[3];
  mac_eval<false, DIM>(g, vel, x, v1, nullptr);
#pragma unroll
  for (int a = 0; a < DIM; ++a) q[a] = x[a] - v1[a] * dt * 0.5f;
  mac_eval<false, DIM>(g, vel, q, v2, nullptr);
  if (ORDER == 2) {
#pragma unroll
    for (int a = 0; a < DIM; ++a) p[a] = x[a] - v2[a] * dt;
    return;
  }
#pragma unroll
  for (int a = 0; a < DIM; ++a) q[a] = x[a] - v2[a] * dt * 0.5f;
  mac_eval<false, DIM>(g, vel, q, v3, nullptr);
#pragma unroll
  for (int a = 0; a < DIM; ++a) q[a] = x[a] - v3[a] * dt;
  mac_eval<false, DIM>(g, vel, q, v4, nullptr);
#pragma unroll
  for (int a = 0; a < DIM; ++a) p[a] = x[a] - (v1[a] + 2.0f * v2[a] + 2.0f * v3[a] + v4[a]) / 6.0f * dt;
}

template <int DIM, int ORDER>
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
    for (int m = 0; m < A.nvec; ++m) A.dst[m][grid][i] = interp<2, false, DIM>(A.src[m].c[grid], sh, bl, p, g.inv_dx, nullptr);
  } else {
    for (int m = 0; m < A.nscal; ++m) A.sdst[m][i] = interp<2, false, DIM>(A.ssrc[m], sh, bl, p, g.inv_dx, nullptr);
  }
}

// ------------------------------------------------------- adjoint sources ----
// Derivative along `axis` of field fd and value of field fv, both on the same
// grid (shape S, offset bl), sampled at p with the B-spline of order K; the
// two fields share every tap and weight.
template <int K, int DIM>
DFM_HD void dual_interp(const float* __restrict__ fd, const float* __restrict__ fv, const int S[3], const float bl[3],
                        const float p[3], float inv_dx, int axis, float& d_out, float& v_out) {
  constexpr int NT = Spline<K>::NT;
  Axis<K, true> ax, ay, az;
  ax.setup(p[0] * inv_dx, bl[0], S[0]);
  ay.setup(p[1] * inv_dx, bl[1], S[1]);
  if (DIM == 3) az.setup(p[2] * inv_dx, bl[2], S[2]);
  // derivative weights along `axis`, value weights elsewhere
  float wx[NT], wy[NT], wz[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    wx[t] = axis == 0 ? ax.dw[t] : ax.w[t];
    wy[t] = axis == 1 ? ay.dw[t] : ay.w[t];
    if (DIM == 3) wz[t] = axis == 2 ? az.dw[t] : az.w[t];
  }
  // Each slice's taps of both fields are loaded and pinned before the
  // first is used (the compiler otherwise waits for every load before the
  // next, keeping registers low); the multiply-adds keep their order.
  float d = 0.f, v = 0.f;
  if (DIM == 2) {
#pragma unroll
    for (int a = 0; a < NT; ++a) {
      const uint32_t rb = __umul24((uint32_t)ax.idx[a], (uint32_t)S[1]);
      float td[NT], tv[NT];
#pragma unroll
      for (int b = 0; b < NT; ++b) {
        td[b] = *at32(fd, rb + (uint32_t)ay.idx[b]);
        tv[b] = *at32(fv, rb + (uint32_t)ay.idx[b]);
      }
      pin(td);
      pin(tv);
      float rd = 0.f, rv = 0.f;
#pragma unroll
      for (int b = 0; b < NT; ++b) {
        rd = fmaf(td[b], wy[b], rd);
        rv = fmaf(tv[b], ay.w[b], rv);
      }
      d = fmaf(rd, wx[a], d);
      v = fmaf(rv, ax.w[a], v);
    }
  } else {
#pragma unroll 1
    for (int a = 0; a < NT; ++a) {
      float td[NT][NT], tv[NT][NT];
#pragma unroll
      for (int b = 0; b < NT; ++b) {
        const uint32_t rb = __umul24(__umul24((uint32_t)ax.idx[a], (uint32_t)S[1]) + (uint32_t)ay.idx[b], (uint32_t)S[2]);
#pragma unroll
        for (int c = 0; c < NT; ++c) {
          td[b][c] = *at32(fd, rb + (uint32_t)az.idx[c]);
          tv[b][c] = *at32(fv, rb + (uint32_t)az.idx[c]);
        }
      }
#pragma unroll
      for (int b = 0; b < NT; ++b) {
        pin(td[b]);
        pin(tv[b]);
      }
      float yd = 0.f, yv = 0.f;
#pragma unroll
      for (int b = 0; b < NT; ++b) {
        float zd = 0.f, zv = 0.f;
#pragma unroll
        for (int c = 0; c < NT; ++c) {
          zd = fmaf(td[b][c], wz[c], zd);
          zv = fmaf(tv[b][c], az.w[c], zv);
        }
        yd = fmaf(zd, wy[b], yd);
        yv = fmaf(zv, ay.w[b], yv);
      }
      d = fmaf(yd, wx[a], d);
      v = fmaf(yv, ax.w[a], v);
    }
  }
  d_out = d * inv_dx;
  v_out = v;
}

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

extern "C" int dfm_march(const dfm_grid* g, const dfm_mac* u, const dfm_flowmap* m, float dt, int order,
                         void* stream) {
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
    if (order == 2) hipLaunchKernelGGL((k_march<2, 2>), grid, blk, 0, st, G, U, dt, S, M);
    else hipLaunchKernelGGL((k_march<2, 4>), grid, blk, 0, st, G, U, dt, S, M);
  } else {
    if (order == 2) hipLaunchKernelGGL((k_march<3, 2>), grid, blk, 0, st, G, U, dt, S, M);
    else hipLaunchKernelGGL((k_march<3, 4>), grid, blk, 0, st, G, U, dt, S, M);
  }
  return launch_check("dfm_march");
}

extern "C" int dfm_pullback_u(const dfm_grid* g, const dfm_mac* src, const dfm_flowmap* m, float alpha, float beta,
                              const dfm_mac* add, const dfm_mac* out, void* stream) {
  Geo G;
  int rc = check_grid(g, G);
  if (rc) return rc;
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
  DFM_DISPATCH_DIM(G, hipLaunchKernelGGL((k_pullback_u<2>), grid, blk, 0, S_(stream), G, U, S, M, E),
                   hipLaunchKernelGGL((k_pullback_u<3>), grid, blk, 0, S_(stream), G, U, S, M, E));
  return launch_check("dfm_pullback_u");
}

extern "C" int dfm_pullback_s(const dfm_grid* g, const float* src, const float* psi, float alpha, float beta,
                              const float* add, float* out, void* stream) {
  Geo G;
  int rc = check_grid(g, G);
  if (rc) return rc;
  if (!src || !psi || !out || (beta != 0.f && !add)) return set_error(DFM_E_NULL, "dfm_pullback_s: NULL pointer");
  int64_t cnt = count_of_host(G, 3);
  dim3 grid(blocks_for(cnt)), blk(kBlock);
  DFM_DISPATCH_DIM(G,
                   hipLaunchKernelGGL((k_pullback_s<2>), grid, blk, 0, S_(stream), G, src, psi, alpha, beta, add, out, cnt),
                   hipLaunchKernelGGL((k_pullback_s<3>), grid, blk, 0, S_(stream), G, src, psi, alpha, beta, add, out, cnt));
  return launch_check("dfm_pullback_s");
}

extern "C" int dfm_bfecc_u(const dfm_grid* g, const dfm_mac* init, const dfm_flowmap* fwd, const dfm_flowmap* bwd,
                           const dfm_mac* err, const dfm_mac* out, void* stream) {
  int rc;
  // out = T.init(psi)
  if ((rc = dfm_pullback_u(g, init, fwd, 1.0f, 0.0f, nullptr, out, stream))) return rc;
  // err = 0.5 F.out(phi) - 0.5 init
  if ((rc = dfm_pullback_u(g, out, bwd, 0.5f, -0.5f, init, err, stream))) return rc;
  // out = out - T.err(psi)
  return dfm_pullback_u(g, err, fwd, -1.0f, 1.0f, out, out, stream);
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

extern "C" int dfm_bfecc_s(const dfm_grid* g, const float* init, const float* psi, const float* phi, float* err,
                           float* out, void* stream) {
  int rc;
  if ((rc = dfm_pullback_s(g, init, psi, 1.0f, 0.0f, nullptr, out, stream))) return rc;
  if ((rc = dfm_pullback_s(g, out, phi, 0.5f, -0.5f, init, err, stream))) return rc;
  return dfm_pullback_s(g, err, psi, -1.0f, 1.0f, out, out, stream);
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

extern "C" int dfm_sl_advect(const dfm_grid* g, const dfm_mac* vel, float dt, int order, int nvec,
                             const dfm_mac* src_vec, const dfm_mac* dst_vec, int nscal, const float* const* src_s,
                             float* const* dst_s, void* stream) {
  Geo G;
  int rc = check_grid(g, G);
  if (rc) return rc;
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
  if (G.dim == 2) {
    if (order == 2) hipLaunchKernelGGL((k_sl_advect<2, 2>), grid, blk, 0, st, G, V, dt, S, A);
    else hipLaunchKernelGGL((k_sl_advect<2, 4>), grid, blk, 0, st, G, V, dt, S, A);
  } else {
    if (order == 2) hipLaunchKernelGGL((k_sl_advect<3, 2>), grid, blk, 0, st, G, V, dt, S, A);
    else hipLaunchKernelGGL((k_sl_advect<3, 4>), grid, blk, 0, st, G, V, dt, S, A);
  }
  return launch_check("dfm_sl_advect");
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
  Sets S = face_sets(G);
  MapPtrs R{};
  for (int a = 0; a < G.dim; ++a) R.psi[a] = r->c[a];
  int64_t total = S.off[S.nset];
  dim3 grid(blocks_for(total)), blk(kBlock);
  Mac U = to_mac(u), W = to_mac(w);
  hipStream_t st = S_(stream);
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
