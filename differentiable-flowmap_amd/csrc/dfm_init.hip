// Vortex initial condition of the 2D control problem and its vector-Jacobian
// product with respect to the control vector theta (gfx950).
//
// Forward (sixteen_vortex_vel_func_with_pos_coef, 2D/vortex/init_conditions.py:368-383,
// with angular_vel_func, :141-145): at every cell centre x = (I + 0.5) dx,
//   u(x) = sum_k w_k g(r_k; a_k) 2.5 (-p_y, p_x),   p = x - c_k,
//   r_k = |p| + 1e-6,  g(r; a) = s (1 - exp(-r^2 / a^2)) / r^2,  s = -0.01,
//   a_k = 0.05 rr_k,
// split to the MAC faces as the mean of the two adjacent cell values with
// clamp-to-edge sampling (split_central_vector, 2D/vortex/taichi_utils.py:158-167).
// theta layout (4K values): [w_0..w_{K-1}, cx.., cy.., rr..].
//
// The reference differentiates this map by one-sided finite differences of
// the initializer per component (calculate_dtheta_ind,
// 2D/vortex/optimize_vortex.py:193-226: <lambda_u, (u(theta + eps e_i) -
// u(theta)) / eps>, eps = 1e-5, unprojected); here the same inner product is
// formed with the analytic derivative of u, in FP64.
#include "dfm_common.hpp"
#include "dfm_internal.hpp"
#include "../../include/dfm.h"

using namespace dfm;

namespace {

constexpr double kSwirl = -0.01;  // angular_vel_func strength
constexpr double kGain = 2.5;
constexpr double kRadiusScale = 0.05;
constexpr int kInitB = 256;
constexpr int kVjpB = 1024;

struct Vortex {
  double w, cx, cy, a;
};

DFM_HD Vortex vortex_of(const float* th, int K, int k) {
  return Vortex{(double)th[k], (double)th[K + k], (double)th[2 * K + k], kRadiusScale * (double)th[3 * K + k]};
}

// velocity of all K vortices at cell centre (x, y)
DFM_HD void cell_velocity(const float* th, int K, double x, double y, double u[2]) {
  u[0] = 0.0;
  u[1] = 0.0;
  for (int k = 0; k < K; ++k) {
    const Vortex v = vortex_of(th, K, k);
    const double px = x - v.cx, py = y - v.cy;
    const double r = sqrt(px * px + py * py) + 1e-6;
    const double g = kSwirl * (1.0 - exp(-(r * r) / (v.a * v.a))) / (r * r);
    const double f = g * v.w * kGain;
    u[0] += f * -py;
    u[1] += f * px;
  }
}

// d u(x, y) / d (w_k, cx_k, cy_k, rr_k): d[param][component]
DFM_HD void cell_velocity_grad(const float* th, int K, int k, double x, double y, double d[4][2]) {
  const Vortex v = vortex_of(th, K, k);
  const double px = x - v.cx, py = y - v.cy;
  const double q = sqrt(px * px + py * py);
  const double r = q + 1e-6;
  const double a2 = v.a * v.a;
  const double e = exp(-(r * r) / a2);
  const double g = kSwirl * (1.0 - e) / (r * r);
  const double dg_dr = kSwirl * (2.0 * e / (a2 * r) - 2.0 * (1.0 - e) / (r * r * r));
  const double dg_da = -2.0 * kSwirl * e / (a2 * v.a);
  const double vx = -py, vy = px;  // swirl direction
  const double wG = v.w * kGain;
  // strength
  d[0][0] = g * kGain * vx;
  d[0][1] = g * kGain * vy;
  // centre: p = x - c, so d/dc = -d/dp; dr/dp = p / |p| (0 at the centre,
  // where the swirl direction vanishes as well)
  const double ex = q > 0.0 ? px / q : 0.0, ey = q > 0.0 ? py / q : 0.0;
  d[1][0] = -wG * (dg_dr * ex * vx);
  d[1][1] = -wG * (dg_dr * ex * vy + g);
  d[2][0] = -wG * (dg_dr * ey * vx - g);
  d[2][1] = -wG * (dg_dr * ey * vy);
  // radius parameter: a = 0.05 rr
  d[3][0] = wG * dg_da * kRadiusScale * vx;
  d[3][1] = wG * dg_da * kRadiusScale * vy;
}

// Face a (0: x-faces (n0+1, n1); 1: y-faces (n0, n1+1)) -> its two adjacent
// cells (clamped), in the reference's sampling order (right/top first).
DFM_HD void face_cells(int a, int i, int j, int n0, int n1, int c[2][2]) {
  if (a == 0) {
    c[0][0] = min(max(i, 0), n0 - 1);
    c[1][0] = min(max(i - 1, 0), n0 - 1);
    c[0][1] = c[1][1] = j;
  } else {
    c[0][0] = c[1][0] = i;
    c[0][1] = min(max(j, 0), n1 - 1);
    c[1][1] = min(max(j - 1, 0), n1 - 1);
  }
}

__global__ void __launch_bounds__(kInitB) k_vortex_faces(int n0, int n1, float dx, const float* __restrict__ th, int K,
                                                         float* __restrict__ ux, float* __restrict__ uy) {
  const uint32_t nfx = (uint32_t)(n0 + 1) * n1, nfy = (uint32_t)n0 * (n1 + 1);
  const uint32_t g = blockIdx.x * kInitB + threadIdx.x;
  if (g >= nfx + nfy) return;
  const int a = g < nfx ? 0 : 1;
  const uint32_t id = a == 0 ? g : g - nfx;
  const int s1 = a == 0 ? n1 : n1 + 1;
  const int i = (int)(id / (uint32_t)s1), j = (int)(id - (uint32_t)i * s1);
  int c[2][2];
  face_cells(a, i, j, n0, n1, c);
  double u0[2], u1[2];
  cell_velocity(th, K, (c[0][0] + 0.5) * dx, (c[0][1] + 0.5) * dx, u0);
  cell_velocity(th, K, (c[1][0] + 0.5) * dx, (c[1][1] + 0.5) * dx, u1);
  const float v = (float)(0.5 * (u0[a] + u1[a]));
  if (a == 0) ux[id] = v;
  else uy[id] = v;
}

// dtheta[p K + k] = sum over faces of lam_f * d u_f / d theta_(p, k).  One
// block per vortex; each thread walks a fixed set of faces and the block sum
// runs in a fixed order, so the result is deterministic.
__global__ void __launch_bounds__(kVjpB) k_vortex_vjp(int n0, int n1, float dx, const float* __restrict__ th, int K,
                                                      const float* __restrict__ lx, const float* __restrict__ ly,
                                                      float* __restrict__ dth) {
  const int k = blockIdx.x;
  const uint32_t nfx = (uint32_t)(n0 + 1) * n1, nfy = (uint32_t)n0 * (n1 + 1);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (uint32_t g = threadIdx.x; g < nfx + nfy; g += kVjpB) {
    const int a = g < nfx ? 0 : 1;
    const uint32_t id = a == 0 ? g : g - nfx;
    const int s1 = a == 0 ? n1 : n1 + 1;
    const int i = (int)(id / (uint32_t)s1), j = (int)(id - (uint32_t)i * s1);
    const double lam = (double)(a == 0 ? lx[id] : ly[id]);
    if (lam == 0.0) continue;
    int c[2][2];
    face_cells(a, i, j, n0, n1, c);
    double d0[4][2], d1[4][2];
    cell_velocity_grad(th, K, k, (c[0][0] + 0.5) * dx, (c[0][1] + 0.5) * dx, d0);
    cell_velocity_grad(th, K, k, (c[1][0] + 0.5) * dx, (c[1][1] + 0.5) * dx, d1);
#pragma unroll
    for (int p = 0; p < 4; ++p) acc[p] += lam * 0.5 * (d0[p][a] + d1[p][a]);
  }
  __shared__ double red[kVjpB / 64][4];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    double v = acc[p];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (l == 0) red[w][p] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double s = 0.0;
    for (int t = 0; t < kVjpB / 64; ++t) s += red[t][threadIdx.x];
    dth[threadIdx.x * K + k] = (float)s;
  }
}

int check_vortex_args(const dfm_grid* g, const float* theta, int K) {
  if (!g || !theta) return set_error(DFM_E_NULL, "vortex initializer: NULL argument");
  if (g->dim != 2) return set_error(DFM_E_ARG, "the vortex initializer is 2D");
  if (g->n[0] < 1 || g->n[1] < 1 || g->n[2] != 1) return set_error(DFM_E_ARG, "bad 2D grid");
  if ((int64_t)(g->n[0] + 1) * (g->n[1] + 1) * 2 >= (int64_t(1) << 31)) return set_error(DFM_E_ARG, "grid too large");
  if (K < 1 || K > 4096) return set_error(DFM_E_ARG, "vortex count must be in [1, 4096]");
  return DFM_OK;
}

}  // namespace

extern "C" int dfm_vortex_velocity_2d(const dfm_grid* g, const float* theta, int K, const dfm_mac* u, void* stream) {
  int rc = check_vortex_args(g, theta, K);
  if (rc != DFM_OK) return rc;
  if (!u || !u->c[0] || !u->c[1]) return set_error(DFM_E_NULL, "dfm_vortex_velocity_2d: NULL velocity");
  const int64_t nf = (int64_t)(g->n[0] + 1) * g->n[1] + (int64_t)g->n[0] * (g->n[1] + 1);
  hipLaunchKernelGGL(k_vortex_faces, dim3((unsigned)((nf + kInitB - 1) / kInitB)), dim3(kInitB), 0,
                     reinterpret_cast<hipStream_t>(stream), g->n[0], g->n[1], g->dx, theta, K, u->c[0], u->c[1]);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(DFM_E_HIP, "dfm_vortex_velocity_2d", hipGetErrorString(e));
  return DFM_OK;
}

extern "C" int dfm_vortex_velocity_2d_vjp(const dfm_grid* g, const float* theta, int K, const dfm_mac* lam,
                                          float* dtheta, void* stream) {
  int rc = check_vortex_args(g, theta, K);
  if (rc != DFM_OK) return rc;
  if (!lam || !lam->c[0] || !lam->c[1] || !dtheta) return set_error(DFM_E_NULL, "dfm_vortex_velocity_2d_vjp: NULL");
  hipLaunchKernelGGL(k_vortex_vjp, dim3(K), dim3(kVjpB), 0, reinterpret_cast<hipStream_t>(stream), g->n[0], g->n[1],
                     g->dx, theta, K, lam->c[0], lam->c[1], dtheta);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(DFM_E_HIP, "dfm_vortex_velocity_2d_vjp", hipGetErrorString(e));
  return DFM_OK;
}
