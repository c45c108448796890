// Per-point device routines shared by the single-grid flow-map kernels and
// the batched ones: the RK march of a point (and its Jacobian column), the
// semi-Lagrangian back-trace, and the paired derivative/value sample of the
// adjoint sources.
#pragma once
#include "dfm_common.hpp"

namespace dfm {

// Velocity sample of the march and back-trace: in 3D the components share
// their axis set-ups (mac_eval_shared, bitwise mac_eval).
template <bool GRAD, int DIM>
DFM_HD void vel_sample(const Geo& g, const Mac& u, const float p[3], float val[3], float J[3][3]) {
  if constexpr (DIM == 3) mac_eval_shared<GRAD>(g, u, p, val, J);
  else mac_eval<GRAD, DIM>(g, u, p, val, J);
}

// RK2 (midpoint) / RK4 march of point p and, with HAS_T, its Jacobian column
// T through velocity u over dt: p += -dt u(p), T += -dt (grad u)(p) T
// (RK_grid / RK_grid_only_psi, 2D/vortex/flowmap.py:125-215).
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
  vel_sample<HAS_T, DIM>(g, u, p, u1, J);
  if (HAS_T) JT(T, k1);
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    q[a] = p[a] + 0.5f * neg * u1[a];
    if (HAS_T) Tq[a] = T[a] + 0.5f * neg * k1[a];
  }
  vel_sample<HAS_T, DIM>(g, u, q, u2, J);
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
  vel_sample<HAS_T, DIM>(g, u, q, u3, J);
  if (HAS_T) JT(Tq, k3);
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    q[a] = p[a] + neg * u3[a];
    if (HAS_T) Tq[a] = T[a] + neg * k3[a];
  }
  vel_sample<HAS_T, DIM>(g, u, q, u4, J);
  if (HAS_T) JT(Tq, k4);
  const float sixth = 1.0f / 6.0f;
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    p[a] = p[a] + neg * sixth * (u1[a] + 2.0f * u2[a] + 2.0f * u3[a] + u4[a]);
    if (HAS_T) T[a] = T[a] + neg * sixth * (k1[a] + 2.0f * k2[a] + 2.0f * k3[a] + k4[a]);
  }
}

// RK back-trace of node x through vel over dt into p (RK2_/RK4_advect_*_grid,
// 2D/vortex/flowmap.py:233-329).
template <int DIM, int ORDER>
DFM_HD void backtrace(const Geo& g, const Mac& vel, float dt, const float x[3], float p[3]) {
  float v1[3], v2[3], v3[3], v4[3], q[3];
  vel_sample<false, DIM>(g, vel, x, v1, nullptr);
#pragma unroll
  for (int a = 0; a < DIM; ++a) q[a] = x[a] - v1[a] * dt * 0.5f;
  vel_sample<false, DIM>(g, vel, q, v2, nullptr);
  if (ORDER == 2) {
#pragma unroll
    for (int a = 0; a < DIM; ++a) p[a] = x[a] - v2[a] * dt;
    return;
  }
#pragma unroll
  for (int a = 0; a < DIM; ++a) q[a] = x[a] - v2[a] * dt * 0.5f;
  vel_sample<false, DIM>(g, vel, q, v3, nullptr);
#pragma unroll
  for (int a = 0; a < DIM; ++a) q[a] = x[a] - v3[a] * dt;
  vel_sample<false, DIM>(g, vel, q, v4, nullptr);
#pragma unroll
  for (int a = 0; a < DIM; ++a) p[a] = x[a] - (v1[a] + 2.0f * v2[a] + 2.0f * v3[a] + v4[a]) / 6.0f * dt;
}

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

}  // namespace dfm
