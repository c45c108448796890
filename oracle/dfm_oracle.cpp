// CPU oracle for the leapfrog flow-map hot path -- TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// this library, as the checker or as the labelled CPU baseline.  The product
// (libdfm.so, HIP) never links, loads or calls it.
//
// It restates the reference's Taichi kernels tap by tap, in the same loop
// shape as the reference (non-separable weights N(x)N(y)N(z) per tap), so it
// shares no structure with the separable HIP kernels it checks.  Every entry
// point exists in a double (`_d`) and a float (`_f`) flavour: FP64 is the
// oracle, FP32 is the "reference arithmetic" mode and the CPU baseline.
//
// Field convention: a field of shape (sx, sy, sz) is C-contiguous with z
// fastest (2D fields use sz = 1 and 2-vectors); point arrays are AoS
// (npts, dim) like the reference's ti.Vector.field.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <omp.h>

namespace {

// B-spline kernels N_k and derivatives dN_k.
// linear:    2D/vortex/taichi_utils.py:210-218
// quadratic: 2D/vortex/taichi_utils.py:381-400 (3D/taichi_utils.py:356-375)
// cubic:     2D/vortex/taichi_utils.py:356-377 (3D/taichi_utils.py:410-431)
// quartic:   2D/vortex/taichi_utils.py:325-353 (3D/taichi_utils.py:379-407)
template <class R> R N1(R x) { return R(1) - std::fabs(x); }
template <class R> R dN1(R x) { return x < R(0) ? R(1) : R(-1); }
template <class R> R N2(R x) {
  R a = std::fabs(x);
  if (a < R(0.5)) return R(0.75) - a * a;
  if (a < R(1.5)) return R(0.5) * (R(1.5) - a) * (R(1.5) - a);
  return R(0);
}
template <class R> R dN2(R x) {
  R a = std::fabs(x), r = 0;
  if (a < R(0.5)) r = R(-2) * a;
  else if (a < R(1.5)) r = R(0.5) * (R(2) * a - R(3));
  return x < R(0) ? -r : r;
}
template <class R> R N3(R x) {
  R a = std::fabs(x);
  if (a < R(1)) return R(0.5) * a * a * a - a * a + R(2) / R(3);
  if (a < R(2)) return (R(2) - a) * (R(2) - a) * (R(2) - a) / R(6);
  return R(0);
}
template <class R> R dN3(R x) {
  R a = std::fabs(x), r = 0;
  if (a < R(1)) r = R(1.5) * a * a - R(2) * a;
  else if (a < R(2)) r = R(-0.5) * (R(2) - a) * (R(2) - a);
  return x < R(0) ? -r : r;
}
template <class R> R N4(R x) {
  R a = std::fabs(x), a2 = a * a, a3 = a2 * a, a4 = a2 * a2;
  if (a < R(0.5)) return R(115) / R(192) - R(5) * a2 / R(8) + a4 / R(4);
  if (a < R(1.5))
    return R(55) / R(96) + R(5) * a / R(24) - R(5) * a2 / R(4) + R(5) * a3 / R(6) - a4 / R(6);
  if (a < R(2.5))
    return R(625) / R(384) - R(125) * a / R(48) + R(25) * a2 / R(16) - R(5) * a3 / R(12) + a4 / R(24);
  return R(0);
}
template <class R> R dN4(R x) {
  R a = std::fabs(x), a2 = a * a, a3 = a2 * a, r = 0;
  if (a < R(0.5)) r = R(-1.25) * a + a3;
  else if (a < R(1.5)) r = R(5) / R(24) - R(2.5) * a + R(2.5) * a2 - R(2) / R(3) * a3;
  else if (a < R(2.5)) r = R(-125) / R(48) + R(25) / R(8) * a - R(1.25) * a2 + a3 / R(6);
  return x < R(0) ? -r : r;
}
template <class R> R Nk(int k, R x) { return k == 1 ? N1(x) : k == 2 ? N2(x) : (k == 3 ? N3(x) : N4(x)); }
template <class R> R dNk(int k, R x) { return k == 1 ? dN1(x) : k == 2 ? dN2(x) : (k == 3 ? dN3(x) : dN4(x)); }

// Tap window [lo, hi] relative to floor(s): linear 0..1, quadratic -1..2,
// cubic -2..3, quartic -3..4 (2D/vortex/taichi_utils.py:230-234,433,470,506).
inline void window(int k, int& lo, int& hi) {
  if (k == 1) { lo = 0; hi = 1; } else if (k == 2) { lo = -1; hi = 2; } else if (k == 3) { lo = -2; hi = 3; } else { lo = -3; hi = 4; }
}

template <class R> struct Fld {
  const R* d;
  int s[3];
};

// sample(): truncating float->int conversion then clamp to the field
// (2D/vortex/taichi_utils.py:125-130, 3D/taichi_utils.py:130-136).
template <class R> R sample(const Fld<R>& f, R u, R v, R w) {
  int i = std::max(0, std::min((int)u, f.s[0] - 1));
  int j = std::max(0, std::min((int)v, f.s[1] - 1));
  int k = std::max(0, std::min((int)w, f.s[2] - 1));
  return f.d[((int64_t)i * f.s[1] + j) * f.s[2] + k];
}

// interp_grad_{2,3,4}: value and gradient of a B-spline interpolant of field f
// whose node (0,..) sits at BL*dx.  The sample coordinate is clamped to
// [1, dim-2-eps] per axis before the weights are formed
// (2D/vortex/taichi_utils.py:413-520, 3D/taichi_utils.py:433-536); the
// linear interp_grad_1 clamps to [0, dim-1-eps] (2D/vortex/taichi_utils.py:249-281).
template <class R>
void interp(int kern, int dim, const Fld<R>& f, const R* bl, const R* p, R dx, R* val, R* grad) {
  const R eps = R(1e-6);
  const R lo_s = kern == 1 ? R(0) : R(1);
  const int hi_off = kern == 1 ? 1 : 2;
  R s[3] = {0, 0, 0}, fl[3] = {0, 0, 0};
  for (int a = 0; a < dim; ++a) {
    R u = p[a] / dx - bl[a];
    s[a] = std::max(lo_s, std::min(u, R(f.s[a] - hi_off) - eps));
    fl[a] = std::floor(s[a]);
  }
  int lo, hi;
  window(kern, lo, hi);
  R v = 0, g[3] = {0, 0, 0};
  if (dim == 2) {
    for (int i = lo; i <= hi; ++i)
      for (int j = lo; j <= hi; ++j) {
        R xi = s[0] - (fl[0] + i), yj = s[1] - (fl[1] + j);
        R value = sample(f, fl[0] + i, fl[1] + j, R(0));
        g[0] += R(1) / dx * (value * dNk(kern, xi) * Nk(kern, yj));
        g[1] += R(1) / dx * (value * Nk(kern, xi) * dNk(kern, yj));
        v += value * Nk(kern, xi) * Nk(kern, yj);
      }
  } else {
    for (int i = lo; i <= hi; ++i)
      for (int j = lo; j <= hi; ++j)
        for (int k = lo; k <= hi; ++k) {
          R xi = s[0] - (fl[0] + i), yj = s[1] - (fl[1] + j), zk = s[2] - (fl[2] + k);
          R value = sample(f, fl[0] + i, fl[1] + j, fl[2] + k);
          g[0] += R(1) / dx * (value * dNk(kern, xi) * Nk(kern, yj) * Nk(kern, zk));
          g[1] += R(1) / dx * (value * Nk(kern, xi) * dNk(kern, yj) * Nk(kern, zk));
          g[2] += R(1) / dx * (value * Nk(kern, xi) * Nk(kern, yj) * dNk(kern, zk));
          v += value * Nk(kern, xi) * Nk(kern, yj) * Nk(kern, zk);
        }
  }
  *val = v;
  if (grad)
    for (int a = 0; a < dim; ++a) grad[a] = g[a];
}

// MAC velocity sampler: component c lives on faces normal to axis c, with
// BL = 0 along c and 0.5 elsewhere (interp_u_MAC_grad,
// 2D/vortex/taichi_utils.py:636-640, 3D/taichi_utils.py:596-601).
template <class R> struct Mac {
  Fld<R> c[3];
  int dim;
  void init(int d, const int* n, const R* const* comps) {
    dim = d;
    for (int a = 0; a < d; ++a) {
      c[a].d = comps[a];
      for (int b = 0; b < 3; ++b) c[a].s[b] = (b < d) ? n[b] + (a == b ? 1 : 0) : 1;
    }
  }
  // u(p) and J[a][b] = d u_a / d x_b
  void eval(int kern, const R* p, R dx, R* u, R J[3][3]) const {
    for (int a = 0; a < dim; ++a) {
      R bl[3] = {R(0.5), R(0.5), R(0.5)};
      bl[a] = 0;
      interp(kern, dim, c[a], bl, p, dx, &u[a], J ? J[a] : nullptr);
    }
  }
};

// RK2/RK4 march of a point and (optionally) one Jacobian column T:
// psi += -dt u(psi);  T += -dt (grad u)(psi) T.
// RK4_grid / RK2_grid / *_only_psi: 2D/vortex/flowmap.py:125-215,
// 3D/flowmap.py:85-179.
template <class R>
void march_point(const Mac<R>& U, int order, R dx, R dt, R* psi, R* T) {
  const int d = U.dim;
  R neg = -dt;
  R J[3][3], u1[3], u2[3], u3[3], u4[3], k1[3], k2[3], k3[3], k4[3];
  R p1[3] = {0, 0, 0}, p2[3] = {0, 0, 0}, p3[3] = {0, 0, 0}, T1[3] = {0, 0, 0}, T2[3] = {0, 0, 0}, T3[3] = {0, 0, 0};
  auto JT = [&](R (*Jm)[3], const R* t, R* out) {
    for (int a = 0; a < d; ++a) {
      out[a] = 0;
      for (int b = 0; b < d; ++b) out[a] += Jm[a][b] * t[b];
    }
  };
  U.eval(2, psi, dx, u1, J);
  if (T) JT(J, T, k1);
  for (int a = 0; a < d; ++a) {
    p1[a] = psi[a] + R(0.5) * neg * u1[a];
    if (T) T1[a] = T[a] + R(0.5) * neg * k1[a];
  }
  U.eval(2, p1, dx, u2, J);
  if (T) JT(J, T1, k2);
  if (order == 2) {
    for (int a = 0; a < d; ++a) {
      psi[a] = psi[a] + neg * u2[a];
      if (T) T[a] = T[a] + neg * k2[a];
    }
    return;
  }
  for (int a = 0; a < d; ++a) {
    p2[a] = psi[a] + R(0.5) * neg * u2[a];
    if (T) T2[a] = T[a] + R(0.5) * neg * k2[a];
  }
  U.eval(2, p2, dx, u3, J);
  if (T) JT(J, T2, k3);
  for (int a = 0; a < d; ++a) {
    p3[a] = psi[a] + neg * u3[a];
    if (T) T3[a] = T[a] + neg * k3[a];
  }
  U.eval(2, p3, dx, u4, J);
  if (T) JT(J, T3, k4);
  for (int a = 0; a < d; ++a) {
    psi[a] = psi[a] + neg * R(1) / R(6) * (u1[a] + R(2) * u2[a] + R(2) * u3[a] + u4[a]);
    if (T) T[a] = T[a] + neg * R(1) / R(6) * (k1[a] + R(2) * k2[a] + R(2) * k3[a] + k4[a]);
  }
}

// Semi-Lagrangian back-trace of a point by RK2/RK4 with velocity u0
// (RK*_advect_u_grid / RK*_advect_scalar_grid, 2D/vortex/flowmap.py:233-329,
// 3D/flowmap.py:202-320).
template <class R> void backtrace(const Mac<R>& U, int order, R dx, R dt, const R* x, R* out) {
  const int d = U.dim;
  R v1[3], v2[3], v3[3], v4[3], p[3];
  U.eval(2, x, dx, v1, nullptr);
  for (int a = 0; a < d; ++a) p[a] = x[a] - v1[a] * dt * R(0.5);
  U.eval(2, p, dx, v2, nullptr);
  if (order == 2) {
    for (int a = 0; a < d; ++a) out[a] = x[a] - v2[a] * dt;
    return;
  }
  for (int a = 0; a < d; ++a) p[a] = x[a] - v2[a] * dt * R(0.5);
  U.eval(2, p, dx, v3, nullptr);
  for (int a = 0; a < d; ++a) p[a] = x[a] - v3[a] * dt;
  U.eval(2, p, dx, v4, nullptr);
  for (int a = 0; a < d; ++a) out[a] = x[a] - (v1[a] + R(2) * v2[a] + R(2) * v3[a] + v4[a]) / R(6) * dt;
}

template <class R> void fld_from(Fld<R>& f, const R* d, const int* shape, int dim) {
  f.d = d;
  for (int b = 0; b < 3; ++b) f.s[b] = b < dim ? shape[b] : 1;
}

// ---------------------------------------------------------------- MGPCG ----
// Restatement of MGPCG (2D/vortex/mgpcg.py:8-282, 3D/mgpcg.py:13-291) on an
// all-Neumann box: A = (#neighbours) I - adjacency; V-cycle preconditioner
// with red-black Gauss-Seidel; restriction by 1/(dim-1); injection
// prolongation; PCG scalars in double on the host (q10).
template <class R> struct Level {
  int n[3];
  std::vector<R> r, z;
  int64_t size() const { return (int64_t)n[0] * n[1] * n[2]; }
};

template <class R> struct MG {
  int dim, L, pre, bottom;
  std::vector<Level<R>> lv;
  std::vector<R> x, p, Ap;

  R nsum(const Level<R>& l, const std::vector<R>& f, int i, int j, int k, R& nn) const {
    R s = 0;
    nn = 0;
    int I[3] = {i, j, k};
    for (int a = 0; a < dim; ++a) {
      int lo[3] = {i, j, k}, hi[3] = {i, j, k};
      lo[a] -= 1;
      hi[a] += 1;
      if (I[a] < l.n[a] - 1) { s += f[((int64_t)hi[0] * l.n[1] + hi[1]) * l.n[2] + hi[2]]; nn += 1; }
      if (I[a] > 0) { s += f[((int64_t)lo[0] * l.n[1] + lo[1]) * l.n[2] + lo[2]]; nn += 1; }
    }
    return s;
  }
  void Amul(const Level<R>& l, const std::vector<R>& in, std::vector<R>& out) const {
#pragma omp parallel for collapse(2)
    for (int i = 0; i < l.n[0]; ++i)
      for (int j = 0; j < l.n[1]; ++j)
        for (int k = 0; k < l.n[2]; ++k) {
          R nn;
          R s = nsum(l, in, i, j, k, nn);
          int64_t id = ((int64_t)i * l.n[1] + j) * l.n[2] + k;
          out[id] = nn * in[id] - s;
        }
  }
  void smooth(int li, int phase) {
    Level<R>& l = lv[li];
#pragma omp parallel for collapse(2)
    for (int i = 0; i < l.n[0]; ++i)
      for (int j = 0; j < l.n[1]; ++j)
        for (int k = 0; k < l.n[2]; ++k) {
          if (((i + j + k) & 1) != phase) continue;
          R nn;
          R s = nsum(l, l.z, i, j, k, nn);
          int64_t id = ((int64_t)i * l.n[1] + j) * l.n[2] + k;
          l.z[id] = (l.r[id] + s) / nn;
        }
  }
  void restrict_(int li) {
    Level<R>& f = lv[li];
    Level<R>& c = lv[li + 1];
    std::fill(c.r.begin(), c.r.end(), R(0));
    std::fill(c.z.begin(), c.z.end(), R(0));
    // serial accumulation keeps the sum order deterministic
    for (int i = 0; i < f.n[0]; ++i)
      for (int j = 0; j < f.n[1]; ++j)
        for (int k = 0; k < f.n[2]; ++k) {
          R nn;
          R s = nsum(f, f.z, i, j, k, nn);
          int64_t id = ((int64_t)i * f.n[1] + j) * f.n[2] + k;
          R res = f.r[id] - (nn * f.z[id] - s);
          int64_t cid = ((int64_t)(i / 2) * c.n[1] + j / 2) * c.n[2] + (dim == 3 ? k / 2 : 0);
          c.r[cid] += res * R(1) / R(dim - 1);
        }
  }
  void prolong(int li) {
    Level<R>& f = lv[li];
    Level<R>& c = lv[li + 1];
#pragma omp parallel for collapse(2)
    for (int i = 0; i < f.n[0]; ++i)
      for (int j = 0; j < f.n[1]; ++j)
        for (int k = 0; k < f.n[2]; ++k) {
          int64_t id = ((int64_t)i * f.n[1] + j) * f.n[2] + k;
          int64_t cid = ((int64_t)(i / 2) * c.n[1] + j / 2) * c.n[2] + (dim == 3 ? k / 2 : 0);
          f.z[id] += c.z[cid];
        }
  }
  // apply_preconditioner, 2D/vortex/mgpcg.py:182-200
  void vcycle() {
    std::fill(lv[0].z.begin(), lv[0].z.end(), R(0));
    for (int l = 0; l < L - 1; ++l) {
      for (int i = 0; i < pre; ++i) { smooth(l, 0); smooth(l, 1); }
      restrict_(l);
    }
    for (int i = 0; i < bottom; ++i) { smooth(L - 1, 0); smooth(L - 1, 1); }
    for (int l = L - 2; l >= 0; --l) {
      prolong(l);
      for (int i = 0; i < pre; ++i) { smooth(l, 1); smooth(l, 0); }
    }
  }
  double dot(const std::vector<R>& a, const std::vector<R>& b) const {
    double s = 0;
    for (size_t i = 0; i < a.size(); ++i) s += (double)a[i] * (double)b[i];
    return s;
  }
  void recenter(std::vector<R>& r) const {
    double m = 0;
    for (size_t i = 0; i < r.size(); ++i) m += (double)r[i];
    m /= (double)r.size();
    for (size_t i = 0; i < r.size(); ++i) r[i] -= (R)m;
  }
};

template <class R>
int mgpcg_solve(int dim, const int* n, int levels, int pre, int bottom, const R* b, R* xout, int max_iters,
                double tol, int relative, int* iters_out, double* res_out) {
  MG<R> mg;
  mg.dim = dim;
  mg.L = levels;
  mg.pre = pre;
  mg.bottom = bottom;
  mg.lv.resize(levels);
  for (int l = 0; l < levels; ++l) {
    for (int a = 0; a < 3; ++a) mg.lv[l].n[a] = a < dim ? (n[a] >> l) : 1;
    mg.lv[l].r.assign(mg.lv[l].size(), R(0));
    mg.lv[l].z.assign(mg.lv[l].size(), R(0));
  }
  int64_t N = mg.lv[0].size();
  mg.x.assign(N, R(0));
  mg.p.assign(N, R(0));
  mg.Ap.assign(N, R(0));
  std::vector<R>& r = mg.lv[0].r;
  std::vector<R>& z = mg.lv[0].z;
  for (int64_t i = 0; i < N; ++i) r[i] = b[i];
  // solve(), 2D/vortex/mgpcg.py:202-282: all-Neumann -> recentre r
  mg.recenter(r);
  double bnorm2 = mg.dot(r, r);
  double thresh = relative ? tol * tol * bnorm2 : tol;
  mg.vcycle();
  for (int64_t i = 0; i < N; ++i) mg.p[i] = z[i];
  double old_zTr = mg.dot(z, r);
  const double eps = 1e-12;
  int it = 0;
  double rTr = bnorm2;
  int status = 1;
  if (bnorm2 == 0) status = 0;
  while (status && it < max_iters) {
    mg.Amul(mg.lv[0], mg.p, mg.Ap);
    double pAp = mg.dot(mg.p, mg.Ap);
    double alpha = old_zTr / (pAp + eps);
    for (int64_t i = 0; i < N; ++i) mg.x[i] += (R)alpha * mg.p[i];
    for (int64_t i = 0; i < N; ++i) r[i] -= (R)alpha * mg.Ap[i];
    rTr = mg.dot(r, r);
    if (rTr < thresh) { status = 0; ++it; break; }
    mg.recenter(r);
    mg.vcycle();
    double new_zTr = mg.dot(z, r);
    double beta = new_zTr / (old_zTr + eps);
    for (int64_t i = 0; i < N; ++i) mg.p[i] = z[i] + (R)beta * mg.p[i];
    old_zTr = new_zTr;
    ++it;
  }
  for (int64_t i = 0; i < N; ++i) xout[i] = mg.x[i];
  if (iters_out) *iters_out = it;
  if (res_out) *res_out = std::sqrt(rTr / (bnorm2 > 0 ? bnorm2 : 1.0));
  return status;
}

// ----------------------------------------------------- point-list drivers ----
template <class R>
void interp_points(int kern, int dim, const R* f, const int* shape, const R* bl, const R* pts, int64_t np, R dx,
                   R* val, R* grad) {
  Fld<R> F;
  fld_from(F, f, shape, dim);
#pragma omp parallel for
  for (int64_t i = 0; i < np; ++i) interp(kern, dim, F, bl, pts + i * dim, dx, val + i, grad ? grad + i * dim : nullptr);
}

template <class R>
void march_points(int dim, int order, const int* n, R dx, R dt, const R* ux, const R* uy, const R* uz, R* psi,
                  R* T, int64_t np) {
  Mac<R> U;
  const R* c[3] = {ux, uy, uz};
  U.init(dim, n, c);
#pragma omp parallel for
  for (int64_t i = 0; i < np; ++i) march_point(U, order, dx, dt, psi + i * dim, T ? T + i * dim : nullptr);
}

// out[i] = T[i] . u0(psi[i])   (advect_u: 2D/vortex/flowmap.py:83-97,
// 3D/flowmap.py:51-69; accumulate_init reuses it with F, phi, f); u0 sampled
// with the B-spline of order kern (2 in the reference's simulators)
template <class R>
void pullback_points(int kern, int dim, const int* n, R dx, const R* ux, const R* uy, const R* uz, const R* psi,
                     const R* T, int64_t np, R* out) {
  Mac<R> U;
  const R* c[3] = {ux, uy, uz};
  U.init(dim, n, c);
#pragma omp parallel for
  for (int64_t i = 0; i < np; ++i) {
    R u[3];
    U.eval(kern, psi + i * dim, dx, u, nullptr);
    R s = 0;
    for (int a = 0; a < dim; ++a) s += T[i * dim + a] * u[a];
    out[i] = s;
  }
}

// back-trace points x by vel (quadratic), then sample src (shape/bl given)
// with the B-spline of order kern
template <class R>
void sl_points(int kern, int dim, int order, const int* n, R dx, R dt, const R* ux, const R* uy, const R* uz,
               const R* src, const int* src_shape, const R* src_bl, const R* x, int64_t np, R* out) {
  Mac<R> U;
  const R* c[3] = {ux, uy, uz};
  U.init(dim, n, c);
  Fld<R> S;
  fld_from(S, src, src_shape, dim);
#pragma omp parallel for
  for (int64_t i = 0; i < np; ++i) {
    R p[3];
    backtrace(U, order, dx, dt, x + i * dim, p);
    interp(kern, dim, S, src_bl, p, dx, out + i, (R*)nullptr);
  }
}

}  // namespace

#define DFM_ORACLE_EXPORTS(R, SFX)                                                                                 \
  extern "C" void oracle_interp_##SFX(int kern, int dim, const R* f, const int* shape, const R* bl, const R* pts, \
                                      int64_t np, R dx, R* val, R* grad) {                                       \
    interp_points<R>(kern, dim, f, shape, bl, pts, np, dx, val, grad);                                           \
  }                                                                                                                \
  extern "C" void oracle_march_##SFX(int dim, int order, const int* n, R dx, R dt, const R* ux, const R* uy,      \
                                     const R* uz, R* psi, R* T, int64_t np) {                                     \
    march_points<R>(dim, order, n, dx, dt, ux, uy, uz, psi, T, np);                                              \
  }                                                                                                                \
  extern "C" void oracle_pullback_##SFX(int kern, int dim, const int* n, R dx, const R* ux, const R* uy,         \
                                        const R* uz, const R* psi, const R* T, int64_t np, R* out) {              \
    pullback_points<R>(kern, dim, n, dx, ux, uy, uz, psi, T, np, out);                                           \
  }                                                                                                                \
  extern "C" void oracle_sl_##SFX(int kern, int dim, int order, const int* n, R dx, R dt, const R* ux,           \
                                  const R* uy, const R* uz, const R* src, const int* src_shape, const R* src_bl,  \
                                  const R* x, int64_t np, R* out) {                                                \
    sl_points<R>(kern, dim, order, n, dx, dt, ux, uy, uz, src, src_shape, src_bl, x, np, out);                   \
  }                                                                                                                \
  extern "C" int oracle_mgpcg_##SFX(int dim, const int* n, int levels, int pre, int bottom, const R* b, R* x,     \
                                    int max_iters, double tol, int relative, int* iters, double* res) {           \
    return mgpcg_solve<R>(dim, n, levels, pre, bottom, b, x, max_iters, tol, relative, iters, res);              \
  }

DFM_ORACLE_EXPORTS(double, d)
DFM_ORACLE_EXPORTS(float, f)

// OpenMP threads of the oracle's parallel loops (the CPU baseline sets them
// explicitly rather than through the environment); returns the count in use.
extern "C" int oracle_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

// B-spline kernel (k = 1,2,3,4) or its derivative at x, for the kernel tests.
extern "C" double oracle_bspline_d(int k, int deriv, double x) { return deriv ? dNk<double>(k, x) : Nk<double>(k, x); }
