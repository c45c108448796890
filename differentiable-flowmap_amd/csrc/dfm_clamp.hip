// BFECC with the extrema clamp (BFECC_clamp = True): BFECC / BFECC_scalar
// followed by clamp_u (2D/vortex/flowmap.py:332-341,440-481; 3D/flowmap.py:323-334,405-452).
//
//   u1 = T.init(psi)                 -> tmp
//   err = (F.u1(phi) - init) / 2     -> err
//   c = u1 - T.err(psi)              -> out
//   out = clamp(c, min, max of u1 at the 2*dim axis neighbours)
//
// The neighbours are sampled with clamp-to-edge indices on the field's own
// grid (each MAC component on its face grid).  The clamp reads u1 from tmp and
// rewrites each out element from itself only, so it runs in place.
#include "dfm_common.hpp"
#include "dfm_internal.hpp"
#include "../../include/dfm.h"

using namespace dfm;

namespace {

constexpr int kCB = 256;

DFM_HD float clamp_at(const float* __restrict__ ref, const int S[3], int dim, int i, int j, int k, float v) {
  const int I[3] = {i, j, k};
  float lo = INFINITY, hi = -INFINITY;
  for (int a = 0; a < dim; ++a) {
    for (int d = -1; d <= 1; d += 2) {
      int J[3] = {I[0], I[1], I[2]};
      J[a] = min(max(I[a] + d, 0), S[a] - 1);
      const float w = ref[lin(S, J[0], J[1], J[2])];
      lo = fminf(lo, w);
      hi = fmaxf(hi, w);
    }
  }
  return fminf(fmaxf(v, lo), hi);
}

// all face grids in one launch
template <int DIM>
__global__ void __launch_bounds__(kCB) k_clamp_faces(Geo g, dfm_mac out, dfm_mac ref) {
  uint32_t nf[3];
  int S[3][3];
  for (int a = 0; a < 3; ++a) {
    shape_of(g, a < DIM ? a : 3, S[a]);
    nf[a] = a < DIM ? (uint32_t)S[a][0] * S[a][1] * S[a][2] : 0u;
  }
  uint32_t id = blockIdx.x * kCB + threadIdx.x;
  int a = 0;
  if (id >= nf[0]) {
    id -= nf[0];
    a = 1;
    if (id >= nf[1]) {
      id -= nf[1];
      a = 2;
    }
  }
  if (a >= DIM || id >= nf[a]) return;
  int i, j, k;
  unlin(S[a], id, i, j, k);
  out.c[a][id] = clamp_at(ref.c[a], S[a], DIM, i, j, k, out.c[a][id]);
}

__global__ void __launch_bounds__(kCB) k_clamp_cells(Geo g, float* __restrict__ out, const float* __restrict__ ref) {
  const uint32_t id = blockIdx.x * kCB + threadIdx.x;
  if (id >= (uint32_t)g.n[0] * g.n[1] * g.n[2]) return;
  int i, j, k;
  unlin(g.n, id, i, j, k);
  out[id] = clamp_at(ref, g.n, g.dim, i, j, k, out[id]);
}

int geo_of(const dfm_grid* g, Geo& G) {
  if (!g) return set_error(DFM_E_NULL, "NULL grid");
  if (g->dim != 2 && g->dim != 3) return set_error(DFM_E_ARG, "dim must be 2 or 3");
  G.dim = g->dim;
  for (int a = 0; a < 3; ++a) G.n[a] = g->n[a];
  G.dx = g->dx;
  G.inv_dx = g->dx > 0.f ? 1.0f / g->dx : 0.f;
  return DFM_OK;
}

int launched(const char* what) {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? DFM_OK : set_error(DFM_E_HIP, what, hipGetErrorString(e));
}

}  // namespace

extern "C" int dfm_bfecc_u_clamp(const dfm_grid* g, const dfm_mac* init, const dfm_flowmap* fwd,
                                 const dfm_flowmap* bwd, const dfm_mac* err, const dfm_mac* tmp, const dfm_mac* out,
                                 void* stream) {
  int rc;
  if (!tmp || !out) return set_error(DFM_E_NULL, "dfm_bfecc_u_clamp: NULL buffer");
  // tmp = u1 = T.init(psi); err = 0.5 F.u1(phi) - 0.5 init; out = u1 - T.err(psi)
  if ((rc = dfm_pullback_u(g, init, fwd, 1.0f, 0.0f, nullptr, tmp, stream))) return rc;
  if ((rc = dfm_pullback_u(g, tmp, bwd, 0.5f, -0.5f, init, err, stream))) return rc;
  if ((rc = dfm_pullback_u(g, err, fwd, -1.0f, 1.0f, tmp, out, stream))) return rc;
  Geo G;
  if ((rc = geo_of(g, G))) return rc;
  int64_t faces = 0;
  for (int a = 0; a < G.dim; ++a) faces += count_of_host(G, a);
  const unsigned blocks = (unsigned)((faces + kCB - 1) / kCB);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (G.dim == 2)
    hipLaunchKernelGGL(k_clamp_faces<2>, dim3(blocks), dim3(kCB), 0, st, G, *out, *tmp);
  else
    hipLaunchKernelGGL(k_clamp_faces<3>, dim3(blocks), dim3(kCB), 0, st, G, *out, *tmp);
  return launched("dfm_bfecc_u_clamp: launch failed");
}

extern "C" int dfm_bfecc_s_clamp(const dfm_grid* g, const float* init, const float* psi, const float* phi, float* err,
                                 float* tmp, float* out, void* stream) {
  int rc;
  if (!tmp || !out) return set_error(DFM_E_NULL, "dfm_bfecc_s_clamp: NULL buffer");
  if ((rc = dfm_pullback_s(g, init, psi, 1.0f, 0.0f, nullptr, tmp, stream))) return rc;
  if ((rc = dfm_pullback_s(g, tmp, phi, 0.5f, -0.5f, init, err, stream))) return rc;
  if ((rc = dfm_pullback_s(g, err, psi, -1.0f, 1.0f, tmp, out, stream))) return rc;
  Geo G;
  if ((rc = geo_of(g, G))) return rc;
  const int64_t cells = (int64_t)G.n[0] * G.n[1] * G.n[2];
  hipLaunchKernelGGL(k_clamp_cells, dim3((unsigned)((cells + kCB - 1) / kCB)), dim3(kCB), 0,
                     static_cast<hipStream_t>(stream), G, out, tmp);
  return launched("dfm_bfecc_s_clamp: launch failed");
}
