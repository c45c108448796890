// Obstacle masks applied to fields (gfx950): mask_velocity_by_boundary,
// mask_passive_by_boundary, mask_adj_velocity_by_boundary and
// mask_adj_passive_by_boundary (2D/vortex/init_conditions.py:548-585,
// 3D/init_conditions.py:535-578).
//
// A cell is solid when mask >= 1; the passive average takes the 3^d
// neighbours with mask <= 0 (the reference's two tests: cells with
// 0 < mask < 1 are neither).  Every face of a solid cell takes that cell's
// boundary velocity (0 for the adjoint form).  A face shared by two solid
// cells is written from both cells in the reference's unordered parallel
// loop; here the cell on the face's high side wins, which agrees whenever the
// two boundary velocities do (always, for the zero boundary velocity the
// simulators use).
#include "dfm_common.hpp"
#include "dfm_internal.hpp"
#include "../../include/dfm.h"

using namespace dfm;

namespace {

constexpr int kMB = 256;

DFM_HD uint32_t lin3(const int s[3], int i, int j, int k) {
  return ((uint32_t)i * (uint32_t)s[1] + (uint32_t)j) * (uint32_t)s[2] + (uint32_t)k;
}

// one thread per face of all dim face grids; bvel: SoA (dim, cells) or nullptr
template <int DIM>
__global__ void __launch_bounds__(kMB) k_mask_faces(Geo g, const float* __restrict__ mask,
                                                    const float* __restrict__ bvel, dfm_mac u) {
  int S[3][3];
  uint32_t nf[3];
  for (int a = 0; a < 3; ++a) {
    shape_of(g, a < DIM ? a : 3, S[a]);
    nf[a] = a < DIM ? (uint32_t)S[a][0] * S[a][1] * S[a][2] : 0u;
  }
  const int C[3] = {g.n[0], g.n[1], g.n[2]};
  const uint32_t cells = (uint32_t)C[0] * C[1] * C[2];
  uint32_t id = blockIdx.x * kMB + threadIdx.x;
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
  int I[3] = {i, j, k};
  // high-side cell = I, low-side cell = I - e_a
  int hiI[3] = {i, j, k}, loI[3] = {i, j, k};
  loI[a] -= 1;
  const bool has_hi = I[a] < C[a], has_lo = I[a] > 0;
  int src = -1;
  if (has_hi && mask[lin3(C, hiI[0], hiI[1], hiI[2])] >= 1.f) src = (int)lin3(C, hiI[0], hiI[1], hiI[2]);
  else if (has_lo && mask[lin3(C, loI[0], loI[1], loI[2])] >= 1.f) src = (int)lin3(C, loI[0], loI[1], loI[2]);
  if (src < 0) return;
  u.c[a][id] = bvel ? bvel[(size_t)a * cells + (uint32_t)src] : 0.f;
}

// solid cells: the mean of the fluid cells of the 3^d neighbourhood (0 if
// none), in the reference's loop order; adjoint form: 0.  Reads only fluid
// cells and writes only solid ones, so the launch is race-free.
template <int DIM>
__global__ void __launch_bounds__(kMB) k_mask_cells(Geo g, const float* __restrict__ mask, float* __restrict__ s,
                                                    int adjoint) {
  const int C[3] = {g.n[0], g.n[1], g.n[2]};
  const uint32_t id = blockIdx.x * kMB + threadIdx.x;
  if (id >= (uint32_t)C[0] * C[1] * C[2]) return;
  if (!(mask[id] >= 1.f)) return;
  if (adjoint) {
    s[id] = 0.f;
    return;
  }
  int i, j, k;
  unlin(C, id, i, j, k);
  float acc = 0.f, w = 0.f;
  for (int di = -1; di <= 1; ++di)
    for (int dj = -1; dj <= 1; ++dj)
      for (int dk = (DIM == 3 ? -1 : 0); dk <= (DIM == 3 ? 1 : 0); ++dk) {
        const int a = i + di, b = j + dj, c = k + dk;
        if (a < 0 || a >= C[0] || b < 0 || b >= C[1] || c < 0 || c >= C[2]) continue;
        const uint32_t q = lin3(C, a, b, c);
        if (mask[q] <= 0.f) {
          acc += s[q];
          w += 1.f;
        }
      }
  s[id] = w > 0.f ? acc / w : 0.f;
}

int check_grid_mask(const dfm_grid* g, const float* mask, Geo& G) {
  if (!g || !mask) return set_error(DFM_E_NULL, "mask: NULL argument");
  if (g->dim != 2 && g->dim != 3) return set_error(DFM_E_ARG, "dim must be 2 or 3");
  for (int a = 0; a < 3; ++a)
    if (g->n[a] < 1) return set_error(DFM_E_ARG, "grid extents must be >= 1");
  if (g->dim == 2 && g->n[2] != 1) return set_error(DFM_E_ARG, "2D grid must have n[2] == 1");
  if ((int64_t)(g->n[0] + 1) * (g->n[1] + 1) * (g->n[2] + 1) * 3 >= (int64_t(1) << 31))
    return set_error(DFM_E_ARG, "grid too large");
  G.dim = g->dim;
  for (int a = 0; a < 3; ++a) G.n[a] = g->n[a];
  G.dx = g->dx;
  G.inv_dx = 1.0f / g->dx;
  return DFM_OK;
}

}  // namespace

extern "C" int dfm_mask_velocity(const dfm_grid* g, const float* mask, const float* bvel, const dfm_mac* u,
                                 void* stream) {
  Geo G;
  int rc = check_grid_mask(g, mask, G);
  if (rc != DFM_OK) return rc;
  if (!u) return set_error(DFM_E_NULL, "dfm_mask_velocity: NULL velocity");
  for (int a = 0; a < g->dim; ++a)
    if (!u->c[a]) return set_error(DFM_E_NULL, "dfm_mask_velocity: NULL velocity component");
  int64_t nf = 0;
  for (int a = 0; a < g->dim; ++a) nf += count_of_host(G, a);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const unsigned grid = (unsigned)((nf + kMB - 1) / kMB);
  if (g->dim == 2) hipLaunchKernelGGL((k_mask_faces<2>), dim3(grid), dim3(kMB), 0, st, G, mask, bvel, *u);
  else hipLaunchKernelGGL((k_mask_faces<3>), dim3(grid), dim3(kMB), 0, st, G, mask, bvel, *u);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(DFM_E_HIP, "dfm_mask_velocity", hipGetErrorString(e));
  return DFM_OK;
}

extern "C" int dfm_mask_scalar(const dfm_grid* g, const float* mask, float* s, int adjoint, void* stream) {
  Geo G;
  int rc = check_grid_mask(g, mask, G);
  if (rc != DFM_OK) return rc;
  if (!s) return set_error(DFM_E_NULL, "dfm_mask_scalar: NULL field");
  const int64_t n = (int64_t)g->n[0] * g->n[1] * g->n[2];
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const unsigned grid = (unsigned)((n + kMB - 1) / kMB);
  if (g->dim == 2) hipLaunchKernelGGL((k_mask_cells<2>), dim3(grid), dim3(kMB), 0, st, G, mask, s, adjoint);
  else hipLaunchKernelGGL((k_mask_cells<3>), dim3(grid), dim3(kMB), 0, st, G, mask, s, adjoint);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(DFM_E_HIP, "dfm_mask_scalar", hipGetErrorString(e));
  return DFM_OK;
}
