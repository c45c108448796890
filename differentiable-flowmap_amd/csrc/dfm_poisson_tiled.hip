// Support for the PoissonSolver-compatible entry (3D/fast_mgpcg.py:121-158,
// 192, 258-273): the 8^3-tiled flat layout that MGPCG_3 hands to the external
// solver, and a check that the operator it passes is the all-Neumann 7-point
// Laplacian the MGPCG handle solves.
//
// Tiled layout (rearrange_tensor_for_cuda, 3D/fast_mgpcg.py:121-145): cell
// (i, j, k) = (8 t0 + l0, 8 t1 + l1, 8 t2 + l2) sits at
//     ((t0 T1 + t1) T2 + t2) * 512 + (l0 * 8 + l1) * 8 + l2,
// T_a = n_a / 8 tiles per axis.
//
// Operator (init / construct_adiag / construct_ax, 3D/fast_mgpcg.py:48-74,
// 162-190): a_diag = 6 - (number of axes on which the cell touches a wall,
// counting each touched end), a_x/a_y/a_z = -1 coupling a cell to its +axis
// neighbour and 0 on the last layer, is_dof = 1 everywhere.
#include "dfm_common.hpp"
#include "dfm_internal.hpp"
#include "../../include/dfm.h"

using namespace dfm;

namespace {

constexpr int kTB = 256;

DFM_HD uint32_t tiled_index(const int n[3], int i, int j, int k) {
  const uint32_t T1 = (uint32_t)n[1] >> 3, T2 = (uint32_t)n[2] >> 3;
  const uint32_t tile = (((uint32_t)i >> 3) * T1 + ((uint32_t)j >> 3)) * T2 + ((uint32_t)k >> 3);
  return tile * 512u + (((uint32_t)i & 7u) * 8u + ((uint32_t)j & 7u)) * 8u + ((uint32_t)k & 7u);
}

// one thread per cell, in dense order (coalesced dense side; each 8-cell run
// of k is one contiguous 32 B run on the tiled side)
__global__ void __launch_bounds__(kTB) k_tile8(Geo g, const float* __restrict__ src, float* __restrict__ dst,
                                              int to_tiles) {
  const uint32_t id = blockIdx.x * kTB + threadIdx.x;
  if (id >= (uint32_t)g.n[0] * g.n[1] * g.n[2]) return;
  int i, j, k;
  unlin(g.n, id, i, j, k);
  const uint32_t t = tiled_index(g.n, i, j, k);
  if (to_tiles)
    dst[t] = src[id];
  else
    dst[id] = src[t];
}

// counts cells whose coefficients differ from the all-Neumann operator
__global__ void __launch_bounds__(kTB) k_operator_check(Geo g, const float* __restrict__ ad,
                                                       const float* __restrict__ ax, const float* __restrict__ ay,
                                                       const float* __restrict__ az,
                                                       const uint8_t* __restrict__ dof, int tiled,
                                                       unsigned* __restrict__ bad) {
  const uint32_t id = blockIdx.x * kTB + threadIdx.x;
  bool wrong = false;
  if (id < (uint32_t)g.n[0] * g.n[1] * g.n[2]) {
    int I[3];
    unlin(g.n, id, I[0], I[1], I[2]);
    const uint32_t q = tiled ? tiled_index(g.n, I[0], I[1], I[2]) : id;
    float diag = 6.f;
    for (int a = 0; a < 3; ++a) {
      if (I[a] == 0 || I[a] == g.n[a] - 1) diag -= 1.f;
    }
    const float* off[3] = {ax, ay, az};
    wrong = ad[q] != diag || dof[q] != 1;
    for (int a = 0; a < 3; ++a) wrong = wrong || off[a][q] != (I[a] == g.n[a] - 1 ? 0.f : -1.f);
  }
  const unsigned long long m = __ballot(wrong);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(bad, (unsigned)__popcll(m));
}

int check_grid8(const dfm_grid* g, Geo& G) {
  if (!g) return set_error(DFM_E_NULL, "tiled Poisson: NULL grid");
  if (g->dim != 3) return set_error(DFM_E_ARG, "the tiled layout is 3D");
  for (int a = 0; a < 3; ++a)
    if (g->n[a] < 8 || g->n[a] % 8) return set_error(DFM_E_ARG, "tiled layout: every n must be a multiple of 8");
  if ((int64_t)g->n[0] * g->n[1] * g->n[2] >= (int64_t(1) << 31)) return set_error(DFM_E_ARG, "grid too large");
  G.dim = 3;
  for (int a = 0; a < 3; ++a) G.n[a] = g->n[a];
  G.dx = g->dx;
  G.inv_dx = g->dx > 0.f ? 1.0f / g->dx : 0.f;
  return DFM_OK;
}

}  // namespace

extern "C" int dfm_tile8_convert(const dfm_grid* g, const float* src, float* dst, int to_tiles, void* stream) {
  Geo G;
  int rc = check_grid8(g, G);
  if (rc != DFM_OK) return rc;
  if (!src || !dst) return set_error(DFM_E_NULL, "dfm_tile8_convert: NULL field");
  if (src == dst) return set_error(DFM_E_ARG, "dfm_tile8_convert: src and dst must differ");
  const int64_t N = (int64_t)G.n[0] * G.n[1] * G.n[2];
  hipLaunchKernelGGL(k_tile8, dim3((unsigned)((N + kTB - 1) / kTB)), dim3(kTB), 0,
                     static_cast<hipStream_t>(stream), G, src, dst, to_tiles ? 1 : 0);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? DFM_OK : set_error(DFM_E_HIP, "dfm_tile8_convert: launch failed", hipGetErrorString(e));
}

extern "C" int dfm_poisson_operator_check(const dfm_grid* g, const float* a_diag, const float* a_x, const float* a_y,
                                          const float* a_z, const uint8_t* is_dof, int tiled, int* mismatches,
                                          void* stream) {
  Geo G;
  int rc = check_grid8(g, G);
  if (rc != DFM_OK) return rc;
  if (!a_diag || !a_x || !a_y || !a_z || !is_dof || !mismatches)
    return set_error(DFM_E_NULL, "dfm_poisson_operator_check: NULL argument");
  hipStream_t st = static_cast<hipStream_t>(stream);
  unsigned* d_bad = nullptr;
  if (hipMallocAsync((void**)&d_bad, sizeof(unsigned), st) != hipSuccess)
    return set_error(DFM_E_HIP, "dfm_poisson_operator_check: allocation failed");
  unsigned h_bad = 0;
  hipError_t e = hipMemsetAsync(d_bad, 0, sizeof(unsigned), st);
  const int64_t N = (int64_t)G.n[0] * G.n[1] * G.n[2];
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_operator_check, dim3((unsigned)((N + kTB - 1) / kTB)), dim3(kTB), 0, st, G, a_diag, a_x,
                       a_y, a_z, is_dof, tiled ? 1 : 0, d_bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&h_bad, d_bad, sizeof(unsigned), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFreeAsync(d_bad, st);
  if (e != hipSuccess) return set_error(DFM_E_HIP, "dfm_poisson_operator_check failed", hipGetErrorString(e));
  *mismatches = (int)h_bad;
  return DFM_OK;
}
