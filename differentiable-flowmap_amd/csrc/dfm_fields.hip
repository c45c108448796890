// MAC <-> cell-centre conversions and the maximum cell speed (gfx950):
// get_central_vector / split_central_vector (2D/vortex/taichi_utils.py:152-167,
// 3D/taichi_utils.py:163-183) and calc_max_speed
// (2D/vortex/lfm_midpoint_simulator.py:328-335, 3D/lfm_midpoint_simulator.py:347-355).
//
// Cell-centred vectors are SoA, component a of cell c at [a * cells + c].
#include "dfm_common.hpp"
#include "dfm_internal.hpp"
#include "../../include/dfm.h"

using namespace dfm;

namespace {

constexpr int kEB = 256;

// central_a(c) = 0.5 * (u_a[c + e_a] + u_a[c])
template <int DIM>
__global__ void __launch_bounds__(kEB) k_central(Geo g, dfm_mac u, float* __restrict__ central) {
#pragma clang fp contract(off)
  const uint32_t cells = (uint32_t)g.n[0] * g.n[1] * g.n[2];
  const uint32_t id = (uint32_t)xcd_block(blockIdx.x, gridDim.x) * kEB + threadIdx.x;
  if (id >= cells) return;
  int I[3];
  unlin(g.n, id, I[0], I[1], I[2]);
  for (int a = 0; a < DIM; ++a) {
    int S[3];
    shape_of(g, a, S);
    int J[3] = {I[0], I[1], I[2]};
    const float lo = u.c[a][lin(S, J[0], J[1], J[2])];
    J[a] += 1;
    const float hi = u.c[a][lin(S, J[0], J[1], J[2])];
    central[(size_t)a * cells + id] = 0.5f * (hi + lo);
  }
}

// u_a(face) = 0.5 * (central_a(cell on the high side) + central_a(cell on the
// low side)), the missing cell clamped to the nearest one (sample's clamp)
template <int DIM>
__global__ void __launch_bounds__(kEB) k_split(Geo g, const float* __restrict__ central, dfm_mac u) {
#pragma clang fp contract(off)
  const uint32_t cells = (uint32_t)g.n[0] * g.n[1] * g.n[2];
  uint32_t nf[3];
  int S[3][3];
  for (int a = 0; a < 3; ++a) {
    shape_of(g, a < DIM ? a : 3, S[a]);
    nf[a] = a < DIM ? (uint32_t)S[a][0] * S[a][1] * S[a][2] : 0u;
  }
  uint32_t id = blockIdx.x * kEB + threadIdx.x;
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
  int I[3];
  unlin(S[a], id, I[0], I[1], I[2]);
  int hi[3] = {I[0], I[1], I[2]}, lo[3] = {I[0], I[1], I[2]};
  hi[a] = min(I[a], g.n[a] - 1);
  lo[a] = max(I[a] - 1, 0);
  const float* c = central + (size_t)a * cells;
  u.c[a][id] = 0.5f * (c[lin(g.n, hi[0], hi[1], hi[2])] + c[lin(g.n, lo[0], lo[1], lo[2])]);
}

__global__ void k_speed_init(unsigned* __restrict__ bits) {
  if (threadIdx.x == 0) *bits = __float_as_uint(1.0e-3f);
}

// max over cells of |central velocity|; speeds are >= 0 (or NaN), so their
// bit patterns order as unsigned integers and atomicMax on them is exact
template <int DIM>
__global__ void __launch_bounds__(kEB) k_max_speed(Geo g, dfm_mac u, unsigned* __restrict__ bits) {
#pragma clang fp contract(off)
  const uint32_t cells = (uint32_t)g.n[0] * g.n[1] * g.n[2];
  unsigned best = 0u;
  for (uint32_t id = blockIdx.x * kEB + threadIdx.x; id < cells; id += gridDim.x * kEB) {
    int I[3];
    unlin(g.n, id, I[0], I[1], I[2]);
    float s2 = 0.f;
    for (int a = 0; a < DIM; ++a) {
      int S[3];
      shape_of(g, a, S);
      int J[3] = {I[0], I[1], I[2]};
      const float lo = u.c[a][lin(S, J[0], J[1], J[2])];
      J[a] += 1;
      const float v = 0.5f * (lo + u.c[a][lin(S, J[0], J[1], J[2])]);
      s2 = s2 + v * v;
    }
    best = max(best, __float_as_uint(sqrtf(s2)));
  }
  for (int o = 32; o > 0; o >>= 1) best = max(best, (unsigned)__shfl_down((int)best, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(bits, best);
}

// Non-finite entries of up to kMaxCount fields, summed into *count (which
// the first launch zeroes): a grid-stride walk per field, one atomic per wave.
constexpr int kMaxCount = 8;
struct CountArgs {
  const float* f[kMaxCount];
  int64_t n[kMaxCount];
  int nf;
};
__global__ void k_count_zero(unsigned* __restrict__ count) {
  if (threadIdx.x == 0) *count = 0u;
}
__global__ void __launch_bounds__(kEB) k_count_nonfinite(CountArgs A, unsigned* __restrict__ count) {
  unsigned c = 0;
  for (int m = 0; m < A.nf; ++m) {
    const float* __restrict__ f = A.f[m];
    for (int64_t i = (int64_t)blockIdx.x * kEB + threadIdx.x; i < A.n[m]; i += (int64_t)gridDim.x * kEB)
      c += isfinite(f[i]) ? 0u : 1u;
  }
  for (int o = 32; o > 0; o >>= 1) c += (unsigned)__shfl_down((int)c, o, 64);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, c);
}

int check_grid_fields(const dfm_grid* g, Geo& G) {
  if (!g) return set_error(DFM_E_NULL, "NULL grid");
  if (g->dim != 2 && g->dim != 3) return set_error(DFM_E_ARG, "dim must be 2 or 3");
  for (int a = 0; a < 3; ++a)
    if (g->n[a] < 1) return set_error(DFM_E_ARG, "grid extents must be >= 1");
  if (g->dim == 2 && g->n[2] != 1) return set_error(DFM_E_ARG, "2D grid must have n[2] == 1");
  if ((int64_t)(g->n[0] + 1) * (g->n[1] + 1) * (g->n[2] + 1) * 3 >= (int64_t(1) << 31))
    return set_error(DFM_E_ARG, "grid too large");
  G.dim = g->dim;
  for (int a = 0; a < 3; ++a) G.n[a] = g->n[a];
  G.dx = g->dx;
  G.inv_dx = g->dx > 0.f ? 1.0f / g->dx : 0.f;
  return DFM_OK;
}

int check_mac_fields(const dfm_mac* u, int dim, const char* what) {
  if (!u) return set_error(DFM_E_NULL, what);
  for (int a = 0; a < dim; ++a)
    if (!u->c[a]) return set_error(DFM_E_NULL, what);
  return DFM_OK;
}

int launched(const char* what) {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? DFM_OK : set_error(DFM_E_HIP, what, hipGetErrorString(e));
}

}  // namespace

extern "C" int dfm_central_vector(const dfm_grid* g, const dfm_mac* u, float* central, void* stream) {
  Geo G;
  int rc = check_grid_fields(g, G);
  if (rc != DFM_OK) return rc;
  if ((rc = check_mac_fields(u, g->dim, "dfm_central_vector: NULL velocity")) != DFM_OK) return rc;
  if (!central) return set_error(DFM_E_NULL, "dfm_central_vector: NULL output");
  const int64_t cells = (int64_t)G.n[0] * G.n[1] * G.n[2];
  const unsigned blocks = (unsigned)((cells + kEB - 1) / kEB);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (G.dim == 2)
    hipLaunchKernelGGL(k_central<2>, dim3(blocks), dim3(kEB), 0, st, G, *u, central);
  else
    hipLaunchKernelGGL(k_central<3>, dim3(blocks), dim3(kEB), 0, st, G, *u, central);
  return launched("dfm_central_vector: launch failed");
}

extern "C" int dfm_split_central_vector(const dfm_grid* g, const float* central, const dfm_mac* u, void* stream) {
  Geo G;
  int rc = check_grid_fields(g, G);
  if (rc != DFM_OK) return rc;
  if ((rc = check_mac_fields(u, g->dim, "dfm_split_central_vector: NULL velocity")) != DFM_OK) return rc;
  if (!central) return set_error(DFM_E_NULL, "dfm_split_central_vector: NULL input");
  int64_t faces = 0;
  for (int a = 0; a < G.dim; ++a) faces += count_of_host(G, a);
  const unsigned blocks = (unsigned)((faces + kEB - 1) / kEB);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (G.dim == 2)
    hipLaunchKernelGGL(k_split<2>, dim3(blocks), dim3(kEB), 0, st, G, central, *u);
  else
    hipLaunchKernelGGL(k_split<3>, dim3(blocks), dim3(kEB), 0, st, G, central, *u);
  return launched("dfm_split_central_vector: launch failed");
}

extern "C" int dfm_max_speed(const dfm_grid* g, const dfm_mac* u, float* max_speed, void* stream) {
  Geo G;
  int rc = check_grid_fields(g, G);
  if (rc != DFM_OK) return rc;
  if ((rc = check_mac_fields(u, g->dim, "dfm_max_speed: NULL velocity")) != DFM_OK) return rc;
  if (!max_speed) return set_error(DFM_E_NULL, "dfm_max_speed: NULL output");
  unsigned* bits = reinterpret_cast<unsigned*>(max_speed);
  const int64_t cells = (int64_t)G.n[0] * G.n[1] * G.n[2];
  int64_t b = (cells + kEB - 1) / kEB;
  const unsigned blocks = (unsigned)(b < 1024 ? b : 1024);
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_speed_init, dim3(1), dim3(64), 0, st, bits);
  if (G.dim == 2)
    hipLaunchKernelGGL(k_max_speed<2>, dim3(blocks), dim3(kEB), 0, st, G, *u, bits);
  else
    hipLaunchKernelGGL(k_max_speed<3>, dim3(blocks), dim3(kEB), 0, st, G, *u, bits);
  return launched("dfm_max_speed: launch failed");
}

extern "C" int dfm_count_nonfinite(int nfields, const float* const* fields, const int64_t* sizes, unsigned* count,
                                   void* stream) {
  if (nfields < 0 || nfields > kMaxCount) return set_error(DFM_E_ARG, "dfm_count_nonfinite: 0..8 fields");
  if (!count || (nfields && (!fields || !sizes))) return set_error(DFM_E_NULL, "dfm_count_nonfinite: NULL argument");
  CountArgs A{};
  A.nf = nfields;
  int64_t most = 0;
  for (int m = 0; m < nfields; ++m) {
    if (!fields[m] || sizes[m] < 0) return set_error(DFM_E_ARG, "dfm_count_nonfinite: NULL field or negative size");
    A.f[m] = fields[m];
    A.n[m] = sizes[m];
    most = sizes[m] > most ? sizes[m] : most;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_count_zero, dim3(1), dim3(64), 0, st, count);
  int64_t b = (most + kEB - 1) / kEB;
  const unsigned blocks = (unsigned)(b < 1 ? 1 : (b < 2048 ? b : 2048));
  hipLaunchKernelGGL(k_count_nonfinite, dim3(blocks), dim3(kEB), 0, st, A, count);
  return launched("dfm_count_nonfinite: launch failed");
}
