// MGPCG pressure projection for the all-Neumann box (MI355X / gfx950).
//
// Algorithm (restated from MGPCG, 2D/vortex/mgpcg.py:8-282 and
// 3D/mgpcg.py:13-291): PCG on A x = b, A = (#neighbours) I - adjacency on
// cells, preconditioned by one multigrid V-cycle with red-black Gauss-Seidel
// (colour = (i+j+k) & 1), residual restriction by summing the 2^d children
// and scaling by 1/(dim-1), piecewise-constant prolongation, `bottom` RBGS
// sweeps on the coarsest level; the residual is recentred to mean zero before
// every preconditioner application.
//
// MI355X mapping:
//  * PCG scalars (alpha, beta, r.r, z.r, mean) live on the device in FP64 and
//    are produced by deterministic grid reductions inside the kernels that
//    compute the summands: every block stores its partial sums, and the last
//    block to arrive adds them in block order and runs the scalar step of the
//    recurrence.  No separate finalize launch and no per-iteration host round
//    trip; a device "done" flag turns every later kernel of the solve into a
//    no-op, so the host only polls the flag.
//  * Fusions: divergence + apply_bc + partial sums; the first pre-smoothing
//    sweep from z = 0 is one kernel (black cells read neighbour r/nn directly);
//    prolongation is fused into the first post-smoothing half-sweep (the
//    colour whose value is overwritten next is never materialised); the
//    recentring of r is applied on the fly by the level-0 V-cycle kernels and
//    written back, together with the z.r reduction and beta, by the last
//    level-0 half-sweep; the coarse levels run in one LDS-resident workgroup.
//  * Cell indices are decomposed with invariant-divisor multiplications.
#include "dfm_common.hpp"
#include "dfm_internal.hpp"
#include "../../include/dfm.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>

using namespace dfm;

namespace {

constexpr int kB = 256;          // threads per block
constexpr int kRedBlocks = 2048;  // most blocks of a reducing kernel (grid-stride beyond)
constexpr int kUnroll = 4;        // cells per step of a long reducing stencil walk, loads in flight together
constexpr unsigned kFluidBit = 1u << 31;  // nbr_load masks: marks a fluid cell
// Coarse levels of at most this many cells run each RBGS sweep as one launch
// (k_sweep2); their half-sweeps are launch-latency bound.
constexpr int64_t kSweep2MaxCells = int64_t(1) << 19;
constexpr unsigned kCtrGroups = 64;  // arrival-counter groups of a grid reduction
constexpr unsigned kCtrStride = 32;  // counters 128 B apart
constexpr unsigned kCtrWords = (kCtrGroups + 1) * kCtrStride;
constexpr int kMaxLevels = 12;
// alpha = z.r / p.Ap and beta = z.r' / z.r with a guard against a zero
// denominator only.  The reference adds eps = 1e-12 to both denominators
// (2D/vortex/mgpcg.py:204,240,272), which is harmless with its absolute stop
// r.r < 1e-12 but not with a relative one: on a nearly divergence-free field
// (b of order 1e-5 per cell) z.r falls below 1e-12 while r.r is still 1e-7
// of b.b, the guard then dominates alpha, and PCG stalls at that level.
DFM_HD double safe_ratio(double num, double den) { return den != 0.0 ? num / den : 0.0; }

enum Slot { S_ZR = 0, S_ALPHA, S_BETA, S_RR, S_BB, S_THRESH, S_MEAN, S_PAP, S_NFLUID, S_NSLOT };
enum Flag { F_DONE = 0, F_ITERS, F_STATUS, F_NTRUE, F_NFLAG };  // F_NTRUE: failed true-residual checks
// The scalars and the flags share one device block (scalars first) and one
// pinned host mirror, so one copy reads both back.
constexpr size_t kCtlBytes = sizeof(double) * S_NSLOT + sizeof(int) * F_NFLAG;

struct Lvl {
  int n[3];
  int64_t size;
  float* r;
  float* z;
  float* z2;  // second iterate buffer of the fused-sweep levels (else nullptr)
};

struct Dims {
  int n[3];
  FastDiv f2;  // / n[2]
  FastDiv f1;  // / n[1]
  FastDiv fh;  // / cell pairs along the innermost axis ((inner + 1) / 2)
  const uint8_t* bm;  // obstacle flags (1 = solid) of this level, or nullptr
  // Colour-split storage (level 0 only; 0 = natural order): cell (i, j, k)
  // of colour c = (i + j + k) & 1 and natural index id is stored at
  // c * hh + id / 2, hh = cells / 2.  Needs an even innermost extent, so
  // each row holds half its cells in each colour and id / 2 numbers a
  // colour's cells densely.  A half-sweep then reads and writes only the
  // cache lines of the colours it uses.
  uint32_t hh;
};

inline Dims make_dims(int n0, int n1, int n2, int dim) {
  Dims d;
  d.n[0] = n0;
  d.n[1] = n1;
  d.n[2] = n2;
  d.f2 = make_fastdiv((uint32_t)n2);
  d.f1 = make_fastdiv((uint32_t)n1);
  const int inner = dim == 3 ? n2 : n1;
  d.fh = make_fastdiv((uint32_t)((inner + 1) >> 1));
  d.bm = nullptr;
  d.hh = 0;
  return d;
}

// 32-bit: every grid (faces included) stays below 2^31 points (checked at creation)
DFM_HD uint32_t cidx(const Dims& d, int i, int j, int k) {
  return ((uint32_t)i * (uint32_t)d.n[1] + (uint32_t)j) * (uint32_t)d.n[2] + (uint32_t)k;
}

template <bool MASK> DFM_HD bool is_fluid(const Dims& d, uint32_t id) { return !MASK || d.bm[id] == 0; }

// Storage index of cell (i, j, k): natural, or colour-split (Dims::hh).
template <bool SPLIT>
DFM_HD uint32_t cell_at(const Dims& d, int i, int j, int k) {
  const uint32_t id = cidx(d, i, j, k);
  if (!SPLIT) return id;
  return (((i + j + k) & 1) ? d.hh : 0u) + (id >> 1);
}

// Storage indices of the 2*DIM neighbours of colour-split cell (i, j, k) at
// storage index id (order -x, +x, -y, +y, -z, +z) and whether each is in the
// grid.  The neighbours have the other colour; along the innermost axis the
// neighbour at natural index id_n +- 1 is at (id_n +- 1) / 2, which is the
// cell's own half index plus (parity - 1) or parity.
template <int DIM>
DFM_HD void split_nbrs(const Dims& d, int i, int j, int k, uint32_t id, uint32_t q[6], bool in[6]) {
  const uint32_t H = d.hh;
  const bool c1 = id >= H;
  const uint32_t m = c1 ? id - H : id;
  const uint32_t o = (c1 ? 0u : H) + m;  // same half index, other colour
  if (DIM == 3) {
    const uint32_t hj = (uint32_t)d.n[2] >> 1, hi = (uint32_t)d.n[1] * hj, kp = (uint32_t)(k & 1);
    q[0] = o - hi; q[1] = o + hi; q[2] = o - hj; q[3] = o + hj; q[4] = o - 1 + kp; q[5] = o + kp;
    in[4] = k > 0;
    in[5] = k < d.n[2] - 1;
  } else {
    const uint32_t hi = (uint32_t)d.n[1] >> 1, jp = (uint32_t)(j & 1);
    q[0] = o - hi; q[1] = o + hi; q[2] = o - 1 + jp; q[3] = o + jp;
    q[4] = q[5] = 0;
    in[4] = in[5] = false;
  }
  in[0] = i > 0;
  in[1] = i < d.n[0] - 1;
  in[2] = j > 0;
  in[3] = j < d.n[1] - 1;
}

// Visit the in-grid neighbours of cell (i, j, k) = id, restricted to fluid
// cells when MASK (num_fluid_neighbors / neighbor_sum, 2D/vortex/mgpcg.py:92-115;
// the obstacle form num_fluid_neighbors2 / neighbor_sum2,
// 2D/vortex/mgpcg_solid.py:140-170), in the order -x, +x, -y, +y, -z, +z.
template <int DIM, bool MASK, typename F>
DFM_HD void for_nbrs(const Dims& d, int i, int j, int k, uint32_t id, F&& fn) {
  const uint32_t sy = (uint32_t)d.n[2], sx = (uint32_t)d.n[1] * sy;
  auto visit = [&](bool inside, uint32_t q, int a, int b, int c) {
    if (inside && is_fluid<MASK>(d, q)) fn(q, a, b, c);
  };
  visit(i > 0, id - sx, i - 1, j, k);
  visit(i < d.n[0] - 1, id + sx, i + 1, j, k);
  visit(j > 0, id - sy, i, j - 1, k);
  visit(j < d.n[1] - 1, id + sy, i, j + 1, k);
  if (DIM == 3) {
    visit(k > 0, id - 1, i, j, k - 1);
    visit(k < d.n[2] - 1, id + 1, i, j, k + 1);
  }
}

// neighbour sum and count
template <int DIM, bool MASK = false>
DFM_HD float nbr_sum(const float* __restrict__ f, const Dims& d, int i, int j, int k, float& nn) {
  float s = 0.f;
  nn = 0.f;
  for_nbrs<DIM, MASK>(d, i, j, k, cidx(d, i, j, k), [&](uint32_t q, int, int, int) {
    s += f[q];
    nn += 1.f;
  });
  return s;
}

template <int DIM, bool MASK = false>
DFM_HD float nbr_count(const Dims& d, int i, int j, int k) {
  if (!MASK) {
    float nn = 0.f;
    nn += (i > 0) + (i < d.n[0] - 1) + (j > 0) + (j < d.n[1] - 1);
    if (DIM == 3) nn += (k > 0) + (k < d.n[2] - 1);
    return nn;
  }
  float nn = 0.f;
  for_nbrs<DIM, MASK>(d, i, j, k, cidx(d, i, j, k), [&](uint32_t, int, int, int) { nn += 1.f; });
  return nn;
}

// The 2*DIM neighbour values of cell (i, j, k) in for_nbrs order, for
// walks that keep several cells' loads in flight: every neighbour is loaded
// unconditionally into v (an absent one reads the cell's own, in-grid entry;
// its value is not used), and bit t of the result is set for each neighbour
// present (in the grid and, MASK, fluid).  The caller passes all of a step's
// loaded values through pin() before using any: otherwise the compiler sinks
// each load into the branch of its only, conditional, use and waits for it
// there, one memory round trip per neighbour.
template <int DIM, bool MASK, bool SPLIT, typename T>
DFM_HD unsigned nbr_load(const Dims& d, const T* __restrict__ f, int i, int j, int k, uint32_t id, T v[2 * DIM]) {
  bool in[6];
  uint32_t q[6];
  if constexpr (SPLIT) {
    split_nbrs<DIM>(d, i, j, k, id, q, in);
  } else {
    const uint32_t sy = (uint32_t)d.n[2], sx = (uint32_t)d.n[1] * sy;
    in[0] = i > 0; in[1] = i < d.n[0] - 1; in[2] = j > 0; in[3] = j < d.n[1] - 1; in[4] = k > 0; in[5] = k < d.n[2] - 1;
    q[0] = id - sx; q[1] = id + sx; q[2] = id - sy; q[3] = id + sy; q[4] = id - 1; q[5] = id + 1;
  }
  unsigned m = 0;
#pragma unroll
  for (int t = 0; t < 2 * DIM; ++t) {
    const uint32_t qs = in[t] ? q[t] : id;
    v[t] = f[qs];
    m |= (in[t] && is_fluid<MASK>(d, qs)) ? 1u << t : 0u;
  }
  return m;
}

// Cell index -> (i, j, k); grids stay below 2^31 cells (checked at creation).
DFM_HD void cell_of(const Dims& d, uint32_t id, int& i, int& j, int& k) {
  uint32_t u = (uint32_t)id;
  uint32_t t = fdiv(u, d.f2);
  k = (int)(u - t * (uint32_t)d.n[2]);
  uint32_t q = fdiv(t, d.f1);
  j = (int)(t - q * (uint32_t)d.n[1]);
  i = (int)q;
}

// Pair t of cells along the innermost axis (cells 2h and 2h + 1 of a row):
// (i, j) of the row, the innermost index `own` of its colour-`phase` cell and
// `other` of the other cell.  Either may equal the row length (outside an
// odd-sized row).
template <int DIM>
DFM_HD void colour_pair(const Dims& d, uint32_t t, int phase, int& i, int& j, int& own, int& other) {
  const uint32_t half = d.fh.d;
  uint32_t o = fdiv(t, d.fh);
  int base = 2 * (int)(t - o * half);
  int off;
  if (DIM == 3) {
    uint32_t q = fdiv(o, d.f1);
    i = (int)q;
    j = (int)(o - q * (uint32_t)d.n[1]);
    off = (phase - i - j) & 1;
  } else {
    i = (int)o;
    j = 0;
    off = (phase - i) & 1;
  }
  own = base + off;
  other = base + 1 - off;
}

// Colour-`phase` cell of pair index t; false when it falls outside an
// odd-sized row.
template <int DIM>
DFM_HD bool colour_cell(const Dims& d, uint32_t t, int phase, int& i, int& j, int& k) {
  int own, other;
  colour_pair<DIM>(d, t, phase, i, j, own, other);
  if (DIM == 3) {
    k = own;
    return k < d.n[2];
  }
  j = own;
  k = 0;
  return j < d.n[1];
}

// Storage index -> (i, j, k) in either layout.
template <int DIM, bool SPLIT>
DFM_HD void cell_of_at(const Dims& d, uint32_t id, int& i, int& j, int& k) {
  if constexpr (SPLIT) {
    const bool c1 = id >= d.hh;
    (void)colour_cell<DIM>(d, c1 ? id - d.hh : id, c1 ? 1 : 0, i, j, k);
  } else {
    cell_of(d, id, i, j, k);
  }
}

// Storage index of the cell with natural index id in the level's layout.
DFM_HD uint32_t level0_at(const Dims& d, uint32_t id) {
  if (!d.hh) return id;
  int i, j, k;
  cell_of(d, id, i, j, k);
  return cell_at<true>(d, i, j, k);
}

// 1/nn for the neighbour counts that occur (1..6): one multiply instead of an
// IEEE divide per update.  A fluid cell walled in by obstacles (nn = 0) has
// no equation; it gets 0 (the reference would divide by zero).
DFM_HD float inv_count(float nn) {
  const int c = (int)nn;
  return c == 6 ? (1.0f / 6.0f)
       : c == 5 ? 0.2f
       : c == 4 ? 0.25f
       : c == 3 ? (1.0f / 3.0f)
       : c == 2 ? 0.5f
       : c == 1 ? 1.0f
                : 0.0f;
}

DFM_HD bool is_done(const int* flags) { return flags[F_DONE] != 0; }

// ---------------------------------------------------------- reductions ----
DFM_HD double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Block sum of up to two values over a BS-thread block; thread 0 gets the totals.
template <int BS = kB>
DFM_HD void block_sum2(double& a, double& b) {
  __shared__ double sa[BS / 64], sb[BS / 64];
  a = wave_sum(a);
  b = wave_sum(b);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) { sa[w] = a; sb[w] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = 0; b = 0;
    for (int t = 0; t < BS / 64; ++t) { a += sa[t]; b += sb[t]; }
  }
}

// Scalar steps of the PCG recurrence, run on the grid totals (a, b).
enum FinOp { FIN_MEAN = 0, FIN_BNORM, FIN_ZR_INIT, FIN_ALPHA, FIN_RR, FIN_BETA, FIN_TRUE, FIN_COUNT };

// Grid-reduction context: partial sums (2 x kRedBlocks), the arrival counters
// (zero between reductions), the FP64 scalars and the solve flags.
struct Red {
  double* part;
  unsigned* ctr;  // kCtrWords arrival counters, zero between reductions
  double* sc;
  int* flags;
  double inv_n;  // 1 / fine cells
  double tol;    // FIN_BNORM only
  int relative;  // FIN_BNORM only
};

DFM_HD void fin_op(const Red& R, int op, double a, double b) {
  double* sc = R.sc;
  int* flags = R.flags;
  switch (op) {
    case FIN_MEAN:  // a = sum r
      sc[S_MEAN] = a * R.inv_n;
      break;
    case FIN_BNORM:  // a = b.b (after recentring, which k_recenter_norm wrote back)
      sc[S_MEAN] = 0.0;
      sc[S_BB] = a;
      sc[S_RR] = a;
      sc[S_THRESH] = R.relative ? R.tol * R.tol * a : R.tol;
      flags[F_ITERS] = 0;
      flags[F_STATUS] = 0;
      flags[F_NTRUE] = 0;
      flags[F_DONE] = (a == 0.0) ? 1 : 0;
      break;
    case FIN_ZR_INIT:  // a = z.r
      sc[S_ZR] = a;
      break;
    case FIN_ALPHA:  // a = p.Ap
      sc[S_PAP] = a;
      sc[S_ALPHA] = safe_ratio(sc[S_ZR], a);
      break;
    case FIN_RR: {  // a = r.r, b = sum r
      sc[S_RR] = a;
      sc[S_MEAN] = b * R.inv_n;
      flags[F_ITERS] += 1;
      if (!isfinite(a)) {
        flags[F_STATUS] = 2;
        flags[F_DONE] = 1;
      } else if (a < sc[S_THRESH]) {
        flags[F_STATUS] = 1;
        flags[F_DONE] = 1;
      }
      break;
    }
    case FIN_BETA: {  // a = z.r (new)
      sc[S_BETA] = safe_ratio(a, sc[S_ZR]);
      sc[S_ZR] = a;
      break;
    }
    case FIN_COUNT:  // a = fluid cells of an obstacle mask
      sc[S_NFLUID] = a;
      break;
    case FIN_TRUE: {  // a = |b - A x|^2 in FP64, b = sum(b - A x): refinement check
      sc[S_RR] = a;
      sc[S_MEAN] = b * R.inv_n;
      if (!isfinite(a)) {
        flags[F_STATUS] = 2;
        flags[F_DONE] = 1;
      } else if (a < sc[S_THRESH]) {
        flags[F_STATUS] = 1;
        flags[F_DONE] = 1;
      } else {
        flags[F_STATUS] = 0;
        flags[F_DONE] = 0;
        flags[F_NTRUE] += 1;
      }
      break;
    }
  }
}

// Deterministic grid reduction of (a, b) followed by fin_op: every block
// stores its block sums and counts itself in; the last block to arrive adds
// the partials in block order and runs the scalar step, then returns the
// counter to zero for the next reduction on the stream.  Called by every
// thread of every block (no thread may have returned).  Grids of reducing
// kernels hold at most kRedBlocks blocks.
//
// Publication without a device-scope fence: a device-scope release fence
// writes back the whole (per-XCD, mutually non-coherent) L2 and costs
// ~100 us over a few thousand blocks.  Only the two partials need publishing,
// so they are stored as device-scope atomics (coherent stores that bypass the
// non-coherent L2 state), the store is waited for before the counter
// increment, and the last block reads them back as device-scope atomic loads.
// Everything else a reducing kernel writes is consumed by later kernels.
// Arrivals are counted in two levels (kCtrGroups group counters on separate
// cache lines, then one top counter) because increments of one address
// serialise (~10 ns each): a single counter would add ~40 us to a
// 4096-block reduction.
DFM_HD void wait_stores() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

template <int BS = kB>
DFM_HD void grid_reduce(const Red& R, int op, double a, double b) {
  static_assert(BS % 64 == 0 && kRedBlocks % BS == 0, "whole waves; the partials split evenly over the block");
  __shared__ int s_last;
  block_sum2<BS>(a, b);
  if (threadIdx.x == 0) {
    __hip_atomic_store(&R.part[blockIdx.x], a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&R.part[kRedBlocks + blockIdx.x], b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    wait_stores();  // both stores complete before the increment is issued
    const unsigned G = gridDim.x, g = blockIdx.x % kCtrGroups;
    const unsigned ngroups = G < kCtrGroups ? G : kCtrGroups;
    const unsigned members = G / kCtrGroups + (g < G % kCtrGroups ? 1u : 0u);
    unsigned* gc = R.ctr + g * kCtrStride;
    unsigned* top = R.ctr + kCtrGroups * kCtrStride;
    int last = 0;
    if (__hip_atomic_fetch_add(gc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == members - 1) {
      __hip_atomic_store(gc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      wait_stores();
      last = __hip_atomic_fetch_add(top, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == ngroups - 1;
      if (last) __hip_atomic_store(top, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    s_last = last;
  }
  __syncthreads();
  if (!s_last) return;
  // all of a thread's partials are requested before any is summed (a rolled
  // loop would wait one memory latency per element)
  constexpr int U = kRedBlocks / BS;
  double va[U], vb[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const unsigned t = threadIdx.x + u * BS;
    va[u] = t < gridDim.x ? __hip_atomic_load(&R.part[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
    vb[u] = t < gridDim.x ? __hip_atomic_load(&R.part[kRedBlocks + t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                          : 0.0;
  }
  double ta = 0, tb = 0;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    ta += va[u];
    tb += vb[u];
  }
  block_sum2<BS>(ta, tb);
  if (threadIdx.x == 0) fin_op(R, op, ta, tb);
}

// ------------------------------------------------------ projection ends ----
struct Faces {
  float* c[3];
};

// apply_bc + divergence (no 1/dx) + b = -div + partial sum of b
// (apply_bc/divergence, 2D/vortex/mgpcg.py:294-336; 3D/mgpcg.py:380-411).
// MASK: the obstacle form (MGPCG_2_solid.apply_bc/divergence,
// 2D/vortex/mgpcg_solid.py:404-431): every face of a solid cell is zeroed and
// a solid cell's divergence is 0.  A fluid cell reads a face next to a solid
// cell as 0, so the zeroing writes of other threads never feed a result.
template <int DIM, bool MASK>
__global__ void __launch_bounds__(kB) k_div(Dims d, Faces u, float* __restrict__ b, Red R) {
  double acc = 0;
  int64_t N = (int64_t)d.n[0] * d.n[1] * d.n[2];
  const uint32_t sy = (uint32_t)d.n[2], sx = (uint32_t)d.n[1] * sy;
  const SlabWalk<uint32_t> w = slab_walk<kB>((uint32_t)N);
  for (uint32_t id = w.first; id < w.end; id += w.stride) {
    int i, j, k;
    cell_of(d, id, i, j, k);
    // face grids: x (n0+1, n1, n2), y (n0, n1+1, n2), z (n0, n1, n2+1)
    Dims fx{{d.n[0] + 1, d.n[1], d.n[2]}}, fy{{d.n[0], d.n[1] + 1, d.n[2]}}, fz{{d.n[0], d.n[1], d.n[2] + 1}};
    const bool solid = !is_fluid<MASK>(d, id);
    // open faces: not a wall, and (MASK) neither adjacent cell solid
    const bool oxl = i > 0 && !solid && is_fluid<MASK>(d, id - sx);
    const bool oxr = i < d.n[0] - 1 && !solid && is_fluid<MASK>(d, id + sx);
    const bool oyl = j > 0 && !solid && is_fluid<MASK>(d, id - sy);
    const bool oyr = j < d.n[1] - 1 && !solid && is_fluid<MASK>(d, id + sy);
    float div = 0.f;
    float xl = oxl ? u.c[0][cidx(fx, i, j, k)] : 0.f;
    float xr = oxr ? u.c[0][cidx(fx, i + 1, j, k)] : 0.f;
    float yl = oyl ? u.c[1][cidx(fy, i, j, k)] : 0.f;
    float yr = oyr ? u.c[1][cidx(fy, i, j + 1, k)] : 0.f;
    div = (xr - xl) + (yr - yl);
    if (i == 0 || solid) u.c[0][cidx(fx, i, j, k)] = 0.f;
    if (i == d.n[0] - 1 || solid) u.c[0][cidx(fx, i + 1, j, k)] = 0.f;
    if (j == 0 || solid) u.c[1][cidx(fy, i, j, k)] = 0.f;
    if (j == d.n[1] - 1 || solid) u.c[1][cidx(fy, i, j + 1, k)] = 0.f;
    if (DIM == 3) {
      const bool ozl = k > 0 && !solid && is_fluid<MASK>(d, id - 1);
      const bool ozr = k < d.n[2] - 1 && !solid && is_fluid<MASK>(d, id + 1);
      float zl = ozl ? u.c[2][cidx(fz, i, j, k)] : 0.f;
      float zr = ozr ? u.c[2][cidx(fz, i, j, k + 1)] : 0.f;
      div += (zr - zl);
      if (k == 0 || solid) u.c[2][cidx(fz, i, j, k)] = 0.f;
      if (k == d.n[2] - 1 || solid) u.c[2][cidx(fz, i, j, k + 1)] = 0.f;
    }
    float bv = solid ? 0.f : -div;
    b[d.hh ? cell_at<true>(d, i, j, k) : id] = bv;  // level-0 layout
    acc += (double)bv;
  }
  grid_reduce(R, FIN_MEAN, acc, 0.0);
}

// copy b into r (for dfm_mgpcg_solve) with partial sum; solid cells get 0
template <bool MASK>
__global__ void __launch_bounds__(kB) k_load_b(Dims d, int64_t N, const float* __restrict__ b, float* __restrict__ r,
                                               const uint8_t* __restrict__ bm, Red R) {
  double acc = 0;
  const SlabWalk<uint32_t> w = slab_walk<kB>((uint32_t)N);
  for (uint32_t id = w.first; id < w.end; id += w.stride) {
    float v = (!MASK || bm[id] == 0) ? b[id] : 0.f;
    r[level0_at(d, id)] = v;
    acc += (double)v;
  }
  grid_reduce(R, FIN_MEAN, acc, 0.0);
}

// r -= mean (kept as b for the refinement residual); partial r.r.  Solid
// cells keep r = 0 (the mean is over fluid cells).  x starts at 0 without
// being written: the first x/r update of a solve (F_ITERS == 0) does not read
// it, and a solve that ends before any iteration (b = 0) zeroes it.
template <bool MASK>
__global__ void __launch_bounds__(kB) k_recenter_norm(int64_t N, float* __restrict__ r, float* __restrict__ bkeep,
                                                      const uint8_t* __restrict__ bm, Red R) {
  float m = (float)R.sc[S_MEAN];
  double acc = 0;
  const SlabWalk<uint32_t> w = slab_walk<kB>((uint32_t)N);
  for (uint32_t id = w.first; id < w.end; id += w.stride) {
    float v = (!MASK || bm[id] == 0) ? r[id] - m : 0.f;
    r[id] = v;
    bkeep[id] = v;
    acc += (double)v * (double)v;
  }
  grid_reduce(R, FIN_BNORM, acc, 0.0);
}

// subtract_grad_p + apply_bc (2D/vortex/mgpcg.py:338-356, 3D/mgpcg.py:413-439;
// obstacle form 2D/vortex/mgpcg_solid.py:433-461): interior faces
// u_a -= p[I] - p[I - e_a]; wall faces and (MASK) faces of solid cells = 0.
// One thread per cell updates the cell's low face on every axis (and the
// wall face at the high end of an axis it closes), so p, in FP64, is read
// once per cell rather than once per face grid.  Face (i, j, k) of axis a
// lies in the (n + e_a) face grid at: x, the cell's own index; y,
// id + i n2; z, id + i n1 + j.
template <int DIM, bool MASK>
__global__ void __launch_bounds__(kB) k_subgrad(Dims d, Faces u, const double* __restrict__ p) {
  const int64_t N = (int64_t)d.n[0] * d.n[1] * d.n[2];
  const uint32_t id = (uint32_t)xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
  if (id >= N) return;
  int i, j, k;
  cell_of(d, id, i, j, k);
  const uint32_t sy = (uint32_t)d.n[2], sx = (uint32_t)d.n[1] * sy;
  const bool fl = is_fluid<MASK>(d, id);
  const bool split = d.hh != 0;  // level-0 layout of p
  const uint32_t c1 = (uint32_t)((i + j + k) & 1);
  auto at = [&](uint32_t lin, uint32_t colour) { return split ? (colour ? d.hh : 0u) + (lin >> 1) : lin; };
  const double pc = p[at(id, c1)];
  // the pressure difference is formed in FP64: p grows like n |u| in these
  // units, so an FP32 p would cost ~n ulps of velocity
  auto face = [&](float* __restrict__ c, uint32_t f, bool inner, uint32_t lo, bool high_wall, uint32_t f_high) {
    const bool open = inner && fl && is_fluid<MASK>(d, lo);
    c[f] = open ? (float)((double)c[f] - (pc - p[at(lo, c1 ^ 1u)])) : 0.f;
    if (high_wall) c[f_high] = 0.f;
  };
  face(u.c[0], id, i > 0, id - sx, i == d.n[0] - 1, id + sx);
  const uint32_t fy = id + (uint32_t)i * sy;
  face(u.c[1], fy, j > 0, id - sy, j == d.n[1] - 1, fy + sy);
  if (DIM == 3) {
    const uint32_t fz = id + (uint32_t)i * (uint32_t)d.n[1] + (uint32_t)j;
    face(u.c[2], fz, k > 0, id - 1, k == d.n[2] - 1, fz + 1);
  }
}

// Iterative-refinement residual r = b - A x evaluated in FP64 (x is FP64;
// b is the recentred FP32 right-hand side); partial |r|^2 and sum r.
// Solid cells: r = 0 (b and the operator row are 0 there).
template <int DIM, bool MASK, int U, bool SPLIT>
__global__ void __launch_bounds__(kB) k_true_resid(Dims d, const double* __restrict__ x, const float* __restrict__ b,
                                                   float* __restrict__ r, Red R, int if_converged) {
  // only after convergence on the recursive residual (uniform over the grid:
  // the flags change only in the last block, after every block has read them)
  if (if_converged && !(R.flags[F_DONE] && R.flags[F_STATUS] == 1)) return;
  double rr = 0, rs = 0;
  int64_t N = (int64_t)d.n[0] * d.n[1] * d.n[2];
  const SlabWalk<uint32_t> w = slab_walk<kB>((uint32_t)N);
  // U cells of the walk per step, all their loads issued before any is
  // used; cells are still consumed in walk order, so U does not change
  // the result
  for (uint32_t id0 = w.first; id0 < w.end; id0 += U * w.stride) {
    double xn[U][2 * DIM], xv[U];
    float bv[U];
    unsigned has[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t id = id0 + u * w.stride;
      has[u] = 0;
      xv[u] = 0.0;
      bv[u] = 0.f;
      if (id < w.end && is_fluid<MASK>(d, id)) {
        int i, j, k;
        cell_of_at<DIM, SPLIT>(d, id, i, j, k);
        has[u] = nbr_load<DIM, MASK, SPLIT>(d, x, i, j, k, id, xn[u]) | kFluidBit;
        xv[u] = x[id];
        bv[u] = b[id];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) pin(xn[u]);
    pin(xv);
    pin(bv);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t id = id0 + u * w.stride;
      if (id < w.end) {
        double v = 0.0;
        if (has[u]) {
          double s = 0.0;  // absent neighbours add +0.0 (see ap_value)
#pragma unroll
          for (int t = 0; t < 2 * DIM; ++t) s += (has[u] >> t & 1u) ? xn[u][t] : 0.0;
          const double nn = (double)__popc(has[u] & ((1u << (2 * DIM)) - 1u));
          v = (double)bv[u] - (nn * xv[u] - s);
        }
        r[id] = (float)v;
        rr += v * v;
        rs += v;
      }
    }
  }
  grid_reduce(R, FIN_TRUE, rr, rs);
}

// FP64 solution -> FP32 output
__global__ void __launch_bounds__(kB) k_to_float(Dims d, int64_t N, const double* __restrict__ x,
                                                 float* __restrict__ y) {
  const SlabWalk<uint32_t> w = slab_walk<kB>((uint32_t)N);
  for (uint32_t id = w.first; id < w.end; id += w.stride) y[id] = (float)x[level0_at(d, id)];
}

// Level-0 pressure in natural cell order (from the colour-split layout)
__global__ void __launch_bounds__(kB) k_unsplit(Dims d, int64_t N, const double* __restrict__ x,
                                                double* __restrict__ y) {
  const SlabWalk<uint32_t> w = slab_walk<kB>((uint32_t)N);
  for (uint32_t id = w.first; id < w.end; id += w.stride) y[id] = x[level0_at(d, id)];
}

// ------------------------------------------------------------ PCG steps ----
// Ap = A p ; partial p.Ap.  MASK: solid rows are 0 (compute_Ap of the
// obstacle solver updates fluid cells only, 2D/vortex/mgpcg_solid.py:172-178).
// A p of one cell from its loaded stencil: formed in FP64 and rounded once.
// For smooth p, nn p - sum(p) cancels all but ~(pi/n)^2 of |p|, and the FP32
// form's error then drifts the recursive residual away from b - A x by up to
// ~1e-2 |b| at 512^2.  k_Ap_dot and k_update_xr both call it on the same
// loads, so the value the update uses is bitwise the one alpha was formed
// with, and A p is never stored.
// The neighbour count is the popcount of the presence bits, and an absent
// neighbour adds +0.0 to the running sum, which (starting from +0.0, the sum
// is never -0.0) leaves it bitwise unchanged: the same value as summing the
// present neighbours only, without a conditional FP64 add per neighbour.
template <int DIM>
DFM_HD float ap_value(unsigned has, const float* pn, float pc) {
  double s = 0.0;
#pragma unroll
  for (int t = 0; t < 2 * DIM; ++t) s += (double)((has >> t & 1u) ? pn[t] : 0.0f);
  const double nn = (double)__popc(has & ((1u << (2 * DIM)) - 1u));
  return (float)(nn * (double)pc - s);
}

template <int DIM, bool MASK, int U, bool SPLIT>
__global__ void __launch_bounds__(kB) k_Ap_dot(Dims d, const float* __restrict__ p, Red R) {
  if (is_done(R.flags)) return;
  double acc = 0;
  int64_t N = (int64_t)d.n[0] * d.n[1] * d.n[2];
  const SlabWalk<uint32_t> w = slab_walk<kB>((uint32_t)N);
  // U cells of the walk per step, all their loads issued before any is
  // used; cells are still consumed in walk order, so U does not change
  // the result
  for (uint32_t id0 = w.first; id0 < w.end; id0 += U * w.stride) {
    float pn[U][2 * DIM], pc[U];
    unsigned has[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t id = id0 + u * w.stride;
      has[u] = 0;
      pc[u] = 0.f;
      if (id < w.end && is_fluid<MASK>(d, id)) {
        int i, j, k;
        cell_of_at<DIM, SPLIT>(d, id, i, j, k);
        has[u] = nbr_load<DIM, MASK, SPLIT>(d, p, i, j, k, id, pn[u]) | kFluidBit;
        pc[u] = p[id];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) pin(pn[u]);
    pin(pc);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t id = id0 + u * w.stride;
      if (id < w.end) {
        const float v = has[u] ? ap_value<DIM>(has[u], pn[u], pc[u]) : 0.f;
        acc += (double)pc[u] * (double)v;
      }
    }
  }
  grid_reduce(R, FIN_ALPHA, acc, 0.0);
}

// x += alpha p (x in FP64) ; r = (r - mean) - alpha A p with A p re-formed
// from p's stencil (ap_value) ; partial r.r and sum r.  Without an obstacle
// mask the stored r still carries the mean the last V-cycle removed on the
// fly (sc[S_MEAN]; 0 right after the set-up's recentring): the V-cycle reads
// r - mean and never writes it back, which saves the write-back in its last
// half-sweep.  (r - mean) in FP32 is the value the write-back stored, so the
// result is bitwise the same.  With a mask r is written back recentred, and
// solid cells hold p = r = 0 throughout.
template <int DIM, bool MASK, int U, bool SPLIT>
__global__ void __launch_bounds__(kB) k_update_xr(Dims d, double* __restrict__ x, float* __restrict__ r,
                                                  const float* __restrict__ p, Red R) {
  if (is_done(R.flags)) return;
  const double alpha_d = R.sc[S_ALPHA];
  const float alpha = (float)alpha_d;
  const float m = MASK ? 0.f : (float)R.sc[S_MEAN];
  // the solve's first update: x = 0 (never written by the set-up), not read
  const bool x_zero = R.flags[F_ITERS] == 0;
  double rr = 0, rs = 0;
  const int64_t N = (int64_t)d.n[0] * d.n[1] * d.n[2];
  const SlabWalk<uint32_t> w = slab_walk<kB>((uint32_t)N);
  // U cells of the walk per step, every load issued before any is used;
  // cells are consumed in walk order (the sums do not depend on U)
  for (uint32_t id0 = w.first; id0 < w.end; id0 += U * w.stride) {
    float pn[U][2 * DIM], pc[U], rv[U];
    double xv[U];
    unsigned has[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t id = id0 + u * w.stride;
      has[u] = 0;
      pc[u] = rv[u] = 0.f;
      xv[u] = 0.0;
      if (id < w.end) {
        if (!x_zero) xv[u] = x[id];
        rv[u] = r[id];
        if (is_fluid<MASK>(d, id)) {
          int i, j, k;
          cell_of_at<DIM, SPLIT>(d, id, i, j, k);
          has[u] = nbr_load<DIM, MASK, SPLIT>(d, p, i, j, k, id, pn[u]) | kFluidBit;
          pc[u] = p[id];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) pin(pn[u]);
    pin(pc);
    pin(rv);
    pin(xv);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t id = id0 + u * w.stride;
      if (id < w.end) {
        const float ap = has[u] ? ap_value<DIM>(has[u], pn[u], pc[u]) : 0.f;
        x[id] = fma(alpha_d, (double)pc[u], xv[u]);
        const float v = fmaf(-alpha, ap, MASK ? rv[u] : rv[u] - m);
        r[id] = v;
        rr += (double)v * (double)v;
        rs += (double)v;
      }
    }
  }
  grid_reduce(R, FIN_RR, rr, rs);
}

// z.r with r - shift and its scalar step `op` (FIN_ZR_INIT or FIN_BETA);
// p = z when init_p; with a mask the recentred r is written back (see
// k_update_xr).  Used when level 0 is not smoothed by the global kernels
// (otherwise the last level-0 half-sweep does this).  Solid cells keep r = 0.
template <bool MASK>
__global__ void __launch_bounds__(kB) k_zr(int64_t N, const float* __restrict__ z, float* __restrict__ r,
                                           const double* shift, const uint8_t* __restrict__ bm, Red R, int op,
                                           int init_p, float* __restrict__ p) {
  if (is_done(R.flags)) return;
  float m = shift ? (float)*shift : 0.f;
  double acc = 0;
  const SlabWalk<uint32_t> w = slab_walk<kB>((uint32_t)N);
  for (uint32_t id = w.first; id < w.end; id += w.stride) {
    const bool fl = !MASK || bm[id] == 0;
    float rv = fl ? r[id] - m : 0.f;
    if (MASK && shift && fl) r[id] = rv;
    float zv = z[id];
    if (init_p) p[id] = zv;
    acc += (double)zv * (double)rv;
  }
  grid_reduce(R, op, acc, 0.0);
}

// p = z + beta p
__global__ void __launch_bounds__(kB) k_update_p(int64_t N, const float* __restrict__ z, float* __restrict__ p,
                                                 const double* sc, const int* flags) {
  if (is_done(flags)) return;
  float beta = (float)sc[S_BETA];
  const SlabWalk<uint32_t> w = slab_walk<kB>((uint32_t)N);
  for (uint32_t id = w.first; id < w.end; id += w.stride)
    p[id] = fmaf(beta, p[id], z[id]);
}

// -------------------------------------------------------------- V-cycle ----
// Storage of a level's r and z: FP32, or BF16 (uint16_t, round to nearest
// even) with FP32 arithmetic -- the mixed-precision mode stores the levels
// >= 2 that way, in global memory and in the LDS tail.
DFM_HD float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
DFM_HD uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (uint16_t)(u >> 16);
}
template <typename T> struct Sto;
template <> struct Sto<float> {
  DFM_HD static float ld(const float* p, uint32_t i) { return p[i]; }
  DFM_HD static void st(float* p, uint32_t i, float v) { p[i] = v; }
};
template <> struct Sto<uint16_t> {
  DFM_HD static float ld(const uint16_t* p, uint32_t i) { return bf16_to_f32(p[i]); }
  DFM_HD static void st(uint16_t* p, uint32_t i, float v) { p[i] = f32_to_bf16(v); }
};
using bf16_t = uint16_t;

// First RBGS sweep from z = 0 with an obstacle mask: colour 0 cells z = r/nn;
// colour 1 cells use the colour-0 neighbours' r/nn.  `shift` (level 0)
// recentres r on the fly.  Solid cells get z = 0 and are not neighbours
// (smooth of the obstacle solver, 2D/vortex/mgpcg_solid.py:229-236).  Without
// a mask the same sweep runs as k_first_half plus an ordinary half-sweep.
template <int DIM, bool MASK>
__global__ void __launch_bounds__(kB) k_first_sweep(Dims d, const float* __restrict__ r, float* __restrict__ z,
                                                    const double* shift, const int* flags) {
#pragma clang fp contract(off)
  if (is_done(flags)) return;
  float m = shift ? (float)*shift : 0.f;
  int64_t N = (int64_t)d.n[0] * d.n[1] * d.n[2];
  uint32_t id = (uint32_t)xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
  if (id >= N) return;
  if (!is_fluid<MASK>(d, id)) {
    z[id] = 0.f;
    return;
  }
  int i, j, k;
  cell_of(d, id, i, j, k);
  float nn = nbr_count<DIM, MASK>(d, i, j, k);
  if (((i + j + k) & 1) == 0) {
    z[id] = (r[id] - m) * inv_count(nn);
    return;
  }
  // every neighbour value is loaded before the first is used
  float rn[2 * DIM];
  const unsigned has = nbr_load<DIM, MASK, false>(d, r, i, j, k, id, rn);
  pin(rn);
  float rv = r[id] - m;
  const int c3[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
  float s = 0.f;
#pragma unroll
  for (int q = 0; q < 2 * DIM; ++q)
    if (has >> q & 1u) s += (rn[q] - m) * inv_count(nbr_count<DIM, MASK>(d, c3[q][0], c3[q][1], c3[q][2]));
  z[id] = (rv + s) * inv_count(nn);
}

// Gauss-Seidel update of cell (i, j, k): (r - m + sum of neighbour z) / nn.
// With zc, the neighbours' values are read as z + zc[coarse parent] (fused
// prolongation; the cell's own value is overwritten, so it never needs it).
template <int DIM, bool MASK, typename ZCT>
DFM_HD float gs_value(const Dims& d, const float* __restrict__ r, const float* __restrict__ z,
                      const ZCT* __restrict__ zc, const Dims& dc, float m, int i, int j, int k, uint32_t id) {
  float nn, s;
  if (zc) {
    s = 0.f;
    nn = 0.f;
    for_nbrs<DIM, MASK>(d, i, j, k, id, [&](uint32_t q, int a, int b, int c) {
      s += z[q] + Sto<ZCT>::ld(zc, cidx(dc, a >> 1, b >> 1, DIM == 3 ? (c >> 1) : 0));
      nn += 1.f;
    });
  } else {
    s = nbr_sum<DIM, MASK>(z, d, i, j, k, nn);
  }
  return (r[id] - m + s) * inv_count(nn);
}

// Pairs per thread of the global half-sweep kernels below, and pairs per
// step of the z.r half-sweep's walk.
constexpr int kSweepU = 2;  // 1 and 4 measured slower at 128^3 and 256^3
constexpr int kZrU = 2;     // 1, 2 and 4 measured alike

// First half-sweep from z = 0 without an obstacle mask: the cells of colour
// `phase` get (r - m) / nn (kSweepU pairs per thread, as k_smooth_u).  The
// other colour then takes an ordinary half-sweep: its neighbour values are
// exactly the (r - m) / nn that k_first_sweep forms in place, so the pair
// gives k_first_sweep's result bitwise while each launch stays
// bandwidth-bound (k_first_sweep re-forms six neighbour counts per cell).
template <int DIM, bool SPLIT>
__global__ void __launch_bounds__(kB) k_first_half(Dims d, const float* __restrict__ r, float* __restrict__ z,
                                                   int phase, const double* shift, const int* flags) {
#pragma clang fp contract(off)
  if (is_done(flags)) return;
  const float m = shift ? (float)*shift : 0.f;
  const int inner = DIM == 3 ? d.n[2] : d.n[1];
  const uint32_t pairs = (uint32_t)(DIM == 3 ? d.n[0] * d.n[1] : d.n[0]) * d.fh.d;
  const uint32_t t0 = (uint32_t)xcd_block(blockIdx.x, gridDim.x) * (kB * kSweepU) + threadIdx.x;
  float rv[kSweepU], nn[kSweepU];
  uint32_t ids[kSweepU];
  bool has[kSweepU];
#pragma unroll
  for (int u = 0; u < kSweepU; ++u) {
    const uint32_t t = t0 + u * kB;
    has[u] = false;
    rv[u] = nn[u] = 0.f;
    ids[u] = 0;
    if (t < pairs) {
      int i, j, own, other;
      colour_pair<DIM>(d, t, phase, i, j, own, other);
      if (own < inner) {
        const int a = DIM == 3 ? j : own, k = DIM == 3 ? own : 0;
        const uint32_t id = SPLIT ? (phase ? d.hh : 0u) + t : cidx(d, i, a, k);
        ids[u] = id;
        has[u] = true;
        rv[u] = r[id];
        nn[u] = nbr_count<DIM>(d, i, a, k);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < kSweepU; ++u)
    if (has[u]) z[ids[u]] = (rv[u] - m) * inv_count(nn[u]);
}

// One half-sweep of colour `phase` without an obstacle mask or a z.r
// reduction, kSweepU pairs per thread (the pairs kB apart, so each step of
// a wave stays coalesced): every load of a thread's pairs is issued before
// the first store, which keeps several cache lines per lane in flight
// (latency, not bandwidth, bounds a one-pair-per-thread half-sweep).  The
// loads read the other colour's z, the coarse correction and r, none of
// which the stores of this colour change.  Each cell is formed by gs_value's
// operations in its order, so the result is bitwise the per-pair kernel's.

template <int DIM, bool SPLIT, typename ZCT = float>
__global__ void __launch_bounds__(kB) k_smooth_u(Dims d, const float* __restrict__ r, float* __restrict__ z,
                                                 int phase, const double* shift, const ZCT* __restrict__ zc,
                                                 Dims dc, const int* flags) {
  if (is_done(flags)) return;
  const float m = shift ? (float)*shift : 0.f;
  const int inner = DIM == 3 ? d.n[2] : d.n[1];
  const uint32_t pairs = (uint32_t)(DIM == 3 ? d.n[0] * d.n[1] : d.n[0]) * d.fh.d;
  const uint32_t t0 = (uint32_t)xcd_block(blockIdx.x, gridDim.x) * (kB * kSweepU) + threadIdx.x;
  float zn[kSweepU][2 * DIM], zp[kSweepU][2 * DIM], rv[kSweepU];
  unsigned has[kSweepU];
  uint32_t ids[kSweepU];
#pragma unroll
  for (int u = 0; u < kSweepU; ++u) {
    const uint32_t t = t0 + u * kB;
    has[u] = 0;
    rv[u] = 0.f;
    ids[u] = 0;
    if (t < pairs) {
      int i, j, own, other;
      colour_pair<DIM>(d, t, phase, i, j, own, other);
      if (own < inner) {
        const int a = DIM == 3 ? j : own, k = DIM == 3 ? own : 0;
        const uint32_t id = SPLIT ? (phase ? d.hh : 0u) + t : cidx(d, i, a, k);
        ids[u] = id;
        const unsigned nb = nbr_load<DIM, false, SPLIT>(d, (const float*)z, i, a, k, id, zn[u]);
        has[u] = nb | kFluidBit;
        rv[u] = r[id];
        if (zc) {  // the coarse parent of each neighbour (an absent one: of the cell)
          const int c3[6][3] = {{i - 1, a, k}, {i + 1, a, k}, {i, a - 1, k}, {i, a + 1, k}, {i, a, k - 1}, {i, a, k + 1}};
#pragma unroll
          for (int q = 0; q < 2 * DIM; ++q) {
            const bool pr = nb >> q & 1u;
            zp[u][q] = Sto<ZCT>::ld(zc, cidx(dc, (pr ? c3[q][0] : i) >> 1, (pr ? c3[q][1] : a) >> 1,
                                         DIM == 3 ? ((pr ? c3[q][2] : k) >> 1) : 0));
          }
        }
      }
    }
  }
#pragma unroll
  for (int u = 0; u < kSweepU; ++u) pin(zn[u]);
  pin(rv);
  if (zc) {
#pragma unroll
    for (int u = 0; u < kSweepU; ++u) pin(zp[u]);
  }
#pragma unroll
  for (int u = 0; u < kSweepU; ++u) {
    if (has[u]) {
      float sum = 0.f, nn = 0.f;
#pragma unroll
      for (int q = 0; q < 2 * DIM; ++q)
        if (has[u] >> q & 1u) {
          sum += zc ? zn[u][q] + zp[u][q] : zn[u][q];
          nn += 1.f;
        }
      z[ids[u]] = (rv[u] - m + sum) * inv_count(nn);
    }
  }
}

// One half-sweep of colour `phase`; threads cover only that colour (the
// innermost axis is walked in steps of two).  ZR: this is the last level-0
// half-sweep of the V-cycle, so it also writes the recentred r back for both
// cells of each pair, forms z.r over them (the other colour's z is final) and
// runs the scalar step `op` on the total (p = z as well when p is given).
// MASK: solid cells are skipped (z and r stay 0 there).
template <int DIM, bool ZR, bool MASK, bool SPLIT>
__global__ void __launch_bounds__(kB) k_smooth(Dims d, float* __restrict__ r, float* __restrict__ z, int phase,
                                               const double* shift, const float* __restrict__ zc, Dims dc, Red R,
                                               int op, float* __restrict__ p) {
  if (is_done(R.flags)) return;
  const float m = shift ? (float)*shift : 0.f;
  const int inner = DIM == 3 ? d.n[2] : d.n[1];
  const uint32_t pairs = (uint32_t)(DIM == 3 ? d.n[0] * d.n[1] : d.n[0]) * d.fh.d;
  double acc = 0;
  const SlabWalk<uint32_t> w = slab_walk<kB>(pairs);
  // U pairs of the walk per step, every load issued before the first store
  // (the loads read the other colour's z, the coarse correction and r, none
  // of which a store of this launch changes); pairs are consumed in walk
  // order with gs_value's operations in its order
  constexpr int U = ZR ? kZrU : 1;
  constexpr unsigned kOwnBit = 1u << 30;  // the own cell lies in the row
  for (uint32_t t0 = w.first; t0 < w.end; t0 += U * w.stride) {
    float zn[U][2 * DIM], zp[U][2 * DIM], rown[U], zoth[U], roth[U];
    uint32_t ido[U], idt[U];
    unsigned has[U];  // kFluidBit: own cell fluid; kOwnBit; bits 0..: neighbours present
    int ostate[U];    // other cell: 0 none (or not ZR), 1 solid, 2 fluid
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t t = t0 + u * w.stride;
      has[u] = 0;
      ostate[u] = 0;
      rown[u] = zoth[u] = roth[u] = 0.f;
      ido[u] = idt[u] = 0;
      if (t >= w.end) continue;
      int i, j, own, other;
      colour_pair<DIM>(d, t, phase, i, j, own, other);
      if (own < inner) {
        const int a = DIM == 3 ? j : own, k = DIM == 3 ? own : 0;
        // colour-split: pair t is half index t of both colours
        const uint32_t id = SPLIT ? (phase ? d.hh : 0u) + t : cidx(d, i, a, k);
        ido[u] = id;
        has[u] = kOwnBit;
        if (is_fluid<MASK>(d, id)) {
          const unsigned nb = nbr_load<DIM, MASK, SPLIT>(d, (const float*)z, i, a, k, id, zn[u]);
          has[u] |= nb | kFluidBit;
          rown[u] = r[id];
          if (zc) {  // the coarse parent of each neighbour (an absent one: of the cell)
            const int c3[6][3] = {{i - 1, a, k}, {i + 1, a, k}, {i, a - 1, k},
                                  {i, a + 1, k}, {i, a, k - 1}, {i, a, k + 1}};
#pragma unroll
            for (int q = 0; q < 2 * DIM; ++q) {
              const bool pr = nb >> q & 1u;
              zp[u][q] = zc[cidx(dc, (pr ? c3[q][0] : i) >> 1, (pr ? c3[q][1] : a) >> 1,
                                 DIM == 3 ? ((pr ? c3[q][2] : k) >> 1) : 0)];
            }
          }
        }
      }
      if (ZR && other < inner) {
        const uint32_t id = SPLIT ? (phase ? 0u : d.hh) + t : DIM == 3 ? cidx(d, i, j, other) : cidx(d, i, other, 0);
        idt[u] = id;
        zoth[u] = z[id];
        ostate[u] = 1;
        if (is_fluid<MASK>(d, id)) {
          ostate[u] = 2;
          roth[u] = r[id];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) pin(zn[u]);
    pin(rown);
    pin(zoth);
    pin(roth);
    if (zc) {
#pragma unroll
      for (int u = 0; u < U; ++u) pin(zp[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (has[u] & kFluidBit) {
        float sum = 0.f, nn = 0.f;
#pragma unroll
        for (int q = 0; q < 2 * DIM; ++q)
          if (has[u] >> q & 1u) {
            sum += zc ? zn[u][q] + zp[u][q] : zn[u][q];
            nn += 1.f;
          }
        const float znew = (rown[u] - m + sum) * inv_count(nn);
        z[ido[u]] = znew;
        if (ZR) {
          const float rv = rown[u] - m;
          if (MASK && shift) r[ido[u]] = rv;
          if (p) p[ido[u]] = znew;
          acc += (double)znew * (double)rv;
        }
      } else if (ZR && (has[u] & kOwnBit) && p) {
        p[ido[u]] = 0.f;
      }
      if (ZR && ostate[u]) {
        if (ostate[u] == 2) {
          const float rv = roth[u] - m;
          if (MASK && shift) r[idt[u]] = rv;
          acc += (double)zoth[u] * (double)rv;
        }
        if (p) p[idt[u]] = zoth[u];
      }
    }
  }
  if (ZR) grid_reduce(R, op, acc, 0.0);
}

// ------------------------------------------------ level-0 row kernels ----
// The colour-split finest level in 3D with an innermost extent that is a
// multiple of 8.  The per-pair kernels above spend ~135 VALU instructions
// per cell on index arithmetic, bounds tests and 64-bit addresses, which
// makes them VALU-bound at 256^3.  Here a thread owns half indices h0 ..
// h0 + 3 of one row (i, j) of colour `phase`, so the arithmetic is done once
// per 4 cells and every row is read with 16-byte loads.  The thread's cells
// sit at natural k = 2h + off, off = (phase - i - j) & 1.  Their x and y
// neighbours are the other colour's cells of the same half indices in rows
// (i +- 1, j) and (i, j +- 1); their -z and +z neighbours are the other
// colour's half indices h - 1 + off and h + off of the own row, i.e. the
// centre row's 4 values plus one beyond the segment (h0 - 1 when off = 0,
// h0 + 4 when off = 1).  Coarse parents follow the same pattern: the parent
// of natural k is k >> 1 = h for the x and y neighbours, h - 1 + off and
// h + off for the z neighbours.  Each cell is formed with the per-pair
// kernels' operations in their order, so the smoothed values are bitwise
// theirs; the z.r sums add the same terms in another order.
constexpr int kRowV = 4;

struct F4 {
  float v[kRowV];
};
DFM_HD F4 ld4(const float* __restrict__ p, uint32_t i) {
  const float4 q = *reinterpret_cast<const float4*>(p + i);
  return F4{{q.x, q.y, q.z, q.w}};
}
DFM_HD F4 ld4(const uint16_t* __restrict__ p, uint32_t i) {
  const uint2 q = *reinterpret_cast<const uint2*>(p + i);
  return F4{{__uint_as_float(q.x << 16), __uint_as_float(q.x & 0xffff0000u), __uint_as_float(q.y << 16),
             __uint_as_float(q.y & 0xffff0000u)}};
}
DFM_HD void st4(float* __restrict__ p, uint32_t i, const F4& a) {
  *reinterpret_cast<float4*>(p + i) = make_float4(a.v[0], a.v[1], a.v[2], a.v[3]);
}

// A thread's row segment.
struct RowSeg {
  int i, j, h0, off;
  uint32_t rb;    // half index of the row's first pair
  unsigned nbr;   // bits 0..3: the -x, +x, -y, +y neighbour rows exist
  bool zm0, zp3;  // cell 0 has a -z neighbour, cell 3 a +z neighbour
};

DFM_HD RowSeg row_seg(const Dims& d, uint32_t t, int phase) {
  RowSeg f;
  const uint32_t half = d.fh.d;
  const uint32_t p0 = t * kRowV;
  const uint32_t row = fdiv(p0, d.fh);
  const uint32_t q = fdiv(row, d.f1);
  f.i = (int)q;
  f.j = (int)(row - q * (uint32_t)d.n[1]);
  f.h0 = (int)(p0 - row * half);
  f.rb = row * half;
  f.off = (phase - f.i - f.j) & 1;
  f.nbr = (f.i > 0 ? 1u : 0u) | (f.i < d.n[0] - 1 ? 2u : 0u) | (f.j > 0 ? 4u : 0u) | (f.j < d.n[1] - 1 ? 8u : 0u);
  f.zm0 = f.off == 1 || f.h0 > 0;
  f.zp3 = f.off == 0 || f.h0 + kRowV < (int)half;
  return f;
}

// Threads of a row launch: one per kRowV pairs.
__host__ __device__ inline uint32_t row_threads(const Dims& d) { return (uint32_t)d.n[0] * (uint32_t)d.n[1] * (d.fh.d / kRowV); }

// The 6 neighbour rows of a segment in array f (x, y rows of the same half
// indices; the own row's 4 values and the one beyond), with c the segment's
// first index in the other colour's half and (sx, sy) the row strides.  An
// absent row reads the centre row (its values are not used).
template <typename T>
struct Nbr6 {
  F4 xm, xp, ym, yp, c;
  float e;
};
template <typename T>
DFM_HD Nbr6<T> nbr6(const T* __restrict__ f, const RowSeg& g, uint32_t c, uint32_t sx, uint32_t sy) {
  Nbr6<T> n;
  n.xm = ld4(f, g.nbr & 1u ? c - sx : c);
  n.xp = ld4(f, g.nbr & 2u ? c + sx : c);
  n.ym = ld4(f, g.nbr & 4u ? c - sy : c);
  n.yp = ld4(f, g.nbr & 8u ? c + sy : c);
  n.c = ld4(f, c);
  const bool ext = g.off ? g.zp3 : g.zm0;
  n.e = Sto<T>::ld(f, ext ? (g.off ? c + kRowV : c - 1u) : c);
  return n;
}
// the -z / +z neighbour value of cell v
template <typename T>
DFM_HD float zm_of(const Nbr6<T>& n, int off, int v) {
  return off ? n.c.v[v] : (v == 0 ? n.e : n.c.v[v - 1]);
}
template <typename T>
DFM_HD float zp_of(const Nbr6<T>& n, int off, int v) {
  return off ? (v == kRowV - 1 ? n.e : n.c.v[v + 1]) : n.c.v[v];
}

// Gauss-Seidel values of the segment's 4 cells (gs_value's operations: the
// neighbours summed -x, +x, -y, +y, -z, +z, each as z + parent correction
// with ZC, then (r - m + sum) / nn).
template <bool ZC, typename ZCT>
DFM_HD F4 row_gs(const Dims& d, const RowSeg& g, const F4& rv, float m, const Nbr6<float>& z, const Nbr6<ZCT>& pc) {
  F4 out;
#pragma unroll
  for (int v = 0; v < kRowV; ++v) {
    float sum = 0.f, nn = 0.f;
    if (g.nbr & 1u) { sum += ZC ? z.xm.v[v] + pc.xm.v[v] : z.xm.v[v]; nn += 1.f; }
    if (g.nbr & 2u) { sum += ZC ? z.xp.v[v] + pc.xp.v[v] : z.xp.v[v]; nn += 1.f; }
    if (g.nbr & 4u) { sum += ZC ? z.ym.v[v] + pc.ym.v[v] : z.ym.v[v]; nn += 1.f; }
    if (g.nbr & 8u) { sum += ZC ? z.yp.v[v] + pc.yp.v[v] : z.yp.v[v]; nn += 1.f; }
    if (v > 0 || g.zm0) {
      const float a = zm_of(z, g.off, v);
      sum += ZC ? a + zm_of(pc, g.off, v) : a;
      nn += 1.f;
    }
    if (v < kRowV - 1 || g.zp3) {
      const float a = zp_of(z, g.off, v);
      sum += ZC ? a + zp_of(pc, g.off, v) : a;
      nn += 1.f;
    }
    out.v[v] = (rv.v[v] - m + sum) * inv_count(nn);
  }
  return out;
}

// The coarse parents of a segment's neighbours: the parent rows of the x and
// y neighbour rows and of the own row, in the coarse level's natural order
// (whose innermost extent is the fine level's pairs per row).
template <typename ZCT>
DFM_HD Nbr6<ZCT> parents6(const ZCT* __restrict__ zc, const Dims& dc, const RowSeg& g) {
  auto row = [&](int ci, int cj) { return ((uint32_t)ci * (uint32_t)dc.n[1] + (uint32_t)cj) * (uint32_t)dc.n[2]; };
  const uint32_t h = (uint32_t)g.h0, own = row(g.i >> 1, g.j >> 1) + h;
  Nbr6<ZCT> n;
  n.xm = ld4(zc, g.nbr & 1u ? row((g.i - 1) >> 1, g.j >> 1) + h : own);
  n.xp = ld4(zc, g.nbr & 2u ? row((g.i + 1) >> 1, g.j >> 1) + h : own);
  n.ym = ld4(zc, g.nbr & 4u ? row(g.i >> 1, (g.j - 1) >> 1) + h : own);
  n.yp = ld4(zc, g.nbr & 8u ? row(g.i >> 1, (g.j + 1) >> 1) + h : own);
  n.c = ld4(zc, own);
  const bool ext = g.off ? g.zp3 : g.zm0;
  n.e = Sto<ZCT>::ld(zc, ext ? (g.off ? own + kRowV : own - 1u) : own);
  return n;
}

// One level-0 half-sweep of colour `phase` (MODE 1; MODE 2 with the coarse
// correction zc added to the neighbour values, the fused prolongation), or
// the first half-sweep from z = 0 (MODE 0: (r - m) / nn, k_first_half).
template <int MODE, typename ZCT = float>
__global__ void __launch_bounds__(kB) k_row_half(Dims d, const float* __restrict__ r, float* __restrict__ z,
                                                 int phase, const double* shift, const ZCT* __restrict__ zc,
                                                 Dims dc, const int* flags) {
#pragma clang fp contract(off)
  if (is_done(flags)) return;
  const uint32_t t = (uint32_t)xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
  if (t >= row_threads(d)) return;
  const float m = shift ? (float)*shift : 0.f;
  const RowSeg g = row_seg(d, t, phase);
  const uint32_t own = (phase ? d.hh : 0u) + g.rb + (uint32_t)g.h0;
  const F4 rv = ld4(r, own);
  F4 out;
  if constexpr (MODE == 0) {
#pragma unroll
    for (int v = 0; v < kRowV; ++v) {
      const int k = 2 * (g.h0 + v) + g.off;
      float nn = 0.f;
      nn += (g.i > 0) + (g.i < d.n[0] - 1) + (g.j > 0) + (g.j < d.n[1] - 1);
      nn += (k > 0) + (k < d.n[2] - 1);
      out.v[v] = (rv.v[v] - m) * inv_count(nn);
    }
  } else {
    const uint32_t c = (phase ? 0u : d.hh) + g.rb + (uint32_t)g.h0;
    const uint32_t sy = d.fh.d, sx = (uint32_t)d.n[1] * sy;
    const Nbr6<float> zn = nbr6(z, g, c, sx, sy);
    if constexpr (MODE == 2) {
      const Nbr6<ZCT> pc = parents6(zc, dc, g);
      out = row_gs<true>(d, g, rv, m, zn, pc);
    } else {
      out = row_gs<false, float>(d, g, rv, m, zn, zn);
    }
  }
  st4(z, own, out);
}

// The 4 cells of a segment's own colour (off = 0: cells 0, 2, 4, 6; off = 1:
// cells 1, 3, 5, 7) and of the other colour, from its two 16-byte halves.
// The cells are picked from named vector components: selects between
// elements of a register array can be folded into a run-time index into
// that array, which then lives in scratch memory.
struct Seg8 {
  float4 lo, hi;
};
DFM_HD Seg8 ld8(const float* __restrict__ p, uint32_t i) {
  return Seg8{*reinterpret_cast<const float4*>(p + i), *reinterpret_cast<const float4*>(p + i + kRowV)};
}
DFM_HD F4 pick(const Seg8& s, bool odd) {
  return F4{{odd ? s.lo.y : s.lo.x, odd ? s.lo.w : s.lo.z, odd ? s.hi.y : s.hi.x, odd ? s.hi.w : s.hi.z}};
}

// The same half-sweeps on a natural-order 3D level (the coarse global levels,
// or level 0 when it is not colour-split): a thread owns the 8 cells of pairs
// h0 .. h0 + 3 of row (i, j) and updates its 4 cells of colour `phase` (at
// 2 (h0 + v) + off).  Each neighbour row is two 16-byte loads; the segment
// is stored back whole, its other-colour cells with the values just read
// (no other thread writes them in a half-sweep).  Bitwise k_smooth_u /
// k_first_half.
// Gauss-Seidel update, in place in z, of the own-colour cells of segment g
// of a natural-order 3D level (with ZC the coarse correction zc added to the
// neighbour values, the fused prolongation).
template <bool ZC, typename ZCT>
DFM_HD void nat_seg_update(const Dims& d, const RowSeg& g, const float* __restrict__ r, float* __restrict__ z,
                           float m, const ZCT* __restrict__ zc, const Dims& dc) {
#pragma clang fp contract(off)
  const uint32_t sy = (uint32_t)d.n[2], sx = (uint32_t)d.n[1] * sy;
  const uint32_t base = 2u * (g.rb + (uint32_t)g.h0);  // natural index of the segment's first cell
  const bool odd = g.off != 0;
  const F4 rv = pick(ld8(r, base), odd);
  const Seg8 c8 = ld8(z, base);
  Nbr6<float> zn;
  zn.xm = pick(ld8(z, g.nbr & 1u ? base - sx : base), odd);
  zn.xp = pick(ld8(z, g.nbr & 2u ? base + sx : base), odd);
  zn.ym = pick(ld8(z, g.nbr & 4u ? base - sy : base), odd);
  zn.yp = pick(ld8(z, g.nbr & 8u ? base + sy : base), odd);
  zn.c = pick(c8, !odd);  // the other colour's cell of each pair
  const bool ext = odd ? g.zp3 : g.zm0;
  zn.e = z[ext ? (odd ? base + 2 * kRowV : base - 1u) : base];
  F4 out;
  if constexpr (ZC) out = row_gs<true>(d, g, rv, m, zn, parents6(zc, dc, g));
  else out = row_gs<false, float>(d, g, rv, m, zn, zn);
  const float4 lo = odd ? make_float4(c8.lo.x, out.v[0], c8.lo.z, out.v[1])
                        : make_float4(out.v[0], c8.lo.y, out.v[1], c8.lo.w);
  const float4 hi = odd ? make_float4(c8.hi.x, out.v[2], c8.hi.z, out.v[3])
                        : make_float4(out.v[2], c8.hi.y, out.v[3], c8.hi.w);
  *reinterpret_cast<float4*>(z + base) = lo;
  *reinterpret_cast<float4*>(z + base + kRowV) = hi;
}

template <int MODE, typename ZCT = float>
__global__ void __launch_bounds__(kB) k_row_half_nat(Dims d, const float* __restrict__ r, float* __restrict__ z,
                                                     int phase, const double* shift, const ZCT* __restrict__ zc,
                                                     Dims dc, const int* flags) {
#pragma clang fp contract(off)
  if (is_done(flags)) return;
  const uint32_t t = (uint32_t)xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
  if (t >= row_threads(d)) return;
  const float m = shift ? (float)*shift : 0.f;
  const RowSeg g = row_seg(d, t, phase);
  if constexpr (MODE == 0) {
    const uint32_t base = 2u * (g.rb + (uint32_t)g.h0);
    const F4 rv = pick(ld8(r, base), g.off != 0);
#pragma unroll
    for (int v = 0; v < kRowV; ++v) {
      const int k = 2 * (g.h0 + v) + g.off;
      float nn = 0.f;
      nn += (g.i > 0) + (g.i < d.n[0] - 1) + (g.j > 0) + (g.j < d.n[1] - 1);
      nn += (k > 0) + (k < d.n[2] - 1);
      z[base + (uint32_t)(2 * v + g.off)] = (rv.v[v] - m) * inv_count(nn);
    }
  } else {
    nat_seg_update<MODE == 2, ZCT>(d, g, r, z, m, zc, dc);
  }
}

// The last level-0 half-sweep of a V-cycle fused with z.r (k_smooth<3, true,
// false, true>): the segment's cells are smoothed, and z.r is summed over
// them and the other colour's cells of the same pairs (whose z is final),
// own cell first in each pair; p = z for both when p is given.  Then the
// scalar step `op` runs on the total.
__global__ void __launch_bounds__(kB) k_row_zr(Dims d, const float* __restrict__ r, float* __restrict__ z, int phase,
                                               const double* shift, Red R, int op, float* __restrict__ p) {
  if (is_done(R.flags)) return;
  const float m = shift ? (float)*shift : 0.f;
  double acc = 0;
  const SlabWalk<uint32_t> w = slab_walk<kB>(row_threads(d));
  const uint32_t sy = d.fh.d, sx = (uint32_t)d.n[1] * sy;
  for (uint32_t t = w.first; t < w.end; t += w.stride) {
    const RowSeg g = row_seg(d, t, phase);
    const uint32_t own = (phase ? d.hh : 0u) + g.rb + (uint32_t)g.h0;
    const uint32_t c = (phase ? 0u : d.hh) + g.rb + (uint32_t)g.h0;
    const F4 rv = ld4(r, own);
    const F4 ro = ld4(r, c);
    const Nbr6<float> zn = nbr6(z, g, c, sx, sy);
    const F4 zo = row_gs<false, float>(d, g, rv, m, zn, zn);
    st4(z, own, zo);
    if (p) {
      st4(p, own, zo);
      st4(p, c, zn.c);
    }
#pragma unroll
    for (int v = 0; v < kRowV; ++v) {
      const float rvo = rv.v[v] - m;
      acc += (double)zo.v[v] * (double)rvo;
      const float rvt = ro.v[v] - m;
      acc += (double)zn.c.v[v] * (double)rvt;
    }
  }
  grid_reduce(R, op, acc, 0.0);
}

// The same segment of the other colour (cells at natural k = 2h + 1 - off).
DFM_HD RowSeg other_colour(const Dims& d, const RowSeg& g) {
  RowSeg f = g;
  f.off = 1 - g.off;
  f.zm0 = f.off == 1 || f.h0 > 0;
  f.zp3 = f.off == 0 || f.h0 + kRowV < (int)d.fh.d;
  return f;
}

// A p of the segment's 4 cells of one colour (ap_value: the neighbour values
// summed in FP64, -x, +x, -y, +y, -z, +z, an absent one as 0; nn pc - sum).
DFM_HD F4 row_Ap(const RowSeg& g, const F4& pc, const Nbr6<float>& n) {
  F4 ap;
#pragma unroll
  for (int v = 0; v < kRowV; ++v) {
    const bool zm = v > 0 || g.zm0, zp = v < kRowV - 1 || g.zp3;
    double sum = 0.0;
    sum += (double)((g.nbr & 1u) ? n.xm.v[v] : 0.0f);
    sum += (double)((g.nbr & 2u) ? n.xp.v[v] : 0.0f);
    sum += (double)((g.nbr & 4u) ? n.ym.v[v] : 0.0f);
    sum += (double)((g.nbr & 8u) ? n.yp.v[v] : 0.0f);
    sum += (double)(zm ? zm_of(n, g.off, v) : 0.0f);
    sum += (double)(zp ? zp_of(n, g.off, v) : 0.0f);
    const double nn = (double)(__popc(g.nbr) + (zm ? 1 : 0) + (zp ? 1 : 0));
    ap.v[v] = (float)(nn * (double)pc.v[v] - sum);
  }
  return ap;
}

// p.Ap over the colour-split finest level (k_Ap_dot): a thread walks row
// segments, both colours of each (8 cells).
__global__ void __launch_bounds__(kB) k_row_Ap_dot(Dims d, const float* __restrict__ p, Red R) {
  if (is_done(R.flags)) return;
  double acc = 0;
  const SlabWalk<uint32_t> w = slab_walk<kB>(row_threads(d));
  const uint32_t sy = d.fh.d, sx = (uint32_t)d.n[1] * sy;
  for (uint32_t t = w.first; t < w.end; t += w.stride) {
    const RowSeg g0 = row_seg(d, t, 0), g1 = other_colour(d, g0);
    const uint32_t c0 = g0.rb + (uint32_t)g0.h0, c1 = d.hh + c0;
    const Nbr6<float> n0 = nbr6(p, g0, c1, sx, sy);  // colour-0 cells' neighbours (colour 1)
    const Nbr6<float> n1 = nbr6(p, g1, c0, sx, sy);
    const F4 a0 = row_Ap(g0, n1.c, n0), a1 = row_Ap(g1, n0.c, n1);
#pragma unroll
    for (int v = 0; v < kRowV; ++v) acc += (double)n1.c.v[v] * (double)a0.v[v];
#pragma unroll
    for (int v = 0; v < kRowV; ++v) acc += (double)n0.c.v[v] * (double)a1.v[v];
  }
  grid_reduce(R, FIN_ALPHA, acc, 0.0);
}

// x += alpha p, r = (r - mean) - alpha A p over the colour-split finest
// level (k_update_xr's operations per cell), both colours of each segment.
__global__ void __launch_bounds__(kB) k_row_update_xr(Dims d, double* __restrict__ x, float* __restrict__ r,
                                                      const float* __restrict__ p, Red R) {
  if (is_done(R.flags)) return;
  const double alpha_d = R.sc[S_ALPHA];
  const float alpha = (float)alpha_d;
  const float m = (float)R.sc[S_MEAN];
  const bool x_zero = R.flags[F_ITERS] == 0;
  double rr = 0, rs = 0;
  const SlabWalk<uint32_t> w = slab_walk<kB>(row_threads(d));
  const uint32_t sy = d.fh.d, sx = (uint32_t)d.n[1] * sy;
  for (uint32_t t = w.first; t < w.end; t += w.stride) {
    const RowSeg g0 = row_seg(d, t, 0), g1 = other_colour(d, g0);
    const uint32_t c0 = g0.rb + (uint32_t)g0.h0, c1 = d.hh + c0;
    const Nbr6<float> n0 = nbr6(p, g0, c1, sx, sy);
    const Nbr6<float> n1 = nbr6(p, g1, c0, sx, sy);
    const F4 r0 = ld4(r, c0), r1 = ld4(r, c1);
    double x0[kRowV], x1[kRowV];
#pragma unroll
    for (int v = 0; v < kRowV; v += 2) {
      double2 a = make_double2(0.0, 0.0), b = make_double2(0.0, 0.0);
      if (!x_zero) {
        a = *reinterpret_cast<const double2*>(x + c0 + v);
        b = *reinterpret_cast<const double2*>(x + c1 + v);
      }
      x0[v] = a.x;
      x0[v + 1] = a.y;
      x1[v] = b.x;
      x1[v + 1] = b.y;
    }
    const F4 a0 = row_Ap(g0, n1.c, n0), a1 = row_Ap(g1, n0.c, n1);
    F4 o0, o1;
#pragma unroll
    for (int v = 0; v < kRowV; ++v) {
      x0[v] = fma(alpha_d, (double)n1.c.v[v], x0[v]);
      x1[v] = fma(alpha_d, (double)n0.c.v[v], x1[v]);
      o0.v[v] = fmaf(-alpha, a0.v[v], r0.v[v] - m);
      o1.v[v] = fmaf(-alpha, a1.v[v], r1.v[v] - m);
    }
#pragma unroll
    for (int v = 0; v < kRowV; v += 2) {
      *reinterpret_cast<double2*>(x + c0 + v) = make_double2(x0[v], x0[v + 1]);
      *reinterpret_cast<double2*>(x + c1 + v) = make_double2(x1[v], x1[v + 1]);
    }
    st4(r, c0, o0);
    st4(r, c1, o1);
#pragma unroll
    for (int v = 0; v < kRowV; ++v) {
      rr += (double)o0.v[v] * (double)o0.v[v];
      rs += (double)o0.v[v];
    }
#pragma unroll
    for (int v = 0; v < kRowV; ++v) {
      rr += (double)o1.v[v] * (double)o1.v[v];
      rs += (double)o1.v[v];
    }
  }
  grid_reduce(R, FIN_RR, rr, rs);
}

// One full RBGS sweep (colour c1, then the other) as a single launch, for the
// levels whose half-sweeps are launch-latency bound.  The second colour's
// update needs the first colour's new values at its neighbours; each such
// thread recomputes them from the previous iterate.  Reading zin and writing
// zout (separate buffers) keeps the launch race-free, and every value is
// formed with the operations of the two half-sweeps in their order, so the
// result is bitwise theirs.  zc (first sweep after prolongation only) is
// added to the neighbour values of the first colour's update, as there.
template <int DIM, bool MASK, typename ZCT = float>
__global__ void __launch_bounds__(kB) k_sweep2(Dims d, const float* __restrict__ r, const float* __restrict__ zin,
                                               float* __restrict__ zout, int c1, const double* shift,
                                               const ZCT* __restrict__ zc, Dims dc, const int* flags) {
#pragma clang fp contract(off)
  if (is_done(flags)) return;
  const float m = shift ? (float)*shift : 0.f;
  const int64_t N = (int64_t)d.n[0] * d.n[1] * d.n[2];
  const uint32_t id = (uint32_t)xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
  if (id >= N) return;
  if (!is_fluid<MASK>(d, id)) {
    zout[id] = 0.f;
    return;
  }
  int i, j, k;
  cell_of(d, id, i, j, k);
  if (((i + j + k) & 1) == c1) {
    zout[id] = gs_value<DIM, MASK, ZCT>(d, r, zin, zc, dc, m, i, j, k, id);
    return;
  }
  float s = 0.f, nn = 0.f;
  for_nbrs<DIM, MASK>(d, i, j, k, id, [&](uint32_t q, int a, int b, int c) {
    s += gs_value<DIM, MASK, ZCT>(d, r, zin, zc, dc, m, a, b, c, q);
    nn += 1.f;
  });
  zout[id] = (r[id] - m + s) * inv_count(nn);
}

// restrict (2D/vortex/mgpcg.py:152-158): coarse r = sum over the 2^d children
// of the fine residual r - A z, times 1/(dim-1).  One thread per coarse cell,
// so the sum order is fixed.  MASK: fluid children only
// (2D/vortex/mgpcg_solid.py:216-224).
template <int DIM, bool MASK, bool SPLIT, typename CT = float>
__global__ void __launch_bounds__(kB) k_restrict(Dims d, const float* __restrict__ r, const float* __restrict__ z,
                                                 Dims dc, CT* __restrict__ rc, const double* shift,
                                                 const int* flags) {
#pragma clang fp contract(off)
  if (is_done(flags)) return;
  float m = shift ? (float)*shift : 0.f;
  int64_t Nc = (int64_t)dc.n[0] * dc.n[1] * dc.n[2];
  uint32_t cid = (uint32_t)xcd_block(blockIdx.x, gridDim.x) * kB + threadIdx.x;
  if (cid >= Nc) return;
  int ci, cj, ck;
  cell_of(dc, cid, ci, cj, ck);
  // the children's values are all loaded before the first is used
  constexpr int NCH = DIM == 3 ? 8 : 4, KZ = DIM == 3 ? 2 : 1;
  float zn[NCH][2 * DIM], zv[NCH], rv[NCH];
  unsigned has[NCH];
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    const int i = 2 * ci + ch / (2 * KZ), j = 2 * cj + (ch / KZ) % 2, k = DIM == 3 ? 2 * ck + ch % KZ : 0;
    const uint32_t id = cell_at<SPLIT>(d, i, j, k);
    has[ch] = 0;
    zv[ch] = rv[ch] = 0.f;
    if (is_fluid<MASK>(d, id)) {
      has[ch] = nbr_load<DIM, MASK, SPLIT>(d, z, i, j, k, id, zn[ch]) | kFluidBit;
      zv[ch] = z[id];
      rv[ch] = r[id];
    }
  }
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) pin(zn[ch]);
  pin(zv);
  pin(rv);
  float acc = 0.f;
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    if (!has[ch]) continue;
    float s = 0.f, nn = 0.f;
#pragma unroll
    for (int q = 0; q < 2 * DIM; ++q)
      if (has[ch] >> q & 1u) {
        s += zn[ch][q];
        nn += 1.f;
      }
    acc += (rv[ch] - m) - (nn * zv[ch] - s);
  }
  Sto<CT>::st(rc, cid, acc * (1.0f / (float)(DIM - 1)));
}

// ---------------------------------------------------- LDS-tiled levels ----
// The coarse global levels are launch-latency bound: five launches per level
// and V-cycle (first sweep, second sweep, restriction; two post sweeps).
// Here one workgroup owns a tile of cells and keeps it in LDS with a halo
// of kTH cells, enough for the dependency depth of the level's 2 x 2
// half-sweeps: half-sweep t runs over the tile plus a halo shrinking by one
// cell per half-sweep, so the interior values come out exactly as the
// global half-sweeps produce them (same operations, same order, no
// multiply-add contraction), with the redundant halo work done in LDS.
//   k_tile_pre:  z from 0 by 2 RBGS sweeps (colour 0 first) + restriction;
//   k_tile_post: 2 RBGS sweeps (colour 1 first) from z, the coarse
//                correction added to the neighbour values of the first.
// Levels without an obstacle mask, with 2 pre/post sweeps, whose extents are
// multiples of the tile, use them.
constexpr int kTH = 4;  // halo: the 4 half-sweeps of 2 sweeps
constexpr int kTB = 512;  // threads of a tiled workgroup: 2 per column of the 3D tile face
template <int DIM> struct MgTile;
template <> struct MgTile<3> { static constexpr int T = 8, E = 8 + 2 * kTH, EK = E, N = E * E * E; };
template <> struct MgTile<2> { static constexpr int T = 16, E = 16 + 2 * kTH, EK = 1, N = E * E; };

template <int DIM>
struct TileFrame {
  int o[3];  // global index of local (0, 0, 0)
};

template <int DIM>
DFM_HD TileFrame<DIM> tile_frame(const Dims& d, uint32_t tile) {
  using M = MgTile<DIM>;
  const int t1 = d.n[1] / M::T, t2 = DIM == 3 ? d.n[2] / M::T : 1;
  const int tk = (int)(tile % (uint32_t)t2), tij = (int)(tile / (uint32_t)t2);
  const int tj = tij % t1, ti = tij / t1;
  TileFrame<DIM> f;
  f.o[0] = ti * M::T - kTH;
  f.o[1] = tj * M::T - kTH;
  f.o[2] = DIM == 3 ? tk * M::T - kTH : 0;
  return f;
}

// Visit the in-grid cells of colour `colour` (colour < 0: every cell) in the
// tile plus a halo of h, calling fn(i, j, k, l) with global indices and the
// LDS slot.  3D: the two threads of column (lj, lk) of the 16 x 16 tile face
// walk it along i on one colour, alternating cells, so no lane idles on the
// other colour and no index is divided; 2D: the 24 x 24 box is enumerated
// with constant divisors.
template <int DIM, class F>
DFM_HD void tile_pass(const Dims& d, const TileFrame<DIM>& f, int h, int colour, F&& fn) {
  using M = MgTile<DIM>;
  const int r0 = kTH - h, r1 = kTH + M::T + h;  // local range of the region on every axis
  if constexpr (DIM == 3) {
    static_assert(2 * M::E * M::E == kTB, "two threads per column of the tile face");
    const int c = threadIdx.x % (M::E * M::E), half = threadIdx.x / (M::E * M::E);
    const int lj = c / M::E, lk = c % M::E;
    const int j = f.o[1] + lj, k = f.o[2] + lk;
    if (lj < r0 || lj >= r1 || lk < r0 || lk >= r1 || j < 0 || j >= d.n[1] || k < 0 || k >= d.n[2]) return;
    int a0 = max(r0, -f.o[0]);
    const int a1 = min(r1, d.n[0] - f.o[0]);
    int step = 1;
    if (colour >= 0) {
      a0 += (colour - (f.o[0] + a0 + j + k)) & 1;
      step = 2;
    }
    // the column's cells alternate between its two threads
    for (int li = a0 + half * step; li < a1; li += 2 * step) fn(f.o[0] + li, j, k, (li * M::E + lj) * M::E + lk);
  } else {
    for (int t = threadIdx.x; t < M::E * M::E; t += kTB) {
      const int li = t / M::E, lj = t % M::E;
      if (li < r0 || li >= r1 || lj < r0 || lj >= r1) continue;
      const int i = f.o[0] + li, j = f.o[1] + lj;
      if (i < 0 || i >= d.n[0] || j < 0 || j >= d.n[1]) continue;
      if (colour >= 0 && ((i + j) & 1) != colour) continue;
      fn(i, j, 0, t);
    }
  }
}

// Load the tile plus its full halo from global field(s) into LDS.  3D: each
// thread's half column of E/2 cells is loaded with every load issued before
// any is stored (two waves per SIMD leave little else to hide the latency);
// cells outside the grid get 0.  ld(i, j, k, slot, inside) is called per cell.
template <int DIM, class L>
DFM_HD void tile_load(const Dims& d, const TileFrame<DIM>& f, L&& ld) {
  using M = MgTile<DIM>;
  if constexpr (DIM == 3) {
    const int c = threadIdx.x % (M::E * M::E), half = threadIdx.x / (M::E * M::E);
    const int lj = c / M::E, lk = c % M::E;
    const int j = f.o[1] + lj, k = f.o[2] + lk;
    const bool col = j >= 0 && j < d.n[1] && k >= 0 && k < d.n[2];
#pragma unroll
    for (int q = 0; q < M::E / 2; ++q) {
      const int li = half * (M::E / 2) + q;
      const int i = f.o[0] + li;
      ld(i, j, k, (li * M::E + lj) * M::E + lk, col && i >= 0 && i < d.n[0]);
    }
  } else {
    for (int t = threadIdx.x; t < M::E * M::E; t += kTB) {
      const int i = f.o[0] + t / M::E, j = f.o[1] + t % M::E;
      ld(i, j, 0, t, i >= 0 && i < d.n[0] && j >= 0 && j < d.n[1]);
    }
  }
}

// Neighbour sum of slot l (in-grid neighbours, for_nbrs order) of values
// v(slot); returns the count in nn.
template <int DIM, class V>
DFM_HD float tile_nbr_sum(const Dims& d, int i, int j, int k, int l, float& nn, V&& v) {
#pragma clang fp contract(off)
  using M = MgTile<DIM>;
  constexpr int SI = M::E * M::EK, SJ = M::EK;
  float s = 0.f;
  nn = 0.f;
  if (i > 0) { s += v(l - SI); nn += 1.f; }
  if (i < d.n[0] - 1) { s += v(l + SI); nn += 1.f; }
  if (j > 0) { s += v(l - SJ); nn += 1.f; }
  if (j < d.n[1] - 1) { s += v(l + SJ); nn += 1.f; }
  if (DIM == 3) {
    if (k > 0) { s += v(l - 1); nn += 1.f; }
    if (k < d.n[2] - 1) { s += v(l + 1); nn += 1.f; }
  }
  return s;
}

template <int DIM, typename FT = float, typename CT = float>
__global__ void __launch_bounds__(kTB) k_tile_pre(Dims d, const FT* __restrict__ r, FT* __restrict__ z, Dims dc,
                                                 CT* __restrict__ rc, const int* flags) {
#pragma clang fp contract(off)
  using M = MgTile<DIM>;
  __shared__ float rl[M::N];
  __shared__ float zl[M::N];
  if (is_done(flags)) return;
  const TileFrame<DIM> f = tile_frame<DIM>(d, blockIdx.x);
  if constexpr (DIM == 3) {  // stage in registers: every load issued before the first store
    float rv[M::E];
    int li = 0;
    tile_load<DIM>(d, f, [&](int i, int j, int k, int, bool in) {
      rv[li++ % M::E] = in ? Sto<FT>::ld(r, cidx(d, i, j, k)) : 0.f;
    });
    li = 0;
    tile_load<DIM>(d, f, [&](int, int, int, int l, bool) {
      rl[l] = rv[li++ % M::E];
      zl[l] = 0.f;
    });
  } else {
    tile_load<DIM>(d, f, [&](int i, int j, int, int l, bool in) {
      rl[l] = in ? Sto<FT>::ld(r, cidx(d, i, j, 0)) : 0.f;
      zl[l] = 0.f;
    });
  }
  __syncthreads();
  const float m = 0.f;
  // first sweep from z = 0: colour 0 gets (r - 0) / count (k_first_sweep)
  tile_pass<DIM>(d, f, kTH, 0, [&](int i, int j, int k, int l) {
    zl[l] = (rl[l] - m) * inv_count(nbr_count<DIM>(d, i, j, k));
  });
  __syncthreads();
  for (int t = 1; t < kTH; ++t) {  // half-sweeps 1..3: colours 1, 0, 1
    tile_pass<DIM>(d, f, kTH - t, t & 1, [&](int i, int j, int k, int l) {
      float nn;
      const float s = tile_nbr_sum<DIM>(d, i, j, k, l, nn, [&](int q) { return zl[q]; });
      zl[l] = (rl[l] - m + s) * inv_count(nn);
    });
    __syncthreads();
  }
  tile_pass<DIM>(d, f, 0, -1, [&](int i, int j, int k, int l) { Sto<FT>::st(z, cidx(d, i, j, k), zl[l]); });
  // restriction of the tile's interior (k_restrict): one thread per fine
  // cell forms its residual term, then one thread per coarse cell adds its
  // children's terms in k_restrict's order (i, j, k offsets, k fastest)
  constexpr int TC = M::T / 2, TCK = DIM == 3 ? TC : 1, KZ = DIM == 3 ? 2 : 1;
  constexpr int SI = M::E * M::EK, SJ = M::EK;
  constexpr int NC = TC * TC * TCK, CH = 4 * KZ;
  static_assert(NC * CH <= kTB, "one thread per fine cell of the tile");
  __syncthreads();  // zl is final; rl slots of the interior are reused below
  float term = 0.f;
  const int t = threadIdx.x;
  if (t < NC * CH) {
    const int cc = t / CH, ch = t % CH;  // coarse cell, child (a, b, c) = (ch / (2 KZ), ch / KZ % 2, ch % KZ)
    const int a = ch / (2 * KZ), b = (ch / KZ) % 2, c = ch % KZ;
    const int ck = (f.o[2] + kTH) / 2 + (DIM == 3 ? cc % TCK : 0);
    const int cj = (f.o[1] + kTH) / 2 + (cc / TCK) % TC;
    const int ci = (f.o[0] + kTH) / 2 + cc / (TCK * TC);
    const int i = 2 * ci + a, j = 2 * cj + b, k = DIM == 3 ? 2 * ck + c : 0;
    const int l = (i - f.o[0]) * SI + (j - f.o[1]) * SJ + (DIM == 3 ? k - f.o[2] : 0);
    float nn;
    const float s = tile_nbr_sum<DIM>(d, i, j, k, l, nn, [&](int q) { return zl[q]; });
    term = (rl[l] - m) - (nn * zl[l] - s);
  }
  __syncthreads();
  float* terms = rl;  // NC * CH <= kTB floats
  if (t < NC * CH) terms[t] = term;
  __syncthreads();
  if (t < NC) {
    float acc = 0.f;
    for (int q = 0; q < CH; ++q) acc += terms[t * CH + q];
    const int ck = (f.o[2] + kTH) / 2 + (DIM == 3 ? t % TCK : 0);
    const int cj = (f.o[1] + kTH) / 2 + (t / TCK) % TC;
    const int ci = (f.o[0] + kTH) / 2 + t / (TCK * TC);
    Sto<CT>::st(rc, cidx(dc, ci, cj, DIM == 3 ? ck : 0), acc * (1.0f / (float)(DIM - 1)));
  }
}

template <int DIM, typename FT = float, typename ZCT = float>
__global__ void __launch_bounds__(kTB) k_tile_post(Dims d, const FT* __restrict__ r, const FT* __restrict__ zin,
                                                  const ZCT* __restrict__ zc, Dims dc, FT* __restrict__ zout,
                                                  const int* flags) {
#pragma clang fp contract(off)
  using M = MgTile<DIM>;
  __shared__ float rl[M::N];
  __shared__ float zl[M::N];
  __shared__ float pl[M::N];  // coarse correction at each cell (its parent's value)
  if (is_done(flags)) return;
  const TileFrame<DIM> f = tile_frame<DIM>(d, blockIdx.x);
  if constexpr (DIM == 3) {  // stage in registers: every load issued before the first store
    float rv[M::E], zv[M::E], pv[M::E];
    int li = 0;
    tile_load<DIM>(d, f, [&](int i, int j, int k, int, bool in) {
      const int q = li++ % M::E;
      const uint32_t id = in ? cidx(d, i, j, k) : 0u;
      rv[q] = in ? Sto<FT>::ld(r, id) : 0.f;
      zv[q] = in ? Sto<FT>::ld(zin, id) : 0.f;
      pv[q] = in ? Sto<ZCT>::ld(zc, cidx(dc, i >> 1, j >> 1, k >> 1)) : 0.f;
    });
    li = 0;
    tile_load<DIM>(d, f, [&](int, int, int, int l, bool) {
      const int q = li++ % M::E;
      rl[l] = rv[q];
      zl[l] = zv[q];
      pl[l] = pv[q];
    });
  } else {
    tile_load<DIM>(d, f, [&](int i, int j, int, int l, bool in) {
      const uint32_t id = in ? cidx(d, i, j, 0) : 0u;
      rl[l] = in ? Sto<FT>::ld(r, id) : 0.f;
      zl[l] = in ? Sto<FT>::ld(zin, id) : 0.f;
      pl[l] = in ? Sto<ZCT>::ld(zc, cidx(dc, i >> 1, j >> 1, 0)) : 0.f;
    });
  }
  __syncthreads();
  const float m = 0.f;
  // first half-sweep (colour 1): neighbour values are z + the coarse correction
  tile_pass<DIM>(d, f, kTH - 1, 1, [&](int i, int j, int k, int l) {
    float nn;
    const float s = tile_nbr_sum<DIM>(d, i, j, k, l, nn, [&](int q) { return zl[q] + pl[q]; });
    zl[l] = (rl[l] - m + s) * inv_count(nn);
  });
  __syncthreads();
  for (int t = 1; t < kTH; ++t) {  // half-sweeps 1..3: colours 0, 1, 0
    tile_pass<DIM>(d, f, kTH - 1 - t, (t + 1) & 1, [&](int i, int j, int k, int l) {
      float nn;
      const float s = tile_nbr_sum<DIM>(d, i, j, k, l, nn, [&](int q) { return zl[q]; });
      zl[l] = (rl[l] - m + s) * inv_count(nn);
    });
    __syncthreads();
  }
  tile_pass<DIM>(d, f, 0, -1, [&](int i, int j, int k, int l) { Sto<FT>::st(zout, cidx(d, i, j, k), zl[l]); });
}

// Obstacle flags of the MG levels: level 0 = mask > 0 (the solver's
// bm[0] <= 0 fluid test); a coarse cell is solid iff all its 2^d children
// are (downsample_bm, 2D/vortex/mgpcg_solid.py:98-112).
__global__ void __launch_bounds__(kB) k_mask_level0(int64_t N, const float* __restrict__ mask,
                                                    uint8_t* __restrict__ bm, Red R) {
  double fluid = 0;
  const SlabWalk<uint32_t> w = slab_walk<kB>((uint32_t)N);
  for (uint32_t id = w.first; id < w.end; id += w.stride) {
    const uint8_t sld = mask[id] > 0.f ? 1 : 0;
    bm[id] = sld;
    fluid += sld ? 0.0 : 1.0;
  }
  grid_reduce(R, FIN_COUNT, fluid, 0.0);
}

template <int DIM>
__global__ void __launch_bounds__(kB) k_mask_coarsen(Dims d, const uint8_t* __restrict__ bm, Dims dc,
                                                     uint8_t* __restrict__ bmc) {
  int64_t Nc = (int64_t)dc.n[0] * dc.n[1] * dc.n[2];
  uint32_t cid = blockIdx.x * kB + threadIdx.x;
  if (cid >= Nc) return;
  int ci, cj, ck;
  cell_of(dc, cid, ci, cj, ck);
  uint8_t all = 1;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < (DIM == 3 ? 2 : 1); ++c)
        all &= bm[cidx(d, 2 * ci + a, 2 * cj + b, DIM == 3 ? 2 * ck + c : 0)];
  bmc[cid] = all;
}

// --------------------------------------------------- coarse tail in LDS ----
// Levels Lt..L-1 of the V-cycle run inside one workgroup with every level's
// r and z resident in LDS (FP32, or BF16 storage with FP32 arithmetic in the
// mixed-precision mode): pre-smoothing from zero, restriction, the `bottom`
// sweeps on the coarsest level, prolongation and post-smoothing -- the launch
// sequence the reference issues per level collapsed into one kernel.

constexpr int kTailMax = 8;
constexpr int64_t kTailAutoCells = 16384;  // largest first tail level chosen automatically
struct TailArgs {
  int nl;             // levels in the tail
  int n[kTailMax][3]; // their dims
  int off[kTailMax];  // element offset of level l's r in LDS (z follows r)
  int pre, bottom;
  int bottom_exact;   // 1: solve the coarsest level exactly (DCT-II) instead of RBGS
  float inv_dm1;      // restriction scale 1/(dim-1)
  Dims d[kTailMax];   // index helpers of each level
};

// Orthonormal DCT-II along one axis of an LDS array (or its inverse, DCT-III):
// out[.., k, ..] = sum_m in[.., m, ..] * c_k cos(pi k (2m+1) / 2N) (forward).
// The all-Neumann Laplacian factorises over axes and each factor is
// diagonalised by this basis with eigenvalues 2 - 2 cos(pi k / N).
template <typename T>
DFM_HD void lds_dct_axis(const T* in, T* out, const int* n, int axis, bool inverse) {
  const int N = n[axis];
  const int stride = axis == 0 ? n[1] * n[2] : (axis == 1 ? n[2] : 1);
  const int total = n[0] * n[1] * n[2];
  const float pi_2n = 3.14159265358979f / (2.0f * (float)N);
  const float c0 = sqrtf(1.0f / (float)N), c1 = sqrtf(2.0f / (float)N);
  for (int id = threadIdx.x; id < total; id += blockDim.x) {
    int coord = (id / stride) % N;
    int base = id - coord * stride;
    float acc = 0.f;
    for (int m = 0; m < N; ++m) {
      // forward: k = coord (output frequency), m = sample; inverse: swapped
      int kf = inverse ? m : coord, xs = inverse ? coord : m;
      float c = (kf == 0 ? c0 : c1) * __cosf(pi_2n * (float)(kf * (2 * xs + 1)));
      acc = fmaf(Sto<T>::ld(in, base + m * stride), c, acc);
    }
    Sto<T>::st(out, id, acc);
  }
  __syncthreads();
}

// z = A^+ r on one LDS level (mean-zero solution of the Neumann problem):
// forward DCT on every axis, divide by the eigenvalue sum (zero mode -> 0),
// inverse DCT.  r is used as scratch; the result lands in z.
template <int DIM, typename T>
DFM_HD void lds_exact_solve(T* r, T* z, const int* n) {
  T* a = r;
  T* b = z;
  for (int ax = 0; ax < DIM; ++ax) {
    lds_dct_axis<T>(a, b, n, ax, false);
    T* t = a; a = b; b = t;
  }
  const int total = n[0] * n[1] * n[2];
  for (int id = threadIdx.x; id < total; id += blockDim.x) {
    int k = id % n[2], t = id / n[2];
    int j = t % n[1], i = t / n[1];
    float lam = 2.0f - 2.0f * __cosf(3.14159265358979f * (float)i / (float)n[0]) +
                2.0f - 2.0f * __cosf(3.14159265358979f * (float)j / (float)n[1]);
    if (DIM == 3) lam += 2.0f - 2.0f * __cosf(3.14159265358979f * (float)k / (float)n[2]);
    float v = Sto<T>::ld(a, id);
    Sto<T>::st(a, id, (i == 0 && j == 0 && k == 0) ? 0.f : v / lam);
  }
  __syncthreads();
  for (int ax = 0; ax < DIM; ++ax) {
    lds_dct_axis<T>(a, b, n, ax, true);
    T* t = a; a = b; b = t;
  }
  if (a != z) {
    for (int id = threadIdx.x; id < total; id += blockDim.x) Sto<T>::st(z, id, Sto<T>::ld(a, id));
    __syncthreads();
  }
}

// One half-sweep of colour `colour` on an LDS-resident level: every thread
// owns one cell of that colour per pass (pairs along the innermost axis).
template <int DIM, typename T>
DFM_HD void lds_half_sweep(T* r, T* z, const Dims& d, int colour) {
  const int* n = d.n;
  const uint32_t pairs = (uint32_t)(DIM == 3 ? n[0] * n[1] : n[0]) * d.fh.d;
  const int sx = n[1] * n[2], sy = n[2];
  for (uint32_t t = threadIdx.x; t < pairs; t += blockDim.x) {
    int i, j, k;
    if (!colour_cell<DIM>(d, t, colour, i, j, k)) continue;
    int id = (i * n[1] + j) * n[2] + k;
    float s = 0.f, nn = 0.f;
    if (i > 0) { s += Sto<T>::ld(z, id - sx); nn += 1.f; }
    if (i < n[0] - 1) { s += Sto<T>::ld(z, id + sx); nn += 1.f; }
    if (j > 0) { s += Sto<T>::ld(z, id - sy); nn += 1.f; }
    if (j < n[1] - 1) { s += Sto<T>::ld(z, id + sy); nn += 1.f; }
    if (DIM == 3) {
      if (k > 0) { s += Sto<T>::ld(z, id - 1); nn += 1.f; }
      if (k < n[2] - 1) { s += Sto<T>::ld(z, id + 1); nn += 1.f; }
    }
    Sto<T>::st(z, id, (Sto<T>::ld(r, id) + s) * inv_count(nn));
  }
  __syncthreads();
}

template <int DIM, typename T>
__global__ void __launch_bounds__(1024) k_tail(TailArgs A, const T* __restrict__ rin, T* __restrict__ zout,
                                               const double* shift, const int* flags) {
  if (is_done(flags)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* smem = reinterpret_cast<T*>(smem_raw);
  const float m = shift ? (float)*shift : 0.f;
  auto size_of = [&](int l) { return A.n[l][0] * A.n[l][1] * A.n[l][2]; };
  {
    T* r0 = smem + A.off[0];
    T* z0 = r0 + size_of(0);
    for (int id = threadIdx.x; id < size_of(0); id += blockDim.x) {
      Sto<T>::st(r0, id, Sto<T>::ld(rin, id) - m);
      Sto<T>::st(z0, id, 0.f);
    }
  }
  __syncthreads();
  // down: smooth from zero, restrict
  for (int l = 0; l < A.nl - 1; ++l) {
    T* r = smem + A.off[l];
    T* z = r + size_of(l);
    const int* n = A.n[l];
    for (int s = 0; s < A.pre; ++s) {
      lds_half_sweep<DIM, T>(r, z, A.d[l], 0);
      lds_half_sweep<DIM, T>(r, z, A.d[l], 1);
    }
    T* rc = smem + A.off[l + 1];
    T* zc = rc + size_of(l + 1);
    const int* nc = A.n[l + 1];
    for (int cid = threadIdx.x; cid < size_of(l + 1); cid += blockDim.x) {
      int ci, cj, ck;
      cell_of(A.d[l + 1], (uint32_t)cid, ci, cj, ck);
      float acc = 0.f;
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
          for (int c = 0; c < (DIM == 3 ? 2 : 1); ++c) {
            int i = 2 * ci + a, j = 2 * cj + b, k = DIM == 3 ? 2 * ck + c : 0;
            int id = (i * n[1] + j) * n[2] + k;
            float s = 0.f, nn = 0.f;
            if (i > 0) { s += Sto<T>::ld(z, id - n[1] * n[2]); nn += 1.f; }
            if (i < n[0] - 1) { s += Sto<T>::ld(z, id + n[1] * n[2]); nn += 1.f; }
            if (j > 0) { s += Sto<T>::ld(z, id - n[2]); nn += 1.f; }
            if (j < n[1] - 1) { s += Sto<T>::ld(z, id + n[2]); nn += 1.f; }
            if (DIM == 3) {
              if (k > 0) { s += Sto<T>::ld(z, id - 1); nn += 1.f; }
              if (k < n[2] - 1) { s += Sto<T>::ld(z, id + 1); nn += 1.f; }
            }
            acc += Sto<T>::ld(r, id) - (nn * Sto<T>::ld(z, id) - s);
          }
      Sto<T>::st(rc, cid, acc * A.inv_dm1);
      Sto<T>::st(zc, cid, 0.f);
    }
    __syncthreads();
  }
  // bottom
  {
    int l = A.nl - 1;
    T* r = smem + A.off[l];
    T* z = r + size_of(l);
    if (A.bottom_exact) {
      lds_exact_solve<DIM, T>(r, z, A.n[l]);
    } else {
      for (int s = 0; s < A.bottom; ++s) {
        lds_half_sweep<DIM, T>(r, z, A.d[l], 0);
        lds_half_sweep<DIM, T>(r, z, A.d[l], 1);
      }
    }
  }
  // up: prolongate, post-smooth (colour 1 first)
  for (int l = A.nl - 2; l >= 0; --l) {
    T* r = smem + A.off[l];
    T* z = r + size_of(l);
    T* zc = smem + A.off[l + 1] + size_of(l + 1);
    const int* n = A.n[l];
    const int* nc = A.n[l + 1];
    for (int id = threadIdx.x; id < size_of(l); id += blockDim.x) {
      int i, j, k;
      cell_of(A.d[l], (uint32_t)id, i, j, k);
      int cid = ((i >> 1) * nc[1] + (j >> 1)) * nc[2] + (DIM == 3 ? (k >> 1) : 0);
      Sto<T>::st(z, id, Sto<T>::ld(z, id) + Sto<T>::ld(zc, cid));
    }
    __syncthreads();
    for (int s = 0; s < A.pre; ++s) {
      lds_half_sweep<DIM, T>(r, z, A.d[l], 1);
      lds_half_sweep<DIM, T>(r, z, A.d[l], 0);
    }
  }
  T* z0 = smem + A.off[0] + size_of(0);
  for (int id = threadIdx.x; id < size_of(0); id += blockDim.x) Sto<T>::st(zout, id, Sto<T>::ld(z0, id));
}

// ------------------------------------------------------------- handle -----
// A solve between its queued start and its completion.
struct Pending {
  int active;
  int it, refinements, max_iters;
  bool use_graph;
  hipStream_t st;
  Faces F;           // velocity of a projection (for the gradient subtraction)
};

struct Solver {
  Geo G;
  int L, pre, bottom;
  Lvl lv[kMaxLevels];
  Dims ld[kMaxLevels];  // index helpers per level
  double* x;  // PCG iterate (pressure), FP64
  float* b;   // recentred right-hand side, kept for the FP64 refinement residual
  int max_refine;
  float* p;
  double* part;   // 2 x kRedBlocks block partials
  unsigned* ctr;  // grid-reduction arrival counters (kCtrWords)
  double* sc;
  int* flags;
  int* h_flags;
  double* h_sc;
  int last_iters;
  double last_res;
  int check_every;
  // coarse tail: levels tail_lt..L-1 run in one workgroup with LDS storage
  int bf16_coarse;   // option: store levels >= 2 in BF16 inside the tail
  int bottom_exact;  // option: exact (DCT) coarsest-level solve in the tail
  int tail_bf16;     // the tail in use is BF16
  int tail_lt;       // first tail level (L = no tail)
  int tail_min;      // option: lowest level the tail may start at (0 = any that fits)
  size_t tail_bytes; // dynamic LDS of the tail kernel
  TailArgs tail;
  float* z0_final;   // level-0 buffer holding M^-1 r after the last V-cycle
  // captured PCG iterations
  int use_graph;
  int predict;       // iterations the previous solve took to first convergence
  hipStream_t cap;
  hipEvent_t rdy;  // polled for the flag readback (spin_sync)
  hipGraphExec_t gexec;
  int gexec_iters;
  Pending pend;      // the solve between project_begin and project_end
  // obstacle mask (MGPCG_2_solid): solid flags per level, fluid cell count
  int masked;
  uint8_t* bm[kMaxLevels];
  double nfluid;
  int bf16_lo;       // first level stored in BF16 in global memory (L: none); see bf16_first_level
  int fused_sweeps;  // option: one launch per RBGS sweep on levels with z2
  int tiled_sweeps;  // option: LDS-tiled pre/post smoothing on eligible coarse levels
  int split_opt;     // option: level 0 colour-split where eligible (set_layout)
  int row_opt;       // option: row half-sweeps (1: where eligible and wide enough, 2: where eligible)
  double* xn;        // natural-order copy of a colour-split pressure (on request)
  hipStream_t last_st;  // stream of the last solve
  int ready;         // partials, counters, control block and x zeroed (ensure_ready)
};

// Largest LDS a workgroup may request on this device.
size_t max_lds_bytes() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 65536;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) return 65536;
  return (size_t)v;
}

// Pick the first tail level: the largest suffix of levels whose r and z fit
// in LDS (BF16 storage allowed only for levels >= 2).
void configure_tail(Solver& s, size_t cap) {
  int best = s.L;
  size_t best_bytes = 0;
  for (int l = s.L - 1; l >= s.tail_min && s.L - l <= kTailMax; --l) {
    // one workgroup smooths the tail: a first level beyond kTailAutoCells is
    // faster on the global (tiled) kernels even where it fits in LDS, as
    // 32^3 does in BF16 (one tail 137 us against 2 tiled launches plus a
    // 16^3 tail, ~45 us); an explicit DFM_MG_TAIL_LEVEL is not capped
    if (s.tail_min == 0 && s.lv[l].size > kTailAutoCells) break;
    const size_t es = (s.bf16_coarse && l >= 2) ? 2 : 4;
    size_t tot = 0;
    for (int k = l; k < s.L; ++k) tot += 2 * (size_t)s.lv[k].size * es;
    if (tot > cap) break;
    best = l;
    best_bytes = tot;
  }
  s.tail_lt = best;
  s.tail_bytes = best_bytes;
  s.tail_bf16 = (s.bf16_coarse && best >= 2) ? 1 : 0;
  if (best >= s.L) return;
  s.tail.nl = s.L - best;
  int off = 0;
  for (int k = 0; k < s.tail.nl; ++k) {
    for (int a = 0; a < 3; ++a) s.tail.n[k][a] = s.lv[best + k].n[a];
    s.tail.d[k] = make_dims(s.tail.n[k][0], s.tail.n[k][1], s.tail.n[k][2], s.G.dim);
    s.tail.off[k] = off;
    off += 2 * (int)s.lv[best + k].size;
  }
  s.tail.pre = s.pre;
  s.tail.bottom = s.bottom;
  s.tail.bottom_exact = s.bottom_exact;
  s.tail.inv_dm1 = 1.0f / (float)(s.G.dim - 1);
}

// Dynamic-LDS opt-in for the tail instantiation this handle launches.
bool enable_tail_lds(const Solver& s) {
  if (s.tail_lt >= s.L) return true;
  const void* fn = s.G.dim == 2 ? (s.tail_bf16 ? (const void*)k_tail<2, uint16_t> : (const void*)k_tail<2, float>)
                                : (s.tail_bf16 ? (const void*)k_tail<3, uint16_t> : (const void*)k_tail<3, float>);
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)s.tail_bytes) == hipSuccess;
}

// Choose the tail for the current options; falls back to the 64 KB static
// limit (then to no tail) if the runtime refuses the larger allocation.
void setup_tail(Solver& s) {
  if (s.masked) {  // the LDS tail and its exact bottom solve assume no obstacles
    s.tail_lt = s.L;
    return;
  }
  configure_tail(s, max_lds_bytes());
  if (!enable_tail_lds(s)) {
    (void)hipGetLastError();
    configure_tail(s, 65536);
    if (!enable_tail_lds(s)) {
      (void)hipGetLastError();
      s.tail_lt = s.L;
    }
  }
}

inline unsigned grid1(int64_t n) { return (unsigned)((n + kB - 1) / kB); }
inline unsigned gridr(int64_t n) {
  int64_t b = (n + kB - 1) / kB;
  return (unsigned)(b < kRedBlocks ? b : kRedBlocks);
}

// Whether a reducing walk over n cells gives each thread more than kUnroll
// cells (then the stencil walks keep kUnroll cells' loads in flight; below
// that the unrolled form only costs occupancy).
inline bool long_walk(int64_t n) { return n > (int64_t)kRedBlocks * kB * kUnroll; }

inline Red red_of(const Solver& s, double tol = 0.0, int relative = 0) {
  const double n = s.masked ? (s.nfluid > 0 ? s.nfluid : 1.0) : (double)s.lv[0].size;
  return Red{s.part, s.ctr, s.sc, s.flags, 1.0 / n, tol, relative};
}

inline uint32_t pairs_of(const Dims& d, int dim) {
  return (uint32_t)(dim == 3 ? d.n[0] * d.n[1] : d.n[0]) * d.fh.d;
}

// z.r of the V-cycle result and its scalar step, fused into the last level-0
// half-sweep when level 0 is smoothed by the global kernels.
struct ZrFuse {
  int op;    // FIN_ZR_INIT or FIN_BETA
  float* p;  // also p = z (start and restart of PCG), else nullptr
};

// `sweeps` RBGS sweeps (colour c1 first, then the other) on level l, starting
// from z = 0 when from_zero (first sweep fused into one kernel; z is set to
// lv[l].z) or else from the iterate in buffer z.  zc_prolong fuses the
// prolongation into the first sweep; zr (if given) is fused into the last
// half-sweep.  Levels with a second buffer run each full sweep as one k_sweep2
// launch, alternating between lv[l].z and lv[l].z2; the others sweep in place.
// On return z names the buffer holding the result.  Returns whether zr was done.
// One global half-sweep of colour c (no z.r): several pairs per thread
// without an obstacle mask, one per thread with it.
// Whether the colour-split level with dims d runs the row kernels (3D, an
// innermost extent that is a multiple of 8; with a coarse correction, a
// coarse innermost extent of half the fine one).
inline bool row_ok(const Solver& s, const Dims& d, const Dims* dc = nullptr) {
  return s.row_opt && d.hh != 0 && d.n[2] % (2 * kRowV) == 0 && (!dc || 2 * dc->n[2] == d.n[2]);
}

inline unsigned row_grid(const Dims& d) { return (row_threads(d) + kB - 1) / kB; }

// The natural-order counterpart of row_ok (k_row_half_nat).  With option
// value 1 only on levels whose row launch has a workgroup for each of the
// 256 CUs: below that (a 64^3 level has 128) the per-pair sweep, with twice
// the threads, measured as fast or faster.
inline bool row_ok_nat(const Solver& s, const Dims& d, const Dims* dc = nullptr) {
  return s.row_opt && d.hh == 0 && d.n[2] % (2 * kRowV) == 0 && (!dc || 2 * dc->n[2] == d.n[2]) &&
         (s.row_opt == 2 || row_grid(d) >= 256);
}

template <int DIM, bool MASK, bool SPLIT>
void half_sweep(Solver& s, hipStream_t st, const Dims& d, float* r, float* z, int c, const double* shift,
                const float* zc, const Dims& dc, const Red& R, bool zc_bf16 = false) {
  const uint32_t pairs = pairs_of(d, DIM);
  if constexpr (!SPLIT && DIM == 3 && !MASK) {
    if (row_ok_nat(s, d, zc ? &dc : nullptr)) {
      const dim3 g(row_grid(d));
      if (!zc)
        hipLaunchKernelGGL((k_row_half_nat<1>), g, dim3(kB), 0, st, d, (const float*)r, z, c, shift,
                           (const float*)nullptr, dc, (const int*)s.flags);
      else if (zc_bf16)
        hipLaunchKernelGGL((k_row_half_nat<2, bf16_t>), g, dim3(kB), 0, st, d, (const float*)r, z, c, shift,
                           reinterpret_cast<const bf16_t*>(zc), dc, (const int*)s.flags);
      else
        hipLaunchKernelGGL((k_row_half_nat<2>), g, dim3(kB), 0, st, d, (const float*)r, z, c, shift, zc, dc,
                           (const int*)s.flags);
      return;
    }
  }
  if constexpr (SPLIT && DIM == 3 && !MASK) {
    // (the finest level's coarse correction, level 1, is never stored in BF16)
    if (!zc_bf16 && row_ok(s, d, zc ? &dc : nullptr)) {
      const dim3 g(row_grid(d));
      if (!zc)
        hipLaunchKernelGGL((k_row_half<1>), g, dim3(kB), 0, st, d, (const float*)r, z, c, shift,
                           (const float*)nullptr, dc, (const int*)s.flags);
      else
        hipLaunchKernelGGL((k_row_half<2>), g, dim3(kB), 0, st, d, (const float*)r, z, c, shift, zc, dc,
                           (const int*)s.flags);
      return;
    }
  }
  if constexpr (!MASK) {
    const unsigned g = (unsigned)((pairs + (uint32_t)(kB * kSweepU) - 1) / (uint32_t)(kB * kSweepU));
    if (zc_bf16)
      hipLaunchKernelGGL((k_smooth_u<DIM, SPLIT, bf16_t>), dim3(g), dim3(kB), 0, st, d, (const float*)r, z, c, shift,
                         reinterpret_cast<const bf16_t*>(zc), dc, (const int*)s.flags);
    else
      hipLaunchKernelGGL((k_smooth_u<DIM, SPLIT>), dim3(g), dim3(kB), 0, st, d, (const float*)r, z, c, shift, zc, dc,
                         (const int*)s.flags);
  } else {
    hipLaunchKernelGGL((k_smooth<DIM, false, MASK, SPLIT>), dim3(grid1(pairs)), dim3(kB), 0, st, d, r, z, c, shift,
                       zc, dc, R, 0, (float*)nullptr);
  }
}

inline bool bf16_level(const Solver& s, int l) { return l >= s.bf16_lo; }

template <int DIM, bool MASK, bool SPLIT>
bool sweeps_global_t(Solver& s, hipStream_t st, int l, int sweeps, int c1, bool from_zero, const double* shift,
                     const float* zc_prolong, const ZrFuse* zr, float*& z) {
  Lvl& L = s.lv[l];
  const bool zc_bf16 = l + 1 < s.L && bf16_level(s, l + 1);  // MASK never stores BF16
  const Dims& d = s.ld[l];
  const Dims& dc = s.ld[l + 1 < s.L ? l + 1 : l];
  const Red R = red_of(s);
  int k0 = 0;
  if (from_zero) {
    z = L.z;
    if constexpr (MASK) {
      hipLaunchKernelGGL((k_first_sweep<DIM, MASK>), dim3(grid1(L.size)), dim3(kB), 0, st, d, L.r, z, shift,
                         s.flags);
    } else {
      // colour c1 from zero, then an ordinary half-sweep of the other colour
      // (k_first_sweep's result bitwise)
      const unsigned g = (unsigned)((pairs_of(d, DIM) + (uint32_t)(kB * kSweepU) - 1) / (uint32_t)(kB * kSweepU));
      if (SPLIT && DIM == 3 && row_ok(s, d))
        hipLaunchKernelGGL((k_row_half<0>), dim3(row_grid(d)), dim3(kB), 0, st, d, (const float*)L.r, z, c1, shift,
                           (const float*)nullptr, dc, (const int*)s.flags);
      else if (!SPLIT && DIM == 3 && row_ok_nat(s, d))
        hipLaunchKernelGGL((k_row_half_nat<0>), dim3(row_grid(d)), dim3(kB), 0, st, d, (const float*)L.r, z, c1,
                           shift, (const float*)nullptr, dc, (const int*)s.flags);
      else
        hipLaunchKernelGGL((k_first_half<DIM, SPLIT>), dim3(g), dim3(kB), 0, st, d, (const float*)L.r, z, c1, shift,
                           (const int*)s.flags);
      half_sweep<DIM, MASK, SPLIT>(s, st, d, L.r, z, 1 - c1, shift, (const float*)nullptr, dc, R);
    }
    k0 = 1;
  }
  const uint32_t pairs = pairs_of(d, DIM);
  const bool fused = !SPLIT && s.fused_sweeps && L.z2 && !zr;
  bool zr_done = false;
  for (int it = k0; it < sweeps; ++it) {
    const float* zc = (it == k0 && zc_prolong) ? zc_prolong : nullptr;
    if (fused) {
      float* zout = z == L.z ? L.z2 : L.z;
      if (zc && zc_bf16)
        hipLaunchKernelGGL((k_sweep2<DIM, MASK, bf16_t>), dim3(grid1(L.size)), dim3(kB), 0, st, d, (const float*)L.r,
                           (const float*)z, zout, c1, shift, reinterpret_cast<const bf16_t*>(zc), dc,
                           (const int*)s.flags);
      else
        hipLaunchKernelGGL((k_sweep2<DIM, MASK>), dim3(grid1(L.size)), dim3(kB), 0, st, d, (const float*)L.r,
                           (const float*)z, zout, c1, shift, zc, dc, (const int*)s.flags);
      z = zout;
      continue;
    }
    half_sweep<DIM, MASK, SPLIT>(s, st, d, L.r, z, c1, shift, zc, dc, R, zc != nullptr && zc_bf16);
    if (zr && it == sweeps - 1) {
      if (SPLIT && DIM == 3 && !MASK && row_ok(s, d))
        hipLaunchKernelGGL(k_row_zr, dim3(gridr(row_threads(d))), dim3(kB), 0, st, d, (const float*)L.r, z, 1 - c1,
                           shift, R, zr->op, zr->p);
      else
        hipLaunchKernelGGL((k_smooth<DIM, true, MASK, SPLIT>), dim3(gridr(pairs)), dim3(kB), 0, st, d, L.r, z,
                           1 - c1, shift, (const float*)nullptr, dc, R, zr->op, zr->p);
      zr_done = true;
    } else {
      half_sweep<DIM, MASK, SPLIT>(s, st, d, L.r, z, 1 - c1, shift, (const float*)nullptr, dc, R);
    }
  }
  return zr_done;
}

// Level 0 in the colour-split layout takes the SPLIT kernels (never with an
// obstacle mask); every other level is stored in natural order.
inline bool split_level(const Solver& s, int l);

template <int DIM, bool MASK>
bool sweeps_global(Solver& s, hipStream_t st, int l, int sweeps, int c1, bool from_zero, const double* shift,
                   const float* zc_prolong, const ZrFuse* zr, float*& z) {
  if constexpr (!MASK) {
    if (split_level(s, l))
      return sweeps_global_t<DIM, false, true>(s, st, l, sweeps, c1, from_zero, shift, zc_prolong, zr, z);
  }
  return sweeps_global_t<DIM, MASK, false>(s, st, l, sweeps, c1, from_zero, shift, zc_prolong, zr, z);
}

// Whether level l runs its pre/post smoothing in the LDS-tiled kernels:
// a coarse global level (not 0, second buffer present), 2 pre/post sweeps,
// extents multiples of the tile, fused sweeps enabled.
inline bool tiled_level(const Solver& s, int l) {
  if (l == 0 || !s.fused_sweeps || s.pre != 2 || !s.lv[l].z2 || s.tiled_sweeps == 0) return false;
  const int T = s.G.dim == 3 ? MgTile<3>::T : MgTile<2>::T;
  for (int a = 0; a < s.G.dim; ++a)
    if (s.lv[l].n[a] % T) return false;
  return true;
}

inline bool split_level(const Solver& s, int l) { return l == 0 && s.ld[0].hh != 0; }

template <int DIM>
inline unsigned tile_count(const Dims& d) {
  const int T = MgTile<DIM>::T;
  return (unsigned)((d.n[0] / T) * (d.n[1] / T) * (DIM == 3 ? d.n[2] / T : 1));
}

// apply_preconditioner (2D/vortex/mgpcg.py:182-200): z0 = M^-1 r0, followed by
// z.r and its scalar step `zr`.  Levels below tail_lt use global kernels;
// levels tail_lt.. run in one LDS kernel.  With an obstacle mask the bottom
// level gets bottom/2 sweeps in the order (0, 1) then bottom/2 in (1, 0)
// (2D/vortex/mgpcg_solid.py:263-268) and no LDS tail.
template <int DIM, bool MASK>
void vcycle(Solver& s, hipStream_t st, const double* shift0, const ZrFuse& zr) {
  float* cur[kMaxLevels] = {nullptr};
  const bool tail = s.tail_lt < s.L;
  const int Lg = tail ? s.tail_lt : s.L - 1;  // levels smoothed with global kernels before the bottom
  bool zr_done = false;
  for (int l = 0; l < Lg; ++l) {
    const double* sh = l == 0 ? shift0 : nullptr;
    if (!MASK && tiled_level(s, l)) {
      cur[l] = s.lv[l].z;
      const dim3 g(tile_count<DIM>(s.ld[l]));
      if (bf16_level(s, l))
        hipLaunchKernelGGL((k_tile_pre<DIM, bf16_t, bf16_t>), g, dim3(kTB), 0, st, s.ld[l],
                           (const bf16_t*)s.lv[l].r, (bf16_t*)cur[l], s.ld[l + 1], (bf16_t*)s.lv[l + 1].r,
                           (const int*)s.flags);
      else if (bf16_level(s, l + 1))
        hipLaunchKernelGGL((k_tile_pre<DIM, float, bf16_t>), g, dim3(kTB), 0, st, s.ld[l], (const float*)s.lv[l].r,
                           cur[l], s.ld[l + 1], (bf16_t*)s.lv[l + 1].r, (const int*)s.flags);
      else
        hipLaunchKernelGGL((k_tile_pre<DIM>), g, dim3(kTB), 0, st, s.ld[l], (const float*)s.lv[l].r, cur[l],
                           s.ld[l + 1], s.lv[l + 1].r, (const int*)s.flags);
      continue;
    }
    sweeps_global<DIM, MASK>(s, st, l, s.pre, 0, true, sh, nullptr, nullptr, cur[l]);
    Lvl& F = s.lv[l];
    Lvl& C = s.lv[l + 1];
    bool split_done = false;
    if constexpr (!MASK) {
      if (split_level(s, l)) {
        if (bf16_level(s, l + 1))
          hipLaunchKernelGGL((k_restrict<DIM, false, true, bf16_t>), dim3(grid1(C.size)), dim3(kB), 0, st, s.ld[l],
                             F.r, cur[l], s.ld[l + 1], (bf16_t*)C.r, sh, s.flags);
        else
          hipLaunchKernelGGL((k_restrict<DIM, false, true>), dim3(grid1(C.size)), dim3(kB), 0, st, s.ld[l], F.r,
                             cur[l], s.ld[l + 1], C.r, sh, s.flags);
        split_done = true;
      }
    }
    if (!split_done) {
      if (bf16_level(s, l + 1))
        hipLaunchKernelGGL((k_restrict<DIM, MASK, false, bf16_t>), dim3(grid1(C.size)), dim3(kB), 0, st, s.ld[l], F.r,
                           cur[l], s.ld[l + 1], (bf16_t*)C.r, sh, s.flags);
      else
        hipLaunchKernelGGL((k_restrict<DIM, MASK, false>), dim3(grid1(C.size)), dim3(kB), 0, st, s.ld[l], F.r,
                           cur[l], s.ld[l + 1], C.r, sh, s.flags);
    }
  }
  if (tail) {
    Lvl& T = s.lv[s.tail_lt];
    const double* sh = s.tail_lt == 0 ? shift0 : nullptr;
    if (s.tail_bf16)
      hipLaunchKernelGGL((k_tail<DIM, bf16_t>), dim3(1), dim3(1024), s.tail_bytes, st, s.tail, (const bf16_t*)T.r,
                         (bf16_t*)T.z, sh, s.flags);
    else
      hipLaunchKernelGGL((k_tail<DIM, float>), dim3(1), dim3(1024), s.tail_bytes, st, s.tail, T.r, T.z, sh, s.flags);
    cur[s.tail_lt] = T.z;
  } else {
    const double* sh = s.L == 1 ? shift0 : nullptr;
    const ZrFuse* zb = s.L == 1 ? &zr : nullptr;
    float*& zb_cur = cur[s.L - 1];
    if (MASK && s.bottom >= 2) {
      const int first = (s.bottom + 1) / 2;
      sweeps_global<DIM, MASK>(s, st, s.L - 1, first, 0, true, sh, nullptr, nullptr, zb_cur);
      zr_done = sweeps_global<DIM, MASK>(s, st, s.L - 1, s.bottom - first, 1, false, sh, nullptr, zb, zb_cur);
    } else {
      zr_done = sweeps_global<DIM, MASK>(s, st, s.L - 1, s.bottom, 0, true, sh, nullptr, zb, zb_cur);
    }
  }
  for (int l = Lg - 1; l >= 0; --l) {
    const double* sh = l == 0 ? shift0 : nullptr;
    if (!MASK && tiled_level(s, l)) {
      float* out = cur[l] == s.lv[l].z ? s.lv[l].z2 : s.lv[l].z;
      const dim3 g(tile_count<DIM>(s.ld[l]));
      if (bf16_level(s, l))
        hipLaunchKernelGGL((k_tile_post<DIM, bf16_t, bf16_t>), g, dim3(kTB), 0, st, s.ld[l],
                           (const bf16_t*)s.lv[l].r, (const bf16_t*)cur[l], (const bf16_t*)cur[l + 1], s.ld[l + 1],
                           (bf16_t*)out, (const int*)s.flags);
      else if (bf16_level(s, l + 1))
        hipLaunchKernelGGL((k_tile_post<DIM, float, bf16_t>), g, dim3(kTB), 0, st, s.ld[l], (const float*)s.lv[l].r,
                           (const float*)cur[l], (const bf16_t*)cur[l + 1], s.ld[l + 1], out, (const int*)s.flags);
      else
        hipLaunchKernelGGL((k_tile_post<DIM>), g, dim3(kTB), 0, st, s.ld[l], (const float*)s.lv[l].r,
                           (const float*)cur[l], (const float*)cur[l + 1], s.ld[l + 1], out, (const int*)s.flags);
      cur[l] = out;
      continue;
    }
    const bool d0 = sweeps_global<DIM, MASK>(s, st, l, s.pre, 1, false, sh, cur[l + 1], l == 0 ? &zr : nullptr,
                                             cur[l]);
    if (l == 0) zr_done = d0;
  }
  s.z0_final = cur[0];
  if (!zr_done) {
    Lvl& L0 = s.lv[0];
    hipLaunchKernelGGL((k_zr<MASK>), dim3(gridr(L0.size)), dim3(kB), 0, st, L0.size, (const float*)s.z0_final, L0.r,
                       shift0, (const uint8_t*)s.bm[0], red_of(s), zr.op, zr.p ? 1 : 0, zr.p);
  }
}

// Ap = A p with p.Ap -> alpha, and the FP64 refinement residual b - A x
// (after convergence only), in the level-0 layout and walk length's form.
template <int DIM, bool MASK, bool SPLIT>
void launch_Ap_dot_t(Solver& s, hipStream_t st, unsigned gr, const Red& R) {
  if (long_walk(s.lv[0].size))
    hipLaunchKernelGGL((k_Ap_dot<DIM, MASK, kUnroll, SPLIT>), dim3(gr), dim3(kB), 0, st, s.ld[0], (const float*)s.p, R);
  else
    hipLaunchKernelGGL((k_Ap_dot<DIM, MASK, 1, SPLIT>), dim3(gr), dim3(kB), 0, st, s.ld[0], (const float*)s.p, R);
}

template <int DIM, bool MASK, bool SPLIT>
void launch_update_xr_t(Solver& s, hipStream_t st, unsigned gr, const Red& R) {
  if (long_walk(s.lv[0].size))
    hipLaunchKernelGGL((k_update_xr<DIM, MASK, kUnroll, SPLIT>), dim3(gr), dim3(kB), 0, st, s.ld[0], s.x, s.lv[0].r,
                       (const float*)s.p, R);
  else
    hipLaunchKernelGGL((k_update_xr<DIM, MASK, 1, SPLIT>), dim3(gr), dim3(kB), 0, st, s.ld[0], s.x, s.lv[0].r,
                       (const float*)s.p, R);
}

template <int DIM, bool MASK>
void launch_update_xr(Solver& s, hipStream_t st, unsigned gr, const Red& R) {
  if constexpr (!MASK) {
    if (DIM == 3 && split_level(s, 0) && row_ok(s, s.ld[0])) {
      hipLaunchKernelGGL(k_row_update_xr, dim3(gridr(row_threads(s.ld[0]))), dim3(kB), 0, st, s.ld[0], s.x,
                         s.lv[0].r, (const float*)s.p, R);
      return;
    }
    if (split_level(s, 0)) return launch_update_xr_t<DIM, false, true>(s, st, gr, R);
  }
  launch_update_xr_t<DIM, MASK, false>(s, st, gr, R);
}

template <int DIM, bool MASK>
void launch_Ap_dot(Solver& s, hipStream_t st, unsigned gr, const Red& R) {
  if constexpr (!MASK) {
    if (DIM == 3 && split_level(s, 0) && row_ok(s, s.ld[0])) {
      hipLaunchKernelGGL(k_row_Ap_dot, dim3(gridr(row_threads(s.ld[0]))), dim3(kB), 0, st, s.ld[0],
                         (const float*)s.p, R);
      return;
    }
    if (split_level(s, 0)) return launch_Ap_dot_t<DIM, false, true>(s, st, gr, R);
  }
  launch_Ap_dot_t<DIM, MASK, false>(s, st, gr, R);
}

template <int DIM, bool MASK, bool SPLIT>
void launch_true_resid_t(Solver& s, hipStream_t st) {
  const unsigned gr = gridr(s.lv[0].size);
  if (long_walk(s.lv[0].size))
    hipLaunchKernelGGL((k_true_resid<DIM, MASK, kUnroll, SPLIT>), dim3(gr), dim3(kB), 0, st, s.ld[0],
                       (const double*)s.x, (const float*)s.b, s.lv[0].r, red_of(s), 1);
  else
    hipLaunchKernelGGL((k_true_resid<DIM, MASK, 1, SPLIT>), dim3(gr), dim3(kB), 0, st, s.ld[0], (const double*)s.x,
                       (const float*)s.b, s.lv[0].r, red_of(s), 1);
}

template <int DIM, bool MASK>
void launch_true_resid(Solver& s, hipStream_t st) {
  if constexpr (!MASK) {
    if (split_level(s, 0)) return launch_true_resid_t<DIM, false, true>(s, st);
  }
  launch_true_resid_t<DIM, MASK, false>(s, st);
}

// Zero-fills n 32-bit words; a kernel rather than a memset, so that a caller
// capturing the first solve into a graph records kernel nodes only.
__global__ void __launch_bounds__(kB) k_zero_words(uint32_t* __restrict__ p, uint64_t n) {
  for (uint64_t id = (uint64_t)blockIdx.x * kB + threadIdx.x; id < n; id += (uint64_t)gridDim.x * kB) p[id] = 0u;
}

// The PCG iteration (solve(), 2D/vortex/mgpcg.py:235-278) is queued rotated
// so that it ends at the convergence test: the head -- Ap and p.Ap -> alpha;
// x/r update and r.r -> convergence flag and mean -- follows the set-up (or a
// restart) once z, p and z.r are set; every later iteration is the
// recentring V-cycle with z.r -> beta, the p update, then the head.  A solve
// that converges in the k-th head therefore queues nothing after it: with
// the V-cycle last, the converging iteration's V-cycle and p update (~25
// launches) ran as no-ops.  Every kernel is a no-op once `done` is set, so
// iterations may still be queued past convergence.
template <int DIM, bool MASK>
void pcg_head(Solver& s, hipStream_t st) {
  const unsigned gr = gridr(s.lv[0].size);
  const Red R = red_of(s);
  launch_Ap_dot<DIM, MASK>(s, st, gr, R);
  launch_update_xr<DIM, MASK>(s, st, gr, R);
}

template <int DIM, bool MASK>
void pcg_iteration(Solver& s, hipStream_t st) {
  Lvl& L0 = s.lv[0];
  int64_t N = L0.size;
  unsigned gr = gridr(N);
  vcycle<DIM, MASK>(s, st, s.sc + S_MEAN, ZrFuse{FIN_BETA, nullptr});
  hipLaunchKernelGGL(k_update_p, dim3(gr), dim3(kB), 0, st, N, (const float*)s.z0_final, s.p, (const double*)s.sc,
                     (const int*)s.flags);
  pcg_head<DIM, MASK>(s, st);
}

void drop_graph(Solver& s) {
  if (s.gexec) (void)hipGraphExecDestroy(s.gexec);
  s.gexec = nullptr;
  s.gexec_iters = 0;
}

// Level 0 is stored colour-split (Dims::hh) unless the option is off, an
// obstacle mask is set, level 0 sits in the LDS tail, or its innermost
// extent is odd.  Decided at the start of each solve; a change drops the
// captured iterations, whose kernels carry the layout.
bool split_eligible(const Solver& s) {
  const int inner = s.G.dim == 3 ? s.lv[0].n[2] : s.lv[0].n[1];
  return s.split_opt && !s.masked && s.tail_lt > 0 && inner % 2 == 0;
}

// Mixed-precision storage: with DFM_MG_BF16_COARSE the levels >= 2 are kept
// in BF16 where the kernels that touch them take that storage: the LDS tail
// (always) and, when every global level between level 2 and the tail runs the
// LDS-tiled pre/post kernels, those levels too (the restriction into the
// first BF16 level and the prolongation out of it convert).  Otherwise only
// the tail's levels are BF16.  No obstacle mask (its levels stay FP32).
int bf16_first_level(const Solver& s) {
  if (!s.bf16_coarse || s.masked || s.tail_lt >= s.L || s.tail_lt < 2) return s.L;
  for (int l = 2; l < s.tail_lt; ++l)
    if (!tiled_level(s, l)) return s.tail_lt;
  return 2;
}

void set_layout(Solver& s) {
  const int lo = bf16_first_level(s);
  if (s.bf16_lo != lo) {
    s.bf16_lo = lo;
    drop_graph(s);
  }
  const uint32_t hh = split_eligible(s) ? (uint32_t)(s.lv[0].size / 2) : 0u;
  if (s.ld[0].hh != hh) {
    s.ld[0].hh = hh;
    drop_graph(s);
  }
}

// Capture check_every PCG iterations into one executable graph (on a private
// capture stream, default priority); replayed on the caller's stream.  The
// graph holds kernel launches only -- every pointer it uses is owned by the
// handle and stable -- so one capture serves every later solve.
template <int DIM, bool MASK>
bool build_graph(Solver& s) {
  drop_graph(s);
  if (!s.cap && hipStreamCreateWithFlags(&s.cap, hipStreamNonBlocking) != hipSuccess) {
    s.cap = nullptr;
    return false;
  }
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(s.cap, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
  for (int q = 0; q < s.check_every; ++q) pcg_iteration<DIM, MASK>(s, s.cap);
  if (hipStreamEndCapture(s.cap, &g) != hipSuccess || !g) {
    (void)hipGetLastError();
    return false;
  }
  bool ok = hipGraphInstantiate(&s.gexec, g, nullptr, nullptr, 0) == hipSuccess;
  (void)hipGraphDestroy(g);
  if (!ok) {
    s.gexec = nullptr;
    (void)hipGetLastError();
    return false;
  }
  s.gexec_iters = s.check_every;
  return true;
}

// Queue k PCG iterations (with `head`, the first is the head alone: the
// set-up or a restart has just formed z, p and z.r; then graph replays and
// single iterations), the FP64 true-residual check that runs only if they
// converged, and an asynchronous copy of the flag block to pinned host memory.
template <int DIM, bool MASK>
int pcg_queue(Solver& s, int k, bool head) {
  Pending& P = s.pend;
  int left = k;
  if (head && left > 0) {
    pcg_head<DIM, MASK>(s, P.st);
    --left;
  }
  while (P.use_graph && left >= s.gexec_iters) {
    if (hipGraphLaunch(s.gexec, P.st) != hipSuccess) return set_error(DFM_E_HIP, "mgpcg: graph launch failed");
    left -= s.gexec_iters;
  }
  for (int q = 0; q < left; ++q) pcg_iteration<DIM, MASK>(s, P.st);
  P.it += k;
  // Converged on the recursive FP32 residual: measure the true residual
  // b - A x in FP64; if it misses the tolerance, PCG restarts from it.
  if (P.refinements < s.max_refine) launch_true_resid<DIM, MASK>(s, P.st);
  if (hipMemcpyAsync(s.h_sc, s.sc, kCtlBytes, hipMemcpyDeviceToHost, P.st) != hipSuccess)
    return set_error(DFM_E_HIP, "mgpcg: flag readback failed");
  return DFM_OK;
}

// PCG from r0 (already in lv[0].r, its mean in sc[S_MEAN]): solve(),
// 2D/vortex/mgpcg.py:202-282, with FP64 iterative refinement.  Split in two
// so that a caller can queue several independent solves on different streams
// before waiting on any: pcg_begin queues the set-up and as many iterations
// as the previous solve needed (consecutive projections converge in nearly
// the same number, and iterations queued past convergence are no-ops);
// pcg_end waits for the flag block and continues until convergence.
template <int DIM, bool MASK>
int pcg_begin(Solver& s, hipStream_t st, double tol, int relative, int max_iters) {
  Lvl& L0 = s.lv[0];
  int64_t N = L0.size;
  unsigned gr = gridr(N);
  hipLaunchKernelGGL((k_recenter_norm<MASK>), dim3(gr), dim3(kB), 0, st, N, L0.r, s.b, (const uint8_t*)s.bm[0],
                     red_of(s, tol, relative));
  vcycle<DIM, MASK>(s, st, nullptr, ZrFuse{FIN_ZR_INIT, s.p});
  Pending& P = s.pend;
  P.st = st;
  P.it = 0;
  P.refinements = 0;
  P.max_iters = max_iters;
  P.use_graph = s.use_graph;
  if (P.use_graph && (!s.gexec || s.gexec_iters != s.check_every)) P.use_graph = build_graph<DIM, MASK>(s);
  int k = s.predict > 0 ? s.predict : s.check_every;
  if (k > max_iters) k = max_iters;
  P.active = 1;
  return pcg_queue<DIM, MASK>(s, k, true);
}

template <int DIM, bool MASK>
int pcg_end(Solver& s) {
  Pending& P = s.pend;
  P.active = 0;
  hipStream_t st = P.st;
  bool done = false;
  for (;;) {
    if (spin_sync(st, s.rdy) != hipSuccess) return set_error(DFM_E_HIP, "mgpcg: flag readback failed");
    if (s.h_flags[F_DONE]) {
      if (s.h_flags[F_STATUS] == 1 && P.refinements == 0) s.predict = s.h_flags[F_ITERS];
      done = true;
      break;
    }
    // the iteration budget counts the iterations the device performed, not
    // those queued: queued ones past a convergence that the true-residual
    // check then rejects were no-ops, and counting them would make a solve
    // that ends at max_iters depend on how many the host queued at a time
    P.it = s.h_flags[F_ITERS];
    bool restart = false;
    if (s.h_flags[F_NTRUE] > P.refinements) {
      ++P.refinements;
      vcycle<DIM, MASK>(s, st, s.sc + S_MEAN, ZrFuse{FIN_ZR_INIT, s.p});
      restart = true;
    }
    if (P.it >= P.max_iters) break;
    int k = s.check_every;
    if (P.it + k > P.max_iters) k = P.max_iters - P.it;
    int rc = pcg_queue<DIM, MASK>(s, k, restart);
    if (rc != DFM_OK) return rc;
  }
  // the scalars came back with the flags of the last synchronised batch; only
  // a solve that stopped at max_iters after queueing a restart reads again
  if (!done && (hipMemcpyAsync(s.h_sc, s.sc, kCtlBytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
                spin_sync(st, s.rdy) != hipSuccess))
    return set_error(DFM_E_HIP, "mgpcg: status readback failed");
  // a solve that ended before its first x/r update (b = 0 after recentring)
  // never wrote x: its solution is x = 0
  if (s.h_flags[F_ITERS] == 0) {
    const uint64_t words = 2ull * (uint64_t)s.lv[0].size, blocks = (words + kB - 1) / kB;
    hipLaunchKernelGGL(k_zero_words, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(kB), 0, st,
                       reinterpret_cast<uint32_t*>(s.x), words);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(DFM_E_HIP, "mgpcg launch", hipGetErrorString(e));
  s.last_iters = s.h_flags[F_ITERS];
  double bb = s.h_sc[S_BB];
  s.last_res = bb > 0 ? std::sqrt(s.h_sc[S_RR] / bb) : 0.0;
  if (s.h_flags[F_STATUS] == 2) return set_error(DFM_E_NONFINITE, "mgpcg: residual is not finite");
  if (!s.h_flags[F_DONE]) return set_error(DFM_E_NOCONV, "mgpcg: max_iters reached without convergence");
  return DFM_OK;
}

template <int DIM, bool MASK>
int pcg(Solver& s, hipStream_t st, double tol, int relative, int max_iters) {
  int rc = pcg_begin<DIM, MASK>(s, st, tol, relative, max_iters);
  if (rc != DFM_OK) {
    s.pend.active = 0;
    return rc;
  }
  return pcg_end<DIM, MASK>(s);
}


// The first call on a handle that launches work zeroes the reduction
// partials and counters, the control block and x, ordered on the caller's
// stream.  Creation itself only allocates: it submits nothing and waits for
// nothing, so it is safe while other threads capture graphs or keep their
// streams busy.
int ensure_ready(Solver& s, hipStream_t st) {
  if (s.ready) return DFM_OK;
  static_assert(kCtlBytes % 4 == 0, "control block is a whole number of words");
  const struct {
    void* p;
    uint64_t words;
  } bufs[4] = {{s.part, 2ull * kRedBlocks * 2}, {s.ctr, kCtrWords}, {s.sc, kCtlBytes / 4},
               {s.x, 2ull * (uint64_t)s.lv[0].size}};
  for (const auto& b : bufs) {
    const uint64_t blocks = (b.words + kB - 1) / kB;
    hipLaunchKernelGGL(k_zero_words, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(kB), 0, st,
                       static_cast<uint32_t*>(b.p), b.words);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(DFM_E_HIP, "mgpcg: zeroing the handle's buffers", hipGetErrorString(e));
  s.ready = 1;
  return DFM_OK;
}

void destroy(Solver* s) {
  if (!s) return;
  for (int l = 0; l < s->L; ++l) {
    if (s->lv[l].r) (void)hipFree(s->lv[l].r);
    if (s->lv[l].z) (void)hipFree(s->lv[l].z);
    if (s->lv[l].z2) (void)hipFree(s->lv[l].z2);
  }
  if (s->x) (void)hipFree(s->x);
  if (s->xn) (void)hipFree(s->xn);
  if (s->b) (void)hipFree(s->b);
  if (s->p) (void)hipFree(s->p);
  if (s->part) (void)hipFree(s->part);
  if (s->ctr) (void)hipFree(s->ctr);
  for (int l = 0; l < kMaxLevels; ++l)
    if (s->bm[l]) (void)hipFree(s->bm[l]);
  if (s->sc) (void)hipFree(s->sc);  // flags live in the same block
  if (s->h_sc) (void)hipHostFree(s->h_sc);
  if (s->rdy) (void)hipEventDestroy(s->rdy);
  drop_graph(*s);
  if (s->cap) (void)hipStreamDestroy(s->cap);
  delete s;
}

}  // namespace

extern "C" int dfm_mgpcg_create(const dfm_grid* g, int levels, int pre_post, int bottom, void** handle) {
  if (!g || !handle) return set_error(DFM_E_NULL, "dfm_mgpcg_create: NULL argument");
  *handle = nullptr;
  if (g->dim != 2 && g->dim != 3) return set_error(DFM_E_ARG, "dim must be 2 or 3");
  if (pre_post < 1 || bottom < 1) return set_error(DFM_E_ARG, "pre_post and bottom must be >= 1");
  int nmin = g->n[0];
  for (int a = 1; a < g->dim; ++a) nmin = g->n[a] < nmin ? g->n[a] : nmin;
  if (nmin < 2) return set_error(DFM_E_ARG, "grid too small");
  if ((int64_t)(g->n[0] + 1) * (g->n[1] + 1) * (g->n[2] + 1) >= (int64_t(1) << 31))
    return set_error(DFM_E_ARG, "grid too large (faces must stay below 2^31 per axis)");
  if (levels <= 0) {
    int lg = 0;
    while ((1 << (lg + 1)) <= nmin) ++lg;
    levels = lg - 3 + 1;
    if (levels < 1) levels = 1;
    // no more levels than the grid halves exactly into (e.g. 100 = 4 x 25: 3)
    for (int a = 0; a < g->dim; ++a)
      while (levels > 1 && g->n[a] % (1 << (levels - 1)) != 0) --levels;
  }
  if (levels > kMaxLevels) return set_error(DFM_E_ARG, "too many levels");
  for (int a = 0; a < g->dim; ++a)
    if (g->n[a] % (1 << (levels - 1)) != 0) return set_error(DFM_E_ARG, "n must be divisible by 2^(levels-1)");
  if (g->dim == 2 && g->n[2] != 1) return set_error(DFM_E_ARG, "2D grid must have n[2] == 1");
  Solver* s = new (std::nothrow) Solver();
  if (!s) return set_error(DFM_E_HIP, "out of host memory");
  std::memset(s, 0, sizeof(Solver));
  s->G.dim = g->dim;
  for (int a = 0; a < 3; ++a) s->G.n[a] = g->n[a];
  s->G.dx = g->dx;
  s->G.inv_dx = 1.0f / g->dx;
  s->L = levels;
  s->bf16_lo = levels;
  s->pre = pre_post;
  s->bottom = bottom;
  s->check_every = 1;
  s->max_refine = 3;
  s->use_graph = 1;
  s->fused_sweeps = 1;
  s->row_opt = 1;
  s->tiled_sweeps = 1;
  s->split_opt = 1;
  // Allocation only (ensure_ready zeroes on the first call that launches
  // work); the first failing call is named in the error.
  const char* failed = nullptr;
  hipError_t err = hipSuccess;
  auto need = [&](hipError_t e, const char* what) {
    if (!failed && e != hipSuccess) {
      failed = what;
      err = e;
    }
    return failed == nullptr;
  };
  for (int l = 0; l < levels && !failed; ++l) {
    for (int a = 0; a < 3; ++a) s->lv[l].n[a] = a < g->dim ? (g->n[a] >> l) : 1;
    s->lv[l].size = (int64_t)s->lv[l].n[0] * s->lv[l].n[1] * s->lv[l].n[2];
    s->ld[l] = make_dims(s->lv[l].n[0], s->lv[l].n[1], s->lv[l].n[2], g->dim);
    need(hipMalloc(&s->lv[l].r, sizeof(float) * s->lv[l].size), "hipMalloc (level residual)") &&
        need(hipMalloc(&s->lv[l].z, sizeof(float) * s->lv[l].size), "hipMalloc (level correction)");
    if (!failed && l > 0 && s->lv[l].size <= kSweep2MaxCells)
      need(hipMalloc(&s->lv[l].z2, sizeof(float) * s->lv[l].size), "hipMalloc (level correction copy)");
  }
  int64_t N = s->lv[0].size;
  need(hipMalloc(&s->x, sizeof(double) * N), "hipMalloc (x)") && need(hipMalloc(&s->b, sizeof(float) * N), "hipMalloc (b)") &&
      need(hipMalloc(&s->p, sizeof(float) * N), "hipMalloc (p)") &&
      need(hipMalloc(&s->part, sizeof(double) * 2 * kRedBlocks), "hipMalloc (partials)") &&
      need(hipMalloc(&s->ctr, sizeof(unsigned) * kCtrWords), "hipMalloc (counters)") &&
      need(hipMalloc(&s->sc, kCtlBytes), "hipMalloc (control block)") &&
      need(hipHostMalloc(&s->h_sc, kCtlBytes, hipHostMallocDefault), "hipHostMalloc (control block)") &&
      need(hipEventCreateWithFlags(&s->rdy, hipEventDisableTiming), "hipEventCreate (readback)");
  if (failed) {
    char detail[160];
    std::snprintf(detail, sizeof(detail), "%s: %s", failed, hipGetErrorString(err));
    destroy(s);
    return set_error(DFM_E_HIP, "dfm_mgpcg_create: device allocation failed", detail);
  }
  s->flags = reinterpret_cast<int*>(reinterpret_cast<char*>(s->sc) + sizeof(double) * S_NSLOT);
  s->h_flags = reinterpret_cast<int*>(reinterpret_cast<char*>(s->h_sc) + sizeof(double) * S_NSLOT);
  setup_tail(*s);
  *handle = s;
  return DFM_OK;
}

extern "C" int dfm_mgpcg_destroy(void* handle) {
  destroy(static_cast<Solver*>(handle));
  return DFM_OK;
}

extern "C" int dfm_mgpcg_set_option(void* handle, int option, int value) {
  Solver* s = static_cast<Solver*>(handle);
  if (!s) return set_error(DFM_E_NULL, "mgpcg handle is NULL");
  if (s->pend.active) return set_error(DFM_E_ARG, "mgpcg: a begun projection has not been ended on this handle");
  switch (option) {
    case DFM_MG_BF16_COARSE:
      s->bf16_coarse = value ? 1 : 0;
      setup_tail(*s);
      drop_graph(*s);
      return DFM_OK;
    case DFM_MG_TAIL:
      if (value) setup_tail(*s);
      else s->tail_lt = s->L;
      drop_graph(*s);
      return DFM_OK;
    case DFM_MG_GRAPH:
      s->use_graph = value ? 1 : 0;
      drop_graph(*s);
      return DFM_OK;
    case DFM_MG_TAIL_LEVEL:
      if (value < 0) return set_error(DFM_E_ARG, "tail level must be >= 0");
      s->tail_min = value;
      setup_tail(*s);
      drop_graph(*s);
      return DFM_OK;
    case DFM_MG_BOTTOM_EXACT:
      s->bottom_exact = value ? 1 : 0;
      setup_tail(*s);
      drop_graph(*s);
      return DFM_OK;
    case DFM_MG_MAX_REFINE:
      if (value < 0) return set_error(DFM_E_ARG, "max_refine must be >= 0");
      s->max_refine = value;
      return DFM_OK;
    case DFM_MG_CHECK_EVERY:
      if (value < 1) return set_error(DFM_E_ARG, "check_every must be >= 1");
      s->check_every = value;
      drop_graph(*s);
      return DFM_OK;
    case DFM_MG_FUSED_SWEEPS:
      // 0: two launches per sweep; 1: one launch per sweep; 2 (default): LDS-tiled
      // pre/post smoothing where eligible, else one launch per sweep
      if (value < 0 || value > 2) return set_error(DFM_E_ARG, "fused sweeps option must be 0, 1 or 2");
      s->fused_sweeps = value ? 1 : 0;
      s->tiled_sweeps = value == 2 ? 1 : 0;
      drop_graph(*s);
      return DFM_OK;
    case DFM_MG_SPLIT_LAYOUT:
      s->split_opt = value ? 1 : 0;
      drop_graph(*s);
      return DFM_OK;
    case DFM_MG_ROW_KERNELS:
      if (value < 0 || value > 2) return set_error(DFM_E_ARG, "row kernels option must be 0, 1 or 2");
      s->row_opt = value;
      drop_graph(*s);
      return DFM_OK;
    default:
      return set_error(DFM_E_ARG, "unknown MGPCG option");
  }
}

extern "C" int dfm_mgpcg_get_option(void* handle, int option, int* value) {
  Solver* s = static_cast<Solver*>(handle);
  if (!s || !value) return set_error(DFM_E_NULL, "dfm_mgpcg_get_option: NULL argument");
  switch (option) {
    case DFM_MG_BF16_COARSE: *value = s->tail_bf16; return DFM_OK;
    case DFM_MG_TAIL: *value = s->tail_lt; return DFM_OK;
    case DFM_MG_MAX_REFINE: *value = s->max_refine; return DFM_OK;
    case DFM_MG_CHECK_EVERY: *value = s->check_every; return DFM_OK;
    case DFM_MG_GRAPH: *value = s->gexec ? 2 : s->use_graph; return DFM_OK;
    case DFM_MG_BOTTOM_EXACT: *value = (s->tail_lt < s->L) ? s->bottom_exact : 0; return DFM_OK;
    case DFM_MG_TAIL_LEVEL: *value = s->tail_lt; return DFM_OK;
    case DFM_MG_FUSED_SWEEPS: *value = s->fused_sweeps + s->tiled_sweeps; return DFM_OK;
    case DFM_MG_SPLIT_LAYOUT: *value = split_eligible(*s) ? 1 : 0; return DFM_OK;
    case DFM_MG_BF16_LEVEL: *value = bf16_first_level(*s); return DFM_OK;
    case DFM_MG_LEVELS: *value = s->L; return DFM_OK;
    case DFM_MG_ROW_KERNELS: *value = s->row_opt; return DFM_OK;
    default: return set_error(DFM_E_ARG, "unknown MGPCG option");
  }
}

extern "C" int dfm_mgpcg_project_begin(void* handle, const dfm_mac* u, double tol, int relative, int max_iters,
                                       void* stream) {
  Solver* s = static_cast<Solver*>(handle);
  if (!s) return set_error(DFM_E_NULL, "mgpcg handle is NULL");
  if (!u) return set_error(DFM_E_NULL, "velocity is NULL");
  for (int a = 0; a < s->G.dim; ++a)
    if (!u->c[a]) return set_error(DFM_E_NULL, "velocity component is NULL");
  if (max_iters < 1) return set_error(DFM_E_ARG, "max_iters must be >= 1");
  if (s->pend.active) return set_error(DFM_E_ARG, "mgpcg: a begun projection has not been ended on this handle");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (int rc0 = ensure_ready(*s, st)) return rc0;
  set_layout(*s);
  s->last_st = st;
  Lvl& L0 = s->lv[0];
  const Dims& d0 = s->ld[0];
  Faces F{{u->c[0], u->c[1], u->c[2]}};
  unsigned gr = gridr(L0.size);
  int rc;
  const Red R = red_of(*s);
  if (s->G.dim == 2) {
    if (s->masked) {
      hipLaunchKernelGGL((k_div<2, true>), dim3(gr), dim3(kB), 0, st, d0, F, L0.r, R);
      rc = pcg_begin<2, true>(*s, st, tol, relative, max_iters);
    } else {
      hipLaunchKernelGGL((k_div<2, false>), dim3(gr), dim3(kB), 0, st, d0, F, L0.r, R);
      rc = pcg_begin<2, false>(*s, st, tol, relative, max_iters);
    }
  } else {
    if (s->masked) {
      hipLaunchKernelGGL((k_div<3, true>), dim3(gr), dim3(kB), 0, st, d0, F, L0.r, R);
      rc = pcg_begin<3, true>(*s, st, tol, relative, max_iters);
    } else {
      hipLaunchKernelGGL((k_div<3, false>), dim3(gr), dim3(kB), 0, st, d0, F, L0.r, R);
      rc = pcg_begin<3, false>(*s, st, tol, relative, max_iters);
    }
  }
  s->pend.F = F;
  if (rc != DFM_OK) s->pend.active = 0;
  return rc;
}

extern "C" int dfm_mgpcg_project_end(void* handle) {
  Solver* s = static_cast<Solver*>(handle);
  if (!s) return set_error(DFM_E_NULL, "mgpcg handle is NULL");
  if (!s->pend.active) return set_error(DFM_E_ARG, "mgpcg: no begun projection on this handle");
  const bool m = s->masked;
  int rc = s->G.dim == 2 ? (m ? pcg_end<2, true>(*s) : pcg_end<2, false>(*s))
                         : (m ? pcg_end<3, true>(*s) : pcg_end<3, false>(*s));
  if (rc != DFM_OK && rc != DFM_E_NOCONV) return rc;
  // subtract the pressure gradient even when max_iters was hit (the reference
  // prints and continues, 2D/vortex/mgpcg.py:280-282)
  hipStream_t st = s->pend.st;
  Faces F = s->pend.F;
  Lvl& L0 = s->lv[0];
  const Dims& d0 = s->ld[0];
  const unsigned gc = grid1(L0.size);  // one thread per cell
  const double* px = s->x;
  if (s->G.dim == 2) {
    if (m) hipLaunchKernelGGL((k_subgrad<2, true>), dim3(gc), dim3(kB), 0, st, d0, F, px);
    else hipLaunchKernelGGL((k_subgrad<2, false>), dim3(gc), dim3(kB), 0, st, d0, F, px);
  } else {
    if (m) hipLaunchKernelGGL((k_subgrad<3, true>), dim3(gc), dim3(kB), 0, st, d0, F, px);
    else hipLaunchKernelGGL((k_subgrad<3, false>), dim3(gc), dim3(kB), 0, st, d0, F, px);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(DFM_E_HIP, "mgpcg subtract_grad_p", hipGetErrorString(e));
  return rc;
}

extern "C" int dfm_mgpcg_project(void* handle, const dfm_mac* u, double tol, int relative, int max_iters,
                                 void* stream) {
  int rc = dfm_mgpcg_project_begin(handle, u, tol, relative, max_iters, stream);
  if (rc != DFM_OK) return rc;
  return dfm_mgpcg_project_end(handle);
}

extern "C" int dfm_mgpcg_solve(void* handle, const float* b, float* x, double tol, int relative, int max_iters,
                               void* stream) {
  Solver* s = static_cast<Solver*>(handle);
  if (!s) return set_error(DFM_E_NULL, "mgpcg handle is NULL");
  if (!b || !x) return set_error(DFM_E_NULL, "b/x is NULL");
  if (max_iters < 1) return set_error(DFM_E_ARG, "max_iters must be >= 1");
  if (s->pend.active) return set_error(DFM_E_ARG, "mgpcg: a begun projection has not been ended on this handle");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (int rc0 = ensure_ready(*s, st)) return rc0;
  set_layout(*s);
  s->last_st = st;
  Lvl& L0 = s->lv[0];
  unsigned gr = gridr(L0.size);
  int rc;
  if (s->masked) {
    hipLaunchKernelGGL((k_load_b<true>), dim3(gr), dim3(kB), 0, st, s->ld[0], L0.size, b, L0.r,
                       (const uint8_t*)s->bm[0], red_of(*s));
    rc = s->G.dim == 2 ? pcg<2, true>(*s, st, tol, relative, max_iters) : pcg<3, true>(*s, st, tol, relative, max_iters);
  } else {
    hipLaunchKernelGGL((k_load_b<false>), dim3(gr), dim3(kB), 0, st, s->ld[0], L0.size, b, L0.r,
                       (const uint8_t*)nullptr, red_of(*s));
    rc = s->G.dim == 2 ? pcg<2, false>(*s, st, tol, relative, max_iters)
                       : pcg<3, false>(*s, st, tol, relative, max_iters);
  }
  if (rc != DFM_OK && rc != DFM_E_NOCONV) return rc;
  hipLaunchKernelGGL(k_to_float, dim3(gr), dim3(kB), 0, st, s->ld[0], L0.size, s->x, x);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(DFM_E_HIP, "mgpcg: solution copy", hipGetErrorString(e));
  return rc;
}

extern "C" int dfm_mgpcg_set_mask(void* handle, const float* mask, void* stream) {
  Solver* s = static_cast<Solver*>(handle);
  if (!s) return set_error(DFM_E_NULL, "mgpcg handle is NULL");
  if (s->pend.active) return set_error(DFM_E_ARG, "mgpcg: a begun projection has not been ended on this handle");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  drop_graph(*s);
  if (!mask) {
    s->masked = 0;
    for (int l = 0; l < s->L; ++l) s->ld[l].bm = nullptr;
    setup_tail(*s);
    return DFM_OK;
  }
  for (int l = 0; l < s->L; ++l)
    if (!s->bm[l] && hipMalloc(&s->bm[l], (size_t)s->lv[l].size) != hipSuccess)
      return set_error(DFM_E_HIP, "dfm_mgpcg_set_mask: device allocation failed");
  if (int rc0 = ensure_ready(*s, st)) return rc0;
  Lvl& L0 = s->lv[0];
  Red R = red_of(*s);
  hipLaunchKernelGGL(k_mask_level0, dim3(gridr(L0.size)), dim3(kB), 0, st, L0.size, mask, s->bm[0], R);
  for (int l = 0; l + 1 < s->L; ++l) {
    Dims d = s->ld[l], dc = s->ld[l + 1];
    if (s->G.dim == 2)
      hipLaunchKernelGGL((k_mask_coarsen<2>), dim3(grid1(s->lv[l + 1].size)), dim3(kB), 0, st, d,
                         (const uint8_t*)s->bm[l], dc, s->bm[l + 1]);
    else
      hipLaunchKernelGGL((k_mask_coarsen<3>), dim3(grid1(s->lv[l + 1].size)), dim3(kB), 0, st, d,
                         (const uint8_t*)s->bm[l], dc, s->bm[l + 1]);
  }
  double nf = 0.0;
  if (hipMemcpyAsync(&nf, s->sc + S_NFLUID, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return set_error(DFM_E_HIP, "dfm_mgpcg_set_mask: fluid count readback failed");
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(DFM_E_HIP, "dfm_mgpcg_set_mask", hipGetErrorString(e));
  s->masked = 1;
  s->nfluid = nf;
  for (int l = 0; l < s->L; ++l) s->ld[l].bm = s->bm[l];
  setup_tail(*s);
  s->predict = 0;
  return DFM_OK;
}

extern "C" int dfm_mgpcg_status(void* handle, int* iters, double* rel_residual) {
  Solver* s = static_cast<Solver*>(handle);
  if (!s) return set_error(DFM_E_NULL, "mgpcg handle is NULL");
  if (iters) *iters = s->last_iters;
  if (rel_residual) *rel_residual = s->last_res;
  return DFM_OK;
}

extern "C" int dfm_mgpcg_pressure(void* handle, double** p) {
  Solver* s = static_cast<Solver*>(handle);
  if (!s || !p) return set_error(DFM_E_NULL, "dfm_mgpcg_pressure: NULL argument");
  if (!s->ready) return set_error(DFM_E_ARG, "dfm_mgpcg_pressure: no solve has run on this handle");
  if (!s->ld[0].hh) {
    *p = s->x;
    return DFM_OK;
  }
  // colour-split level 0: reorder into natural order behind the last solve
  const int64_t N = s->lv[0].size;
  if (!s->xn && hipMalloc(&s->xn, sizeof(double) * N) != hipSuccess) {
    s->xn = nullptr;
    return set_error(DFM_E_HIP, "dfm_mgpcg_pressure: device allocation failed");
  }
  hipLaunchKernelGGL(k_unsplit, dim3(gridr(N)), dim3(kB), 0, s->last_st, s->ld[0], N, (const double*)s->x, s->xn);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(DFM_E_HIP, "dfm_mgpcg_pressure: reorder", hipGetErrorString(e));
  *p = s->xn;
  return DFM_OK;
}
