// Adjoint source terms at the face nodes of a 3D MAC grid, LDS-tiled.
//
//   r_a(X) = sum_m (d/dx_a fd_m)(X) * fv_m(X)   at every node X of face grid a,
//
// both fields of pair m sampled with the B-spline of order K on their own
// lattice: face grid c for the velocity pairs (u_c, w_c) of
// calculate_nabla_u_w (3D/flowmap.py:359-380), the cell grid for the scalar
// pairs (s, lambda_s) of calculate_nabla_scalar_adjoint (3D/flowmap.py:382-403).
//
// X is a node of a fixed grid, so along each axis a sample's taps and weights
// depend only on that axis' output index: they are the Axis<K>::setup values
// of the gather kernels (coordinate clamp and tap clamp included), and a
// sample is the tensor product of three 1D stencils.  One workgroup owns a
// TX x TY x TZ tile of outputs.  Per field it loads the lattice window its
// taps cover into LDS once, contracts z, then y, then x (the order of the
// gather kernels' dual_interp, so every result is bitwise theirs) and keeps
// the product in registers.  Global traffic is each field's window once per
// tile (the halo re-reads mostly hit L2) instead of 64 gathered taps per
// sample and field.
#pragma once
#include "dfm_common.hpp"

namespace dfm {
namespace tiled {

constexpr int TX = 8, TY = 8, TZ = 32;
constexpr int THREADS = TY * TZ;  // the x pass gives each thread one (y, z) column of the tile
constexpr int kMaxPairs = 8;
constexpr int kMaxPos = 32;       // max(TX, TY, TZ)
static_assert(TX <= kMaxPos && TY <= kMaxPos && TZ <= kMaxPos, "tile edge");

struct Pairs {
  int n;
  int slot[kMaxPairs];  // table slot of the pair's lattice
  int nslot;
  int slot_lattice[3];  // face axis 0..2 or 3 = cells
  const float* fd[kMaxPairs];
  const float* fv[kMaxPairs];
};

struct Tiles {
  int nt[3][3];  // tiles along each axis of output grid a
  int off[4];    // first block of each output grid
  float* out[3];
};

// Workgroup barrier that orders LDS accesses only: __syncthreads() also
// waits for every outstanding global load of the thread, which would end the
// window prefetch at the first barrier.
DFM_HD void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int K>
struct Smem {
  static constexpr int NT = Spline<K>::NT;
  static constexpr int EX = TX + NT, EY = TY + NT, EZ = TZ + NT;  // window bounds (see tables())
  static constexpr int PZ = EZ | 1;                                // odd row pitch
  static_assert(EZ <= 64, "a window row is one wave wide");
  // lattice window; the z pass overwrites each row with its outputs
  __attribute__((aligned(16))) float A[EX * EY * PZ];
  float Y[EX * TY * TZ];  // y-pass output
  static constexpr int NW = (NT + 3) & ~3;  // table rows padded to 16 bytes (vector reads)
  static constexpr int TXY = TX > TY ? TX : TY;
  // [slot][output][tap] value weights along z and along x, y; derivative
  // weights along the output grid's axis; tap indices relative to the window
  __attribute__((aligned(16))) float wvz[3][TZ][NW];
  __attribute__((aligned(16))) float wvxy[3][2][TXY][NW];
  __attribute__((aligned(16))) float wd[3][kMaxPos][NW];
  __attribute__((aligned(16))) uint8_t riz[3][TZ][NW];
  __attribute__((aligned(16))) uint8_t rixy[3][2][TXY][NW];
  int lo[3][3], ext[3][3];
  DFM_HD float& wv(int slot, int b, int pos, int q) { return b == 2 ? wvz[slot][pos][q] : wvxy[slot][b][pos][q]; }
  DFM_HD uint8_t& ri(int slot, int b, int pos, int q) { return b == 2 ? riz[slot][pos][q] : rixy[slot][b][pos][q]; }
  static_assert(EX < 256 && EY < 256 && EZ < 256, "relative tap indices are bytes");
};

// Tables of one (slot, axis, output position).  Along an axis the sample
// coordinate rises by one per output (up to the rounding of the node
// position, < 1/2) or stays clamped, so the first tap rises by at most one
// per output plus one rounding step: a tile of T outputs touches at most
// T + NT lattice nodes.  The window extent is clamped to that bound anyway,
// and every relative index to the window, so no LDS access can leave the
// arrays.
//
// Cubic and linear taps are read as base + q.  The base is never below 0,
// but it can be S - 3 (in FP32 the clamp bound S - 2 - 1e-6 rounds to S - 2
// once S is large), and the fourth tap then lies past the lattice: the
// gather path clamps its index to S - 1, where its weight is exactly 0.  So
// for these splines the window is sized from the unclamped taps and loaded
// with every index clamped to S - 1: window slot r holds f[min(lo + r, S - 1)],
// exactly the value the gather reads for that tap.
template <int K>
__device__ void table_entry(Smem<K>& sm, const Geo& g, int a, int slot, int lat, int b, int pos, const int o0[3],
                            const int cnt[3]) {
  constexpr int NT = Spline<K>::NT;
  const int lim = b == 0 ? TX : (b == 1 ? TY : TZ);
  if (pos >= lim) return;
  const int cap = lim + NT;
  int S[3];
  shape_of(g, lat, S);
  // node position along b of output grid a (node_pos) and the lattice's offset
  const float off = b == a ? 0.0f : 0.5f;
  const float bl = (lat < 3 && b == lat) ? 0.0f : 0.5f;
  auto setup = [&](int p, Axis<K, true>& ax) {
    const int I = o0[b] + (p < cnt[b] ? p : cnt[b] - 1);
    const float pb = ((float)I + off) * g.dx;
    ax.setup(pb * g.inv_dx, bl, S[b]);
  };
  Axis<K, true> first, last, mine;
  setup(0, first);
  setup(cnt[b] - 1, last);
  setup(pos, mine);
  constexpr bool kConsec = K != 4;  // taps base + q (see above)
  const int lo = first.idx[0];
  int ext = kConsec ? last.idx[0] + NT - lo : last.idx[NT - 1] - lo + 1;
  ext = ext < cap ? ext : cap;
#pragma unroll
  for (int q = 0; q < NT; ++q) {
    sm.wv(slot, b, pos, q) = mine.w[q];
    if (b == a) sm.wd[slot][pos][q] = mine.dw[q];
    const int r = (kConsec ? mine.idx[0] + q : mine.idx[q]) - lo;
    sm.ri(slot, b, pos, q) = (uint8_t)(r < 0 ? 0 : (r >= ext ? ext - 1 : r));
  }
  if (pos == 0) {
    sm.lo[slot][b] = lo;
    sm.ext[slot][b] = ext;
  }
}

template <int K>
__global__ void __launch_bounds__(THREADS) k_src_tiled(Geo g, Pairs P, Tiles T) {
  constexpr int NT = Spline<K>::NT;
  using SM = Smem<K>;
  __shared__ SM sm;
  const int bid = (int)xcd_block(blockIdx.x, gridDim.x);
  const int a = bid >= T.off[1] ? (bid >= T.off[2] ? 2 : 1) : 0;
  int t = bid - T.off[a];
  const int tz = t % T.nt[a][2];
  t /= T.nt[a][2];
  const int ty = t % T.nt[a][1];
  const int tx = t / T.nt[a][1];
  int So[3];
  shape_of(g, a, So);
  const int o0[3] = {tx * TX, ty * TY, tz * TZ};
  const int cnt[3] = {min(TX, So[0] - o0[0]), min(TY, So[1] - o0[1]), min(TZ, So[2] - o0[2])};
  const int tid = threadIdx.x;
  for (int e = tid; e < P.nslot * 3 * kMaxPos; e += THREADS) {
    const int slot = e / (3 * kMaxPos), b = (e / kMaxPos) % 3;
    table_entry<K>(sm, g, a, slot, P.slot_lattice[slot], b, e % kMaxPos, o0, cnt);
  }
  // z pass: output z fixed per thread; y and x passes: (y, z) fixed per thread
  const int zo = tid % TZ, yo = tid / TZ;
  float r[TX], d[TX];
#pragma unroll
  for (int q = 0; q < TX; ++q) r[q] = 0.f;

  // Field i of the sequence is fd (i even) or fv (i odd) of pair i / 2.
  // With the cubic (and linear) spline the taps of a sample are consecutive
  // nodes (the coordinate clamp keeps them inside the lattice), so each pass
  // reads them at one base address plus immediate offsets; the quartic
  // spline's taps can repeat at a wall and go through the index tables.
  constexpr bool kConsec = K != 4;
  constexpr int XR = (SM::EX + THREADS / 64 - 1) / (THREADS / 64);  // window x-planes per wave
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  float v[XR][SM::EY];
  // Window of field i into v: wave w loads the x-planes w, w + 4, ...; lanes
  // along z (coalesced).  The loads stay in flight until store() (the
  // barriers between the passes fence LDS only).
  auto issue = [&](int i) {
    const int m = i >> 1, slot = P.slot[m];
    int S[3];
    shape_of(g, P.slot_lattice[slot], S);
    const float* src = (i & 1) ? P.fv[m] : P.fd[m];
    const int ex = __builtin_amdgcn_readfirstlane(sm.ext[slot][0]);
    const int ey = __builtin_amdgcn_readfirstlane(sm.ext[slot][1]);
    const int ez = __builtin_amdgcn_readfirstlane(sm.ext[slot][2]);
    const int lx = __builtin_amdgcn_readfirstlane(sm.lo[slot][0]);
    const int ly = __builtin_amdgcn_readfirstlane(sm.lo[slot][1]);
    const int lz = __builtin_amdgcn_readfirstlane(sm.lo[slot][2]);
    // window indices clamped to the lattice (slot r holds f[min(lo + r, S - 1)])
    const uint32_t S2 = (uint32_t)S[2];
    const uint32_t zl = (uint32_t)min(lz + lane, S[2] - 1);
#pragma unroll
    for (int kx = 0; kx < XR; ++kx) {
      const int x = wave + (THREADS / 64) * kx;
      const uint32_t xg = (uint32_t)min(lx + x, S[0] - 1);
#pragma unroll
      for (int y = 0; y < SM::EY; ++y) {
        const uint32_t yg = (uint32_t)min(ly + y, S[1] - 1);
        v[kx][y] = (x < ex && y < ey && lane < ez) ? *at32(src, (xg * (uint32_t)S[1] + yg) * S2 + zl) : 0.f;
      }
    }
  };
  auto store = [&](int i) {
    const int slot = P.slot[i >> 1];
    const int ex = __builtin_amdgcn_readfirstlane(sm.ext[slot][0]);
    const int ey = __builtin_amdgcn_readfirstlane(sm.ext[slot][1]);
    const int ez = __builtin_amdgcn_readfirstlane(sm.ext[slot][2]);
#pragma unroll
    for (int kx = 0; kx < XR; ++kx) {
      const int x = wave + (THREADS / 64) * kx;
#pragma unroll
      for (int y = 0; y < SM::EY; ++y)
        if (x < ex && y < ey && lane < ez) sm.A[(x * ey + y) * SM::PZ + lane] = v[kx][y];
    }
  };
  __syncthreads();  // tables
  const int nf = 2 * P.n;
  // Three barriers per field: the z pass runs in place in A (each window row
  // is read by the 32 lanes of one wave before they write its outputs), the
  // y pass writes Y, and the next field's window goes into A during the x
  // pass, which reads Y.
  if (nf > 0) {
    issue(0);
    store(0);
    lds_barrier();
    if (nf > 1) issue(1);
  }
#pragma unroll 1
  for (int i = 0; i < nf; ++i) {
    const int slot = P.slot[i >> 1];
    const bool deriv = !(i & 1);
    const int ex = __builtin_amdgcn_readfirstlane(sm.ext[slot][0]);
    const int ey = __builtin_amdgcn_readfirstlane(sm.ext[slot][1]);
    // z pass: A (ex*ey rows) -> A, in place
    {
      float w[NT];
      int rz[NT];
#pragma unroll
      for (int q = 0; q < NT; ++q) {
        w[q] = (deriv && a == 2) ? sm.wd[slot][zo][q] : sm.wvz[slot][zo][q];
        rz[q] = sm.riz[slot][zo][q];
      }
      const float* in = sm.A + yo * SM::PZ + (kConsec ? rz[0] : 0);
      float* out = sm.A + yo * SM::PZ + zo;
#pragma unroll 6
      for (int row = yo; row < ex * ey; row += THREADS / TZ) {
        float acc = 0.f;
#pragma unroll
        for (int q = 0; q < NT; ++q) acc = fmaf(in[kConsec ? q : rz[q]], w[q], acc);
        *out = acc;
        in += (THREADS / TZ) * SM::PZ;
        out += (THREADS / TZ) * SM::PZ;
      }
    }
    lds_barrier();
    // y pass: A -> Y as (x, yo, z), x over the window
    {
      float w[NT];
      int ry[NT];
#pragma unroll
      for (int q = 0; q < NT; ++q) {
        w[q] = (deriv && a == 1) ? sm.wd[slot][yo][q] : sm.wvxy[slot][1][yo][q];
        ry[q] = sm.rixy[slot][1][yo][q];
      }
      const float* in = sm.A + zo + (kConsec ? ry[0] * SM::PZ : 0);
      float* out = sm.Y + yo * TZ + zo;
#pragma unroll 4
      for (int x = 0; x < ex; ++x) {
        float acc = 0.f;
#pragma unroll
        for (int q = 0; q < NT; ++q) acc = fmaf(in[(kConsec ? q : ry[q]) * SM::PZ], w[q], acc);
        *out = acc;
        in += ey * SM::PZ;
        out += TY * TZ;
      }
    }
    lds_barrier();
    if (i + 1 < nf) store(i + 1);  // A is free; the x pass reads Y
    // x pass: Y -> registers, outputs (xo, yo, zo); the position's weights
    // and tap indices are block-uniform
    {
      const float* col = sm.Y + yo * TZ + zo;
#pragma unroll
      for (int xo = 0; xo < TX; ++xo) {
        float w[NT];
        int rx[NT];
        const float* wsrc = (deriv && a == 0) ? sm.wd[slot][xo] : sm.wvxy[slot][0][xo];
#pragma unroll
        for (int q = 0; q < NT; ++q) {
          w[q] = wsrc[q];
          rx[q] = sm.rixy[slot][0][xo][q];
        }
        float acc = 0.f;
        const float* in = col + (kConsec ? rx[0] * (TY * TZ) : 0);
#pragma unroll
        for (int q = 0; q < NT; ++q) acc = fmaf(in[(kConsec ? q : rx[q]) * (TY * TZ)], w[q], acc);
        if (deriv) d[xo] = acc * g.inv_dx;
        else r[xo] = fmaf(d[xo], acc, r[xo]);
      }
    }
    lds_barrier();  // Y is free, A holds the next window
    if (i + 2 < nf) issue(i + 2);
  }
  if (yo < cnt[1] && zo < cnt[2]) {
    float* out = T.out[a];
    const uint32_t S1 = (uint32_t)So[1], S2 = (uint32_t)So[2];
#pragma unroll
    for (int xo = 0; xo < TX; ++xo)
      if (xo < cnt[0]) out[((uint32_t)(o0[0] + xo) * S1 + (uint32_t)(o0[1] + yo)) * S2 + (uint32_t)(o0[2] + zo)] = r[xo];
  }
}

}  // namespace tiled
}  // namespace dfm
