// Shared device helpers for the flow-map kernels (gfx950 / CDNA4).
//
// Layout: every scalar field is a dense (sx, sy, sz) array, z fastest (2D uses
// sz = 1).  The MAC velocity component a lives on a (n + e_a) face grid.
// Flow-map point sets (psi, psi_a, T_a) are SoA: component c of point i is at
// [c * npoints + i], so a wavefront reads each component fully coalesced.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "dfm_fastdiv.hpp"

#define DFM_HD __device__ __forceinline__

namespace dfm {

struct Geo {
  int dim;
  int n[3];  // cells per axis (n[2] = 1 in 2D)
  float dx;
  float inv_dx;
};

// Shape of a face grid (axis a < dim) or of the cell grid (a == 3).
DFM_HD void shape_of(const Geo& g, int a, int s[3]) {
  for (int b = 0; b < 3; ++b) s[b] = g.n[b] + ((b == a && b < g.dim) ? 1 : 0);
}
inline void shape_of_host(const Geo& g, int a, int s[3]) {
  for (int b = 0; b < 3; ++b) s[b] = g.n[b] + ((b == a && b < g.dim) ? 1 : 0);
}
inline int64_t count_of_host(const Geo& g, int a) {
  int s[3];
  shape_of_host(g, a, s);
  return (int64_t)s[0] * s[1] * s[2];
}

DFM_HD int64_t lin(const int s[3], int i, int j, int k) { return ((int64_t)i * s[1] + j) * s[2] + k; }

// Inverse of lin(); grids stay below 2^32 points, so 32-bit division is used.
DFM_HD void unlin(const int s[3], int64_t id, int& i, int& j, int& k) {
  uint32_t u = (uint32_t)id, s2 = (uint32_t)s[2], s1 = (uint32_t)s[1];
  uint32_t t = u / s2;
  k = (int)(u - t * s2);
  uint32_t q = t / s1;
  j = (int)(t - q * s1);
  i = (int)q;
}

// Node position of index (i,j,k) on grid `a` (face axis or 3 = cells):
// offset 0 along the face normal, 0.5 elsewhere.
DFM_HD void node_pos(const Geo& g, int a, int i, int j, int k, float p[3]) {
  int I[3] = {i, j, k};
  for (int b = 0; b < 3; ++b) p[b] = ((float)I[b] + ((b == a) ? 0.0f : 0.5f)) * g.dx;
}

// XCD-aware block remap: the hardware deals consecutive workgroups round-robin
// over the 8 XCDs; give each XCD one contiguous range of logical blocks so the
// spatial neighbourhood a gather touches stays in that XCD's L2.
DFM_HD int64_t xcd_block(int64_t bid, int64_t nblocks) {
  int64_t nb8 = nblocks & ~int64_t(7);
  if (bid >= nb8) return bid;
  return (bid & 7) * (nb8 >> 3) + (bid >> 3);
}

// An empty asm that takes x: the load that produced x is then used
// unconditionally at this point, so the compiler issues it together with
// the loads before it instead of moving it next to its first use and
// waiting for it there (one memory round trip per load).
template <typename T>
DFM_HD void pin(T& x) {
  asm volatile("" : "+v"(x));
}
template <typename T, int N>
DFM_HD void pin(T (&a)[N]) {
#pragma unroll
  for (int t = 0; t < N; ++t) pin(a[t]);
}

// f + o for a 32-bit element offset o into a field of fewer than 2^30
// values.  The byte offset stays 32-bit, so a load from a wave-uniform f is
// one base register pair plus one offset register, with no 64-bit multiply
// or add per tap.
DFM_HD const float* at32(const float* f, uint32_t o) {
  return reinterpret_cast<const float*>(reinterpret_cast<const char*>(f) + (o << 2));
}
// The same for a 32-bit byte offset.
DFM_HD const float* at32b(const float* f, uint32_t ob) {
  return reinterpret_cast<const float*>(reinterpret_cast<const char*>(f) + ob);
}

// Grid-stride walk over [0, n) for kernels whose grid may be smaller than the
// index space: the blocks on one XCD sweep one contiguous slab of it (n / 8,
// rounded up to whole blocks), so the neighbouring planes a stencil reads are
// the ones that XCD's L2 has just loaded.  for (id = first; id < end; id +=
// stride).  Grids that are not a multiple of 8 blocks walk the plain way.
template <typename I>
struct SlabWalk {
  I first, end, stride;
};

template <int BLOCK, typename I = uint32_t>
__device__ inline SlabWalk<I> slab_walk(I n) {
  const I G = gridDim.x;
  if (G & 7) return SlabWalk<I>{(I)blockIdx.x * BLOCK + threadIdx.x, n, G * BLOCK};
  I chunk = (n + 7) / 8;
  chunk = (chunk + BLOCK - 1) / BLOCK * BLOCK;
  const I lo = (I)(blockIdx.x & 7) * chunk;
  const I hi = lo + chunk < n ? lo + chunk : n;
  return SlabWalk<I>{lo + (I)(blockIdx.x >> 3) * BLOCK + threadIdx.x, hi, (G >> 3) * BLOCK};
}

// ----------------------------------------------------------- B-splines ----
// Piecewise polynomials of the uniform B-splines of order 1 (linear, support
// 1), 2 (quadratic, support 1.5), 3 (cubic, support 2) and 4 (quartic,
// support 2.5), written from their closed forms; derivative sign follows x
// (the linear one's derivative at x = 0 is -1, as N_1's derivative in
// 2D/vortex/taichi_utils.py:213-218).
//
// Only the taps inside the support are visited: NT = K + 1 consecutive nodes
// starting at base(s).  The reference loops over 4 / 6 / 8 taps per axis
// (2D/vortex/taichi_utils.py:433,470,506); every tap it visits beyond these
// has |x| >= support and weight exactly 0, so the sums are the same while a
// 3D quadratic sample gathers 27 values instead of 64 (cubic 64 instead of 216).
//
// Within the support, which polynomial piece a tap uses is fixed by its
// position, so the NT weights are evaluated branch-free as polynomials of the
// sample's offset t from the support's reference node REF = base + REF_OFF
// (tap x = t + REF_OFF - tap index); the pieces are those of w(x) below.
template <int K> struct Spline;

template <> struct Spline<1> {
  static constexpr int NT = 2;       // nodes floor(s) .. floor(s)+1
  static constexpr int REF_OFF = 0;  // t = s - floor(s) in [0, 1)
  DFM_HD static int base(float s) { return (int)floorf(s); }
  DFM_HD static float w(float x) { return fmaxf(0.0f, 1.0f - fabsf(x)); }
  DFM_HD static float dw(float x) { return fabsf(x) < 1.0f ? (x < 0.0f ? 1.0f : -1.0f) : 0.0f; }
  DFM_HD static void weights(float t, float* wt) {
    wt[0] = 1.0f - t;
    wt[1] = t;
  }
  DFM_HD static void dweights(float, float* d) {
    d[0] = -1.0f;
    d[1] = 1.0f;
  }
};

template <> struct Spline<2> {
  static constexpr int NT = 3;       // nodes round(s)-1 .. round(s)+1
  static constexpr int REF_OFF = 1;  // t = s - round(s) in [-1/2, 1/2)
  DFM_HD static int base(float s) { return (int)floorf(s + 0.5f) - 1; }
  DFM_HD static float w(float x) {
    float a = fabsf(x);
    return a < 0.5f ? 0.75f - a * a : (a < 1.5f ? 0.5f * (1.5f - a) * (1.5f - a) : 0.0f);
  }
  DFM_HD static float dw(float x) {
    float a = fabsf(x);
    float r = a < 0.5f ? -2.0f * a : (a < 1.5f ? a - 1.5f : 0.0f);
    return x < 0.0f ? -r : r;
  }
  DFM_HD static void weights(float t, float* wt) {
    const float l = 0.5f - t, h = 0.5f + t;
    wt[0] = 0.5f * l * l;
    wt[1] = 0.75f - t * t;
    wt[2] = 0.5f * h * h;
  }
  DFM_HD static void dweights(float t, float* d) {
    d[0] = t - 0.5f;
    d[1] = -2.0f * t;
    d[2] = t + 0.5f;
  }
};

template <> struct Spline<3> {
  static constexpr int NT = 4;       // nodes floor(s)-1 .. floor(s)+2
  static constexpr int REF_OFF = 1;  // t = s - floor(s) in [0, 1)
  DFM_HD static int base(float s) { return (int)floorf(s) - 1; }
  DFM_HD static float w(float x) {
    float a = fabsf(x);
    if (a < 1.0f) return (0.5f * a - 1.0f) * a * a + (2.0f / 3.0f);
    if (a < 2.0f) {
      float t = 2.0f - a;
      return t * t * t * (1.0f / 6.0f);
    }
    return 0.0f;
  }
  DFM_HD static float dw(float x) {
    float a = fabsf(x), r = 0.0f;
    if (a < 1.0f) r = (1.5f * a - 2.0f) * a;
    else if (a < 2.0f) {
      float t = 2.0f - a;
      r = -0.5f * t * t;
    }
    return x < 0.0f ? -r : r;
  }
  DFM_HD static void weights(float t, float* wt) {
    const float u = 1.0f - t;
    wt[0] = u * u * u * (1.0f / 6.0f);              // x = t + 1
    wt[1] = (0.5f * t - 1.0f) * t * t + (2.0f / 3.0f);  // x = t
    wt[2] = (0.5f * u - 1.0f) * u * u + (2.0f / 3.0f);  // x = t - 1
    wt[3] = t * t * t * (1.0f / 6.0f);              // x = t - 2
  }
  DFM_HD static void dweights(float t, float* d) {
    const float u = 1.0f - t;
    d[0] = -0.5f * u * u;
    d[1] = (1.5f * t - 2.0f) * t;
    d[2] = -(1.5f * u - 2.0f) * u;
    d[3] = 0.5f * t * t;
  }
};

template <> struct Spline<4> {
  static constexpr int NT = 5;       // nodes round(s)-2 .. round(s)+2
  static constexpr int REF_OFF = 2;  // t = s - round(s) in [-1/2, 1/2)
  DFM_HD static int base(float s) { return (int)floorf(s + 0.5f) - 2; }
  DFM_HD static float inner(float a) {  // piece 1/2 <= |x| < 3/2
    return (55.0f / 96.0f) + a * ((5.0f / 24.0f) + a * (-1.25f + a * ((5.0f / 6.0f) - a * (1.0f / 6.0f))));
  }
  DFM_HD static float dinner(float a) {
    return (5.0f / 24.0f) + a * (-2.5f + a * (2.5f - a * (2.0f / 3.0f)));
  }
  DFM_HD static float w(float x) {
    float a = fabsf(x), a2 = a * a;
    if (a < 0.5f) return (115.0f / 192.0f) + a2 * (-0.625f + 0.25f * a2);
    if (a < 1.5f) return inner(a);
    if (a < 2.5f) {
      float t = 2.5f - a;
      float t2 = t * t;
      return t2 * t2 * (1.0f / 24.0f);
    }
    return 0.0f;
  }
  DFM_HD static float dw(float x) {
    float a = fabsf(x), r = 0.0f;
    if (a < 0.5f) r = a * (-1.25f + a * a);
    else if (a < 1.5f) r = dinner(a);
    else if (a < 2.5f) {
      float t = 2.5f - a;
      r = -t * t * t * (1.0f / 6.0f);
    }
    return x < 0.0f ? -r : r;
  }
  DFM_HD static void weights(float t, float* wt) {
    const float l = 0.5f - t, h = 0.5f + t, t2 = t * t;
    wt[0] = l * l * l * l * (1.0f / 24.0f);                   // x = t + 2
    wt[1] = inner(1.0f + t);                                  // x = t + 1
    wt[2] = (115.0f / 192.0f) + t2 * (-0.625f + 0.25f * t2);  // x = t
    wt[3] = inner(1.0f - t);                                  // x = t - 1
    wt[4] = h * h * h * h * (1.0f / 24.0f);                   // x = t - 2
  }
  DFM_HD static void dweights(float t, float* d) {
    const float l = 0.5f - t, h = 0.5f + t;
    d[0] = -l * l * l * (1.0f / 6.0f);
    d[1] = dinner(1.0f + t);
    d[2] = t * (-1.25f + t * t);
    d[3] = -dinner(1.0f - t);
    d[4] = h * h * h * (1.0f / 6.0f);
  }
};

// Clamped sample coordinate s in [1, S-2-1e-6] (the reference's clamp,
// 2D/vortex/taichi_utils.py:420-421; [0, S-1-1e-6] for the linear spline,
// :225-226) and the first node of the support.  For the quadratic spline s in
// [1, S-2] puts the 3 nodes inside [0, S-1]; the base is clamped anyway so
// that no input (NaN included) can leave the field.  The other splines clamp
// every node index to the field, as sample() does.
template <int K> DFM_HD float clamp_lo() { return K == 1 ? 0.0f : 1.0f; }
template <int K> DFM_HD float clamp_hi(int S) { return (float)(S - (K == 1 ? 1 : 2)) - 1e-6f; }
template <int K> DFM_HD float clamp_coord(float p_over_dx, float bl, int S) {
  return fmaxf(clamp_lo<K>(), fminf(p_over_dx - bl, clamp_hi<K>(S)));
}
// 1 where the sample coordinate p/dx - bl is strictly inside the clamp range
// (the sample then moves with p), 0 where it is clamped.
template <int K> DFM_HD float clamp_live(float p_over_dx, float bl, int S) {
  const float raw = p_over_dx - bl;
  return (raw > clamp_lo<K>() && raw < clamp_hi<K>(S)) ? 1.0f : 0.0f;
}
template <int K> DFM_HD int support_base(float s, int S) {
  int b = Spline<K>::base(s);
  if (K == 2) b = b < 0 ? 0 : (b > S - 3 ? S - 3 : b);
  return b;
}
template <int K> DFM_HD int tap_index(int b, int t, int S) {
  int id = b + t;
  if (K != 2) id = id < 0 ? 0 : (id > S - 1 ? S - 1 : id);
  return id;
}

// Per-axis tap set of one sample: node indices and the NT weights (and
// derivative weights).
template <int K, bool D> struct Axis {
  int idx[Spline<K>::NT];
  float w[Spline<K>::NT];
  float dw[D ? Spline<K>::NT : 1];
  DFM_HD void setup(float p_over_dx, float bl, int S) {
    float s = clamp_coord<K>(p_over_dx, bl, S);
    int b = support_base<K>(s, S);
    const float t = s - (float)(b + Spline<K>::REF_OFF);
    Spline<K>::weights(t, w);
    if (D) Spline<K>::dweights(t, dw);
#pragma unroll
    for (int q = 0; q < Spline<K>::NT; ++q) idx[q] = tap_index<K>(b, q, S);
  }
};

// Weight (and derivative) of tap a of the support for offset t, with a a
// runtime (wave-uniform) loop index: quadratic taps from a closed form in
// the signed tap position, wider splines from the piecewise form at
// x = t + REF_OFF - a.  Register arrays indexed by a loop counter would be
// placed in scratch memory.
template <int K> DFM_HD float tap_w(float t, int a) {
  if (K == 2) {
    const float h = 0.5f + (float)(a - 1) * t;  // a = 0: 1/2 - t, a = 2: 1/2 + t
    return a == 1 ? 0.75f - t * t : 0.5f * h * h;
  }
  return Spline<K>::w(t + (float)(Spline<K>::REF_OFF - a));
}
template <int K> DFM_HD float tap_dw(float t, int a) {
  if (K == 2) return a == 1 ? -2.0f * t : t + 0.5f * (float)(a - 1);
  return Spline<K>::dw(t + (float)(Spline<K>::REF_OFF - a));
}

// x set-up of a quadratic sample: support base bx and offset tx (the x
// weights are formed per slice from tx, tap_w, rather than held in arrays).
DFM_HD void xsetup2(float p_over_dx, float bl, int S, int& bx, float& tx) {
  const float sx = clamp_coord<2>(p_over_dx, bl, S);
  bx = support_base<2>(sx, S);
  tx = sx - (float)(bx + Spline<2>::REF_OFF);
}

// Quadratic 3D sample (and optionally gradient) of field f with planes of
// S1 x S2 values, from its axis set-ups: x as (bx, tx), y and z as Axis.
// All 27 taps (9 rows of 3, each one vector load) are requested before the
// first is used, one memory round trip per sample; then the taps are
// contracted z, then y, then x.  Quadratic taps are unclamped (the support
// base is), so tap (a, b, c) sits at o00 + (a S1 + b) S2 + c.  o00 is one
// 32-bit offset formed with 24-bit multiplies (a plane holds fewer than 2^24
// values, a field fewer than 2^30); each row load adds a wave-uniform row
// stride to it in 32 bits.
template <bool GRAD>
DFM_HD float interp2_3d(const float* __restrict__ f, int S1i, int S2i, int bx, float tx, const Axis<2, GRAD>& ay,
                        const Axis<2, GRAD>& az, float inv_dx, float g[3]) {
  const uint32_t S1 = (uint32_t)S1i, S2 = (uint32_t)S2i;
  const uint32_t o00 = __umul24(__umul24((uint32_t)bx, S1) + (uint32_t)ay.idx[0], S2) + (uint32_t)az.idx[0];
  float tv[3][3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const float* row = at32b(f, (o00 << 2) + (((uint32_t)a * S1 + (uint32_t)b) * S2 << 2));
#pragma unroll
      for (int c = 0; c < 3; ++c) tv[a][b][c] = row[c];
    }
  }
  float v = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float wxa = tap_w<2>(tx, a);
    float yv = 0.f, yd = 0.f, yz = 0.f;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      float zv = 0.f, zd = 0.f;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float val = tv[a][b][c];
        zv = fmaf(val, az.w[c], zv);
        if (GRAD) zd = fmaf(val, az.dw[c], zd);
      }
      yv = fmaf(zv, ay.w[b], yv);
      if (GRAD) {
        yd = fmaf(zv, ay.dw[b], yd);
        yz = fmaf(zd, ay.w[b], yz);
      }
    }
    v = fmaf(yv, wxa, v);
    if (GRAD) {
      gx = fmaf(yv, tap_dw<2>(tx, a), gx);
      gy = fmaf(yd, wxa, gy);
      gz = fmaf(yz, wxa, gz);
    }
  }
  if (GRAD) {
    g[0] = gx * inv_dx;
    g[1] = gy * inv_dx;
    g[2] = gz * inv_dx;
  }
  return v;
}

// Value (and optionally the full gradient) of the B-spline interpolant of
// field f (shape S, node offset bl) at physical point p.  Separable: the z
// taps are contracted first, then y, then x.  Row addresses are signed, so
// the consecutive z taps of a row fold into one address plus immediate offsets.
template <int K, bool GRAD, int DIM>
DFM_HD float interp(const float* __restrict__ f, const int S[3], const float bl[3], const float p[3], float inv_dx,
                    float g[3]) {
  constexpr int NT = Spline<K>::NT;
  Axis<K, GRAD> ay;
  ay.setup(p[1] * inv_dx, bl[1], S[1]);
  if (DIM == 2) {
    Axis<K, GRAD> ax;
    ax.setup(p[0] * inv_dx, bl[0], S[0]);
    float v = 0.f, gx = 0.f, gy = 0.f;
    // quadratic: tap (a, b) at o00 + a S1 + b
    const uint32_t o00 = __umul24((uint32_t)ax.idx[0], (uint32_t)S[1]) + (uint32_t)ay.idx[0];
#pragma unroll
    for (int a = 0; a < NT; ++a) {
      const uint32_t row = __umul24((uint32_t)ax.idx[a], (uint32_t)S[1]);
      float rv = 0.f, rd = 0.f;
#pragma unroll
      for (int b = 0; b < NT; ++b) {
        float val = K == 2 ? at32b(f, (o00 << 2) + ((uint32_t)a * (uint32_t)S[1] << 2))[b]
                           : *at32(f, row + (uint32_t)ay.idx[b]);
        rv = fmaf(val, ay.w[b], rv);
        if (GRAD) rd = fmaf(val, ay.dw[b], rd);
      }
      v = fmaf(rv, ax.w[a], v);
      if (GRAD) {
        gx = fmaf(rv, ax.dw[a], gx);
        gy = fmaf(rd, ax.w[a], gy);
      }
    }
    if (GRAD) {
      g[0] = gx * inv_dx;
      g[1] = gy * inv_dx;
    }
    return v;
  } else {
    Axis<K, GRAD> az;
    az.setup(p[2] * inv_dx, bl[2], S[2]);
    if constexpr (K == 2) {
      int bx;
      float tx;
      xsetup2(p[0] * inv_dx, bl[0], S[0], bx, tx);
      return interp2_3d<GRAD>(f, S[1], S[2], bx, tx, ay, az, inv_dx, g);
    } else {
      // the x weights are formed per slice (tap_w) rather than held in
      // arrays; one slice of taps in flight at a time (register pressure)
      const float sx = clamp_coord<K>(p[0] * inv_dx, bl[0], S[0]);
      const int bx = support_base<K>(sx, S[0]);
      const float tx = sx - (float)(bx + Spline<K>::REF_OFF);
      float v = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
      for (int a = 0; a < NT; ++a) {
        const float wxa = tap_w<K>(tx, a);
        const int ia = tap_index<K>(bx, a, S[0]);
        float yv = 0.f, yd = 0.f, yz = 0.f;
#pragma unroll
        for (int b = 0; b < NT; ++b) {
          const float* row = f + ((int64_t)ia * S[1] + ay.idx[b]) * S[2];
          float zv = 0.f, zd = 0.f;
#pragma unroll
          for (int c = 0; c < NT; ++c) {
            const float val = row[az.idx[c]];
            zv = fmaf(val, az.w[c], zv);
            if (GRAD) zd = fmaf(val, az.dw[c], zd);
          }
          yv = fmaf(zv, ay.w[b], yv);
          if (GRAD) {
            yd = fmaf(zv, ay.dw[b], yd);
            yz = fmaf(zd, ay.w[b], yz);
          }
        }
        v = fmaf(yv, wxa, v);
        if (GRAD) {
          gx = fmaf(yv, tap_dw<K>(tx, a), gx);
          gy = fmaf(yd, wxa, gy);
          gz = fmaf(yz, wxa, gz);
        }
      }
      if (GRAD) {
        g[0] = gx * inv_dx;
        g[1] = gy * inv_dx;
        g[2] = gz * inv_dx;
      }
      return v;
    }
  }
}

// MAC velocity: component pointers and their face-grid shapes.
struct Mac {
  const float* c[3];
};

// u(p) for the MAC field (spline order K, quadratic by default), optionally
// J[a][b] = d u_a / d x_b.  The component loop stays rolled; its results are
// stored through wave-uniform branches (a dynamically indexed val[a] / J[a]
// would live in scratch memory).
template <bool GRAD, int DIM, int K = 2>
DFM_HD void mac_eval(const Geo& g, const Mac& u, const float p[3], float val[3], float J[3][3]) {
#pragma unroll 1
  for (int a = 0; a < DIM; ++a) {
    int S[3];
    shape_of(g, a, S);
    const float bl[3] = {a == 0 ? 0.0f : 0.5f, a == 1 ? 0.0f : 0.5f, a == 2 ? 0.0f : 0.5f};
    float gr[3] = {0.f, 0.f, 0.f};
    const float v = interp<K, GRAD, DIM>(u.c[a], S, bl, p, g.inv_dx, gr);
    if (a == 0) {
      val[0] = v;
      if (GRAD) { J[0][0] = gr[0]; J[0][1] = gr[1]; J[0][2] = gr[2]; }
    } else if (a == 1) {
      val[1] = v;
      if (GRAD) { J[1][0] = gr[0]; J[1][1] = gr[1]; J[1][2] = gr[2]; }
    } else {
      val[2] = v;
      if (GRAD) { J[2][0] = gr[0]; J[2][1] = gr[1]; J[2][2] = gr[2]; }
    }
  }
}

// mac_eval<GRAD, 3> (quadratic) with the components sharing their axis
// set-ups: along axis b, component b samples at offset 0 on n[b] + 1 nodes
// and the other two at offset 1/2 on n[b] nodes, so 6 set-ups serve the 9
// (component, axis) pairs.  Each component's arithmetic is interp<2, GRAD,
// 3>'s, so the results are bitwise mac_eval's.  The components are unrolled,
// which takes more registers than mac_eval's rolled loop: it serves the
// kernels whose per-point work is these samples (march, back-trace).
template <bool GRAD>
DFM_HD void mac_eval_shared(const Geo& g, const Mac& u, const float p[3], float val[3], float J[3][3]) {
  const float px = p[0] * g.inv_dx, py = p[1] * g.inv_dx, pz = p[2] * g.inv_dx;
  const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
  int bxN, bxC;
  float txN, txC;
  xsetup2(px, 0.0f, n0 + 1, bxN, txN);
  xsetup2(px, 0.5f, n0, bxC, txC);
  Axis<2, GRAD> yN, yC, zN, zC;
  yN.setup(py, 0.0f, n1 + 1);
  yC.setup(py, 0.5f, n1);
  zN.setup(pz, 0.0f, n2 + 1);
  zC.setup(pz, 0.5f, n2);
  float gr[3] = {0.f, 0.f, 0.f};
  val[0] = interp2_3d<GRAD>(u.c[0], n1, n2, bxN, txN, yC, zC, g.inv_dx, gr);
  if (GRAD) { J[0][0] = gr[0]; J[0][1] = gr[1]; J[0][2] = gr[2]; }
  val[1] = interp2_3d<GRAD>(u.c[1], n1 + 1, n2, bxC, txC, yN, zC, g.inv_dx, gr);
  if (GRAD) { J[1][0] = gr[0]; J[1][1] = gr[1]; J[1][2] = gr[2]; }
  val[2] = interp2_3d<GRAD>(u.c[2], n1, n2 + 1, bxC, txC, yC, zN, g.inv_dx, gr);
  if (GRAD) { J[2][0] = gr[0]; J[2][1] = gr[1]; J[2][2] = gr[2]; }
}

}  // namespace dfm
