// Division by a runtime-invariant divisor via one high multiply, for the
// cell-index decompositions of the stencil kernels (an integer divide by a
// non-constant costs ~40 VALU instructions on CDNA).
//
// Method: Granlund & Montgomery, "Division by invariant integers using
// multiplication" (PLDI 1994), unsigned case.  With s = ceil(log2 d) and
// m = floor(2^32 (2^s - d) / d) + 1 (which fits in 32 bits),
//   floor(n / d) = (mulhi(n, m) + n) >> s      for 0 <= n < 2^31,
// the bound on n keeping mulhi(n, m) + n inside 32 bits.
//
// Plain C++ so that host code and a host-compiled test share it; on the GPU
// the 64-bit product's high half compiles to v_mul_hi_u32.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define DFM_FD_HD __host__ __device__ __forceinline__
#else
#define DFM_FD_HD inline
#endif

namespace dfm {

struct FastDiv {
  uint32_t d;  // divisor (>= 1)
  uint32_t m;  // magic multiplier
  uint32_t s;  // shift
};

inline FastDiv make_fastdiv(uint32_t d) {
  if (d == 0) d = 1;
  uint32_t s = 0;
  while ((uint64_t(1) << s) < d) ++s;
  const uint64_t m = ((uint64_t(1) << 32) * ((uint64_t(1) << s) - d)) / d + 1;
  return FastDiv{d, (uint32_t)m, s};
}

DFM_FD_HD uint32_t fdiv(uint32_t n, const FastDiv& f) {
  const uint32_t hi = (uint32_t)(((uint64_t)n * f.m) >> 32);
  return (hi + n) >> f.s;
}

}  // namespace dfm
