// Device helpers of the split (FP16X2, BF16X2, BF16X3) contractions, shared
// by the contraction of the stored AO operand (k_gemm_mixed.hip) and the
// fused AO-and-contraction kernel (k_fused.hip): the operand scale, the
// split of FP32 values into fp16 / bf16 terms in MFMA fragment order, and
// the 16x16x32 MFMA of each form.
#pragma once

#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace qmckl_amd {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

// power-of-two scale bringing max |x| into [2^13, 2^14); the exponent is
// clamped so that the scale and its inverse stay normal floats
__device__ inline float pow2_scale(float m) {
  if (!(m > 0.0f) || !isfinite(m)) return 1.0f;
  return ldexpf(1.0f, min(100, max(-100, 13 - ilogbf(m))));
}

template <int SP>
struct Terms;  // number of split terms per operand
template <>
struct Terms<QMCKL_AMD_SPLIT_FP16X2> {
  static constexpr int n = 2;
};
template <>
struct Terms<QMCKL_AMD_SPLIT_BF16X2> {
  static constexpr int n = 2;
};
template <>
struct Terms<QMCKL_AMD_SPLIT_BF16X3> {
  static constexpr int n = 3;
};

template <int SP>
__device__ inline f4 mfma(const uint4 &a, const uint4 &b, f4 c) {
  if constexpr (SP == QMCKL_AMD_SPLIT_FP16X2)
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8, a), __builtin_bit_cast(h8, b), c,
                                                  0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8, a), __builtin_bit_cast(b8, b), c,
                                                 0, 0, 0);
}

template <int SP>
__device__ inline void split8(const float (&v)[8], uint4 (&out)[Terms<SP>::n]) {
  if constexpr (SP == QMCKL_AMD_SPLIT_FP16X2) {
    h8 hi, lo;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const _Float16 h = (_Float16)v[e];
      hi[e] = h;
      lo[e] = (_Float16)(v[e] - (float)h);
    }
    out[0] = __builtin_bit_cast(uint4, hi);
    out[1] = __builtin_bit_cast(uint4, lo);
  } else {
    b8 hi, mid, lo;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const __bf16 h = (__bf16)v[e];
      const float r = v[e] - (float)h;
      const __bf16 m = (__bf16)r;
      hi[e] = h;
      mid[e] = m;
      if constexpr (SP == QMCKL_AMD_SPLIT_BF16X3) lo[e] = (__bf16)(r - (float)m);
    }
    out[0] = __builtin_bit_cast(uint4, hi);
    out[1] = __builtin_bit_cast(uint4, mid);
    if constexpr (SP == QMCKL_AMD_SPLIT_BF16X3) out[2] = __builtin_bit_cast(uint4, lo);
  }
}

}  // namespace qmckl_amd
