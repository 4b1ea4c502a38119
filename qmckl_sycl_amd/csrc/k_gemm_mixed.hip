// Mixed-precision MO contraction on the matrix cores.
//
// Same product as the FP64 kernel, mo[(p,c)][i] = sum_k ao[(p,c)][k] C[i][k]
// (src/qmckl_mo_sycl.cpp:23-50), with the AO operand stored in FP32
// (rounded once from the FP64 assembly) and the MO coefficients pre-split.
// Forms:
//   FP16X2: each operand is scaled by a power of two (AO per (p,c) row from
//           its max |value|, C per MO from its max |coefficient|) and split
//           into fp16 hi + lo; three products hi*hi + hi*lo + lo*hi.
//   BF16X2: bf16 hi + lo, the same three products (no scaling needed).
//   BF16X3: bf16 hi + mid + lo, six products (hh, hm, mh, hl, lh, mm).
//   FP32:   v_mfma_f32_16x16x4_f32 on the FP32 operands.
// The MFMA accumulates in FP32 over `flush` consecutive 32-wide K blocks
// and is then added into FP64 accumulators, which bounds the FP32
// accumulation error to that of short partial sums.  K blocks whose AO tile
// is all zero for the workgroup's 64 rows (screened AOs) are skipped using
// the nonzero flags the AO kernel writes.
//
// Tiling: 256 threads = 2 x 2 wavefronts; a wavefront owns 32 rows x
// 16*CBW columns (2 x CBW tiles of 16x16), the workgroup 64 rows x 32*CBW.
// Per K block the 64x32 FP32 AO slab is loaded once (two 16-byte loads per
// thread), scaled and split into LDS; the C slab (32*CBW MOs x 32 AOs, K
// contiguous per MO) is copied into LDS.  16x16x32 operands: lane l holds
// A[l%16][8(l/16)..8(l/16)+7] and B[8(l/16)..+7][l%16] and receives
// D[4(l/16)+j][l%16], j = 0..3.
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace qmckl_amd {

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kBM = 64;
constexpr int kKB = 32;
constexpr int kPad = 8;  // halves of padding per LDS row

// power-of-two scale bringing max |x| into [2^13, 2^14)
__device__ inline float pow2_scale(float m) {
  if (!(m > 0.0f) || !isfinite(m)) return 1.0f;
  return ldexpf(1.0f, 13 - ilogbf(m));
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
__device__ inline void split_store(float x, uint16_t *dst, int stride) {
  if constexpr (SP == QMCKL_AMD_SPLIT_FP16X2) {
    const _Float16 h = (_Float16)x;
    const _Float16 l = (_Float16)(x - (float)h);
    dst[0] = __builtin_bit_cast(uint16_t, h);
    dst[stride] = __builtin_bit_cast(uint16_t, l);
  } else {
    const __bf16 h = (__bf16)x;
    const float r = x - (float)h;
    const __bf16 m = (__bf16)r;
    dst[0] = __builtin_bit_cast(uint16_t, h);
    dst[stride] = __builtin_bit_cast(uint16_t, m);
    if constexpr (SP == QMCKL_AMD_SPLIT_BF16X3) {
      const __bf16 l = (__bf16)(r - (float)m);
      dst[2 * stride] = __builtin_bit_cast(uint16_t, l);
    }
  }
}

template <int SP>
__device__ inline f4 mfma(const uint4 &a, const uint4 &b, f4 c) {
  if constexpr (SP == QMCKL_AMD_SPLIT_FP16X2)
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8, a), __builtin_bit_cast(h8, b), c,
                                                  0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8, a), __builtin_bit_cast(b8, b), c,
                                                 0, 0, 0);
}

// Split GEMM (FP16X2, BF16X2, BF16X3).
//   A   [R][lda] FP32 AO rows, lda a multiple of 32, entries past K zero
//   rowmax [R] max |A row| (scale source)
//   Bs  [T][Ncol][ldb] split C terms (pre-scaled by colscale), K contiguous
//   colinv [Ncol] inverse column scales
template <int SP, int CBW>
__global__ void __launch_bounds__(kThreads)
gemm_split_kernel(int64_t R, int64_t K, int64_t Ncol, const float *__restrict__ A, int64_t lda,
                  const float *__restrict__ rowmax, const uint16_t *__restrict__ Bs, int64_t ldb,
                  const float *__restrict__ colinv, const uint8_t *__restrict__ flags, int nkb,
                  int flush, double *__restrict__ C, int64_t ldc) {
  constexpr int T = Terms<SP>::n;
  constexpr int BN = 32 * CBW;
  constexpr int LDR = kKB + kPad;  // LDS row length in halves
  __shared__ uint16_t sA[T][kBM * LDR];
  __shared__ uint16_t sB[T][BN * LDR];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t row0 = (int64_t)blockIdx.x * kBM;
  const int64_t col0 = (int64_t)blockIdx.y * BN;
  const int64_t nblk = (K + kKB - 1) / kKB;
  const int64_t Bterm = Ncol * ldb;

  // this thread's A load slot: row r_ld, 8 consecutive k at seg
  const int r_ld = tid >> 2, seg = (tid & 3) * 8;
  const int64_t grow = row0 + r_ld;
  const float sc = grow < R ? pow2_scale(rowmax[grow]) : 1.0f;

  f4 d[2][CBW];
  double acc[2][CBW][4];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      d[rb][cb] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[rb][cb][j] = 0.0;
    }

  int pending = 0;
  for (int64_t kb = 0; kb < nblk; ++kb) {
    if (flags != nullptr && flags[(int64_t)blockIdx.x * nkb + kb] == 0) continue;
    const int64_t k0 = kb * kKB;
    // A slab: FP32 -> scaled split terms
    {
      float v[8];
      if (grow < R) {
        const float4 *src = reinterpret_cast<const float4 *>(A + grow * lda + k0 + seg);
        const float4 x0 = src[0], x1 = src[1];
        v[0] = x0.x; v[1] = x0.y; v[2] = x0.z; v[3] = x0.w;
        v[4] = x1.x; v[5] = x1.y; v[6] = x1.z; v[7] = x1.w;
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = 0.f;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j)
        split_store<SP>(v[j] * sc, &sA[0][r_ld * LDR + seg + j], kBM * LDR);
    }
    // C slab: BN columns x 32 k of each term, 16 bytes per copy
    for (int i = tid; i < T * BN * 4; i += kThreads) {
      const int t = i / (BN * 4), rem = i - t * BN * 4;
      const int c = rem >> 2, s8 = (rem & 3) * 8;
      const int64_t gc = col0 + c;
      uint4 val = {0u, 0u, 0u, 0u};
      if (gc < Ncol) val = *reinterpret_cast<const uint4 *>(Bs + t * Bterm + gc * ldb + k0 + s8);
      *reinterpret_cast<uint4 *>(&sB[t][c * LDR + s8]) = val;
    }
    __syncthreads();
    uint4 a[T][2];
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
        a[t][rb] = *reinterpret_cast<const uint4 *>(&sA[t][(wr * 32 + rb * 16 + li) * LDR + 8 * lk]);
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      uint4 b[T];
#pragma unroll
      for (int t = 0; t < T; ++t)
        b[t] = *reinterpret_cast<const uint4 *>(&sB[t][(wc * 16 * CBW + cb * 16 + li) * LDR + 8 * lk]);
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        f4 x = d[rb][cb];
        x = mfma<SP>(a[0][rb], b[0], x);
        x = mfma<SP>(a[0][rb], b[1], x);
        x = mfma<SP>(a[1][rb], b[0], x);
        if constexpr (T == 3) {
          x = mfma<SP>(a[0][rb], b[T - 1], x);
          x = mfma<SP>(a[T - 1][rb], b[0], x);
          x = mfma<SP>(a[1][rb], b[1], x);
        }
        d[rb][cb] = x;
      }
    }
    __syncthreads();
    if (++pending >= flush) {
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < CBW; ++cb) {
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[rb][cb][j] += (double)d[rb][cb][j];
          d[rb][cb] = f4{0.f, 0.f, 0.f, 0.f};
        }
      pending = 0;
    }
  }
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      const int64_t col = col0 + wc * 16 * CBW + cb * 16 + li;
      if (col >= Ncol) continue;
      const double ci = (double)colinv[col];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t row = row0 + wr * 32 + rb * 16 + 4 * lk + j;
        if (row >= R) continue;
        const double ri = 1.0 / (double)pow2_scale(rowmax[row]);
        C[row * ldc + col] = (acc[rb][cb][j] + (double)d[rb][cb][j]) * ri * ci;
      }
    }
}

// FP32 form: v_mfma_f32_16x16x4_f32, lane l holds A[l%16][l/16], B[l/16][l%16].
//   Bf [Ncol][ldb] FP32 C (unscaled), K contiguous.
template <int CBW>
__global__ void __launch_bounds__(kThreads)
gemm_f32_kernel(int64_t R, int64_t K, int64_t Ncol, const float *__restrict__ A, int64_t lda,
                const float *__restrict__ Bf, int64_t ldb, const uint8_t *__restrict__ flags, int nkb,
                int flush, double *__restrict__ C, int64_t ldc) {
  constexpr int BN = 32 * CBW;
  constexpr int LDF = kKB + 4;  // LDS row length in floats
  __shared__ float sA[kBM * LDF];
  __shared__ float sB[BN * LDF];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t row0 = (int64_t)blockIdx.x * kBM;
  const int64_t col0 = (int64_t)blockIdx.y * BN;
  const int64_t nblk = (K + kKB - 1) / kKB;
  const int r_ld = tid >> 2, seg = (tid & 3) * 8;
  const int64_t grow = row0 + r_ld;

  f4 d[2][CBW];
  double acc[2][CBW][4];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      d[rb][cb] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[rb][cb][j] = 0.0;
    }
  int pending = 0;
  for (int64_t kb = 0; kb < nblk; ++kb) {
    if (flags != nullptr && flags[(int64_t)blockIdx.x * nkb + kb] == 0) continue;
    const int64_t k0 = kb * kKB;
    {
      float4 x0 = {0.f, 0.f, 0.f, 0.f}, x1 = x0;
      if (grow < R) {
        const float4 *src = reinterpret_cast<const float4 *>(A + grow * lda + k0 + seg);
        x0 = src[0];
        x1 = src[1];
      }
      float *dst = &sA[r_ld * LDF + seg];
      dst[0] = x0.x; dst[1] = x0.y; dst[2] = x0.z; dst[3] = x0.w;
      dst[4] = x1.x; dst[5] = x1.y; dst[6] = x1.z; dst[7] = x1.w;
    }
    for (int i = tid; i < BN * 8; i += kThreads) {
      const int c = i >> 3, s4 = (i & 7) * 4;
      const int64_t gc = col0 + c;
      float4 val = {0.f, 0.f, 0.f, 0.f};
      if (gc < Ncol) val = *reinterpret_cast<const float4 *>(Bf + gc * ldb + k0 + s4);
      float *dst = &sB[c * LDF + s4];
      dst[0] = val.x; dst[1] = val.y; dst[2] = val.z; dst[3] = val.w;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kKB; kk += 4) {
      float a[2];
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) a[rb] = sA[(wr * 32 + rb * 16 + li) * LDF + kk + lk];
#pragma unroll
      for (int cb = 0; cb < CBW; ++cb) {
        const float b = sB[(wc * 16 * CBW + cb * 16 + li) * LDF + kk + lk];
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) d[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rb], b, d[rb][cb], 0, 0, 0);
      }
    }
    __syncthreads();
    if (++pending >= flush) {
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < CBW; ++cb) {
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[rb][cb][j] += (double)d[rb][cb][j];
          d[rb][cb] = f4{0.f, 0.f, 0.f, 0.f};
        }
      pending = 0;
    }
  }
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      const int64_t col = col0 + wc * 16 * CBW + cb * 16 + li;
      if (col >= Ncol) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t row = row0 + wr * 32 + rb * 16 + 4 * lk + j;
        if (row < R) C[row * ldc + col] = acc[rb][cb][j] + (double)d[rb][cb][j];
      }
    }
}

// One workgroup per MO i: column scale from max_k |C[i][k]| and the split
// terms of C[i][k] * scale, zero past K.
template <int SP>
__global__ void __launch_bounds__(kThreads)
split_coef_kernel(int64_t M, int64_t K, const double *__restrict__ coef, int64_t ldb,
                  uint16_t *__restrict__ Bs, float *__restrict__ colinv) {
  __shared__ float red[kThreads];
  const int64_t i = blockIdx.x;
  const double *row = coef + i * K;
  float m = 0.f;
  for (int64_t k = threadIdx.x; k < K; k += kThreads) m = fmaxf(m, fabsf((float)row[k]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  const float sc = SP == QMCKL_AMD_SPLIT_FP32 ? 1.0f : pow2_scale(red[0]);
  if (threadIdx.x == 0) colinv[i] = 1.0f / sc;
  for (int64_t k = threadIdx.x; k < ldb; k += kThreads) {
    const float x = k < K ? (float)row[k] * sc : 0.0f;
    if constexpr (SP == QMCKL_AMD_SPLIT_FP32) {
      reinterpret_cast<float *>(Bs)[i * ldb + k] = x;
    } else {
      split_store<SP>(x, Bs + i * ldb + k, (int)(M * ldb));
    }
  }
}

template <int SP, int CBW>
hipError_t launch_split(hipStream_t s, int64_t R, int64_t K, int64_t Ncol, const float *A, int64_t lda,
                        const float *rowmax, const void *B, int64_t ldb, const float *colinv,
                        const uint8_t *flags, int nkb, int flush, double *C, int64_t ldc) {
  constexpr int BN = 32 * CBW;
  const int64_t gx = (R + kBM - 1) / kBM, gy = (Ncol + BN - 1) / BN;
  if (gx > 0x7fffffff || gy > 65535) return hipErrorInvalidValue;
  if constexpr (SP == QMCKL_AMD_SPLIT_FP32)
    hipLaunchKernelGGL(gemm_f32_kernel<CBW>, dim3((unsigned)gx, (unsigned)gy), dim3(kThreads), 0, s, R,
                       K, Ncol, A, lda, (const float *)B, ldb, flags, nkb, flush, C, ldc);
  else
    hipLaunchKernelGGL((gemm_split_kernel<SP, CBW>), dim3((unsigned)gx, (unsigned)gy), dim3(kThreads), 0, s,
                       R, K, Ncol, A, lda, rowmax, (const uint16_t *)B, ldb, colinv, flags, nkb, flush, C,
                       ldc);
  return hipGetLastError();
}

template <int SP>
hipError_t launch_form(hipStream_t s, int64_t R, int64_t K, int64_t Ncol, const float *A, int64_t lda,
                       const float *rowmax, const void *B, int64_t ldb, const float *colinv,
                       const uint8_t *flags, int nkb, int flush, double *C, int64_t ldc) {
  const int64_t nb = (Ncol + 31) / 32;  // 32-column groups
  if (nb <= 1) return launch_split<SP, 1>(s, R, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, flush, C, ldc);
  if (nb <= 2) return launch_split<SP, 2>(s, R, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, flush, C, ldc);
  if (nb <= 3) return launch_split<SP, 3>(s, R, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, flush, C, ldc);
  if (nb <= 4) return launch_split<SP, 4>(s, R, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, flush, C, ldc);
  return launch_split<SP, 5>(s, R, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, flush, C, ldc);
}

}  // namespace

int split_terms(int form) {
  switch (form) {
    case QMCKL_AMD_SPLIT_FP16X2:
    case QMCKL_AMD_SPLIT_BF16X2: return 2;
    case QMCKL_AMD_SPLIT_BF16X3: return 3;
    case QMCKL_AMD_SPLIT_FP32: return 2;  // one float = two 16-bit slots
    default: return 0;
  }
}

hipError_t launch_split_coef(hipStream_t s, int form, int64_t M, int64_t K, const double *coef,
                             int64_t ldb, uint16_t *B, float *colinv) {
  if (M <= 0) return hipSuccess;
  switch (form) {
    case QMCKL_AMD_SPLIT_FP16X2:
      hipLaunchKernelGGL(split_coef_kernel<QMCKL_AMD_SPLIT_FP16X2>, dim3((unsigned)M), dim3(kThreads), 0, s, M, K, coef, ldb, B, colinv);
      break;
    case QMCKL_AMD_SPLIT_BF16X2:
      hipLaunchKernelGGL(split_coef_kernel<QMCKL_AMD_SPLIT_BF16X2>, dim3((unsigned)M), dim3(kThreads), 0, s, M, K, coef, ldb, B, colinv);
      break;
    case QMCKL_AMD_SPLIT_BF16X3:
      hipLaunchKernelGGL(split_coef_kernel<QMCKL_AMD_SPLIT_BF16X3>, dim3((unsigned)M), dim3(kThreads), 0, s, M, K, coef, ldb, B, colinv);
      break;
    case QMCKL_AMD_SPLIT_FP32:
      hipLaunchKernelGGL(split_coef_kernel<QMCKL_AMD_SPLIT_FP32>, dim3((unsigned)M), dim3(kThreads), 0, s, M, K, coef, ldb, B, colinv);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_mo_gemm_mixed(hipStream_t s, int form, int64_t R, int64_t K, int64_t Ncol,
                                const float *A, int64_t lda, const float *rowmax, const void *B,
                                int64_t ldb, const float *colinv, const uint8_t *flags, int nkb,
                                int flush, double *C, int64_t ldc) {
  if (R <= 0 || Ncol <= 0) return hipSuccess;
  if (K <= 0 || (lda % kKB) != 0 || (ldb % kKB) != 0 || lda < K || ldb < K || flush < 1)
    return hipErrorInvalidValue;
  switch (form) {
    case QMCKL_AMD_SPLIT_FP16X2:
      return launch_form<QMCKL_AMD_SPLIT_FP16X2>(s, R, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, flush, C, ldc);
    case QMCKL_AMD_SPLIT_BF16X2:
      return launch_form<QMCKL_AMD_SPLIT_BF16X2>(s, R, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, flush, C, ldc);
    case QMCKL_AMD_SPLIT_BF16X3:
      return launch_form<QMCKL_AMD_SPLIT_BF16X3>(s, R, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, flush, C, ldc);
    case QMCKL_AMD_SPLIT_FP32:
      return launch_form<QMCKL_AMD_SPLIT_FP32>(s, R, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, flush, C, ldc);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace qmckl_amd
