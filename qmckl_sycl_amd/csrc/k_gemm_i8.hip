// INT8X5: the accurate mixed form of the MO contraction, on the int8
// matrix cores with exact integer sums.
//
// Same product as the other contractions, mo[(p,c)][i] = sum_k ao[(p,c)][k]
// C[i][k] (src/qmckl_mo_sycl.cpp:23-50).  The FP32 forms lose accuracy on
// badly conditioned bases because an MFMA rounds its FP32 partial sums, at
// about 2^-24 of sum |C||ao|, whatever the precision of its operands.  Here
// each operand is cut into five signed slices in fixed point against a
// power-of-two scale: y = x 2^e ~ t0 + t1 2^-7 + ... + t4 2^-28, the leading
// slice in [-64, 64], the others in [-64, 64] too (balanced digits: each is
// rint of the remainder times 128).  The scale is per (AO row, 64-AO stage)
// for the AO operand, whose value is read as the FP32 value plus its FP32
// residual (the FP64 value to 2^-48), and per MO for the coefficients.  The
// 15 slice products of order i + j <= 4 run as v_mfma_i32_16x16x64_i8, one
// per 64-K stage; the products of one order share one int32 sum (at most 5
// products x 64 terms x 64 x 64 < 2^21: exact).  Each stage is folded into
// FP64 as
//   x = (D0*128 + D1) + (D2*128 + D3) 2^-14 + D4 2^-21   (exact in FP64)
// times the row's stage scale 2^-(e+7); the MO scale is applied at the
// store.  What is dropped (slice products of order >= 5, the slices'
// truncation, the residual's rounding) stays below about 2^-33 of
// (max |ao| of the row's stage) x (max |C| of the MO) per term.  Modelled
// in numpy on the reference's ChBrClF basis: 4.7e-8 of each row's maximum
// against 4e-6 for the FP32-class forms; 1.5e-9 on the 64-water system.
//
// 16x16x64 int8 operands: lane 16g + r holds A[r][16g .. 16g+15] and
// B[16g .. 16g+15][r] and receives D[4g + j][r], j = 0..3 (checked on the
// device; the K order within a stage only has to agree between A and B).
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace qmckl_amd {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kSl = 5;        // slices per operand
constexpr int kStageK = 64;   // K per stage: one 16x16x64 MFMA
constexpr int kRows = 32;     // AO rows per workgroup (half a 64-row tile)
constexpr int kMaxSt = kMaxFlagBlocks / 2;

// y ~ t[0] + t[1] 2^-7 + ... + t[4] 2^-28 for |y| <= 64 (every step exact)
__device__ inline void slices5(double y, int (&t)[kSl]) {
#pragma unroll
  for (int i = 0; i < kSl; ++i) {
    const double r = rint(y);
    t[i] = (int)r;
    y = (y - r) * 128.0;
  }
}

// the exponent e with m 2^e in [32, 64), clamped so the scales stay normal
__device__ inline int slice_exponent(double m) {
  if (!(m > 0.0) || !isfinite(m)) return 0;
  return min(900, max(-900, 5 - ilogb(m)));
}

// One workgroup per prepared column i (MO col_mo[i]; zeros for i >= M):
// colinv[i] = 2^-e with max_k |C[i][k]| 2^e in [32, 64), and the slices of
// C[i][k] 2^e in the contraction's fragment order: fragment
// ((i / 16) * nst + k / 64) * 5 + s of 1 KB, lane 16 ((k % 64) / 16) + i % 16,
// byte k % 16.
__global__ void __launch_bounds__(kThreads)
slice_coef_i8_kernel(int64_t M, int64_t K, const double *__restrict__ coef, int64_t ldb,
                     const int32_t *__restrict__ col_mo, int8_t *__restrict__ Bq, float *__restrict__ colinv) {
  __shared__ double red[kThreads];
  const int64_t i = blockIdx.x;
  const bool real = i < M;
  const double *row = coef + (real ? (col_mo != nullptr ? (int64_t)col_mo[i] : i) : 0) * K;
  double m = 0.0;
  if (real)
    for (int64_t k = threadIdx.x; k < K; k += kThreads) m = fmax(m, fabs(row[k]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  const int e = min(120, max(-120, slice_exponent(red[0])));  // colinv is a float
  if (threadIdx.x == 0 && real) colinv[i] = (float)ldexp(1.0, -e);
  const int64_t nst = ldb / kStageK;
  for (int64_t k = threadIdx.x; k < ldb; k += kThreads) {
    int t[kSl];
    slices5(real && k < K ? ldexp(row[k], e) : 0.0, t);
    const int64_t lane = 16 * ((k % kStageK) / 16) + i % 16;
    const int64_t f0 = ((i / 16) * nst + k / kStageK) * kSl;
#pragma unroll
    for (int s = 0; s < kSl; ++s) Bq[(f0 + s) * 1024 + lane * 16 + k % 16] = (int8_t)t[s];
  }
}

__device__ inline v4i mfma8(const v4i &a, const v4i &b, const v4i &c) {
  return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

// Workgroup: kRows AO rows of one component (half a 64-row tile of the
// blocked operand) x 64*CBW columns (all MOs up to 320), four wavefronts;
// wavefront w owns the column blocks w, w + 4, ... (CBW of them) for both
// 16-row blocks.  The live-stage list and per-stage live bits are those of
// the live-block split kernel (nonzero AO flags of the tile, nonzero
// coefficient blocks of the column plan), one bit per (stage, column
// block); a 32-AO half whose flag is zero is taken as zeros, since the AO
// operand kernel may leave it unwritten.  Per stage: every thread slices 8
// consecutive AOs of one row (the row's 8 threads are neighbouring lanes,
// so the row's stage maximum
// is three shuffles) into the double-buffered LDS slab [slice][row][64 K];
// one barrier; the next stage's AO values are requested; each wavefront
// reads its A fragments from the slab and, per live column block, its five
// coefficient fragments (the next block's are requested while the current
// one computes) and runs 2 x 15 MFMAs and the fold.
template <int CBW>
__global__ void __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(2, 2)))
gemm_i8_kernel(int64_t P, int64_t npt, int64_t K, int64_t Ncol, const float *__restrict__ A,
               const float *__restrict__ Alo, int64_t lda, const int8_t *__restrict__ Bq, int64_t ldb,
               const float *__restrict__ colinv, const uint8_t *__restrict__ flags, int nkb,
               const int32_t *__restrict__ perm, double *__restrict__ C, int64_t ldc, ColPlan plan) {
  constexpr int kSlabBytes = kSl * kRows * kStageK;  // one buffer
  constexpr int kOutBytes = 16 * 64 * CBW * 8;       // 16 output rows in FP64
  constexpr int kBufBytes = 2 * kSlabBytes > kOutBytes ? 2 * kSlabBytes : kOutBytes;
  __shared__ __attribute__((aligned(16))) unsigned char s_buf[kBufBytes];
  __shared__ double s_rs[2][kRows];  // 2^-(e + 7) of each row, per slab buffer
  __shared__ int s_list[kMaxSt];
  __shared__ uint32_t s_live[kMaxSt];
  __shared__ int s_count;
  __shared__ int32_t s_perm[kRows];
  __shared__ int32_t s_mloc[64 * CBW];
  __shared__ float s_ci[64 * CBW];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int cq = __builtin_amdgcn_readfirstlane(wave);
  // workgroup t: its 32-row half, component, 64-point tile and column
  // group, in that order from the fastest (the ten workgroups of a point
  // tile have the same live stages and are dispatched together)
  const int64_t t = blockIdx.x + (int64_t)gridDim.x * blockIdx.y;
  const int half = (int)(t % 2);
  const int64_t comp = (t / 2) % 5, pt = (t / 10) % npt, by = t / (10 * npt);
  const int64_t rt = comp * npt + pt;
  const int r0 = half * kRows;
  const int64_t p0 = pt * 64 + r0;
  const int64_t gb0 = by * 4 * CBW;
  const int nst = (int)((K + kStageK - 1) / kStageK);

  // epilogue tables
  if (tid < kRows) {
    const int64_t i = p0 + tid;
    s_perm[tid] = i < P ? (perm != nullptr ? perm[i] : (int32_t)i) : -1;
  }
  for (int c = tid; c < 64 * CBW; c += kThreads) {
    const int64_t col = gb0 * 16 + c;
    s_mloc[c] = col < Ncol ? (int32_t)((plan.col_mo != nullptr ? (int64_t)plan.col_mo[col] : col) - gb0 * 16) : -1;
    s_ci[c] = col < Ncol ? colinv[col] : 0.0f;
  }

  // live stages, one bit per column block of the workgroup
  if (wave == 0) {
    const uint8_t *f = flags != nullptr ? flags + rt * nkb : nullptr;
    int base = 0;
    for (int s0 = 0; s0 < nst; s0 += 64) {
      const int st = s0 + lane;
      const bool h0 = st < nst && (f == nullptr || f[2 * st] != 0);
      const bool h1 = st < nst && (f == nullptr || (2 * st + 1 < nkb && f[2 * st + 1] != 0));
      uint32_t lv = 0;
      if (h0 || h1) {
        if (plan.blk != nullptr) {
          const int64_t g = 4 * (int64_t)st;
          uint64_t w[4 * CBW];
#pragma unroll
          for (int b = 0; b < 4 * CBW; ++b) {
            const int64_t blk = gb0 + b < plan.nblk ? gb0 + b : plan.nblk - 1;
            w[b] = plan.blk[blk * plan.nw + (g >> 6)];
          }
#pragma unroll
          for (int b = 0; b < 4 * CBW; ++b) {
            const uint64_t nib = gb0 + b < plan.nblk ? (w[b] >> (g & 63)) & 15ull : 0ull;
            const bool live = (h0 && (nib & 3ull) != 0) || (h1 && (nib & 12ull) != 0);
            lv |= (live ? 1u : 0u) << b;
          }
        } else {
          lv = (1u << (4 * CBW)) - 1u;
        }
      }
      const bool live = lv != 0;
      const uint64_t m = __ballot(live);
      if (live) {
        const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
        s_list[pos] = st;
        // bits 30, 31: the stage's two 32-AO halves hold values (the AO
        // operand kernel may leave a half whose flag is zero unwritten)
        s_live[pos] = lv | (h0 ? 1u << 30 : 0u) | (h1 ? 1u << 31 : 0u);
      }
      base += __popcll(m);
    }
    if (lane == 0) s_count = base;
  }
  __syncthreads();
  const int count = s_count;

  // AO slab: thread (row ar, group q) holds AOs 8q .. 8q+7 of the stage
  const int ar = tid >> 3, q = tid & 7;
  const float *Ap = A + rt * (lda / 4) * 256 + (r0 + ar) * 4 + 2 * q * 256;
  const float *Lp = Alo + rt * (lda / 4) * 256 + (r0 + ar) * 4 + 2 * q * 256;
  f4 ra[2], rl[2];
  auto load_a = [&](int64_t k0) {
    const int64_t o = (k0 / 4) * 256;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      ra[h] = __builtin_nontemporal_load(reinterpret_cast<const f4 *>(Ap + o + h * 256));
      rl[h] = __builtin_nontemporal_load(reinterpret_cast<const f4 *>(Lp + o + h * 256));
    }
  };
  auto write_slab = [&](int buf, uint32_t halves) {
    if (!((halves >> (q >> 2)) & 1u)) {  // this thread's half is unwritten: zeros
#pragma unroll
      for (int h = 0; h < 2; ++h) ra[h] = rl[h] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    float m = 0.0f;
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int j = 0; j < 4; ++j) m = fmaxf(m, fabsf(ra[h][j]));
    m = fmaxf(m, __shfl_xor(m, 1));
    m = fmaxf(m, __shfl_xor(m, 2));
    m = fmaxf(m, __shfl_xor(m, 4));
    // |hi + lo| <= |hi| (1 + 2^-24), so the leading slice stays within 64
    const int e = slice_exponent((double)m);
    const double sc = ldexp(1.0, e);
    uint32_t pk[kSl][2];
#pragma unroll
    for (int s = 0; s < kSl; ++s) pk[s][0] = pk[s][1] = 0u;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      int t[kSl];
      slices5(((double)ra[j / 4][j % 4] + (double)rl[j / 4][j % 4]) * sc, t);
#pragma unroll
      for (int s = 0; s < kSl; ++s) pk[s][j / 4] |= ((uint32_t)t[s] & 255u) << (8 * (j % 4));
    }
    unsigned char *slab = s_buf + buf * kSlabBytes;
#pragma unroll
    for (int s = 0; s < kSl; ++s)
      *reinterpret_cast<uint2 *>(slab + (s * kRows + ar) * kStageK + 8 * q) = make_uint2(pk[s][0], pk[s][1]);
    if (q == 0) s_rs[buf][ar] = ldexp(1.0, -(e + 7));
  };

  // coefficient fragments through a buffer resource over the workgroup's
  // 4*CBW column blocks; a dead block's loads get an offset past the end,
  // so they return zeros without a memory access
  const int blk_bytes = nst * kSl * 1024;
  const __amdgpu_buffer_rsrc_t brs = __builtin_amdgcn_make_buffer_rsrc(
      (void *)(Bq + gb0 * (int64_t)blk_bytes), (short)0, 4 * CBW * blk_bytes, 0x00020000);
  auto load_b = [&](v4i (&b)[kSl], int st, int cb, bool live) {
    const int vo = live ? 16 * lane : (int)0x80000000;
    const int so = (cq + 4 * cb) * blk_bytes + st * (kSl * 1024);
#pragma unroll
    for (int s = 0; s < kSl; ++s)
      b[s] = __builtin_bit_cast(v4i, __builtin_amdgcn_raw_buffer_load_b128(brs, vo, so + s * 1024, 0));
  };
  auto wave_live = [&](int s) -> uint32_t { return (uint32_t)__builtin_amdgcn_readfirstlane(s_live[s]) >> cq; };
  auto stage_of = [&](int s) -> int { return __builtin_amdgcn_readfirstlane(s_list[s]); };

  double acc[2][CBW][4];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[r][cb][j] = 0.0;

  if (count > 0) load_a((int64_t)stage_of(0) * kStageK);
  for (int s = 0; s < count; ++s) {
    const int buf = s & 1;
    const int st = stage_of(s);
    const uint32_t live = wave_live(s);
    v4i bc[kSl];
    load_b(bc, st, 0, (live & 1u) != 0);  // in flight while the slab is cut
    write_slab(buf, (uint32_t)__builtin_amdgcn_readfirstlane(s_live[s]) >> 30);
    __syncthreads();
    if (s + 1 < count) load_a((int64_t)stage_of(s + 1) * kStageK);
    const unsigned char *slab = s_buf + buf * kSlabBytes;
    v4i a[2][kSl];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int sl = 0; sl < kSl; ++sl)
        a[r][sl] = *reinterpret_cast<const v4i *>(slab + (sl * kRows + 16 * r + li) * kStageK + 16 * lg);
    double rs[2][4];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int j = 0; j < 4; ++j) rs[r][j] = s_rs[buf][16 * r + 4 * lg + j];
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      v4i bn[kSl];
      if (cb + 1 < CBW) load_b(bn, st, cb + 1, ((live >> (4 * (cb + 1))) & 1u) != 0);
      if ((live >> (4 * cb)) & 1u) {
        // d[r][L]: the products of order L (slice i of A times slice L - i
        // of C), issued round by round over both row blocks, so that
        // consecutive MFMAs feed different sums
        const v4i z = {0, 0, 0, 0};
        v4i d[2][kSl];
#pragma unroll
        for (int L = 0; L < kSl; ++L)
#pragma unroll
          for (int r = 0; r < 2; ++r) d[r][L] = mfma8(a[r][0], bc[L], z);
#pragma unroll
        for (int i = 1; i < kSl; ++i)
#pragma unroll
          for (int L = i; L < kSl; ++L)
#pragma unroll
            for (int r = 0; r < 2; ++r) d[r][L] = mfma8(a[r][i], bc[L - i], d[r][L]);
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int e0 = d[r][0][j] * 128 + d[r][1][j];
            const int e1 = d[r][2][j] * 128 + d[r][3][j];
            double x = fma((double)e1, 0x1p-14, (double)e0);
            x = fma((double)d[r][4][j], 0x1p-21, x);
            acc[r][cb][j] = fma(x, rs[r][j], acc[r][cb][j]);
          }
      }
      if (cb + 1 < CBW) {
#pragma unroll
        for (int sl = 0; sl < kSl; ++sl) bc[sl] = bn[sl];
      }
    }
  }

  // epilogue: 16 rows at a time through LDS, in MO order, MO scales applied
  const int64_t g_lo = gb0 * 16;
  const int g_n = (int)(Ncol - g_lo < 64 * CBW ? Ncol - g_lo : 64 * CBW);
  double *so = reinterpret_cast<double *>(s_buf);
  int mloc[CBW];
  double ci[CBW];
#pragma unroll
  for (int cb = 0; cb < CBW; ++cb) {
    const int c = (cq + 4 * cb) * 16 + li;
    mloc[cb] = s_mloc[c];
    ci[cb] = (double)s_ci[c];
  }
  const bool vec = (ldc % 2) == 0 && (g_n % 2) == 0 && (reinterpret_cast<uintptr_t>(C) % 16) == 0;
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    __syncthreads();  // the slab (or the previous pass) has been read
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (mloc[cb] >= 0) so[(4 * lg + j) * (64 * CBW) + mloc[cb]] = acc[pass][cb][j] * ci[cb];
    __syncthreads();
    for (int row = 0; row < 16; ++row) {
      const int32_t p = s_perm[pass * 16 + row];
      if (p < 0) break;  // rows past the last point
      double *dst = C + ((int64_t)p * 5 + comp) * ldc + g_lo;
      if (vec) {
        typedef double d2 __attribute__((ext_vector_type(2)));
        for (int cp = tid; cp < g_n / 2; cp += kThreads) {
          const d2 v = {so[row * (64 * CBW) + 2 * cp], so[row * (64 * CBW) + 2 * cp + 1]};
          __builtin_nontemporal_store(v, reinterpret_cast<d2 *>(dst + 2 * cp));
        }
      } else {
        for (int c = tid; c < g_n; c += kThreads) dst[c] = so[row * (64 * CBW) + c];
      }
    }
  }
}

template <int CBW>
hipError_t launch_i8(hipStream_t s, int64_t P, int64_t npt, int64_t K, int64_t Ncol, const float *A,
                     const float *Alo, int64_t lda, const int8_t *Bq, int64_t ldb, const float *colinv,
                     const uint8_t *flags, int nkb, const int32_t *perm, double *C, int64_t ldc, ColPlan plan) {
  const int64_t gx = 5 * npt * 2, gy = (Ncol + 64 * CBW - 1) / (64 * CBW);
  const int64_t nst = ldb / kStageK;
  if (gx > 0x7fffffff || gy > 65535) return hipErrorInvalidValue;
  if ((K + kStageK - 1) / kStageK > kMaxSt) return hipErrorInvalidValue;  // the stage list lives in LDS
  // the workgroup's coefficient buffer is addressed with 32-bit offsets
  if ((int64_t)4 * CBW * nst * kSl * 1024 >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
  hipLaunchKernelGGL((gemm_i8_kernel<CBW>), dim3((unsigned)gx, (unsigned)gy), dim3(kThreads), 0, s, P, npt, K, Ncol,
                     A, Alo, lda, Bq, ldb, colinv, flags, nkb, perm, C, ldc, plan);
  return hipGetLastError();
}

}  // namespace

int64_t slice_coef_i8_bytes(int64_t M, int64_t ldb) {
  return (int64_t)kSl * split_cols_padded(QMCKL_AMD_INT8X5, M) * ldb;
}

hipError_t launch_slice_coef_i8(hipStream_t s, int64_t M, int64_t K, const double *coef, int64_t ldb, int8_t *Bq,
                                float *colinv, const int32_t *col_mo) {
  if (M <= 0) return hipSuccess;
  if (ldb % kStageK != 0 || ldb < K) return hipErrorInvalidValue;
  const int64_t g = split_cols_padded(QMCKL_AMD_INT8X5, M);
  hipLaunchKernelGGL(slice_coef_i8_kernel, dim3((unsigned)g), dim3(kThreads), 0, s, M, K, coef, ldb, col_mo, Bq,
                     colinv);
  return hipGetLastError();
}

hipError_t launch_mo_gemm_i8(hipStream_t s, int64_t P, int64_t K, int64_t Ncol, const float *A, const float *Alo,
                             int64_t lda, const int8_t *Bq, int64_t ldb, const float *colinv, const uint8_t *flags,
                             int nkb, const int32_t *perm, double *C, int64_t ldc, ColPlan plan) {
  if (P <= 0 || Ncol <= 0) return hipSuccess;
  if (K <= 0 || A == nullptr || Alo == nullptr || (lda % kStageK) != 0 || ldb != lda || lda < K)
    return hipErrorInvalidValue;
  if (flags != nullptr && (nkb > kMaxFlagBlocks || (int64_t)nkb * 32 < K)) return hipErrorInvalidValue;
  if (plan.blk != nullptr && (plan.nw * 64 * 16 < K || plan.nblk < 0)) return hipErrorInvalidValue;
  const int64_t npt = (P + 63) / 64;
  switch (split_wg_cols(QMCKL_AMD_INT8X5, Ncol) / 64) {
    case 1: return launch_i8<1>(s, P, npt, K, Ncol, A, Alo, lda, Bq, ldb, colinv, flags, nkb, perm, C, ldc, plan);
    case 2: return launch_i8<2>(s, P, npt, K, Ncol, A, Alo, lda, Bq, ldb, colinv, flags, nkb, perm, C, ldc, plan);
    case 3: return launch_i8<3>(s, P, npt, K, Ncol, A, Alo, lda, Bq, ldb, colinv, flags, nkb, perm, C, ldc, plan);
    case 4: return launch_i8<4>(s, P, npt, K, Ncol, A, Alo, lda, Bq, ldb, colinv, flags, nkb, perm, C, ldc, plan);
    case 5: return launch_i8<5>(s, P, npt, K, Ncol, A, Alo, lda, Bq, ldb, colinv, flags, nkb, perm, C, ldc, plan);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace qmckl_amd
