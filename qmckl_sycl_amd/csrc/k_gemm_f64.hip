// FP64 MO contraction on the matrix cores.
//
// mo[(p,c)][i] = sum_k ao[(p,c)][k] * coefficient_t[k][i]
// (src/qmckl_mo_sycl.cpp:23-50: per point, k ascending, 5 components,
// coefficient_t [A][M] from 176-184) is a row-major GEMM
//   C[R x Ncol] = A[R x K] * B[K x Ncol],  R = 5P (or P for values).
//
// v_mfma_f64_16x16x4_f64: lane l holds A[l%16][l/16] and B[l/16][l%16] and
// receives D[(l/16)+4j][l%16] in its j-th register pair, j = 0..3 (measured
// on gfx950: rows interleave across the four lane groups).  A workgroup (4 wavefronts) owns a
// 64-row x (64*CBW)-column tile: the 64x16 A slab is shared by the four
// wavefronts through LDS (stored k-major so a fragment read is 16
// consecutive doubles), and wavefront w owns column blocks w, w+4, ...; the
// 16 x 64*CBW slab of B goes through LDS too so every global load is a
// contiguous row segment.  With CBW = 5 one workgroup covers 320 MOs, so
// for mo_num <= 320 each AO row is read from HBM once.  With nonzero flags
// from the AO kernel, K tiles whose AOs are all screened on the 64 rows are
// skipped (adding exact zeros changes nothing).
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace qmckl_amd {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kBM = 64;
constexpr int kBK = 16;
constexpr int kThreads = 256;

// TRANSA: the A operand is stored K-major (A[k*lda + row]); the LDS slab is
// k-major either way.  gridDim.z splits K into chunks of kchunk; split z
// writes its partial product to C + z*R*ldc.
template <int CBW, bool TRANSA>
__global__ void __launch_bounds__(kThreads)
gemm_f64_kernel(int64_t R, int64_t K, int64_t Ncol, const double *__restrict__ A, int64_t lda,
                const double *__restrict__ B, int64_t ldb, double *__restrict__ C, int64_t ldc,
                const uint8_t *__restrict__ flags, int nkb, int64_t kchunk) {
  constexpr int NB = 4 * CBW;   // 16-column blocks per workgroup
  constexpr int BN = 16 * NB;
  __shared__ double As[kBK][kBM + 1];
  __shared__ double Bs[kBK][BN];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t row0 = (int64_t)blockIdx.x * kBM;
  const int64_t col0 = (int64_t)blockIdx.y * BN;
  const int li = lane & 15, lk = lane >> 4;

  d4 acc[4][CBW];
#pragma unroll
  for (int rb = 0; rb < 4; ++rb)
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) acc[rb][cb] = d4{0.0, 0.0, 0.0, 0.0};

  const int64_t kbeg = (int64_t)blockIdx.z * kchunk;
  const int64_t kend = kbeg + kchunk < K ? kbeg + kchunk : K;
  C += (int64_t)blockIdx.z * R * ldc;
  for (int64_t k0 = kbeg; k0 < kend; k0 += kBK) {
    // skip K tiles whose 32-AO block is zero on all 64 rows (uniform branch)
    if (flags != nullptr && flags[(int64_t)blockIdx.x * nkb + (k0 >> 5)] == 0) continue;
    for (int i = tid; i < kBM * kBK; i += kThreads) {
      if (TRANSA) {
        const int kk = i / kBM, r = i - kk * kBM;
        const int64_t gr = row0 + r, gk = k0 + kk;
        As[kk][r] = (gr < R && gk < kend) ? A[gk * lda + gr] : 0.0;
      } else {
        const int r = i / kBK, kk = i - r * kBK;
        const int64_t gr = row0 + r, gk = k0 + kk;
        As[kk][r] = (gr < R && gk < kend) ? A[gr * lda + gk] : 0.0;
      }
    }
    for (int i = tid; i < kBK * BN; i += kThreads) {
      const int kk = i / BN, cc = i - kk * BN;
      const int64_t gk = k0 + kk, gc = col0 + cc;
      Bs[kk][cc] = (gk < kend && gc < Ncol) ? B[gk * ldb + gc] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 4) {
      double a[4];
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) a[rb] = As[kk + lk][rb * 16 + li];
#pragma unroll
      for (int cb = 0; cb < CBW; ++cb) {
        const double b = Bs[kk + lk][(wave + 4 * cb) * 16 + li];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
          acc[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb], b, acc[rb][cb], 0, 0, 0);
      }
    }
    __syncthreads();
  }

#pragma unroll
  for (int cb = 0; cb < CBW; ++cb) {
    const int64_t col = col0 + (wave + 4 * cb) * 16 + li;
    if (col >= Ncol) continue;
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t row = row0 + rb * 16 + lk + 4 * j;
        if (row < R) C[row * ldc + col] = acc[rb][cb][j];
      }
  }
}

template <int CBW, bool TRANSA>
hipError_t launch_cbw(hipStream_t s, int64_t R, int64_t K, int64_t Ncol, const double *A,
                      int64_t lda, const double *B, int64_t ldb, double *C, int64_t ldc,
                      const uint8_t *flags, int nkb, int splits) {
  constexpr int BN = 64 * CBW;
  const int64_t gx = (R + kBM - 1) / kBM, gy = (Ncol + BN - 1) / BN;
  if (gx > 0x7fffffff || gy > 65535 || splits < 1 || splits > 65535) return hipErrorInvalidValue;
  // K chunk per split, a multiple of the 32-wide flag blocks
  const int64_t kchunk = ((K + splits - 1) / splits + 31) / 32 * 32;
  hipLaunchKernelGGL((gemm_f64_kernel<CBW, TRANSA>), dim3((unsigned)gx, (unsigned)gy, (unsigned)splits),
                     dim3(kThreads), 0, s, R, K, Ncol, A, lda, B, ldb, C, ldc, flags, nkb, kchunk);
  return hipGetLastError();
}

template <bool TRANSA>
hipError_t dispatch(hipStream_t s, int64_t R, int64_t K, int64_t Ncol, const double *A, int64_t lda,
                    const double *B, int64_t ldb, double *C, int64_t ldc, const uint8_t *flags,
                    int nkb, int splits) {
  if (R <= 0 || Ncol <= 0) return hipSuccess;
  if (K <= 0) return hipErrorInvalidValue;
  const int64_t nb = (Ncol + 15) / 16;  // 16-column blocks needed
  if (nb <= 4) return launch_cbw<1, TRANSA>(s, R, K, Ncol, A, lda, B, ldb, C, ldc, flags, nkb, splits);
  if (nb <= 8) return launch_cbw<2, TRANSA>(s, R, K, Ncol, A, lda, B, ldb, C, ldc, flags, nkb, splits);
  if (nb <= 12) return launch_cbw<3, TRANSA>(s, R, K, Ncol, A, lda, B, ldb, C, ldc, flags, nkb, splits);
  if (nb <= 16) return launch_cbw<4, TRANSA>(s, R, K, Ncol, A, lda, B, ldb, C, ldc, flags, nkb, splits);
  return launch_cbw<5, TRANSA>(s, R, K, Ncol, A, lda, B, ldb, C, ldc, flags, nkb, splits);
}

// out[i] = sum_z part[z][i], z ascending (fixed order: deterministic)
__global__ void __launch_bounds__(kThreads)
sum_splits_kernel(int64_t n, int splits, const double *__restrict__ part, double *__restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double t = 0.0;
    for (int z = 0; z < splits; ++z) t += part[(int64_t)z * n + i];
    out[i] = t;
  }
}

}  // namespace

hipError_t launch_mo_gemm_f64(hipStream_t s, int64_t R, int64_t K, int64_t Ncol, const double *A,
                              int64_t lda, const double *B, int64_t ldb, double *C, int64_t ldc,
                              const uint8_t *flags, int nkb) {
  return dispatch<false>(s, R, K, Ncol, A, lda, B, ldb, C, ldc, flags, nkb, 1);
}

int gemm_tn_splits(int64_t K) {
  int64_t z = K / 2048;
  return (int)(z < 1 ? 1 : (z > 64 ? 64 : z));
}

hipError_t launch_gemm_tn_f64(hipStream_t s, int64_t R, int64_t K, int64_t Ncol, const double *A,
                              int64_t lda, const double *B, int64_t ldb, double *C, int64_t ldc,
                              double *partial, int splits) {
  if (R <= 0 || Ncol <= 0) return hipSuccess;
  if (ldc != Ncol) return hipErrorInvalidValue;
  if (splits <= 1) return dispatch<true>(s, R, K, Ncol, A, lda, B, ldb, C, ldc, nullptr, 0, 1);
  hipError_t e = dispatch<true>(s, R, K, Ncol, A, lda, B, ldb, partial, ldc, nullptr, 0, splits);
  if (e != hipSuccess) return e;
  const int64_t n = R * Ncol;
  int64_t g = (n + kThreads - 1) / kThreads;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(sum_splits_kernel, dim3((unsigned)g), dim3(kThreads), 0, s, n, splits, partial, C);
  return hipGetLastError();
}

}  // namespace qmckl_amd
