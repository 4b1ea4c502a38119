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
// That kernel serves the low-level launchers (coefficients given as a plain
// [A][M] array); the mo_vgl / mo_value providers and the adjoint use
// gemm_f64_frag_kernel below, which takes its B operand in MFMA fragment
// order (36.0 -> 14.3 ms at 409,600 x 4,160 x 320 on MI355X).
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

// The forward MO contraction with the coefficients in MFMA fragment order.
// For v_mfma_f64_16x16x4_f64 lane l holds B[4*ks + l/16][16*cb + l%16] of
// column block cb at K step ks; the prepared copy stores that fragment at
//   Bf[(cb * nk4 + ks) * 64 + l],   nk4 = K padded to 16, over 4
// (zeros past K and past the last MO), so a wavefront loads each fragment
// as 512 contiguous bytes straight from L2 into registers.  Only the AO
// slab goes through LDS (all four wavefronts use its rows).  A workgroup
// owns 64 rows x 64*CBW columns (wavefront w: column blocks w, w+4, ...);
// K advances in stages of 16 with the next stage's AO slab and coefficient
// fragments requested while the current stage computes; stages whose 32-AO
// block is screened on all 64 rows are dropped from the stage list up front.
// With a column plan each listed stage also carries the live bits of the
// workgroup's column blocks: a wavefront issues neither the loads nor the
// MFMAs of a column block whose coefficients are zero on the stage's 16
// AOs, and a stage with no live block is dropped.  Adding exact zeros
// changes nothing, so results equal the dense kernel's bit for bit; the
// plan's column order only moves whole columns.
constexpr int kMaxStages16 = 2 * kMaxFlagBlocks;
constexpr int kOutRows = 8;  // output rows per pass of the planned-order epilogue

__global__ void __launch_bounds__(kThreads)
frag_b_f64_kernel(int64_t K, int64_t Ncol, const double *__restrict__ B, int64_t ldb, int64_t nk4,
                  int64_t total, const int32_t *__restrict__ col_mo, const int32_t *__restrict__ row_map,
                  double *__restrict__ Bf) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t frag = i >> 6;
    const int l = (int)(i & 63);
    const int64_t cb = frag / nk4, ks = frag - cb * nk4;
    const int64_t k = 4 * ks + l / 16, col = 16 * cb + l % 16;
    Bf[i] = (k < K && col < Ncol) ? B[(row_map != nullptr ? (int64_t)row_map[k] : k) * ldb +
                                      (col_mo != nullptr ? (int64_t)col_mo[col] : col)]
                                  : 0.0;
  }
}

// out[r][j] = in[rmap[r]][cmap[j]] (a null map: the identity)
__global__ void __launch_bounds__(kThreads)
gather_kernel(int64_t rows, int64_t cols, const double *__restrict__ in, int64_t ld_in,
              const int32_t *__restrict__ rmap, const int32_t *__restrict__ cmap, double *__restrict__ out,
              int64_t ld_out) {
  const int64_t n = rows * cols;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / cols, j = i - r * cols;
    const int64_t rs = rmap != nullptr ? (int64_t)rmap[r] : r;
    out[r * ld_out + j] = in[rs * ld_in + (cmap != nullptr ? (int64_t)cmap[j] : j)];
  }
}

// one wavefront per (MO, 64 AO blocks of 16): lane s tests its 16 AOs
__global__ void __launch_bounds__(kThreads)
coef_blocks_kernel(int64_t M, int64_t K, const double *__restrict__ coef, int64_t nw,
                   uint64_t *__restrict__ words) {
  const int lane = threadIdx.x & 63;
  const int64_t wid = (int64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  if (wid >= M * nw) return;  // whole wavefronts
  const int64_t m = wid / nw, s = (wid - m * nw) * 64 + lane;
  bool nz = false;
  for (int j = 0; j < 16; ++j) {
    const int64_t k = 16 * s + j;
    if (k < K && !(coef[m * K + k] == 0.0)) nz = true;
  }
  const uint64_t b = __ballot(nz);
  if (lane == 0) words[wid] = b;
}

// TRANSA: the A operand is stored K-major (A[k*lda + row]).  gridDim.z
// splits K into chunks of kchunk (a multiple of 16); split z writes its
// partial product to C + z*R*ldc.  flags are for the unsplit provider path.
// tflags (tnkb bytes per tile of 320 rows of the 64-point tiles' (point,
// component) rows): with TRANSA the K rows' tiles, flags over the output
// rows in 16-row blocks: a stage's rows are neither loaded nor multiplied
// into the 16-row blocks whose flag is zero, and a stage with none left is
// dropped; without, the output rows' tiles, flags over the output columns
// in 16-column blocks: a block whose flag is zero is neither computed nor
// written (its entries keep what they held; the caller reads none of
// them).
template <int CBW, bool TRANSA>
__global__ void __launch_bounds__(kThreads, 2)
gemm_f64_frag_kernel(int64_t R, int64_t K, int64_t Ncol, const double *__restrict__ A, int64_t lda,
                     const double *__restrict__ Bf, int64_t nk4, double *__restrict__ C, int64_t ldc,
                     const uint8_t *__restrict__ flags, int nkb, int64_t kchunk, ColPlan plan,
                     const int32_t *__restrict__ kmap, const uint8_t *__restrict__ tflags, int tnkb) {
  __shared__ double As[kBK][kBM + 1];
  __shared__ int s_list[kMaxStages16];
  __shared__ uint32_t s_live[kMaxStages16];
  __shared__ int s_count;
  __shared__ double s_out[TRANSA ? 1 : kOutRows * 64 * CBW];  // output rows in MO order (planned columns)

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wv = __builtin_amdgcn_readfirstlane(wave);
  const int li = lane & 15, lk = lane >> 4;
  // The workgroup's tile (bx, by, bz).  With a K split the tiles are
  // numbered so that each XCD takes consecutive ones: the workgroups are
  // handed to the 8 XCDs in turn by their linear index, and with the tiles
  // in grid order the row tiles of one K chunk, which all read that
  // chunk's B fragments, would be spread over every XCD's L2 at once.
  int64_t bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
  if (TRANSA && gridDim.z > 1) {
    const int64_t n = (int64_t)gridDim.x * gridDim.y * gridDim.z;
    const int64_t lin = blockIdx.x + (int64_t)gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z);
    const int64_t q = n / 8, r = n % 8, xcd = lin % 8, slot = lin / 8;
    const int64_t t = xcd < r ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
    bx = t % gridDim.x;
    by = (t / gridDim.x) % gridDim.y;
    bz = t / ((int64_t)gridDim.x * gridDim.y);
  }
  const int64_t row0 = bx * kBM;
  const int64_t gb0 = by * 4 * CBW;  // the workgroup's first column block
  const int64_t cb0 = gb0 + wv;      // this wavefront's first column block
  const int64_t kbeg = bz * kchunk;
  const int64_t kend = kbeg + kchunk < K ? kbeg + kchunk : K;
  const int nst = kend > kbeg ? (int)((kend - kbeg + kBK - 1) / kBK) : 0;
  C += bz * R * ldc;

  // output column blocks gb0 + b that the tile's flags keep
  uint32_t omask = 0xffffffffu;
  if (!TRANSA && tflags != nullptr) {
    const uint8_t *tf = tflags + (row0 / 320) * tnkb;
    omask = 0;
#pragma unroll
    for (int b = 0; b < 4 * CBW; ++b)
      if ((gb0 + b) * 16 < Ncol && tf[gb0 + b] != 0) omask |= 1u << b;
  }
  // stage list: live stages and, per stage, bit b for live column block gb0 + b
  const bool listed = flags != nullptr || plan.blk != nullptr || tflags != nullptr;
  int count = nst;
  if (listed) {
    if (wave == 0) {
      const uint8_t *f = flags != nullptr ? flags + bx * nkb : nullptr;
      const int64_t g0 = kbeg / kBK;  // 16-AO block of the chunk's first stage
      int base = 0;
      for (int s0 = 0; s0 < nst; s0 += 64) {
        const int st = s0 + lane;
        bool live = st < nst && (f == nullptr || f[st >> 1] != 0);
        uint32_t lv = omask;
        if (TRANSA && tflags != nullptr && live) {
          // the K rows of stage st lie in one 320-row tile (20 stages); the
          // workgroup's 64 output rows are 16-row blocks 4 bx .. 4 bx + 3;
          // lv: bit rb for live row block rb
          const uint8_t *tf = tflags + ((g0 + st) / 20) * tnkb;
          const int64_t b0 = 4 * bx;
          lv = 0;
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (b0 + q < tnkb && tf[b0 + q] != 0) lv |= 1u << q;
          live = lv != 0;
        }
        if (!TRANSA && tflags != nullptr) live = live && omask != 0;
        if (live && plan.blk != nullptr) {
          // all loads first (independent, no branches), then the bits
          const int64_t g = g0 + st;
          uint64_t w[4 * CBW];
#pragma unroll
          for (int b = 0; b < 4 * CBW; ++b) {
            const int64_t blk = gb0 + b < plan.nblk ? gb0 + b : plan.nblk - 1;
            w[b] = plan.blk[blk * plan.nw + (g >> 6)];
          }
          uint32_t bits = 0;
#pragma unroll
          for (int b = 0; b < 4 * CBW; ++b)
            bits |= (gb0 + b < plan.nblk ? (uint32_t)((w[b] >> (g & 63)) & 1ull) : 0u) << b;
          lv &= bits;
          live = lv != 0;
        }
        const uint64_t m = __ballot(live);
        if (live) {
          const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
          s_list[pos] = st;
          s_live[pos] = lv;
        }
        base += __popcll(m);
      }
      if (lane == 0) s_count = base;
    }
    __syncthreads();
    count = s_count;
  }
  // this wavefront's column blocks that hold columns: bit 4j for block
  // cb0 + 4j (the zero padding of the fragments past Ncol is neither
  // loaded nor multiplied)
  const int64_t nvb = (Ncol + 15) / 16 - gb0;
  uint32_t vmask = 0;
#pragma unroll
  for (int cb = 0; cb < CBW; ++cb)
    if (wv + 4 * cb < nvb) vmask |= 1u << (4 * cb);
  // this wavefront's live column blocks at stage s: bit 4j for block cb0 + 4j
  // (TRANSA: all that hold columns, and the live 16-row blocks in row_live)
  auto wave_live = [&](int s) -> uint32_t {
    if (TRANSA || !listed) return vmask;
    return ((uint32_t)__builtin_amdgcn_readfirstlane(s_live[s]) >> wv) & vmask;
  };
  auto row_live = [&](int s) -> uint32_t {
    if (!TRANSA || tflags == nullptr) return 0xfu;
    return (uint32_t)__builtin_amdgcn_readfirstlane(s_live[s]);
  };

  // A slot: row ar, 4 consecutive K at aq (K-major storage: 4 consecutive
  // rows ar4.. at K aq4); rows past R read zero or row R-1, K past the
  // chunk reads zero
  const int ar = tid >> 2, aq = (tid & 3) * 4;
  const int aq4 = tid >> 4, ar4 = (tid & 15) * 4;
  const double *Ap = TRANSA ? A + row0 + ar4 : A + (row0 + ar < R ? row0 + ar : R - 1) * lda + aq;
  const double *Bp = Bf + cb0 * nk4 * 64 + lane;
  const int64_t cb_stride = 4 * nk4 * 64;  // the wavefront's next column block

  double ra[4];
  double b[4][CBW];
  auto stage_k0 = [&](int s) {
    return kbeg + (int64_t)__builtin_amdgcn_readfirstlane(listed ? s_list[s] : s) * kBK;
  };
  // with a K-row map, this thread's mapped row of the stage after the next
  // one is fetched a stage ahead, so the slab load does not wait on it
  int32_t kr_next = 0;
  auto map_row = [&](int s) -> int32_t {
    if (kmap == nullptr || s >= count) return 0;
    const int64_t k = stage_k0(s) + aq4;
    return k < kend ? kmap[k] : 0;
  };
  auto load_a = [&](int64_t k0, int32_t kr_mapped, uint32_t rows) {
    if constexpr (TRANSA) {
      const bool kin = k0 + aq4 < kend && ((rows >> (ar4 >> 4)) & 1u);  // rows of a dead 16-row block read zero
      const int64_t kr = kmap != nullptr ? (int64_t)kr_mapped : k0 + aq4;
#pragma unroll
      for (int j = 0; j < 4; ++j) ra[j] = kin && row0 + ar4 + j < R ? Ap[kr * lda + j] : 0.0;
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) ra[j] = k0 + aq + j < kend ? Ap[k0 + j] : 0.0;
    }
  };
  auto store_a = [&]() {
    if constexpr (TRANSA) {
#pragma unroll
      for (int j = 0; j < 4; ++j) As[aq4][ar4 + j] = ra[j];
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) As[aq + j][ar] = ra[j];
    }
  };
  // The forward product loads its fragments through a buffer resource over
  // the workgroup's 4*CBW column blocks (lane offset 8*l bytes, block and
  // step offset in a scalar register).  A dead block gets a lane offset
  // past the end of the buffer, so its load returns zero without a memory
  // access, and every stage issues the same loads (the waits on them stay
  // exact).
  const int blk_bytes = (int)(nk4 * 512);
  const __amdgpu_buffer_rsrc_t brs =
      __builtin_amdgcn_make_buffer_rsrc((void *)(Bf + gb0 * nk4 * 64), (short)0, TRANSA ? 0 : 4 * CBW * blk_bytes,
                                        0x00020000);
  auto load_b = [&](int64_t k0, int ks, uint32_t live) {
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      if constexpr (TRANSA) {
        // loaded whatever the live bits (a conditional load measured 0.2 ms
        // slower on the adjoint's dC product); the fragments past Ncol are
        // zero padding, in bounds, and not multiplied
        b[ks][cb] = Bp[cb * cb_stride + (k0 / 4 + ks) * 64];
      } else {
        const int vo = (live >> (4 * cb)) & 1u ? 8 * lane : (int)0x80000000;
        const int so = (wv + 4 * cb) * blk_bytes + (int)(k0 / 4 + ks) * 512;
        b[ks][cb] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(brs, vo, so, 0));
      }
    }
  };

  d4 acc[4][CBW];
#pragma unroll
  for (int rb = 0; rb < 4; ++rb)
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) acc[rb][cb] = d4{0.0, 0.0, 0.0, 0.0};

  uint32_t live0 = 0, rows0 = 0;
  if (count > 0) {
    const int64_t k0 = stage_k0(0);
    live0 = wave_live(0);
    rows0 = row_live(0);
    load_a(k0, map_row(0), rows0);
    kr_next = map_row(1);
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) load_b(k0, ks, live0);
  }
  for (int s = 0; s < count; ++s) {
    __syncthreads();  // every wavefront is done reading the previous AO slab
    store_a();
    __syncthreads();
    const bool more = s + 1 < count;
    const int64_t k1 = more ? stage_k0(s + 1) : 0;
    const uint32_t live1 = more ? wave_live(s + 1) : 0u;
    const uint32_t rows1 = more ? row_live(s + 1) : 0u;
    if (more) {
      load_a(k1, kr_next, rows1);
      kr_next = map_row(s + 2);
    }
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      double a[4];
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) a[rb] = As[4 * ks + lk][rb * 16 + li];
#pragma unroll
      for (int cb = 0; cb < CBW; ++cb)
        if ((live0 >> (4 * cb)) & 1u) {
#pragma unroll
          for (int rb = 0; rb < 4; ++rb)
            if (!TRANSA || ((rows0 >> rb) & 1u))
              acc[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb], b[ks][cb], acc[rb][cb], 0, 0, 0);
        }
      if (more) load_b(k1, ks, live1);  // this K step's fragments are free again
    }
    live0 = live1;
    rows0 = rows1;
  }
  if (TRANSA || plan.col_mo == nullptr) {
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      const int64_t col = (cb0 + 4 * cb) * 16 + li;
      if (col >= Ncol || !((omask >> (wv + 4 * cb)) & 1u)) continue;
#pragma unroll
      for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int64_t row = row0 + rb * 16 + lk + 4 * j;
          if (row < R) C[row * ldc + col] = acc[rb][cb][j];
        }
    }
    return;
  }
  // Planned column order: the workgroup's columns are the MOs of its group
  // [g_lo, g_lo + g_n) in another order.  Rows go through LDS kOutRows at a
  // time, put back in MO order, and are written as contiguous runs.
  const int64_t g_lo = gb0 * 16;
  const int g_n = (int)(Ncol - g_lo < 64 * CBW ? Ncol - g_lo : 64 * CBW);
  int mloc[CBW];
#pragma unroll
  for (int cb = 0; cb < CBW; ++cb) {
    const int64_t col = (cb0 + 4 * cb) * 16 + li;
    mloc[cb] = col < Ncol ? (int)(plan.col_mo[col] - g_lo) : -1;
  }
  for (int pass = 0; pass < kBM / kOutRows; ++pass) {
    __syncthreads();  // the previous pass has been read out
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb)
#pragma unroll
      for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int row = rb * 16 + lk + 4 * j;
          if (mloc[cb] >= 0 && row / kOutRows == pass) s_out[(row % kOutRows) * (64 * CBW) + mloc[cb]] = acc[rb][cb][j];
        }
    __syncthreads();
    for (int rl = 0; rl < kOutRows; ++rl) {
      const int64_t row = row0 + pass * kOutRows + rl;
      if (row >= R) break;
      double *dst = C + row * ldc + g_lo;
      for (int c = tid; c < g_n; c += kThreads) dst[c] = s_out[rl * (64 * CBW) + c];
    }
  }
}

template <int CBW, bool TRANSA>
hipError_t launch_frag(hipStream_t s, int64_t R, int64_t K, int64_t Ncol, const double *A, int64_t lda,
                       const double *Bf, int64_t nk4, double *C, int64_t ldc, const uint8_t *flags,
                       int nkb, int splits, ColPlan plan, const int32_t *kmap, const uint8_t *tflags,
                       int tnkb) {
  const int64_t gx = (R + kBM - 1) / kBM, gy = (Ncol + 64 * CBW - 1) / (64 * CBW);
  if (gx > 0x7fffffff || gy > 65535 || splits < 1 || splits > 65535) return hipErrorInvalidValue;
  if (flags != nullptr && ((int64_t)nkb * 32 < K || nkb > kMaxFlagBlocks || splits != 1))
    return hipErrorInvalidValue;
  const int64_t kchunk = ((K + splits - 1) / splits + kBK - 1) / kBK * kBK;
  if (plan.blk != nullptr &&
      ((kchunk + kBK - 1) / kBK > kMaxStages16 || plan.nw * 64 * kBK < K || plan.nblk < 0))
    return hipErrorInvalidValue;
  // the forward product addresses one workgroup's column blocks with 32-bit offsets
  if (!TRANSA && (int64_t)4 * CBW * nk4 * 512 >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
  if (TRANSA && (plan.col_mo != nullptr || plan.blk != nullptr)) return hipErrorInvalidValue;
  if (!TRANSA && (kmap != nullptr || (tflags != nullptr && plan.col_mo != nullptr))) return hipErrorInvalidValue;
  if (tflags != nullptr && ((int64_t)tnkb * 16 < (TRANSA ? R : Ncol) || (kchunk + kBK - 1) / kBK > kMaxStages16))
    return hipErrorInvalidValue;
  hipLaunchKernelGGL((gemm_f64_frag_kernel<CBW, TRANSA>), dim3((unsigned)gx, (unsigned)gy, (unsigned)splits),
                     dim3(kThreads), 0, s, R, K, Ncol, A, lda, Bf, nk4, C, ldc, flags, nkb, kchunk, plan, kmap,
                     tflags, tnkb);
  return hipGetLastError();
}

template <bool TRANSA>
hipError_t dispatch_frag(hipStream_t s, int64_t R, int64_t K, int64_t Ncol, const double *A, int64_t lda,
                         const double *Bf, int64_t nk4, double *C, int64_t ldc, const uint8_t *flags, int nkb,
                         int splits, ColPlan plan, const int32_t *kmap = nullptr,
                         const uint8_t *tflags = nullptr, int tnkb = 0) {
  switch (f64_frag_cbw(Ncol)) {
    case 1: return launch_frag<1, TRANSA>(s, R, K, Ncol, A, lda, Bf, nk4, C, ldc, flags, nkb, splits, plan, kmap, tflags, tnkb);
    case 2: return launch_frag<2, TRANSA>(s, R, K, Ncol, A, lda, Bf, nk4, C, ldc, flags, nkb, splits, plan, kmap, tflags, tnkb);
    case 3: return launch_frag<3, TRANSA>(s, R, K, Ncol, A, lda, Bf, nk4, C, ldc, flags, nkb, splits, plan, kmap, tflags, tnkb);
    case 4: return launch_frag<4, TRANSA>(s, R, K, Ncol, A, lda, Bf, nk4, C, ldc, flags, nkb, splits, plan, kmap, tflags, tnkb);
    default: return launch_frag<5, TRANSA>(s, R, K, Ncol, A, lda, Bf, nk4, C, ldc, flags, nkb, splits, plan, kmap, tflags, tnkb);
  }
}

// out[m][a] = sum_z part[z][a][m], z ascending (fixed order: deterministic).
// A workgroup sums a 64 x 64 tile of (a, m) with lanes along m (reads of
// part in whole rows), then writes it transposed through LDS with lanes
// along a (writes of out in whole rows).
__global__ void __launch_bounds__(kThreads)
sum_splits_t_kernel(int64_t rows, int64_t cols, int splits, const double *__restrict__ part,
                    double *__restrict__ out) {
  __shared__ double tile[64][65];
  const int64_t a0 = (int64_t)blockIdx.y * 64, m0 = (int64_t)blockIdx.x * 64;
  const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;  // 4 rows of 64 per pass
  const int64_t n = rows * cols;
  for (int r = ly; r < 64; r += kThreads / 64) {
    const int64_t a = a0 + r, m = m0 + lx;
    double t = 0.0;
    if (a < rows && m < cols)
      for (int z = 0; z < splits; ++z) t += part[(int64_t)z * n + a * cols + m];
    tile[r][lx] = t;
  }
  __syncthreads();
  for (int r = ly; r < 64; r += kThreads / 64) {
    const int64_t m = m0 + r, a = a0 + lx;
    if (m < cols && a < rows) out[m * rows + a] = tile[lx][r];
  }
}

}  // namespace

hipError_t launch_mo_gemm_f64(hipStream_t s, int64_t R, int64_t K, int64_t Ncol, const double *A,
                              int64_t lda, const double *B, int64_t ldb, double *C, int64_t ldc,
                              const uint8_t *flags, int nkb) {
  return dispatch<false>(s, R, K, Ncol, A, lda, B, ldb, C, ldc, flags, nkb, 1);
}

int f64_frag_cbw(int64_t Ncol) {
  const int64_t nb = (Ncol + 15) / 16;
  return nb <= 4 ? 1 : nb <= 8 ? 2 : nb <= 12 ? 3 : nb <= 16 ? 4 : 5;
}

hipError_t launch_frag_b_f64(hipStream_t s, int64_t K, int64_t Ncol, const double *B, int64_t ldb,
                             double *Bf, int64_t *nk4, const int32_t *col_mo, const int32_t *row_map) {
  *nk4 = (K + kBK - 1) / kBK * (kBK / 4);
  const int64_t npad = round_up(Ncol, 64 * f64_frag_cbw(Ncol));
  const int64_t total = (npad / 16) * (*nk4) * 64;
  if (total <= 0) return hipSuccess;
  int64_t g = (total + kThreads - 1) / kThreads;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(frag_b_f64_kernel, dim3((unsigned)g), dim3(kThreads), 0, s, K, Ncol, B, ldb, *nk4, total,
                     col_mo, row_map, Bf);
  return hipGetLastError();
}

hipError_t launch_gather(hipStream_t s, int64_t rows, int64_t cols, const double *in, int64_t ld_in,
                         const int32_t *rmap, const int32_t *cmap, double *out, int64_t ld_out) {
  const int64_t n = rows * cols;
  if (n <= 0) return hipSuccess;
  int64_t g = (n + kThreads - 1) / kThreads;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(gather_kernel, dim3((unsigned)g), dim3(kThreads), 0, s, rows, cols, in, ld_in, rmap, cmap,
                     out, ld_out);
  return hipGetLastError();
}

hipError_t launch_coef_blocks(hipStream_t s, int64_t M, int64_t K, const double *coef, int64_t nw,
                              uint64_t *words) {
  if (M <= 0 || nw <= 0) return hipSuccess;
  if (nw * 64 * 16 < K) return hipErrorInvalidValue;
  const int64_t g = (M * nw + kThreads / 64 - 1) / (kThreads / 64);
  if (g > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(coef_blocks_kernel, dim3((unsigned)g), dim3(kThreads), 0, s, M, K, coef, nw, words);
  return hipGetLastError();
}

hipError_t launch_mo_gemm_f64_frag(hipStream_t s, int64_t R, int64_t K, int64_t Ncol, const double *A,
                                   int64_t lda, const double *Bf, int64_t nk4, double *C, int64_t ldc,
                                   const uint8_t *flags, int nkb, ColPlan plan, const uint8_t *ocols, int onkb) {
  if (R <= 0 || Ncol <= 0) return hipSuccess;
  if (K <= 0 || nk4 * 4 < K) return hipErrorInvalidValue;
  return dispatch_frag<false>(s, R, K, Ncol, A, lda, Bf, nk4, C, ldc, flags, nkb, 1, plan, nullptr, ocols, onkb);
}

int gemm_tn_frag_splits(int64_t R, int64_t K) {
  // about 2048 workgroups, at least 64 K steps of 16 per split
  const int64_t tiles = (R + kBM - 1) / kBM;
  int64_t sp = (2048 + tiles - 1) / tiles;
  const int64_t maxsp = (K + 1023) / 1024;
  if (sp > maxsp) sp = maxsp;
  if (sp < 1) sp = 1;
  if (sp > 256) sp = 256;
  return (int)sp;
}

hipError_t launch_gemm_tn_f64_frag(hipStream_t s, int64_t R, int64_t K, int64_t Ncol, const double *A,
                                   int64_t lda, const double *Bf, int64_t nk4, double *out_t,
                                   double *partial, int splits, const int32_t *kmap, const uint8_t *tflags,
                                   int tnkb) {
  if (R <= 0 || Ncol <= 0) return hipSuccess;
  if (K <= 0 || nk4 * 4 < K || splits < 1) return hipErrorInvalidValue;
  hipError_t e = dispatch_frag<true>(s, R, K, Ncol, A, lda, Bf, nk4, partial, Ncol, nullptr, 0, splits, ColPlan(),
                                     kmap, tflags, tnkb);
  if (e != hipSuccess) return e;
  const int64_t gm = (Ncol + 63) / 64, ga = (R + 63) / 64;
  if (gm > 0x7fffffff || ga > 65535) return hipErrorInvalidValue;
  hipLaunchKernelGGL(sum_splits_t_kernel, dim3((unsigned)gm, (unsigned)ga), dim3(kThreads), 0, s, R, Ncol, splits,
                     partial, out_t);
  return hipGetLastError();
}

}  // namespace qmckl_amd
