// Mixed-precision MO contraction on the matrix cores.
//
// Same product as the FP64 kernel, mo[(p,c)][i] = sum_k ao[(p,c)][k] C[i][k]
// (src/qmckl_mo_sycl.cpp:23-50), with the AO operand stored in FP32
// (rounded once from the FP64 assembly) and the MO coefficients pre-split.
// Forms:
//   FP16X2: each operand is scaled by a power of two and split into fp16
//           hi + lo; three products hi*hi + hi*lo + lo*hi.  The AO scale is
//           per (row, 64-wide K stage), chosen in the kernel from the FP32
//           values; the C scale is per MO, from its max |coefficient|.
//   BF16X2: bf16 hi + lo, the same three products.
//   BF16X3: bf16 hi + mid + lo, six products (hh, hm, mh, hl, lh, mm).
//   FP32:   v_mfma_f32_16x16x4_f32 on the FP32 operands.
// The MFMA accumulates in FP32 over 64 consecutive K and is then added into
// FP64 accumulators, which bounds the FP32 accumulation error to that of
// short partial sums.  Power-of-two scales are exact.  K blocks whose AO
// tile is all zero for the workgroup's rows (screened AOs) are skipped using
// the nonzero flags the AO kernel writes.
//
// 16x16x32 operands: lane l holds A[l%16][8(l/16)..8(l/16)+7] and
// B[8(l/16)..+7][l%16] and receives D[4(l/16)+j][l%16], j = 0..3.
#include <hip/hip_runtime.h>


#include "internal.hpp"
#include "split_ops.hpp"

namespace qmckl_amd {

namespace {

constexpr int kThreads = 256;
constexpr int kBM = 64;
constexpr int kKB = 32;

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

// Split GEMM (FP16X2, BF16X2, BF16X3).
//   A      the FP32 AO operand in the blocked layout of launch_ao_rows_f32:
//          row tile rt = (component c, 64 points), [rt][lda/4][64][4]
//   rowmax [5 * tiles * 64] row maxima
//   Bs     split C terms (pre-scaled by their column scale) in the MFMA
//          fragment order of split_coef_kernel (frag = 16), MOs padded to
//          the workgroup width with zeros
//   colinv [Ncol] inverse column scales
//   flags  [5 * tiles][nkb] nonzero flags per (row tile, 32-wide K block),
//          or null
//   perm   the point of each tile row (row (c, i) holds point perm[i]),
//          or null for the identity
// Workgroup tile WR rows (one half of a 64-row tile, or all of it) x
// 64*CBW columns (all MOs for mo_num <= 320, so each FP32 AO row is read
// from HBM once), WR/8 wavefronts at two per SIMD: wavefront w owns rows
// [32*(w/4), +32) and the column blocks w%4, w%4 + 4, ... of the
// workgroup's 4*CBW blocks of 16, 2 x CBW MFMA tiles.
// Each row is scaled by the power of two that brings
// its largest |entry| into [2^13, 2^14).  K advances in stages of 64;
// stages whose two flags are zero are dropped from the workgroup's stage
// list up front.  With a column plan each listed stage carries, per 32-K
// half, the live bits of the workgroup's column blocks (nonzero
// coefficients on those AOs and, from the flags, nonzero AO values): the
// coefficient loads of a dead (block, half) return zeros without touching
// memory, and a stage with nothing live is dropped.  Per stage:
//   - the FP32 A slab (WR x 64, loaded into registers while the previous
//     stage computed) is scaled, split and written to LDS, shared by the
//     workgroup's wavefronts;
//   - each wavefront reads its C terms straight from L2 into registers (the
//     two wavefronts of a column quarter read the same bytes); in fragment
//     order every such load is one contiguous kilobyte; the next stage's
//     half-stage of C is requested as soon as the current one has been
//     consumed;
//   - 2 x 2 x CBW x (3 or 6) MFMAs accumulate in FP32 over the 64 K, and the
//     stage's sums are added into FP64 accumulators (unscaled, exactly, at
//     the store).
// The row (c, i) is written to mo_vgl[perm[i]][c][:].
// Measured on MI355X at 409,600 x 4,160 x 320 (water64, 128 walkers):
// storing C in fragment order instead of [MO][term][K] took the kernel from
// 5.6 to 4.2 ms (a load of 16 columns' 64-byte runs touched 16 cache
// lines); splitting each wavefront's 64 rows between two wavefronts per
// SIMD (the FP64 accumulators then fit 256 registers) from 4.3 to 3.7 ms;
// WR = 32 (two independent 4-wavefront workgroups per CU) equals WR = 64
// for FP16X2 and is 2.5% faster for BF16X3.
// With the points ordered into compact tiles the kernel ran 3.2 ms, and
// 1.9 ms with its global loads taken out (a timing variant): 32 cycles of
// wall time per 16x16x32 MFMA per SIMD, twice the 16 cycles each keeps the
// matrix pipe busy (SQ_VALU_MFMA_BUSY_CYCLES), so the loop leaves the pipe
// idle half the time (FP64 accumulation, A split, LDS reads) and the
// operand loads add what the two wavefronts per SIMD cannot hide.
// Measured slower: 64 x 96 tiles of v_mfma_f32_32x32x16 per wavefront at
// one per SIMD (MOs padded to 384; 4.0 ms, 2.7 ms without loads: 75 cycles
// per 32x32x16), 32 x 96 tiles of it at two per SIMD (3.9 ms), software-pipelined single-barrier loops (with and without
// interleaving hints), persistent workgroups, four-wavefront workgroups of
// 64 x 160 with the column halves on alternate XCDs, FP32 sums over two or
// four stages before the FP64 addition (1% faster, twice the error), and
// skipping the MFMAs of zero 32-K halves inside a live stage.
constexpr int kStage = 64;        // K per stage
constexpr int kLDS = kStage + 8;  // LDS row length in halves (bank spread)
constexpr int kMaxStages = kMaxFlagBlocks / 2;
constexpr int kOutRows = 8;  // output rows per pass of the planned-order epilogue
// Wavefronts per SIMD (= resident workgroups per CU) of the persistent
// kernel by tile width: narrow tiles hold few accumulators, and more of
// them in flight hide the AO-slab loads, which a tile of 64 or 128 MOs
// follows with few MFMAs (the most that fit without spilling).
constexpr int persistent_waves(int cbw, int terms) {
  return cbw == 1 ? (terms == 2 ? 4 : 3) : (cbw == 2 && terms == 2) ? 3 : 2;
}
// The same for the dense per-context kernel (fewer registers: no job loop).
constexpr int dense_waves(int cbw, int terms) { return cbw == 1 || (cbw == 2 && terms == 2) ? 4 : 2; }

// Stage list of row tile rt for the column blocks gb0 .. gb0 + 4 CBW - 1,
// built by wavefront 0 (lane = this thread's lane): the live 64-AO stages
// in s_list[0, *s_count) and, per listed stage, bit b (first 32-AO half)
// and 32 + b (second half) of s_live for live column block gb0 + b.  A
// stage or half is live when the AO flags of the tile (null: all) mark it
// and, with a column plan, some block has a nonzero coefficient there.
template <int CBW>
__device__ __forceinline__ void build_stage_list(int lane, int nst, int64_t rt, const uint8_t *__restrict__ flags,
                                                 int nkb, const ColPlan &plan, int64_t gb0, int *s_list,
                                                 uint64_t *s_live, int *s_count) {
  const uint8_t *f = flags != nullptr ? flags + rt * nkb : nullptr;
  int base = 0;
  for (int s0 = 0; s0 < nst; s0 += 64) {
    const int st = s0 + lane;
    const bool h0 = st < nst && (f == nullptr || f[2 * st] != 0);
    const bool h1 = st < nst && (f == nullptr || (2 * st + 1 < nkb && f[2 * st + 1] != 0));
    uint64_t lv = (h0 ? 0xffffffffull : 0ull) | (h1 ? 0xffffffff00000000ull : 0ull);
    if ((h0 || h1) && plan.blk != nullptr) {
      // all loads first (independent, no branches), then the bits
      const int64_t g = 4 * (int64_t)st;  // first 16-AO block of the stage
      uint64_t w[4 * CBW];
#pragma unroll
      for (int b = 0; b < 4 * CBW; ++b) {
        const int64_t blk = gb0 + b < plan.nblk ? gb0 + b : plan.nblk - 1;
        w[b] = plan.blk[blk * plan.nw + (g >> 6)];
      }
      uint64_t bits = 0;
#pragma unroll
      for (int b = 0; b < 4 * CBW; ++b) {
        const uint64_t nib = gb0 + b < plan.nblk ? (w[b] >> (g & 63)) & 15ull : 0ull;
        bits |= ((nib & 3ull) != 0 ? 1ull : 0ull) << b;
        bits |= ((nib & 12ull) != 0 ? 1ull : 0ull) << (32 + b);
      }
      lv &= bits;
    }
    const bool live = lv != 0;
    const uint64_t m = __ballot(live);
    if (live) {
      const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
      s_list[pos] = st;
      s_live[pos] = lv;
    }
    base += __popcll(m);
  }
  if (lane == 0) *s_count = base;
}

// The row-tile part bx (H parts of 64 / H rows per 64-row tile) and the
// column group by of the t-th workgroup (or work item) of a split
// contraction: the parts and the five components of one 64-point tile take
// consecutive numbers, then the point tiles, then the column groups.  The
// ten workgroups of a point tile have the same live stages and read the
// same coefficient fragments, so they are dispatched together (in row
// order, component-major, FP16X2 on 64 waters at 128 walkers ran 2.40 ms,
// in this order 2.35 ms).
__device__ __forceinline__ void split_item(int64_t t, int64_t npt, int64_t H, int64_t &bx, int64_t &by) {
  const int64_t part = t % H;
  int64_t u = t / H;
  const int64_t comp = u % 5;
  u /= 5;
  const int64_t pt = u % npt;
  by = u / npt;
  bx = (comp * npt + pt) * H + part;
}

// One 32-row x 64*CBW-column tile (bx, by) of a split contraction; the
// workgroup's LDS comes in as pointers so that one kernel can hold tiles of
// several widths (the persistent batch kernel below).
template <int SP, int CBW, int WR>
__device__ __forceinline__ void split_tile(int64_t bx, int64_t by, int64_t P, int64_t npt, int64_t K,
                                           int64_t Ncol, const float *__restrict__ A, int64_t lda,
                                           const float *__restrict__ rowmax, const uint16_t *__restrict__ Bs,
                                           int64_t ldb, const float *__restrict__ colinv,
                                           const uint8_t *__restrict__ flags, int nkb,
                                           const int32_t *__restrict__ perm, double *__restrict__ C, int64_t ldc,
                                           ColPlan plan, uint16_t *sA, float *s_inv, int *s_list, uint64_t *s_live,
                                           int *s_count, double *s_out) {
  constexpr int T = Terms<SP>::n;
  constexpr int RB = 2;            // 16-row blocks per wavefront
  constexpr int KQ = 8;            // K per thread in the A slab

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int cq = __builtin_amdgcn_readfirstlane(wave & 3), rh = wave >> 2;  // column quarter, row half
  const int64_t rt = bx / (64 / WR);  // row tile: component c, tile rows 64*pt ..
  const int r0 = (int)(bx % (64 / WR)) * WR;  // this workgroup's first row of the tile
  const int64_t comp = rt / npt, p0 = (rt - comp * npt) * 64 + r0;
  const int64_t gb0 = by * 4 * CBW;  // the workgroup's first column block
  const int nst = (int)((K + kStage - 1) / kStage);
  // this wavefront's column blocks cq + 4 cb that hold MOs (the prepared
  // coefficients are zero-padded to the workgroup width: on the 48-molecule
  // batch 22% of the 64- and 128-wide tiles' columns are padding); the
  // others are neither loaded nor multiplied
  const int64_t nvb = (Ncol + 15) / 16 - gb0;
  const int ncb = nvb <= cq ? 0 : (int)((nvb - cq + 3) / 4);

  // stage list: live stages and, per stage, bit b (first half) and 32 + b
  // (second half) for live column block gb0 + b
  const bool listed = flags != nullptr || plan.blk != nullptr;
  int count = nst;
  if (listed) {
    if (wave == 0) build_stage_list<CBW>(lane, nst, rt, flags, nkb, plan, gb0, s_list, s_live, s_count);
    __syncthreads();
    count = *s_count;
  }
  // this wavefront's live column blocks in half ks of stage s: bit 4j for
  // block gb0 + cq + 4j
  auto wave_live = [&](int s, int ks) -> uint32_t {
    if (!listed) return 0xffffffffu;
    const uint64_t lv = s_live[s];
    const uint32_t w = ks == 0 ? (uint32_t)lv : (uint32_t)(lv >> 32);
    return (uint32_t)__builtin_amdgcn_readfirstlane(w) >> cq;
  };

  // A slot: row ar = lane of the tile, KQ consecutive K at aq = KQ*wave,
  // read as 16-byte pieces of the blocked operand (consecutive lanes,
  // consecutive pieces); 16 consecutive lanes take 16 consecutive rows, so
  // the LDS stores below do not conflict.  Each row has one power-of-two
  // scale (from its max |entry| over all K).
  const int ar = tid % WR, aq = (tid / WR) * KQ;
  const float *Ap = A + rt * (lda / 4) * 256 + (aq / 4) * 256 + (r0 + ar) * 4;
  const float sc = pow2_scale(rowmax[rt * 64 + r0 + ar]);  // from the row's max |entry|
  if (aq == 0) s_inv[ar] = 1.0f / sc;
  // C operand in fragment order (split_coef_kernel): the 16-byte piece of
  // lane l for column tile ct, 32-K chunk kc, term t is at
  // ((ct * nkc + kc) * T + t) * 512 + 8 * l halves, so each load of a
  // wavefront reads one contiguous kilobyte.  The loads go through a buffer
  // resource over the workgroup's 4*CBW column tiles: lane offset 16*l
  // bytes, tile and stage offset in a scalar register.  A dead (tile, half)
  // gets a lane offset past the end of the buffer, so its loads return
  // zeros without a memory access, and every stage issues the same loads
  // (the waits on them stay exact).
  const int64_t nkc = ldb / 32;
  const int tile_bytes = (int)(nkc * T * 1024);
  const __amdgpu_buffer_rsrc_t brs = __builtin_amdgcn_make_buffer_rsrc(
      (void *)(Bs + gb0 * nkc * T * 512), (short)0, 4 * CBW * tile_bytes, 0x00020000);
  const int lane_off = 16 * lane;

  // two register sets of the A slab, stages s and s + 1 in flight while
  // stage s - 1 computes (the loop is unrolled by two so that each stage's
  // set is fixed at compile time)
  float4 ra0[KQ / 4], ra1[KQ / 4];
  uint4 bq[2][CBW][T];
  auto stage_k0 = [&](int s) {
    return (int64_t)__builtin_amdgcn_readfirstlane(listed ? s_list[s] : s) * kStage;
  };
  auto load_a = [&](float4 (&ra)[KQ / 4], int64_t k0) {
    const float *src = Ap + (k0 / 4) * 256;
#pragma unroll
    for (int q = 0; q < KQ / 4; ++q) ra[q] = *reinterpret_cast<const float4 *>(src + q * 256);
  };
  auto load_b = [&](int64_t k0, int ks, uint32_t live) {
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      if (cb >= ncb) continue;  // padding columns (wavefront-uniform)
      const int vo = (live >> (4 * cb)) & 1u ? lane_off : (int)0x80000000;
      const int so = (cq + 4 * cb) * tile_bytes + (int)(k0 / 32 + ks) * T * 1024;
#pragma unroll
      for (int t = 0; t < T; ++t)
        bq[ks][cb][t] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(brs, vo + t * 1024, so, 0));
    }
  };

  f4 d[RB][CBW];
  double acc[RB][CBW][4];
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[r][cb][j] = 0.0;

  auto stage = [&](float4 (&ra)[KQ / 4], int s) {
    __syncthreads();  // every wavefront is done reading the previous A slab
#pragma unroll
    for (int h = 0; h < KQ / 8; ++h) {
      const float v[8] = {ra[2 * h].x * sc,     ra[2 * h].y * sc,     ra[2 * h].z * sc,
                          ra[2 * h].w * sc,     ra[2 * h + 1].x * sc, ra[2 * h + 1].y * sc,
                          ra[2 * h + 1].z * sc, ra[2 * h + 1].w * sc};
      uint4 o[T];
      split8<SP>(v, o);
#pragma unroll
      for (int t = 0; t < T; ++t) *reinterpret_cast<uint4 *>(&sA[t * (WR * kLDS) + ar * kLDS + aq + 8 * h]) = o[t];
    }
    __syncthreads();
    if (s + 2 < count) load_a(ra, stage_k0(s + 2));  // this set is free again
    const bool more = s + 1 < count;
    const int64_t k1 = more ? stage_k0(s + 1) : 0;
    const uint32_t live1[2] = {more ? wave_live(s + 1, 0) : 0u, more ? wave_live(s + 1, 1) : 0u};

#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      uint4 a[T][RB];
#pragma unroll
      for (int t = 0; t < T; ++t)
#pragma unroll
        for (int r = 0; r < RB; ++r)
          a[t][r] = *reinterpret_cast<const uint4 *>(&sA[t * (WR * kLDS) + ((rh * RB + r) * 16 + li) * kLDS + ks * 32 + 8 * lk]);
      // product by product over all RB x CBW tiles: RB*CBW independent
      // accumulators between dependent MFMAs.  Dead tiles run on the zeros
      // their loads returned (branching around their MFMAs measured slower:
      // 3.28 against 2.97 ms on water64, and around the FP64 additions of
      // stage-dead tiles below 3.10 ms).
      constexpr int NP = T == 3 ? 6 : 3;
      constexpr int pa[6] = {0, 0, 1, 0, 2, 1}, pb[6] = {0, 1, 0, 2, 0, 1};  // (A term, C term)
#pragma unroll
      for (int pr = 0; pr < NP; ++pr)
#pragma unroll
        for (int cb = 0; cb < CBW; ++cb) {
          if (cb >= ncb) continue;
#pragma unroll
          for (int r = 0; r < RB; ++r)
            d[r][cb] = mfma<SP>(a[pa[pr]][r], bq[ks][cb][pb[pr]],
                                pr == 0 && ks == 0 ? f4{0.f, 0.f, 0.f, 0.f} : d[r][cb]);
        }
      if (more) load_b(k1, ks, live1[ks]);  // this half-stage's C registers are free again
    }
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int cb = 0; cb < CBW; ++cb) {
        if (cb >= ncb) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[r][cb][j] += (double)d[r][cb][j];
      }
  };

  if (count > 0) {
    const int64_t k0 = stage_k0(0);
    load_a(ra0, k0);
    if (count > 1) load_a(ra1, stage_k0(1));
    load_b(k0, 0, wave_live(0, 0));
    load_b(k0, 1, wave_live(0, 1));
  } else {
    __syncthreads();  // s_inv is read by every wavefront in the epilogue
  }
  for (int s = 0; s < count; s += 2) {
    stage(ra0, s);
    if (s + 1 < count) stage(ra1, s + 1);
  }
  if (plan.col_mo == nullptr) {
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      const int64_t col = (gb0 + cq + 4 * cb) * 16 + li;
      if (col >= Ncol) continue;
      const double ci = (double)colinv[col];
#pragma unroll
      for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int row = (rh * RB + r) * 16 + 4 * lk + j;
          const int64_t i = p0 + row;
          if (i < P) {
            const int64_t p = perm != nullptr ? (int64_t)perm[i] : i;
            C[(p * 5 + comp) * ldc + col] = acc[r][cb][j] * ((double)s_inv[row] * ci);
          }
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
  double ci[CBW];
#pragma unroll
  for (int cb = 0; cb < CBW; ++cb) {
    const int64_t col = (gb0 + cq + 4 * cb) * 16 + li;
    mloc[cb] = col < Ncol ? (int)(plan.col_mo[col] - g_lo) : -1;
    ci[cb] = col < Ncol ? (double)colinv[col] : 0.0;
  }
  for (int pass = 0; pass < WR / kOutRows; ++pass) {
    __syncthreads();  // the previous pass has been read out
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb)
#pragma unroll
      for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int row = (rh * RB + r) * 16 + 4 * lk + j;
          if (mloc[cb] >= 0 && row / kOutRows == pass)
            s_out[(row % kOutRows) * (64 * CBW) + mloc[cb]] = acc[r][cb][j] * ((double)s_inv[row] * ci[cb]);
        }
    __syncthreads();
    for (int rl = 0; rl < kOutRows; ++rl) {
      const int64_t i = p0 + pass * kOutRows + rl;
      if (i >= P) break;
      const int64_t p = perm != nullptr ? (int64_t)perm[i] : i;
      double *dst = C + (p * 5 + comp) * ldc + g_lo;
      for (int c = tid; c < g_n; c += WR * 8) dst[c] = s_out[rl * (64 * CBW) + c];
    }
  }
}


template <int SP, int CBW, int WR>
__global__ void __attribute__((amdgpu_flat_work_group_size(WR * 8, WR * 8),
                               amdgpu_waves_per_eu(dense_waves(CBW, Terms<SP>::n), dense_waves(CBW, Terms<SP>::n))))
gemm_split_kernel(int64_t P, int64_t npt, int64_t K, int64_t Ncol, const float *__restrict__ A,
                  int64_t lda, const float *__restrict__ rowmax, const uint16_t *__restrict__ Bs,
                  int64_t ldb, const float *__restrict__ colinv, const uint8_t *__restrict__ flags, int nkb,
                  const int32_t *__restrict__ perm, double *__restrict__ C, int64_t ldc, ColPlan plan) {
  __shared__ uint16_t sA[Terms<SP>::n * WR * kLDS];
  __shared__ float s_inv[WR];
  __shared__ int s_list[kMaxStages];
  __shared__ uint64_t s_live[kMaxStages];
  __shared__ int s_count;
  __shared__ double s_out[kOutRows * 64 * CBW];  // output rows in MO order (planned columns)
  int64_t bx, by;
  split_item(blockIdx.x + (int64_t)gridDim.x * blockIdx.y, npt, 64 / WR, bx, by);
  split_tile<SP, CBW, WR>(bx, by, P, npt, K, Ncol, A, lda, rowmax, Bs, ldb, colinv, flags, nkb, perm, C, ldc, plan,
                          sA, s_inv, s_list, s_live, &s_count, s_out);
}

// Live-block form of the split tile (the default for the split forms).
// Same operands, sums, scales and FP64 accumulation per stage as
// split_tile, so the results are the same bit for bit; what differs is
// what is issued.  Counters on MI355X (water64, 128 walkers) showed the
// four-wavefront kernel above bound by the vector-memory return path, 85%
// busy, nine tenths of its bytes coefficient fragments, and the dead ones
// (zero blocks of the column plan, AOs screened for the tile) costing as
// much there as live ones although their loads touch no memory.  Here:
//   - a (column block, 32-K half) that the stage list marks dead is
//     neither loaded nor multiplied: its loads are not issued and each
//     wavefront branches (uniformly) around its MFMAs; the first live half
//     of a block in a stage starts its FP32 sum from zero, and only blocks
//     live in the stage are added into the FP64 accumulators;
//   - the FP32 AO slab is read with nontemporal loads (NT), so that the
//     streamed operand does not push the fragments out of L2, and is
//     double-buffered in LDS (one barrier per stage);
//   - what the epilogue needs per row (the point the row is stored at) and
//     per column (its place in the MO group, its inverse scale) is fetched
//     into LDS at the start, so the output stores wait on no load; rows go
//     out 16 at a time through the idle slab buffer, as 16-byte
//     nontemporal runs when the rows are 16-byte aligned.
// Measured at 409,600 x 4,160 x 320 (water64, 128 walkers, FP16X2): 2.36 ms
// against 2.97 ms for split_tile; 52% of its MFMAs are issued.  RB = 4
// (the whole 64-row tile per wavefront, one wavefront per SIMD) halves the
// fragment bytes per MFMA but runs 3.1 ms: one wavefront per SIMD hides
// too little latency.  Counters (same workload, 2.38 ms): texture-data
// unit 85% busy, 32 GB through L1 per launch (nine tenths fragments, L2
// latency 242 cycles), matrix pipe 23% busy.  Measured slower: eight-
// wavefront workgroups of a whole 64-row tile fetching each live fragment
// once into LDS for both row halves (3.40 ms: the fragment reads, 160 KB
// per stage and workgroup, then bind on LDS bandwidth).  Not faster: the
// whole 64-row tile per wavefront (half the fragment bytes per MFMA) in
// eight-wavefront workgroups of 2-3 column blocks each, two wavefronts per
// SIMD (2.39 ms).
template <int SP, int CBW, int RB, bool NT>
__device__ __forceinline__ void split_tile_live(int64_t bx, int64_t by, int64_t P, int64_t npt, int64_t K,
                                                int64_t Ncol, const float *__restrict__ A, int64_t lda,
                                                const float *__restrict__ rowmax,
                                                const uint16_t *__restrict__ Bs, int64_t ldb,
                                                const float *__restrict__ colinv,
                                                const uint8_t *__restrict__ flags, int nkb,
                                                const int32_t *__restrict__ perm, double *__restrict__ C,
                                                int64_t ldc, ColPlan plan, uint16_t *sA, float *s_inv,
                                                int *s_list, uint64_t *s_live, int *s_count, int32_t *s_perm,
                                                int32_t *s_mloc, float *s_ci) {
  constexpr int T = Terms<SP>::n;
  constexpr int WR = 16 * RB;
  constexpr int KQ = WR / 4;  // K per thread in the A slab (256 threads, 64 K)
  constexpr int SLAB = T * WR * kLDS;  // halves per LDS slab

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int cq = __builtin_amdgcn_readfirstlane(wave);  // column quarter
  const int64_t rt = bx / (64 / WR);
  const int r0 = (int)(bx % (64 / WR)) * WR;
  const int64_t comp = rt / npt, p0 = (rt - comp * npt) * 64 + r0;
  const int64_t gb0 = by * 4 * CBW;
  const int nst = (int)((K + kStage - 1) / kStage);

  // epilogue tables
  if (tid < WR) {
    const int64_t i = p0 + tid;
    s_perm[tid] = i < P ? (perm != nullptr ? perm[i] : (int32_t)i) : -1;
  }
  for (int c = tid; c < 64 * CBW; c += 256) {
    const int64_t col = gb0 * 16 + c;
    s_mloc[c] = col < Ncol ? (int32_t)((plan.col_mo != nullptr ? (int64_t)plan.col_mo[col] : col) - gb0 * 16) : -1;
    s_ci[c] = col < Ncol ? colinv[col] : 0.0f;
  }

  // stage list (as in split_tile)
  const bool listed = flags != nullptr || plan.blk != nullptr;
  int count = nst;
  if (listed) {
    if (wave == 0) build_stage_list<CBW>(lane, nst, rt, flags, nkb, plan, gb0, s_list, s_live, s_count);
    __syncthreads();
    count = *s_count;
  }
  // this wavefront's column blocks that hold MOs: bit 4j for block
  // gb0 + cq + 4j (the padding past the last MO is neither loaded nor
  // multiplied)
  const int64_t nvb = (Ncol + 15) / 16 - gb0;
  uint32_t vmask = 0;
#pragma unroll
  for (int cb = 0; cb < CBW; ++cb)
    if (cq + 4 * cb < nvb) vmask |= 1u << (4 * cb);
  // this wavefront's live blocks in half ks of stage s: bit 4j for block
  // gb0 + cq + 4j
  auto wave_live = [&](int s, int ks) -> uint32_t {
    if (!listed) return vmask;
    const uint64_t lv = s_live[s];
    const uint32_t w = ks == 0 ? (uint32_t)lv : (uint32_t)(lv >> 32);
    return ((uint32_t)__builtin_amdgcn_readfirstlane(w) >> cq) & vmask;
  };

  const int ar = tid % WR, aq = (tid / WR) * KQ;
  const float *Ap = A + rt * (lda / 4) * 256 + (aq / 4) * 256 + (r0 + ar) * 4;
  const float sc = pow2_scale(rowmax[rt * 64 + r0 + ar]);
  if (aq == 0) s_inv[ar] = 1.0f / sc;
  const int64_t nkc = ldb / 32;
  const int tile_bytes = (int)(nkc * T * 1024);
  const __amdgpu_buffer_rsrc_t brs = __builtin_amdgcn_make_buffer_rsrc(
      (void *)(Bs + gb0 * nkc * T * 512), (short)0, 4 * CBW * tile_bytes, 0x00020000);
  const int lane_off = 16 * lane;

  float4 ra[KQ / 4];
  uint4 bq[2][CBW][T];
  auto stage_k0 = [&](int s) {
    return (int64_t)__builtin_amdgcn_readfirstlane(listed ? s_list[s] : s) * kStage;
  };
  auto load_a = [&](int64_t k0) {
    const float *src = Ap + (k0 / 4) * 256;
#pragma unroll
    for (int q = 0; q < KQ / 4; ++q) {
      if constexpr (NT) {
        const f4 v = __builtin_nontemporal_load(reinterpret_cast<const f4 *>(src + q * 256));
        ra[q] = make_float4(v[0], v[1], v[2], v[3]);
      } else {
        ra[q] = *reinterpret_cast<const float4 *>(src + q * 256);
      }
    }
  };
  // fragments of half ks of the stage at k0 for this wavefront's live blocks
  auto load_b = [&](int64_t k0, int ks, uint32_t live) {
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      if ((live >> (4 * cb)) & 1u) {
        const int so = (cq + 4 * cb) * tile_bytes + (int)(k0 / 32 + ks) * T * 1024;
#pragma unroll
        for (int t = 0; t < T; ++t)
          bq[ks][cb][t] =
              __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(brs, lane_off + t * 1024, so, 0));
      }
    }
  };
  auto write_slab = [&](int buf) {
    uint16_t *dst = sA + buf * SLAB;
#pragma unroll
    for (int h = 0; h < KQ / 8; ++h) {
      const float v[8] = {ra[2 * h].x * sc,     ra[2 * h].y * sc,     ra[2 * h].z * sc,
                          ra[2 * h].w * sc,     ra[2 * h + 1].x * sc, ra[2 * h + 1].y * sc,
                          ra[2 * h + 1].z * sc, ra[2 * h + 1].w * sc};
      uint4 o[T];
      split8<SP>(v, o);
#pragma unroll
      for (int t = 0; t < T; ++t) *reinterpret_cast<uint4 *>(&dst[t * (WR * kLDS) + ar * kLDS + aq + 8 * h]) = o[t];
    }
  };

  f4 d[RB][CBW];
  double acc[RB][CBW][4];
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[r][cb][j] = 0.0;

  if (count > 0) {
    const int64_t k0 = stage_k0(0);
    load_a(k0);
    load_b(k0, 0, wave_live(0, 0));
    load_b(k0, 1, wave_live(0, 1));
  } else {
    __syncthreads();
  }
  for (int s = 0; s < count; ++s) {
    const int buf = s & 1;
    write_slab(buf);
    __syncthreads();
    const bool more = s + 1 < count;
    const int64_t k1 = more ? stage_k0(s + 1) : 0;
    const uint32_t live1[2] = {more ? wave_live(s + 1, 0) : 0u, more ? wave_live(s + 1, 1) : 0u};
    if (more) load_a(k1);
    const uint32_t live0[2] = {wave_live(s, 0), wave_live(s, 1)};
    const uint16_t *slab = sA + buf * SLAB;
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      uint4 a[T][RB];
#pragma unroll
      for (int t = 0; t < T; ++t)
#pragma unroll
        for (int r = 0; r < RB; ++r)
          a[t][r] = *reinterpret_cast<const uint4 *>(&slab[t * (WR * kLDS) + (r * 16 + li) * kLDS + ks * 32 + 8 * lk]);
      constexpr int NP = T == 3 ? 6 : 3;
      constexpr int pa[6] = {0, 0, 1, 0, 2, 1}, pb[6] = {0, 1, 0, 2, 0, 1};  // (A term, C term)
#pragma unroll
      for (int cb = 0; cb < CBW; ++cb) {
        if (!((live0[ks] >> (4 * cb)) & 1u)) continue;
        const bool fresh = ks == 0 || !((live0[0] >> (4 * cb)) & 1u);  // first live half of the stage
        if (fresh) {
#pragma unroll
          for (int pr = 0; pr < NP; ++pr)
#pragma unroll
            for (int r = 0; r < RB; ++r)
              d[r][cb] = mfma<SP>(a[pa[pr]][r], bq[ks][cb][pb[pr]], pr == 0 ? f4{0.f, 0.f, 0.f, 0.f} : d[r][cb]);
        } else {
#pragma unroll
          for (int pr = 0; pr < NP; ++pr)
#pragma unroll
            for (int r = 0; r < RB; ++r) d[r][cb] = mfma<SP>(a[pa[pr]][r], bq[ks][cb][pb[pr]], d[r][cb]);
        }
      }
      if (more) load_b(k1, ks, live1[ks]);  // this half's fragment registers are free again
    }
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      if (!(((live0[0] | live0[1]) >> (4 * cb)) & 1u)) continue;
#pragma unroll
      for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[r][cb][j] += (double)d[r][cb][j];
    }
  }

  // epilogue: 16 rows at a time through the slab buffer, in MO order
  const int64_t g_lo = gb0 * 16;
  const int g_n = (int)(Ncol - g_lo < 64 * CBW ? Ncol - g_lo : 64 * CBW);
  constexpr int OR = 16;
  double *so = reinterpret_cast<double *>(sA);
  int mloc[CBW];
  double ci[CBW];
#pragma unroll
  for (int cb = 0; cb < CBW; ++cb) {
    const int c = (cq + 4 * cb) * 16 + li;  // the lane's column within the group
    mloc[cb] = s_mloc[c];
    ci[cb] = (double)s_ci[c];
  }
  const bool vec = (ldc % 2) == 0 && (g_n % 2) == 0 && (reinterpret_cast<uintptr_t>(C) % 16) == 0;
#pragma unroll
  for (int pass = 0; pass < WR / OR; ++pass) {  // pass = 16-row block r
    __syncthreads();  // the slab (or the previous pass) has been read
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (mloc[cb] >= 0)
          so[(4 * lk + j) * (64 * CBW) + mloc[cb]] = acc[pass][cb][j] * ((double)s_inv[pass * 16 + 4 * lk + j] * ci[cb]);
    __syncthreads();
    for (int row = 0; row < OR; ++row) {
      const int32_t p = s_perm[pass * OR + row];
      if (p < 0) break;  // rows past the last point
      double *dst = C + ((int64_t)p * 5 + comp) * ldc + g_lo;
      if (vec) {
        typedef double d2 __attribute__((ext_vector_type(2)));
        for (int cp = tid; cp < g_n / 2; cp += 256) {
          const d2 v = {so[row * (64 * CBW) + 2 * cp], so[row * (64 * CBW) + 2 * cp + 1]};
          __builtin_nontemporal_store(v, reinterpret_cast<d2 *>(dst + 2 * cp));
        }
      } else {
        for (int c = tid; c < g_n; c += 256) dst[c] = so[row * (64 * CBW) + c];
      }
    }
  }
}

// LDS of one split_tile_live workgroup (the slab buffer doubles as the
// epilogue's 16-row output buffer)
template <int SP, int CBW, int RB>
struct LiveTileLds {
  static constexpr int kSlab = 2 * Terms<SP>::n * 16 * RB * kLDS;  // halves
  static constexpr int kOut = 16 * 64 * CBW * 4;                   // halves
  static constexpr int kA = kSlab > kOut ? kSlab : kOut;
};

template <int SP, int CBW, int RB, bool NT>
__global__ void __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(RB == 4 ? 1 : 2, RB == 4 ? 1 : 2)))
gemm_split_live_kernel(int64_t P, int64_t npt, int64_t K, int64_t Ncol, const float *__restrict__ A, int64_t lda,
                       const float *__restrict__ rowmax, const uint16_t *__restrict__ Bs, int64_t ldb,
                       const float *__restrict__ colinv, const uint8_t *__restrict__ flags, int nkb,
                       const int32_t *__restrict__ perm, double *__restrict__ C, int64_t ldc, ColPlan plan) {
  using L = LiveTileLds<SP, CBW, RB>;
  __shared__ __attribute__((aligned(16))) uint16_t sA[L::kA];
  __shared__ float s_inv[16 * RB];
  __shared__ int s_list[kMaxStages];
  __shared__ uint64_t s_live[kMaxStages];
  __shared__ int s_count;
  __shared__ int32_t s_perm[16 * RB];
  __shared__ int32_t s_mloc[64 * CBW];
  __shared__ float s_ci[64 * CBW];
  int64_t bx, by;
  split_item(blockIdx.x + (int64_t)gridDim.x * blockIdx.y, npt, 64 / (16 * RB), bx, by);
  split_tile_live<SP, CBW, RB, NT>(bx, by, P, npt, K, Ncol, A, lda, rowmax, Bs, ldb, colinv, flags, nkb, perm, C,
                                   ldc, plan, sA, s_inv, s_list, s_live, &s_count, s_perm, s_mloc, s_ci);
}

// Persistent form over a batch: 2 workgroups per CU take (molecule, row
// tile, column group) items from an atomic counter until the batch is done;
// jobs[j] describes molecule j's contraction (all of tile width CBW) and
// owns items [item0, item0 + items), in the order of split_item.
// Every workgroup leaves the loop once the counter passes `total`.  Each
// item runs split_tile_live (LIVE) or split_tile.
template <int SP, int CBW, bool LIVE>
__global__ void __attribute__((amdgpu_flat_work_group_size(256, 256),
                               amdgpu_waves_per_eu(persistent_waves(CBW, Terms<SP>::n),
                                                   persistent_waves(CBW, Terms<SP>::n))))
gemm_split_persistent_kernel(const MixedGemmJob *__restrict__ jobs, int njobs, int64_t total, int *counter) {
  constexpr int WR = 32;
  using L = LiveTileLds<SP, CBW, 2>;
  __shared__ __attribute__((aligned(16))) uint16_t sA[LIVE ? L::kA : Terms<SP>::n * WR * kLDS];
  __shared__ float s_inv[WR];
  __shared__ int s_list[kMaxStages];
  __shared__ uint64_t s_live[kMaxStages];
  __shared__ int s_count;
  __shared__ double s_out[LIVE ? 1 : kOutRows * 64 * CBW];
  __shared__ int32_t s_perm[LIVE ? WR : 1];
  __shared__ int32_t s_mloc[LIVE ? 64 * CBW : 1];
  __shared__ float s_ci[LIVE ? 64 * CBW : 1];
  __shared__ int64_t s_item;
  __shared__ int s_job;
  for (;;) {
    if (threadIdx.x == 0) {
      const int64_t item = atomicAdd(counter, 1);
      int j = -1;
      if (item < total) {  // the last job whose first item is <= item
        int lo = 0, hi = njobs - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) / 2;
          if (jobs[mid].item0 <= item) lo = mid;
          else hi = mid - 1;
        }
        j = lo;
      }
      s_item = item;
      s_job = j;
    }
    __syncthreads();  // also: the previous tile is finished with the LDS
    const int64_t item = s_item;
    const int j = __builtin_amdgcn_readfirstlane(s_job);
    __syncthreads();  // s_item is read before the next fetch overwrites it
    if (item >= total) return;
    const MixedGemmJob &g = jobs[j];
    const int64_t local = item - g.item0;
    int64_t bx, by;
    split_item(local, g.npt, 2, bx, by);
    if constexpr (LIVE)
      split_tile_live<SP, CBW, 2, true>(bx, by, g.P, g.npt, g.K, g.Ncol, g.A, g.lda, g.rowmax, g.Bs, g.ldb,
                                        g.colinv, g.flags, g.nkb, g.perm, g.C, g.ldc, g.plan, sA, s_inv, s_list,
                                        s_live, &s_count, s_perm, s_mloc, s_ci);
    else
      split_tile<SP, CBW, WR>(bx, by, g.P, g.npt, g.K, g.Ncol, g.A, g.lda, g.rowmax, g.Bs, g.ldb, g.colinv, g.flags,
                              g.nkb, g.perm, g.C, g.ldc, g.plan, sA, s_inv, s_list, s_live, &s_count, s_out);
  }
}

// FP32 form: v_mfma_f32_16x16x4_f32, lane l holds A[l%16][l/16], B[l/16][l%16].
//   A  the blocked FP32 operand (as for the split forms)
//   Bf [Ncol][ldb] FP32 C (unscaled), K contiguous.
template <int CBW>
__global__ void __launch_bounds__(kThreads)
gemm_f32_kernel(int64_t P, int64_t npt, int64_t K, int64_t Ncol, const float *__restrict__ A,
                int64_t lda, const float *__restrict__ Bf, int64_t ldb, const uint8_t *__restrict__ flags,
                int nkb, int flush, const int32_t *__restrict__ perm, double *__restrict__ C, int64_t ldc) {
  constexpr int BN = 32 * CBW;
  constexpr int LDF = kKB + 4;  // LDS row length in floats
  __shared__ float sA[kBM * LDF];
  __shared__ float sB[BN * LDF];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t rt = blockIdx.x;  // row tile: component c, points 64*pt ..
  const int64_t comp = rt / npt, p0 = (rt - comp * npt) * 64;
  const int64_t col0 = (int64_t)blockIdx.y * BN;
  const int64_t nblk = (K + kKB - 1) / kKB;
  const int r_ld = tid >> 2, seg = (tid & 3) * 8;
  const float *Ap = A + rt * (lda / 4) * 256 + r_ld * 4;

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
  // d holds the FP32 sums of the aligned group of `flush` blocks kb / flush
  // (groups fixed by K, not by which blocks are skipped, so a row's result
  // does not depend on the other rows of its tile)
  int64_t group = -1;
  for (int64_t kb = 0; kb < nblk; ++kb) {
    if (flags != nullptr && flags[rt * nkb + kb] == 0) continue;
    if (kb / flush != group) {
      if (group >= 0) {
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
          for (int cb = 0; cb < CBW; ++cb) {
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[rb][cb][j] += (double)d[rb][cb][j];
            d[rb][cb] = f4{0.f, 0.f, 0.f, 0.f};
          }
      }
      group = kb / flush;
    }
    const int64_t k0 = kb * kKB;
    {
      const float *src = Ap + ((k0 + seg) / 4) * 256;
      const float4 x0 = *reinterpret_cast<const float4 *>(src);
      const float4 x1 = *reinterpret_cast<const float4 *>(src + 256);
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
  }
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < CBW; ++cb) {
      const int64_t col = col0 + wc * 16 * CBW + cb * 16 + li;
      if (col >= Ncol) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t i = p0 + wr * 32 + rb * 16 + 4 * lk + j;
        if (i < P) {
          const int64_t p = perm != nullptr ? (int64_t)perm[i] : i;
          C[(p * 5 + comp) * ldc + col] = acc[rb][cb][j] + (double)d[rb][cb][j];
        }
      }
    }
}

// One workgroup per MO i: column scale from max_k |C[i][k]| and the split
// terms of C[i][k] * scale, zero past K and for the padding MOs i >= M.
// The terms are stored in the MFMA fragment order of the contraction that
// will read them (frag = 16: 16x16x32 tiles; 32 would suit 32x32x16): element
// (i, k) of term t goes to lane l = i % frag + frag * ((k % (512/frag)) / 8),
// slot k % 8, of fragment ((i / frag) * (ldb / (512/frag)) + k / (512/frag))
// * T + t, 512 halves per fragment.  FP32 form: scaled-by-one values at
// Bs[i*ldb + k] (row layout).
template <int SP>
__global__ void __launch_bounds__(kThreads)
split_coef_kernel(int64_t M, int64_t K, const double *__restrict__ coef, int64_t ldb, int frag,
                  const int32_t *__restrict__ col_mo, uint16_t *__restrict__ Bs, float *__restrict__ colinv) {
  __shared__ float red[kThreads];
  const int64_t i = blockIdx.x;  // prepared column
  const bool real = i < M;
  const double *row = coef + (real ? (col_mo != nullptr ? (int64_t)col_mo[i] : i) : 0) * K;
  float m = 0.f;
  if (real)
    for (int64_t k = threadIdx.x; k < K; k += kThreads) m = fmaxf(m, fabsf((float)row[k]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  const float sc = SP == QMCKL_AMD_SPLIT_FP32 ? 1.0f : pow2_scale(red[0]);
  if (threadIdx.x == 0 && real) colinv[i] = 1.0f / sc;
  for (int64_t k = threadIdx.x; k < ldb; k += kThreads) {
    const float x = real && k < K ? (float)row[k] * sc : 0.0f;
    if constexpr (SP == QMCKL_AMD_SPLIT_FP32) {
      reinterpret_cast<float *>(Bs)[i * ldb + k] = x;
    } else {
      constexpr int T = Terms<SP>::n;
      const int kw = 512 / frag;  // K per fragment
      const int64_t lane = i % frag + frag * ((k % kw) / 8);
      const int64_t fragment = ((i / frag) * (ldb / kw) + k / kw) * T;
      split_store<SP>(x, Bs + fragment * 512 + lane * 8 + k % 8, 512);
    }
  }
}

template <int CBW>
hipError_t launch_f32(hipStream_t s, int64_t P, int64_t npt, int64_t K, int64_t Ncol, const float *A,
                      int64_t lda, const void *B, int64_t ldb, const uint8_t *flags, int nkb,
                      const int32_t *perm, double *C, int64_t ldc) {
  constexpr int BN = 32 * CBW;
  const int64_t gx = 5 * npt, gy = (Ncol + BN - 1) / BN;
  if (gx > 0x7fffffff || gy > 65535) return hipErrorInvalidValue;
  hipLaunchKernelGGL(gemm_f32_kernel<CBW>, dim3((unsigned)gx, (unsigned)gy), dim3(kThreads), 0, s, P, npt,
                     K, Ncol, A, lda, (const float *)B, ldb, flags, nkb, 2, perm, C, ldc);
  return hipGetLastError();
}

template <int SP, int CBW, int WR = 32>
hipError_t launch_split(hipStream_t s, int64_t P, int64_t npt, int64_t K, int64_t Ncol, const float *A,
                        int64_t lda, const float *rowmax, const void *B, int64_t ldb,
                        const float *colinv, const uint8_t *flags, int nkb, const int32_t *perm, double *C,
                        int64_t ldc, ColPlan plan) {
  const int64_t gx = 5 * npt * (64 / WR), gy = (Ncol + 64 * CBW - 1) / (64 * CBW);
  if (gx > 0x7fffffff || gy > 65535 || (lda % kStage) != 0 || (ldb % kStage) != 0)
    return hipErrorInvalidValue;
  if (flags != nullptr && (nkb > kMaxFlagBlocks || (int64_t)nkb * 32 < K)) return hipErrorInvalidValue;
  if (plan.blk != nullptr && ((K + kStage - 1) / kStage > kMaxStages || plan.nw * 64 * 16 < K || plan.nblk < 0))
    return hipErrorInvalidValue;
  // the coefficient buffer of one workgroup's column tiles is addressed with 32-bit offsets
  if ((int64_t)4 * CBW * (ldb / 32) * Terms<SP>::n * 1024 >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
  hipLaunchKernelGGL((gemm_split_kernel<SP, CBW, WR>), dim3((unsigned)gx, (unsigned)gy), dim3(WR * 8), 0, s, P, npt,
                     K, Ncol, A, lda, rowmax, (const uint16_t *)B, ldb, colinv, flags, nkb, perm, C,
                     ldc, plan);
  return hipGetLastError();
}

template <int SP, int CBW, int RB, bool NT>
hipError_t launch_split_live(hipStream_t s, int64_t P, int64_t npt, int64_t K, int64_t Ncol, const float *A,
                             int64_t lda, const float *rowmax, const void *B, int64_t ldb, const float *colinv,
                             const uint8_t *flags, int nkb, const int32_t *perm, double *C, int64_t ldc,
                             ColPlan plan) {
  constexpr int WR = 16 * RB;
  const int64_t gx = 5 * npt * (64 / WR), gy = (Ncol + 64 * CBW - 1) / (64 * CBW);
  if (gx > 0x7fffffff || gy > 65535 || (lda % kStage) != 0 || (ldb % kStage) != 0)
    return hipErrorInvalidValue;
  if (flags != nullptr && (nkb > kMaxFlagBlocks || (int64_t)nkb * 32 < K)) return hipErrorInvalidValue;
  if (plan.blk != nullptr && ((K + kStage - 1) / kStage > kMaxStages || plan.nw * 64 * 16 < K || plan.nblk < 0))
    return hipErrorInvalidValue;
  if ((int64_t)4 * CBW * (ldb / 32) * Terms<SP>::n * 1024 >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
  hipLaunchKernelGGL((gemm_split_live_kernel<SP, CBW, RB, NT>), dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s,
                     P, npt, K, Ncol, A, lda, rowmax, (const uint16_t *)B, ldb, colinv, flags, nkb, perm, C, ldc,
                     plan);
  return hipGetLastError();
}

template <int SP>
hipError_t launch_form(hipStream_t s, int64_t P, int64_t npt, int64_t K, int64_t Ncol, const float *A,
                       int64_t lda, const float *rowmax, const void *B, int64_t ldb,
                       const float *colinv, const uint8_t *flags, int nkb, const int32_t *perm, double *C,
                       int64_t ldc, ColPlan plan, int gemm_tile) {
  if constexpr (SP == QMCKL_AMD_SPLIT_FP32) {
    if (plan.col_mo != nullptr || plan.blk != nullptr) return hipErrorInvalidValue;
    const int64_t nb = (Ncol + 31) / 32;  // 32-column groups
    if (nb <= 1) return launch_f32<1>(s, P, npt, K, Ncol, A, lda, B, ldb, flags, nkb, perm, C, ldc);
    if (nb <= 2) return launch_f32<2>(s, P, npt, K, Ncol, A, lda, B, ldb, flags, nkb, perm, C, ldc);
    if (nb <= 3) return launch_f32<3>(s, P, npt, K, Ncol, A, lda, B, ldb, flags, nkb, perm, C, ldc);
    if (nb <= 4) return launch_f32<4>(s, P, npt, K, Ncol, A, lda, B, ldb, flags, nkb, perm, C, ldc);
    return launch_f32<5>(s, P, npt, K, Ncol, A, lda, B, ldb, flags, nkb, perm, C, ldc);
  } else {
    const int cbw = split_wg_cols(SP, Ncol) / 64;
    if (gemm_tile == QMCKL_GEMM_TILE_DENSE64_DEVICE) {
      if (cbw == 1) return launch_split<SP, 1, 64>(s, P, npt, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, perm, C, ldc, plan);
      if (cbw == 2) return launch_split<SP, 2, 64>(s, P, npt, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, perm, C, ldc, plan);
      if (cbw == 3) return launch_split<SP, 3, 64>(s, P, npt, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, perm, C, ldc, plan);
      if constexpr (Terms<SP>::n == 2) {
        if (cbw == 4) return launch_split<SP, 4, 64>(s, P, npt, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, perm, C, ldc, plan);
        if (cbw == 5) return launch_split<SP, 5, 64>(s, P, npt, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, perm, C, ldc, plan);
      }
      return hipErrorInvalidValue;
    }
    if (gemm_tile != QMCKL_GEMM_TILE_DENSE_DEVICE) {
#define QMCKL_LIVE(W)                                                                                              \
  switch (gemm_tile) {                                                                                            \
    case QMCKL_GEMM_TILE_LIVE_CACHED_DEVICE:                                                                     \
      return launch_split_live<SP, W, 2, false>(s, P, npt, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, perm, C, ldc, plan); \
    case QMCKL_GEMM_TILE_LIVE64_DEVICE:                                                                          \
      return launch_split_live<SP, W, 4, true>(s, P, npt, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, perm, C, ldc, plan); \
    default:                                                                                                      \
      return launch_split_live<SP, W, 2, true>(s, P, npt, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, perm, C, ldc, plan); \
  }
      if (cbw == 1) { QMCKL_LIVE(1) }
      if (cbw == 2) { QMCKL_LIVE(2) }
      if (cbw == 3) { QMCKL_LIVE(3) }
      if constexpr (Terms<SP>::n == 2) {
        if (cbw == 4) { QMCKL_LIVE(4) }
        if (cbw == 5) { QMCKL_LIVE(5) }
      }
#undef QMCKL_LIVE
      return hipErrorInvalidValue;
    }
    if (cbw == 1) return launch_split<SP, 1>(s, P, npt, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, perm, C, ldc, plan);
    if (cbw == 2) return launch_split<SP, 2>(s, P, npt, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, perm, C, ldc, plan);
    if (cbw == 3) return launch_split<SP, 3>(s, P, npt, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, perm, C, ldc, plan);
    if constexpr (Terms<SP>::n == 2) {
      if (cbw == 4) return launch_split<SP, 4>(s, P, npt, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, perm, C, ldc, plan);
      if (cbw == 5) return launch_split<SP, 5>(s, P, npt, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, perm, C, ldc, plan);
    }
    return hipErrorInvalidValue;
  }
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
                             int64_t ldb, uint16_t *B, float *colinv, const int32_t *col_mo) {
  if (M <= 0) return hipSuccess;
  if (ldb % 64 != 0 || (form == QMCKL_AMD_SPLIT_FP32 && col_mo != nullptr)) return hipErrorInvalidValue;
  const unsigned g = (unsigned)split_cols_padded(form, M);
  const int frag = 16;  // fragment order of gemm_split_kernel's 16x16x32 tiles
  switch (form) {
    case QMCKL_AMD_SPLIT_FP16X2:
      hipLaunchKernelGGL(split_coef_kernel<QMCKL_AMD_SPLIT_FP16X2>, dim3(g), dim3(kThreads), 0, s, M, K, coef, ldb, frag, col_mo, B, colinv);
      break;
    case QMCKL_AMD_SPLIT_BF16X2:
      hipLaunchKernelGGL(split_coef_kernel<QMCKL_AMD_SPLIT_BF16X2>, dim3(g), dim3(kThreads), 0, s, M, K, coef, ldb, frag, col_mo, B, colinv);
      break;
    case QMCKL_AMD_SPLIT_BF16X3:
      hipLaunchKernelGGL(split_coef_kernel<QMCKL_AMD_SPLIT_BF16X3>, dim3(g), dim3(kThreads), 0, s, M, K, coef, ldb, frag, col_mo, B, colinv);
      break;
    case QMCKL_AMD_SPLIT_FP32:
      hipLaunchKernelGGL(split_coef_kernel<QMCKL_AMD_SPLIT_FP32>, dim3(g), dim3(kThreads), 0, s, M, K, coef, ldb, frag, col_mo, B, colinv);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t make_mixed_gemm_job(int form, int64_t P, int64_t K, int64_t Ncol, const float *A, int64_t lda,
                               const float *rowmax, const void *B, int64_t ldb, const float *colinv,
                               const uint8_t *flags, int nkb, const int32_t *perm, double *C, int64_t ldc,
                               ColPlan plan, MixedGemmJob *job, int gemm_tile) {
  if (form != QMCKL_AMD_SPLIT_FP16X2 && form != QMCKL_AMD_SPLIT_BF16X2 && form != QMCKL_AMD_SPLIT_BF16X3)
    return hipErrorInvalidValue;
  if (P <= 0 || Ncol <= 0 || K <= 0 || (lda % 64) != 0 || (ldb % 64) != 0 || lda < K || ldb < K)
    return hipErrorInvalidValue;
  if (flags != nullptr && (nkb > kMaxFlagBlocks || (int64_t)nkb * 32 < K)) return hipErrorInvalidValue;
  if (plan.blk != nullptr && ((K + kStage - 1) / kStage > kMaxStages || plan.nw * 64 * 16 < K || plan.nblk < 0))
    return hipErrorInvalidValue;
  const int T = form == QMCKL_AMD_SPLIT_BF16X3 ? 3 : 2;
  const int cbw = split_wg_cols(form, Ncol) / 64;
  if ((int64_t)4 * cbw * (ldb / 32) * T * 1024 >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
  const int64_t npt = (P + 63) / 64;
  job->P = P;
  job->npt = npt;
  job->K = K;
  job->Ncol = Ncol;
  job->A = A;
  job->lda = lda;
  job->rowmax = rowmax;
  job->Bs = static_cast<const uint16_t *>(B);
  job->ldb = ldb;
  job->colinv = colinv;
  job->flags = flags;
  job->nkb = nkb;
  job->perm = perm;
  job->C = C;
  job->ldc = ldc;
  job->plan = plan;
  job->cbw = cbw;
  job->dense = gemm_tile == QMCKL_GEMM_TILE_DENSE_DEVICE || gemm_tile == QMCKL_GEMM_TILE_DENSE64_DEVICE ? 1 : 0;
  job->item0 = 0;
  job->gx = 5 * npt * 2;  // 32-row halves of the 64-row tiles
  job->items = job->gx * ((Ncol + 64 * cbw - 1) / (64 * cbw));
  return hipSuccess;
}

namespace {
__global__ void zero_int_kernel(int *p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}
}  // namespace

hipError_t launch_mo_gemm_mixed_persistent(hipStream_t s, int form, int cbw, int dense, const MixedGemmJob *jobs,
                                           int njobs, int64_t total, int *counter) {
  if (njobs <= 0 || total <= 0) return hipSuccess;
  // tile widths up to 192 MOs: wider persistent tiles need more registers
  // than two workgroups per CU leave (the 320-wide one spills to scratch)
  if (cbw < 1 || cbw > 3) return hipErrorInvalidValue;
  if (total >= 0x7fffffff - 65536) return hipErrorInvalidValue;  // the counter is an int
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  const unsigned grid =
      (unsigned)(persistent_waves(cbw, split_terms(form)) * (cus > 0 ? cus : 1));  // as many as fit
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(64), 0, s, counter);
#define QMCKL_PERSISTENT(SPV, W)                                                                                   \
  do {                                                                                                            \
    if (dense)                                                                                                    \
      hipLaunchKernelGGL((gemm_split_persistent_kernel<SPV, W, false>), dim3(grid), dim3(256), 0, s, jobs, njobs, \
                         total, counter);                                                                         \
    else                                                                                                          \
      hipLaunchKernelGGL((gemm_split_persistent_kernel<SPV, W, true>), dim3(grid), dim3(256), 0, s, jobs, njobs,  \
                         total, counter);                                                                         \
  } while (0)
  switch (form * 8 + cbw) {
    case QMCKL_AMD_SPLIT_FP16X2 * 8 + 1: QMCKL_PERSISTENT(QMCKL_AMD_SPLIT_FP16X2, 1); break;
    case QMCKL_AMD_SPLIT_FP16X2 * 8 + 2: QMCKL_PERSISTENT(QMCKL_AMD_SPLIT_FP16X2, 2); break;
    case QMCKL_AMD_SPLIT_FP16X2 * 8 + 3: QMCKL_PERSISTENT(QMCKL_AMD_SPLIT_FP16X2, 3); break;
    case QMCKL_AMD_SPLIT_BF16X2 * 8 + 1: QMCKL_PERSISTENT(QMCKL_AMD_SPLIT_BF16X2, 1); break;
    case QMCKL_AMD_SPLIT_BF16X2 * 8 + 2: QMCKL_PERSISTENT(QMCKL_AMD_SPLIT_BF16X2, 2); break;
    case QMCKL_AMD_SPLIT_BF16X2 * 8 + 3: QMCKL_PERSISTENT(QMCKL_AMD_SPLIT_BF16X2, 3); break;
    case QMCKL_AMD_SPLIT_BF16X3 * 8 + 1: QMCKL_PERSISTENT(QMCKL_AMD_SPLIT_BF16X3, 1); break;
    case QMCKL_AMD_SPLIT_BF16X3 * 8 + 2: QMCKL_PERSISTENT(QMCKL_AMD_SPLIT_BF16X3, 2); break;
    case QMCKL_AMD_SPLIT_BF16X3 * 8 + 3: QMCKL_PERSISTENT(QMCKL_AMD_SPLIT_BF16X3, 3); break;
    default: return hipErrorInvalidValue;
  }
#undef QMCKL_PERSISTENT
  return hipGetLastError();
}

hipError_t launch_mo_gemm_mixed(hipStream_t s, int form, int64_t P, int64_t K, int64_t Ncol,
                                const float *A, int64_t lda, const float *rowmax, const void *B, int64_t ldb,
                                const float *colinv, const uint8_t *flags, int nkb, const int32_t *perm,
                                double *C, int64_t ldc, ColPlan plan, int gemm_tile) {
  if (P <= 0 || Ncol <= 0) return hipSuccess;
  if (K <= 0 || (lda % 64) != 0 || (ldb % 64) != 0 || lda < K || ldb < K)
    return hipErrorInvalidValue;
  const int64_t npt = (P + 63) / 64;
  switch (form) {
    case QMCKL_AMD_SPLIT_FP16X2:
      return launch_form<QMCKL_AMD_SPLIT_FP16X2>(s, P, npt, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, perm, C, ldc, plan, gemm_tile);
    case QMCKL_AMD_SPLIT_BF16X2:
      return launch_form<QMCKL_AMD_SPLIT_BF16X2>(s, P, npt, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, perm, C, ldc, plan, gemm_tile);
    case QMCKL_AMD_SPLIT_BF16X3:
      return launch_form<QMCKL_AMD_SPLIT_BF16X3>(s, P, npt, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, perm, C, ldc, plan, gemm_tile);
    case QMCKL_AMD_SPLIT_FP32:
      return launch_form<QMCKL_AMD_SPLIT_FP32>(s, P, npt, K, Ncol, A, lda, rowmax, B, ldb, colinv, flags, nkb, perm, C, ldc, plan, gemm_tile);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace qmckl_amd
