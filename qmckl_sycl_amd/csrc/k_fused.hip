// Fused AO evaluation and split MO contraction (FP16X2, BF16X2, BF16X3)
// for narrow MO sets: one kernel computes mo_vgl from the coordinates,
// without the FP32 AO operand that the two-kernel chain writes to HBM and
// reads back (the same arithmetic as that chain: src/qmckl_ao_sycl.cpp:
// 40-95, 291-331, 367-422 for the AOs, src/qmckl_mo_sycl.cpp:23-50 for the
// contraction).
//
// Workgroup: one 64-point tile (lane = point in the AO evaluation) x one
// group of 64 MOs, eight wavefronts, one workgroup per CU.  Its 320 rows
// are the five components of the 64 points, component-major, so a 16-row
// MFMA block is 16 points of one component; wavefront w owns MO block w % 4
// and row blocks 10 (w / 4) .. + 9: 80 FP64 and 40 FP32 accumulator
// registers per lane, which is why the kernel is for up to 64 MOs per
// workgroup (wider MO sets take one workgroup per group of 64, each
// evaluating the AOs again).
//   - Screen, once: the wavefronts take the shells 64 at a time (one shell
//     per lane for the loads, broadcast with readlane) and mark in local
//     memory the 64-AO stages with a shell nonzero at some point of the
//     tile.
//   - Per live stage: the wavefronts split the stage's shells (wavefront w
//     takes every eighth; their data loaded one shell per lane, each
//     shell's primitives one per lane, a shell ahead) and write their AOs'
//     five components, rounded to FP32, into the LDS slab [row][64 AOs];
//     each row is scaled by the power of two of its own stage maximum and
//     split into fp16 / bf16 terms, in place (8 lanes of one wavefront per
//     row: the row maximum is three shuffles); each wavefront runs its
//     10 x 2 x (3 or 6) MFMAs with the coefficient fragments of its MO
//     block (the split coefficients in the fragment order of the
//     stored-operand kernels, loaded at the start of the stage) and adds
//     the stage's FP32 sums, times the rows' inverse stage scales, into
//     FP64.
// Results round differently from the two-kernel chain, whose rows have one
// scale over all AOs; every result depends only on its own point's inputs,
// so the tile composition (point order) changes no bit.
//
// Measured on MI355X against the chain (AO operand kernel + dense split
// contraction), FP16X2: a 605-AO, 120-MO molecule at 91,537 points 1.41 ms
// against 0.27 + 0.47 ms; 280 AOs, 60 MOs at 47,600 points 0.21 against
// 0.09 + 0.12 ms.  One workgroup per CU runs its evaluation, split and
// MFMA phases one after the other with nothing to overlap them, where the
// chain keeps five AO wavefronts per SIMD and four contraction workgroups
// per CU resident.
#include <hip/hip_runtime.h>

#include "ao_eval.hpp"
#include "internal.hpp"
#include "split_ops.hpp"

namespace qmckl_amd {

namespace {

constexpr int kFT = 512;          // threads: eight wavefronts
constexpr int kFP = 64;           // points per workgroup
constexpr int kFRows = 5 * kFP;   // rows (component, point)
// LDS row stride in floats: a row holds its 64 AOs and, once split in
// place, their T terms of 64 halves each (2 terms: 64 floats, 3 terms: 96),
// plus a pad of 4 floats; a multiple of 4 for 16-byte accesses
template <int T>
constexpr int fstride() {
  return T == 3 ? 100 : 68;
}
constexpr int kFStage = 64;       // AOs per stage
constexpr int kFMaxStages = kFusedMaxLd / kFStage;  // stages of the live mask

// v at lane j, for every lane (j uniform)
__device__ __attribute__((always_inline)) inline double bcast(double v, int j) {
  const int64_t b = __builtin_bit_cast(int64_t, v);
  const int lo = __builtin_amdgcn_readlane((int)b, j), hi = __builtin_amdgcn_readlane((int)(b >> 32), j);
  return __builtin_bit_cast(double, ((int64_t)hi << 32) | (uint32_t)lo);
}

// Writes the AOs of this lane's point that fall in [k_lo, k_hi) into the
// slab rows (c, lane), rounded to FP32; the others are passed over.
struct SlabWriter {
  float *slab;
  int stride;
  int lane;
  int k, k_lo, k_hi;
  __device__ __attribute__((always_inline)) void push(const double (&v)[5]) {
    if (k >= k_lo && k < k_hi) {
#pragma unroll
      for (int c = 0; c < 5; ++c) slab[(c * kFP + lane) * stride + (k - k_lo)] = (float)v[c];
    }
    ++k;
  }
};

template <int SP>
__global__ void __attribute__((amdgpu_flat_work_group_size(kFT, kFT), amdgpu_waves_per_eu(2, 2)))
fused_ao_mo_kernel(int64_t P, int N, int64_t S, const double *__restrict__ coord,
                   const int32_t *__restrict__ perm, const double *__restrict__ nucl,
                   const ShellDesc *__restrict__ shells, const int32_t *__restrict__ stage_shell,
                   const double *__restrict__ ao_factor, const double *__restrict__ expo,
                   const double *__restrict__ coef, const double *__restrict__ nucl_range, int64_t A,
                   int64_t ld, const uint16_t *__restrict__ Bs, const float *__restrict__ colinv, int64_t M,
                   double *__restrict__ C, int64_t ldc) {
  constexpr int T = Terms<SP>::n;
  constexpr int kFStride = fstride<T>();
  __shared__ __attribute__((aligned(16))) float slab[kFRows * kFStride];
  __shared__ float s_inv[kFRows];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wv = __builtin_amdgcn_readfirstlane(wave);
  const int cq = wv & 3;            // this wavefront's MO block of 16
  const int rb0 = (wv >> 2) * 10;   // its first row block
  const int64_t pt = blockIdx.x, gb = blockIdx.y;

  // this lane's point for the AO evaluation
  const int64_t i = pt * kFP + lane;
  const bool inside = i < P;
  const int64_t p = inside ? (perm != nullptr ? (int64_t)perm[i] : i) : 0;
  const double px = inside ? coord[p] : 0.0;
  const double py = inside ? coord[P + p] : 0.0;
  const double pz = inside ? coord[2 * P + p] : 0.0;

  // coefficient fragments of the workgroup's four MO blocks (buffer
  // resource; lane offset 16 l bytes, block and 32-K chunk in the scalar
  // offset)
  const int64_t nkc = ld / 32;
  const int tile_bytes = (int)(nkc * T * 1024);
  const __amdgpu_buffer_rsrc_t brs = __builtin_amdgcn_make_buffer_rsrc(
      (void *)(Bs + gb * 4 * nkc * T * 512), (short)0, 4 * tile_bytes, 0x00020000);

  double acc[10][4];
#pragma unroll
  for (int r = 0; r < 10; ++r)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[r][j] = 0.0;

  // The screen, once per workgroup: a 64-AO stage is live for the tile if
  // some shell with an AO in it is nonzero at some of the 64 points.  Each
  // wavefront takes chunks of 64 shells, one shell per lane for the loads,
  // broadcast with readlane; a live shell sets the bits of the (at most two)
  // stages it touches.
  const int nst_a = (int)((A + kFStage - 1) / kFStage);
  __shared__ uint32_t s_live[kFMaxStages / 32];
  if (tid < kFMaxStages / 32) s_live[tid] = 0u;
  __syncthreads();
  for (int64_t s0 = (int64_t)wv * 64; s0 < S; s0 += (kFT / 64) * 64) {
    const int64_t sl = s0 + lane;
    double nx = 0.0, ny = 0.0, nz = 0.0, lim = -1.0, amin = 0.0;
    int stages = 0;  // first stage | last stage << 16
    if (sl < S) {
      const ShellDesc d = shells[sl];
      nx = nucl[d.nucl];
      ny = nucl[N + d.nucl];
      nz = nucl[2 * N + d.nucl];
      lim = kCutoff * nucl_range[d.nucl];
      amin = d.a_min;
      const int nc = (d.l + 1) * (d.l + 2) / 2;
      stages = (d.ao_start / kFStage) | (((d.ao_start + nc - 1) / kFStage) << 16);
    }
    const int cnt = (int)min<int64_t>(64, S - s0);
    for (int j = 0; j < cnt; ++j) {
      const double x = px - bcast(nx, j), y = py - bcast(ny, j), z = pz - bcast(nz, j);
      const double r2 = fma(x, x, fma(y, y, z * z));
      const bool live = inside && !(r2 > bcast(lim, j)) && !(bcast(amin, j) * r2 > kCutoff);
      if (__ballot(live) != 0 && lane == 0) {
        const int sb = __builtin_amdgcn_readlane(stages, j);
        const int a = sb & 0xffff, b = sb >> 16;
        atomicOr(&s_live[a >> 5], 1u << (a & 31));
        if (b != a) atomicOr(&s_live[b >> 5], 1u << (b & 31));
      }
    }
  }
  __syncthreads();

  for (int st = 0; st < nst_a; ++st) {
    const int k_lo = kFStage * st, k_end = k_lo + kFStage;
    if (((s_live[st >> 5] >> (st & 31)) & 1u) == 0u) continue;  // all its AOs are zero for the tile
    // the stage's coefficient fragments, loaded now for use after the AO
    // evaluation and split
    uint4 bq[2][T];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
      for (int t = 0; t < T; ++t)
        bq[ks][t] = __builtin_bit_cast(
            uint4, __builtin_amdgcn_raw_buffer_load_b128(brs, 16 * lane + t * 1024,
                                                         cq * tile_bytes + (2 * st + ks) * T * 1024, 0));
    const int64_t s_first = stage_shell[st];

    __syncthreads();  // the previous stage's slab has been read
    if (k_end > A) {  // AOs past ao_num in the last stage are zeros
      for (int e = tid; e < kFRows * kFStage; e += kFT) {
        const int r = e / kFStage, k = e - r * kFStage;
        if (k_lo + k >= A) slab[r * kFStride + k] = 0.0f;
      }
    }
    SlabWriter w;
    w.slab = slab;
    w.stride = kFStride;
    w.lane = lane;
    w.k_lo = k_lo;
    w.k_hi = k_end;
    // this wavefront's shells of the stage, s_first + wv + 8 i, one per lane
    // for the loads (a stage has at most 64 + 1 shells)
    const int64_t sl = s_first + wv + (int64_t)(kFT / 64) * lane;
    double nx = 0.0, ny = 0.0, nz = 0.0, lim = -1.0, amin = 0.0;
    int ao0 = 0x7fffffff, lsh = 0, pst = 0, pnum = 0;
    if (sl < S) {
      const ShellDesc d = shells[sl];
      nx = nucl[d.nucl];
      ny = nucl[N + d.nucl];
      nz = nucl[2 * N + d.nucl];
      lim = kCutoff * nucl_range[d.nucl];
      amin = d.a_min;
      ao0 = d.ao_start;
      lsh = d.l;
      pst = d.prim_start;
      pnum = d.prim_num;
    }
    // the primitives of the wavefront's current shell, one per lane, loaded
    // one shell ahead
    int cur_ps = __builtin_amdgcn_readlane(pst, 0), cur_pn = __builtin_amdgcn_readlane(pnum, 0);
    double ex = 0.0, cf = 0.0;
    if (lane < cur_pn && __builtin_amdgcn_readlane(ao0, 0) < k_end) {
      ex = expo[cur_ps + lane];
      cf = coef[cur_ps + lane];
    }
    for (int j = 0; j < 64; ++j) {
      const int a0 = __builtin_amdgcn_readlane(ao0, j);
      if (a0 >= k_end) break;
      // the next shell's primitives
      double ex_n = 0.0, cf_n = 0.0;
      int nx_ps = 0, nx_pn = 0;
      if (j + 1 < 64) {
        nx_ps = __builtin_amdgcn_readlane(pst, j + 1);
        nx_pn = __builtin_amdgcn_readlane(pnum, j + 1);
        if (lane < nx_pn && __builtin_amdgcn_readlane(ao0, j + 1) < k_end) {
          ex_n = expo[nx_ps + lane];
          cf_n = coef[nx_ps + lane];
        }
      }
      w.k = a0;
      const int l = __builtin_amdgcn_readlane(lsh, j);
      const double x = px - bcast(nx, j), y = py - bcast(ny, j), z = pz - bcast(nz, j);
      const double r2 = fma(x, x, fma(y, y, z * z));
      const bool live = inside && !(r2 > bcast(lim, j));
      double R[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      if (__ballot(live && !(bcast(amin, j) * r2 > kCutoff)) != 0) {
        for (int q = 0; q < cur_pn; ++q) {
          const double a = q < 64 ? bcast(ex, q) : expo[cur_ps + q];
          const double c = q < 64 ? bcast(cf, q) : coef[cur_ps + q];
          const double ar2 = a * r2;
          if (live && !(ar2 > kCutoff)) {
            const double v = c * exp_neg_f32(ar2);
            const double two_a = -2.0 * a * v;
            R[0] += v;
            R[1] += two_a * x;
            R[2] += two_a * y;
            R[3] += two_a * z;
            R[4] += two_a * (3.0 - 2.0 * ar2);
          }
        }
        switch (l) {
          case 0: shell_aos<0>(w, ao_factor, x, y, z, R); break;
          case 1: shell_aos<1>(w, ao_factor, x, y, z, R); break;
          case 2: shell_aos<2>(w, ao_factor, x, y, z, R); break;
          case 3: shell_aos<3>(w, ao_factor, x, y, z, R); break;
          default: shell_aos_any(w, ao_factor, x, y, z, R, l); break;
        }
      } else {  // screened for every point of the tile: zeros
        const double zero[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        const int nc = (l + 1) * (l + 2) / 2;
        for (int c = 0; c < nc; ++c) w.push(zero);
      }
      ex = ex_n;
      cf = cf_n;
      cur_ps = nx_ps;
      cur_pn = nx_pn;
    }
    __syncthreads();

    // each row scaled by its stage maximum and split in place: 8 lanes of
    // one wavefront per row, 8 AOs each (all 8 read before any writes:
    // the writes depend on the shuffled maximum)
#pragma unroll
    for (int rep = 0; rep < kFRows * 8 / kFT; ++rep) {
      const int chunk = tid + rep * kFT;
      const int row = chunk >> 3, q = chunk & 7;
      float *rp = slab + row * kFStride;
      const float4 v0 = *reinterpret_cast<const float4 *>(rp + 8 * q);
      const float4 v1 = *reinterpret_cast<const float4 *>(rp + 8 * q + 4);
      float m = fmaxf(fmaxf(fmaxf(fabsf(v0.x), fabsf(v0.y)), fmaxf(fabsf(v0.z), fabsf(v0.w))),
                      fmaxf(fmaxf(fabsf(v1.x), fabsf(v1.y)), fmaxf(fabsf(v1.z), fabsf(v1.w))));
      m = fmaxf(m, __shfl_xor(m, 1));
      m = fmaxf(m, __shfl_xor(m, 2));
      m = fmaxf(m, __shfl_xor(m, 4));
      const float sc = pow2_scale(m);
      const float v[8] = {v0.x * sc, v0.y * sc, v0.z * sc, v0.w * sc, v1.x * sc, v1.y * sc, v1.z * sc, v1.w * sc};
      uint4 o[T];
      split8<SP>(v, o);
      // term t of the row's 64 AOs at halves [64 t, 64 t + 64) of the row
      uint16_t *hp = reinterpret_cast<uint16_t *>(rp);
#pragma unroll
      for (int t = 0; t < T; ++t) *reinterpret_cast<uint4 *>(hp + 64 * t + 8 * q) = o[t];
      if (q == 0) s_inv[row] = 1.0f / sc;
    }
    __syncthreads();

    // the wavefront's MFMAs over the stage's two 32-AO halves, five row
    // blocks at a time, each product issued over the five before the next
    // (five independent accumulators between dependent MFMAs)
    f4 dacc[10];
    constexpr int NP = T == 3 ? 6 : 3;
    constexpr int pa[6] = {0, 0, 1, 0, 2, 1}, pb[6] = {0, 1, 0, 2, 0, 1};  // (A term, C term)
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        uint4 a[5][T];
#pragma unroll
        for (int r = 0; r < 5; ++r) {
          const uint16_t *hp = reinterpret_cast<const uint16_t *>(slab + ((rb0 + 5 * g + r) * 16 + li) * kFStride);
#pragma unroll
          for (int t = 0; t < T; ++t) a[r][t] = *reinterpret_cast<const uint4 *>(hp + 64 * t + 32 * ks + 8 * lk);
        }
#pragma unroll
        for (int pr = 0; pr < NP; ++pr)
#pragma unroll
          for (int r = 0; r < 5; ++r)
            dacc[5 * g + r] = mfma<SP>(a[r][pa[pr]], bq[ks][pb[pr]],
                                       pr == 0 && ks == 0 ? f4{0.f, 0.f, 0.f, 0.f} : dacc[5 * g + r]);
      }
    }
#pragma unroll
    for (int r = 0; r < 10; ++r)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[r][j] += (double)dacc[r][j] * (double)s_inv[(rb0 + r) * 16 + 4 * lk + j];
  }

  // the rows (c, point) of this wavefront's blocks, its MO column
  const int64_t col = gb * 64 + cq * 16 + li;
  if (col >= M) return;
  const double ci = (double)colinv[col];
#pragma unroll
  for (int r = 0; r < 10; ++r)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = (rb0 + r) * 16 + 4 * lk + j;
      const int c = row / kFP, pl = row % kFP;
      const int64_t ip = pt * kFP + pl;
      if (ip < P) {
        const int64_t pp = perm != nullptr ? (int64_t)perm[ip] : ip;
        C[(pp * 5 + c) * ldc + col] = acc[r][j] * ci;
      }
    }
}

}  // namespace

hipError_t launch_fused_ao_mo(hipStream_t s, int form, const DeviceBasis &b, int64_t P, const double *coord,
                              const int32_t *perm, const double *nucl, const double *ao_factor, int64_t ld,
                              const uint16_t *Bs, const float *colinv, int64_t M, double *C, int64_t ldc) {
  if (P <= 0 || M <= 0) return hipSuccess;
  const int64_t A = b.ao_num;
  if (b.max_l > kMaxL || ld % kFStage != 0 || ld < A || ld > kFusedMaxLd || b.stage_shell == nullptr)
    return hipErrorInvalidValue;
  // the coefficient buffer holds M rounded up to the split width (a multiple of 64)
  if (split_cols_padded(form, M) % 64 != 0 || (int64_t)4 * (ld / 32) * 3 * 1024 >= ((int64_t)1 << 31))
    return hipErrorInvalidValue;
  const int64_t gx = (P + kFP - 1) / kFP, gy = (M + 63) / 64;
  if (gx > 0x7fffffff || gy > 65535 || P > 0x7fffffff) return hipErrorInvalidValue;
  const dim3 grid((unsigned)gx, (unsigned)gy);
  switch (form) {
    case QMCKL_AMD_SPLIT_FP16X2:
      hipLaunchKernelGGL(fused_ao_mo_kernel<QMCKL_AMD_SPLIT_FP16X2>, grid, dim3(kFT), 0, s, P, (int)b.nucl_num,
                         b.shell_num, coord, perm, nucl, b.shells, b.stage_shell, ao_factor, b.expo, b.coef_norm,
                         b.nucl_range, A, ld, Bs, colinv, M, C, ldc);
      break;
    case QMCKL_AMD_SPLIT_BF16X2:
      hipLaunchKernelGGL(fused_ao_mo_kernel<QMCKL_AMD_SPLIT_BF16X2>, grid, dim3(kFT), 0, s, P, (int)b.nucl_num,
                         b.shell_num, coord, perm, nucl, b.shells, b.stage_shell, ao_factor, b.expo, b.coef_norm,
                         b.nucl_range, A, ld, Bs, colinv, M, C, ldc);
      break;
    case QMCKL_AMD_SPLIT_BF16X3:
      hipLaunchKernelGGL(fused_ao_mo_kernel<QMCKL_AMD_SPLIT_BF16X3>, grid, dim3(kFT), 0, s, P, (int)b.nucl_num,
                         b.shell_num, coord, perm, nucl, b.shells, b.stage_shell, ao_factor, b.expo, b.coef_norm,
                         b.nucl_range, A, ld, Bs, colinv, M, C, ldc);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace qmckl_amd
