// AO operand of the mixed-precision MO contraction, one lane per point.
//
// Same arithmetic as the FP64 AO kernel (src/qmckl_ao_sycl.cpp: radial
// shell sums with the nucleus screen r^2 > cutoff*range and the primitive
// screen alpha r^2 > cutoff, 40-95; Cartesian monomials x^a y^b z^c in order
// of descending a then b, 291-331; product-rule assembly times ao_factor,
// 367-422), evaluated in FP64 and rounded once to FP32 at the store.
//
// The operand is internal to the contraction, so its layout is chosen for
// the two kernels that touch it:
//   buf[c][pt][kq][pl][4]   c = VGL component (5), pt = 64-point tile,
//                           kq = AO index / 4, pl = point in tile,
//                           the 4 consecutive AOs 4kq .. 4kq+3
// i.e. the GEMM row (c, point) of a 64-row tile has its AOs in 16-byte
// pieces, 64 of which (one per point of the tile) form a contiguous
// kilobyte for each group of 4 AOs.
// AOs past ao_num up to ld (a multiple of 64) are written as zeros.
// Also written: rowmax[c][tile row] = max |entry| of the row (each lane owns
// its five rows, so a wavefront needs no reduction; the AO groups of a tile
// combine theirs with atomic integer maxima of the bit patterns), and
// flags per (row tile, 32-AO block): 1 where some shell overlapping the
// block passes the screen for some point of the tile (the same flag for the
// five components), 0 where every value of the block is zero by the screen.
//
// The 64 points of tile t are perm[64t .. 64t+63] (the caller orders the
// points so that a tile is spatially compact).  With flags, a 64-AO stage
// whose shells are all screened for the tile's 64 points is neither
// evaluated nor written; its flags are zero, so the contraction skips it.
//
// Work decomposition: a wavefront owns 64 points and one group of 64-AO
// stages (enough groups per point tile for about four wavefronts per SIMD)
// and walks each live stage's shells in AO order (a shell that straddles a
// stage boundary is evaluated for each stage, each writing its own AOs);
// the shell loop is
// wavefront-uniform (shell data, nucleus and
// primitives come through scalar loads; a primitive screened for all 64
// points costs no exp), and for l <= 3 the shell's Cartesian components
// are unrolled at compile time, so each AO costs a few FP64 multiplies on
// powers computed once per (point, shell).  No barriers: the four
// wavefronts of a workgroup are independent, each with its own LDS ring for
// assembling 16-byte pieces.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ao_eval.hpp"
#include "internal.hpp"

namespace qmckl_amd {

namespace {

constexpr int kThreads = 256;

// Running state of one lane: the AO counter of the wavefront, row maxima
// and the screen of the current stage's 32-AO halves.  Values are staged in the
// wavefront's LDS ring [component][AO mod 4][lane] (the slot is a runtime
// index, which registers cannot take without spilling); after every fourth
// AO each lane reads its four values back and stores them as one 16-byte
// piece, so each store instruction writes one contiguous kilobyte.  With LO
// (the INT8X5 form) the residual (float)(v - (double)(float)v) of every
// value goes the same way into a second array of the same layout, so that
// hi + lo carries the FP64 value to 2^-48 relative.
template <bool LO>
struct RowWriter {
  float mx[5];
  uint32_t hscreen;  // bit h: half h of the current stage passes the screen (its flag)
  int k;             // next AO index (wavefront-uniform)
  float *base;       // &buf[0][pt][0][lane][0]
  int64_t cstride;   // floats between components
  uint8_t *flags;    // flags + pt * nkb, component stride fstride
  int64_t fstride;
  float *ring;       // this wavefront's [5][4][64] LDS ring
  float *ring_lo;    // the residuals' ring (LO)
  float *base_lo;    // &buf_lo[0][pt][0][lane][0] (LO)
  int lane;
  uint32_t hlive;    // bit h: 32-AO half h of the current stage is stored

  int k_lo, k_hi;    // AOs of the current stage

  __device__ __attribute__((always_inline)) void push(const double (&v)[5]) {
    if (k < k_lo || k >= k_hi) {  // part of a shell outside the current stage
      ++k;
      return;
    }
    const int slot = k & 3;
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      const float w = (float)v[c];
      ring[(c * 4 + slot) * 64 + lane] = w;
      if constexpr (LO) ring_lo[(c * 4 + slot) * 64 + lane] = (float)(v[c] - (double)w);
      mx[c] = fmaxf(mx[c], fabsf(w));
    }
    if (slot == 3 && ((hlive >> ((k >> 5) & 1)) & 1u)) {
      // nontemporal stores: the operand is far larger than the caches and is
      // read once, later, by the contraction (1% on water64 and on the
      // 48-molecule batch against ordinary stores)
      typedef float v4f __attribute__((ext_vector_type(4)));
      float *o = base + (int64_t)(k >> 2) * 256;
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        const v4f v = {ring[(c * 4 + 0) * 64 + lane], ring[(c * 4 + 1) * 64 + lane],
                       ring[(c * 4 + 2) * 64 + lane], ring[(c * 4 + 3) * 64 + lane]};
        __builtin_nontemporal_store(v, reinterpret_cast<v4f *>(o + c * cstride));
      }
      if constexpr (LO) {
        float *ol = base_lo + (int64_t)(k >> 2) * 256;
#pragma unroll
        for (int c = 0; c < 5; ++c) {
          const v4f v = {ring_lo[(c * 4 + 0) * 64 + lane], ring_lo[(c * 4 + 1) * 64 + lane],
                         ring_lo[(c * 4 + 2) * 64 + lane], ring_lo[(c * 4 + 3) * 64 + lane]};
          __builtin_nontemporal_store(v, reinterpret_cast<v4f *>(ol + c * cstride));
        }
      }
    }
    // the flags of a finished 32-AO block, from the screen (per-value
    // nonzero tests cost three VALU instructions per value)
    if ((k & 31) == 31 && flags != nullptr && lane < 5)
      flags[lane * fstride + (k >> 5)] = (uint8_t)((hscreen >> ((k >> 5) & 1)) & 1u);
    ++k;
  }
};

// Is any AO of the shells from s_first that start below k_end nonzero for
// some point of the wavefront?  Same screen as eval_stage: a shell is zero
// at r^2 when a * r^2 > cutoff for all of its exponents a (a_min * r^2
// rounds no higher than a * r^2), or past the nucleus range.
__device__ __attribute__((always_inline)) inline bool stage_live(int64_t s_first, int64_t S, int k_end, int N,
                                                                 bool inside, double px, double py, double pz,
                                                                 const double *__restrict__ nucl,
                                                                 const ShellDesc *__restrict__ shells,
                                                                 const double *__restrict__ nucl_range) {
  bool any = false;
  for (int64_t s = s_first; s < S; ++s) {
    const ShellDesc d = shells[s];
    if (d.ao_start >= k_end) break;
    const double x = px - nucl[d.nucl];
    const double y = py - nucl[N + d.nucl];
    const double z = pz - nucl[2 * N + d.nucl];
    const double r2 = fma(x, x, fma(y, y, z * z));
    any |= inside && !(r2 > kCutoff * nucl_range[d.nucl]) && !(d.a_min * r2 > kCutoff);
  }
  return __ballot(any) != 0;
}

// This lane's screen of the 64-AO stage [k_lo, k_end) from shell s_first
// on, per 32-AO half: bit 0 (1) when some shell overlapping the first
// (second) half passes it.  A shell is zero at r^2 when a * r^2 > cutoff for
// all of its exponents a (a_min * r^2 rounds no higher than a * r^2), or
// past the nucleus range.
__device__ __attribute__((always_inline)) inline uint32_t lane_screen(int64_t s_first, int64_t S, int k_lo,
                                                                      int k_end, int N,
                                                                      bool inside, double px, double py, double pz,
                                                                      const double *__restrict__ nucl,
                                                                      const ShellDesc *__restrict__ shells,
                                                                      const double *__restrict__ nucl_range) {
  uint32_t h = 0;
  const int k_mid = k_lo + 32;
  for (int64_t s = s_first; s < S; ++s) {
    const ShellDesc d = shells[s];
    if (d.ao_start >= k_end) break;
    const double x = px - nucl[d.nucl];
    const double y = py - nucl[N + d.nucl];
    const double z = pz - nucl[2 * N + d.nucl];
    const double r2 = fma(x, x, fma(y, y, z * z));
    const bool live = inside && !(r2 > kCutoff * nucl_range[d.nucl]) && !(d.a_min * r2 > kCutoff);
    const int a_end = d.ao_start + (d.l + 1) * (d.l + 2) / 2;
    h |= (live && d.ao_start < k_mid ? 1u : 0u) | (live && a_end > k_mid ? 2u : 0u);
  }
  return h;
}

// Walks the shells from s_first that start below k_end in AO order for
// this lane's point and pushes their AOs into w (the writer keeps those of
// its current stage).
template <typename W>
__device__ __attribute__((always_inline)) inline void eval_stage(W &w, int64_t s_first, int64_t S, int k_end,
                                                                 int N, bool inside, double px, double py,
                                                                 double pz, const double *__restrict__ nucl,
                                                                 const ShellDesc *__restrict__ shells,
                                                                 const double *__restrict__ ao_factor,
                                                                 const double *__restrict__ expo,
                                                                 const double *__restrict__ coef,
                                                                 const double *__restrict__ nucl_range) {
  for (int64_t s = s_first; s < S; ++s) {
    const ShellDesc d = shells[s];
    if (d.ao_start >= k_end) break;
    const double x = px - nucl[d.nucl];
    const double y = py - nucl[N + d.nucl];
    const double z = pz - nucl[2 * N + d.nucl];
    const double r2 = fma(x, x, fma(y, y, z * z));
    const bool live = inside && !(r2 > kCutoff * nucl_range[d.nucl]);
    double R[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const double *ex = expo + d.prim_start;
    const double *cf = coef + d.prim_start;
    for (int j = 0; j < d.prim_num; ++j) {
      const double a = ex[j];
      const double ar2 = a * r2;
      if (live && !(ar2 > kCutoff)) {
        const double v = cf[j] * exp(-ar2);
        const double two_a = -2.0 * a * v;
        R[0] += v;
        R[1] += two_a * x;
        R[2] += two_a * y;
        R[3] += two_a * z;
        R[4] += two_a * (3.0 - 2.0 * ar2);
      }
    }
    switch (d.l) {
      case 0: shell_aos<0>(w, ao_factor, x, y, z, R); break;
      case 1: shell_aos<1>(w, ao_factor, x, y, z, R); break;
      case 2: shell_aos<2>(w, ao_factor, x, y, z, R); break;
      case 3: shell_aos<3>(w, ao_factor, x, y, z, R); break;
      default: shell_aos_any(w, ao_factor, x, y, z, R, d.l); break;
    }
  }
}

// Pointer types of the read-only inputs: the per-context kernel's
// arguments, or, in the batched kernel (whose pointers come from a table in
// memory), the constant address space, so that the wavefront-uniform reads
// of the basis tables stay scalar loads.
struct ArgPtrs {
  template <class T>
  using p = const T *__restrict__;
};
struct TablePtrs {
#if defined(__HIP_DEVICE_COMPILE__)
  template <class T>
  using p = const __attribute__((address_space(4))) T *;
#else  // the host pass compiles the kernel bodies without address spaces
  template <class T>
  using p = const T *;
#endif
};

// The work of wavefront wi (< npt * groups) of one operand: point tile
// wi / groups, AO group wi % groups; ring / ring_lo are its LDS rings.
template <bool LO, class PT>
__device__ __attribute__((always_inline)) inline void ao_rows_wave(
    int64_t wi, int64_t P, int N, int64_t S, typename PT::template p<double> coord,
    typename PT::template p<int32_t> perm, typename PT::template p<double> nucl,
    typename PT::template p<ShellDesc> shells, typename PT::template p<int32_t> stage_shell,
    typename PT::template p<double> ao_factor, typename PT::template p<double> expo,
    typename PT::template p<double> coef, typename PT::template p<double> nucl_range, int64_t A, int64_t ld,
    int64_t npt, int groups, float *__restrict__ buf, float *__restrict__ buf_lo, float *__restrict__ rowmax,
    uint8_t *__restrict__ flags, int nkb, int skip_halves, float *ring, float *ring_lo) {
  const int lane = threadIdx.x & 63;
  const int64_t pt = wi / groups;
  const int g = (int)(wi - pt * groups);
  if (pt >= npt) return;  // wavefront-uniform
  const int64_t i = pt * 64 + lane;
  const bool inside = i < P;
  const int64_t p = inside ? (perm != nullptr ? (int64_t)perm[i] : i) : 0;
  const double px = inside ? coord[p] : 0.0;
  const double py = inside ? coord[P + p] : 0.0;
  const double pz = inside ? coord[2 * P + p] : 0.0;

  // this wavefront's AOs: 64-AO stages [g*nst/groups, (g+1)*nst/groups)
  const int64_t nst = ld / 64, nst_a = (A + 63) / 64;
  const int64_t st_lo = g * nst / groups, st_hi = (g + 1) * nst / groups;
  RowWriter<LO> w;
#pragma unroll
  for (int c = 0; c < 5; ++c) w.mx[c] = 0.0f;
  w.hscreen = 3u;
  w.cstride = npt * (ld / 4) * 256;
  w.base = buf + pt * (ld / 4) * 256 + lane * 4;
  w.flags = flags != nullptr ? flags + pt * nkb : nullptr;
  w.fstride = npt * nkb;
  w.lane = lane;
  w.ring = ring;
  w.ring_lo = ring_lo;
  w.base_lo = LO ? buf_lo + pt * (ld / 4) * 256 + lane * 4 : nullptr;
  w.hlive = 3u;

  // Stage by stage; a shell that straddles a stage boundary is evaluated
  // for each stage it reaches (each writes its own AOs).  The screen and the
  // shell walk are the same as stage_live and eval_stage, written out here:
  // through those functions the kernel takes 105 VGPRs instead of 94, four
  // wavefronts per SIMD instead of five, and 1.32 instead of 1.06 ms on
  // water64 at 128 walkers.
  for (int64_t st = st_lo; st < st_hi; ++st) {
    const int k_end = (int)(64 * (st + 1));
    const int64_t s_first = st < nst_a ? stage_shell[st] : S;
    if (flags != nullptr) {
      // the same screen per 32-AO half (a shell counts for each half it
      // overlaps): with skip_halves a half zero for all 64 points is not
      // stored either (its flags come out zero, and the contractions that
      // take skip_halves never read a half whose flag is zero)
      bool any0 = false, any1 = false;
      const int k_mid = k_end - 32;
      for (int64_t s = s_first; s < S; ++s) {
        const ShellDesc d = shells[s];
        if (d.ao_start >= k_end) break;
        const double x = px - nucl[d.nucl];
        const double y = py - nucl[N + d.nucl];
        const double z = pz - nucl[2 * N + d.nucl];
        const double r2 = fma(x, x, fma(y, y, z * z));
        const bool live = inside && !(r2 > kCutoff * nucl_range[d.nucl]) && !(d.a_min * r2 > kCutoff);
        const int a_end = d.ao_start + (d.l + 1) * (d.l + 2) / 2;
        any0 |= live && d.ao_start < k_mid;
        any1 |= live && a_end > k_mid;
      }
      const bool h0 = __ballot(any0) != 0, h1 = __ballot(any1) != 0;
      if (!h0 && !h1) {
        if (lane < 10) flags[pt * nkb + (lane >> 1) * w.fstride + 2 * st + (lane & 1)] = 0;
        continue;
      }
      w.hscreen = (h0 ? 1u : 0u) | (h1 ? 2u : 0u);
      w.hlive = skip_halves ? w.hscreen : 3u;
    }
    w.k_lo = (int)(64 * st);
    w.k_hi = k_end;
    w.k = s_first < S ? shells[s_first].ao_start : w.k_lo;
    for (int64_t s = s_first; s < S; ++s) {
      const ShellDesc d = shells[s];
      if (d.ao_start >= k_end) break;
      const double x = px - nucl[d.nucl];
      const double y = py - nucl[N + d.nucl];
      const double z = pz - nucl[2 * N + d.nucl];
      const double r2 = fma(x, x, fma(y, y, z * z));
      const bool live = inside && !(r2 > kCutoff * nucl_range[d.nucl]);
      double R[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      const auto ex = expo + d.prim_start;
      const auto cf = coef + d.prim_start;
      for (int j = 0; j < d.prim_num; ++j) {
        const double a = ex[j];
        const double ar2 = a * r2;
        if (live && !(ar2 > kCutoff)) {
          const double v = cf[j] * (LO ? exp(-ar2) : exp_neg_f32(ar2));
          const double two_a = -2.0 * a * v;
          R[0] += v;
          R[1] += two_a * x;
          R[2] += two_a * y;
          R[3] += two_a * z;
          R[4] += two_a * (3.0 - 2.0 * ar2);
        }
      }
      switch (d.l) {
        case 0: shell_aos<0>(w, ao_factor, x, y, z, R); break;
        case 1: shell_aos<1>(w, ao_factor, x, y, z, R); break;
        case 2: shell_aos<2>(w, ao_factor, x, y, z, R); break;
        case 3: shell_aos<3>(w, ao_factor, x, y, z, R); break;
        default: shell_aos_any(w, ao_factor, x, y, z, R, d.l); break;
      }
    }
    // zero AOs from ao_num up to the stage's end (the padding the
    // contraction reads, and the last 32-block's flags)
    if (w.k < w.k_lo) w.k = w.k_lo;
    const double zero[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    while (w.k < w.k_hi) w.push(zero);
  }
  // row maxima [5][npt*64]: the max over the AO groups, as an integer max
  // of the bit patterns (the order of non-negative floats); the array was
  // zeroed by fill_zero_kernel
  uint32_t *rm = reinterpret_cast<uint32_t *>(rowmax) + pt * 64 + lane;
#pragma unroll
  for (int c = 0; c < 5; ++c)
    if (inside && w.mx[c] > 0.0f) atomicMax(rm + c * npt * 64, __float_as_uint(w.mx[c]));
}

template <bool LO>
__global__ void __attribute__((amdgpu_flat_work_group_size(kThreads, kThreads), amdgpu_waves_per_eu(LO ? 4 : 5)))
ao_rows_kernel(int64_t P, int N, int64_t S, const double *__restrict__ coord, const int32_t *__restrict__ perm,
               const double *__restrict__ nucl, const ShellDesc *__restrict__ shells,
               const int32_t *__restrict__ stage_shell, const double *__restrict__ ao_factor,
               const double *__restrict__ expo, const double *__restrict__ coef,
               const double *__restrict__ nucl_range, int64_t A, int64_t ld, int64_t npt, int groups,
               float *__restrict__ buf, float *__restrict__ buf_lo, float *__restrict__ rowmax,
               uint8_t *__restrict__ flags, int nkb, int skip_halves) {
  __shared__ float s_ring[kThreads / 64][5 * 4 * 64];
  __shared__ float s_ring_lo[LO ? kThreads / 64 : 1][LO ? 5 * 4 * 64 : 1];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t wi = (int64_t)blockIdx.x * (kThreads / 64) + wave;
  ao_rows_wave<LO, ArgPtrs>(wi, P, N, S, coord, perm, nucl, shells, stage_shell, ao_factor, expo, coef, nucl_range,
                            A, ld, npt, groups, buf, buf_lo, rowmax, flags, nkb, skip_halves, s_ring[wave],
                            LO ? s_ring_lo[wave] : nullptr);
}

// The operands of a batch of contexts in one launch: wavefront wi belongs to
// the last job whose first wavefront is <= wi (jobs in order of wave0).
// One launch instead of one per molecule: a small molecule's kernel runs a
// few hundred wavefronts, each walking all shells of its stages, so the
// per-molecule launches left the GPU mostly idle behind four hardware queues.
__global__ void __attribute__((amdgpu_flat_work_group_size(kThreads, kThreads), amdgpu_waves_per_eu(5)))
ao_rows_batch_kernel(const AoRowsJob *__restrict__ jobs, int njobs, int64_t total) {
  __shared__ float s_ring[kThreads / 64][5 * 4 * 64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t wi = (int64_t)blockIdx.x * (kThreads / 64) + wave;
  if (wi >= total) return;
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (jobs[mid].wave0 <= wi) lo = mid;
    else hi = mid - 1;
  }
  const AoRowsJob &J = jobs[lo];
  using K = TablePtrs;
  ao_rows_wave<false, K>(wi - J.wave0, J.P, J.N, J.S, (K::p<double>)J.coord, (K::p<int32_t>)J.perm,
                         (K::p<double>)J.nucl, (K::p<ShellDesc>)J.shells, (K::p<int32_t>)J.stage_shell,
                         (K::p<double>)J.ao_factor, (K::p<double>)J.expo, (K::p<double>)J.coef,
                         (K::p<double>)J.nucl_range, J.A, J.ld, J.npt, J.groups, J.buf, nullptr, J.rowmax, J.flags,
                         J.nkb, J.skip_halves, s_ring[wave], nullptr);
}

// Zeroes the row maxima of every job (blockIdx.y = job).
__global__ void __launch_bounds__(kThreads) fill_rowmax_batch_kernel(const AoRowsJob *__restrict__ jobs) {
  const AoRowsJob &J = jobs[blockIdx.y];
  const int64_t n = 5 * J.npt * 64;
  uint32_t *p = reinterpret_cast<uint32_t *>(J.rowmax);
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) p[i] = 0u;
}

__global__ void __launch_bounds__(kThreads) fill_zero_kernel(int64_t n, uint32_t *__restrict__ p) {
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    p[i] = 0u;
}

// ---------------------------------------------------------------------------
// FP64 ao_vgl in the caller's layout ao_vgl[point][component][AO] (NC = 5),
// or ao_value[point][AO] (NC = 1: a 4 KB ring, five wavefronts per SIMD,
// 0.93 against 1.86 ms for the lane-per-AO kernel on water64), with the
// same lane-per-point evaluation.  A lane's five rows are 5 A doubles apart
// and its own stores would touch one cache line per lane, so the values go
// through a ring in LDS, [component][AO mod kRing64][point], and every
// kRing64-th AO the wavefront writes the ring out with lanes re-assigned:
// lane l moves AO slot l % kRing64 of point l / kRing64 (+ 64 / kRing64 per
// instruction), so each store instruction writes 64/kRing64 rows of
// kRing64 * 8 contiguous bytes, as nontemporal stores (the array is far
// larger than the caches).  The wavefront's 64 points are perm[64 t ..]
// (spatially compact tiles, the order of the mixed path) or 64 consecutive
// points; a stage that the screen zeroes for all 64 points is written as
// zeros without evaluation.  The nonzero flags per (64 rows of the
// [point * 5 + component] row space, 32-AO block) that the FP64
// contraction reads are zeroed first; each lane then marks the (at most
// two) row tiles of its point that hold a nonzero entry of a block.
// The host uses it for bases of 1024 AOs or more whose rows are whole 64-byte
// runs (A % 8 == 0); the lane-per-AO kernel of k_ao.hip, which stores 512-byte
// runs of a row per instruction, writes the others.  On unaligned rows each
// 8-AO run touches two partial lines: on the 48-molecule ragged batch
// (30-620 AOs) this kernel took 22.3 ms of AO time, 15.4 with plain stores,
// the lane-per-AO kernel 10.0.
// Measured on water64 at 128 walkers (81,920 points x 4,160 AOs, 13.6 GB
// written): 4.00 ms, against 5.26 ms for the lane-per-AO kernel.  The same launch without its stores takes 1.93 ms; the
// stores add the rest because the 20 KB ring leaves room for eight
// wavefronts per CU, too few to hide them.  Runs of 4 AOs took 10.4 ms,
// runs of 16 AOs (40 KB ring, four wavefronts per CU) 6.2 ms; holding one
// 8-AO flush in registers to write whole 128-byte lines spilled to scratch;
// plain stores instead of nontemporal ones 4.30 ms; the caller's point order
// instead of compact tiles 4.40 ms; 16-byte stores (two AOs per lane, half
// the store instructions) 1% faster, which would need a second path for
// callers' arrays that are only 8-byte aligned.
#ifndef QMCKL_AO_F64_RING
#define QMCKL_AO_F64_RING 8
#endif
#ifndef QMCKL_AO_F64_GROUP_WAVES
#define QMCKL_AO_F64_GROUP_WAVES 64
#endif
constexpr int kRing64 = QMCKL_AO_F64_RING;  // AOs per ring flush
constexpr int kAoF64GroupWaves = QMCKL_AO_F64_GROUP_WAVES;  // wavefronts per CU a launch aims at
// ring row: 64 points, the point index XOR-swizzled by 8 * (slot mod 8) so
// that the flush reads of one instruction (8 slots of 8 points) fall on
// distinct bank pairs
constexpr int kRingLd = 64;
__device__ __attribute__((always_inline)) inline int ring_col(int slot, int q) { return q ^ (8 * (slot & 7)); }
static_assert(64 % kRing64 == 0 && kRing64 >= 2 && kRing64 <= 16, "ring length");

template <int NC>
struct F64Writer {
  uint32_t hs;       // bit h: this lane's point passes the screen in 32-AO half h of the current stage (NC = 5)
  int64_t t0;        // row tile of this lane's first row (point * 5)
  int cs;            // first component whose row is in tile t0 + 1 (5: none)
  int k;             // next AO index (wavefront-uniform)
  int k_lo, k_hi;    // AOs of the current stage
  double *ring;      // this wavefront's [NC][kRing64][kRingLd] LDS ring
  int p;             // this lane's point (-1 past P)
  int pq[kRing64];   // the points whose slots this lane moves in a flush
  double *out;
  int64_t A;
  uint8_t *flags;    // zeroed [tiles][nkb], or null
  int nkb, lane;

  // the ring's first n slots hold AOs kb .. kb + n - 1: read them out (lane
  // l: slot l % kRing64 of kRing64 points) and store them; the five
  // components half the points at a time (half the registers: 141 instead
  // of 195 with the screen flags below)
  __device__ __attribute__((always_inline)) void flush(int kb, int n) {
    __builtin_amdgcn_wave_barrier();
    const int s = lane % kRing64;
    constexpr int NH = NC == 5 ? 2 : 1;
    constexpr int H = kRing64 / NH;
#pragma unroll
    for (int half = 0; half < NH; ++half) {
      double v[H][NC];
#pragma unroll
      for (int j = 0; j < H; ++j) {
        const int q = (half * H + j) * (64 / kRing64) + lane / kRing64;
#pragma unroll
        for (int c = 0; c < NC; ++c) v[j][c] = ring[(c * kRing64 + s) * kRingLd + ring_col(s, q)];
      }
      if (half == NH - 1) __builtin_amdgcn_wave_barrier();  // every read of the ring is issued
#pragma unroll
      for (int j = 0; j < H; ++j) {
        if (pq[half * H + j] < 0) continue;
        double *row = out + (int64_t)pq[half * H + j] * NC * A + s;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          if (s < n) __builtin_nontemporal_store(v[j][c], row + c * A + kb);
        }
      }
    }
  }
  // the flags of 32-AO block b, from the screen: a point whose shells there
  // pass it marks the (one or two) row tiles of its rows
  __device__ __attribute__((always_inline)) void flag_block(int b) {
    if constexpr (NC != 5) return;  // the flags serve the MO contraction of ao_vgl
    if (flags != nullptr && ((hs >> (b & 1)) & 1u)) {
      flags[t0 * nkb + b] = 1;
      if (cs < 5) flags[(t0 + 1) * nkb + b] = 1;
    }
  }
  __device__ __attribute__((always_inline)) void push(const double (&v)[5]) {
    if (k < k_lo || k >= k_hi) {  // part of a shell outside the current stage
      ++k;
      return;
    }
    const int slot = k % kRing64;
#pragma unroll
    for (int c = 0; c < NC; ++c) ring[(c * kRing64 + slot) * kRingLd + ring_col(slot, lane)] = v[c];
    if (slot == kRing64 - 1) flush(k - (kRing64 - 1), kRing64);
    if ((k & 31) == 31) flag_block(k >> 5);
    ++k;
  }
  // end of a stage whose last AO is k_end - 1: the part-filled ring and
  // 32-block at the end of the AO range
  __device__ __attribute__((always_inline)) void finish(int k_end) {
    if (k_end % kRing64 != 0) flush(k_end - k_end % kRing64, k_end % kRing64);
    if (k_end % 32 != 0) flag_block(k_end >> 5);
  }
  // a stage screened out for every point: zeros (its flags stay zero)
  __device__ __attribute__((always_inline)) void zeros(int k_end) {
    const int n = k_end - k_lo;
    for (int q = 0; q < 64; ++q) {
      const int pu = __builtin_amdgcn_readlane(p, q);
      if (pu < 0) break;  // the points past P are the last ones
      double *row = out + (int64_t)pu * NC * A + k_lo + lane;
#pragma unroll
      for (int c = 0; c < NC; ++c)
        if (lane < n) __builtin_nontemporal_store(0.0, row + c * A);
    }
  }
};

template <int NC>
__global__ void __attribute__((amdgpu_flat_work_group_size(64, 64), amdgpu_waves_per_eu(NC == 1 ? 4 : 2)))
ao_vgl_f64_kernel(int64_t P, int N, int64_t S, const double *__restrict__ coord, const int32_t *__restrict__ perm,
                  const double *__restrict__ nucl, const ShellDesc *__restrict__ shells,
                  const int32_t *__restrict__ stage_shell, const double *__restrict__ ao_factor,
                  const double *__restrict__ expo, const double *__restrict__ coef,
                  const double *__restrict__ nucl_range, int64_t A, int groups, double *__restrict__ out,
                  uint8_t *__restrict__ flags, int nkb) {
  __shared__ double s_ring[NC * kRing64 * kRingLd];
  const int lane = threadIdx.x;
  const int64_t pt = blockIdx.x / groups;
  const int g = (int)(blockIdx.x - pt * groups);
  const int64_t i = pt * 64 + lane;
  const bool inside = i < P;
  const int64_t p = inside ? (perm != nullptr ? (int64_t)perm[i] : i) : 0;
  const double px = inside ? coord[p] : 0.0;
  const double py = inside ? coord[P + p] : 0.0;
  const double pz = inside ? coord[2 * P + p] : 0.0;

  // this wavefront's AOs: 64-AO stages [g*nst/groups, (g+1)*nst/groups)
  const int64_t nst = (A + 63) / 64;
  const int64_t st_lo = g * nst / groups, st_hi = (g + 1) * nst / groups;
  F64Writer<NC> w;
  w.hs = 0;
  w.t0 = (5 * p) >> 6;
  w.cs = inside && 64 - (int)((5 * p) & 63) < 5 ? 64 - (int)((5 * p) & 63) : 5;
  w.ring = s_ring;
  w.p = inside ? (int)p : -1;
#pragma unroll
  for (int j = 0; j < kRing64; ++j) w.pq[j] = __shfl(w.p, j * (64 / kRing64) + lane / kRing64);
  w.out = out;
  w.A = A;
  w.flags = flags;
  w.nkb = nkb;
  w.lane = lane;

  for (int64_t st = st_lo; st < st_hi; ++st) {
    const int k_end = (int)(64 * (st + 1) < A ? 64 * (st + 1) : A);
    const int64_t s_first = stage_shell[st];
    w.k_lo = (int)(64 * st);
    w.k_hi = k_end;
    // per-lane halves for the flags (NC = 5); whole-stage liveness otherwise
    bool live;
    if constexpr (NC == 5) {
      w.hs = lane_screen(s_first, S, w.k_lo, k_end, N, inside, px, py, pz, nucl, shells, nucl_range);
      live = __ballot(w.hs != 0) != 0;
    } else {
      live = stage_live(s_first, S, k_end, N, inside, px, py, pz, nucl, shells, nucl_range);
    }
    if (!live) {
      w.zeros(k_end);
      continue;
    }
    w.k = shells[s_first].ao_start;
    eval_stage(w, s_first, S, k_end, N, inside, px, py, pz, nucl, shells, ao_factor, expo, coef, nucl_range);
    w.finish(k_end);
  }
}

// Screen flags per (64-point tile, 16-AO block): wavefront = (tile, group
// of 64-AO stages), lane = point; per stage the lane's screen of each
// quarter (a shell counts for every quarter its AOs overlap, under the
// test of lane_screen), ORed over the tile's points by a ballot.  The
// adjoint leaves the blocks of G whose flag is zero unwritten and its AO
// kernel (k_adjoint.hip) reads G where its own test passes: the two tests
// are the same expressions on the same inputs (r^2 by the same fmas), so
// they agree bit for bit.
__global__ void __launch_bounds__(kThreads)
tile_screen_kernel(int64_t P, int N, int64_t S, const double *__restrict__ coord, const int32_t *__restrict__ perm,
                   const double *__restrict__ nucl, const ShellDesc *__restrict__ shells,
                   const int32_t *__restrict__ stage_shell, const double *__restrict__ nucl_range, int64_t A,
                   int groups, uint8_t *__restrict__ tflags, int nkb) {
  const int lane = threadIdx.x & 63;
  const int64_t wi = (int64_t)blockIdx.x * (kThreads / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t npt = (P + 63) / 64;
  const int64_t pt = wi / groups;
  const int g = (int)(wi - pt * groups);
  if (pt >= npt) return;  // wavefront-uniform
  const int64_t i = pt * 64 + lane;
  const bool inside = i < P;
  const int64_t p = inside ? (perm != nullptr ? (int64_t)perm[i] : i) : 0;
  const double px = inside ? coord[p] : 0.0;
  const double py = inside ? coord[P + p] : 0.0;
  const double pz = inside ? coord[2 * P + p] : 0.0;
  const int64_t nst = (A + 63) / 64;
  for (int64_t st = g * nst / groups; st < (g + 1) * nst / groups; ++st) {
    const int k_lo = (int)(64 * st);
    const int k_end = (int)(64 * (st + 1) < A ? 64 * (st + 1) : A);
    uint32_t qs = 0;
    for (int64_t si = stage_shell[st]; si < S; ++si) {
      const ShellDesc d = shells[si];
      if (d.ao_start >= k_end) break;
      const double x = px - nucl[d.nucl];
      const double y = py - nucl[N + d.nucl];
      const double z = pz - nucl[2 * N + d.nucl];
      const double r2 = fma(x, x, fma(y, y, z * z));
      const bool live = inside && !(r2 > kCutoff * nucl_range[d.nucl]) && !(d.a_min * r2 > kCutoff);
      if (live) {
        const int a_end = d.ao_start + (d.l + 1) * (d.l + 2) / 2;
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (d.ao_start < k_lo + 16 * (q + 1) && a_end > k_lo + 16 * q) qs |= 1u << q;
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const bool any = __ballot((qs >> q) & 1u) != 0;
      if (lane == 0 && 4 * st + q < nkb) tflags[pt * nkb + 4 * st + q] = any ? 1 : 0;
    }
  }
}

__global__ void __launch_bounds__(kThreads)
point_rows_map_kernel(int64_t P, const int32_t *__restrict__ perm, int32_t *__restrict__ kmap) {
  for (int64_t k = blockIdx.x * (int64_t)kThreads + threadIdx.x; k < 5 * P; k += (int64_t)gridDim.x * kThreads) {
    const int64_t i = k / 5;
    kmap[k] = (int32_t)(5 * (int64_t)perm[i] + (k - 5 * i));
  }
}

__global__ void __launch_bounds__(kThreads) fill_zero_bytes_kernel(int64_t n, uint8_t *__restrict__ p) {
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    p[i] = 0;
}

}  // namespace

hipError_t launch_ao_vgl_f64(hipStream_t s, const DeviceBasis &b, int64_t P, const double *coord,
                             const int32_t *perm, const double *nucl, const double *ao_factor, double *out,
                             int ncomp, uint8_t *flags, int nkb) {
  if (P <= 0) return hipSuccess;
  const int64_t A = b.ao_num;
  if (b.max_l > kMaxL || A <= 0 || b.stage_shell == nullptr || (ncomp != 5 && ncomp != 1))
    return hipErrorInvalidValue;
  if (flags != nullptr && (ncomp != 5 || (int64_t)nkb != (A + 31) / 32)) return hipErrorInvalidValue;
  const int64_t npt = (P + 63) / 64;
  const int groups = ao_rows_groups(P, (A + 63) / 64 * 64, kAoF64GroupWaves);
  const int64_t blocks = npt * groups;
  if (blocks > 0x7fffffff || P > 0x7fffffff || A > 0x7fffffff) return hipErrorInvalidValue;
  if (flags != nullptr) {
    // a kernel, not a memset: the launch may be captured into a graph
    const int64_t n = (5 * P + 63) / 64 * nkb;
    hipLaunchKernelGGL(fill_zero_bytes_kernel, dim3((unsigned)std::min<int64_t>((n + kThreads - 1) / kThreads, 4096)),
                       dim3(kThreads), 0, s, n, flags);
  }
  if (ncomp == 5)
    hipLaunchKernelGGL(ao_vgl_f64_kernel<5>, dim3((unsigned)blocks), dim3(64), 0, s, P, (int)b.nucl_num,
                       b.shell_num, coord, perm, nucl, b.shells, b.stage_shell, ao_factor, b.expo, b.coef_norm,
                       b.nucl_range, A, groups, out, flags, nkb);
  else
    hipLaunchKernelGGL(ao_vgl_f64_kernel<1>, dim3((unsigned)blocks), dim3(64), 0, s, P, (int)b.nucl_num,
                       b.shell_num, coord, perm, nucl, b.shells, b.stage_shell, ao_factor, b.expo, b.coef_norm,
                       b.nucl_range, A, groups, out, nullptr, 0);
  return hipGetLastError();
}

int ao_rows_groups(int64_t P, int64_t ld, int group_waves) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  const int64_t npt = (P + 63) / 64, nst = ld / 64;
  // default 256 wavefronts per CU (measured on water64 at 128 walkers with
  // ordered tiles: 1.29 ms at 16, 1.25 at 32, 1.15 at 64, 1.09 at 128, 1.06
  // at 256; the contraction after it is unaffected)
  const int64_t want = (group_waves > 0 ? group_waves : 256) * (int64_t)cus;
  int64_t g = (want + npt - 1) / npt;
  if (g > nst) g = nst;
  if (g < 1) g = 1;
  return (int)g;
}

hipError_t launch_ao_rows_f32(hipStream_t s, const DeviceBasis &b, int64_t P, const double *coord,
                              const int32_t *perm, const double *nucl, const double *ao_factor, int64_t ld,
                              int groups, float *buf, float *rowmax, uint8_t *flags, int nkb, float *buf_lo,
                              bool skip_halves) {
  if (P <= 0) return hipSuccess;
  if (b.max_l > kMaxL || ld % 64 != 0 || ld < b.ao_num || groups < 1 || groups > ld / 64 ||
      b.stage_shell == nullptr)
    return hipErrorInvalidValue;
  if (flags != nullptr && (int64_t)nkb * 32 != ld) return hipErrorInvalidValue;
  const int64_t npt = (P + 63) / 64;
  const int64_t waves = npt * groups;
  const int64_t blocks = (waves + kThreads / 64 - 1) / (kThreads / 64);
  if (blocks > 0x7fffffff || ld > 0x7fffffff) return hipErrorInvalidValue;
  const int64_t nmax = 5 * npt * 64;
  hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)std::min<int64_t>((nmax + kThreads - 1) / kThreads, 4096)),
                     dim3(kThreads), 0, s, nmax, reinterpret_cast<uint32_t *>(rowmax));
  if (buf_lo != nullptr)
    hipLaunchKernelGGL(ao_rows_kernel<true>, dim3((unsigned)blocks), dim3(kThreads), 0, s, P, (int)b.nucl_num,
                       b.shell_num, coord, perm, nucl, b.shells, b.stage_shell, ao_factor, b.expo, b.coef_norm,
                       b.nucl_range, b.ao_num, ld, npt, groups, buf, buf_lo, rowmax, flags, nkb,
                       skip_halves ? 1 : 0);
  else
    hipLaunchKernelGGL(ao_rows_kernel<false>, dim3((unsigned)blocks), dim3(kThreads), 0, s, P, (int)b.nucl_num,
                       b.shell_num, coord, perm, nucl, b.shells, b.stage_shell, ao_factor, b.expo, b.coef_norm,
                       b.nucl_range, b.ao_num, ld, npt, groups, buf, nullptr, rowmax, flags, nkb,
                       skip_halves ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_tile_screen(hipStream_t s, const DeviceBasis &b, int64_t P, const double *coord,
                              const int32_t *perm, const double *nucl, uint8_t *tflags, int nkb) {
  if (P <= 0) return hipSuccess;
  const int64_t A = b.ao_num;
  if (A <= 0 || b.stage_shell == nullptr || (int64_t)nkb != (A + 15) / 16 || P > 0x7fffffff)
    return hipErrorInvalidValue;
  const int64_t npt = (P + 63) / 64, nst = (A + 63) / 64;
  int64_t groups = (64 * 256 + npt - 1) / npt;  // about 16,384 wavefronts
  if (groups > nst) groups = nst;
  if (groups < 1) groups = 1;
  const int64_t blocks = (npt * groups + kThreads / 64 - 1) / (kThreads / 64);
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(tile_screen_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, P, (int)b.nucl_num,
                     b.shell_num, coord, perm, nucl, b.shells, b.stage_shell, b.nucl_range, A, (int)groups, tflags,
                     nkb);
  return hipGetLastError();
}

hipError_t launch_point_rows_map(hipStream_t s, int64_t P, const int32_t *perm, int32_t *kmap) {
  if (P <= 0) return hipSuccess;
  if (5 * P > 0x7fffffff) return hipErrorInvalidValue;
  int64_t g = (5 * P + kThreads - 1) / kThreads;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(point_rows_map_kernel, dim3((unsigned)g), dim3(kThreads), 0, s, P, perm, kmap);
  return hipGetLastError();
}

hipError_t make_ao_rows_job(const DeviceBasis &b, int64_t P, const double *coord, const int32_t *perm,
                            const double *nucl, const double *ao_factor, int64_t ld, int groups, float *buf,
                            float *rowmax, uint8_t *flags, int nkb, bool skip_halves, AoRowsJob *job) {
  if (P <= 0) return hipErrorInvalidValue;
  if (b.max_l > kMaxL || ld % 64 != 0 || ld < b.ao_num || groups < 1 || groups > ld / 64 ||
      b.stage_shell == nullptr || ld > 0x7fffffff)
    return hipErrorInvalidValue;
  if (flags != nullptr && (int64_t)nkb * 32 != ld) return hipErrorInvalidValue;
  job->P = P;
  job->N = (int)b.nucl_num;
  job->S = b.shell_num;
  job->coord = coord;
  job->perm = perm;
  job->nucl = nucl;
  job->shells = b.shells;
  job->stage_shell = b.stage_shell;
  job->ao_factor = ao_factor;
  job->expo = b.expo;
  job->coef = b.coef_norm;
  job->nucl_range = b.nucl_range;
  job->A = b.ao_num;
  job->ld = ld;
  job->npt = (P + 63) / 64;
  job->groups = groups;
  job->buf = buf;
  job->rowmax = rowmax;
  job->flags = flags;
  job->nkb = flags != nullptr ? nkb : 0;
  job->skip_halves = skip_halves ? 1 : 0;
  job->wave0 = 0;
  job->waves = job->npt * groups;
  return hipSuccess;
}

hipError_t launch_ao_rows_batch(hipStream_t s, const AoRowsJob *jobs, int njobs, int64_t total) {
  if (njobs <= 0 || total <= 0) return hipSuccess;
  if (njobs > 65535) return hipErrorInvalidValue;
  const int64_t blocks = (total + kThreads / 64 - 1) / (kThreads / 64);
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(fill_rowmax_batch_kernel, dim3(64, (unsigned)njobs), dim3(kThreads), 0, s, jobs);
  hipLaunchKernelGGL(ao_rows_batch_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, jobs, njobs, total);
  return hipGetLastError();
}

}  // namespace qmckl_amd
