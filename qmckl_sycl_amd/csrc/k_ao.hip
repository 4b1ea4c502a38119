// Shell/AO VGL kernels (FP64 arithmetic; FP64 or FP32 stores).
//
// Arithmetic of src/qmckl_ao_sycl.cpp: radial shell sums with the nucleus
// screen r^2 > cutoff*range and the primitive screen alpha r^2 > cutoff
// (40-95), Cartesian monomials x^a y^b z^c in order of descending a then b
// (291-331), and the product-rule assembly times ao_factor (367-422).
// Screened entries are exact zeros.
//
// Work decomposition (one workgroup = 64 points x all AOs, 4 wavefronts).
// The shells, in AO order, are cut at finalize into chunks of at most 16
// shells and 64 AOs.  For each chunk:
//   - phase 1: each wavefront takes whole shells, lane = point.  The shell
//     index is wavefront-uniform, so the shell descriptor, the nucleus and
//     the (exponent, coefficient) pairs come through scalar loads, and a
//     primitive screened for all 64 points costs no exp (the branch is
//     skipped when no lane's predicate holds).  R, dR/dx, dR/dy, dR/dz and
//     lap R land in LDS [shell][component][point];
//   - phase 2: lane = AO of the chunk (its shell, exponents (a,b,c), factor
//     and nucleus are loaded once), wavefront w walks points w, w+4, ...;
//     every store of a wavefront is one contiguous run of a [p][c][*] row,
//     through a pointer advanced by a constant stride.  Each lane ORs its
//     AO's nonzero entries into a bit per 64-row tile; one wavefront ballot
//     per tile and chunk turns them into the chunk's nonzero flags, written
//     to LDS once per chunk and wavefront.
// Two barriers per chunk: phase 1 -> phase 2, and phase 2 -> the next
// chunk's phase 1 (which overwrites the radial tile).
// r^2 is fma(x,x,fma(y,y,z*z)), as in the CPU oracle, so the screens decide
// identically.  The kernel is bound by the HBM write of ao_vgl:
// 40 B (FP64) or 20 B (FP32) per (point, AO) plus 24 B per point.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.hpp"

namespace qmckl_amd {

namespace {

constexpr int kTP = 64;            // points per workgroup
constexpr int kThreads = 256;      // 4 wavefronts
constexpr int kRS = 5 * kTP + 2;   // LDS stride between shells (bank spread)
#ifndef QMCKL_AO_TARGET_GROUPS
#define QMCKL_AO_TARGET_GROUPS 4096
#endif
constexpr int64_t kAoTargetGroups = QMCKL_AO_TARGET_GROUPS;  // workgroups a launch aims at

struct Mono {
  double v, gx, gy, gz, lap;
};

// (x^a, a x^(a-1), a(a-1) x^(a-2)) by the recurrence
// (f, f', f'') -> (x f, f + x f', 2 f' + x f'').
__device__ inline void pow3(double x, int a, double &f0, double &f1, double &f2) {
  double p0 = 1.0, p1 = 0.0, p2 = 0.0;
  for (int i = 0; i < a; ++i) {
    const double n2 = 2.0 * p1 + x * p2;
    const double n1 = p0 + x * p1;
    p0 = x * p0;
    p1 = n1;
    p2 = n2;
  }
  f0 = p0;
  f1 = p1;
  f2 = p2;
}

__device__ inline Mono monomial(int a, int b, int c, double x, double y, double z) {
  double X0, X1, X2, Y0, Y1, Y2, Z0, Z1, Z2;
  pow3(x, a, X0, X1, X2);
  pow3(y, b, Y0, Y1, Y2);
  pow3(z, c, Z0, Z1, Z2);
  Mono m;
  const double yz = Y0 * Z0;
  m.v = X0 * yz;
  m.gx = X1 * yz;
  m.gy = X0 * Y1 * Z0;
  m.gz = X0 * Y0 * Z1;
  m.lap = X2 * yz + X0 * Y2 * Z0 + X0 * Y0 * Z2;
  return m;
}

// NCOMP = 5 (VGL) or 1 (value).  FROM_TABLE: radial parts are read from a
// precomputed shell_vgl [P][5][S] instead of being summed from primitives.
// OutT = double (the plain form) or float (the AO operand of the mixed MO
// contraction: the FP64 value rounded once at the store).
template <int NCOMP, bool FROM_TABLE, typename OutT>
__global__ void __launch_bounds__(kThreads)
ao_kernel(int64_t P, int N, const double *__restrict__ coord, const double *__restrict__ nucl,
          const ShellDesc *__restrict__ shells, const int32_t *__restrict__ chunks, int nchunks,
          const int32_t *__restrict__ ao_shell, const int32_t *__restrict__ ao_abc,
          const double *__restrict__ ao_factor, const double *__restrict__ expo,
          const double *__restrict__ coef, const double *__restrict__ nucl_range,
          const double *__restrict__ shell_table, int S_table, AoOut<OutT> ao, int cgroups) {
  extern __shared__ unsigned char s_flag[];  // [NCOMP row tiles][nkb] nonzero flags
  __shared__ double rad[kChunkShells * kRS];
  __shared__ double pt[3][kTP];
  // this workgroup's chunks: group blockIdx.y of cgroups
  const int ch_begin = (int)((int64_t)blockIdx.y * nchunks / cgroups);
  const int ch_end = (int)((int64_t)(blockIdx.y + 1) * nchunks / cgroups);

  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t p0 = (int64_t)blockIdx.x * kTP;
  const int np = (int)((P - p0) < kTP ? (P - p0) : kTP);
  const bool track_flags = ao.flags != nullptr;

  if (tid < kTP) {
    const int64_t p = p0 + tid;
    const bool ok = tid < np;
    pt[0][tid] = ok ? coord[p] : 0.0;
    pt[1][tid] = ok ? coord[P + p] : 0.0;
    pt[2][tid] = ok ? coord[2 * P + p] : 0.0;
  }
  if (track_flags)
    for (int i = tid; i < NCOMP * ao.nkb; i += kThreads) s_flag[i] = 0;
  __syncthreads();

  const double px = pt[0][lane], py = pt[1][lane], pz = pt[2][lane];
  const int64_t ostep = (int64_t)(kThreads / 64) * NCOMP * ao.ld;  // stride between a wavefront's points

  for (int ch = ch_begin; ch < ch_end; ++ch) {
    const int s0 = chunks[ch];
    const int ns = chunks[ch + 1] - s0;

    // phase 1: radial parts, wavefront per shell, lane per point
    for (int sl = wave; sl < ns; sl += kThreads / 64) {
      const ShellDesc d = shells[s0 + sl];
      double t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      if (FROM_TABLE) {
        if (lane < np) {
          const double *row = shell_table + (p0 + lane) * 5 * (int64_t)S_table + d.shell;
#pragma unroll
          for (int c = 0; c < NCOMP; ++c) t[c] = row[(int64_t)c * S_table];
        }
      } else {
        const double x = px - nucl[d.nucl];
        const double y = py - nucl[N + d.nucl];
        const double z = pz - nucl[2 * N + d.nucl];
        const double r2 = fma(x, x, fma(y, y, z * z));
        const bool live = lane < np && !(r2 > kCutoff * nucl_range[d.nucl]);
        const double *ex = expo + d.prim_start;
        const double *cf = coef + d.prim_start;
        for (int j = 0; j < d.prim_num; ++j) {
          const double a = ex[j];
          const double ar2 = a * r2;
          if (live && !(ar2 > kCutoff)) {
            const double v = cf[j] * exp(-ar2);
            t[0] += v;
            if (NCOMP == 5) {
              const double two_a = -2.0 * a * v;
              t[1] += two_a * x;
              t[2] += two_a * y;
              t[3] += two_a * z;
              t[4] += two_a * (3.0 - 2.0 * ar2);
            }
          }
        }
      }
#pragma unroll
      for (int c = 0; c < NCOMP; ++c) rad[sl * kRS + c * kTP + lane] = t[c];
    }
    __syncthreads();

    // phase 2: lane = AO of the chunk, wavefront walks its points
    const ShellDesc first = shells[s0], last = shells[s0 + ns - 1];
    const int ka = first.ao_start;
    const int nao = last.ao_start + (last.l + 1) * (last.l + 2) / 2 - ka;
    const bool act = lane < nao;
    const int k = ka + (act ? lane : 0);
    const int sl = ao_shell[k] - s0;
    const int abc = ao_abc[k];
    const int ea = abc & 0xff, eb = (abc >> 8) & 0xff, ec = (abc >> 16) & 0xff;
    const double f = ao_factor[k];
    const int n = shells[s0 + sl].nucl;
    const double nx = nucl[n], ny = nucl[N + n], nz = nucl[2 * N + n];
    const double *rs = rad + sl * kRS;
    // lanes of the chunk's (up to three) 32-AO flag blocks
    const int kb0 = ka >> 5;
    const int cut0 = 32 * (kb0 + 1) - ka, cut1 = cut0 + 32;
    const uint64_t act_mask = nao >= 64 ? ~0ull : ((1ull << nao) - 1ull);
    const uint64_t bm0 = act_mask & ((1ull << cut0) - 1ull);
    const uint64_t bm1 = act_mask & ~((1ull << cut0) - 1ull) & (cut1 >= 64 ? ~0ull : ((1ull << cut1) - 1ull));
    const uint64_t bm2 = act_mask & (cut1 >= 64 ? 0ull : ~((1ull << cut1) - 1ull));
    uint32_t nzt = 0;  // this lane: bit tile = some entry of its AO in 64-row tile `tile` is nonzero

    OutT *o = ao.out + (p0 + wave) * NCOMP * ao.ld + k;
    for (int pl = wave; pl < np; pl += kThreads / 64, o += ostep) {
      const Mono m = monomial(ea, eb, ec, pt[0][pl] - nx, pt[1][pl] - ny, pt[2][pl] - nz);
      const double R = rs[pl];
      OutT w[NCOMP];
      w[0] = (OutT)(m.v * R * f);
      if (NCOMP == 5) {
        const double Rx = rs[1 * kTP + pl], Ry = rs[2 * kTP + pl];
        const double Rz = rs[3 * kTP + pl], RL = rs[4 * kTP + pl];
        w[1 % NCOMP] = (OutT)((m.gx * R + m.v * Rx) * f);
        w[2 % NCOMP] = (OutT)((m.gy * R + m.v * Ry) * f);
        w[3 % NCOMP] = (OutT)((m.gz * R + m.v * Rz) * f);
        w[4 % NCOMP] = (OutT)((m.lap * R + m.v * RL + 2.0 * (m.gx * Rx + m.gy * Ry + m.gz * Rz)) * f);
      }
      if (act) {
#pragma unroll
        for (int c = 0; c < NCOMP; ++c) o[c * ao.ld] = w[c];
      }
      if (track_flags) {
#pragma unroll
        for (int c = 0; c < NCOMP; ++c) nzt |= (w[c] != (OutT)0 ? 1u : 0u) << ((pl * NCOMP + c) >> 6);
      }
    }
    if (track_flags) {
      // one ballot per row tile per chunk: which of the chunk's (up to three)
      // 32-AO blocks hold a nonzero entry of that tile
      uint32_t fmask = 0;  // bit 3*tile + block
#pragma unroll
      for (int tile = 0; tile < NCOMP; ++tile) {
        const uint64_t nzm = __ballot(act && ((nzt >> tile) & 1u));
        fmask |= ((nzm & bm0) ? 1u : 0u) << (3 * tile);
        fmask |= ((nzm & bm1) ? 2u : 0u) << (3 * tile);
        fmask |= ((nzm & bm2) ? 4u : 0u) << (3 * tile);
      }
      if (lane == 0 && fmask != 0)
        for (int tile = 0; tile < NCOMP; ++tile)
          for (int b = 0; b < 3; ++b)
            if (((fmask >> (3 * tile + b)) & 1u) && kb0 + b < ao.nkb) s_flag[tile * ao.nkb + kb0 + b] = 1;
    }
    __syncthreads();  // phase 2 is done with the radial tile
  }
  if (track_flags) {
    // this workgroup's rows [NCOMP*p0, NCOMP*(p0+np)) are whole 64-row
    // tiles.  One group: it owns their flags and writes them all; several
    // groups (chunk ranges) share a tile's flags, which the launcher has
    // zeroed, and each sets only the ones it found (a 32-AO block can
    // straddle two groups' chunks).
    const int64_t rt0 = (p0 * NCOMP) >> 6;
    const int64_t rt_end = (NCOMP * (p0 + np) + 63) >> 6;
    for (int i = tid; i < NCOMP * ao.nkb; i += kThreads) {
      const int64_t rt = rt0 + i / ao.nkb;
      if (rt < rt_end && (cgroups == 1 || s_flag[i] != 0)) ao.flags[rt * ao.nkb + (i % ao.nkb)] = s_flag[i];
    }
  }
}

// shell_vgl [P][5][S] in the user's shell numbering; thread = (point, shell).
__global__ void __launch_bounds__(kThreads)
shell_vgl_kernel(int64_t P, int S, int N, const double *__restrict__ coord,
                 const double *__restrict__ nucl, const ShellDesc *__restrict__ shells,
                 const double *__restrict__ expo, const double *__restrict__ coef,
                 const double *__restrict__ nucl_range, double *__restrict__ out) {
  const int64_t total = P * (int64_t)S;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = i / S;
    const ShellDesc d = shells[i - p * S];
    const double x = coord[p] - nucl[d.nucl];
    const double y = coord[P + p] - nucl[N + d.nucl];
    const double z = coord[2 * P + p] - nucl[2 * N + d.nucl];
    const double r2 = fma(x, x, fma(y, y, z * z));
    double t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (!(r2 > kCutoff * nucl_range[d.nucl])) {
      for (int j = d.prim_start; j < d.prim_start + d.prim_num; ++j) {
        const double ar2 = expo[j] * r2;
        if (ar2 > kCutoff) continue;
        const double v = coef[j] * exp(-ar2);
        const double two_a = -2.0 * expo[j] * v;
        t[0] += v;
        t[1] += two_a * x;
        t[2] += two_a * y;
        t[3] += two_a * z;
        t[4] += two_a * (3.0 - 2.0 * ar2);
      }
    }
    double *o = out + p * 5 * S + d.shell;
#pragma unroll
    for (int c = 0; c < 5; ++c) o[(int64_t)c * S] = t[c];
  }
}

__global__ void __launch_bounds__(kThreads) zero_bytes_kernel(int64_t n, uint8_t *__restrict__ p) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

}  // namespace

template <int NCOMP, bool FROM_TABLE, typename OutT>
hipError_t launch_ao(hipStream_t s, const DeviceBasis &b, int64_t P, const double *coord,
                     const double *nucl, const double *ao_factor, const double *shell_table,
                     int64_t S_table, AoOut<OutT> ao) {
  if (P <= 0) return hipSuccess;
  const int64_t blocks = (P + kTP - 1) / kTP;
  if (blocks > 0x7fffffff || b.chunks == nullptr || b.nchunks <= 0) return hipErrorInvalidValue;
  if (ao.flags != nullptr && (ao.nkb <= 0 || ao.nkb > kMaxFlagBlocks)) return hipErrorInvalidValue;
  const size_t lds = ao.flags != nullptr ? (size_t)NCOMP * ao.nkb : 0;
  // Split each 64-point tile's chunks over several workgroups when there are
  // few tiles, so the launch has enough workgroups to keep every CU busy to
  // the end (three fit per CU; a tile's chunks are a long serial loop).
  int64_t groups = (kAoTargetGroups + blocks - 1) / blocks;
  if (groups > b.nchunks) groups = b.nchunks;
  if (groups > 65535) groups = 65535;
  if (groups < 1) groups = 1;
  if (groups > 1 && ao.flags != nullptr) {
    // a kernel, not a memset: the launch may be captured into a graph
    const int64_t n = (NCOMP * P + 63) / 64 * ao.nkb;
    hipLaunchKernelGGL(zero_bytes_kernel, dim3((unsigned)std::min<int64_t>((n + kThreads - 1) / kThreads, 4096)),
                       dim3(kThreads), 0, s, n, ao.flags);
  }
  hipLaunchKernelGGL((ao_kernel<NCOMP, FROM_TABLE, OutT>), dim3((unsigned)blocks, (unsigned)groups),
                     dim3(kThreads), lds, s, P, (int)b.nucl_num, coord, nucl, b.shells, b.chunks, b.nchunks,
                     b.ao_shell, b.ao_abc, ao_factor, b.expo, b.coef_norm, b.nucl_range, shell_table,
                     (int)S_table, ao, (int)groups);
  return hipGetLastError();
}

hipError_t launch_ao_vgl(hipStream_t s, const DeviceBasis &b, int64_t P, const double *coord,
                         const double *nucl, const double *ao_factor, double *out, int ncomp,
                         uint8_t *flags, int nkb, const int32_t *perm) {
  if (ao_vgl_lane_kernel(b.ao_num))
    return launch_ao_vgl_f64(s, b, P, coord, perm, nucl, ao_factor, out, ncomp, ncomp == 5 ? flags : nullptr,
                             ncomp == 5 ? nkb : 0);
  AoOut<double> ao{out, b.ao_num, flags, nkb};
  if (ncomp == 5) return launch_ao<5, false, double>(s, b, P, coord, nucl, ao_factor, nullptr, 0, ao);
  ao.flags = nullptr;
  return launch_ao<1, false, double>(s, b, P, coord, nucl, ao_factor, nullptr, 0, ao);
}

hipError_t launch_ao_from_shell_table(hipStream_t s, const DeviceBasis &b, int64_t P,
                                      const double *coord, const double *nucl,
                                      const double *ao_factor, const double *shell_vgl,
                                      int64_t S_table, double *out, int ncomp) {
  AoOut<double> ao{out, b.ao_num, nullptr, 0};
  if (ncomp == 5)
    return launch_ao<5, true, double>(s, b, P, coord, nucl, ao_factor, shell_vgl, S_table, ao);
  return launch_ao<1, true, double>(s, b, P, coord, nucl, ao_factor, shell_vgl, S_table, ao);
}

hipError_t launch_shell_vgl(hipStream_t s, const DeviceBasis &b, int64_t P, const double *coord,
                            const double *nucl, double *shell_vgl) {
  if (P <= 0) return hipSuccess;
  const int64_t work = P * b.shell_num;
  int64_t blocks = (work + kThreads - 1) / kThreads;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(shell_vgl_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, P,
                     (int)b.shell_num, (int)b.nucl_num, coord, nucl, b.shells, b.expo, b.coef_norm,
                     b.nucl_range, shell_vgl);
  return hipGetLastError();
}

}  // namespace qmckl_amd
