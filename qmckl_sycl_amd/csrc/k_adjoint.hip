// Adjoint of L = sum_{p,c,i} w[p][c][i] mo_vgl[p][c][i] with respect to the
// nuclear coordinates and the primitive exponents (FP64).
//
// With G[p][c][k] = sum_i w[p][c][i] C[i][k] (one GEMM, done by the caller),
// L depends on a nucleus and an exponent only through the AOs:
//   dL/dR_n     = - sum_p sum_{k on n} sum_c G[p][c][k] d/dr AO_c[p][k]
//   dL/dalpha_j =   sum_p sum_{k in shell(j)} sum_c G[p][c][k] dAO_c/dalpha_j
// (AO_c depends on r - R_n; the exponent derivative holds the normalised
// contraction coefficient fixed, as the basis arrays are independent inputs).
// AO_k = f P(d) g(s), d = r - R_n, s = |d|^2, g = sum_j c_j e^{-a_j s}:
//   d_i g = 2 d_i g', d_ij g = 4 d_i d_j g'' + 2 delta_ij g',
//   lap g = 4 s g'' + 6 g', d_j lap g = 2 d_j (10 g'' + 4 s g'''),
// and the exponent derivative replaces g by h_j = -s c_j e^{-a_j s}
// (h' = -c e^{-as}(1 - a s), h'' = c a e^{-as}(2 - a s)), which enters
// linearly through three per-shell sums U0, U1, U2.
//
// The primitives are recomputed here, not stored by the forward pass.
// Work: workgroup = 64 points and a contiguous range of the shells (the
// shells of a point block are split over several workgroups so the launch
// has enough wavefronts), wavefront = one shell at a time, lane =
// point; per (point, shell) the lane evaluates g..g''', loops over the
// shell's AOs (monomials to third order), and reduces over the wavefront in
// a fixed order.  The 64 points of a block are those of a spatially compact
// tile when the caller orders them (perm), so that most shells are
// screened for every point of the block and skipped whole: on 64 waters at
// 32 walkers 26% of the (block, shell) pairs stay live in tile order, 47% in
// walker order, against the 87% the nucleus screen alone leaves.  Each
// (workgroup, shell) writes one partial dR and
// each (workgroup, primitive) one partial dalpha; a second kernel sums the
// partials in a fixed order, so the result is bitwise reproducible.
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace qmckl_amd {

namespace {

constexpr int kTP = 64;
constexpr int kThreads = 256;

// x^a and its first three derivatives: (f, f', f'', f''') ->
// (x f, f + x f', 2 f' + x f'', 3 f'' + x f''').
__device__ inline void pow4(double x, int a, double *p) {
  double f0 = 1.0, f1 = 0.0, f2 = 0.0, f3 = 0.0;
  for (int i = 0; i < a; ++i) {
    const double n3 = 3.0 * f2 + x * f3;
    const double n2 = 2.0 * f1 + x * f2;
    const double n1 = f0 + x * f1;
    f0 = x * f0;
    f1 = n1;
    f2 = n2;
    f3 = n3;
  }
  p[0] = f0;
  p[1] = f1;
  p[2] = f2;
  p[3] = f3;
}

// Sum over the wavefront, the same value in every lane, without LDS:
// neighbours by DPP quad permutes (xor 1, xor 2), the mirrors of 8 and 16
// lanes, then v_permlane16_swap / v_permlane32_swap for the 16-lane rows
// and the 32-lane halves.  At every step the two partners add the same two
// values, so all lanes hold the same bits; the order is fixed.  (The
// shuffles through ds_bpermute it replaces cost six dependent LDS round
// trips per sum, and the kernel takes three sums per shell and one per
// primitive.)
template <int CTRL>
__device__ __attribute__((always_inline)) inline double dpp_mov(double x) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = __builtin_amdgcn_update_dpp(0u, (uint32_t)u, CTRL, 0xf, 0xf, false);
  const uint32_t hi = __builtin_amdgcn_update_dpp(0u, (uint32_t)(u >> 32), CTRL, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
// x of this lane plus x of the lane in the partner row (16-lane rows 0<->1,
// 2<->3) or half (lanes 0-31 <-> 32-63)
template <bool HALVES>
__device__ __attribute__((always_inline)) inline double swap_add(double x) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint32_t l = (uint32_t)u, h = (uint32_t)(u >> 32);
  const auto lo = HALVES ? __builtin_amdgcn_permlane32_swap(l, l, false, false)
                         : __builtin_amdgcn_permlane16_swap(l, l, false, false);
  const auto hi = HALVES ? __builtin_amdgcn_permlane32_swap(h, h, false, false)
                         : __builtin_amdgcn_permlane16_swap(h, h, false, false);
  // [0]: the lower row (half) everywhere, [1]: the upper one
  return __builtin_bit_cast(double, ((uint64_t)hi[0] << 32) | lo[0]) +
         __builtin_bit_cast(double, ((uint64_t)hi[1] << 32) | lo[1]);
}
__device__ __attribute__((always_inline)) inline double wave_sum(double v) {
  v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_mov<0x141>(v);  // row_half_mirror
  v += dpp_mov<0x140>(v);  // row_mirror
  v = swap_add<false>(v);
  return swap_add<true>(v);
}

__global__ void __launch_bounds__(kThreads)
ao_adjoint_kernel(int64_t P, int S, int A, int N, const double *__restrict__ coord,
                  const int32_t *__restrict__ perm, const double *__restrict__ nucl,
                  const ShellDesc *__restrict__ shells,
                  const int32_t *__restrict__ ao_abc, const double *__restrict__ ao_factor,
                  const double *__restrict__ expo, const double *__restrict__ coef,
                  const double *__restrict__ nucl_range, const double *__restrict__ G,
                  double *__restrict__ partR, double *__restrict__ partA, int64_t Np, int sgroups) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t blk = blockIdx.x;
  const int64_t i = blk * kTP + lane;  // this lane's G rows are 5 i + c
  const bool inb = i < P;
  const int64_t p = inb && perm != nullptr ? (int64_t)perm[i] : i;
  const double rx = inb ? coord[p] : 0.0;
  const double ry = inb ? coord[P + p] : 0.0;
  const double rz = inb ? coord[2 * P + p] : 0.0;
  double *myR = partR + blk * (int64_t)S * 3;
  double *myA = partA + blk * Np;

  // shells of this workgroup: a contiguous range (the sgroups ranges split
  // the shells evenly), its four wavefronts taking every fourth, so that
  // the G rows' AO runs of neighbouring shells are read by one workgroup
  // (interleaving the shells over workgroups made each 128-byte line of G
  // come from HBM several times: 12.7 GB fetched for 3.4 GB of G on water64
  // at 32 walkers)
  const int s_lo = (int)((int64_t)blockIdx.y * S / sgroups), s_hi = (int)((int64_t)(blockIdx.y + 1) * S / sgroups);
  for (int si = s_lo + wave; si < s_hi; si += kThreads / 64) {
    const ShellDesc d = shells[si];
    const double dx = rx - nucl[d.nucl];
    const double dy = ry - nucl[N + d.nucl];
    const double dz = rz - nucl[2 * N + d.nucl];
    const double s = fma(dx, dx, fma(dy, dy, dz * dz));
    // A lane is live when the nucleus screen passes and some primitive is
    // not screened: a_min s > cutoff means a s > cutoff for every exponent
    // of the shell (a >= a_min, and rounding is monotonic), so every term of
    // a lane that is not live is zero.  The test (and s) are those of the
    // tile screen (k_ao_rows.hip), whose zero blocks of G are not written.
    const bool live = inb && !(s > kCutoff * nucl_range[d.nucl]) && !(d.a_min * s > kCutoff);
    if (!__any(live)) {
      // what the sums below give for a shell that no lane reaches: -(+0)
      // for the coordinates, +0 for the exponents
      if (lane == 0) {
        myR[(int64_t)d.shell * 3 + 0] = -0.0;
        myR[(int64_t)d.shell * 3 + 1] = -0.0;
        myR[(int64_t)d.shell * 3 + 2] = -0.0;
        for (int j = 0; j < d.prim_num; ++j) myA[d.prim_start + j] = 0.0;
      }
      continue;
    }
    // G of the shell's AOs is read one AO ahead of its use, the first AO's
    // before the primitive sums, so that the loads (64 rows 5 A doubles
    // apart per instruction) are in flight while the lane computes
    const double *Gp = G + (i * 5) * (int64_t)A;
    double wn[5];
    auto load_g = [&](int k) {
#pragma unroll
      for (int c = 0; c < 5; ++c) wn[c] = live ? Gp[(int64_t)c * A + k] : 0.0;
    };
    load_g(d.ao_start);
    double g0 = 0.0, g1 = 0.0, g2 = 0.0, g3 = 0.0;
    for (int j = 0; j < d.prim_num; ++j) {
      const double a = expo[d.prim_start + j];
      const double as = a * s;
      if (live && !(as > kCutoff)) {
        const double v = coef[d.prim_start + j] * exp(-as);
        g0 += v;
        g1 -= a * v;
        g2 += a * a * v;
        g3 -= a * a * a * v;
      }
    }
    double gR0 = 0.0, gR1 = 0.0, gR2 = 0.0, U0 = 0.0, U1 = 0.0, U2 = 0.0;
    const int ncart = (d.l + 1) * (d.l + 2) / 2;
    {
      const double dd[3] = {dx, dy, dz};
      const double gi[3] = {2.0 * dx * g1, 2.0 * dy * g1, 2.0 * dz * g1};
      const double lapg = 4.0 * s * g2 + 6.0 * g1;
      const double t3 = 2.0 * (10.0 * g2 + 4.0 * s * g3);
      for (int kk = 0; kk < ncart; ++kk) {
        const int k = d.ao_start + kk;
        const int abc = ao_abc[k];
        double X[4], Y[4], Z[4];
        pow4(dx, abc & 0xff, X);
        pow4(dy, (abc >> 8) & 0xff, Y);
        pow4(dz, (abc >> 16) & 0xff, Z);
        const double m0 = X[0] * Y[0] * Z[0];
        const double Pi[3] = {X[1] * Y[0] * Z[0], X[0] * Y[1] * Z[0], X[0] * Y[0] * Z[1]};
        const double lapP = X[2] * Y[0] * Z[0] + X[0] * Y[2] * Z[0] + X[0] * Y[0] * Z[2];
        const double Pij[3][3] = {{X[2] * Y[0] * Z[0], X[1] * Y[1] * Z[0], X[1] * Y[0] * Z[1]},
                                  {X[1] * Y[1] * Z[0], X[0] * Y[2] * Z[0], X[0] * Y[1] * Z[1]},
                                  {X[1] * Y[0] * Z[1], X[0] * Y[1] * Z[1], X[0] * Y[0] * Z[2]}};
        const double dlapP[3] = {X[3] * Y[0] * Z[0] + X[1] * Y[2] * Z[0] + X[1] * Y[0] * Z[2],
                                 X[2] * Y[1] * Z[0] + X[0] * Y[3] * Z[0] + X[0] * Y[1] * Z[2],
                                 X[2] * Y[0] * Z[1] + X[0] * Y[2] * Z[1] + X[0] * Y[0] * Z[3]};
        const double f = ao_factor[k];
        double w[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) w[c] = wn[c] * f;
        if (kk + 1 < ncart) load_g(k + 1);
        // r-gradient of the five components, contracted with w
        double gR[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          double gij[3];
#pragma unroll
          for (int i = 0; i < 3; ++i) gij[i] = 4.0 * dd[i] * dd[q] * g2 + (i == q ? 2.0 * g1 : 0.0);
          double acc = w[0] * (Pi[q] * g0 + m0 * gi[q]);
          double cross = 0.0;
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            acc += w[1 + i] * (Pij[i][q] * g0 + Pi[i] * gi[q] + Pi[q] * gi[i] + m0 * gij[i]);
            cross += Pij[i][q] * gi[i] + Pi[i] * gij[i];
          }
          acc += w[4] * (dlapP[q] * g0 + lapP * gi[q] + Pi[q] * lapg + m0 * dd[q] * t3 + 2.0 * cross);
          gR[q] = acc;
        }
        gR0 += gR[0];
        gR1 += gR[1];
        gR2 += gR[2];
        const double md = Pi[0] * dx + Pi[1] * dy + Pi[2] * dz;
        U0 += w[0] * m0 + w[1] * Pi[0] + w[2] * Pi[1] + w[3] * Pi[2] + w[4] * lapP;
        U1 += 2.0 * m0 * (w[1] * dx + w[2] * dy + w[3] * dz) + w[4] * (6.0 * m0 + 4.0 * md);
        U2 += 4.0 * s * w[4] * m0;
      }
    }
    gR0 = wave_sum(gR0);
    gR1 = wave_sum(gR1);
    gR2 = wave_sum(gR2);
    if (lane == 0) {
      myR[(int64_t)d.shell * 3 + 0] = -gR0;
      myR[(int64_t)d.shell * 3 + 1] = -gR1;
      myR[(int64_t)d.shell * 3 + 2] = -gR2;
    }
    for (int j = 0; j < d.prim_num; ++j) {
      const double a = expo[d.prim_start + j];
      const double as = a * s;
      double v = 0.0;
      if (live && !(as > kCutoff)) {
        const double e = coef[d.prim_start + j] * exp(-as);
        v = -s * e * U0 - e * (1.0 - as) * U1 + e * a * (2.0 - as) * U2;
      }
      v = wave_sum(v);
      if (lane == 0) myA[d.prim_start + j] = v;
    }
  }
}

// dR[n][q] = sum over blocks b and the shells s of nucleus n of
// partR[b][s][q], dA[t] = sum over b of partA[b][t], in a fixed order:
// kRedLanes threads per output each sum the blocks b = j, j + kRedLanes, ...
// (b ascending, and for dR the shells ascending within each b), then one
// thread adds the kRedLanes partial sums in j order.  256 threads handle
// 256 / kRedLanes outputs.
constexpr int kRedLanes = 16;
__global__ void __launch_bounds__(kThreads)
adjoint_reduce_kernel(int64_t nblk, int S, int N, int64_t Np, const int64_t *__restrict__ nucleus_index,
                      const int64_t *__restrict__ nucleus_shell_num, const double *__restrict__ partR,
                      const double *__restrict__ partA, double *__restrict__ dR, double *__restrict__ dA) {
  __shared__ double part[kThreads];
  constexpr int per = kThreads / kRedLanes;  // outputs per workgroup
  const int j = threadIdx.x % kRedLanes, o = threadIdx.x / kRedLanes;
  const int64_t nR = dR != nullptr ? 3 * (int64_t)N : 0;
  const int64_t t = blockIdx.x * (int64_t)per + o;  // output: dR entries first, then dA
  double acc = 0.0;
  if (t < nR) {
    const int64_t n = t / 3, q = t - 3 * n;
    const int64_t s0 = nucleus_index[n], s1 = s0 + nucleus_shell_num[n];
    for (int64_t b = j; b < nblk; b += kRedLanes)
      for (int64_t s = s0; s < s1; ++s) acc += partR[(b * S + s) * 3 + q];
  } else if (dA != nullptr && t - nR < Np) {
    const int64_t a = t - nR;
    for (int64_t b = j; b < nblk; b += kRedLanes) acc += partA[b * Np + a];
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  if (j == 0) {
    double sum = 0.0;
#pragma unroll
    for (int i = 0; i < kRedLanes; ++i) sum += part[o * kRedLanes + i];
    if (t < nR) dR[t] = sum;
    else if (dA != nullptr && t - nR < Np) dA[t - nR] = sum;
  }
}

}  // namespace

hipError_t launch_ao_adjoint(hipStream_t s, const DeviceBasis &b, int64_t P, const double *coord,
                             const int32_t *perm, const double *nucl, const double *ao_factor, const double *G,
                             double *partR, double *partA) {
  if (P <= 0) return hipSuccess;
  const int64_t blocks = (P + kTP - 1) / kTP;
  // about 4,096 workgroups, at most one wavefront per shell
  int64_t groups = (4096 + blocks - 1) / blocks;
  const int64_t gmax = (b.shell_num + kThreads / 64 - 1) / (kThreads / 64);
  if (groups > gmax) groups = gmax;
  if (groups < 1) groups = 1;
  if (blocks > 0x7fffffff || groups > 65535) return hipErrorInvalidValue;
  hipLaunchKernelGGL(ao_adjoint_kernel, dim3((unsigned)blocks, (unsigned)groups), dim3(kThreads), 0, s, P,
                     (int)b.shell_num, (int)b.ao_num, (int)b.nucl_num, coord, perm, nucl, b.shells, b.ao_abc,
                     ao_factor, b.expo, b.coef_norm, b.nucl_range, G, partR, partA, b.prim_num, (int)groups);
  return hipGetLastError();
}

int64_t ao_adjoint_blocks(int64_t P) { return (P + kTP - 1) / kTP; }

hipError_t launch_adjoint_reduce(hipStream_t s, const DeviceBasis &b, int64_t nblk,
                                 const int64_t *nucleus_index, const int64_t *nucleus_shell_num,
                                 const double *partR, const double *partA, double *dR, double *dA) {
  const int64_t n = (dR != nullptr ? 3 * b.nucl_num : 0) + (dA != nullptr ? b.prim_num : 0);
  if (n <= 0) return hipSuccess;
  constexpr int per = kThreads / kRedLanes;
  const int64_t g = (n + per - 1) / per;
  if (g > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(adjoint_reduce_kernel, dim3((unsigned)g), dim3(kThreads), 0, s, nblk,
                     (int)b.shell_num, (int)b.nucl_num, b.prim_num, nucleus_index, nucleus_shell_num,
                     partR, partA, dR, dA);
  return hipGetLastError();
}

}  // namespace qmckl_amd
