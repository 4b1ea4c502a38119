// Per-lane AO evaluation shared by the AO kernels (k_ao_rows.hip) and the
// fused AO-and-contraction kernel (k_fused.hip): the Cartesian AOs of one
// shell from its radial sums (src/qmckl_ao_sycl.cpp:291-331 for the
// monomial order, 367-422 for the product rule times ao_factor) pushed in
// AO order into a writer, and an FP32-accuracy e^-x.
#pragma once

#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace qmckl_amd {

// Evaluates the l = L Cartesian AOs of one shell for this lane's point and
// pushes them in the reference order (a descending, then b descending).
template <int L, typename W, typename FP>
__device__ __attribute__((always_inline)) inline void shell_aos(W &w, FP ao_factor, double x, double y, double z,
                                                                const double (&R)[5]) {
  double px[L + 1], py[L + 1], pz[L + 1];
  px[0] = py[0] = pz[0] = 1.0;
#pragma unroll
  for (int e = 1; e <= L; ++e) {
    px[e] = px[e - 1] * x;
    py[e] = py[e - 1] * y;
    pz[e] = pz[e - 1] * z;
  }
  const int k0 = w.k;
  int i = 0;
#pragma unroll
  for (int a = L; a >= 0; --a) {
#pragma unroll
    for (int b = L - a; b >= 0; --b, ++i) {
      const int c = L - a - b;
      const double X0 = px[a], Y0 = py[b], Z0 = pz[c];
      const double X1 = a > 0 ? a * px[a > 0 ? a - 1 : 0] : 0.0;
      const double Y1 = b > 0 ? b * py[b > 0 ? b - 1 : 0] : 0.0;
      const double Z1 = c > 0 ? c * pz[c > 0 ? c - 1 : 0] : 0.0;
      const double X2 = a > 1 ? a * (a - 1) * px[a > 1 ? a - 2 : 0] : 0.0;
      const double Y2 = b > 1 ? b * (b - 1) * py[b > 1 ? b - 2 : 0] : 0.0;
      const double Z2 = c > 1 ? c * (c - 1) * pz[c > 1 ? c - 2 : 0] : 0.0;
      const double yz = Y0 * Z0;
      const double m = X0 * yz;
      const double gx = X1 * yz, gy = X0 * Y1 * Z0, gz = X0 * Y0 * Z1;
      const double lap = X2 * yz + X0 * (Y2 * Z0 + Y0 * Z2);
      const double f = ao_factor[k0 + i];
      double v[5];
      v[0] = m * R[0] * f;
      v[1] = (gx * R[0] + m * R[1]) * f;
      v[2] = (gy * R[0] + m * R[2]) * f;
      v[3] = (gz * R[0] + m * R[3]) * f;
      v[4] = (lap * R[0] + m * R[4] + 2.0 * (gx * R[1] + gy * R[2] + gz * R[3])) * f;
      w.push(v);
    }
  }
}

// (x^e, e x^(e-1), e(e-1) x^(e-2)) by the recurrence
// (f, f', f'') -> (x f, f + x f', 2 f' + x f'').
__device__ inline void pow3(double x, int e, double &f0, double &f1, double &f2) {
  double p0 = 1.0, p1 = 0.0, p2 = 0.0;
  for (int i = 0; i < e; ++i) {
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

// Shells of l > 3 (rare): the same AOs with the exponents as (wavefront-
// uniform) loop counts instead of unrolled code, which keeps the register
// footprint of the kernel at that of l <= 3.
template <typename W, typename FP>
__device__ __attribute__((always_inline)) inline void shell_aos_any(W &w, FP ao_factor, double x, double y,
                                                                    double z, const double (&R)[5], int l) {
  const int k0 = w.k;
  int i = 0;
  for (int a = l; a >= 0; --a) {
    for (int b = l - a; b >= 0; --b, ++i) {
      const int c = l - a - b;
      double X0, X1, X2, Y0, Y1, Y2, Z0, Z1, Z2;
      pow3(x, a, X0, X1, X2);
      pow3(y, b, Y0, Y1, Y2);
      pow3(z, c, Z0, Z1, Z2);
      const double yz = Y0 * Z0;
      const double m = X0 * yz;
      const double gx = X1 * yz, gy = X0 * Y1 * Z0, gz = X0 * Y0 * Z1;
      const double lap = X2 * yz + X0 * (Y2 * Z0 + Y0 * Z2);
      const double f = ao_factor[k0 + i];
      double v[5];
      v[0] = m * R[0] * f;
      v[1] = (gx * R[0] + m * R[1]) * f;
      v[2] = (gy * R[0] + m * R[2]) * f;
      v[3] = (gz * R[0] + m * R[3]) * f;
      v[4] = (lap * R[0] + m * R[4] + 2.0 * (gx * R[1] + gy * R[2] + gz * R[3])) * f;
      w.push(v);
    }
  }
}

// e^-x for 0 <= x <= cutoff with FP32 accuracy, for the FP32 operand of
// the FP32-class forms (which round every value to FP32 anyway): -x log2 e
// = n + f in FP64 (n integer, |f| <= 1/2), 2^f by v_exp_f32 (about 1 ulp),
// times 2^n exactly.  A third of the FP64 exp's instructions.
__device__ __attribute__((always_inline)) inline double exp_neg_f32(double x) {
  const double t = x * -1.4426950408889634074;
  const double n = rint(t);
  return ldexp((double)__builtin_amdgcn_exp2f((float)(t - n)), (int)n);
}

}  // namespace qmckl_amd
