/*
 * FP64 CPU oracle for en_distance -> ao_vgl -> mo_vgl and its adjoint.
 *
 * TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check and the
 * benchmark's CPU-baseline leg use it as the checker.  Nothing in the
 * product library links or calls it.
 *
 * It restates the arithmetic of the reference path, file:line below
 * (paths relative to the reference checkout):
 *   en_distance   src/qmckl_electron.cpp:302-331 and 536-549 ('T','T' call,
 *                 output [P][N]); the plain Euclidean distance is computed,
 *                 not the racy in-place sqrt loop of 544-547.
 *   finalize      src/qmckl_ao_sycl.cpp:1134-1146 (normalised coefficients),
 *                 1168-1179 (nucleus_max_ang_mom), 1209-1228 (nucleus_range).
 *   shell_vgl     src/qmckl_ao_sycl.cpp:40-95 (nucleus screen r^2 >
 *                 cutoff*range at 53, primitive screen alpha r^2 > cutoff at
 *                 72-75, radial sums at 77-84).  Screened pairs are written
 *                 as exact zeros (the reference leaves them unwritten).
 *   ao_vgl        src/qmckl_ao_sycl.cpp:195-334 (monomials in order of
 *                 descending a then descending b) and 337-422 (product-rule
 *                 assembly times ao_factor); AO numbering per
 *                 src/qmckl_ao_sycl.cpp:170-186 (shells in order, AOs of a
 *                 shell contiguous).
 *   mo_vgl        src/qmckl_mo_sycl.cpp:23-50 (per point, k ascending,
 *                 mo[c][i] += coefficient_t[k][i] * ao[c][k]) with
 *                 coefficient_t[k][i] = coefficient[i][k]
 *                 (src/qmckl_mo_sycl.cpp:176-184).  The sum is dense: the
 *                 reference's skip of AOs whose value is exactly 0
 *                 (src/qmckl_mo_sycl.cpp:40) changes nothing for screened
 *                 AOs, whose five components are all 0.
 *   adjoint       no reference counterpart; analytic derivatives of the
 *                 expressions above (radial part R(s), s = |r - R_n|^2).
 *
 * r^2 is formed as fma(x,x,fma(y,y,z*z)) so that the screening decisions
 * are bit-identical with the device kernels, which form it the same way.
 */
#include "qmckl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NCART_MAX (((QO_MAX_L) + 1) * ((QO_MAX_L) + 2) / 2)

static int g_threads = 0;

int qo_set_num_threads(int n) {
  g_threads = n;
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#endif
  return 0;
}

int qo_get_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static inline double r2_of(double x, double y, double z) {
  return fma(x, x, fma(y, y, z * z));
}

int qo_en_distance(int64_t P, int64_t N, const double *coord,
                   const double *nucl_coord, double *en_distance) {
  if (P < 0 || N <= 0) return 1;
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < P; ++p) {
    for (int64_t n = 0; n < N; ++n) {
      double x = coord[p] - nucl_coord[n];
      double y = coord[P + p] - nucl_coord[N + n];
      double z = coord[2 * P + p] - nucl_coord[2 * N + n];
      en_distance[p * N + n] = sqrt(r2_of(x, y, z));
    }
  }
  return 0;
}

/* Electron-electron distances per walker (src/qmckl_electron.cpp:333-371:
 * qmckl_distance('T','T') of each walker's electrons against themselves):
 * ee_distance[w][j][i] = |r_(w,i) - r_(w,j)|, coord [3][W*E], point w*E + i. */
int qo_ee_distance(int64_t E, int64_t W, const double *coord, double *ee_distance) {
  if (E <= 0 || W <= 0) return 1;
  const int64_t P = E * W;
#pragma omp parallel for schedule(static)
  for (int64_t w = 0; w < W; ++w) {
    for (int64_t j = 0; j < E; ++j) {
      for (int64_t i = 0; i < E; ++i) {
        const int64_t a = w * E + i, b = w * E + j;
        double x = coord[a] - coord[b];
        double y = coord[P + a] - coord[P + b];
        double z = coord[2 * P + a] - coord[2 * P + b];
        ee_distance[(w * E + j) * E + i] = sqrt(r2_of(x, y, z));
      }
    }
  }
  return 0;
}

int qo_finalize_ao(int64_t N, int64_t S, int64_t Np,
                   const int64_t *nucleus_index, const int64_t *nucleus_shell_num,
                   const int32_t *shell_ang_mom, const int64_t *shell_prim_num,
                   const int64_t *shell_prim_index, const double *shell_factor,
                   const double *exponent, const double *coefficient,
                   const double *prim_factor, double *coef_norm,
                   int32_t *nucleus_max_ang_mom, double *nucleus_range) {
  for (int64_t s = 0; s < S; ++s) {
    if (shell_ang_mom[s] < 0 || shell_ang_mom[s] > QO_MAX_L) return 2;
    if (shell_prim_index[s] < 0 || shell_prim_index[s] + shell_prim_num[s] > Np) return 3;
    for (int64_t j = shell_prim_index[s]; j < shell_prim_index[s] + shell_prim_num[s]; ++j)
      coef_norm[j] = coefficient[j] * prim_factor[j] * shell_factor[s];
  }
  for (int64_t n = 0; n < N; ++n) {
    int32_t lmax = 0;
    double range = 0.0;
    if (nucleus_index[n] < 0 || nucleus_index[n] + nucleus_shell_num[n] > S) return 4;
    for (int64_t s = nucleus_index[n]; s < nucleus_index[n] + nucleus_shell_num[n]; ++s) {
      if (shell_ang_mom[s] > lmax) lmax = shell_ang_mom[s];
      for (int64_t j = shell_prim_index[s]; j < shell_prim_index[s] + shell_prim_num[s]; ++j) {
        double r = 1.0 / exponent[j];
        if (r > range) range = r;
      }
    }
    nucleus_max_ang_mom[n] = lmax;
    nucleus_range[n] = range;
  }
  return 0;
}

/* Cartesian exponents of shell l, reference order: a from l down to 0, then
 * b from l-a down to 0, c = l-a-b.  Returns the count. */
static int cart_table(int l, int abc[][3]) {
  int n = 0;
  for (int a = l; a >= 0; --a)
    for (int b = l - a; b >= 0; --b) {
      abc[n][0] = a;
      abc[n][1] = b;
      abc[n][2] = l - a - b;
      ++n;
    }
  return n;
}

/* d^k/dx^k x^a = a!/(a-k)! x^(a-k) */
static inline double dpow(int a, int k, const double *pw) {
  if (a < k) return 0.0;
  double f = 1.0;
  for (int i = 0; i < k; ++i) f *= (double)(a - i);
  return f * pw[a - k];
}

/* Monomial x^a y^b z^c and its derivatives.
 * m[0] value, m[1..3] gradient, m[4] Laplacian,
 * m[5..10] Hessian xx xy xz yy yz zz, m[11..13] gradient of the Laplacian. */
static void monomial(int a, int b, int c, const double *px, const double *py,
                     const double *pz, int order, double *m) {
  double X0 = px[a], Y0 = py[b], Z0 = pz[c];
  double X1 = dpow(a, 1, px), Y1 = dpow(b, 1, py), Z1 = dpow(c, 1, pz);
  double X2 = dpow(a, 2, px), Y2 = dpow(b, 2, py), Z2 = dpow(c, 2, pz);
  m[0] = X0 * Y0 * Z0;
  m[1] = X1 * Y0 * Z0;
  m[2] = X0 * Y1 * Z0;
  m[3] = X0 * Y0 * Z1;
  m[4] = X2 * Y0 * Z0 + X0 * Y2 * Z0 + X0 * Y0 * Z2;
  if (order < 3) return;
  double X3 = dpow(a, 3, px), Y3 = dpow(b, 3, py), Z3 = dpow(c, 3, pz);
  m[5] = X2 * Y0 * Z0;
  m[6] = X1 * Y1 * Z0;
  m[7] = X1 * Y0 * Z1;
  m[8] = X0 * Y2 * Z0;
  m[9] = X0 * Y1 * Z1;
  m[10] = X0 * Y0 * Z2;
  m[11] = X3 * Y0 * Z0 + X1 * Y2 * Z0 + X1 * Y0 * Z2;
  m[12] = X2 * Y1 * Z0 + X0 * Y3 * Z0 + X0 * Y1 * Z2;
  m[13] = X2 * Y0 * Z1 + X0 * Y2 * Z1 + X0 * Y0 * Z3;
}

typedef struct {
  int64_t N, S, Np, A;
  const double *nucl_coord;
  const int64_t *nucleus_index, *nucleus_shell_num, *shell_prim_num, *shell_prim_index;
  const int32_t *shell_ang_mom;
  const double *exponent, *ao_factor;
  double *coef_norm, *nucleus_range;
  int32_t *nucleus_max_ang_mom;
  int64_t *shell_ao_start; /* first AO of each shell */
} basis_t;

static int basis_init(basis_t *B, int64_t N, int64_t S, int64_t Np, int64_t A,
                      const double *nucl_coord, const int64_t *nucleus_index,
                      const int64_t *nucleus_shell_num, const int32_t *shell_ang_mom,
                      const int64_t *shell_prim_num, const int64_t *shell_prim_index,
                      const double *shell_factor, const double *exponent,
                      const double *coefficient, const double *prim_factor,
                      const double *ao_factor) {
  memset(B, 0, sizeof(*B));
  B->N = N; B->S = S; B->Np = Np; B->A = A;
  B->nucl_coord = nucl_coord;
  B->nucleus_index = nucleus_index;
  B->nucleus_shell_num = nucleus_shell_num;
  B->shell_prim_num = shell_prim_num;
  B->shell_prim_index = shell_prim_index;
  B->shell_ang_mom = shell_ang_mom;
  B->exponent = exponent;
  B->ao_factor = ao_factor;
  B->coef_norm = (double *)malloc(sizeof(double) * (size_t)(Np > 0 ? Np : 1));
  B->nucleus_range = (double *)malloc(sizeof(double) * (size_t)N);
  B->nucleus_max_ang_mom = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
  B->shell_ao_start = (int64_t *)malloc(sizeof(int64_t) * (size_t)(S + 1));
  if (!B->coef_norm || !B->nucleus_range || !B->nucleus_max_ang_mom || !B->shell_ao_start)
    return 5;
  int rc = qo_finalize_ao(N, S, Np, nucleus_index, nucleus_shell_num, shell_ang_mom,
                          shell_prim_num, shell_prim_index, shell_factor, exponent,
                          coefficient, prim_factor, B->coef_norm,
                          B->nucleus_max_ang_mom, B->nucleus_range);
  if (rc) return rc;
  /* AO numbering walks nuclei in order and, within a nucleus, its shells
   * (src/qmckl_ao_sycl.cpp:170-186). */
  int64_t k = 0;
  for (int64_t s = 0; s < S; ++s) B->shell_ao_start[s] = -1;
  for (int64_t n = 0; n < N; ++n)
    for (int64_t s = nucleus_index[n]; s < nucleus_index[n] + nucleus_shell_num[n]; ++s) {
      int l = shell_ang_mom[s];
      B->shell_ao_start[s] = k;
      k += (l + 1) * (l + 2) / 2;
    }
  B->shell_ao_start[S] = k;
  if (k != A) return 6;
  return 0;
}

static void basis_free(basis_t *B) {
  free(B->coef_norm);
  free(B->nucleus_range);
  free(B->nucleus_max_ang_mom);
  free(B->shell_ao_start);
}

/* Radial sums of one shell at one point (src/qmckl_ao_sycl.cpp:70-85).
 * t[0] R, t[1..3] grad R, t[4] lap R.  mag[] gets the same sums over
 * absolute values of the terms.  Returns 0 if every primitive is screened. */
static int radial(const basis_t *B, int64_t s, double x, double y, double z,
                  double r2, double *t, double *mag) {
  int any = 0;
  for (int i = 0; i < 5; ++i) t[i] = mag[i] = 0.0;
  for (int64_t j = B->shell_prim_index[s]; j < B->shell_prim_index[s] + B->shell_prim_num[s]; ++j) {
    double ar2 = B->exponent[j] * r2;
    if (ar2 > QO_CUTOFF) continue;
    any = 1;
    double v = B->coef_norm[j] * exp(-ar2);
    double two_a = -2.0 * B->exponent[j] * v;
    double q = 3.0 - 2.0 * ar2;
    t[0] += v;
    t[1] += two_a * x;
    t[2] += two_a * y;
    t[3] += two_a * z;
    t[4] += two_a * q;
    mag[0] += fabs(v);
    mag[1] += fabs(two_a * x);
    mag[2] += fabs(two_a * y);
    mag[3] += fabs(two_a * z);
    mag[4] += fabs(two_a) * (3.0 + 2.0 * ar2);
  }
  return any;
}

int qo_shell_vgl(int64_t P, int64_t N, int64_t S, int64_t Np, const double *coord,
                 const double *nucl_coord, const int64_t *nucleus_index,
                 const int64_t *nucleus_shell_num, const int32_t *shell_ang_mom,
                 const int64_t *shell_prim_num, const int64_t *shell_prim_index,
                 const double *shell_factor, const double *exponent,
                 const double *coefficient, const double *prim_factor,
                 double *shell_vgl) {
  basis_t B;
  /* ao_num is not needed here; count it so basis_init's check passes. */
  int64_t A = 0;
  for (int64_t s = 0; s < S; ++s) A += (shell_ang_mom[s] + 1) * (shell_ang_mom[s] + 2) / 2;
  int rc = basis_init(&B, N, S, Np, A, nucl_coord, nucleus_index, nucleus_shell_num,
                      shell_ang_mom, shell_prim_num, shell_prim_index, shell_factor,
                      exponent, coefficient, prim_factor, NULL);
  if (rc) { basis_free(&B); return rc; }
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < P; ++p) {
    double *out = shell_vgl + p * 5 * S;
    for (int64_t n = 0; n < N; ++n) {
      double x = coord[p] - nucl_coord[n];
      double y = coord[P + p] - nucl_coord[N + n];
      double z = coord[2 * P + p] - nucl_coord[2 * N + n];
      double r2 = r2_of(x, y, z);
      int far = r2 > QO_CUTOFF * B.nucleus_range[n];
      for (int64_t s = nucleus_index[n]; s < nucleus_index[n] + nucleus_shell_num[n]; ++s) {
        double t[5] = {0, 0, 0, 0, 0}, mag[5];
        if (!far) radial(&B, s, x, y, z, r2, t, mag);
        for (int c = 0; c < 5; ++c) out[c * S + s] = t[c];
      }
    }
  }
  basis_free(&B);
  return 0;
}

/* Shared forward AO evaluation at one point; writes [5][A] rows (value-only
 * if ncomp == 1). */
static void ao_point(const basis_t *B, const double *r, int ncomp, double *ao,
                     int64_t ld, double *aomag) {
  const int64_t N = B->N;
  for (int64_t n = 0; n < N; ++n) {
    double x = r[0] - B->nucl_coord[n];
    double y = r[1] - B->nucl_coord[N + n];
    double z = r[2] - B->nucl_coord[2 * N + n];
    double r2 = r2_of(x, y, z);
    int far = r2 > QO_CUTOFF * B->nucleus_range[n];
    double px[QO_MAX_L + 1], py[QO_MAX_L + 1], pz[QO_MAX_L + 1];
    px[0] = py[0] = pz[0] = 1.0;
    for (int i = 1; i <= QO_MAX_L; ++i) {
      px[i] = px[i - 1] * x;
      py[i] = py[i - 1] * y;
      pz[i] = pz[i - 1] * z;
    }
    for (int64_t s = B->nucleus_index[n]; s < B->nucleus_index[n] + B->nucleus_shell_num[n]; ++s) {
      int l = B->shell_ang_mom[s];
      int abc[NCART_MAX][3];
      int nc = cart_table(l, abc);
      int64_t k0 = B->shell_ao_start[s];
      double t[5], tm[5];
      int any = far ? 0 : radial(B, s, x, y, z, r2, t, tm);
      for (int j = 0; j < nc; ++j) {
        int64_t k = k0 + j;
        if (!any) {
          for (int c = 0; c < ncomp; ++c) {
            ao[c * ld + k] = 0.0;
            if (aomag) aomag[c * ld + k] = 0.0;
          }
          continue;
        }
        double m[14];
        monomial(abc[j][0], abc[j][1], abc[j][2], px, py, pz, 2, m);
        double f = B->ao_factor[k];
        ao[k] = m[0] * t[0] * f;
        if (aomag) aomag[k] = fabs(m[0]) * tm[0] * fabs(f);
        if (ncomp == 1) continue;
        for (int i = 1; i <= 3; ++i) {
          ao[i * ld + k] = (m[i] * t[0] + m[0] * t[i]) * f;
          if (aomag) aomag[i * ld + k] = (fabs(m[i]) * tm[0] + fabs(m[0]) * tm[i]) * fabs(f);
        }
        ao[4 * ld + k] = (m[4] * t[0] + m[0] * t[4] +
                          2.0 * (m[1] * t[1] + m[2] * t[2] + m[3] * t[3])) * f;
        if (aomag)
          aomag[4 * ld + k] = (fabs(m[4]) * tm[0] + fabs(m[0]) * tm[4] +
                               2.0 * (fabs(m[1]) * tm[1] + fabs(m[2]) * tm[2] +
                                      fabs(m[3]) * tm[3])) * fabs(f);
      }
    }
  }
}

int qo_ao_vgl(int64_t P, int64_t N, int64_t S, int64_t Np, int64_t A,
              const double *coord, const double *nucl_coord,
              const int64_t *nucleus_index, const int64_t *nucleus_shell_num,
              const int32_t *shell_ang_mom, const int64_t *shell_prim_num,
              const int64_t *shell_prim_index, const double *shell_factor,
              const double *exponent, const double *coefficient,
              const double *prim_factor, const double *ao_factor,
              double *ao_vgl, double *ao_mag) {
  basis_t B;
  int rc = basis_init(&B, N, S, Np, A, nucl_coord, nucleus_index, nucleus_shell_num,
                      shell_ang_mom, shell_prim_num, shell_prim_index, shell_factor,
                      exponent, coefficient, prim_factor, ao_factor);
  if (rc) { basis_free(&B); return rc; }
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t p = 0; p < P; ++p) {
    double r[3] = {coord[p], coord[P + p], coord[2 * P + p]};
    ao_point(&B, r, 5, ao_vgl + p * 5 * A, A, ao_mag ? ao_mag + p * 5 * A : NULL);
  }
  basis_free(&B);
  return 0;
}

int qo_ao_value(int64_t P, int64_t N, int64_t S, int64_t Np, int64_t A,
                const double *coord, const double *nucl_coord,
                const int64_t *nucleus_index, const int64_t *nucleus_shell_num,
                const int32_t *shell_ang_mom, const int64_t *shell_prim_num,
                const int64_t *shell_prim_index, const double *shell_factor,
                const double *exponent, const double *coefficient,
                const double *prim_factor, const double *ao_factor,
                double *ao_value) {
  basis_t B;
  int rc = basis_init(&B, N, S, Np, A, nucl_coord, nucleus_index, nucleus_shell_num,
                      shell_ang_mom, shell_prim_num, shell_prim_index, shell_factor,
                      exponent, coefficient, prim_factor, ao_factor);
  if (rc) { basis_free(&B); return rc; }
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t p = 0; p < P; ++p) {
    double r[3] = {coord[p], coord[P + p], coord[2 * P + p]};
    ao_point(&B, r, 1, ao_value + p * A, A, NULL);
  }
  basis_free(&B);
  return 0;
}

int qo_mo_vgl(int64_t P, int64_t A, int64_t M, const double *mo_coef,
              const double *ao_vgl, double *mo_vgl, const double *ao_mag,
              double *mo_mag) {
  double *ct = (double *)malloc(sizeof(double) * (size_t)(A * M));
  double *cta = (ao_mag && mo_mag) ? (double *)malloc(sizeof(double) * (size_t)(A * M)) : NULL;
  if (!ct) return 5;
  for (int64_t k = 0; k < A; ++k)
    for (int64_t i = 0; i < M; ++i) {
      ct[k * M + i] = mo_coef[i * A + k];
      if (cta) cta[k * M + i] = fabs(mo_coef[i * A + k]);
    }
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t p = 0; p < P; ++p) {
    const double *ao = ao_vgl + p * 5 * A;
    double *mo = mo_vgl + p * 5 * M;
    for (int64_t i = 0; i < 5 * M; ++i) mo[i] = 0.0;
    for (int64_t k = 0; k < A; ++k) {
      const double *ck = ct + k * M;
      for (int c = 0; c < 5; ++c) {
        double a = ao[c * A + k];
        double *row = mo + c * M;
        for (int64_t i = 0; i < M; ++i) row[i] += ck[i] * a;
      }
    }
    if (cta) {
      const double *am = ao_mag + p * 5 * A;
      double *mm = mo_mag + p * 5 * M;
      for (int64_t i = 0; i < 5 * M; ++i) mm[i] = 0.0;
      for (int64_t k = 0; k < A; ++k) {
        const double *ck = cta + k * M;
        for (int c = 0; c < 5; ++c) {
          double a = am[c * A + k];
          double *row = mm + c * M;
          for (int64_t i = 0; i < M; ++i) row[i] += ck[i] * a;
        }
      }
    }
  }
  free(ct);
  free(cta);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Adjoint.                                                                */
/*                                                                         */
/* With d = r - R_n, s = |d|^2 and a radial function g(s), the AO          */
/* components are f*(P g), f*d_i(P g), f*lap(P g).  Their r-derivatives    */
/* need the Hessian and the gradient of the Laplacian of P*g:              */
/*   d_i g = 2 d_i g',  d_ij g = 4 d_i d_j g'' + 2 delta_ij g',            */
/*   lap g = 4 s g'' + 6 g',  d_j lap g = 2 d_j (10 g'' + 4 s g''').       */
/* dAO/dR_n = -dAO/dr.  The exponent derivative at fixed normalised        */
/* coefficient replaces g by h_j = -s c_j e^{-a_j s} (one primitive):      */
/*   h' = -c e^{-as}(1 - a s),  h'' = c a e^{-as}(2 - a s).                */
/* ---------------------------------------------------------------------- */

/* Value (0), gradient (1..3) and Laplacian (4) of P*g given monomial data m
 * and radial derivatives g0 = g, g1 = g', g2 = g''. */
static void vgl_of(const double *m, double dx, double dy, double dz, double s,
                   double g0, double g1, double g2, double *o) {
  double gi[3] = {2.0 * dx * g1, 2.0 * dy * g1, 2.0 * dz * g1};
  double lapg = 4.0 * s * g2 + 6.0 * g1;
  o[0] = m[0] * g0;
  o[1] = m[1] * g0 + m[0] * gi[0];
  o[2] = m[2] * g0 + m[0] * gi[1];
  o[3] = m[3] * g0 + m[0] * gi[2];
  o[4] = m[4] * g0 + m[0] * lapg + 2.0 * (m[1] * gi[0] + m[2] * gi[1] + m[3] * gi[2]);
}

/* r-gradient of the five components: out[c][j] = d_j component_c. */
static void dvgl_dr(const double *m, const double *d, double s, double g0,
                    double g1, double g2, double g3, double out[5][3]) {
  double gi[3], gij[3][3], dlapg[3];
  for (int i = 0; i < 3; ++i) gi[i] = 2.0 * d[i] * g1;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      gij[i][j] = 4.0 * d[i] * d[j] * g2 + (i == j ? 2.0 * g1 : 0.0);
  double lapg = 4.0 * s * g2 + 6.0 * g1;
  for (int j = 0; j < 3; ++j) dlapg[j] = 2.0 * d[j] * (10.0 * g2 + 4.0 * s * g3);
  /* monomial Hessian in full form and gradient of its Laplacian */
  double Pij[3][3] = {{m[5], m[6], m[7]}, {m[6], m[8], m[9]}, {m[7], m[9], m[10]}};
  const double *Pi = m + 1;
  const double *dlapP = m + 11;
  double lapP = m[4];
  for (int j = 0; j < 3; ++j) {
    out[0][j] = Pi[j] * g0 + m[0] * gi[j];
    for (int i = 0; i < 3; ++i)
      out[1 + i][j] = Pij[i][j] * g0 + Pi[i] * gi[j] + Pi[j] * gi[i] + m[0] * gij[i][j];
    double cross = 0.0;
    for (int i = 0; i < 3; ++i) cross += Pij[i][j] * gi[i] + Pi[i] * gij[i][j];
    out[4][j] = dlapP[j] * g0 + lapP * gi[j] + Pi[j] * lapg + m[0] * dlapg[j] + 2.0 * cross;
  }
}

int qo_mo_vgl_adjoint(int64_t P, int64_t N, int64_t S, int64_t Np, int64_t A,
                      int64_t M, const double *coord, const double *nucl_coord,
                      const int64_t *nucleus_index, const int64_t *nucleus_shell_num,
                      const int32_t *shell_ang_mom, const int64_t *shell_prim_num,
                      const int64_t *shell_prim_index, const double *shell_factor,
                      const double *exponent, const double *coefficient,
                      const double *prim_factor, const double *ao_factor,
                      const double *mo_coef, const double *w,
                      double *d_nucl_coord, double *d_exponent, double *d_mo_coef) {
  basis_t B;
  int rc = basis_init(&B, N, S, Np, A, nucl_coord, nucleus_index, nucleus_shell_num,
                      shell_ang_mom, shell_prim_num, shell_prim_index, shell_factor,
                      exponent, coefficient, prim_factor, ao_factor);
  if (rc) { basis_free(&B); return rc; }
  int nth = qo_get_num_threads();
  double *accR = (double *)calloc((size_t)nth * N * 3, sizeof(double));
  double *accA = (double *)calloc((size_t)nth * (Np > 0 ? Np : 1), sizeof(double));
  double *accC = d_mo_coef ? (double *)calloc((size_t)nth * M * A, sizeof(double)) : NULL;
  if (!accR || !accA || (d_mo_coef && !accC)) { basis_free(&B); return 5; }

#pragma omp parallel num_threads(nth)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double *myR = accR + (size_t)tid * N * 3;
    double *myA = accA + (size_t)tid * (Np > 0 ? Np : 1);
    double *myC = accC ? accC + (size_t)tid * M * A : NULL;
    double *ao = (double *)malloc(sizeof(double) * 5 * A);
    double *G = (double *)malloc(sizeof(double) * 5 * A);
#pragma omp for schedule(dynamic, 4)
    for (int64_t p = 0; p < P; ++p) {
      const double *wp = w + p * 5 * M;
      double r[3] = {coord[p], coord[P + p], coord[2 * P + p]};
      /* G[c][k] = sum_i w[p][c][i] C[i][k] */
      for (int64_t i = 0; i < 5 * A; ++i) G[i] = 0.0;
      for (int c = 0; c < 5; ++c)
        for (int64_t i = 0; i < M; ++i) {
          double wi = wp[c * M + i];
          if (wi == 0.0) continue;
          const double *ci = mo_coef + i * A;
          double *gr = G + c * A;
          for (int64_t k = 0; k < A; ++k) gr[k] += wi * ci[k];
        }
      if (myC) {
        ao_point(&B, r, 5, ao, A, NULL);
        for (int c = 0; c < 5; ++c)
          for (int64_t i = 0; i < M; ++i) {
            double wi = wp[c * M + i];
            if (wi == 0.0) continue;
            const double *ar = ao + c * A;
            double *dc = myC + i * A;
            for (int64_t k = 0; k < A; ++k) dc[k] += wi * ar[k];
          }
      }
      for (int64_t n = 0; n < N; ++n) {
        double d[3] = {r[0] - nucl_coord[n], r[1] - nucl_coord[N + n], r[2] - nucl_coord[2 * N + n]};
        double s = r2_of(d[0], d[1], d[2]);
        if (s > QO_CUTOFF * B.nucleus_range[n]) continue;
        double px[QO_MAX_L + 1], py[QO_MAX_L + 1], pz[QO_MAX_L + 1];
        px[0] = py[0] = pz[0] = 1.0;
        for (int i = 1; i <= QO_MAX_L; ++i) {
          px[i] = px[i - 1] * d[0];
          py[i] = py[i - 1] * d[1];
          pz[i] = pz[i - 1] * d[2];
        }
        for (int64_t sh = nucleus_index[n]; sh < nucleus_index[n] + nucleus_shell_num[n]; ++sh) {
          double g0 = 0, g1 = 0, g2 = 0, g3 = 0;
          int any = 0;
          for (int64_t j = shell_prim_index[sh]; j < shell_prim_index[sh] + shell_prim_num[sh]; ++j) {
            double a = exponent[j], as = a * s;
            if (as > QO_CUTOFF) continue;
            any = 1;
            double v = B.coef_norm[j] * exp(-as);
            g0 += v;
            g1 -= a * v;
            g2 += a * a * v;
            g3 -= a * a * a * v;
          }
          if (!any) continue;
          int l = shell_ang_mom[sh];
          int abc[NCART_MAX][3];
          int nc = cart_table(l, abc);
          int64_t k0 = B.shell_ao_start[sh];
          double gR[3] = {0, 0, 0};
          /* per-AO monomials and G weights, reused for the exponent part */
          double mm[NCART_MAX][14];
          double gw[NCART_MAX][5];
          for (int jj = 0; jj < nc; ++jj) {
            int64_t k = k0 + jj;
            monomial(abc[jj][0], abc[jj][1], abc[jj][2], px, py, pz, 3, mm[jj]);
            double f = ao_factor[k];
            for (int c = 0; c < 5; ++c) gw[jj][c] = G[c * A + k] * f;
            double dv[5][3];
            dvgl_dr(mm[jj], d, s, g0, g1, g2, g3, dv);
            for (int c = 0; c < 5; ++c)
              for (int q = 0; q < 3; ++q) gR[q] += gw[jj][c] * dv[c][q];
          }
          for (int q = 0; q < 3; ++q) myR[n * 3 + q] -= gR[q];
          for (int64_t j = shell_prim_index[sh]; j < shell_prim_index[sh] + shell_prim_num[sh]; ++j) {
            double a = exponent[j], as = a * s;
            if (as > QO_CUTOFF) continue;
            double e = B.coef_norm[j] * exp(-as);
            double h0 = -s * e;
            double h1 = -e * (1.0 - as);
            double h2 = e * a * (2.0 - as);
            double acc = 0.0;
            for (int jj = 0; jj < nc; ++jj) {
              double o[5];
              vgl_of(mm[jj], d[0], d[1], d[2], s, h0, h1, h2, o);
              for (int c = 0; c < 5; ++c) acc += gw[jj][c] * o[c];
            }
            myA[j] += acc;
          }
        }
      }
    }
    free(ao);
    free(G);
  }
  if (d_nucl_coord) {
    for (int64_t i = 0; i < N * 3; ++i) {
      double t = 0.0;
      for (int th = 0; th < nth; ++th) t += accR[(size_t)th * N * 3 + i];
      d_nucl_coord[i] = t;
    }
  }
  if (d_exponent) {
    for (int64_t i = 0; i < Np; ++i) {
      double t = 0.0;
      for (int th = 0; th < nth; ++th) t += accA[(size_t)th * (Np > 0 ? Np : 1) + i];
      d_exponent[i] = t;
    }
  }
  if (d_mo_coef) {
    for (int64_t i = 0; i < M * A; ++i) {
      double t = 0.0;
      for (int th = 0; th < nth; ++th) t += accC[(size_t)th * M * A + i];
      d_mo_coef[i] = t;
    }
  }
  free(accR);
  free(accA);
  free(accC);
  basis_free(&B);
  return 0;
}
