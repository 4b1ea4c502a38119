/*
 * FP64 CPU oracle for the per-electron basis path (en_distance -> ao_vgl ->
 * mo_vgl, and the adjoint of a weighted sum of mo_vgl).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the parity tests
 * and the benchmark's CPU-baseline leg compare against.  It is never linked
 * into, loaded by or called from the product library; the product fails
 * loudly when its HIP library is missing instead of falling back here.
 *
 * Conventions (the reference's, see the .c file for file:line citations):
 *   coord        [3][P]  point coordinates, structure of arrays
 *   nucl_coord   [3][N]  nucleus coordinates, structure of arrays
 *   en_distance  [P][N]
 *   shell_vgl    [P][5][S]
 *   ao_vgl       [P][5][A]   component 0 value, 1..3 d/dx d/dy d/dz, 4 Laplacian
 *   mo_coef      [M][A]      (the layout set_mo_basis_coefficient takes)
 *   mo_vgl       [P][5][M]
 */
#ifndef QMCKL_ORACLE_H
#define QMCKL_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* -ln(1e-12): primitives with alpha*r^2 above this are skipped. */
#define QO_CUTOFF 27.631021115928547
/* Highest shell angular momentum the oracle evaluates. */
#define QO_MAX_L 7

int qo_set_num_threads(int n);
int qo_get_num_threads(void);

int qo_en_distance(int64_t P, int64_t N, const double *coord,
                   const double *nucl_coord, double *en_distance);

/* ee_distance [W][E][E], coord [3][W*E] (point w*E + i). */
int qo_ee_distance(int64_t E, int64_t W, const double *coord, double *ee_distance);

/* Derived basis data computed at finalize time.  coef_norm[Np],
 * nucleus_max_ang_mom[N], nucleus_range[N]. */
int qo_finalize_ao(int64_t N, int64_t S, int64_t Np,
                   const int64_t *nucleus_index, const int64_t *nucleus_shell_num,
                   const int32_t *shell_ang_mom, const int64_t *shell_prim_num,
                   const int64_t *shell_prim_index, const double *shell_factor,
                   const double *exponent, const double *coefficient,
                   const double *prim_factor, double *coef_norm,
                   int32_t *nucleus_max_ang_mom, double *nucleus_range);

/* Radial shell VGL, [P][5][S]; screened (point, shell) pairs are 0. */
int qo_shell_vgl(int64_t P, int64_t N, int64_t S, int64_t Np, const double *coord,
                 const double *nucl_coord, const int64_t *nucleus_index,
                 const int64_t *nucleus_shell_num, const int32_t *shell_ang_mom,
                 const int64_t *shell_prim_num, const int64_t *shell_prim_index,
                 const double *shell_factor, const double *exponent,
                 const double *coefficient, const double *prim_factor,
                 double *shell_vgl);

/* AO VGL, [P][5][A]. ao_mag (nullable) receives, per entry, the sum of the
 * absolute values of the terms that make it up (the scale a rounding error
 * is measured against). */
int qo_ao_vgl(int64_t P, int64_t N, int64_t S, int64_t Np, int64_t A,
              const double *coord, const double *nucl_coord,
              const int64_t *nucleus_index, const int64_t *nucleus_shell_num,
              const int32_t *shell_ang_mom, const int64_t *shell_prim_num,
              const int64_t *shell_prim_index, const double *shell_factor,
              const double *exponent, const double *coefficient,
              const double *prim_factor, const double *ao_factor,
              double *ao_vgl, double *ao_mag);

/* AO value only, [P][A] (component 0 of ao_vgl). */
int qo_ao_value(int64_t P, int64_t N, int64_t S, int64_t Np, int64_t A,
                const double *coord, const double *nucl_coord,
                const int64_t *nucleus_index, const int64_t *nucleus_shell_num,
                const int32_t *shell_ang_mom, const int64_t *shell_prim_num,
                const int64_t *shell_prim_index, const double *shell_factor,
                const double *exponent, const double *coefficient,
                const double *prim_factor, const double *ao_factor,
                double *ao_value);

/* MO VGL, [P][5][M] = sum_k mo_coef[i][k] * ao_vgl[p][c][k].  If ao_mag and
 * mo_mag are both non-null, mo_mag[p][c][i] = sum_k |mo_coef[i][k]| ao_mag. */
int qo_mo_vgl(int64_t P, int64_t A, int64_t M, const double *mo_coef,
              const double *ao_vgl, double *mo_vgl, const double *ao_mag,
              double *mo_mag);

/* Adjoint of L = sum_{p,c,i} w[p][c][i] * mo_vgl[p][c][i]:
 * d_nucl_coord[N][3], d_exponent[Np] (at fixed coefficient, prim_factor and
 * shell_factor), d_mo_coef[M][A].  Any output pointer may be null. */
int qo_mo_vgl_adjoint(int64_t P, int64_t N, int64_t S, int64_t Np, int64_t A,
                      int64_t M, const double *coord, const double *nucl_coord,
                      const int64_t *nucleus_index, const int64_t *nucleus_shell_num,
                      const int32_t *shell_ang_mom, const int64_t *shell_prim_num,
                      const int64_t *shell_prim_index, const double *shell_factor,
                      const double *exponent, const double *coefficient,
                      const double *prim_factor, const double *ao_factor,
                      const double *mo_coef, const double *w,
                      double *d_nucl_coord, double *d_exponent, double *d_mo_coef);

#ifdef __cplusplus
}
#endif
#endif
