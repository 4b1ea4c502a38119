/*
 * qmckl_gpu.h -- C ABI of the MI355X-native per-electron basis path.
 *
 * Drop-in for the en_distance -> ao_vgl -> mo_vgl path of the QMCkl GPU
 * context/provider API as declared by the reference in
 * include/qmckl_gpu.hpp (paths relative to the reference checkout).  Every
 * entry point keeps the reference's name, argument order and meaning and
 * its exit codes; the comment above each cites the declaration it
 * replaces.  Differences:
 *   - the header is plain C with C linkage (the reference's symbols are
 *     C++-mangled), so cgo / JNI / ctypes / Fortran bind it directly;
 *   - qmckl_context_create_device takes a HIP device ordinal instead of a
 *     sycl::queue (the context owns one HIP stream on that device);
 *     device_id == QMCKL_NO_DEVICE creates a context with no device bound,
 *     on which argument validation, masks and scalar setters work and every
 *     device operation fails with QMCKL_FAILURE_DEVICE;
 *   - getters that the reference lacks are added where the path needs them
 *     (qmckl_get_electron_en_distance_device, qmckl_get_error_device);
 *   - the precision, batched and adjoint forms are new entry points at the
 *     end of this file.
 *
 * All array arguments are device pointers (HIP device memory) unless stated
 * otherwise; scalar outputs are host pointers.  Setters deep-copy their
 * inputs; getters copy into a caller buffer after checking size_max and
 * return with the data complete (the call synchronises the context's
 * stream).  One context is not safe for concurrent compute from several
 * threads (as in the reference, the mutex guards the memory table and the
 * error record only).
 */
#ifndef QMCKL_GPU_H
#define QMCKL_GPU_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The library is built with hidden visibility; these are its exports. */
#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

/* ---- types and codes (include/qmckl_types.hpp:12-116) ------------------ */

typedef int32_t qmckl_exit_code_device;
typedef int64_t qmckl_context_device;

#define QMCKL_NULL_CONTEXT_DEVICE ((qmckl_context_device)0)
#define QMCKL_NO_DEVICE (-1)

#define QMCKL_SUCCESS_DEVICE ((qmckl_exit_code_device)0)
#define QMCKL_INVALID_ARG_1_DEVICE ((qmckl_exit_code_device)1)
#define QMCKL_INVALID_ARG_2_DEVICE ((qmckl_exit_code_device)2)
#define QMCKL_INVALID_ARG_3_DEVICE ((qmckl_exit_code_device)3)
#define QMCKL_INVALID_ARG_4_DEVICE ((qmckl_exit_code_device)4)
#define QMCKL_INVALID_ARG_5_DEVICE ((qmckl_exit_code_device)5)
#define QMCKL_INVALID_ARG_6_DEVICE ((qmckl_exit_code_device)6)
#define QMCKL_INVALID_ARG_7_DEVICE ((qmckl_exit_code_device)7)
#define QMCKL_INVALID_ARG_8_DEVICE ((qmckl_exit_code_device)8)
#define QMCKL_INVALID_ARG_9_DEVICE ((qmckl_exit_code_device)9)
#define QMCKL_INVALID_ARG_10_DEVICE ((qmckl_exit_code_device)10)
#define QMCKL_INVALID_ARG_11_DEVICE ((qmckl_exit_code_device)11)
#define QMCKL_INVALID_ARG_12_DEVICE ((qmckl_exit_code_device)12)
#define QMCKL_INVALID_ARG_13_DEVICE ((qmckl_exit_code_device)13)
#define QMCKL_INVALID_ARG_14_DEVICE ((qmckl_exit_code_device)14)
#define QMCKL_INVALID_ARG_15_DEVICE ((qmckl_exit_code_device)15)
#define QMCKL_INVALID_ARG_16_DEVICE ((qmckl_exit_code_device)16)
#define QMCKL_INVALID_ARG_17_DEVICE ((qmckl_exit_code_device)17)
#define QMCKL_INVALID_ARG_18_DEVICE ((qmckl_exit_code_device)18)
#define QMCKL_INVALID_ARG_19_DEVICE ((qmckl_exit_code_device)19)
#define QMCKL_INVALID_ARG_20_DEVICE ((qmckl_exit_code_device)20)
#define QMCKL_FAILURE_DEVICE ((qmckl_exit_code_device)101)
#define QMCKL_ERRNO_DEVICE ((qmckl_exit_code_device)102)
#define QMCKL_INVALID_CONTEXT_DEVICE ((qmckl_exit_code_device)103)
#define QMCKL_ALLOCATION_FAILED_DEVICE ((qmckl_exit_code_device)104)
#define QMCKL_DEALLOCATION_FAILED_DEVICE ((qmckl_exit_code_device)105)
#define QMCKL_NOT_PROVIDED_DEVICE ((qmckl_exit_code_device)106)
#define QMCKL_OUT_OF_BOUNDS_DEVICE ((qmckl_exit_code_device)107)
#define QMCKL_ALREADY_SET_DEVICE ((qmckl_exit_code_device)108)
#define QMCKL_INVALID_EXIT_CODE_DEVICE ((qmckl_exit_code_device)109)

#define QMCKL_MAX_FUN_LEN_DEVICE 256
#define QMCKL_MAX_MSG_LEN_DEVICE 1024

/* include/qmckl_types.hpp:100-103 */
typedef struct qmckl_memory_info_struct_device {
  size_t size;
  void *pointer;
} qmckl_memory_info_struct_device;

/* ---- error (include/qmckl_gpu.hpp:384-387; src/qmckl_basic_functions.cpp:42-79) */

qmckl_exit_code_device qmckl_failwith_device(qmckl_context_device context,
                                             qmckl_exit_code_device exit_code,
                                             const char *function, const char *message);
/* src/qmckl_basic_functions.cpp:42-48 (returns a real description here). */
const char *qmckl_string_of_error_device(qmckl_exit_code_device error);
/* New: read back the last error recorded by qmckl_failwith_device.  Any
 * output pointer may be NULL; function_name must hold
 * QMCKL_MAX_FUN_LEN_DEVICE chars and message QMCKL_MAX_MSG_LEN_DEVICE. */
qmckl_exit_code_device qmckl_get_error_device(qmckl_context_device context,
                                              qmckl_exit_code_device *exit_code,
                                              char *function_name, char *message);

/* ---- context (include/qmckl_gpu.hpp:473-497) ----------------------------- */

/* qmckl_gpu.hpp:483 qmckl_context_create_device(sycl::queue) */
qmckl_context_device qmckl_context_create_device(int device_id);
/* qmckl_gpu.hpp:484-485 */
qmckl_exit_code_device qmckl_context_destroy_device(qmckl_context_device context);
/* qmckl_gpu.hpp:473-474 */
qmckl_exit_code_device qmckl_context_touch_device(qmckl_context_device context);
/* qmckl_gpu.hpp:493-494 */
qmckl_context_device qmckl_context_check_device(qmckl_context_device context);
/* qmckl_gpu.hpp:496-497 */
void qmckl_lock_device(qmckl_context_device context);
void qmckl_unlock_device(qmckl_context_device context);
/* qmckl_gpu.hpp:487-491 (static inline in the reference) */
int64_t qmckl_get_device_id(qmckl_context_device context);
/* qmckl_gpu.hpp:476-481 */
qmckl_exit_code_device qmckl_init_point_device(qmckl_context_device context);
qmckl_exit_code_device qmckl_init_ao_basis_device(qmckl_context_device context);
qmckl_exit_code_device qmckl_init_mo_basis_device(qmckl_context_device context);

/* ---- memory (include/qmckl_gpu.hpp:397-423) ---------------------------- */

/* qmckl_gpu.hpp:401-402 */
void *qmckl_malloc_host(qmckl_context_device context, qmckl_memory_info_struct_device info);
/* qmckl_gpu.hpp:404 */
void *qmckl_malloc_device(qmckl_context_device context, size_t size);
/* qmckl_gpu.hpp:406-410 */
qmckl_exit_code_device qmckl_free_host(qmckl_context_device context, void *ptr);
qmckl_exit_code_device qmckl_free_device(qmckl_context_device context, void *ptr);
/* qmckl_gpu.hpp:414-423 */
qmckl_exit_code_device qmckl_memcpy_H2D(qmckl_context_device context, void *dest,
                                        void *src, size_t size);
qmckl_exit_code_device qmckl_memcpy_D2H(qmckl_context_device context, void *dest,
                                        void *src, size_t size);
qmckl_exit_code_device qmckl_memcpy_D2D(qmckl_context_device context, void *dest,
                                        void *src, size_t size);

/* ---- point (include/qmckl_gpu.hpp:513-519) ------------------------------- */

qmckl_exit_code_device qmckl_set_point_device(qmckl_context_device context, char transp,
                                              int64_t num, double *coord, int64_t size_max);
qmckl_exit_code_device qmckl_get_point_device(qmckl_context_device context, char transp,
                                              double *coord, int64_t size_max);

/* ---- nucleus (include/qmckl_gpu.hpp:525-556) ----------------------------- */

bool qmckl_nucleus_provided_device(qmckl_context_device context);
qmckl_exit_code_device qmckl_get_nucleus_num_device(qmckl_context_device context, int64_t *num);
qmckl_exit_code_device qmckl_set_nucleus_num_device(qmckl_context_device context, int64_t num);
qmckl_exit_code_device qmckl_set_nucleus_charge_device(qmckl_context_device context,
                                                       double *charge, int64_t size_max);
qmckl_exit_code_device qmckl_set_nucleus_coord_device(qmckl_context_device context, char transp,
                                                      double *coord, int64_t size_max);
qmckl_exit_code_device qmckl_get_nucleus_coord_device(qmckl_context_device context, char transp,
                                                      double *coord, int64_t size_max);
qmckl_exit_code_device qmckl_get_nucleus_charge_device(qmckl_context_device context,
                                                       double *charge, int64_t size_max);

/* ---- electron (include/qmckl_gpu.hpp:562-580; include/qmckl_electron.hpp) */

bool qmckl_electron_provided_device(qmckl_context_device context);
qmckl_exit_code_device qmckl_set_electron_num_device(qmckl_context_device context,
                                                     int64_t up_num, int64_t down_num);
qmckl_exit_code_device qmckl_get_electron_num_device(qmckl_context_device context, int64_t *num);
qmckl_exit_code_device qmckl_set_electron_coord_device(qmckl_context_device context, char transp,
                                                       int64_t walk_num, double *coord,
                                                       int64_t size_max);
qmckl_exit_code_device qmckl_get_electron_coord_device(qmckl_context_device context, char transp,
                                                       double *coord, int64_t size_max);

/* ---- electron-nucleus distances ------------------------------------------ */

/* qmckl_gpu.hpp:1045-1046 */
qmckl_exit_code_device qmckl_provide_en_distance_device(qmckl_context_device context);
/* include/qmckl_electron.hpp:30-32 (src/qmckl_electron.cpp:302-331).
 * elec_coord [3][point_num], nucl_coord [3][nucl_num], en_distance
 * [point_num][nucl_num]. */
qmckl_exit_code_device qmckl_compute_en_distance_device(qmckl_context_device context,
                                                        int64_t point_num, int64_t nucl_num,
                                                        const double *elec_coord,
                                                        const double *nucl_coord,
                                                        double *en_distance);
/* include/qmckl_electron.hpp:37-40 (src/qmckl_electron.cpp:373-553).
 * C[i + j*ldc] = |a_i - b_j|, i < m, j < n, on device arrays; a_i is
 * (A[3i], A[3i+1], A[3i+2]) for transa 'N' (lda >= 3) and (A[i], A[i+lda],
 * A[i+2*lda]) for 'T' (lda >= m); b_j likewise from B, transb, n, ldb.
 * Codes: m <= 0 ARG_4, n <= 0 ARG_5, flag not N/n/T/t ARG_1, lda ARG_7,
 * ldb ARG_9, ldc < m ARG_11, NULL A/B/C ARG_6/8/10. */
qmckl_exit_code_device qmckl_distance_device(qmckl_context_device context, char transa, char transb,
                                             int64_t m, int64_t n, const double *A, int64_t lda,
                                             const double *B, int64_t ldb, double *C, int64_t ldc);
/* New getter (the reference has none): en_distance [P][N], size_max >= P*N. */
qmckl_exit_code_device qmckl_get_electron_en_distance_device(qmckl_context_device context,
                                                             double *en_distance,
                                                             int64_t size_max);

/* ---- electron-electron distances ------------------------------------------ */

/* qmckl_gpu.hpp:1043-1044 (src/qmckl_electron.cpp:192-240) */
qmckl_exit_code_device qmckl_provide_ee_distance_device(qmckl_context_device context);
/* include/qmckl_electron.hpp:34-35 (src/qmckl_electron.cpp:333-371).
 * coord [3][walk_num*elec_num] (point w*elec_num + i), ee_distance
 * [walk_num][elec_num][elec_num] with ee_distance[w][j][i] = |r_wi - r_wj|. */
qmckl_exit_code_device qmckl_compute_ee_distance_device(qmckl_context_device context,
                                                        int64_t elec_num, int64_t walk_num,
                                                        const double *coord, double *ee_distance);
/* New getter (the reference has none): ee_distance [walk][elec][elec],
 * size_max >= elec_num^2 * walk_num. */
qmckl_exit_code_device qmckl_get_electron_ee_distance_device(qmckl_context_device context,
                                                             double *ee_distance,
                                                             int64_t size_max);

/* ---- AO basis (include/qmckl_gpu.hpp:586-752) ---------------------------- */

bool qmckl_ao_basis_provided_device(qmckl_context_device context);

/* setters, qmckl_gpu.hpp:664-697 */
qmckl_exit_code_device qmckl_set_ao_basis_type_device(qmckl_context_device context, char basis_type);
qmckl_exit_code_device qmckl_set_ao_basis_shell_num_device(qmckl_context_device context, int64_t shell_num);
qmckl_exit_code_device qmckl_set_ao_basis_prim_num_device(qmckl_context_device context, int64_t prim_num);
qmckl_exit_code_device qmckl_set_ao_basis_ao_num_device(qmckl_context_device context, int64_t ao_num);
qmckl_exit_code_device qmckl_set_ao_basis_nucleus_index_device(qmckl_context_device context,
                                                               int64_t *nucleus_index, int64_t size_max);
qmckl_exit_code_device qmckl_set_ao_basis_nucleus_shell_num_device(qmckl_context_device context,
                                                                   int64_t *nucleus_shell_num,
                                                                   int64_t size_max);
qmckl_exit_code_device qmckl_set_ao_basis_shell_ang_mom_device(qmckl_context_device context,
                                                               int32_t *shell_ang_mom, int64_t size_max);
qmckl_exit_code_device qmckl_set_ao_basis_shell_prim_num_device(qmckl_context_device context,
                                                                int64_t *shell_prim_num, int64_t size_max);
qmckl_exit_code_device qmckl_set_ao_basis_shell_prim_index_device(qmckl_context_device context,
                                                                  int64_t *shell_prim_index,
                                                                  int64_t size_max);
qmckl_exit_code_device qmckl_set_ao_basis_shell_factor_device(qmckl_context_device context,
                                                              double *shell_factor, int64_t size_max);
qmckl_exit_code_device qmckl_set_ao_basis_exponent_device(qmckl_context_device context,
                                                          double *exponent, int64_t size_max);
qmckl_exit_code_device qmckl_set_ao_basis_coefficient_device(qmckl_context_device context,
                                                             double *coefficient, int64_t size_max);
qmckl_exit_code_device qmckl_set_ao_basis_prim_factor_device(qmckl_context_device context,
                                                             double *prim_factor, int64_t size_max);
qmckl_exit_code_device qmckl_set_ao_basis_ao_factor_device(qmckl_context_device context,
                                                           double *ao_factor, int64_t size_max);

/* getters, qmckl_gpu.hpp:641-652 and 702-746 */
qmckl_exit_code_device qmckl_get_ao_basis_type_device(qmckl_context_device context, char *basis_type);
qmckl_exit_code_device qmckl_get_ao_basis_shell_num_device(qmckl_context_device context, int64_t *shell_num);
qmckl_exit_code_device qmckl_get_ao_basis_prim_num_device(qmckl_context_device context, int64_t *prim_num);
qmckl_exit_code_device qmckl_get_ao_basis_ao_num_device(qmckl_context_device context, int64_t *ao_num);
qmckl_exit_code_device qmckl_get_ao_basis_nucleus_index_device(qmckl_context_device context,
                                                               int64_t *nucleus_index, int64_t size_max);
qmckl_exit_code_device qmckl_get_ao_basis_nucleus_shell_num_device(qmckl_context_device context,
                                                                   int64_t *nucleus_shell_num,
                                                                   int64_t size_max);
qmckl_exit_code_device qmckl_get_ao_basis_shell_ang_mom_device(qmckl_context_device context,
                                                               int32_t *shell_ang_mom, int64_t size_max);
qmckl_exit_code_device qmckl_get_ao_basis_shell_prim_num_device(qmckl_context_device context,
                                                                int64_t *shell_prim_num, int64_t size_max);
qmckl_exit_code_device qmckl_get_ao_basis_shell_prim_index_device(qmckl_context_device context,
                                                                  int64_t *shell_prim_index,
                                                                  int64_t size_max);
qmckl_exit_code_device qmckl_get_ao_basis_shell_factor_device(qmckl_context_device context,
                                                              double *shell_factor, int64_t size_max);
qmckl_exit_code_device qmckl_get_ao_basis_exponent_device(qmckl_context_device context,
                                                          double *exponent, int64_t size_max);
qmckl_exit_code_device qmckl_get_ao_basis_coefficient_device(qmckl_context_device context,
                                                             double *coefficient, int64_t size_max);
qmckl_exit_code_device qmckl_get_ao_basis_prim_factor_device(qmckl_context_device context,
                                                             double *prim_factor, int64_t size_max);
qmckl_exit_code_device qmckl_get_ao_basis_ao_factor_device(qmckl_context_device context,
                                                           double *ao_factor, int64_t size_max);

/* finalize, qmckl_gpu.hpp:658-659 (src/qmckl_ao_sycl.cpp:1067-1236) */
qmckl_exit_code_device qmckl_finalize_ao_basis_device(qmckl_context_device context);

/* providers and getters of the computed arrays, qmckl_gpu.hpp:620-639, 748-750 */
qmckl_exit_code_device qmckl_provide_ao_basis_shell_vgl_device(qmckl_context_device context);
qmckl_exit_code_device qmckl_provide_ao_basis_ao_vgl_device(qmckl_context_device context);
qmckl_exit_code_device qmckl_provide_ao_basis_ao_value_device(qmckl_context_device context);
/* src/qmckl_ao.cpp:207-239: shell_vgl [P][5][S] */
qmckl_exit_code_device qmckl_get_ao_basis_shell_vgl_device(qmckl_context_device context,
                                                           double *shell_vgl, int64_t size_max);
/* ao_vgl [P][5][A], size_max >= 5*A*P */
qmckl_exit_code_device qmckl_get_ao_basis_ao_vgl_device(qmckl_context_device context,
                                                        double *ao_vgl, int64_t size_max);
qmckl_exit_code_device qmckl_get_ao_basis_ao_vgl_inplace_device(qmckl_context_device context,
                                                                double *ao_vgl, int64_t size_max);
/* ao_value [P][A], size_max >= A*P */
qmckl_exit_code_device qmckl_get_ao_basis_ao_value_device(qmckl_context_device context,
                                                          double *ao_value, int64_t size_max);
qmckl_exit_code_device qmckl_get_ao_basis_ao_value_inplace_device(qmckl_context_device context,
                                                                  double *ao_value, int64_t size_max);

/* low-level launchers, qmckl_gpu.hpp:586-613 (src/qmckl_ao_sycl.cpp:12-102,
 * 106-434).  coord [3][point_num], nucl_coord [3][nucl_num];
 * coef_normalized and nucleus_range as produced by the finalize step.
 * compute_ao_vgl_gaussian recomputes the radial parts itself and ignores
 * the shell_vgl argument's contents (it may be NULL). */
qmckl_exit_code_device qmckl_compute_ao_basis_shell_gaussian_vgl_device(
    qmckl_context_device context, int prim_num, int shell_num, int point_num, int nucl_num,
    int64_t *nucleus_shell_num, int64_t *nucleus_index, double *nucleus_range,
    int64_t *shell_prim_index, int64_t *shell_prim_num, double *coord, double *nucl_coord,
    double *expo, double *coef_normalized, double *shell_vgl);
qmckl_exit_code_device qmckl_compute_ao_vgl_gaussian_device(
    qmckl_context_device context, int64_t ao_num, int64_t shell_num, int64_t point_num,
    int64_t nucl_num, const double *coord, const double *nucl_coord,
    const int64_t *nucleus_index, const int64_t *nucleus_shell_num,
    const double *nucleus_range, const int32_t *nucleus_max_ang_mom,
    const int32_t *shell_ang_mom, const double *ao_factor, double *shell_vgl, double *ao_vgl);
/* Allocation-free form of qmckl_compute_ao_basis_shell_gaussian_vgl_device
 * (same arguments, same results).  It has no counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture,
 * serving loops).  It builds its shell table on the device in a fixed
 * per-GPU device array instead of allocating, so it calls no hipMalloc or
 * hipFree.  Limit: shell_num <= 65536 (ARG_3 beyond).  Every call on one GPU
 * shares that array: calls that overlap on one GPU (other threads, other
 * contexts) race and must be serialised by the caller.  Invalid index arrays
 * return QMCKL_FAILURE_DEVICE before the radial kernel runs. */
qmckl_exit_code_device qmckl_compute_ao_basis_shell_gaussian_vgl_noalloc_device(
    qmckl_context_device context, int prim_num, int shell_num, int point_num, int nucl_num,
    const int64_t *nucleus_shell_num, const int64_t *nucleus_index, const double *nucleus_range,
    const int64_t *shell_prim_index, const int64_t *shell_prim_num, const double *coord,
    const double *nucl_coord, const double *expo, const double *coef_normalized, double *shell_vgl);

/* ---- MO basis (include/qmckl_gpu.hpp:758-801) ---------------------------- */

bool qmckl_mo_basis_provided_device(qmckl_context_device context);
qmckl_exit_code_device qmckl_get_mo_basis_mo_num_device(qmckl_context_device context, int64_t *mo_num);
qmckl_exit_code_device qmckl_set_mo_basis_mo_num_device(qmckl_context_device context, int64_t mo_num);
/* coefficient [mo_num][ao_num]; the reference takes no size (qmckl_mo.cpp:526-583) */
qmckl_exit_code_device qmckl_set_mo_basis_coefficient_device(qmckl_context_device context,
                                                             double *coefficient);
qmckl_exit_code_device qmckl_finalize_mo_basis_device(qmckl_context_device context);
qmckl_exit_code_device qmckl_provide_mo_basis_mo_vgl_device(qmckl_context_device context);
qmckl_exit_code_device qmckl_provide_mo_basis_mo_value_device(qmckl_context_device context);
/* src/qmckl_mo.cpp:31-106 (defined there, declared in no header; its
 * result type is bool there although it returns exit codes).  Keeps the MOs
 * i with keep[i] != 0, in order; keep is a HOST array of size_max >= mo_num
 * entries.  Selecting no MO is rejected (QMCKL_INVALID_ARG_2_DEVICE). */
qmckl_exit_code_device qmckl_mo_basis_select_mo_device(qmckl_context_device context, int32_t *keep,
                                                       int64_t size_max);
/* mo_vgl [P][5][M], size_max >= 5*M*P */
qmckl_exit_code_device qmckl_get_mo_basis_mo_vgl_device(qmckl_context_device context,
                                                        double *mo_vgl, int64_t size_max);
qmckl_exit_code_device qmckl_get_mo_basis_mo_vgl_inplace_device(qmckl_context_device context,
                                                                double *mo_vgl, int64_t size_max);
/* mo_value [P][M], size_max >= M*P.  Always the FP64 contraction of
 * ao_value, whatever the precision form (in the FP64 form a current mo_vgl's
 * value rows are copied), so it does not depend on the order of calls. */
qmckl_exit_code_device qmckl_get_mo_basis_mo_value_device(qmckl_context_device context,
                                                          double *mo_value, int64_t size_max);
qmckl_exit_code_device qmckl_get_mo_basis_mo_value_inplace_device(qmckl_context_device context,
                                                                  double *mo_value, int64_t size_max);
/* qmckl_gpu.hpp:790-793 (src/qmckl_mo_sycl.cpp:11-56): coefficient_t
 * [ao_num][mo_num], ao_vgl [P][5][A] -> mo_vgl [P][5][M], FP64. */
qmckl_exit_code_device qmckl_compute_mo_basis_mo_vgl_device(qmckl_context_device context,
                                                            int64_t ao_num, int64_t mo_num,
                                                            int64_t point_num, double *coefficient_t,
                                                            double *ao_vgl, double *mo_vgl);
/* qmckl_gpu.hpp:785-788 (src/qmckl_mo_sycl.cpp:60-134) */
qmckl_exit_code_device qmckl_compute_mo_basis_mo_value_device(qmckl_context_device context,
                                                              int64_t ao_num, int64_t mo_num,
                                                              int64_t point_num, double *coefficient_t,
                                                              double *ao_value, double *mo_value);

/* ---- precision form of the MO contraction (new; no counterpart) -----------
 * FP64 (default) is the plain form.  The others keep FP64 for distances,
 * exponent arguments and the AO assembly, round the AO operand to FP32, and
 * run the contraction on the matrix cores:
 *   FP32    FP32 MFMA;
 *   FP16X2  power-of-two scaled fp16 hi+lo split of both operands, products
 *           hi*hi + hi*lo + lo*hi;
 *   BF16X2  bf16 hi+lo split, the same three products;
 *   BF16X3  bf16 hi+mid+lo split, six products;
 * all accumulating in FP32 over short K blocks and in FP64 across them;
 *   INT8X5  the accurate form: the AO operand kept as FP32 value plus FP32
 *           residual, both operands cut into five int8 fixed-point slices
 *           (per (AO row, 64-AO block) and per MO power-of-two scales), the
 *           15 slice products of order <= 4 on the int8 matrix cores with
 *           exact int32 sums, folded into FP64 per 64-AO block: MO error
 *           about 1e-9 of each row's maximum on well-conditioned bases and
 *           below 1e-7 on the badly conditioned ChBrClF basis, where the
 *           FP32-accumulating forms stop at a few 1e-6.
 * Results stay FP64 arrays [P][5][M]; changing the form invalidates mo_vgl. */
typedef int32_t qmckl_precision_device;
#define QMCKL_PRECISION_FP64_DEVICE ((qmckl_precision_device)0)
#define QMCKL_PRECISION_FP32_DEVICE ((qmckl_precision_device)1)
#define QMCKL_PRECISION_FP16X2_DEVICE ((qmckl_precision_device)2)
#define QMCKL_PRECISION_BF16X2_DEVICE ((qmckl_precision_device)3)
#define QMCKL_PRECISION_BF16X3_DEVICE ((qmckl_precision_device)4)
#define QMCKL_PRECISION_INT8X5_DEVICE ((qmckl_precision_device)5)
qmckl_exit_code_device qmckl_set_mo_basis_precision_device(qmckl_context_device context,
                                                           qmckl_precision_device form);
qmckl_exit_code_device qmckl_get_mo_basis_precision_device(qmckl_context_device context,
                                                           qmckl_precision_device *form);

/* ---- differentiation form (new; no counterpart in the reference) ----------
 * Adjoint of L = sum_{p,c,i} weight[p][c][i] * mo_vgl[p][c][i] (FP64) for the
 * current points: d_nucl_coord [N][3] (dL/dR), d_exponent [prim_num]
 * (dL/dalpha at fixed coefficient, prim_factor and shell_factor),
 * d_mo_coefficient [mo_num][ao_num] (the layout of
 * qmckl_set_mo_basis_coefficient_device).  weight is a device array
 * [P][5][mo_num] (size_max >= 5*P*mo_num); any output may be NULL.  AO
 * primitives are recomputed, not stored; reductions over points run in a
 * fixed order (bitwise reproducible). */
qmckl_exit_code_device qmckl_get_mo_basis_mo_vgl_adjoint_device(qmckl_context_device context,
                                                                const double *weight,
                                                                int64_t size_max,
                                                                double *d_nucl_coord,
                                                                double *d_exponent,
                                                                double *d_mo_coefficient);

/* ---- batching form (new; no counterpart in the reference) -----------------
 * Evaluates mo_vgl for count contexts (one molecule each, on one device) in
 * one submission: for each i it does what
 * qmckl_get_mo_basis_mo_vgl_inplace_device(contexts[i], mo_vgl[i],
 * size_max[i]) does (touch, then distances, AOs and MOs in the context's
 * precision form, written to mo_vgl[i] [P_i][5][M_i]), but the work of every
 * context is enqueued on that context's own stream before any of them is
 * waited for, so molecules of different sizes run concurrently on the GPU.
 * The MO contractions of contexts in a split precision form (up to 192 MOs)
 * run as persistent launches over all of them (an atomic counter hands out
 * (molecule, row tile, MO tile) items), one per form and tile width, on
 * contexts[0]'s stream after every context's AO work.
 * Returns when every output is complete.  contexts, mo_vgl and size_max are
 * host arrays of count entries (mo_vgl[i] a device pointer).  Errors: count
 * <= 0 -> ARG_1; a NULL array -> ARG_2/3/4; an invalid context ->
 * INVALID_CONTEXT; contexts on different devices -> ARG_2; otherwise each
 * context's own code (recorded in that context), the first one returned. */
qmckl_exit_code_device qmckl_get_mo_basis_mo_vgl_batch_device(int64_t count,
                                                              const qmckl_context_device *contexts,
                                                              double *const *mo_vgl,
                                                              const int64_t *size_max);

/* ---- tile choices and autotuner (new; no counterpart in the reference) -----
 * The mixed precision forms have two tile choices that change speed, not
 * results: whether points are ordered spatially into 64-point tiles
 * (point_order -1 automatic, 0 off, 1 on; it also applies to the FP64
 * ao_vgl kernel used for bases of 1024 AOs or more and to the adjoint's
 * point visits) and how many wavefronts
 * per CU the AO operand kernel aims at (group_waves 0 automatic, or 1..1024).
 * set: ARG_2 / ARG_3 for values out of range.  get: ARG_2 / ARG_3 for NULL.
 * autotune: times the context's mixed chain on its current points for
 * point_order 0/1 x group_waves 16/32/64/128/256, keeps the fastest, then
 * times the contraction kernels of qmckl_set_mo_basis_gemm_tile_device
 * below with it and keeps the fastest (a process-wide cache reuses the
 * choices for contexts of the same device, point, AO, MO and nucleus counts
 * and form); FP64 form: nothing to choose.
 * Needs the basis and points (NOT_PROVIDED otherwise).  It overwrites the
 * context's cached mo_vgl, which is recomputed on the next get. */
qmckl_exit_code_device qmckl_set_mo_basis_tile_config_device(qmckl_context_device context, int32_t point_order,
                                                             int32_t group_waves);
qmckl_exit_code_device qmckl_get_mo_basis_tile_config_device(qmckl_context_device context, int32_t *point_order,
                                                             int32_t *group_waves);
qmckl_exit_code_device qmckl_mo_basis_autotune_device(qmckl_context_device context);
/* A third tile choice of the split forms (FP16X2, BF16X2, BF16X3): the
 * contraction kernel.  LIVE (1): blocks of zero coefficients or screened
 * AOs are neither loaded nor multiplied, the AO operand is streamed with
 * nontemporal loads; LIVE_CACHED (2): the same with cached AO loads; DENSE
 * (3): every block of a live 64-AO stage is multiplied, dead ones on zeros;
 * LIVE64 (4): LIVE with 64 rows per wavefront (one wavefront per SIMD);
 * DENSE64 (6): DENSE with workgroups of a whole 64-row tile (eight
 * wavefronts).  These five contract the stored FP32 AO operand and give the
 * same results bit for bit.  FUSED (5): one kernel evaluates the AOs of 64 points into
 * local memory and contracts them with 64 MOs (each group of 64 MOs
 * evaluates the AOs again); it allocates no AO operand (about 20 bytes per
 * point and AO) and its results round
 * differently, within the same error bounds (operand rows scaled per 64-AO
 * stage instead of once).  It is slower than the stored-operand kernels
 * (on MI355X, a 605-AO, 120-MO molecule at 91,537 points: 1.41 ms against
 * 0.74), so AUTO never takes it; it applies to MOs in their given order (no
 * column plan) and at most 8,192 AOs, and is taken as AUTO elsewhere.
 * AUTO (0): LIVE where blocks can be dead (the column plan is in use, or
 * 1024 AOs or more) and DENSE otherwise.  The FP32 and INT8X5 forms have
 * one kernel each and ignore the choice.  set: ARG_2 for an unknown value;
 * get: ARG_2 for NULL.  The autotuner above also times LIVE, LIVE_CACHED,
 * DENSE, LIVE64 and DENSE64 (after the point order and AO wavefronts) and
 * keeps the fastest; a context set to FUSED keeps FUSED. */
#define QMCKL_GEMM_TILE_AUTO_DEVICE ((int32_t)0)
#define QMCKL_GEMM_TILE_LIVE_DEVICE ((int32_t)1)
#define QMCKL_GEMM_TILE_LIVE_CACHED_DEVICE ((int32_t)2)
#define QMCKL_GEMM_TILE_DENSE_DEVICE ((int32_t)3)
#define QMCKL_GEMM_TILE_LIVE64_DEVICE ((int32_t)4)
#define QMCKL_GEMM_TILE_FUSED_DEVICE ((int32_t)5)
#define QMCKL_GEMM_TILE_DENSE64_DEVICE ((int32_t)6)
qmckl_exit_code_device qmckl_set_mo_basis_gemm_tile_device(qmckl_context_device context, int32_t gemm_tile);
qmckl_exit_code_device qmckl_get_mo_basis_gemm_tile_device(qmckl_context_device context, int32_t *gemm_tile);

/* ---- MO column plan (new; no counterpart in the reference) ----------------
 * At MO finalize the library maps the coefficients' zeros in blocks of 16
 * AOs and orders the MO columns of its prepared coefficient copies in groups
 * of 16 that share zero blocks; the contraction kernels (FP64 and the split
 * forms) skip those blocks and store every column at its MO's place, so the
 * order changes no result.  The plan is kept when at most 90% of the
 * (16-MO, 16-AO) blocks stay live, else the MO order is used as given.
 * col_mo (may be NULL): receives mo_num entries, the MO of each prepared
 * column (0, 1, ... without a plan); live (may be NULL): the fraction of
 * blocks left live by the planned order (kept or not).
 * NOT_PROVIDED before the MO basis is finalized. */
qmckl_exit_code_device qmckl_get_mo_basis_column_plan_device(qmckl_context_device context, int32_t *col_mo,
                                                             double *live);

/* ---- batching form as a HIP graph (new; no counterpart in the reference) --
 * create: runs the batch above once (same arguments and errors), then
 * records the same work of all count contexts (every kernel of every
 * context's chain, each context on its own stream, forked from and joined
 * into one capture stream) into a HIP graph and returns its handle in
 * *graph (ARG_5 if graph is NULL).
 * launch: replays the graph and returns when every output is complete.  It
 * reads the contexts' current point coordinates, so a caller moves the
 * points with the usual setters (same point counts) between launches.  A
 * change that alters what the recorded kernels take as arguments (point or
 * basis sizes, precision form, MO coefficients, a reallocated buffer, a
 * destroyed context) makes launch fail (QMCKL_FAILURE_DEVICE, or
 * INVALID_CONTEXT) instead of running stale work: create a new graph.
 * An unknown handle -> ARG_1.
 * destroy: releases the graph (ARG_1 for an unknown handle). */
qmckl_exit_code_device qmckl_mo_basis_mo_vgl_batch_graph_create_device(int64_t count,
                                                                       const qmckl_context_device *contexts,
                                                                       double *const *mo_vgl,
                                                                       const int64_t *size_max,
                                                                       int64_t *graph);
qmckl_exit_code_device qmckl_mo_basis_mo_vgl_batch_graph_launch_device(int64_t graph);
qmckl_exit_code_device qmckl_mo_basis_mo_vgl_batch_graph_destroy_device(int64_t graph);

/* ---- kernel timing (new; no counterpart in the reference) ----------------
 * When enabled, every kernel the context launches is bracketed by HIP events
 * on the context's stream and its elapsed time is accumulated under the
 * kernel's name ("en_distance", "ao_vgl", "ao_value", "shell_vgl",
 * "mo_gemm_f64", ...).  For benchmarks and profiling; off by default. */
qmckl_exit_code_device qmckl_set_kernel_timing_device(qmckl_context_device context, bool enable);
qmckl_exit_code_device qmckl_reset_kernel_timing_device(qmckl_context_device context);
qmckl_exit_code_device qmckl_get_kernel_timing_device(qmckl_context_device context,
                                                      const char *kernel, int64_t *launches,
                                                      double *total_ms);

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif

#ifdef __cplusplus
}
#endif
#endif /* QMCKL_GPU_H */
