// Points, nuclei, electrons and electron-nucleus distances.
//
// Semantics follow src/qmckl_point.cpp (set 68-135, get 5-66),
// src/qmckl_nucleus.cpp (setters/getters 359-664), src/qmckl_electron.cpp
// (set_electron_num 8-48, set_electron_coord 50-90, get_electron_coord
// 127-170, provide_en_distance 242-296, compute_en_distance 302-331).
// Coordinates are stored structure-of-arrays ([3][P] and [3][N]).  The
// en_distance array is reallocated only when it must grow, and the result is
// the plain Euclidean distance [P][N].
#include "internal.hpp"

using namespace qmckl_amd;

namespace {

// Store `num` points given in layout `transp` ('N': [num][3], 'T': [3][num])
// into ctx->point, reallocating if the size changes, then touch.
qmckl_exit_code_device store_points(qmckl_context_device context, Context *c, char transp,
                                    int64_t num, const double *coord, const char *fn) {
  if (!has_device(c)) return fail(context, QMCKL_FAILURE_DEVICE, fn, "no device bound");
  if (num > c->point.capacity || c->point.coord == nullptr) {
    if (!drealloc(c, reinterpret_cast<void **>(&c->point.coord),
                  sizeof(double) * 3 * (size_t)(num > 0 ? num : 1)))
      return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
    c->point.capacity = num;
  }
  c->point.num = num;
  hipError_t e = hipSuccess;
  if (num > 0) {
    if (transp == 'T')
      e = hipMemcpyAsync(c->point.coord, coord, sizeof(double) * 3 * num,
                         hipMemcpyDeviceToDevice, c->stream);
    else
      e = launch_soa_from_aos(c->stream, num, coord, c->point.coord);
    if (e == hipSuccess) e = sync(c);
  }
  if (e != hipSuccess) return hip_check(context, e, fn);
  return qmckl_context_touch_device(context);
}

}  // namespace

extern "C" {

qmckl_exit_code_device qmckl_set_point_device(qmckl_context_device context, char transp,
                                              int64_t num, double *coord, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_set_point_device";
  if (size_max < 3 * num) return fail(context, QMCKL_INVALID_ARG_4_DEVICE, fn, "Array too small");
  if (transp != 'N' && transp != 'T')
    return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "transp should be 'N' or 'T'");
  if (coord == nullptr) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "coord is a NULL pointer");
  if (num < 0) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "num < 0");
  return store_points(context, c, transp, num, coord, fn);
}

qmckl_exit_code_device qmckl_get_point_device(qmckl_context_device context, char transp,
                                              double *coord, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_get_point_device";
  if (coord == nullptr) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "coord is a null pointer");
  if (transp != 'N' && transp != 'T')
    return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "transp should be 'N' or 'T'");
  int64_t num = c->point.num;
  if (num == 0 || c->point.coord == nullptr) return QMCKL_NOT_PROVIDED_DEVICE;
  if (size_max < 3 * num) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "size_max too small");
  hipError_t e;
  if (transp == 'N')
    e = launch_aos_from_soa(c->stream, num, c->point.coord, coord);
  else
    e = hipMemcpyAsync(coord, c->point.coord, sizeof(double) * 3 * num, hipMemcpyDeviceToDevice,
                       c->stream);
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

// ---- nucleus ----

bool qmckl_nucleus_provided_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  return c ? c->nucleus.provided : false;
}

qmckl_exit_code_device qmckl_get_nucleus_num_device(qmckl_context_device context, int64_t *num) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  if (num == nullptr)
    return fail(context, QMCKL_INVALID_ARG_2_DEVICE, "qmckl_get_nucleus_num", "num is a null pointer");
  if (c->nucleus.uninitialized & 1) {
    *num = 0;
    return fail(context, QMCKL_NOT_PROVIDED_DEVICE, "qmckl_get_nucleus_num",
                "nucleus data is not provided");
  }
  *num = c->nucleus.num;
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_set_nucleus_num_device(qmckl_context_device context, int64_t num) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const int32_t mask = 1 << 0;
  if (!(c->nucleus.uninitialized & mask))
    return fail(context, QMCKL_ALREADY_SET_DEVICE, "qmckl_set_nucleus_*", nullptr);
  if (num <= 0) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, "qmckl_set_nucleus_num", "num <= 0");
  c->nucleus.num = num;
  c->nucleus.uninitialized &= ~mask;
  c->nucleus.provided = (c->nucleus.uninitialized == 0);
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_set_nucleus_charge_device(qmckl_context_device context,
                                                       double *charge, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_set_nucleus_charge_device";
  const int32_t mask = 1 << 1;
  if (!(c->nucleus.uninitialized & mask))
    return fail(context, QMCKL_ALREADY_SET_DEVICE, fn, nullptr);
  if (charge == nullptr) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "charge is a null pointer");
  int64_t num = 0;
  qmckl_exit_code_device rc = qmckl_get_nucleus_num_device(context, &num);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  if (num > size_max) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "Array too small");
  if (!drealloc(c, reinterpret_cast<void **>(&c->nucleus.charge), sizeof(double) * num))
    return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
  hipError_t e = hipMemcpyAsync(c->nucleus.charge, charge, sizeof(double) * num,
                                hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = sync(c);
  if (e != hipSuccess) return hip_check(context, e, fn);
  c->nucleus.uninitialized &= ~mask;
  c->nucleus.provided = (c->nucleus.uninitialized == 0);
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_set_nucleus_coord_device(qmckl_context_device context, char transp,
                                                      double *coord, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_set_nucleus_coord_device";
  const int32_t mask = 1 << 2;
  if (!(c->nucleus.uninitialized & mask))
    return fail(context, QMCKL_ALREADY_SET_DEVICE, "qmckl_set_nucleus_*", nullptr);
  if (transp != 'N' && transp != 'T')
    return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "transp should be 'N' or 'T'");
  if (coord == nullptr) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "coord is a null pointer");
  const int64_t n = c->nucleus.num;
  if (n <= 0) return fail(context, QMCKL_FAILURE_DEVICE, fn, "nucl_num is not set");
  if (size_max < 3 * n) return fail(context, QMCKL_INVALID_ARG_4_DEVICE, fn, "Array too small");
  if (!drealloc(c, reinterpret_cast<void **>(&c->nucleus.coord), sizeof(double) * 3 * n))
    return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
  hipError_t e;
  if (transp == 'T')
    e = hipMemcpyAsync(c->nucleus.coord, coord, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice,
                       c->stream);
  else
    e = launch_soa_from_aos(c->stream, n, coord, c->nucleus.coord);
  if (e == hipSuccess) e = sync(c);
  if (e != hipSuccess) return hip_check(context, e, fn);
  c->nucleus.uninitialized &= ~mask;
  c->nucleus.provided = (c->nucleus.uninitialized == 0);
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_get_nucleus_coord_device(qmckl_context_device context, char transp,
                                                      double *coord, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_get_nucleus_coord_device";
  if (transp != 'N' && transp != 'T')
    return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "transp should be 'N' or 'T'");
  if (coord == nullptr) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "coord is a null pointer");
  if (c->nucleus.uninitialized & (1 << 2))
    return fail(context, QMCKL_NOT_PROVIDED_DEVICE, fn, "nucleus data is not provided");
  const int64_t n = c->nucleus.num;
  if (size_max < 3 * n) return fail(context, QMCKL_INVALID_ARG_4_DEVICE, fn, "Array too small");
  hipError_t e;
  if (transp == 'N')
    e = launch_aos_from_soa(c->stream, n, c->nucleus.coord, coord);
  else
    e = hipMemcpyAsync(coord, c->nucleus.coord, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice,
                       c->stream);
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

qmckl_exit_code_device qmckl_get_nucleus_charge_device(qmckl_context_device context,
                                                       double *charge, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_get_nucleus_charge_device";
  if (charge == nullptr) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "charge is a null pointer");
  if (c->nucleus.uninitialized & (1 << 1))
    return fail(context, QMCKL_NOT_PROVIDED_DEVICE, fn, "nucleus data is not provided");
  if (c->nucleus.num > size_max) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "Array too small");
  hipError_t e = hipMemcpyAsync(charge, c->nucleus.charge, sizeof(double) * c->nucleus.num,
                                hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

// ---- electrons ----

bool qmckl_electron_provided_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  return c ? c->electron.provided : false;
}

qmckl_exit_code_device qmckl_set_electron_num_device(qmckl_context_device context,
                                                     int64_t up_num, int64_t down_num) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const int32_t mask = 1 << 0;
  if (!(c->electron.uninitialized & mask))
    return fail(context, QMCKL_ALREADY_SET_DEVICE, "qmckl_set_electron_*", nullptr);
  if (up_num <= 0)
    return fail(context, QMCKL_INVALID_ARG_2_DEVICE, "qmckl_set_electron_num_device", "up_num <= 0");
  if (down_num < 0)
    return fail(context, QMCKL_INVALID_ARG_3_DEVICE, "qmckl_set_electron_num_device", "down_num < 0");
  c->electron.up_num = up_num;
  c->electron.down_num = down_num;
  c->electron.num = up_num + down_num;
  c->electron.uninitialized &= ~mask;
  c->electron.provided = (c->electron.uninitialized == 0);
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_get_electron_num_device(qmckl_context_device context, int64_t *num) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  if (num == nullptr)
    return fail(context, QMCKL_INVALID_ARG_2_DEVICE, "qmckl_get_electron_num", "num is a null pointer");
  if (c->electron.uninitialized & 1) return QMCKL_NOT_PROVIDED_DEVICE;
  *num = c->electron.num;
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_set_electron_coord_device(qmckl_context_device context, char transp,
                                                       int64_t walk_num, double *coord,
                                                       int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_set_electron_coord_device";
  const int64_t elec_num = c->electron.num;
  if (elec_num == 0) return fail(context, QMCKL_FAILURE_DEVICE, fn, "elec_num is not set");
  if (walk_num <= 0) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "walk_num <= 0");
  // Remaining checks and their codes are those of qmckl_set_point_device,
  // which the reference delegates to (src/qmckl_electron.cpp:82,
  // src/qmckl_point.cpp:78-92).
  const int64_t num = walk_num * elec_num;
  if (size_max < 3 * num) return fail(context, QMCKL_INVALID_ARG_4_DEVICE, fn, "Array too small");
  if (transp != 'N' && transp != 'T')
    return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "transp should be 'N' or 'T'");
  if (coord == nullptr) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "coord is a NULL pointer");
  qmckl_exit_code_device rc = store_points(context, c, transp, num, coord, fn);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  c->electron.walker_old = c->electron.walker;
  c->electron.walker.num = walk_num;
  c->electron.walker.point = c->point;
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_get_electron_coord_device(qmckl_context_device context, char transp,
                                                       double *coord, int64_t size_max) {
  const char *fn = "qmckl_get_electron_coord_device";
  if (transp != 'N' && transp != 'T')
    return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "transp should be 'N' or 'T'");
  if (coord == nullptr) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "coord is a null pointer");
  if (size_max <= 0) return fail(context, QMCKL_INVALID_ARG_4_DEVICE, fn, "size_max should be > 0");
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  if (!c->electron.provided) return fail(context, QMCKL_NOT_PROVIDED_DEVICE, fn, nullptr);
  return qmckl_get_point_device(context, transp, coord, size_max);
}

// ---- electron-nucleus distances ----

qmckl_exit_code_device qmckl_compute_en_distance_device(qmckl_context_device context,
                                                        int64_t point_num, int64_t nucl_num,
                                                        const double *elec_coord,
                                                        const double *nucl_coord,
                                                        double *en_distance) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_compute_en_distance_device";
  if (point_num <= 0) return QMCKL_INVALID_ARG_2_DEVICE;
  if (nucl_num <= 0) return QMCKL_INVALID_ARG_3_DEVICE;
  if (!elec_coord) return fail(context, QMCKL_INVALID_ARG_4_DEVICE, fn, "NULL pointer");
  if (!nucl_coord) return fail(context, QMCKL_INVALID_ARG_5_DEVICE, fn, "NULL pointer");
  if (!en_distance) return fail(context, QMCKL_INVALID_ARG_6_DEVICE, fn, "NULL pointer");
  if (!has_device(c)) return fail(context, QMCKL_FAILURE_DEVICE, fn, "no device bound");
  hipError_t e;
  {
    KernelTimer t(c, "en_distance");
    e = launch_en_distance(c->stream, point_num, nucl_num, elec_coord, nucl_coord, en_distance);
    t.stop();
  }
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

// General distance launcher: C[i + j ldc] = |a_i - b_j|, a_i the i-th point of
// A ('N': A[3i..3i+2], lda >= 3; 'T': A[i], A[i+lda], A[i+2lda], lda >= m), b_j
// likewise from B with n and ldb (src/qmckl_electron.cpp:373-553).  Exit codes
// as the reference's: m <= 0 ARG_4, n <= 0 ARG_5, a transposition flag other
// than N/n/T/t ARG_1, lda ARG_7, ldb ARG_9, ldc < m ARG_11.  Each leading
// dimension is checked against its own operand's flag; the reference's checks
// at 430-470 also test lda against transb and ldb against transa, so every
// call it accepts is accepted here.  Its kernel takes the square root of a
// whole column from each work item while others still write it; the result
// here is the distance that kernel is meant to produce (and what its callers,
// en_distance at 326 and ee_distance at 361, use).
qmckl_exit_code_device qmckl_distance_device(qmckl_context_device context, char transa, char transb, int64_t m,
                                             int64_t n, const double *A, int64_t lda, const double *B, int64_t ldb,
                                             double *C, int64_t ldc) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_distance_device";
  if (m <= 0) return QMCKL_INVALID_ARG_4_DEVICE;
  if (n <= 0) return QMCKL_INVALID_ARG_5_DEVICE;
  const bool na = transa == 'N' || transa == 'n', ta = transa == 'T' || transa == 't';
  const bool nb = transb == 'N' || transb == 'n', tb = transb == 'T' || transb == 't';
  if (!(na || ta) || !(nb || tb)) return QMCKL_INVALID_ARG_1_DEVICE;
  if (lda < (ta ? m : 3)) return QMCKL_INVALID_ARG_7_DEVICE;
  if (ldb < (tb ? n : 3)) return QMCKL_INVALID_ARG_9_DEVICE;
  if (ldc < m) return QMCKL_INVALID_ARG_11_DEVICE;
  if (!A) return fail(context, QMCKL_INVALID_ARG_6_DEVICE, fn, "NULL pointer");
  if (!B) return fail(context, QMCKL_INVALID_ARG_8_DEVICE, fn, "NULL pointer");
  if (!C) return fail(context, QMCKL_INVALID_ARG_10_DEVICE, fn, "NULL pointer");
  if (!has_device(c)) return fail(context, QMCKL_FAILURE_DEVICE, fn, "no device bound");
  hipError_t e = launch_distance(c->stream, ta, tb, m, n, A, lda, B, ldb, C, ldc);
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

qmckl_exit_code_device qmckl_provide_en_distance_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_provide_en_distance_device";
  if (!c->nucleus.provided) return fail(context, QMCKL_NOT_PROVIDED_DEVICE, fn, nullptr);
  if (c->point.num <= 0 || c->point.coord == nullptr)
    return fail(context, QMCKL_NOT_PROVIDED_DEVICE, fn, "no points set");
  if (c->point.date > c->electron.en_distance_date || c->electron.en_distance == nullptr) {
    const int64_t need = c->point.num * c->nucleus.num;
    if (need > c->electron.en_distance_cap) {
      if (!drealloc(c, reinterpret_cast<void **>(&c->electron.en_distance), sizeof(double) * need))
        return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
      c->electron.en_distance_cap = need;
    }
    qmckl_exit_code_device rc = qmckl_compute_en_distance_device(
        context, c->point.num, c->nucleus.num, c->point.coord, c->nucleus.coord,
        c->electron.en_distance);
    if (rc != QMCKL_SUCCESS_DEVICE) return rc;
    c->electron.en_distance_date = c->date;
  }
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_get_electron_en_distance_device(qmckl_context_device context,
                                                             double *en_distance,
                                                             int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_get_electron_en_distance_device";
  if (en_distance == nullptr) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "NULL pointer");
  qmckl_exit_code_device rc = qmckl_provide_en_distance_device(context);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  const int64_t sze = c->point.num * c->nucleus.num;
  if (size_max < sze) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "input array too small");
  hipError_t e = hipMemcpyAsync(en_distance, c->electron.en_distance, sizeof(double) * sze,
                                hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

// ---- electron-electron distances ----

qmckl_exit_code_device qmckl_compute_ee_distance_device(qmckl_context_device context, int64_t elec_num,
                                                        int64_t walk_num, const double *coord,
                                                        double *ee_distance) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_compute_ee_distance_device";
  if (elec_num <= 0) return QMCKL_INVALID_ARG_2_DEVICE;
  if (walk_num <= 0) return QMCKL_INVALID_ARG_3_DEVICE;
  if (!coord) return fail(context, QMCKL_INVALID_ARG_4_DEVICE, fn, "NULL pointer");
  if (!ee_distance) return fail(context, QMCKL_INVALID_ARG_5_DEVICE, fn, "NULL pointer");
  if (!has_device(c)) return fail(context, QMCKL_FAILURE_DEVICE, fn, "no device bound");
  hipError_t e;
  {
    KernelTimer t(c, "ee_distance");
    e = launch_ee_distance(c->stream, elec_num, walk_num, coord, ee_distance);
    t.stop();
  }
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

qmckl_exit_code_device qmckl_provide_ee_distance_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_provide_ee_distance_device";
  if (!c->electron.provided) return fail(context, QMCKL_NOT_PROVIDED_DEVICE, fn, nullptr);
  const int64_t E = c->electron.num, W = c->electron.walker.num;
  if (W <= 0 || c->point.coord == nullptr || c->point.num != E * W)
    return fail(context, QMCKL_NOT_PROVIDED_DEVICE, fn, "no electron coordinates set");
  if (c->point.date > c->electron.ee_distance_date || c->electron.ee_distance == nullptr) {
    const int64_t need = E * E * W;
    if (need > c->electron.ee_distance_cap) {
      if (!drealloc(c, reinterpret_cast<void **>(&c->electron.ee_distance), sizeof(double) * need))
        return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
      c->electron.ee_distance_cap = need;
    }
    qmckl_exit_code_device rc =
        qmckl_compute_ee_distance_device(context, E, W, c->point.coord, c->electron.ee_distance);
    if (rc != QMCKL_SUCCESS_DEVICE) return rc;
    c->electron.ee_distance_date = c->date;
  }
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_get_electron_ee_distance_device(qmckl_context_device context,
                                                             double *ee_distance, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_get_electron_ee_distance_device";
  if (ee_distance == nullptr) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "NULL pointer");
  qmckl_exit_code_device rc = qmckl_provide_ee_distance_device(context);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  const int64_t sze = c->electron.num * c->electron.num * c->electron.walker.num;
  if (size_max < sze) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "input array too small");
  hipError_t e = hipMemcpyAsync(ee_distance, c->electron.ee_distance, sizeof(double) * sze,
                                hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

}  // extern "C"
