// Low-level AO launchers of the reference API, which take raw device arrays
// instead of a finalized context:
//   qmckl_compute_ao_basis_shell_gaussian_vgl_device  (src/qmckl_ao_sycl.cpp:12-102)
//   qmckl_compute_ao_vgl_gaussian_device              (src/qmckl_ao_sycl.cpp:106-434)
// Both copy the small index arrays to the host, build temporary device
// tables and run the same kernels as the providers; the second one reads
// the radial parts from the caller's shell_vgl [P][5][S], as the reference
// does.
#include <vector>

#include "internal.hpp"

using namespace qmckl_amd;

namespace {

template <typename T>
bool d2h(Context *c, std::vector<T> &dst, const T *src, int64_t n) {
  dst.resize((size_t)(n > 0 ? n : 0));
  if (n <= 0) return true;
  return hipMemcpyAsync(dst.data(), src, sizeof(T) * n, hipMemcpyDeviceToHost, c->stream) ==
             hipSuccess &&
         sync(c) == hipSuccess;
}

}  // namespace

extern "C" {

qmckl_exit_code_device qmckl_compute_ao_basis_shell_gaussian_vgl_device(
    qmckl_context_device context, int prim_num, int shell_num, int point_num, int nucl_num,
    int64_t *nucleus_shell_num, int64_t *nucleus_index, double *nucleus_range,
    int64_t *shell_prim_index, int64_t *shell_prim_num, double *coord, double *nucl_coord,
    double *expo, double *coef_normalized, double *shell_vgl) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_compute_ao_basis_shell_gaussian_vgl_device";
  if (prim_num <= 0) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "prim_num <= 0");
  if (shell_num <= 0) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "shell_num <= 0");
  if (point_num < 0) return fail(context, QMCKL_INVALID_ARG_4_DEVICE, fn, "point_num < 0");
  if (nucl_num <= 0) return fail(context, QMCKL_INVALID_ARG_5_DEVICE, fn, "nucl_num <= 0");
  if (!nucleus_shell_num || !nucleus_index || !nucleus_range || !shell_prim_index ||
      !shell_prim_num || !coord || !nucl_coord || !expo || !coef_normalized || !shell_vgl)
    return fail(context, QMCKL_INVALID_ARG_6_DEVICE, fn, "NULL pointer");
  if (!has_device(c)) return fail(context, QMCKL_FAILURE_DEVICE, fn, "no device bound");
  std::vector<int64_t> nsn, nidx, spi, spn;
  if (!d2h(c, nsn, (const int64_t *)nucleus_shell_num, nucl_num) ||
      !d2h(c, nidx, (const int64_t *)nucleus_index, nucl_num) ||
      !d2h(c, spi, (const int64_t *)shell_prim_index, shell_num) ||
      !d2h(c, spn, (const int64_t *)shell_prim_num, shell_num))
    return fail(context, QMCKL_FAILURE_DEVICE, fn, "copy to the host failed");
  // Angular momenta are not needed for the radial parts: treat every shell
  // as l = 0 so the table builder's AO count is shell_num.
  std::vector<int32_t> l0((size_t)shell_num, 0);
  std::vector<double> ones_s((size_t)shell_num, 1.0), ones_p((size_t)prim_num, 1.0);
  DeviceBasis b;
  std::string msg = build_basis_tables(c, b, nucl_num, shell_num, prim_num, shell_num, nidx, nsn, l0,
                                       spn, spi, ones_s, ones_p, ones_p, ones_p, expo);
  if (!msg.empty()) {
    release_basis_tables(c, b);
    return fail(context, QMCKL_FAILURE_DEVICE, fn, msg.c_str());
  }
  DeviceBasis view = b;
  view.coef_norm = coef_normalized;
  view.nucl_range = nucleus_range;
  hipError_t e = launch_shell_vgl(c->stream, view, point_num, coord, nucl_coord, shell_vgl);
  if (e == hipSuccess) e = sync(c);
  release_basis_tables(c, b);
  return hip_check(context, e, fn);
}

qmckl_exit_code_device qmckl_compute_ao_basis_shell_gaussian_vgl_noalloc_device(
    qmckl_context_device context, int prim_num, int shell_num, int point_num, int nucl_num,
    const int64_t *nucleus_shell_num, const int64_t *nucleus_index, const double *nucleus_range,
    const int64_t *shell_prim_index, const int64_t *shell_prim_num, const double *coord,
    const double *nucl_coord, const double *expo, const double *coef_normalized, double *shell_vgl) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_compute_ao_basis_shell_gaussian_vgl_noalloc_device";
  if (prim_num <= 0) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "prim_num <= 0");
  if (shell_num <= 0) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "shell_num <= 0");
  if ((int64_t)shell_num > kNoAllocMaxShells)
    return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn,
                "shell_num exceeds the 65536 shells of the allocation-free scratch table");
  if (point_num < 0) return fail(context, QMCKL_INVALID_ARG_4_DEVICE, fn, "point_num < 0");
  if (nucl_num <= 0) return fail(context, QMCKL_INVALID_ARG_5_DEVICE, fn, "nucl_num <= 0");
  if (!nucleus_shell_num || !nucleus_index || !nucleus_range || !shell_prim_index ||
      !shell_prim_num || !coord || !nucl_coord || !expo || !coef_normalized || !shell_vgl)
    return fail(context, QMCKL_INVALID_ARG_6_DEVICE, fn, "NULL pointer");
  if (!has_device(c)) return fail(context, QMCKL_FAILURE_DEVICE, fn, "no device bound");
  if (hipSetDevice(c->device_id) != hipSuccess) return hip_check(context, hipGetLastError(), fn);
  int table_err = 0;
  hipError_t e = launch_shell_vgl_noalloc(c->stream, c->device_id, point_num, nucl_num, shell_num, prim_num,
                                          nucleus_shell_num, nucleus_index, nucleus_range, shell_prim_index,
                                          shell_prim_num, coord, nucl_coord, expo, coef_normalized, shell_vgl,
                                          &table_err);
  if (e == hipSuccess && table_err != 0) {
    // the messages of the allocating launcher's host-side checks
    const char *msg = (table_err & 1) ? "nucleus_index/nucleus_shell_num out of range"
                      : (table_err & 4) ? "shell_prim_index/shell_prim_num out of range"
                                        : "a shell belongs to no nucleus or to two";
    return fail(context, QMCKL_FAILURE_DEVICE, fn, msg);
  }
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

qmckl_exit_code_device qmckl_compute_ao_vgl_gaussian_device(
    qmckl_context_device context, int64_t ao_num, int64_t shell_num, int64_t point_num,
    int64_t nucl_num, const double *coord, const double *nucl_coord,
    const int64_t *nucleus_index, const int64_t *nucleus_shell_num,
    const double *nucleus_range, const int32_t *nucleus_max_ang_mom,
    const int32_t *shell_ang_mom, const double *ao_factor, double *shell_vgl, double *ao_vgl) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_compute_ao_vgl_gaussian_device";
  if (ao_num <= 0) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "ao_num <= 0");
  if (shell_num <= 0) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "shell_num <= 0");
  if (point_num < 0) return fail(context, QMCKL_INVALID_ARG_4_DEVICE, fn, "point_num < 0");
  if (nucl_num <= 0) return fail(context, QMCKL_INVALID_ARG_5_DEVICE, fn, "nucl_num <= 0");
  if (!coord || !nucl_coord || !nucleus_index || !nucleus_shell_num || !shell_ang_mom ||
      !ao_factor || !shell_vgl || !ao_vgl)
    return fail(context, QMCKL_INVALID_ARG_6_DEVICE, fn, "NULL pointer");
  (void)nucleus_range;        // screening already applied in shell_vgl (zeros)
  (void)nucleus_max_ang_mom;  // derived from shell_ang_mom
  if (!has_device(c)) return fail(context, QMCKL_FAILURE_DEVICE, fn, "no device bound");
  std::vector<int64_t> nidx, nsn;
  std::vector<int32_t> sam;
  if (!d2h(c, nidx, nucleus_index, nucl_num) || !d2h(c, nsn, nucleus_shell_num, nucl_num) ||
      !d2h(c, sam, shell_ang_mom, shell_num))
    return fail(context, QMCKL_FAILURE_DEVICE, fn, "copy to the host failed");
  // No primitives are involved: one dummy primitive per shell.
  std::vector<int64_t> spn((size_t)shell_num, 1), spi((size_t)shell_num);
  for (int64_t s = 0; s < shell_num; ++s) spi[s] = s;
  std::vector<double> ones_s((size_t)shell_num, 1.0);
  DeviceBasis b;
  std::string msg = build_basis_tables(c, b, nucl_num, shell_num, shell_num, ao_num, nidx, nsn, sam,
                                       spn, spi, ones_s, ones_s, ones_s, ones_s, nullptr);
  if (!msg.empty()) {
    release_basis_tables(c, b);
    return fail(context, QMCKL_FAILURE_DEVICE, fn, msg.c_str());
  }
  hipError_t e = launch_ao_from_shell_table(c->stream, b, point_num, coord, nucl_coord, ao_factor,
                                            shell_vgl, shell_num, ao_vgl, 5);
  if (e == hipSuccess) e = sync(c);
  release_basis_tables(c, b);
  return hip_check(context, e, fn);
}

}  // extern "C"
