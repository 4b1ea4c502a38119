// AO basis: setters/getters, finalize and the shell_vgl / ao_vgl / ao_value
// providers.
//
// Follows src/qmckl_ao.cpp: init mask 6-24 (14 bits), providers 47-199 and
// 748-849 of src/qmckl_ao_sycl.cpp, getters 207-396, setters 402-1236 (each
// re-finalises once all 14 bits are clear), getters 1238-1765, and
// src/qmckl_ao_sycl.cpp:1067-1236 for the derived data (normalised
// coefficients, nucleus_max_ang_mom, nucleus_range).  Finalize also builds
// the device tables the kernels read (one ShellDesc per shell in AO order,
// per-AO shell index and Cartesian exponents) after validating the index
// arrays on the host; cached results are invalidated when the basis
// changes.
#include <algorithm>
#include <cmath>
#include <string>

#include "internal.hpp"

using namespace qmckl_amd;

namespace {

template <typename T>
qmckl_exit_code_device set_array(qmckl_context_device context, Context *c, int bit, T **slot,
                                 const T *src, int64_t count, int64_t size_max, const char *fn,
                                 qmckl_exit_code_device size_code, const char *unset_msg) {
  if (count <= 0) return fail(context, QMCKL_FAILURE_DEVICE, fn, unset_msg);
  if (size_max < count) return fail(context, size_code, fn, "input array too small");
  if (src == nullptr) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "NULL pointer");
  if (!has_device(c)) return fail(context, QMCKL_FAILURE_DEVICE, fn, "no device bound");
  if (!drealloc(c, reinterpret_cast<void **>(slot), sizeof(T) * count))
    return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
  hipError_t e = hipMemcpyAsync(*slot, src, sizeof(T) * count, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = sync(c);
  if (e != hipSuccess) return hip_check(context, e, fn);
  c->ao.uninitialized &= ~(int32_t)(1 << bit);
  c->ao.provided = (c->ao.uninitialized == 0);
  if (c->ao.provided) return qmckl_finalize_ao_basis_device(context);
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device mark_scalar(qmckl_context_device context, Context *c, int bit) {
  c->ao.uninitialized &= ~(int32_t)(1 << bit);
  c->ao.provided = (c->ao.uninitialized == 0);
  if (c->ao.provided) return qmckl_finalize_ao_basis_device(context);
  return QMCKL_SUCCESS_DEVICE;
}

template <typename T>
qmckl_exit_code_device get_array(qmckl_context_device context, Context *c, int bit, const T *src,
                                 T *dst, int64_t count, int64_t size_max, const char *fn) {
  if (c->ao.uninitialized & (1 << bit)) return fail(context, QMCKL_NOT_PROVIDED_DEVICE, fn, nullptr);
  if (dst == nullptr) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "NULL pointer");
  if (size_max < count) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "Array too small");
  hipError_t e = hipMemcpyAsync(dst, src, sizeof(T) * count, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

template <typename T>
bool d2h(Context *c, std::vector<T> &dst, const T *src, int64_t n) {
  dst.resize((size_t)n);
  if (n == 0) return true;
  return hipMemcpyAsync(dst.data(), src, sizeof(T) * n, hipMemcpyDeviceToHost, c->stream) ==
             hipSuccess &&
         sync(c) == hipSuccess;
}

template <typename T>
T *h2d_new(Context *c, const std::vector<T> &src) {
  T *p = static_cast<T *>(dmalloc(c, sizeof(T) * std::max<size_t>(src.size(), 1)));
  if (!p) return nullptr;
  if (!src.empty() && hipMemcpyAsync(p, src.data(), sizeof(T) * src.size(),
                                     hipMemcpyHostToDevice, c->stream) != hipSuccess) {
    dfree(c, p);
    return nullptr;
  }
  return p;
}

void free_tables(Context *c, DeviceBasis &b) {
  dfree(c, b.shells);
  dfree(c, b.ao_shell);
  dfree(c, b.ao_abc);
  dfree(c, b.coef_norm);
  dfree(c, b.nucl_range);
  dfree(c, b.nucl_max_l);
  dfree(c, b.nucl_prim_index);
  dfree(c, b.chunks);
  dfree(c, b.stage_shell);
  b.chunks = nullptr;
  b.nchunks = 0;
  b.stage_shell = nullptr;
  b.shells = nullptr;
  b.ao_shell = b.ao_abc = b.nucl_max_l = nullptr;
  b.coef_norm = b.nucl_range = nullptr;
  b.nucl_prim_index = nullptr;
  b.expo = nullptr;
}

}  // namespace

namespace qmckl_amd {

// Host-side construction of the device tables from host copies of the
// basis arrays.  Returns an error message (empty on success).
std::string build_basis_tables(Context *c, DeviceBasis &b, int64_t N, int64_t S, int64_t Np,
                               int64_t A, const std::vector<int64_t> &nucleus_index,
                               const std::vector<int64_t> &nucleus_shell_num,
                               const std::vector<int32_t> &shell_ang_mom,
                               const std::vector<int64_t> &shell_prim_num,
                               const std::vector<int64_t> &shell_prim_index,
                               const std::vector<double> &shell_factor,
                               const std::vector<double> &exponent,
                               const std::vector<double> &coefficient,
                               const std::vector<double> &prim_factor, double *expo_dev) {
  std::vector<ShellDesc> shells;
  std::vector<int32_t> ao_shell, ao_abc, nucl_max_l((size_t)N, 0);
  std::vector<double> coef_norm((size_t)Np, 0.0), nucl_range((size_t)N, 0.0);
  std::vector<int64_t> nucl_prim_index((size_t)N + 1, 0);
  std::vector<char> seen((size_t)S, 0);
  int max_prims = 0, max_l = 0;
  if (N > INT32_MAX || S > INT32_MAX || Np > INT32_MAX || A > INT32_MAX)
    return "basis too large for 32-bit device indices";
  for (int64_t n = 0; n < N; ++n) {
    const int64_t s0 = nucleus_index[n], ns = nucleus_shell_num[n];
    if (ns < 0 || s0 < 0 || s0 + ns > S) return "nucleus_index/nucleus_shell_num out of range";
    nucl_prim_index[n] = ns > 0 ? shell_prim_index[s0] : 0;
    for (int64_t s = s0; s < s0 + ns; ++s) {
      if (seen[s]) return "a shell belongs to two nuclei";
      seen[s] = 1;
      const int l = shell_ang_mom[s];
      if (l < 0 || l > kMaxL) return "shell angular momentum out of the supported range";
      const int64_t p0 = shell_prim_index[s], np = shell_prim_num[s];
      if (np < 0 || p0 < 0 || p0 + np > Np) return "shell_prim_index/shell_prim_num out of range";
      ShellDesc d;
      d.nucl = (int32_t)n;
      d.prim_start = (int32_t)p0;
      d.prim_num = (int32_t)np;
      d.ao_start = (int32_t)ao_shell.size();
      d.l = l;
      d.shell = (int32_t)s;
      d.a_min = np > 0 ? exponent[p0] : 0.0;
      for (int64_t j = p0; j < p0 + np; ++j) d.a_min = std::min(d.a_min, exponent[j]);
      const int32_t pos = (int32_t)shells.size();
      shells.push_back(d);
      for (int a = l; a >= 0; --a)
        for (int bb = l - a; bb >= 0; --bb) {
          ao_shell.push_back(pos);
          ao_abc.push_back(a | (bb << 8) | ((l - a - bb) << 16));
        }
      max_prims = std::max<int>(max_prims, (int)np);
      max_l = std::max(max_l, l);
      nucl_max_l[n] = std::max(nucl_max_l[n], l);
      for (int64_t j = p0; j < p0 + np; ++j) {
        coef_norm[j] = coefficient[j] * prim_factor[j] * shell_factor[s];
        const double r = 1.0 / exponent[j];
        if (r > nucl_range[n]) nucl_range[n] = r;
      }
    }
  }
  nucl_prim_index[N] = Np;
  if ((int64_t)shells.size() != S) return "some shells belong to no nucleus";
  if ((int64_t)ao_shell.size() != A) return "ao_num does not match the shells' Cartesian components";
  b.nucl_num = N;
  b.shell_num = (int64_t)shells.size();
  b.prim_num = Np;
  b.ao_num = A;
  b.max_shell_prims = max_prims;
  b.max_l = max_l;
  b.expo = expo_dev;
  b.shells = h2d_new(c, shells);
  b.ao_shell = h2d_new(c, ao_shell);
  b.ao_abc = h2d_new(c, ao_abc);
  b.coef_norm = h2d_new(c, coef_norm);
  b.nucl_range = h2d_new(c, nucl_range);
  b.nucl_max_l = h2d_new(c, nucl_max_l);
  b.nucl_prim_index = h2d_new(c, nucl_prim_index);
  // chunks of whole shells for the AO kernel: <= kChunkShells shells and
  // <= kChunkAos AOs each
  std::vector<int32_t> chunks{0};
  int ns = 0, na = 0;
  for (size_t i = 0; i < shells.size(); ++i) {
    const int nc = (shells[i].l + 1) * (shells[i].l + 2) / 2;
    if (ns > 0 && (ns == kChunkShells || na + nc > kChunkAos)) {
      chunks.push_back((int32_t)i);
      ns = na = 0;
    }
    ++ns;
    na += nc;
  }
  chunks.push_back((int32_t)shells.size());
  b.nchunks = (int)chunks.size() - 1;
  b.chunks = h2d_new(c, chunks);
  // first shell holding an AO of each 64-AO stage (S past the last AO), for
  // kernels that split the AOs of a point between workgroups
  const int64_t nst = (A + 63) / 64;
  std::vector<int32_t> stage_shell((size_t)nst + 1);
  for (int64_t st = 0; st <= nst; ++st)
    stage_shell[(size_t)st] = 64 * st < A ? ao_shell[(size_t)(64 * st)] : (int32_t)shells.size();
  b.stage_shell = h2d_new(c, stage_shell);
  if (!b.shells || !b.ao_shell || !b.ao_abc || !b.coef_norm || !b.nucl_range || !b.nucl_max_l ||
      !b.nucl_prim_index || !b.chunks || !b.stage_shell)
    return "device allocation failed";
  if (sync(c) != hipSuccess) return "copy to device failed";
  return "";
}

void release_basis_tables(Context *c, DeviceBasis &b) { free_tables(c, b); }

hipError_t tile_point_order(Context *c, const int32_t **perm) {
  *perm = nullptr;
  const int64_t P = c->point.num, ld = round_up(c->ao.ao_num, 64);
  const int32_t cfg = c->mo.tile_point_order;
  if (!(cfg >= 0 ? cfg == 1 : ld >= 1024 && P >= 4096)) return hipSuccess;
  AoBasis &ao = c->ao;
  if (!grow_bytes(c, reinterpret_cast<void **>(&ao.point_perm), &ao.point_perm_cap, (int64_t)sizeof(int32_t) * P) ||
      !grow_bytes(c, &ao.order_ws, &ao.order_ws_cap, point_order_workspace_bytes(P)))
    return hipErrorOutOfMemory;
  KernelTimer t(c, "point_order");
  const hipError_t e =
      launch_point_order(c->stream, P, c->point.coord, c->nucleus.num, c->nucleus.coord, ao.order_ws, ao.point_perm);
  t.stop();
  if (e == hipSuccess) *perm = ao.point_perm;
  return e;
}

}  // namespace qmckl_amd

extern "C" {

bool qmckl_ao_basis_provided_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  return c ? c->ao.provided : false;
}

qmckl_exit_code_device qmckl_finalize_ao_basis_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_finalize_ao_basis_device";
  if (!c->ao.provided) return fail(context, QMCKL_NOT_PROVIDED_DEVICE, fn, nullptr);
  int64_t N = 0;
  qmckl_exit_code_device rc = qmckl_get_nucleus_num_device(context, &N);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  AoBasis &ao = c->ao;
  const int64_t S = ao.shell_num, Np = ao.prim_num, A = ao.ao_num;
  std::vector<int64_t> nidx, nsn, spn, spi;
  std::vector<int32_t> sam;
  std::vector<double> sf, ex, co, pf;
  if (!d2h(c, nidx, ao.nucleus_index, N) || !d2h(c, nsn, ao.nucleus_shell_num, N) ||
      !d2h(c, sam, ao.shell_ang_mom, S) || !d2h(c, spn, ao.shell_prim_num, S) ||
      !d2h(c, spi, ao.shell_prim_index, S) || !d2h(c, sf, ao.shell_factor, S) ||
      !d2h(c, ex, ao.exponent, Np) || !d2h(c, co, ao.coefficient, Np) ||
      !d2h(c, pf, ao.prim_factor, Np))
    return fail(context, QMCKL_FAILURE_DEVICE, fn, "copy of the basis to the host failed");
  free_tables(c, ao.dev);
  std::string msg = build_basis_tables(c, ao.dev, N, S, Np, A, nidx, nsn, sam, spn, spi, sf, ex,
                                       co, pf, ao.exponent);
  if (!msg.empty()) {
    free_tables(c, ao.dev);
    return fail(context, QMCKL_FAILURE_DEVICE, fn, msg.c_str());
  }
  // A new basis invalidates every cached AO/MO result and the mixed-form
  // operands (their padded row length depends on ao_num).
  c->basis_generation += 1;
  ao.shell_vgl_date = ao.ao_vgl_date = ao.ao_value_date = 0;
  c->mo.mo_vgl_date = c->mo.mo_value_date = 0;
  dfree(c, ao.ao_f32);
  ao.ao_f32 = nullptr;
  ao.ao_f32_cap = 0;
  dfree(c, ao.ao_f32_lo);
  ao.ao_f32_lo = nullptr;
  ao.ao_f32_lo_cap = 0;
  dfree(c, c->mo.coef_split);
  c->mo.coef_split = nullptr;
  dfree(c, c->mo.coef_i8);
  c->mo.coef_i8 = nullptr;
  c->mo.coef_split_form = QMCKL_PRECISION_FP64_DEVICE;
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_set_ao_basis_type_device(qmckl_context_device context, char basis_type) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  if (basis_type != 'G' && basis_type != 'S')
    return fail(context, QMCKL_INVALID_ARG_2_DEVICE, "qmckl_set_ao_basis_type_device", nullptr);
  c->ao.type = basis_type;
  return mark_scalar(context, c, 0);
}

qmckl_exit_code_device qmckl_set_ao_basis_shell_num_device(qmckl_context_device context,
                                                           int64_t shell_num) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_set_ao_basis_shell_num_device";
  if (shell_num <= 0) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "shell_num <= 0");
  if (0 < c->ao.prim_num && c->ao.prim_num < shell_num)
    return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "shell_num > prim_num");
  c->ao.shell_num = shell_num;
  return mark_scalar(context, c, 1);
}

qmckl_exit_code_device qmckl_set_ao_basis_prim_num_device(qmckl_context_device context,
                                                          int64_t prim_num) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_set_ao_basis_prim_num_device";
  if (prim_num <= 0) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "prim_num must be positive");
  if (c->ao.shell_num <= 0) return fail(context, QMCKL_FAILURE_DEVICE, fn, "shell_num is not set");
  if (prim_num < c->ao.shell_num) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "prim_num < shell_num");
  c->ao.prim_num = prim_num;
  return mark_scalar(context, c, 2);
}

qmckl_exit_code_device qmckl_set_ao_basis_ao_num_device(qmckl_context_device context, int64_t ao_num) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_set_ao_basis_ao_num_device";
  if (ao_num <= 0) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "ao_num must be positive");
  if (c->ao.shell_num <= 0) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "shell_num is not set");
  if (ao_num < c->ao.shell_num) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "ao_num < shell_num");
  c->ao.ao_num = ao_num;
  return mark_scalar(context, c, 12);
}

qmckl_exit_code_device qmckl_set_ao_basis_nucleus_index_device(qmckl_context_device context,
                                                               int64_t *nucleus_index, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  // The reference reports a short array with QMCKL_FAILURE here
  // (src/qmckl_ao.cpp:598-602) and with INVALID_ARG_3 everywhere else.
  return set_array(context, c, 4, &c->ao.nucleus_index, (const int64_t *)nucleus_index,
                   c->nucleus.num, size_max, "qmckl_set_ao_basis_nucleus_index_device",
                   QMCKL_FAILURE_DEVICE, "nucl_num is not set");
}

qmckl_exit_code_device qmckl_set_ao_basis_nucleus_shell_num_device(qmckl_context_device context,
                                                                   int64_t *nucleus_shell_num,
                                                                   int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return set_array(context, c, 3, &c->ao.nucleus_shell_num, (const int64_t *)nucleus_shell_num,
                   c->nucleus.num, size_max, "qmckl_set_ao_basis_nucleus_shell_num_device",
                   QMCKL_INVALID_ARG_3_DEVICE, "nucl_num is not set");
}

qmckl_exit_code_device qmckl_set_ao_basis_shell_ang_mom_device(qmckl_context_device context,
                                                               int32_t *shell_ang_mom, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return set_array(context, c, 5, &c->ao.shell_ang_mom, (const int32_t *)shell_ang_mom,
                   c->ao.shell_num, size_max, "qmckl_set_ao_basis_shell_ang_mom_device",
                   QMCKL_INVALID_ARG_3_DEVICE, "shell_num is not set");
}

qmckl_exit_code_device qmckl_set_ao_basis_shell_prim_num_device(qmckl_context_device context,
                                                                int64_t *shell_prim_num, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return set_array(context, c, 6, &c->ao.shell_prim_num, (const int64_t *)shell_prim_num,
                   c->ao.shell_num, size_max, "qmckl_set_ao_basis_shell_prim_num_device",
                   QMCKL_INVALID_ARG_3_DEVICE, "shell_num is not set");
}

qmckl_exit_code_device qmckl_set_ao_basis_shell_prim_index_device(qmckl_context_device context,
                                                                  int64_t *shell_prim_index,
                                                                  int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return set_array(context, c, 7, &c->ao.shell_prim_index, (const int64_t *)shell_prim_index,
                   c->ao.shell_num, size_max, "qmckl_set_ao_basis_shell_prim_index_device",
                   QMCKL_INVALID_ARG_3_DEVICE, "shell_num is not set");
}

qmckl_exit_code_device qmckl_set_ao_basis_shell_factor_device(qmckl_context_device context,
                                                              double *shell_factor, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return set_array(context, c, 8, &c->ao.shell_factor, (const double *)shell_factor,
                   c->ao.shell_num, size_max, "qmckl_set_ao_basis_shell_factor_device",
                   QMCKL_INVALID_ARG_3_DEVICE, "shell_num is not set");
}

qmckl_exit_code_device qmckl_set_ao_basis_exponent_device(qmckl_context_device context,
                                                          double *exponent, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return set_array(context, c, 9, &c->ao.exponent, (const double *)exponent, c->ao.prim_num,
                   size_max, "qmckl_set_ao_basis_exponent_device", QMCKL_INVALID_ARG_3_DEVICE,
                   "prim_num is not set");
}

qmckl_exit_code_device qmckl_set_ao_basis_coefficient_device(qmckl_context_device context,
                                                             double *coefficient, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return set_array(context, c, 10, &c->ao.coefficient, (const double *)coefficient,
                   c->ao.prim_num, size_max, "qmckl_set_ao_basis_coefficient_device",
                   QMCKL_INVALID_ARG_3_DEVICE, "prim_num is not set");
}

qmckl_exit_code_device qmckl_set_ao_basis_prim_factor_device(qmckl_context_device context,
                                                             double *prim_factor, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return set_array(context, c, 11, &c->ao.prim_factor, (const double *)prim_factor,
                   c->ao.prim_num, size_max, "qmckl_set_ao_basis_prim_factor_device",
                   QMCKL_INVALID_ARG_3_DEVICE, "prim_num is not set");
}

qmckl_exit_code_device qmckl_set_ao_basis_ao_factor_device(qmckl_context_device context,
                                                           double *ao_factor, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return set_array(context, c, 13, &c->ao.ao_factor, (const double *)ao_factor, c->ao.ao_num,
                   size_max, "qmckl_set_ao_basis_ao_factor_device", QMCKL_INVALID_ARG_3_DEVICE,
                   "ao_num is not set");
}

// ---- scalar getters (src/qmckl_ao.cpp:1238-1340) ----

qmckl_exit_code_device qmckl_get_ao_basis_type_device(qmckl_context_device context, char *basis_type) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  if (c->ao.uninitialized & 1)
    return fail(context, QMCKL_NOT_PROVIDED_DEVICE, "qmckl_get_ao_basis_type_device", nullptr);
  if (!basis_type) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, "qmckl_get_ao_basis_type_device", "NULL pointer");
  *basis_type = c->ao.type;
  return QMCKL_SUCCESS_DEVICE;
}

static qmckl_exit_code_device get_scalar(qmckl_context_device context, int bit, int64_t value,
                                         int64_t *out, const char *fn) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  if (c->ao.uninitialized & (1 << bit)) return fail(context, QMCKL_NOT_PROVIDED_DEVICE, fn, nullptr);
  if (!out) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "NULL pointer");
  *out = value;
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_get_ao_basis_shell_num_device(qmckl_context_device context, int64_t *shell_num) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return get_scalar(context, 1, c->ao.shell_num, shell_num, "qmckl_get_ao_basis_shell_num_device");
}

qmckl_exit_code_device qmckl_get_ao_basis_prim_num_device(qmckl_context_device context, int64_t *prim_num) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return get_scalar(context, 2, c->ao.prim_num, prim_num, "qmckl_get_ao_basis_prim_num_device");
}

qmckl_exit_code_device qmckl_get_ao_basis_ao_num_device(qmckl_context_device context, int64_t *ao_num) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return get_scalar(context, 12, c->ao.ao_num, ao_num, "qmckl_get_ao_basis_ao_num_device");
}

// ---- array getters (src/qmckl_ao.cpp:1342-1765) ----

qmckl_exit_code_device qmckl_get_ao_basis_nucleus_index_device(qmckl_context_device context,
                                                               int64_t *nucleus_index, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return get_array(context, c, 4, c->ao.nucleus_index, nucleus_index, c->nucleus.num, size_max,
                   "qmckl_get_ao_basis_nucleus_index_device");
}

qmckl_exit_code_device qmckl_get_ao_basis_nucleus_shell_num_device(qmckl_context_device context,
                                                                   int64_t *nucleus_shell_num,
                                                                   int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return get_array(context, c, 3, c->ao.nucleus_shell_num, nucleus_shell_num, c->nucleus.num,
                   size_max, "qmckl_get_ao_basis_nucleus_shell_num_device");
}

qmckl_exit_code_device qmckl_get_ao_basis_shell_ang_mom_device(qmckl_context_device context,
                                                               int32_t *shell_ang_mom, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return get_array(context, c, 5, c->ao.shell_ang_mom, shell_ang_mom, c->ao.shell_num, size_max,
                   "qmckl_get_ao_basis_shell_ang_mom_device");
}

qmckl_exit_code_device qmckl_get_ao_basis_shell_prim_num_device(qmckl_context_device context,
                                                                int64_t *shell_prim_num, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return get_array(context, c, 6, c->ao.shell_prim_num, shell_prim_num, c->ao.shell_num, size_max,
                   "qmckl_get_ao_basis_shell_prim_num_device");
}

qmckl_exit_code_device qmckl_get_ao_basis_shell_prim_index_device(qmckl_context_device context,
                                                                  int64_t *shell_prim_index,
                                                                  int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return get_array(context, c, 7, c->ao.shell_prim_index, shell_prim_index, c->ao.shell_num,
                   size_max, "qmckl_get_ao_basis_shell_prim_index_device");
}

qmckl_exit_code_device qmckl_get_ao_basis_shell_factor_device(qmckl_context_device context,
                                                              double *shell_factor, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return get_array(context, c, 8, c->ao.shell_factor, shell_factor, c->ao.shell_num, size_max,
                   "qmckl_get_ao_basis_shell_factor_device");
}

qmckl_exit_code_device qmckl_get_ao_basis_exponent_device(qmckl_context_device context,
                                                          double *exponent, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return get_array(context, c, 9, c->ao.exponent, exponent, c->ao.prim_num, size_max,
                   "qmckl_get_ao_basis_exponent_device");
}

qmckl_exit_code_device qmckl_get_ao_basis_coefficient_device(qmckl_context_device context,
                                                             double *coefficient, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return get_array(context, c, 10, c->ao.coefficient, coefficient, c->ao.prim_num, size_max,
                   "qmckl_get_ao_basis_coefficient_device");
}

qmckl_exit_code_device qmckl_get_ao_basis_prim_factor_device(qmckl_context_device context,
                                                             double *prim_factor, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return get_array(context, c, 11, c->ao.prim_factor, prim_factor, c->ao.prim_num, size_max,
                   "qmckl_get_ao_basis_prim_factor_device");
}

qmckl_exit_code_device qmckl_get_ao_basis_ao_factor_device(qmckl_context_device context,
                                                           double *ao_factor, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  return get_array(context, c, 13, c->ao.ao_factor, ao_factor, c->ao.ao_num, size_max,
                   "qmckl_get_ao_basis_ao_factor_device");
}

// ---- providers ----

static qmckl_exit_code_device ensure_points(qmckl_context_device context, Context *c, const char *fn) {
  if (!c->ao.provided) return fail(context, QMCKL_NOT_PROVIDED_DEVICE, fn, nullptr);
  if (c->ao.type != 'G')
    return fail(context, QMCKL_FAILURE_DEVICE, fn, "Not yet implemented for basis type != 'G'");
  if (!c->nucleus.provided) return fail(context, QMCKL_NOT_PROVIDED_DEVICE, fn, "nucleus not provided");
  if (c->point.num <= 0 || c->point.coord == nullptr)
    return fail(context, QMCKL_NOT_PROVIDED_DEVICE, fn, "no points set");
  if (!has_device(c)) return fail(context, QMCKL_FAILURE_DEVICE, fn, "no device bound");
  return QMCKL_SUCCESS_DEVICE;
}

static bool grow(Context *c, double **slot, int64_t *cap, int64_t need) {
  if (*slot != nullptr && *cap >= need) return true;
  if (!drealloc(c, reinterpret_cast<void **>(slot), sizeof(double) * (size_t)std::max<int64_t>(need, 1)))
    return false;
  *cap = need;
  return true;
}

qmckl_exit_code_device qmckl_provide_ao_basis_shell_vgl_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_provide_ao_basis_shell_vgl_device";
  qmckl_exit_code_device rc = ensure_points(context, c, fn);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  if (c->point.date > c->ao.shell_vgl_date || c->ao.shell_vgl_date == 0) {
    const int64_t need = c->point.num * 5 * c->ao.shell_num;
    if (!grow(c, &c->ao.shell_vgl, &c->ao.shell_vgl_cap, need))
      return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
    hipError_t e;
    {
      KernelTimer t(c, "shell_vgl");
      e = launch_shell_vgl(c->stream, c->ao.dev, c->point.num, c->point.coord, c->nucleus.coord,
                           c->ao.shell_vgl);
      t.stop();
    }
    if (e == hipSuccess) e = sync(c);
    if (e != hipSuccess) return hip_check(context, e, fn);
    c->ao.shell_vgl_date = c->date;
  }
  return QMCKL_SUCCESS_DEVICE;
}

// Compute ao_vgl (ncomp = 5) or ao_value (ncomp = 1) for the current points
// into `dst`.  The lane-per-point kernel runs over spatially compact point
// tiles where tile_point_order says so (each point's rows are the same
// either way).
static qmckl_exit_code_device compute_ao(qmckl_context_device context, Context *c, double *dst,
                                         int ncomp, const char *fn, uint8_t *flags = nullptr,
                                         int nkb = 0) {
  const int32_t *perm = nullptr;
  hipError_t e = ao_vgl_lane_kernel(c->ao.ao_num) ? tile_point_order(c, &perm) : hipSuccess;
  if (e == hipErrorOutOfMemory) return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
  if (e == hipSuccess) {
    KernelTimer t(c, ncomp == 5 ? "ao_vgl" : "ao_value");
    e = launch_ao_vgl(c->stream, c->ao.dev, c->point.num, c->point.coord, c->nucleus.coord,
                      c->ao.ao_factor, dst, ncomp, flags, nkb, perm);
    t.stop();
  }
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

qmckl_exit_code_device qmckl_provide_ao_basis_ao_vgl_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_provide_ao_basis_ao_vgl_device";
  qmckl_exit_code_device rc = ensure_points(context, c, fn);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  if (c->point.date > c->ao.ao_vgl_date || c->ao.ao_vgl_date == 0) {
    const int64_t need = c->point.num * 5 * c->ao.ao_num;
    if (!grow(c, &c->ao.ao_vgl, &c->ao.ao_vgl_cap, need))
      return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
    // nonzero flags per (64 rows, 32 AOs) let the MO contraction skip
    // screened blocks
    const int nkb = (int)((c->ao.ao_num + 31) / 32);
    uint8_t *flags = nullptr;
    if (nkb <= kMaxFlagBlocks) {
      const int64_t nflags = (5 * c->point.num + 63) / 64 * nkb;
      if (!grow_bytes(c, reinterpret_cast<void **>(&c->ao.ao_vgl_flags), &c->ao.ao_vgl_flags_cap, nflags))
        return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
      flags = c->ao.ao_vgl_flags;
    }
    rc = compute_ao(context, c, c->ao.ao_vgl, 5, fn, flags, flags ? nkb : 0);
    if (rc != QMCKL_SUCCESS_DEVICE) return rc;
    c->ao.ao_vgl_date = c->date;
  }
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_provide_ao_basis_ao_value_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_provide_ao_basis_ao_value_device";
  qmckl_exit_code_device rc = ensure_points(context, c, fn);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  if (c->point.date > c->ao.ao_value_date || c->ao.ao_value_date == 0) {
    const int64_t P = c->point.num, A = c->ao.ao_num;
    if (!grow(c, &c->ao.ao_value, &c->ao.ao_value_cap, P * A))
      return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
    if (c->ao.ao_vgl_date != 0 && c->point.date <= c->ao.ao_vgl_date && c->ao.ao_vgl) {
      // ao_vgl is current: take its value rows (src/qmckl_ao_sycl.cpp:788-807).
      hipError_t e = launch_strided_copy(c->stream, P, A, c->ao.ao_vgl, 5 * A, c->ao.ao_value, A);
      if (e == hipSuccess) e = sync(c);
      if (e != hipSuccess) return hip_check(context, e, fn);
    } else {
      rc = compute_ao(context, c, c->ao.ao_value, 1, fn);
      if (rc != QMCKL_SUCCESS_DEVICE) return rc;
    }
    c->ao.ao_value_date = c->date;
  }
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_get_ao_basis_shell_vgl_device(qmckl_context_device context,
                                                           double *shell_vgl, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_get_ao_basis_shell_vgl_device";
  if (!shell_vgl) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "NULL pointer");
  qmckl_exit_code_device rc = qmckl_provide_ao_basis_shell_vgl_device(context);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  const int64_t sze = c->ao.shell_num * 5 * c->point.num;
  if (size_max < sze) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "input array too small");
  hipError_t e = hipMemcpyAsync(shell_vgl, c->ao.shell_vgl, sizeof(double) * sze,
                                hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

qmckl_exit_code_device qmckl_get_ao_basis_ao_vgl_device(qmckl_context_device context, double *ao_vgl,
                                                        int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_get_ao_basis_ao_vgl_device";
  if (!ao_vgl) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "NULL pointer");
  qmckl_exit_code_device rc = qmckl_provide_ao_basis_ao_vgl_device(context);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  const int64_t sze = c->ao.ao_num * 5 * c->point.num;
  if (size_max < sze) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "input array too small");
  hipError_t e = hipMemcpyAsync(ao_vgl, c->ao.ao_vgl, sizeof(double) * sze, hipMemcpyDeviceToDevice,
                                c->stream);
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

qmckl_exit_code_device qmckl_get_ao_basis_ao_vgl_inplace_device(qmckl_context_device context,
                                                                double *ao_vgl, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_get_ao_basis_ao_vgl_inplace_device";
  if (!ao_vgl) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "NULL pointer");
  const int64_t sze = c->ao.ao_num * 5 * c->point.num;
  if (size_max < sze) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "input array too small");
  qmckl_exit_code_device rc = qmckl_context_touch_device(context);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  rc = ensure_points(context, c, fn);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  // Computed straight into the caller's buffer; the context's own cache is
  // left stale (the touch above invalidated it).
  return compute_ao(context, c, ao_vgl, 5, fn);
}

qmckl_exit_code_device qmckl_get_ao_basis_ao_value_device(qmckl_context_device context,
                                                          double *ao_value, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_get_ao_basis_ao_value_device";
  if (!ao_value) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "NULL pointer");
  qmckl_exit_code_device rc = qmckl_provide_ao_basis_ao_value_device(context);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  const int64_t sze = c->ao.ao_num * c->point.num;
  if (size_max < sze) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "input array too small");
  hipError_t e = hipMemcpyAsync(ao_value, c->ao.ao_value, sizeof(double) * sze,
                                hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

qmckl_exit_code_device qmckl_get_ao_basis_ao_value_inplace_device(qmckl_context_device context,
                                                                  double *ao_value, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_get_ao_basis_ao_value_inplace_device";
  if (!ao_value) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "NULL pointer");
  const int64_t sze = c->ao.ao_num * c->point.num;
  if (size_max < sze) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "input array too small");
  qmckl_exit_code_device rc = qmckl_context_touch_device(context);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  rc = ensure_points(context, c, fn);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  return compute_ao(context, c, ao_value, 1, fn);
}

}  // extern "C"
