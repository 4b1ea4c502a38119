// MO basis: setters/getters, finalize and the mo_vgl / mo_value providers.
//
// Follows src/qmckl_mo.cpp (provided 10-23, provide_mo_vgl 114-181,
// provide_mo_value 185-286, getters 294-453, set_mo_num 486-524,
// set_coefficient 526-583) and src/qmckl_mo_sycl.cpp (contraction 11-56,
// finalize 136-190: coefficient_t[A][M] = coefficient[M][A]).  The
// contraction mo_vgl[(p,c)][i] = sum_k ao_vgl[(p,c)][k] coefficient_t[k][i]
// is one GEMM with 5P rows, run by the FP64 MFMA kernel on a copy of the
// coefficients in MFMA fragment order made at finalize.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "internal.hpp"

using namespace qmckl_amd;

namespace {

bool grow(Context *c, double **slot, int64_t *cap, int64_t need) {
  if (*slot != nullptr && *cap >= need) return true;
  if (!drealloc(c, reinterpret_cast<void **>(slot), sizeof(double) * (size_t)std::max<int64_t>(need, 1)))
    return false;
  *cap = need;
  return true;
}

qmckl_exit_code_device ready(qmckl_context_device context, Context *c, const char *fn) {
  if (!c->mo.provided) return fail(context, QMCKL_NOT_PROVIDED_DEVICE, fn, nullptr);
  if (!c->ao.provided) return fail(context, QMCKL_NOT_PROVIDED_DEVICE, fn, "ao_basis not provided");
  if (!has_device(c)) return fail(context, QMCKL_FAILURE_DEVICE, fn, "no device bound");
  return QMCKL_SUCCESS_DEVICE;
}

// Planned column orders (plan_mo_columns) are kept when at most kPlanLive
// of the (16-MO, 16-AO) blocks stay live; otherwise the MO order is used.
constexpr double kPlanLive = 0.9;
constexpr int64_t kPlanMaxAo = 16 * 4096;  // the GEMM kernels' stage lists hold 4,096 blocks of 16 AOs

void free_col_plans(Context *c) {
  for (ColPlanSlot &p : c->mo.col_plans) {
    dfree(c, p.col_mo);
    dfree(c, p.blk);
    p = ColPlanSlot();
  }
}

// The nonzero map of the coefficients at 16-AO granularity, on the host.
hipError_t map_coef_zeros(Context *c) {
  MoBasis &mo = c->mo;
  free_col_plans(c);
  mo.nz_words.clear();
  mo.nz_nw = 0;
  const int64_t A = c->ao.ao_num, M = mo.mo_num;
  if (A <= 0 || M <= 0 || A > kPlanMaxAo) return hipSuccess;
  const int64_t nw = ((A + 15) / 16 + 63) / 64;
  uint64_t *dwords = static_cast<uint64_t *>(dmalloc(c, sizeof(uint64_t) * (size_t)(M * nw)));
  if (dwords == nullptr) return hipErrorOutOfMemory;
  std::vector<uint64_t> words((size_t)(M * nw));
  hipError_t e = launch_coef_blocks(c->stream, M, A, mo.coefficient, nw, dwords);
  if (e == hipSuccess)
    e = hipMemcpyAsync(words.data(), dwords, sizeof(uint64_t) * words.size(), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = sync(c);
  dfree(c, dwords);
  if (e != hipSuccess) return e;
  mo.nz_words.swap(words);
  mo.nz_nw = nw;
  return hipSuccess;
}

// The column plan for workgroups of `group` columns: the MOs of each group
// of `group` consecutive MOs are ordered among themselves.  Built on first
// use after finalize (slot 0 holds the FP64 contraction's width).
hipError_t col_plan_for(Context *c, int64_t group, ColPlan *out) {
  MoBasis &mo = c->mo;
  *out = ColPlan();
  if (mo.nz_nw <= 0 || group <= 0 || group % 16 != 0) return hipSuccess;
  ColPlanSlot *slot = nullptr;
  for (ColPlanSlot &p : mo.col_plans)
    if (p.group == group) slot = &p;
  if (slot == nullptr) {
    slot = mo.col_plans[0].group == 0 ? &mo.col_plans[0] : &mo.col_plans[1];
    dfree(c, slot->col_mo);
    dfree(c, slot->blk);
    *slot = ColPlanSlot();
    const int64_t M = mo.mo_num, nw = mo.nz_nw, nkb16 = (c->ao.ao_num + 15) / 16;
    const int64_t nblk = (M + 15) / 16;
    std::vector<int32_t> order((size_t)M);
    std::vector<uint64_t> blk((size_t)(nblk * nw), 0);
    int64_t live = 0;
    for (int64_t g0 = 0; g0 < M; g0 += group) {
      const int64_t n = std::min<int64_t>(group, M - g0);
      std::vector<int32_t> o;
      std::vector<uint64_t> b;
      live += plan_mo_columns(mo.nz_words.data() + g0 * nw, n, nw, nkb16, o, b);
      for (int64_t j = 0; j < n; ++j) order[(size_t)(g0 + j)] = (int32_t)(g0 + o[(size_t)j]);
      std::copy(b.begin(), b.end(), blk.begin() + (g0 / 16) * nw);
    }
    slot->group = group;
    slot->live = (double)live / ((double)nblk * (double)nkb16);
    if (slot->live <= kPlanLive) {
      slot->col_mo = static_cast<int32_t *>(dmalloc(c, sizeof(int32_t) * order.size()));
      slot->blk = static_cast<uint64_t *>(dmalloc(c, sizeof(uint64_t) * blk.size()));
      if (slot->col_mo == nullptr || slot->blk == nullptr) return hipErrorOutOfMemory;
      hipError_t e = hipMemcpyAsync(slot->col_mo, order.data(), sizeof(int32_t) * order.size(),
                                    hipMemcpyHostToDevice, c->stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(slot->blk, blk.data(), sizeof(uint64_t) * blk.size(), hipMemcpyHostToDevice, c->stream);
      if (e == hipSuccess) e = sync(c);
      if (e != hipSuccess) return e;
      slot->host_blk.swap(blk);
    }
  }
  if (slot->col_mo != nullptr) {
    out->col_mo = slot->col_mo;
    out->blk = slot->blk;
    out->nw = mo.nz_nw;
    out->nblk = (mo.mo_num + 15) / 16;
  }
  return hipSuccess;
}

// the FP64 contraction's workgroup width (launch_mo_gemm_f64_frag)
int64_t f64_group(int64_t M) { return 64 * (int64_t)f64_frag_cbw(M); }

}  // namespace

namespace qmckl_amd {

hipError_t mo_f64_column_plan(Context *c, ColPlan *plan, const std::vector<uint64_t> **host_blk) {
  const hipError_t e = col_plan_for(c, f64_group(c->mo.mo_num), plan);
  *host_blk = nullptr;
  if (e == hipSuccess && plan->col_mo != nullptr)
    for (const ColPlanSlot &p : c->mo.col_plans)
      if (p.col_mo == plan->col_mo) *host_blk = &p.host_blk;
  return e;
}

}  // namespace qmckl_amd

namespace {

const char *gemm_timer_name(qmckl_precision_device form) {
  switch (form) {
    case QMCKL_PRECISION_FP32_DEVICE: return "mo_gemm_fp32";
    case QMCKL_PRECISION_FP16X2_DEVICE: return "mo_gemm_fp16x2";
    case QMCKL_PRECISION_BF16X2_DEVICE: return "mo_gemm_bf16x2";
    case QMCKL_PRECISION_BF16X3_DEVICE: return "mo_gemm_bf16x3";
    case QMCKL_PRECISION_INT8X5_DEVICE: return "mo_gemm_int8x5";
    default: return "mo_gemm_f64";
  }
}

const char *fused_timer_name(qmckl_precision_device form) {
  switch (form) {
    case QMCKL_PRECISION_FP16X2_DEVICE: return "ao_mo_fused_fp16x2";
    case QMCKL_PRECISION_BF16X2_DEVICE: return "ao_mo_fused_bf16x2";
    default: return "ao_mo_fused_bf16x3";
  }
}

// The contraction kernel that the context's choice (QMCKL_GEMM_TILE_*) gives
// form: FUSED only for the split forms with MOs in their given order and at
// most kFusedMaxLd AOs (AUTO's choice elsewhere); AUTO: LIVE where blocks
// can be dead (a column plan, or 1024 AOs or more), DENSE otherwise.
// Measured on MI355X: water64 (4,160 AOs, planned columns) 2.38 ms with the
// live kernel against 2.97 with the dense one; the 48-molecule ragged batch
// (30-620 AOs, dense coefficients) 144 M points/s with the dense kernel
// against 123 M with the live one.
int32_t resolve_gemm_tile(int32_t choice, qmckl_precision_device form, bool planned, int64_t ld) {
  const bool fusable = !planned && ld <= kFusedMaxLd &&
                       (form == QMCKL_PRECISION_FP16X2_DEVICE || form == QMCKL_PRECISION_BF16X2_DEVICE ||
                        form == QMCKL_PRECISION_BF16X3_DEVICE);
  if (choice == QMCKL_GEMM_TILE_FUSED_DEVICE) choice = fusable ? choice : QMCKL_GEMM_TILE_AUTO_DEVICE;
  if (choice != QMCKL_GEMM_TILE_AUTO_DEVICE) return choice;
  return planned || ld >= 1024 ? QMCKL_GEMM_TILE_LIVE_DEVICE : QMCKL_GEMM_TILE_DENSE_DEVICE;
}

// Mixed forms: the FP64 AO assembly stored as FP32 in the blocked operand
// layout with row maxima and nonzero flags, then the split MFMA contraction
// into dst; or, for the fused kernel, both in one launch.
qmckl_exit_code_device compute_mo_vgl_mixed(qmckl_context_device context, Context *c, double *dst,
                                            const char *fn) {
  if (!c->ao.provided || c->point.num <= 0 || c->point.coord == nullptr || !c->nucleus.provided)
    return fail(context, QMCKL_NOT_PROVIDED_DEVICE, "qmckl_ao_basis", nullptr);
  const qmckl_precision_device form = c->mo.precision;
  const int64_t P = c->point.num, A = c->ao.ao_num, M = c->mo.mo_num;
  const int64_t ld = round_up(A, 64);
  const int nkb = (int)(ld / 32);
  const int64_t rows = 5 * round_up(P, 64);  // (component, point) rows, points padded to 64
  AoBasis &ao = c->ao;
  MoBasis &mo = c->mo;
  const bool i8 = form == QMCKL_PRECISION_INT8X5_DEVICE;
  // the split forms use the column plan of their workgroup width; the FP32
  // form keeps the MO order
  ColPlan plan;
  hipError_t e = hipSuccess;
  if (form != QMCKL_PRECISION_FP32_DEVICE) e = col_plan_for(c, split_wg_cols(form, M), &plan);
  if (e != hipSuccess) return hip_check(context, e, fn);
  const int32_t gemm_tile = resolve_gemm_tile(mo.tile_gemm, form, plan.col_mo != nullptr, ld);
  const bool fused = gemm_tile == QMCKL_GEMM_TILE_FUSED_DEVICE;
  if (!fused && (!grow_bytes(c, reinterpret_cast<void **>(&ao.ao_f32), &ao.ao_f32_cap,
                             (int64_t)sizeof(float) * rows * ld) ||
                 !grow_bytes(c, reinterpret_cast<void **>(&ao.ao_f32_rowmax), &ao.ao_f32_rowmax_cap,
                             (int64_t)sizeof(float) * rows) ||
                 (i8 && !grow_bytes(c, reinterpret_cast<void **>(&ao.ao_f32_lo), &ao.ao_f32_lo_cap,
                                    (int64_t)sizeof(float) * rows * ld))))
    return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
  uint8_t *flags = nullptr;
  if (!fused && nkb <= kMaxFlagBlocks) {
    if (!grow_bytes(c, reinterpret_cast<void **>(&ao.ao_f32_flags), &ao.ao_f32_flags_cap, rows / 64 * nkb))
      return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
    flags = ao.ao_f32_flags;
  }
  // Tiles of spatially close points screen out far more 64-AO stages per
  // tile; worth its four small launches for large bases and many points
  // (water64 at 128 walkers: 94% -> 74% of the stages live).
  const int32_t *perm = nullptr;
  if (i8 && (mo.coef_i8 == nullptr || mo.coef_split_form != form)) {
    // the coefficients cut into int8 slices, rebuilt whenever the prepared
    // form or the MO set changed
    dfree(c, mo.coef_i8);
    mo.coef_i8 = static_cast<int8_t *>(dmalloc(c, slice_coef_i8_bytes(M, ld)));
    if (!drealloc(c, reinterpret_cast<void **>(&mo.coef_colinv), sizeof(float) * M) || !mo.coef_i8)
      return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
    e = launch_slice_coef_i8(c->stream, M, A, mo.coefficient, ld, mo.coef_i8, mo.coef_colinv, plan.col_mo);
    if (e == hipSuccess) mo.coef_split_form = form;
  } else if (!i8 && (mo.coef_split == nullptr || mo.coef_split_form != form)) {
    dfree(c, mo.coef_split);
    mo.coef_split = static_cast<uint16_t *>(
        dmalloc(c, sizeof(uint16_t) * split_terms(form) * split_cols_padded(form, M) * ld));
    if (!drealloc(c, reinterpret_cast<void **>(&mo.coef_colinv), sizeof(float) * M) || !mo.coef_split)
      return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
    e = launch_split_coef(c->stream, form, M, A, mo.coefficient, ld, mo.coef_split, mo.coef_colinv, plan.col_mo);
    if (e == hipSuccess) mo.coef_split_form = form;
  }
  if (e == hipSuccess && (flags != nullptr || (fused && nkb <= kMaxFlagBlocks))) {
    e = tile_point_order(c, &perm);
    if (e == hipErrorOutOfMemory) return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
  }
  if (e == hipSuccess && fused) {
    KernelTimer t(c, fused_timer_name(form));
    e = launch_fused_ao_mo(c->stream, form, ao.dev, P, c->point.coord, perm, c->nucleus.coord, ao.ao_factor, ld,
                           mo.coef_split, mo.coef_colinv, M, dst, M);
    t.stop();
    if (e == hipSuccess) e = sync(c);
    return hip_check(context, e, fn);
  }
  // Every contraction but the dense split kernel reads no 32-AO half whose
  // flag is zero, so the AO kernel need not store such halves.
  const bool skip_halves =
      flags != nullptr && (i8 || form == QMCKL_PRECISION_FP32_DEVICE ||
                           (gemm_tile != QMCKL_GEMM_TILE_DENSE_DEVICE && gemm_tile != QMCKL_GEMM_TILE_DENSE64_DEVICE));
  // In an eager batch, split forms of tile width <= 3 leave their AO
  // operand to the batch's single AO launch and their contraction to its
  // persistent launches (wider tiles need more registers than the
  // persistent kernel has to spare).
  const bool collect = c->collect_jobs != nullptr && form != QMCKL_PRECISION_FP32_DEVICE && !i8 &&
                       split_wg_cols(form, M) <= 192;
  if (e == hipSuccess && collect && c->collect_ao != nullptr) {
    AoRowsJob aj{};
    e = make_ao_rows_job(ao.dev, P, c->point.coord, perm, c->nucleus.coord, ao.ao_factor, ld,
                         ao_rows_groups(P, ld, mo.tile_group_waves), ao.ao_f32, ao.ao_f32_rowmax, flags,
                         flags ? nkb : 0, skip_halves, &aj);
    if (e == hipSuccess) c->collect_ao->push_back(aj);
  } else if (e == hipSuccess) {
    KernelTimer t(c, i8 ? "ao_vgl_f32_lo" : "ao_vgl_f32");
    e = launch_ao_rows_f32(c->stream, ao.dev, P, c->point.coord, perm, c->nucleus.coord, ao.ao_factor, ld,
                           ao_rows_groups(P, ld, mo.tile_group_waves), ao.ao_f32, ao.ao_f32_rowmax, flags,
                           flags ? nkb : 0, i8 ? ao.ao_f32_lo : nullptr, skip_halves);
    t.stop();
  }
  if (e == hipSuccess && i8) {
    KernelTimer t(c, gemm_timer_name(form));
    e = launch_mo_gemm_i8(c->stream, P, A, M, ao.ao_f32, ao.ao_f32_lo, ld, mo.coef_i8, ld, mo.coef_colinv, flags,
                          flags ? nkb : 0, perm, dst, M, plan);
    t.stop();
  } else if (e == hipSuccess && collect) {
    MixedGemmJob job;
    e = make_mixed_gemm_job(form, P, A, M, ao.ao_f32, ld, ao.ao_f32_rowmax, mo.coef_split, ld, mo.coef_colinv,
                            flags, flags ? nkb : 0, perm, dst, M, plan, &job, gemm_tile);
    if (e == hipSuccess) c->collect_jobs->push_back(job);
  } else if (e == hipSuccess) {
    KernelTimer t(c, gemm_timer_name(form));
    e = launch_mo_gemm_mixed(c->stream, form, P, A, M, ao.ao_f32, ld, ao.ao_f32_rowmax, mo.coef_split, ld,
                             mo.coef_colinv, flags, flags ? nkb : 0, perm, dst, M, plan, gemm_tile);
    t.stop();
  }
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

// mo_vgl for the current points into dst: FP64 from the context's ao_vgl, or
// one of the mixed forms.
qmckl_exit_code_device compute_mo_vgl(qmckl_context_device context, Context *c, double *dst,
                                      const char *fn) {
  if (c->mo.precision != QMCKL_PRECISION_FP64_DEVICE) return compute_mo_vgl_mixed(context, c, dst, fn);
  qmckl_exit_code_device rc = qmckl_provide_ao_basis_ao_vgl_device(context);
  if (rc != QMCKL_SUCCESS_DEVICE) return fail(context, QMCKL_NOT_PROVIDED_DEVICE, "qmckl_ao_basis", nullptr);
  const int64_t P = c->point.num, A = c->ao.ao_num, M = c->mo.mo_num;
  const int nkb = (int)((A + 31) / 32);
  const uint8_t *flags = nkb <= kMaxFlagBlocks ? c->ao.ao_vgl_flags : nullptr;
  hipError_t e;
  {
    KernelTimer t(c, "mo_gemm_f64");
    ColPlan plan;
    e = col_plan_for(c, f64_group(M), &plan);
    if (e == hipSuccess)
      e = launch_mo_gemm_f64_frag(c->stream, 5 * P, A, M, c->ao.ao_vgl, A, c->mo.coef_frag,
                                  c->mo.coef_frag_nk4, dst, M, flags, flags ? nkb : 0, plan);
    t.stop();
  }
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

qmckl_exit_code_device compute_mo_value(qmckl_context_device context, Context *c, double *dst,
                                        const char *fn) {
  const int64_t P = c->point.num, A = c->ao.ao_num, M = c->mo.mo_num;
  hipError_t e;
  // mo_value is always the FP64 value-only contraction, whatever the
  // precision form, so that it does not depend on whether mo_vgl was asked
  // for first; in the FP64 form a current mo_vgl holds the same sums, and
  // its value rows are copied (src/qmckl_mo.cpp:230-256)
  if (c->mo.precision == QMCKL_PRECISION_FP64_DEVICE && c->mo.mo_vgl_date != 0 &&
      c->mo.mo_vgl_date == c->point.date && c->mo.mo_vgl) {
    e = launch_strided_copy(c->stream, P, M, c->mo.mo_vgl, 5 * M, dst, M);
  } else {
    qmckl_exit_code_device rc = qmckl_provide_ao_basis_ao_value_device(context);
    if (rc != QMCKL_SUCCESS_DEVICE)
      return fail(context, QMCKL_NOT_PROVIDED_DEVICE, "qmckl_ao_basis_ao_value_device", nullptr);
    KernelTimer t(c, "mo_value_gemm_f64");
    ColPlan plan;
    e = col_plan_for(c, f64_group(M), &plan);
    if (e == hipSuccess)
      e = launch_mo_gemm_f64_frag(c->stream, P, A, M, c->ao.ao_value, A, c->mo.coef_frag,
                                  c->mo.coef_frag_nk4, dst, M, nullptr, 0, plan);
    t.stop();
  }
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

}  // namespace

extern "C" {

bool qmckl_mo_basis_provided_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  return c ? c->mo.provided : false;
}

qmckl_exit_code_device qmckl_get_mo_basis_mo_num_device(qmckl_context_device context, int64_t *mo_num) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  if (c->mo.uninitialized & 1)
    return fail(context, QMCKL_NOT_PROVIDED_DEVICE, "qmckl_get_mo_basis_mo_num", nullptr);
  if (!mo_num) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, "qmckl_get_mo_basis_mo_num", "NULL pointer");
  *mo_num = c->mo.mo_num;
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_finalize_mo_basis_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  c->basis_generation += 1;
  const char *fn = "qmckl_finalize_mo_basis_device";
  if (c->mo.coefficient == nullptr) return fail(context, QMCKL_NOT_PROVIDED_DEVICE, fn, nullptr);
  const int64_t A = c->ao.ao_num, M = c->mo.mo_num;
  if (!drealloc(c, reinterpret_cast<void **>(&c->mo.coefficient_t), sizeof(double) * A * M))
    return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
  // the same coefficients in the fragment order of the FP64 contraction
  const int64_t frag_len = round_up(M, 64 * f64_frag_cbw(M)) * round_up(A, 16);
  if (!drealloc(c, reinterpret_cast<void **>(&c->mo.coef_frag), sizeof(double) * frag_len))
    return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
  ColPlan plan;
  hipError_t e = map_coef_zeros(c);
  if (e == hipSuccess) e = col_plan_for(c, f64_group(M), &plan);
  if (e == hipSuccess) e = launch_transpose(c->stream, M, A, c->mo.coefficient, c->mo.coefficient_t);
  if (e == hipSuccess)
    e = launch_frag_b_f64(c->stream, A, M, c->mo.coefficient_t, M, c->mo.coef_frag, &c->mo.coef_frag_nk4,
                          plan.col_mo);
  if (e == hipSuccess) e = sync(c);
  if (e != hipSuccess) return hip_check(context, e, fn);
  c->mo.mo_vgl_date = c->mo.mo_value_date = 0;
  c->mo.coef_split_form = QMCKL_PRECISION_FP64_DEVICE;  // rebuilt on next mixed use
  c->mo.adj_cfrag_ok = false;                            // and on next adjoint
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_set_mo_basis_mo_num_device(qmckl_context_device context, int64_t mo_num) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const int32_t mask = 1;
  if (!(c->mo.uninitialized & mask))
    return fail(context, QMCKL_ALREADY_SET_DEVICE, "qmckl_set_mo_basis_mo_num_device", nullptr);
  if (mo_num <= 0)
    return fail(context, QMCKL_INVALID_ARG_2_DEVICE, "qmckl_set_mo_basis_mo_num_device", "mo_num <= 0");
  c->mo.mo_num = mo_num;
  c->mo.uninitialized &= ~mask;
  c->mo.provided = (c->mo.uninitialized == 0);
  if (c->mo.provided) return qmckl_finalize_mo_basis_device(context);
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_set_mo_basis_coefficient_device(qmckl_context_device context,
                                                             double *coefficient) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_set_mo_basis_coefficient_device";
  const int32_t mask = 1 << 1;
  if (!(c->mo.uninitialized & mask)) return fail(context, QMCKL_ALREADY_SET_DEVICE, fn, nullptr);
  if (coefficient == nullptr) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "NULL pointer");
  // The reference copies ao_num*mo_num values unchecked (src/qmckl_mo.cpp:558-569).
  if (c->ao.ao_num <= 0) return fail(context, QMCKL_FAILURE_DEVICE, fn, "ao_num is not set");
  if (c->mo.mo_num <= 0) return fail(context, QMCKL_FAILURE_DEVICE, fn, "mo_num is not set");
  if (!has_device(c)) return fail(context, QMCKL_FAILURE_DEVICE, fn, "no device bound");
  const int64_t n = c->ao.ao_num * c->mo.mo_num;
  if (!drealloc(c, reinterpret_cast<void **>(&c->mo.coefficient), sizeof(double) * n))
    return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
  hipError_t e = hipMemcpyAsync(c->mo.coefficient, coefficient, sizeof(double) * n,
                                hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = sync(c);
  if (e != hipSuccess) return hip_check(context, e, fn);
  c->mo.uninitialized &= ~mask;
  c->mo.provided = (c->mo.uninitialized == 0);
  if (c->mo.provided) return qmckl_finalize_mo_basis_device(context);
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_provide_mo_basis_mo_vgl_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_provide_mo_basis_mo_vgl_device";
  qmckl_exit_code_device rc = ready(context, c, fn);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  if (c->point.date > c->mo.mo_vgl_date || c->mo.mo_vgl_date == 0) {
    if (!grow(c, &c->mo.mo_vgl, &c->mo.mo_vgl_cap, 5 * c->mo.mo_num * c->point.num))
      return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
    rc = compute_mo_vgl(context, c, c->mo.mo_vgl, fn);
    if (rc != QMCKL_SUCCESS_DEVICE) return rc;
    c->mo.mo_vgl_date = c->date;
  }
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_provide_mo_basis_mo_value_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_provide_mo_basis_mo_value_device";
  qmckl_exit_code_device rc = ready(context, c, fn);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  if (c->point.date > c->mo.mo_value_date || c->mo.mo_value_date == 0) {
    if (!grow(c, &c->mo.mo_value, &c->mo.mo_value_cap, c->mo.mo_num * c->point.num))
      return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
    rc = compute_mo_value(context, c, c->mo.mo_value, fn);
    if (rc != QMCKL_SUCCESS_DEVICE) return rc;
    c->mo.mo_value_date = c->date;
  }
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_get_mo_basis_mo_vgl_device(qmckl_context_device context, double *mo_vgl,
                                                        int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_get_mo_basis_mo_vgl_device";
  if (!mo_vgl) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "NULL pointer");
  qmckl_exit_code_device rc = qmckl_provide_mo_basis_mo_vgl_device(context);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  const int64_t sze = 5 * c->point.num * c->mo.mo_num;
  if (size_max < sze) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "input array too small");
  hipError_t e = hipMemcpyAsync(mo_vgl, c->mo.mo_vgl, sizeof(double) * sze, hipMemcpyDeviceToDevice,
                                c->stream);
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

qmckl_exit_code_device qmckl_get_mo_basis_mo_vgl_inplace_device(qmckl_context_device context,
                                                                double *mo_vgl, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_get_mo_basis_mo_vgl_inplace_device";
  if (!mo_vgl) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "NULL pointer");
  const int64_t sze = 5 * c->mo.mo_num * c->point.num;
  if (size_max < sze) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "input array too small");
  qmckl_exit_code_device rc = qmckl_context_touch_device(context);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  rc = ready(context, c, fn);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  return compute_mo_vgl(context, c, mo_vgl, fn);
}

qmckl_exit_code_device qmckl_get_mo_basis_mo_value_device(qmckl_context_device context,
                                                          double *mo_value, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_get_mo_basis_mo_value_device";
  if (!mo_value) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "NULL pointer");
  qmckl_exit_code_device rc = qmckl_provide_mo_basis_mo_value_device(context);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  const int64_t sze = c->point.num * c->mo.mo_num;
  if (size_max < sze) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "input array too small");
  hipError_t e = hipMemcpyAsync(mo_value, c->mo.mo_value, sizeof(double) * sze,
                                hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

qmckl_exit_code_device qmckl_get_mo_basis_mo_value_inplace_device(qmckl_context_device context,
                                                                  double *mo_value, int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_get_mo_basis_mo_value_inplace_device";
  if (!mo_value) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "NULL pointer");
  const int64_t sze = c->mo.mo_num * c->point.num;
  if (size_max < sze) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "input array too small");
  qmckl_exit_code_device rc = qmckl_context_touch_device(context);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  rc = ready(context, c, fn);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  return compute_mo_value(context, c, mo_value, fn);
}

qmckl_exit_code_device qmckl_compute_mo_basis_mo_vgl_device(qmckl_context_device context,
                                                            int64_t ao_num, int64_t mo_num,
                                                            int64_t point_num, double *coefficient_t,
                                                            double *ao_vgl, double *mo_vgl) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_compute_mo_basis_mo_vgl_device";
  if (ao_num <= 0) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "ao_num <= 0");
  if (mo_num <= 0) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "mo_num <= 0");
  if (point_num < 0) return fail(context, QMCKL_INVALID_ARG_4_DEVICE, fn, "point_num < 0");
  if (!coefficient_t) return fail(context, QMCKL_INVALID_ARG_5_DEVICE, fn, "NULL pointer");
  if (!ao_vgl) return fail(context, QMCKL_INVALID_ARG_6_DEVICE, fn, "NULL pointer");
  if (!mo_vgl) return fail(context, QMCKL_INVALID_ARG_7_DEVICE, fn, "NULL pointer");
  if (!has_device(c)) return fail(context, QMCKL_FAILURE_DEVICE, fn, "no device bound");
  hipError_t e = launch_mo_gemm_f64(c->stream, 5 * point_num, ao_num, mo_num, ao_vgl, ao_num,
                                    coefficient_t, mo_num, mo_vgl, mo_num);
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

qmckl_exit_code_device qmckl_compute_mo_basis_mo_value_device(qmckl_context_device context,
                                                              int64_t ao_num, int64_t mo_num,
                                                              int64_t point_num, double *coefficient_t,
                                                              double *ao_value, double *mo_value) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_compute_mo_basis_mo_value_device";
  if (ao_num <= 0) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "ao_num <= 0");
  if (mo_num <= 0) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "mo_num <= 0");
  if (point_num < 0) return fail(context, QMCKL_INVALID_ARG_4_DEVICE, fn, "point_num < 0");
  if (!coefficient_t) return fail(context, QMCKL_INVALID_ARG_5_DEVICE, fn, "NULL pointer");
  if (!ao_value) return fail(context, QMCKL_INVALID_ARG_6_DEVICE, fn, "NULL pointer");
  if (!mo_value) return fail(context, QMCKL_INVALID_ARG_7_DEVICE, fn, "NULL pointer");
  if (!has_device(c)) return fail(context, QMCKL_FAILURE_DEVICE, fn, "no device bound");
  hipError_t e = launch_mo_gemm_f64(c->stream, point_num, ao_num, mo_num, ao_value, ao_num,
                                    coefficient_t, mo_num, mo_value, mo_num);
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

qmckl_exit_code_device qmckl_set_mo_basis_precision_device(qmckl_context_device context,
                                                           qmckl_precision_device form) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  if (form < QMCKL_PRECISION_FP64_DEVICE || form > QMCKL_PRECISION_INT8X5_DEVICE)
    return fail(context, QMCKL_INVALID_ARG_2_DEVICE, "qmckl_set_mo_basis_precision_device",
                "unknown precision form");
  if (form != c->mo.precision) {
    c->mo.precision = form;
    c->mo.mo_vgl_date = c->mo.mo_value_date = 0;
  }
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_get_mo_basis_precision_device(qmckl_context_device context,
                                                           qmckl_precision_device *form) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  if (!form)
    return fail(context, QMCKL_INVALID_ARG_2_DEVICE, "qmckl_get_mo_basis_precision_device", "NULL pointer");
  *form = c->mo.precision;
  return QMCKL_SUCCESS_DEVICE;
}

// MO subset: keeps the MOs i with keep[i] != 0 (host array), in order
// (src/qmckl_mo.cpp:31-106), then re-finalizes the MO basis.
qmckl_exit_code_device qmckl_mo_basis_select_mo_device(qmckl_context_device context, int32_t *keep,
                                                       int64_t size_max) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_mo_basis_select_mo_device";
  if (!c->mo.provided) return fail(context, QMCKL_NOT_PROVIDED_DEVICE, fn, nullptr);
  if (keep == nullptr) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "NULL pointer");
  const int64_t M = c->mo.mo_num, A = c->ao.ao_num;
  if (size_max < M) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "Array too small: expected mo_num.");
  int64_t kept = 0;
  for (int64_t i = 0; i < M; ++i) kept += keep[i] != 0;
  if (kept == 0) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "no MO selected");
  if (!has_device(c)) return fail(context, QMCKL_FAILURE_DEVICE, fn, "no device bound");
  double *nc = static_cast<double *>(dmalloc(c, sizeof(double) * A * kept));
  if (!nc) return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
  // copy runs of consecutive kept rows of coefficient [M][A]
  hipError_t e = hipSuccess;
  int64_t dst = 0;
  for (int64_t i = 0; i < M && e == hipSuccess;) {
    if (keep[i] == 0) {
      ++i;
      continue;
    }
    int64_t j = i;
    while (j < M && keep[j] != 0) ++j;
    e = hipMemcpyAsync(nc + dst * A, c->mo.coefficient + i * A, sizeof(double) * A * (j - i),
                       hipMemcpyDeviceToDevice, c->stream);
    dst += j - i;
    i = j;
  }
  if (e == hipSuccess) e = sync(c);
  if (e != hipSuccess) {
    dfree(c, nc);
    return hip_check(context, e, fn);
  }
  dfree(c, c->mo.coefficient);
  c->mo.coefficient = nc;
  c->mo.mo_num = kept;
  return qmckl_finalize_mo_basis_device(context);
}

}  // extern "C"

namespace {

// The split-form contractions a batch collected, grouped by (form, tile
// width): one persistent launch per group, items numbered from 0 within each.
struct PersistentGroup {
  int form, cbw, dense;
  size_t first, n;
  int64_t total;
};
struct PersistentPlan {
  std::vector<MixedGemmJob> sorted;
  std::vector<PersistentGroup> groups;
  size_t table_bytes() const { return sizeof(MixedGemmJob) * sorted.size() + sizeof(int) * groups.size(); }
};

PersistentPlan plan_persistent(const std::vector<MixedGemmJob> &jobs, const std::vector<Context *> &job_ctx) {
  PersistentPlan pp;
  std::vector<size_t> order(jobs.size());
  for (size_t i = 0; i < jobs.size(); ++i) order[i] = i;
  // one job per collecting context, in context order: its form is that
  // context's precision
  auto key = [&](size_t i) { return ((int)job_ctx[i]->mo.precision * 8 + jobs[i].cbw) * 2 + jobs[i].dense; };
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return key(x) < key(y); });
  for (size_t k = 0; k < order.size(); ++k) {
    MixedGemmJob j = jobs[order[k]];
    const int form = (int)job_ctx[order[k]]->mo.precision;
    if (pp.groups.empty() || pp.groups.back().form != form || pp.groups.back().cbw != j.cbw ||
        pp.groups.back().dense != j.dense)
      pp.groups.push_back(PersistentGroup{form, j.cbw, j.dense, k, 0, 0});
    PersistentGroup &g = pp.groups.back();
    j.item0 = g.total;
    g.total += j.items;
    g.n += 1;
    pp.sorted.push_back(j);
  }
  return pp;
}

// The AO operands a batch collected, largest first (so that the small ones
// fill the end of the launch), with their first wavefronts; returns the
// launch's wavefronts.
int64_t plan_ao_jobs(std::vector<AoRowsJob> &ao) {
  std::stable_sort(ao.begin(), ao.end(), [](const AoRowsJob &x, const AoRowsJob &y) {
    return x.npt * x.ld > y.npt * y.ld;
  });
  int64_t total = 0;
  for (AoRowsJob &j : ao) {
    j.wave0 = total;
    total += j.waves;
  }
  return total;
}

bool same_ao_jobs(const std::vector<AoRowsJob> &x, const std::vector<AoRowsJob> &y) {
  if (x.size() != y.size()) return false;
  for (size_t k = 0; k < x.size(); ++k)
    if (std::memcmp(&x[k], &y[k], sizeof(AoRowsJob)) != 0) return false;
  return true;
}

bool same_job(const MixedGemmJob &x, const MixedGemmJob &y) {
  return x.P == y.P && x.npt == y.npt && x.K == y.K && x.Ncol == y.Ncol && x.A == y.A && x.lda == y.lda &&
         x.rowmax == y.rowmax && x.Bs == y.Bs && x.ldb == y.ldb && x.colinv == y.colinv && x.flags == y.flags &&
         x.nkb == y.nkb && x.perm == y.perm && x.C == y.C && x.ldc == y.ldc && x.plan.col_mo == y.plan.col_mo &&
         x.plan.blk == y.plan.blk && x.plan.nw == y.plan.nw && x.plan.nblk == y.plan.nblk && x.cbw == y.cbw &&
         x.dense == y.dense &&
         x.item0 == y.item0 && x.gx == y.gx && x.items == y.items;
}

// Enqueues on c0's stream, after every collecting context's own work (its
// point order and coefficient preparation), the batch's AO operands in one
// launch (ao_table: njobs_ao jobs, ao_waves wavefronts) and then the plan's
// persistent contractions; table holds the sorted jobs followed by one int
// counter per group.  Both tables are already on the device.
// Eager (not captured): a context whose stream has nothing pending is not
// waited for (the event record and wait of 47 contexts took about 0.6 ms of
// host time, during which the GPU idled).
hipError_t launch_persistent(Context *c0, const PersistentPlan &pp, MixedGemmJob *table,
                             const std::vector<Context *> &job_ctx, const AoRowsJob *ao_table, int njobs_ao,
                             int64_t ao_waves, bool eager) {
  int *counters = reinterpret_cast<int *>(reinterpret_cast<char *>(table) + sizeof(MixedGemmJob) * pp.sorted.size());
  hipError_t e = hipSuccess;
  for (Context *c : job_ctx) {
    if (e != hipSuccess || c == c0) continue;
    if (eager) {
      const hipError_t q = hipStreamQuery(c->stream);
      if (q == hipSuccess) continue;
      if (q != hipErrorNotReady) {
        e = q;
        continue;
      }
    }
    if (c->batch_event == nullptr) e = hipEventCreateWithFlags(&c->batch_event, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(c->batch_event, c->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(c0->stream, c->batch_event, 0);
  }
  if (e == hipSuccess && njobs_ao > 0) {
    KernelTimer t(c0, "ao_rows_batch");
    e = launch_ao_rows_batch(c0->stream, ao_table, njobs_ao, ao_waves);
    t.stop();
  }
  for (size_t g = 0; g < pp.groups.size() && e == hipSuccess; ++g) {
    KernelTimer t(c0, "mo_gemm_batch_persistent");
    e = launch_mo_gemm_mixed_persistent(c0->stream, pp.groups[g].form, pp.groups[g].cbw, pp.groups[g].dense,
                                        table + pp.groups[g].first,
                                        (int)pp.groups[g].n, pp.groups[g].total, counters + g);
    t.stop();
  }
  return e;
}

// Eager batch: the tables go to c0's memory (their host copies stay in c0
// for a graph built from this batch), the launches complete before
// returning.
hipError_t run_persistent_contractions(Context *c0, const std::vector<MixedGemmJob> &jobs,
                                       const std::vector<Context *> &job_ctx, std::vector<AoRowsJob> ao_jobs) {
  const PersistentPlan pp = plan_persistent(jobs, job_ctx);
  const int64_t ao_waves = plan_ao_jobs(ao_jobs);
  const size_t gemm_bytes = sizeof(MixedGemmJob) * pp.sorted.size();
  const size_t ao_bytes = sizeof(AoRowsJob) * ao_jobs.size();
  if (!grow_bytes(c0, reinterpret_cast<void **>(&c0->batch_jobs), &c0->batch_jobs_cap,
                  (int64_t)pp.table_bytes()) ||
      !grow_bytes(c0, reinterpret_cast<void **>(&c0->batch_ao_jobs), &c0->batch_ao_jobs_cap,
                  (int64_t)std::max<size_t>(ao_bytes, 1)))
    return hipErrorOutOfMemory;
  // staged through pinned host memory: a copy from pageable memory is
  // staged by the runtime (about 0.3 ms for a 48-molecule table on MI355X)
  const size_t bytes = gemm_bytes + ao_bytes;
  hipError_t e = hipSuccess;
  if (c0->batch_jobs_pinned_cap < bytes) {
    if (c0->batch_jobs_pinned) (void)hipHostFree(c0->batch_jobs_pinned);
    c0->batch_jobs_pinned = nullptr;
    c0->batch_jobs_pinned_cap = 0;
    e = hipHostMalloc(&c0->batch_jobs_pinned, bytes, hipHostMallocDefault);
    if (e != hipSuccess) return e;
    c0->batch_jobs_pinned_cap = bytes;
  }
  char *staging = static_cast<char *>(c0->batch_jobs_pinned);
  std::memcpy(staging, pp.sorted.data(), gemm_bytes);
  if (ao_bytes) std::memcpy(staging + gemm_bytes, ao_jobs.data(), ao_bytes);
  e = hipMemcpyAsync(c0->batch_jobs, staging, gemm_bytes, hipMemcpyHostToDevice, c0->stream);
  if (e == hipSuccess && ao_bytes)
    e = hipMemcpyAsync(c0->batch_ao_jobs, staging + gemm_bytes, ao_bytes, hipMemcpyHostToDevice, c0->stream);
  if (e == hipSuccess)
    e = launch_persistent(c0, pp, c0->batch_jobs, job_ctx, c0->batch_ao_jobs, (int)ao_jobs.size(), ao_waves,
                          true);
  // the job tables are read by the launches: wait before they can be reused
  if (e == hipSuccess) e = hipStreamSynchronize(c0->stream);
  if (e == hipSuccess) {
    c0->batch_jobs_host = pp.sorted;
    c0->batch_ao_host = ao_jobs;
  }
  return e;
}

}  // namespace

extern "C" {

// Batching form: each context's chain is enqueued on its own stream with the
// chain's synchronisations deferred, then every stream is waited for once.
// The contractions of split-form contexts (tile width <= 3) run as one
// persistent launch per form and width over all of them.
qmckl_exit_code_device qmckl_get_mo_basis_mo_vgl_batch_device(int64_t count,
                                                              const qmckl_context_device *contexts,
                                                              double *const *mo_vgl,
                                                              const int64_t *size_max) {
  const char *fn = "qmckl_get_mo_basis_mo_vgl_batch_device";
  if (count <= 0) return QMCKL_INVALID_ARG_1_DEVICE;
  if (contexts == nullptr) return QMCKL_INVALID_ARG_2_DEVICE;
  if (mo_vgl == nullptr) return QMCKL_INVALID_ARG_3_DEVICE;
  if (size_max == nullptr) return QMCKL_INVALID_ARG_4_DEVICE;
  Context *c0 = ctx_of(contexts[0]);
  if (!c0) return QMCKL_INVALID_CONTEXT_DEVICE;
  for (int64_t i = 0; i < count; ++i) {
    Context *c = ctx_of(contexts[i]);
    if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
    if (c->device_id != c0->device_id)
      return fail(contexts[i], QMCKL_INVALID_ARG_2_DEVICE, fn, "contexts on different devices");
    if (mo_vgl[i] == nullptr) return fail(contexts[i], QMCKL_INVALID_ARG_3_DEVICE, fn, "NULL pointer");
    if (size_max[i] < 5 * c->mo.mo_num * c->point.num)
      return fail(contexts[i], QMCKL_INVALID_ARG_4_DEVICE, fn, "output array too small");
  }
  for (int64_t i = 0; i < count; ++i) {
    qmckl_exit_code_device rc = ready(contexts[i], ctx_of(contexts[i]), fn);
    if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  }
  qmckl_exit_code_device first = QMCKL_SUCCESS_DEVICE;
  int64_t enqueued = 0;
  std::vector<MixedGemmJob> jobs;
  std::vector<AoRowsJob> ao_jobs;
  std::vector<Context *> job_ctx;
  c0->batch_jobs_host.clear();
  c0->batch_ao_host.clear();
  for (; enqueued < count; ++enqueued) {
    Context *c = ctx_of(contexts[enqueued]);
    c->defer_sync = true;
    const size_t before = jobs.size();
    c->collect_jobs = &jobs;
    c->collect_ao = &ao_jobs;
    qmckl_exit_code_device rc = qmckl_context_touch_device(contexts[enqueued]);
    if (rc == QMCKL_SUCCESS_DEVICE) rc = compute_mo_vgl(contexts[enqueued], c, mo_vgl[enqueued], fn);
    c->collect_jobs = nullptr;
    c->collect_ao = nullptr;
    c->defer_sync = false;
    if (jobs.size() > before) job_ctx.push_back(c);
    if (rc != QMCKL_SUCCESS_DEVICE) {
      first = rc;
      ++enqueued;
      break;
    }
  }
  if (first == QMCKL_SUCCESS_DEVICE && !jobs.empty()) {
    const hipError_t e = run_persistent_contractions(c0, jobs, job_ctx, ao_jobs);
    if (e != hipSuccess) first = hip_check(contexts[0], e, fn);
  }
  for (int64_t i = 0; i < enqueued; ++i) {
    Context *c = ctx_of(contexts[i]);
    const qmckl_exit_code_device rc = hip_check(contexts[i], sync(c), fn);
    if (first == QMCKL_SUCCESS_DEVICE) first = rc;
  }
  return first;
}

}  // extern "C"

namespace {

// Autotuned tile choices per (device, points, AOs, MOs, nuclei, form), shared
// by every context of the process with that shape.
struct TuneKey {
  int device;
  int64_t P, A, M, N;
  int form;
  bool fused;  // the context keeps the fused contraction kernel
  bool operator==(const TuneKey &o) const {
    return device == o.device && P == o.P && A == o.A && M == o.M && N == o.N && form == o.form &&
           fused == o.fused;
  }
};
struct TuneKeyHash {
  size_t operator()(const TuneKey &k) const {
    size_t h = (size_t)k.device * 1000003u ^ (size_t)k.form ^ ((size_t)k.fused << 20);
    for (int64_t v : {k.P, k.A, k.M, k.N}) h = h * 1000003u ^ (size_t)v;
    return h;
  }
};
struct TuneChoice {
  int32_t point_order, group_waves, gemm_tile;
};
std::mutex g_tune_mutex;
std::unordered_map<TuneKey, TuneChoice, TuneKeyHash> &tune_cache() {
  static std::unordered_map<TuneKey, TuneChoice, TuneKeyHash> m;
  return m;
}

}  // namespace

extern "C" {

qmckl_exit_code_device qmckl_set_mo_basis_tile_config_device(qmckl_context_device context, int32_t point_order,
                                                             int32_t group_waves) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_set_mo_basis_tile_config_device";
  if (point_order < -1 || point_order > 1) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "point_order not in {-1, 0, 1}");
  if (group_waves < 0 || group_waves > 1024) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "group_waves not in [0, 1024]");
  c->mo.tile_point_order = point_order;
  c->mo.tile_group_waves = group_waves;
  c->mo.tile_tuned = false;
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_set_mo_basis_gemm_tile_device(qmckl_context_device context, int32_t gemm_tile) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  if (gemm_tile < QMCKL_GEMM_TILE_AUTO_DEVICE || gemm_tile > QMCKL_GEMM_TILE_DENSE64_DEVICE)
    return fail(context, QMCKL_INVALID_ARG_2_DEVICE, "qmckl_set_mo_basis_gemm_tile_device",
                "gemm_tile not one of QMCKL_GEMM_TILE_*");
  c->mo.tile_gemm = gemm_tile;
  c->mo.tile_tuned = false;
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_get_mo_basis_gemm_tile_device(qmckl_context_device context, int32_t *gemm_tile) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  if (gemm_tile == nullptr)
    return fail(context, QMCKL_INVALID_ARG_2_DEVICE, "qmckl_get_mo_basis_gemm_tile_device", "NULL pointer");
  *gemm_tile = c->mo.tile_gemm;
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_get_mo_basis_tile_config_device(qmckl_context_device context, int32_t *point_order,
                                                             int32_t *group_waves) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_get_mo_basis_tile_config_device";
  if (point_order == nullptr) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "NULL pointer");
  if (group_waves == nullptr) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "NULL pointer");
  *point_order = c->mo.tile_point_order;
  *group_waves = c->mo.tile_group_waves;
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_get_mo_basis_column_plan_device(qmckl_context_device context, int32_t *col_mo,
                                                             double *live) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_get_mo_basis_column_plan_device";
  if (!c->mo.provided || c->mo.coef_frag == nullptr) return fail(context, QMCKL_NOT_PROVIDED_DEVICE, fn, nullptr);
  const int64_t M = c->mo.mo_num;
  ColPlan plan;
  hipError_t e = col_plan_for(c, f64_group(M), &plan);
  if (e != hipSuccess) return hip_check(context, e, fn);
  double lv = 1.0;
  for (const ColPlanSlot &p : c->mo.col_plans)
    if (p.group == f64_group(M)) lv = p.live;
  if (live != nullptr) *live = lv;
  if (col_mo == nullptr) return QMCKL_SUCCESS_DEVICE;
  if (plan.col_mo == nullptr) {
    for (int64_t j = 0; j < M; ++j) col_mo[j] = (int32_t)j;
    return QMCKL_SUCCESS_DEVICE;
  }
  e = hipMemcpyAsync(col_mo, plan.col_mo, sizeof(int32_t) * (size_t)M, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = sync(c);
  return hip_check(context, e, fn);
}

qmckl_exit_code_device qmckl_mo_basis_autotune_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_mo_basis_autotune_device";
  qmckl_exit_code_device rc = ready(context, c, fn);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  MoBasis &mo = c->mo;
  if (mo.precision == QMCKL_PRECISION_FP64_DEVICE) {  // nothing to choose for the FP64 form
    mo.tile_tuned = true;
    return QMCKL_SUCCESS_DEVICE;
  }
  if (c->point.num <= 0 || c->point.coord == nullptr) return fail(context, QMCKL_NOT_PROVIDED_DEVICE, fn, "no points set");
  const TuneKey key{c->device_id, c->point.num, c->ao.ao_num, mo.mo_num, c->nucleus.num, (int)mo.precision,
                    mo.tile_gemm == QMCKL_GEMM_TILE_FUSED_DEVICE};
  {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    auto it = tune_cache().find(key);
    if (it != tune_cache().end()) {
      mo.tile_point_order = it->second.point_order;
      mo.tile_group_waves = it->second.group_waves;
      mo.tile_gemm = it->second.gemm_tile;
      mo.tile_tuned = true;
      return QMCKL_SUCCESS_DEVICE;
    }
  }
  const int64_t need = 5 * c->point.num * mo.mo_num;
  if (!grow(c, &mo.mo_vgl, &mo.mo_vgl_cap, need)) return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
  mo.mo_vgl_date = 0;  // the timing runs write into the provider's buffer
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  const bool can_order = round_up(c->ao.ao_num, 64) / 32 <= kMaxFlagBlocks;
  // the time of two evaluations of the mixed chain with the current choices
  auto time_chain = [&](float *ms) -> bool {
    rc = compute_mo_vgl_mixed(context, c, mo.mo_vgl, fn);  // warm-up
    if (rc != QMCKL_SUCCESS_DEVICE) return false;
    e = hipEventRecord(e0, c->stream);
    for (int r = 0; r < 2 && rc == QMCKL_SUCCESS_DEVICE; ++r) rc = compute_mo_vgl_mixed(context, c, mo.mo_vgl, fn);
    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    return e == hipSuccess && rc == QMCKL_SUCCESS_DEVICE;
  };
  // point order x AO wavefronts with the default contraction, then the
  // contraction kernels with the best of those (the FP32 form has one)
  float best = -1.0f;
  // a context set to the fused kernel keeps it (the other kernels round
  // differently); its point order and AO wavefronts are timed with it
  const int32_t gemm0 = mo.tile_gemm;
  const bool keep_fused = gemm0 == QMCKL_GEMM_TILE_FUSED_DEVICE;
  int32_t best_order = 0, best_waves = 256, best_gemm = keep_fused ? gemm0 : QMCKL_GEMM_TILE_AUTO_DEVICE;
  mo.tile_gemm = best_gemm;
  bool ok = e == hipSuccess;
  for (int32_t order = 0; order <= (can_order ? 1 : 0) && ok; ++order)
    for (int32_t waves : {16, 32, 64, 128, 256}) {
      mo.tile_point_order = order;
      mo.tile_group_waves = waves;
      float ms = 0.0f;
      if (!(ok = time_chain(&ms))) break;
      if (best < 0.0f || ms < best) {
        best = ms;
        best_order = order;
        best_waves = waves;
      }
    }
  if (ok && !keep_fused && mo.precision != QMCKL_PRECISION_FP32_DEVICE &&
      mo.precision != QMCKL_PRECISION_INT8X5_DEVICE) {
    mo.tile_point_order = best_order;
    mo.tile_group_waves = best_waves;
    for (int32_t g : {QMCKL_GEMM_TILE_LIVE_DEVICE, QMCKL_GEMM_TILE_LIVE_CACHED_DEVICE, QMCKL_GEMM_TILE_DENSE_DEVICE,
                      QMCKL_GEMM_TILE_LIVE64_DEVICE, QMCKL_GEMM_TILE_DENSE64_DEVICE}) {
      mo.tile_gemm = g;
      float ms = 0.0f;
      if (!(ok = time_chain(&ms))) break;
      if (ms < best) {
        best = ms;
        best_gemm = g;
      }
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (!ok) mo.tile_gemm = gemm0;
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  if (e != hipSuccess) return hip_check(context, e, fn);
  mo.tile_point_order = best_order;
  mo.tile_group_waves = best_waves;
  mo.tile_gemm = best_gemm;
  mo.tile_tuned = true;
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  tune_cache()[key] = {best_order, best_waves, best_gemm};
  return QMCKL_SUCCESS_DEVICE;
}

}  // extern "C"

namespace {

// A captured batch: the graph, the stream it is launched on, and for every
// context what the captured kernels read as arguments (sizes, precision,
// buffer and table addresses), so that a launch after a change that would
// make the recorded arguments wrong is refused instead of run.
struct BatchGraph {
  int device = 0;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<qmckl_context_device> contexts;
  std::vector<std::vector<uint64_t>> signature;
  MixedGemmJob *table = nullptr;  // the persistent launches' job table and counters (owned)
  AoRowsJob *ao_table = nullptr;  // the AO launch's job table (owned)
};

std::mutex g_graph_mutex;
std::unordered_map<int64_t, BatchGraph *> &graphs() {
  static std::unordered_map<int64_t, BatchGraph *> m;
  return m;
}
int64_t g_next_graph = 1;

std::vector<uint64_t> batch_signature(const Context *c, const double *out) {
  auto u = [](const void *p) { return (uint64_t)(uintptr_t)p; };
  const AoBasis &ao = c->ao;
  const MoBasis &mo = c->mo;
  const DeviceBasis &b = ao.dev;
  // entry 6 is the output (the launch reads it back); the kernels' launch
  // arguments that a new basis or other tile choices change come last
  return {(uint64_t)c->point.num, (uint64_t)c->nucleus.num, (uint64_t)ao.ao_num, (uint64_t)mo.mo_num,
          (uint64_t)mo.precision, (uint64_t)mo.coef_split_form, u(out), u(c->point.coord), u(c->nucleus.coord),
          u(ao.ao_factor), u(ao.ao_vgl), u(ao.ao_vgl_flags), u(ao.ao_f32), u(ao.ao_f32_rowmax),
          u(ao.ao_f32_flags), u(ao.point_perm), u(ao.order_ws), u(b.shells), u(b.chunks), u(b.stage_shell),
          u(b.expo), u(b.coef_norm), u(b.nucl_range), u(b.ao_shell), u(b.ao_abc), u(mo.coef_frag),
          u(mo.coef_split), u(mo.coef_colinv), u(mo.col_plans[0].col_mo), u(mo.col_plans[0].blk),
          u(mo.col_plans[1].col_mo), u(mo.col_plans[1].blk), (uint64_t)ao.shell_num, (uint64_t)b.nchunks,
          c->basis_generation, (uint64_t)(uint32_t)mo.tile_point_order, (uint64_t)(uint32_t)mo.tile_group_waves,
          (uint64_t)(uint32_t)mo.tile_gemm, u(ao.ao_f32_lo), u(mo.coef_i8)};
}

}  // namespace

extern "C" {

qmckl_exit_code_device qmckl_mo_basis_mo_vgl_batch_graph_create_device(int64_t count,
                                                                       const qmckl_context_device *contexts,
                                                                       double *const *mo_vgl,
                                                                       const int64_t *size_max,
                                                                       int64_t *graph) {
  const char *fn = "qmckl_mo_basis_mo_vgl_batch_graph_create_device";
  if (graph == nullptr) return QMCKL_INVALID_ARG_5_DEVICE;
  // one eager run: validates the arguments, allocates every buffer and
  // prepares the coefficients, so that the capture below only records
  // kernels
  qmckl_exit_code_device rc = qmckl_get_mo_basis_mo_vgl_batch_device(count, contexts, mo_vgl, size_max);
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  Context *c0 = ctx_of(contexts[0]);
  auto *g = new BatchGraph;
  g->device = c0->device_id;
  g->contexts.assign(contexts, contexts + count);
  hipError_t e = hipSetDevice(g->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
  // the eager run's persistent job table, copied to memory the graph owns
  // (a later batch may reuse the context's table)
  const std::vector<MixedGemmJob> eager_jobs = c0->batch_jobs_host;
  const std::vector<AoRowsJob> eager_ao = c0->batch_ao_host;
  if (e == hipSuccess && !eager_jobs.empty()) {
    const size_t bytes = sizeof(MixedGemmJob) * eager_jobs.size() + 64 * sizeof(int);
    e = hipMalloc(reinterpret_cast<void **>(&g->table), bytes);
    if (e == hipSuccess)
      e = hipMemcpy(g->table, eager_jobs.data(), sizeof(MixedGemmJob) * eager_jobs.size(), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess && !eager_ao.empty()) {
    e = hipMalloc(reinterpret_cast<void **>(&g->ao_table), sizeof(AoRowsJob) * eager_ao.size());
    if (e == hipSuccess)
      e = hipMemcpy(g->ao_table, eager_ao.data(), sizeof(AoRowsJob) * eager_ao.size(), hipMemcpyHostToDevice);
  }
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> joins((size_t)count, nullptr);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
  for (int64_t i = 0; i < count && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&joins[(size_t)i], hipEventDisableTiming);
  bool capturing = false;
  if (e == hipSuccess) {
    e = hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal);
    capturing = e == hipSuccess;
  }
  if (e == hipSuccess) e = hipEventRecord(fork, g->stream);
  // every context's stream joins the capture, enqueues its chain, and is
  // joined back; split-form contractions are collected as in the eager run
  // and recorded as its persistent launches on the first context's stream
  std::vector<MixedGemmJob> jobs;
  std::vector<AoRowsJob> ao_jobs;
  std::vector<Context *> job_ctx;
  for (int64_t i = 0; i < count && e == hipSuccess && rc == QMCKL_SUCCESS_DEVICE; ++i) {
    Context *c = ctx_of(contexts[i]);
    e = hipStreamWaitEvent(c->stream, fork, 0);
    if (e != hipSuccess) break;
    c->defer_sync = true;
    c->capturing = true;
    const size_t before = jobs.size();
    c->collect_jobs = &jobs;
    c->collect_ao = &ao_jobs;
    rc = qmckl_context_touch_device(contexts[i]);
    if (rc == QMCKL_SUCCESS_DEVICE) rc = compute_mo_vgl(contexts[i], c, mo_vgl[i], fn);
    c->collect_jobs = nullptr;
    c->collect_ao = nullptr;
    c->defer_sync = false;
    c->capturing = false;
    if (jobs.size() > before) job_ctx.push_back(c);
    if (rc == QMCKL_SUCCESS_DEVICE) e = hipEventRecord(joins[(size_t)i], c->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(g->stream, joins[(size_t)i], 0);
  }
  hipEvent_t pjoin = nullptr;
  if (e == hipSuccess && rc == QMCKL_SUCCESS_DEVICE && !jobs.empty()) {
    const PersistentPlan pp = plan_persistent(jobs, job_ctx);
    const int64_t ao_waves = plan_ao_jobs(ao_jobs);
    // the capture must record exactly the tables the eager run uploaded
    bool same = g->table != nullptr && pp.sorted.size() == eager_jobs.size() && pp.groups.size() <= 64 &&
                same_ao_jobs(ao_jobs, eager_ao) && (ao_jobs.empty() || g->ao_table != nullptr);
    for (size_t k = 0; same && k < pp.sorted.size(); ++k) same = same_job(pp.sorted[k], eager_jobs[k]);
    if (!same) e = hipErrorInvalidValue;
    c0->capturing = true;  // no kernel timers inside the capture
    if (e == hipSuccess)
      e = launch_persistent(c0, pp, g->table, job_ctx, g->ao_table, (int)ao_jobs.size(), ao_waves, false);
    c0->capturing = false;
    if (e == hipSuccess) e = hipEventCreateWithFlags(&pjoin, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(pjoin, c0->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(g->stream, pjoin, 0);
  }
  if (capturing) {
    const hipError_t e2 = hipStreamEndCapture(g->stream, &g->graph);
    if (e == hipSuccess) e = e2;
  }
  if (e == hipSuccess && rc == QMCKL_SUCCESS_DEVICE) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (fork) (void)hipEventDestroy(fork);
  if (pjoin) (void)hipEventDestroy(pjoin);
  for (hipEvent_t j : joins)
    if (j) (void)hipEventDestroy(j);
  if (e == hipSuccess && rc == QMCKL_SUCCESS_DEVICE) {
    for (int64_t i = 0; i < count; ++i) g->signature.push_back(batch_signature(ctx_of(contexts[i]), mo_vgl[i]));
    std::lock_guard<std::mutex> lock(g_graph_mutex);
    *graph = g_next_graph++;
    graphs()[*graph] = g;
    return QMCKL_SUCCESS_DEVICE;
  }
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  if (g->table) (void)hipFree(g->table);
  if (g->ao_table) (void)hipFree(g->ao_table);
  delete g;
  (void)hipGetLastError();
  if (rc != QMCKL_SUCCESS_DEVICE) return rc;
  return fail(contexts[0], QMCKL_FAILURE_DEVICE, fn, hipGetErrorString(e));
}

qmckl_exit_code_device qmckl_mo_basis_mo_vgl_batch_graph_launch_device(int64_t graph) {
  const char *fn = "qmckl_mo_basis_mo_vgl_batch_graph_launch_device";
  BatchGraph *g = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_graph_mutex);
    auto it = graphs().find(graph);
    if (it == graphs().end()) return QMCKL_INVALID_ARG_1_DEVICE;
    g = it->second;
  }
  for (size_t i = 0; i < g->contexts.size(); ++i) {
    Context *c = ctx_of(g->contexts[i]);
    if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
    std::vector<uint64_t> now = batch_signature(c, reinterpret_cast<const double *>(g->signature[i][6]));
    if (now != g->signature[i])
      return fail(g->contexts[i], QMCKL_FAILURE_DEVICE, fn,
                  "the context's sizes, precision form, basis or buffers changed since the capture");
  }
  hipError_t e = hipSetDevice(g->device);
  if (e == hipSuccess) e = hipGraphLaunch(g->exec, g->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
  if (e != hipSuccess) return fail(g->contexts[0], QMCKL_FAILURE_DEVICE, fn, hipGetErrorString(e));
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_mo_basis_mo_vgl_batch_graph_destroy_device(int64_t graph) {
  BatchGraph *g = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_graph_mutex);
    auto it = graphs().find(graph);
    if (it == graphs().end()) return QMCKL_INVALID_ARG_1_DEVICE;
    g = it->second;
    graphs().erase(it);
  }
  (void)hipSetDevice(g->device);
  (void)hipStreamSynchronize(g->stream);
  (void)hipGraphExecDestroy(g->exec);
  (void)hipGraphDestroy(g->graph);
  (void)hipStreamDestroy(g->stream);
  if (g->table) (void)hipFree(g->table);
  if (g->ao_table) (void)hipFree(g->ao_table);
  delete g;
  return QMCKL_SUCCESS_DEVICE;
}

}  // extern "C"

