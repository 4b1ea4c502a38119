// Differentiation form: adjoint of a weighted sum of mo_vgl with respect to
// nuclear coordinates, primitive exponents and MO coefficients.
//   G      = w C                  (FP64 MFMA GEMM, [5P][A], C in fragment
//                                  order; with a column plan, K runs over
//                                  the MOs in the plan's order and zero
//                                  coefficient blocks are skipped)
//   dL/dC  = (AO^T w)^T           (FP64 MFMA GEMM with AO read K-major and w
//                                  in fragment order, split over points,
//                                  fixed-order reduction)
//   dL/dR, dL/dalpha              (AO adjoint kernel recomputing primitives,
//                                  per-block partials, fixed-order reduction)
// There is no counterpart in the reference; the forward arithmetic it
// differentiates is src/qmckl_ao_sycl.cpp:40-95 and 291-422 and
// src/qmckl_mo_sycl.cpp:23-50.
#include <vector>

#include "internal.hpp"

using namespace qmckl_amd;

namespace {

bool grow_d(Context *c, double **slot, int64_t *cap, int64_t n) {
  return grow_bytes(c, reinterpret_cast<void **>(slot), cap, n * (int64_t)sizeof(double));
}

}  // namespace

extern "C" {

qmckl_exit_code_device qmckl_get_mo_basis_mo_vgl_adjoint_device(qmckl_context_device context,
                                                                const double *weight,
                                                                int64_t size_max,
                                                                double *d_nucl_coord,
                                                                double *d_exponent,
                                                                double *d_mo_coefficient) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  const char *fn = "qmckl_get_mo_basis_mo_vgl_adjoint_device";
  if (weight == nullptr) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "NULL weight");
  if (!c->mo.provided || !c->ao.provided || !c->nucleus.provided)
    return fail(context, QMCKL_NOT_PROVIDED_DEVICE, fn, nullptr);
  if (c->ao.type != 'G') return fail(context, QMCKL_FAILURE_DEVICE, fn, "basis type != 'G'");
  if (c->point.num <= 0 || c->point.coord == nullptr)
    return fail(context, QMCKL_NOT_PROVIDED_DEVICE, fn, "no points set");
  const int64_t P = c->point.num, A = c->ao.ao_num, M = c->mo.mo_num, N = c->nucleus.num;
  const int64_t S = c->ao.shell_num, Np = c->ao.prim_num;
  if (size_max < 5 * P * M) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "weight array too small");
  if (!has_device(c)) return fail(context, QMCKL_FAILURE_DEVICE, fn, "no device bound");
  MoBasis &mo = c->mo;
  hipError_t e = hipSuccess;
  // Both parts visit the points in spatially compact 64-point tiles where
  // they are ordered for ao_vgl (bases of 1024 AOs or more, 4,096 points or
  // more, or as the context's point-order setting says).  The order sets
  // the sums' rounding, so it comes from a stable sort of the points' cell
  // codes (within a cell by point index) and the gradients stay bitwise
  // reproducible from call to call.
  const int32_t cfg = c->mo.tile_point_order;
  const bool order = cfg >= 0 ? cfg == 1 : round_up(A, 64) >= 1024 && P >= 4096;
  const int32_t *perm = nullptr;  // point of position i
  const int32_t *kmap = nullptr;  // (point, component) row of row 5 i + c
  if (order) {
    if (!grow_bytes(c, reinterpret_cast<void **>(&mo.adj_perm), &mo.adj_perm_cap, (int64_t)sizeof(int32_t) * P) ||
        !grow_bytes(c, &mo.adj_order_ws, &mo.adj_order_ws_cap, point_order_stable_workspace_bytes(P)) ||
        !grow_bytes(c, reinterpret_cast<void **>(&mo.adj_kmap), &mo.adj_kmap_cap, (int64_t)sizeof(int32_t) * 5 * P))
      return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
    e = launch_point_order_stable(c->stream, P, c->point.coord, c->nucleus.num, c->nucleus.coord, mo.adj_order_ws,
                                  mo.adj_perm);
    if (e == hipSuccess) e = launch_point_rows_map(c->stream, P, mo.adj_perm, mo.adj_kmap);
    if (e != hipSuccess) return hip_check(context, e, fn);
    perm = mo.adj_perm;
    kmap = mo.adj_kmap;
  }
  // the AO screen per (64-point tile, 16-AO block): where it is zero, every
  // AO of the block vanishes at every point of the tile
  const int tnkb = (int)((A + 15) / 16);
  const uint8_t *tflags = nullptr;
  if (perm != nullptr) {
    if (!grow_bytes(c, reinterpret_cast<void **>(&mo.adj_tflags), &mo.adj_tflags_cap, (P + 63) / 64 * tnkb))
      return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
    e = launch_tile_screen(c->stream, c->ao.dev, P, c->point.coord, perm, c->nucleus.coord, mo.adj_tflags, tnkb);
    if (e != hipSuccess) return hip_check(context, e, fn);
    tflags = mo.adj_tflags;
  }
  if (d_nucl_coord != nullptr || d_exponent != nullptr) {
    const int64_t nblk = ao_adjoint_blocks(P);
    if (!grow_d(c, &mo.adj_G, &mo.adj_G_cap, 5 * P * A) ||
        !grow_d(c, &mo.adj_partR, &mo.adj_partR_cap, nblk * S * 3) ||
        !grow_d(c, &mo.adj_partA, &mo.adj_partA_cap, nblk * Np))
      return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
    ColPlan plan;
    const std::vector<uint64_t> *host_blk = nullptr;
    e = mo_f64_column_plan(c, &plan, &host_blk);
    if (e != hipSuccess) return hip_check(context, e, fn);
    if (!mo.adj_cfrag_ok) {
      // The coefficients [M][A] as the B operand in fragment order.  With a
      // column plan, K runs over the MOs in the plan's order (rows gathered),
      // and the plan's live bits, transposed to (16-AO column block, 16-MO K
      // block), let the contraction skip the zero blocks.
      mo.adj_planned = plan.col_mo != nullptr && host_blk != nullptr;
      const int64_t len = round_up(A, 64 * f64_frag_cbw(A)) * round_up(M, 16);
      if (!grow_d(c, &mo.adj_cfrag, &mo.adj_cfrag_cap, len))
        return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
      e = launch_frag_b_f64(c->stream, M, A, mo.coefficient, A, mo.adj_cfrag, &mo.adj_cfrag_nk4, nullptr,
                            mo.adj_planned ? plan.col_mo : nullptr);
      if (e == hipSuccess && mo.adj_planned) {
        const int64_t nmb = (M + 15) / 16, nab = (A + 15) / 16, nwt = (nmb + 63) / 64;
        std::vector<uint64_t> t((size_t)(nab * nwt), 0);
        for (int64_t mb = 0; mb < nmb; ++mb)
          for (int64_t ab = 0; ab < nab; ++ab)
            if (((*host_blk)[(size_t)(mb * plan.nw + ab / 64)] >> (ab % 64)) & 1ull)
              t[(size_t)(ab * nwt + mb / 64)] |= 1ull << (mb % 64);
        dfree(c, mo.adj_blk);
        mo.adj_blk = static_cast<uint64_t *>(dmalloc(c, sizeof(uint64_t) * t.size()));
        if (mo.adj_blk == nullptr) return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
        e = hipMemcpyAsync(mo.adj_blk, t.data(), sizeof(uint64_t) * t.size(), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = sync(c);
        mo.adj_blk_nw = nwt;
      }
      if (e != hipSuccess) return hip_check(context, e, fn);
      mo.adj_cfrag_ok = true;
    }
    // w's rows in tile order and its columns in the plan's MO order, then
    // G = w' C' (rows in tile order), skipping zero coefficient blocks and,
    // in tile order, the (tile, 16-AO) blocks of G the AO screen zeroes: the
    // AO adjoint reads G only where a shell passes the same screen for its
    // point, so those blocks are neither computed nor stored.
    const bool gather = mo.adj_planned || kmap != nullptr;
    const uint8_t *gcols = (M + 15) / 16 <= 2 * kMaxFlagBlocks ? tflags : nullptr;  // within the stage list
    if (gather && !grow_d(c, &mo.adj_wperm, &mo.adj_wperm_cap, 5 * P * M))
      return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
    {
      KernelTimer t(c, "adjoint_g_gemm_f64");
      if (gather)
        e = launch_gather(c->stream, 5 * P, M, weight, M, kmap, mo.adj_planned ? plan.col_mo : nullptr,
                          mo.adj_wperm, M);
      const double *wa = gather ? mo.adj_wperm : weight;
      if (e == hipSuccess && mo.adj_planned) {
        ColPlan gp;
        gp.blk = mo.adj_blk;
        gp.nw = mo.adj_blk_nw;
        gp.nblk = (A + 15) / 16;
        e = launch_mo_gemm_f64_frag(c->stream, 5 * P, M, A, wa, M, mo.adj_cfrag, mo.adj_cfrag_nk4, mo.adj_G, A,
                                    nullptr, 0, gp, gcols, tnkb);
      } else if (e == hipSuccess) {
        e = launch_mo_gemm_f64_frag(c->stream, 5 * P, M, A, wa, M, mo.adj_cfrag, mo.adj_cfrag_nk4, mo.adj_G, A,
                                    nullptr, 0, ColPlan(), gcols, tnkb);
      }
      t.stop();
    }
    // primitives that belong to no shell keep a zero partial
    if (e == hipSuccess) e = hipMemsetAsync(mo.adj_partA, 0, sizeof(double) * nblk * Np, c->stream);
    if (e == hipSuccess) {
      KernelTimer t(c, "ao_adjoint");
      e = launch_ao_adjoint(c->stream, c->ao.dev, P, c->point.coord, perm, c->nucleus.coord, c->ao.ao_factor,
                            mo.adj_G, mo.adj_partR, mo.adj_partA);
      t.stop();
    }
    if (e == hipSuccess)
      e = launch_adjoint_reduce(c->stream, c->ao.dev, nblk, c->ao.nucleus_index, c->ao.nucleus_shell_num,
                                mo.adj_partR, mo.adj_partA, d_nucl_coord, d_exponent);
    if (e == hipSuccess) e = sync(c);
    if (e != hipSuccess) return hip_check(context, e, fn);
    (void)N;
  }
  if (d_mo_coefficient != nullptr) {
    qmckl_exit_code_device rc = qmckl_provide_ao_basis_ao_vgl_device(context);
    if (rc != QMCKL_SUCCESS_DEVICE) return rc;
    // dL/dC^T [A][M] = AO^T w: AO [5P][A] read K-major, the weights as the
    // B operand in fragment order, split over the 5P points and components.
    // Where the points are ordered into spatially compact 64-point tiles
    // (as for ao_vgl), K runs over the (point, component) rows in tile order
    // and the 320 rows of a tile are skipped for the 64 AOs of a workgroup
    // that the AO screen zeroes on all of its points (21% of them on the
    // 64-water system at 32 walkers).
    const int splits = gemm_tn_frag_splits(A, 5 * P);
    const int64_t wlen = round_up(M, 64 * f64_frag_cbw(M)) * round_up(5 * P, 16);
    if (!grow_d(c, &mo.adj_partC, &mo.adj_partC_cap, (int64_t)splits * M * A) ||
        !grow_d(c, &mo.adj_wfrag, &mo.adj_wfrag_cap, wlen))
      return fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, fn, nullptr);
    const int64_t kchunk = ((5 * P + splits - 1) / splits + 15) / 16 * 16;
    const bool tiles = tflags != nullptr && kchunk / 16 <= 2 * kMaxFlagBlocks;
    int64_t nk4w = 0;
    if (e == hipSuccess)
      e = launch_frag_b_f64(c->stream, 5 * P, M, weight, M, mo.adj_wfrag, &nk4w, nullptr, tiles ? kmap : nullptr);
    if (e == hipSuccess) {
      KernelTimer t(c, "adjoint_dc_gemm_f64");
      e = launch_gemm_tn_f64_frag(c->stream, A, 5 * P, M, c->ao.ao_vgl, A, mo.adj_wfrag, nk4w, d_mo_coefficient,
                                  mo.adj_partC, splits, tiles ? kmap : nullptr, tiles ? tflags : nullptr, tnkb);
      t.stop();
    }
    if (e == hipSuccess) e = sync(c);
    if (e != hipSuccess) return hip_check(context, e, fn);
  }
  return QMCKL_SUCCESS_DEVICE;
}

}  // extern "C"
