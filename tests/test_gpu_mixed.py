"""Precision forms of the MO contraction on the GPU, through the C ABI,
against the FP64 oracle.

Error measure (written before tuning): for each (point, component) row,
max_i |mo - mo_ref| / max_i |mo_ref|, maximised over rows, reported
separately for value (c=0), gradient (c=1..3) and Laplacian (c=4).
Bars: 1e-6 for FP32, FP16X2 and BF16X3; the two-term bf16 split (BF16X2)
carries ~2^-17 relative representation error per operand and is held to
1e-4 (its measured level is reported); the int8-slice form INT8X5, whose
sums are exact, is held to 1e-7, and to 1e-6 on the badly conditioned
ChBrClF basis where the FP32-accumulating forms stop at a few 1e-6."""
import ctypes

import numpy as np
import pytest

from qmckl_sycl_amd import systems as S
from qmckl_sycl_amd.context import Context, DeviceArray, lib

from helpers import say

pytestmark = pytest.mark.gpu

FORMS = {"fp32": 1, "fp16x2": 2, "bf16x2": 3, "bf16x3": 4, "int8x5": 5}
BAR = {"fp32": 1e-6, "fp16x2": 1e-6, "bf16x2": 1e-4, "bf16x3": 1e-6, "int8x5": 1e-7}


def row_err(got, ref):
    d = np.abs(got - ref)
    rm = np.abs(ref).max(axis=2, keepdims=True)
    e = (d / np.where(rm > 0, rm, 1.0)).max(axis=(0, 2))
    return {"value": float(e[0]), "gradient": float(e[1:4].max()), "laplacian": float(e[4])}


def run_form(ctx, form, P):
    ctx.call("qmckl_set_mo_basis_precision_device", FORMS[form])
    out = ctx.mo_vgl(P)
    f = ctypes.c_int32(-1)
    ctx.call("qmckl_get_mo_basis_precision_device", ctypes.byref(f))
    assert f.value == FORMS[form]
    return out


@pytest.fixture(scope="module")
def chb():
    return S.chbrclf()


# ChBrClF is worse conditioned: Br core exponents up to 1e7 make single
# terms of an MO row up to ~1e2 times the row's largest value, so every
# FP32-class form lands at a few 1e-6 (measured 4-7e-6).  Bars 10x looser.
# INT8X5 sums exactly and is held to the 1e-6 of the well-conditioned bases.
BAR_CHBRCLF = {"fp32": 1e-5, "fp16x2": 1e-5, "bf16x2": 1e-3, "bf16x3": 1e-5, "int8x5": 1e-6}


@pytest.mark.parametrize("form", list(FORMS))
def test_chbrclf_forms(chb, form):
    ref = S.chbrclf_mo_reference()
    with Context(0) as c:
        c.set_system(chb)
        got = run_form(c, form, 136)
    e = row_err(got, ref)
    say(f"chbrclf {form}: {e}")
    assert max(e.values()) <= BAR_CHBRCLF[form], e


@pytest.mark.parametrize("form", list(FORMS))
def test_chbrclf_sized_forms(oracle_lib, form):
    s = S.chbrclf_sized()
    with Context(0) as c:
        c.set_system(s)
        got = run_form(c, form, s.point_num)
    ref = oracle_lib.mo_vgl(s)
    e = row_err(got, ref)
    say(f"chbrclf_sized {form}: {e}")
    assert max(e.values()) <= BAR[form], e


def test_water64_forms_sampled(oracle_lib):
    """64-water, 128 walkers (81,920 points): every form on the full batch,
    checked on 512 sampled points."""
    s = S.water64(walk_num=128)
    rng = np.random.default_rng(1)
    idx = np.sort(rng.choice(s.point_num, size=512, replace=False))
    coord = np.ascontiguousarray(s.coord_T()[:, idx])
    ref = oracle_lib.mo_vgl_from_ao(s.mo_coef, oracle_lib.ao_vgl(s, coord))
    errs = {}
    with Context(0) as c:
        c.set_system(s)
        for form in FORMS:
            got = run_form(c, form, s.point_num)[idx]
            errs[form] = row_err(got, ref)
            say(f"water64 {form}: {errs[form]}")
            assert np.all(np.isfinite(got))
        c.call("qmckl_set_mo_basis_precision_device", 0)
        got = c.mo_vgl(s.point_num)[idx]
        errs["fp64"] = row_err(got, ref)
        say(f"water64 fp64: {errs['fp64']}")
    for form, e in errs.items():
        assert max(e.values()) <= BAR.get(form, 1e-12), (form, e)


def test_water64_fp16x2_every_point(oracle_lib):
    """The headline workload in full: 64-water, 128 walkers, FP16X2 and
    FP64, every one of the 81,920 points against the oracle (evaluated in
    chunks of 8,192 points), the row error measure over all 409,600 rows."""
    s = S.water64(walk_num=128)
    P = s.point_num
    coord = s.coord_T()
    with Context(0) as c:
        c.set_system(s)
        got16 = run_form(c, "fp16x2", P)
        c.call("qmckl_set_mo_basis_precision_device", 0)
        got64 = c.mo_vgl(P)
    worst16 = {"value": 0.0, "gradient": 0.0, "laplacian": 0.0}
    worst64 = dict(worst16)
    for p0 in range(0, P, 8192):
        sl = slice(p0, min(P, p0 + 8192))
        ref = oracle_lib.mo_vgl_from_ao(s.mo_coef, oracle_lib.ao_vgl(s, np.ascontiguousarray(coord[:, sl])))
        for worst, got in ((worst16, got16), (worst64, got64)):
            e = row_err(got[sl], ref)
            for k in worst:
                worst[k] = max(worst[k], e[k])
        del ref
    say(f"water64 every point: fp16x2 {worst16}; fp64 {worst64}")
    assert max(worst16.values()) <= BAR["fp16x2"], worst16
    assert max(worst64.values()) <= 1e-12, worst64


@pytest.mark.parametrize("form", list(FORMS))
def test_point_order_invariance(form):
    """Each point's MOs do not depend on which points share its tile (the
    library orders points spatially into tiles and skips screened stages
    per tile): the same points given in another order give bit-identical
    rows."""
    s = S.water64(walk_num=8)
    P = s.point_num
    perm = np.random.default_rng(7).permutation(P)
    X = np.asarray(s.elec_coord).reshape(P, 3)
    with Context(0) as c:
        c.set_system(s)
        a = run_form(c, form, P)
        c.set_electron_coord(X[perm].reshape(np.shape(s.elec_coord)))
        b = run_form(c, form, P)
    assert np.array_equal(b, a[perm])


@pytest.mark.parametrize("form", ["fp16x2", "bf16x2", "bf16x3"])
def test_tile_choices_and_autotune_keep_results(form):
    """The tile choices (spatial point order on/off, AO-operand wavefronts
    per CU, the contraction kernel) change speed only: every choice, and the
    autotuner's, gives the same rows bitwise."""
    s = S.water64(walk_num=8)
    P = s.point_num
    with Context(0) as c:
        c.set_system(s)
        ref = run_form(c, form, P)
        for order in (0, 1):
            for waves in (16, 128):
                c.call("qmckl_set_mo_basis_tile_config_device", order, waves)
                c.touch()
                assert np.array_equal(c.mo_vgl(P), ref), (order, waves)
        c.call("qmckl_set_mo_basis_tile_config_device", -1, 0)
        for gemm in (1, 2, 3, 4, 6):  # LIVE, LIVE_CACHED, DENSE, LIVE64, DENSE64
            c.call("qmckl_set_mo_basis_gemm_tile_device", gemm)
            c.touch()
            assert np.array_equal(c.mo_vgl(P), ref), gemm
        c.call("qmckl_set_mo_basis_gemm_tile_device", 0)
        c.call("qmckl_mo_basis_autotune_device")
        po, gw, gt = ctypes.c_int32(-5), ctypes.c_int32(-5), ctypes.c_int32(-5)
        c.call("qmckl_get_mo_basis_tile_config_device", ctypes.byref(po), ctypes.byref(gw))
        c.call("qmckl_get_mo_basis_gemm_tile_device", ctypes.byref(gt))
        say(f"autotuned {form}: point_order={po.value} group_waves={gw.value} gemm_tile={gt.value}")
        assert po.value in (0, 1) and gw.value in (16, 32, 64, 128, 256) and gt.value in (0, 1, 2, 3, 4, 6)
        assert np.array_equal(c.mo_vgl(P), ref)


def test_more_mos_than_one_workgroup_covers(oracle_lib):
    """mo_num = 450 needs two column tiles in every contraction (320 MOs per
    workgroup); 8 waters, 4 walkers (320 points), all points checked."""
    s = S.water_cluster(n_side=2, walk_num=4, mo_num=450, name="water8_450mo")
    ref = oracle_lib.mo_vgl(s)
    with Context(0) as c:
        c.set_system(s)
        for form in FORMS:
            e = row_err(run_form(c, form, s.point_num), ref)
            say(f"water8 450 MOs {form}: {e}")
            assert max(e.values()) <= BAR[form], (form, e)
        c.call("qmckl_set_mo_basis_precision_device", 0)
        e = row_err(c.mo_vgl(s.point_num), ref)
        say(f"water8 450 MOs fp64: {e}")
        assert max(e.values()) <= 1e-12, e


def test_switching_forms_recomputes(chb):
    with Context(0) as c:
        c.set_system(chb)
        a = c.mo_vgl(136)
        b = run_form(c, "fp16x2", 136)
        assert not np.array_equal(a, b)
        c.call("qmckl_set_mo_basis_precision_device", 0)
        assert np.array_equal(c.mo_vgl(136), a)
        assert c.rc("qmckl_set_mo_basis_precision_device", 9) == 2


def test_block_skipping_is_exact(oracle_lib):
    """The FP64 contraction with screened blocks skipped (provider path) is
    bit-identical to the dense low-level launcher on the same ao_vgl."""
    s = S.water64(walk_num=2)
    with Context(0) as c:
        c.set_system(s)
        P, A, M = s.point_num, s.ao_num, s.mo_num
        mo = c.mo_vgl(P)
        ao = DeviceArray(c, (P, 5, A))
        c.call("qmckl_get_ao_basis_ao_vgl_device", ao.ptr, ao.size)
        ct = DeviceArray.from_numpy(c, np.ascontiguousarray(s.mo_coef.T))
        out = DeviceArray(c, (P, 5, M))
        c.check(lib().qmckl_compute_mo_basis_mo_vgl_device(c.handle, A, M, P, ct.ptr, ao.ptr, out.ptr))
        dense = out.numpy()
    assert np.array_equal(mo, dense)


def column_plan(c, M):
    col = (ctypes.c_int32 * M)()
    live = ctypes.c_double(-1.0)
    c.call("qmckl_get_mo_basis_column_plan_device", col, ctypes.byref(live))
    return np.frombuffer(col, dtype=np.int32).copy(), live.value


def test_column_plan_follows_coefficient_zeros(chb):
    """Localised MOs (the 64-water system: 70% of the coefficients zero by
    locality) get a planned column order that leaves at most 90% of the
    16-MO x 16-AO blocks live; ChBrClF's dense coefficients keep the MO
    order."""
    s = S.water64(walk_num=1)
    with Context(0) as c:
        c.set_system(s)
        col, live = column_plan(c, s.mo_num)
    say(f"water64 column plan: {live:.3f} of the blocks live")
    assert sorted(col.tolist()) == list(range(s.mo_num))
    assert not np.array_equal(col, np.arange(s.mo_num))
    assert live <= 0.9
    with Context(0) as c:
        c.set_system(chb)
        col, live = column_plan(c, chb.mo_num)
    assert np.array_equal(col, np.arange(chb.mo_num)) and live > 0.9


@pytest.mark.parametrize("mo_num", [37, 333])
def test_planned_columns_ragged_mo_counts(oracle_lib, mo_num):
    """Planned column order with a partial last 16-column block (37 MOs) and
    with more MOs than one workgroup's 320 columns (333): FP64 bitwise equal
    to the dense low-level contraction and within 1e-12 of the oracle, the
    value-only contraction equal to the oracle's value rows, every split
    form within its bar."""
    s = S.water_cluster(n_side=2, walk_num=2, mo_num=mo_num, name=f"water8_{mo_num}mo")
    ref = oracle_lib.mo_vgl(s)
    P, A, M = s.point_num, s.ao_num, s.mo_num
    with Context(0) as c:
        c.set_system(s)
        col, live = column_plan(c, M)
        say(f"water8 {mo_num} MOs: {live:.3f} of the blocks live")
        assert live <= 0.9 and sorted(col.tolist()) == list(range(M))
        mv = c.mo_value(P)  # value-only contraction (mo_vgl not computed yet)
        mo = c.mo_vgl(P)
        e = row_err(mo, ref)
        say(f"water8 {mo_num} MOs fp64: {e}")
        assert max(e.values()) <= 1e-12, e
        rv = ref[:, 0, :]
        ev = np.max(np.abs(mv - rv) / np.maximum(np.abs(rv).max(axis=1, keepdims=True), 1e-300))
        assert ev <= 1e-12, ev
        ao = DeviceArray(c, (P, 5, A))
        c.call("qmckl_get_ao_basis_ao_vgl_device", ao.ptr, ao.size)
        ct = DeviceArray.from_numpy(c, np.ascontiguousarray(s.mo_coef.T))
        out = DeviceArray(c, (P, 5, M))
        c.check(lib().qmckl_compute_mo_basis_mo_vgl_device(c.handle, A, M, P, ct.ptr, ao.ptr, out.ptr))
        assert np.array_equal(mo, out.numpy())
        for form in ("fp16x2", "bf16x2", "bf16x3"):
            e = row_err(run_form(c, form, P), ref)
            say(f"water8 {mo_num} MOs {form}: {e}")
            assert max(e.values()) <= BAR[form], (form, e)


@pytest.mark.parametrize("form", ["fp16x2", "bf16x3"])
def test_mo_value_independent_of_call_order(oracle_lib, form):
    """In a mixed-precision context mo_value is the FP64 value-only
    contraction whether or not mo_vgl (in the mixed form) was computed for
    the same points first: bitwise the same both ways, and within 1e-12 of
    the oracle's value rows (8 waters, 4 walkers)."""
    s = S.water_cluster(n_side=2, walk_num=4, name="water8")
    P = s.point_num
    ref = oracle_lib.mo_vgl(s)[:, 0, :]
    with Context(0) as c:
        c.set_system(s)
        c.call("qmckl_set_mo_basis_precision_device", FORMS[form])
        c.touch()
        first = c.mo_value(P)            # mo_vgl not current
        c.touch()
        c.mo_vgl(P)                      # mixed-form mo_vgl current ...
        second = c.mo_value(P)           # ... and mo_value asked for after it
    assert np.array_equal(first, second)
    ev = np.max(np.abs(first - ref) / np.maximum(np.abs(ref).max(axis=1, keepdims=True), 1e-300))
    say(f"{form} context mo_value vs oracle: {ev:.2e}")
    assert ev <= 1e-12
