"""C-ABI checks that need no GPU: the library loads, exports every symbol
``include/qmckl_gpu.h`` declares, and its host logic (context validity,
masks, set-once fields, setter ordering, argument checks and error
records) behaves as the reference's host code does.  The contexts here are
created with QMCKL_NO_DEVICE, so no HIP call is made."""
import ctypes
import subprocess

import numpy as np
import pytest

from qmckl_sycl_amd import _lib
from qmckl_sycl_amd.context import (ALREADY_SET, FAILURE, INVALID_CONTEXT, NO_DEVICE,
                                    NOT_PROVIDED, Context, QmcklError, lib)


def test_every_declared_symbol_is_exported():
    protos = _lib.parse_header()
    assert len(protos) > 80
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = sorted(set(protos) - exported)
    assert not missing, missing
    L = lib()
    for name in protos:
        assert getattr(L, name) is not None


def test_exports_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    names = [l.split()[-1] for l in out.splitlines() if "qmckl_" in l]
    assert all(not n.startswith("_Z") for n in names)


def test_error_strings():
    L = lib()
    assert L.qmckl_string_of_error_device(0) == b"Success"
    assert L.qmckl_string_of_error_device(3) == b"Invalid argument 3"
    assert L.qmckl_string_of_error_device(108).startswith(b"Attempt to modify")
    assert L.qmckl_string_of_error_device(12345) == b"Unknown error"


def test_invalid_contexts():
    L = lib()
    assert L.qmckl_context_check_device(0) == 0
    assert L.qmckl_context_destroy_device(0) == INVALID_CONTEXT
    assert L.qmckl_context_touch_device(0) == INVALID_CONTEXT
    assert L.qmckl_set_nucleus_num_device(0, 3) == INVALID_CONTEXT
    assert L.qmckl_failwith_device(0, 5, b"f", b"m") == INVALID_CONTEXT
    # a context whose memory no longer carries the valid tag
    buf = (ctypes.c_uint64 * 64)()
    assert L.qmckl_context_check_device(ctypes.addressof(buf)) == 0
    assert L.qmckl_ao_basis_provided_device(0) is False


def test_create_on_missing_device_returns_null():
    # device ordinals beyond the devices present fail cleanly
    assert lib().qmckl_context_create_device(4096) == 0


@pytest.fixture
def ctx():
    c = Context(NO_DEVICE)
    yield c
    c.destroy()


def test_touch_and_date(ctx):
    ctx.touch()
    ctx.touch()
    assert lib().qmckl_get_device_id(ctx.handle) == NO_DEVICE


def test_failwith_records_error(ctx):
    L = lib()
    assert L.qmckl_failwith_device(ctx.handle, 7, b"my_function", b"my message") == 7
    code, fn, msg = ctx.last_error()
    assert (code, fn, msg) == (7, "my_function", "my message")
    assert L.qmckl_failwith_device(ctx.handle, 0, b"x", b"y") == 0
    assert ctx.last_error()[0] == 7   # success does not overwrite the record


def test_nucleus_num_set_once(ctx):
    assert ctx.rc("qmckl_set_nucleus_num_device", 0) == 2
    assert ctx.rc("qmckl_set_nucleus_num_device", 3) == 0
    assert ctx.rc("qmckl_set_nucleus_num_device", 4) == ALREADY_SET
    assert ctx.nucleus_num() == 3
    assert not ctx.provided("nucleus")


def test_nucleus_num_not_provided(ctx):
    with pytest.raises(QmcklError) as e:
        ctx.nucleus_num()
    assert e.value.code == NOT_PROVIDED


def test_electron_num(ctx):
    assert ctx.rc("qmckl_set_electron_num_device", 0, 1) == 2
    assert ctx.rc("qmckl_set_electron_num_device", 2, -1) == 3
    assert not ctx.provided("electron")
    assert ctx.rc("qmckl_set_electron_num_device", 5, 4) == 0
    assert ctx.provided("electron")
    assert ctx.electron_num() == 9
    assert ctx.rc("qmckl_set_electron_num_device", 5, 4) == ALREADY_SET


def test_electron_coord_checks(ctx):
    # elec_num not set -> FAILURE
    assert ctx.rc("qmckl_set_electron_coord_device", b"N", 1, 1234, 3) == FAILURE
    ctx.call("qmckl_set_electron_num_device", 2, 1)
    # array too small -> INVALID_ARG_4 (set_point's code), bad transp -> 2, NULL -> 3
    assert ctx.rc("qmckl_set_electron_coord_device", b"N", 2, 1234, 17) == 4
    assert ctx.rc("qmckl_set_electron_coord_device", b"X", 2, 1234, 18) == 2
    assert ctx.rc("qmckl_set_electron_coord_device", b"N", 2, None, 18) == 3
    assert ctx.rc("qmckl_set_electron_coord_device", b"N", 0, 1234, 18) == 3
    # valid arguments on a context with no device: FAILURE, not a crash
    assert ctx.rc("qmckl_set_electron_coord_device", b"N", 2, 1234, 18) == FAILURE
    assert ctx.rc("qmckl_get_electron_coord_device", b"Q", 1234, 18) == 2
    assert ctx.rc("qmckl_get_electron_coord_device", b"N", None, 18) == 3
    assert ctx.rc("qmckl_get_electron_coord_device", b"N", 1234, 0) == 4


def test_distance_checks(ctx):
    """qmckl_distance_device's codes (src/qmckl_electron.cpp:373-470)."""
    D = "qmckl_distance_device"
    assert ctx.rc(D, b"T", b"T", 0, 5, 1234, 4, 1234, 5, 1234, 4) == 4
    assert ctx.rc(D, b"T", b"T", 4, 0, 1234, 4, 1234, 5, 1234, 4) == 5
    assert ctx.rc(D, b"X", b"T", 4, 5, 1234, 4, 1234, 5, 1234, 4) == 1
    assert ctx.rc(D, b"T", b"y", 4, 5, 1234, 4, 1234, 5, 1234, 4) == 1
    assert ctx.rc(D, b"T", b"T", 4, 5, 1234, 3, 1234, 5, 1234, 4) == 7    # 'T': lda >= m
    assert ctx.rc(D, b"N", b"T", 4, 5, 1234, 2, 1234, 5, 1234, 4) == 7    # 'N': lda >= 3
    assert ctx.rc(D, b"T", b"T", 4, 5, 1234, 4, 1234, 4, 1234, 4) == 9    # 'T': ldb >= n
    assert ctx.rc(D, b"T", b"n", 4, 5, 1234, 4, 1234, 2, 1234, 4) == 9    # 'N': ldb >= 3
    assert ctx.rc(D, b"T", b"T", 4, 5, 1234, 4, 1234, 5, 1234, 3) == 11
    assert ctx.rc(D, b"T", b"T", 4, 5, None, 4, 1234, 5, 1234, 4) == 6
    assert ctx.rc(D, b"T", b"T", 4, 5, 1234, 4, None, 5, 1234, 4) == 8
    assert ctx.rc(D, b"T", b"T", 4, 5, 1234, 4, 1234, 5, None, 4) == 10
    # 'N' for A with lda = 3 and m > 3, 'T' for B: accepted (a leading
    # dimension is checked against its own flag only); no device -> FAILURE
    assert ctx.rc(D, b"n", b"t", 40, 5, 1234, 3, 1234, 5, 1234, 40) == FAILURE


def test_point_checks(ctx):
    assert ctx.rc("qmckl_set_point_device", b"N", 10, 1234, 29) == 4
    assert ctx.rc("qmckl_set_point_device", b"Z", 10, 1234, 30) == 2
    assert ctx.rc("qmckl_set_point_device", b"N", 10, None, 30) == 3
    assert ctx.rc("qmckl_get_point_device", b"N", 1234, 30) == NOT_PROVIDED


def test_ao_scalar_setters_and_order(ctx):
    # prim_num before shell_num fails; shell_num must be positive
    assert ctx.rc("qmckl_set_ao_basis_prim_num_device", 10) == FAILURE
    assert ctx.rc("qmckl_set_ao_basis_shell_num_device", 0) == 2
    assert ctx.rc("qmckl_set_ao_basis_type_device", b"X") == 2
    assert ctx.rc("qmckl_set_ao_basis_type_device", b"G") == 0
    assert ctx.rc("qmckl_set_ao_basis_shell_num_device", 4) == 0
    assert ctx.rc("qmckl_set_ao_basis_prim_num_device", 3) == 2       # prim_num < shell_num
    assert ctx.rc("qmckl_set_ao_basis_prim_num_device", 9) == 0
    assert ctx.rc("qmckl_set_ao_basis_shell_num_device", 10) == 2     # shell_num > prim_num
    assert ctx.rc("qmckl_set_ao_basis_ao_num_device", 3) == 2          # ao_num < shell_num
    assert ctx.rc("qmckl_set_ao_basis_ao_num_device", 12) == 0
    assert (ctx.shell_num(), ctx.prim_num(), ctx.ao_num()) == (4, 9, 12)
    t = ctypes.c_char(b" ")
    ctx.call("qmckl_get_ao_basis_type_device", ctypes.byref(t))
    assert t.value == b"G"
    assert not ctx.provided("ao_basis")


def test_ao_array_setters_need_sizes(ctx):
    # nucleus-indexed arrays need nucl_num first
    assert ctx.rc("qmckl_set_ao_basis_nucleus_index_device", 1234, 5) == FAILURE
    ctx.call("qmckl_set_nucleus_num_device", 5)
    # too small: FAILURE for nucleus_index (as the reference), ARG_3 elsewhere
    assert ctx.rc("qmckl_set_ao_basis_nucleus_index_device", 1234, 4) == FAILURE
    assert ctx.rc("qmckl_set_ao_basis_nucleus_shell_num_device", 1234, 4) == 3
    # per-shell arrays need shell_num
    assert ctx.rc("qmckl_set_ao_basis_shell_ang_mom_device", 1234, 10) == FAILURE
    ctx.call("qmckl_set_ao_basis_shell_num_device", 10)
    assert ctx.rc("qmckl_set_ao_basis_shell_ang_mom_device", 1234, 9) == 3
    # a correctly sized array on a device-less context fails without a crash
    assert ctx.rc("qmckl_set_ao_basis_shell_ang_mom_device", 1234, 10) == FAILURE


def test_ao_getters_not_provided(ctx):
    for name in ["qmckl_get_ao_basis_shell_ang_mom_device", "qmckl_get_ao_basis_exponent_device",
                 "qmckl_get_ao_basis_ao_factor_device"]:
        assert ctx.rc(name, 1234, 100) == NOT_PROVIDED
    assert ctx.rc("qmckl_get_ao_basis_ao_vgl_device", 1234, 100) == NOT_PROVIDED
    assert ctx.rc("qmckl_provide_ao_basis_ao_vgl_device") == NOT_PROVIDED
    assert ctx.rc("qmckl_provide_en_distance_device") == NOT_PROVIDED


def test_mo_setters(ctx):
    assert ctx.rc("qmckl_set_mo_basis_mo_num_device", 0) == 2
    assert ctx.rc("qmckl_set_mo_basis_mo_num_device", 7) == 0
    assert ctx.rc("qmckl_set_mo_basis_mo_num_device", 7) == ALREADY_SET
    assert ctx.mo_num() == 7
    # coefficient needs ao_num
    assert ctx.rc("qmckl_set_mo_basis_coefficient_device", 1234) == FAILURE
    assert ctx.rc("qmckl_set_mo_basis_coefficient_device", None) == 2
    assert ctx.rc("qmckl_get_mo_basis_mo_vgl_device", 1234, 10) == NOT_PROVIDED


def test_memory_without_device(ctx):
    L = lib()
    assert not L.qmckl_malloc_device(ctx.handle, 64)
    assert ctx.last_error()[0] == 104
    assert L.qmckl_free_device(ctx.handle, 1234) == 2     # not tracked
    info = _lib.MemoryInfo(128, None)
    p = L.qmckl_malloc_host(ctx.handle, info)
    assert p
    ctypes.memset(p, 1, 128)
    assert L.qmckl_free_host(ctx.handle, p) == 0
    assert L.qmckl_free_host(ctx.handle, p) == 2          # second free: untracked
    assert L.qmckl_memcpy_H2D(ctx.handle, 1, 2, 8) == FAILURE
    assert L.qmckl_memcpy_D2H(ctx.handle, None, 2, 8) == 2
    assert L.qmckl_memcpy_D2D(ctx.handle, 1, None, 8) == 3


def test_low_level_launchers_validate(ctx):
    L = lib()
    h = ctx.handle
    assert L.qmckl_compute_mo_basis_mo_vgl_device(h, 0, 4, 4, 1, 1, 1) == 2
    assert L.qmckl_compute_mo_basis_mo_vgl_device(h, 4, 0, 4, 1, 1, 1) == 3
    assert L.qmckl_compute_mo_basis_mo_vgl_device(h, 4, 4, 4, None, 1, 1) == 5
    assert L.qmckl_compute_mo_basis_mo_vgl_device(h, 4, 4, 4, 1, 1, 1) == FAILURE
    assert L.qmckl_compute_en_distance_device(h, 0, 3, 1, 1, 1) == 2
    assert L.qmckl_compute_en_distance_device(h, 3, 0, 1, 1, 1) == 3
    assert L.qmckl_compute_en_distance_device(h, 3, 3, 1, 1, 1) == FAILURE
    na = L.qmckl_compute_ao_basis_shell_gaussian_vgl_noalloc_device
    ptrs = [1] * 10
    assert na(h, 0, 4, 4, 1, *ptrs) == 2
    assert na(h, 4, 0, 4, 1, *ptrs) == 3
    assert na(h, 4, 65537, 4, 1, *ptrs) == 3       # past the fixed scratch table
    assert na(h, 4, 4, -1, 1, *ptrs) == 4
    assert na(h, 4, 4, 4, 0, *ptrs) == 5
    assert na(h, 4, 4, 4, 1, *([1] * 9 + [None])) == 6
    assert na(h, 4, 4, 4, 1, *ptrs) == FAILURE       # no device bound


def test_destroy_frees_host_allocations():
    c = Context(NO_DEVICE)
    L = lib()
    for _ in range(200):   # beyond the reference's initial 128-entry table
        assert L.qmckl_malloc_host(c.handle, _lib.MemoryInfo(16, None))
    c.destroy()
    assert c.handle == 0


def test_lock_unlock_recursive(ctx):
    L = lib()
    L.qmckl_lock_device(ctx.handle)
    L.qmckl_lock_device(ctx.handle)
    L.qmckl_unlock_device(ctx.handle)
    L.qmckl_unlock_device(ctx.handle)


def test_precision_form_setting(ctx):
    L = lib()
    f = ctypes.c_int32(-1)
    ctx.call("qmckl_get_mo_basis_precision_device", ctypes.byref(f))
    assert f.value == 0                       # FP64 plain by default
    assert ctx.rc("qmckl_set_mo_basis_precision_device", 6) == 2
    assert ctx.rc("qmckl_set_mo_basis_precision_device", -1) == 2
    for form in (1, 2, 3, 4, 5, 0):
        ctx.call("qmckl_set_mo_basis_precision_device", form)
        ctx.call("qmckl_get_mo_basis_precision_device", ctypes.byref(f))
        assert f.value == form
    assert L.qmckl_get_mo_basis_precision_device(ctx.handle, None) == 2


def test_adjoint_validation(ctx):
    assert ctx.rc("qmckl_get_mo_basis_mo_vgl_adjoint_device", None, 10, 1, 1, 1) == 2
    assert ctx.rc("qmckl_get_mo_basis_mo_vgl_adjoint_device", 1234, 10, 1, 1, 1) == NOT_PROVIDED


def test_kernel_timing_api(ctx):
    n = ctypes.c_int64(5)
    ms = ctypes.c_double(1.0)
    ctx.call("qmckl_set_kernel_timing_device", True)
    ctx.call("qmckl_get_kernel_timing_device", b"ao_vgl", ctypes.byref(n), ctypes.byref(ms))
    assert (n.value, ms.value) == (0, 0.0)
    ctx.call("qmckl_reset_kernel_timing_device")
    assert ctx.rc("qmckl_get_kernel_timing_device", None, None, None) == 2


def test_batch_validation(ctx):
    L = lib()
    f = L.qmckl_get_mo_basis_mo_vgl_batch_device
    one = (ctypes.c_int64 * 1)(ctx.handle)
    ptr = (ctypes.c_void_p * 1)(1234)
    size = (ctypes.c_int64 * 1)(10)
    assert f(0, one, ptr, size) == 1
    assert f(-3, one, ptr, size) == 1
    assert f(1, None, ptr, size) == 2
    assert f(1, one, None, size) == 3
    assert f(1, one, ptr, None) == 4
    bad = (ctypes.c_int64 * 2)(ctx.handle, 77)
    ptr2 = (ctypes.c_void_p * 2)(1234, 1234)
    size2 = (ctypes.c_int64 * 2)(10, 10)
    assert f(2, bad, ptr2, size2) == INVALID_CONTEXT
    nullout = (ctypes.c_void_p * 1)(None)
    assert f(1, one, nullout, size) == 3
    assert ctx.last_error()[1] == "qmckl_get_mo_basis_mo_vgl_batch_device"
    # an unprovided basis is reported as the single-context getter reports it
    assert f(1, one, ptr, size) == NOT_PROVIDED


def test_destroyed_context_is_invalid():
    c = Context(NO_DEVICE)
    h = c.handle
    c.destroy()
    L = lib()
    n = ctypes.c_int64(0)
    assert L.qmckl_get_nucleus_num_device(h, ctypes.byref(n)) == INVALID_CONTEXT
    assert L.qmckl_context_destroy_device(h) == INVALID_CONTEXT


def test_select_mo_and_ee_distance_validation(ctx):
    L = lib()
    keep = (ctypes.c_int32 * 4)(1, 0, 1, 1)
    assert L.qmckl_mo_basis_select_mo_device(ctx.handle, keep, 4) == NOT_PROVIDED
    assert L.qmckl_mo_basis_select_mo_device(0, keep, 4) == INVALID_CONTEXT
    assert L.qmckl_provide_ee_distance_device(ctx.handle) == NOT_PROVIDED
    assert L.qmckl_get_electron_ee_distance_device(ctx.handle, None, 10) == 2
    assert L.qmckl_compute_ee_distance_device(ctx.handle, 0, 1, 1234, 1234) == 2
    assert L.qmckl_compute_ee_distance_device(ctx.handle, 3, 0, 1234, 1234) == 3
    assert L.qmckl_compute_ee_distance_device(ctx.handle, 3, 1, None, 1234) == 4
    assert L.qmckl_compute_ee_distance_device(ctx.handle, 3, 1, 1234, None) == 5


def test_tile_config_and_autotune_validation(ctx):
    L = lib()
    po, gw = ctypes.c_int32(9), ctypes.c_int32(9)
    assert L.qmckl_get_mo_basis_tile_config_device(ctx.handle, ctypes.byref(po), ctypes.byref(gw)) == 0
    assert (po.value, gw.value) == (-1, 0)  # automatic by default
    assert L.qmckl_set_mo_basis_tile_config_device(ctx.handle, 2, 0) == 2
    assert L.qmckl_set_mo_basis_tile_config_device(ctx.handle, 0, -1) == 3
    assert L.qmckl_set_mo_basis_tile_config_device(ctx.handle, 0, 2000) == 3
    assert L.qmckl_set_mo_basis_tile_config_device(ctx.handle, 1, 32) == 0
    assert L.qmckl_get_mo_basis_tile_config_device(ctx.handle, ctypes.byref(po), ctypes.byref(gw)) == 0
    assert (po.value, gw.value) == (1, 32)
    assert L.qmckl_get_mo_basis_tile_config_device(ctx.handle, None, ctypes.byref(gw)) == 2
    assert L.qmckl_get_mo_basis_tile_config_device(ctx.handle, ctypes.byref(po), None) == 3
    assert L.qmckl_set_mo_basis_tile_config_device(0, 0, 0) == INVALID_CONTEXT
    gt = ctypes.c_int32(9)
    assert L.qmckl_get_mo_basis_gemm_tile_device(ctx.handle, ctypes.byref(gt)) == 0 and gt.value == 0
    assert L.qmckl_set_mo_basis_gemm_tile_device(ctx.handle, 7) == 2
    assert L.qmckl_set_mo_basis_gemm_tile_device(ctx.handle, 6) == 0
    assert L.qmckl_set_mo_basis_gemm_tile_device(ctx.handle, 5) == 0
    assert L.qmckl_set_mo_basis_gemm_tile_device(ctx.handle, -1) == 2
    assert L.qmckl_set_mo_basis_gemm_tile_device(ctx.handle, 3) == 0
    assert L.qmckl_get_mo_basis_gemm_tile_device(ctx.handle, ctypes.byref(gt)) == 0 and gt.value == 3
    assert L.qmckl_get_mo_basis_gemm_tile_device(ctx.handle, None) == 2
    assert L.qmckl_set_mo_basis_gemm_tile_device(0, 1) == INVALID_CONTEXT
    assert L.qmckl_mo_basis_autotune_device(ctx.handle) == NOT_PROVIDED
    assert L.qmckl_mo_basis_autotune_device(0) == INVALID_CONTEXT
    live = ctypes.c_double(0.0)
    assert L.qmckl_get_mo_basis_column_plan_device(ctx.handle, None, ctypes.byref(live)) == NOT_PROVIDED
    assert L.qmckl_get_mo_basis_column_plan_device(0, None, None) == INVALID_CONTEXT
    g = ctypes.c_int64(0)
    assert L.qmckl_mo_basis_mo_vgl_batch_graph_launch_device(12345) == 1
    assert L.qmckl_mo_basis_mo_vgl_batch_graph_destroy_device(12345) == 1
    assert L.qmckl_mo_basis_mo_vgl_batch_graph_create_device(1, None, None, None, None) == 5
    handles = (ctypes.c_int64 * 1)(ctx.handle)
    ptrs = (ctypes.c_void_p * 1)(1234)
    sizes = (ctypes.c_int64 * 1)(10)
    assert L.qmckl_mo_basis_mo_vgl_batch_graph_create_device(0, handles, ptrs, sizes, ctypes.byref(g)) == 1
