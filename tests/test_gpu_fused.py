"""Fused AO evaluation and split contraction (QMCKL_GEMM_TILE_FUSED_DEVICE)
on the GPU through the C ABI: that it runs when chosen (and AUTO keeps the
stored-operand kernels), its parity with the FP64 oracle at each split
form's bar (one and several 64-MO groups, point counts that are not
multiples of 64), and that a point's MOs do not depend on the points
sharing its tile.  Its operand rows are scaled per 64-AO stage where the
stored-operand kernels scale them once, so the two agree to the forms'
error bars rather than bitwise."""
import ctypes

import numpy as np
import pytest

from qmckl_sycl_amd import systems as S
from qmckl_sycl_amd.context import Context

from helpers import say

pytestmark = pytest.mark.gpu

SPLIT = {"fp16x2": 2, "bf16x2": 3, "bf16x3": 4}
# the bars of the mixed forms (max over rows of max_i |d| / max_i |ref|)
BAR = {"fp16x2": 1e-6, "bf16x2": 1e-4, "bf16x3": 1e-6}
FUSED = 5  # QMCKL_GEMM_TILE_FUSED_DEVICE


def row_err(got, ref):
    d = np.abs(got - ref)
    rm = np.abs(ref).max(axis=2, keepdims=True)
    return float((d / np.where(rm > 0, rm, 1.0)).max())


def launches(c, name):
    n, ms = ctypes.c_int64(0), ctypes.c_double(0.0)
    c.call("qmckl_get_kernel_timing_device", name.encode(), ctypes.byref(n), ctypes.byref(ms))
    return n.value


def timed_mo_vgl(c, P):
    """mo_vgl and the launches of the fused kernel and of the AO operand
    kernel that computing it took."""
    c.call("qmckl_reset_kernel_timing_device")
    c.call("qmckl_set_kernel_timing_device", True)
    c.touch()
    out = c.mo_vgl(P)
    c.call("qmckl_set_kernel_timing_device", False)
    fused = sum(launches(c, f"ao_mo_fused_{f}") for f in SPLIT)
    return out, fused, launches(c, "ao_vgl_f32")


def dense_waters(mo_num, walk_num=3):
    # 8 waters (520 AOs), coefficients without zeros so that no column plan
    # is kept; 80 electrons x 3 walkers = 240 points (3.75 tiles of 64)
    return S.water_cluster(n_side=2, walk_num=walk_num, mo_num=mo_num, zero_fraction=0.0,
                           name=f"water8_dense_{mo_num}mo")


@pytest.fixture(scope="module")
def ref40(oracle_lib):
    s = dense_waters(40)
    return s, oracle_lib.mo_vgl(s)


@pytest.fixture(scope="module")
def ref100(oracle_lib):
    s = dense_waters(100)
    return s, oracle_lib.mo_vgl(s)


@pytest.fixture(scope="module")
def ref450(oracle_lib):
    s = dense_waters(450)
    return s, oracle_lib.mo_vgl(s)


@pytest.mark.parametrize("form", list(SPLIT))
@pytest.mark.parametrize("which", ["ref40", "ref100"])
def test_fused_matches_oracle(form, which, request):
    """40 MOs (one partial group) and 100 MOs (two groups) of a 520-AO
    basis: AUTO runs the AO operand kernel and a stored-operand contraction,
    FUSED runs the fused kernel alone; both within the form's bar of the
    oracle and of each other."""
    s, ref = request.getfixturevalue(which)
    P = s.point_num
    with Context(0) as c:
        c.set_system(s)
        c.call("qmckl_set_mo_basis_precision_device", SPLIT[form])
        stored, nf, nao = timed_mo_vgl(c, P)
        assert nf == 0 and nao == 1, (nf, nao)
        c.call("qmckl_set_mo_basis_gemm_tile_device", FUSED)
        got, nf, nao = timed_mo_vgl(c, P)
        assert nf == 1 and nao == 0, (nf, nao)
    e, e_stored, e_between = row_err(got, ref), row_err(stored, ref), row_err(got, stored)
    say(f"{s.name} {form}: fused {e:.2e}, stored {e_stored:.2e}, between {e_between:.2e}")
    assert e <= BAR[form] and e_between <= BAR[form], (e, e_between)


@pytest.mark.parametrize("form", list(SPLIT))
def test_forced_fused_many_groups(form, ref450):
    """450 MOs: FUSED runs eight MO groups (the last of two MOs), each
    evaluating the AOs again."""
    s, ref = ref450
    P = s.point_num
    with Context(0) as c:
        c.set_system(s)
        c.call("qmckl_set_mo_basis_precision_device", SPLIT[form])
        auto, nf, nao = timed_mo_vgl(c, P)
        assert nf == 0 and nao == 1, (nf, nao)
        c.call("qmckl_set_mo_basis_gemm_tile_device", FUSED)
        got, nf, nao = timed_mo_vgl(c, P)
        assert nf == 1 and nao == 0, (nf, nao)
    e = row_err(got, ref)
    say(f"{s.name} {form} fused, 8 MO groups: {e:.2e} (stored {row_err(auto, ref):.2e})")
    assert e <= BAR[form], e


@pytest.mark.parametrize("form", list(SPLIT))
def test_fused_point_order_invariance(form, ref100):
    """The same points in another order, and with the spatial point order
    on, give bit-identical rows: each row depends only on its own point."""
    s, _ = ref100
    P = s.point_num
    perm = np.random.default_rng(11).permutation(P)
    X = np.asarray(s.elec_coord).reshape(P, 3)
    with Context(0) as c:
        c.set_system(s)
        c.call("qmckl_set_mo_basis_precision_device", SPLIT[form])
        c.call("qmckl_set_mo_basis_gemm_tile_device", FUSED)
        a = c.mo_vgl(P)
        c.call("qmckl_set_mo_basis_tile_config_device", 1, 0)
        c.touch()
        assert np.array_equal(c.mo_vgl(P), a)
        c.set_electron_coord(X[perm].reshape(np.shape(s.elec_coord)))
        b = c.mo_vgl(P)
    assert np.array_equal(b, a[perm])


def test_fused_choice_ignored_elsewhere(ref40):
    """The FP32 and INT8X5 forms have one kernel each: FUSED gives their
    AUTO results bitwise.  The autotuner keeps FUSED on a context set to it
    (results unchanged bitwise) and leaves an AUTO context off it."""
    s, _ = ref40
    P = s.point_num
    with Context(0) as c:
        c.set_system(s)
        for form in (1, 5):
            c.call("qmckl_set_mo_basis_precision_device", form)
            c.call("qmckl_set_mo_basis_gemm_tile_device", 0)
            a, nf, _ = timed_mo_vgl(c, P)
            c.call("qmckl_set_mo_basis_gemm_tile_device", FUSED)
            b, nf2, _ = timed_mo_vgl(c, P)
            assert nf == 0 and nf2 == 0 and np.array_equal(a, b), form
        c.call("qmckl_set_mo_basis_precision_device", SPLIT["fp16x2"])
        gt = ctypes.c_int32(-1)
        for choice in (FUSED, 0):
            c.call("qmckl_set_mo_basis_gemm_tile_device", choice)
            c.touch()
            ref = c.mo_vgl(P)
            c.call("qmckl_mo_basis_autotune_device")
            c.call("qmckl_get_mo_basis_gemm_tile_device", ctypes.byref(gt))
            got, nf, _ = timed_mo_vgl(c, P)
            say(f"autotuned from {choice}: gemm_tile {gt.value}, fused launches {nf}")
            assert (gt.value == FUSED) == (choice == FUSED) and nf == (choice == FUSED)
            assert np.array_equal(got, ref)
