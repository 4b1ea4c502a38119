"""Rows ranked next after the hot path, on the GPU through the C ABI:
electron-electron distances (provider, getter and low-level launcher)
against the oracle, and MO subsetting (select_mo) against the full MO set."""
import ctypes

import numpy as np
import pytest

from qmckl_sycl_amd import systems as S
from qmckl_sycl_amd.context import Context, DeviceArray, lib

from helpers import say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["n2_like", "chbrclf", "water64_2w"])
def test_ee_distance_vs_oracle(oracle_lib, name):
    s = {"n2_like": lambda: S.n2_like(walk_num=3), "chbrclf": S.chbrclf,
         "water64_2w": lambda: S.water64(walk_num=2)}[name]()
    W, E = s.walk_num, s.elec_num
    ref = oracle_lib.ee_distance(s)
    with Context(0) as c:
        c.set_system(s)
        out = DeviceArray(c, (W, E, E))
        c.call("qmckl_get_electron_ee_distance_device", out.ptr, out.size)
        got = out.numpy()
        assert c.rc("qmckl_get_electron_ee_distance_device", out.ptr, out.size - 1) == 3
        # the low-level launcher on the context's own coordinates
        coord = DeviceArray.from_numpy(c, s.coord_T())
        low = DeviceArray(c, (W, E, E))
        c.check(lib().qmckl_compute_ee_distance_device(c.handle, E, W, coord.ptr, low.ptr))
        assert np.array_equal(low.numpy(), got)
    nz = ref > 0
    rel = np.max(np.abs(got[nz] - ref[nz]) / ref[nz])
    say(f"ee_distance {name}: W={W} E={E} max rel diff {rel:.2e}")
    assert rel <= 2.3e-16
    assert np.all(got[~nz] == 0.0)


def test_ee_distance_follows_new_coordinates(oracle_lib):
    s = S.n2_like(walk_num=2)
    with Context(0) as c:
        c.set_system(s)
        out = DeviceArray(c, (s.walk_num, s.elec_num, s.elec_num))
        c.call("qmckl_get_electron_ee_distance_device", out.ptr, out.size)
        moved = s.elec_coord + 0.25
        c.set_electron_coord(moved)
        c.call("qmckl_get_electron_ee_distance_device", out.ptr, out.size)
        got = out.numpy()
    ref = oracle_lib.ee_distance(s.with_walkers(moved))
    assert np.allclose(got, ref, rtol=1e-15, atol=1e-15)


@pytest.mark.parametrize("ta,tb", [(b"T", b"T"), (b"N", b"T"), (b"T", b"N"), (b"n", b"n")])
def test_distance_launcher(oracle_lib, ta, tb):
    """qmckl_distance_device in its four layouts, with leading dimensions
    larger than needed (the padding holds NaN): with the nuclei as A and the
    points as B, C[p][n] is the oracle's en_distance [P][N] in every layout,
    as the reference's compute_en_distance obtains it with 'T','T'
    (src/qmckl_electron.cpp:326-328); the padding columns of C stay
    untouched."""
    s = S.chbrclf()
    P, N = s.point_num, s.nucl_num
    pts = s.coord_T()              # [3][P]
    nuc = np.asarray(s.nucl_coord)  # [3][N]
    ref = oracle_lib.en_distance(s)  # [P][N]

    def layout(x, t, ld):  # x [3][k] -> 'T' [3][ld] or 'N' [k][ld]
        k = x.shape[1]
        if t in (b"T", b"t"):
            out = np.full((3, ld), np.nan)
            out[:, :k] = x
        else:
            out = np.full((k, ld), np.nan)
            out[:, :3] = x.T
        return np.ascontiguousarray(out)

    lda = N + 3 if ta in (b"T", b"t") else 5
    ldb = P + 7 if tb in (b"T", b"t") else 4
    ldc = N + 2
    with Context(0) as c:
        A = DeviceArray.from_numpy(c, layout(nuc, ta, lda))
        B = DeviceArray.from_numpy(c, layout(pts, tb, ldb))
        C = DeviceArray.from_numpy(c, np.full((P, ldc), -1.0))
        c.check(lib().qmckl_distance_device(c.handle, ta, tb, N, P, A.ptr, lda, B.ptr, ldb, C.ptr, ldc))
        got = C.numpy()
    assert np.all(got[:, N:] == -1.0)  # padding of C untouched
    d = got[:, :N]
    rel = np.max(np.abs(d - ref) / np.maximum(ref, 1e-300))
    say(f"distance {ta.decode()}{tb.decode()}: max rel diff vs oracle {rel:.2e}")
    assert rel <= 2.3e-16


@pytest.mark.parametrize("form", [0, 2])
def test_select_mo(form):
    s = S.chbrclf()
    M = s.mo_num
    keep_np = np.array([1 if i % 3 != 1 else 0 for i in range(M)], dtype=np.int32)
    kept = np.nonzero(keep_np)[0]
    with Context(0) as c:
        c.set_system(s)
        c.call("qmckl_set_mo_basis_precision_device", form)
        full = c.mo_vgl(s.point_num)
        L = lib()
        keep = keep_np.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        assert L.qmckl_mo_basis_select_mo_device(c.handle, None, M) == 2
        assert L.qmckl_mo_basis_select_mo_device(c.handle, keep, M - 1) == 3
        none = np.zeros(M, dtype=np.int32)
        assert L.qmckl_mo_basis_select_mo_device(
            c.handle, none.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), M) == 2
        c.check(L.qmckl_mo_basis_select_mo_device(c.handle, keep, M))
        assert c.mo_num() == len(kept)
        sub = c.mo_vgl(s.point_num)
        value = c.mo_value(s.point_num)
    assert sub.shape == (s.point_num, 5, len(kept))
    if form == 0:
        assert np.array_equal(sub, full[:, :, kept])
    else:
        d = np.abs(sub - full[:, :, kept]).max() / np.abs(full).max()
        say(f"select_mo fp16x2: max diff vs full set {d:.2e}")
        assert d <= 1e-6
    if form == 0:
        assert np.array_equal(value, sub[:, 0, :]) or np.allclose(value, sub[:, 0, :], rtol=0, atol=1e-12)
    else:
        # mo_value is the FP64 contraction in every precision form
        assert np.abs(value - full[:, 0, kept]).max() / np.abs(full).max() <= 1e-6
