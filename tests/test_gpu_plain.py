"""FP64 plain path on the GPU, through the C ABI, against the oracle and the
reference's golden vectors.

Tolerances: the reference's 1e-14 absolute on its ChBrClF known answers;
elsewhere 1e-12 relative (1e-10 on Laplacians) to the sum of absolute
values of each entry's terms (see helpers.RTOL_VGL)."""
import ctypes

import numpy as np
import pytest

from qmckl_sycl_amd import systems as S
from qmckl_sycl_amd.context import Context, DeviceArray, lib

from helpers import assert_vgl_close, device_chain, say

pytestmark = pytest.mark.gpu

AO_VALUE_KAT = [1.020298798341620e-08, 1.516643537739178e-08,
                -4.686370882518819e-09, 7.514816980753531e-09]
AO_VGL_KAT = [1.020298798341620e-08, -4.928035238010602e-08, -4.691009312035986e-08,
              1.449504046436699e-08, 4.296442111843973e-07]


@pytest.fixture(scope="module")
def chb():
    return S.chbrclf()


@pytest.fixture(scope="module")
def chb_ctx(chb):
    say("setting up ChBrClF context")
    c = Context(0).set_system(chb)
    yield c
    c.destroy()


def test_chbrclf_ao_vgl_kat(chb_ctx):
    ao = chb_ctx.ao_vgl(136)
    err = np.max(np.abs(ao[26, :, 219] - AO_VGL_KAT))
    say("ao_vgl KAT max abs err", err)
    assert err <= 1e-14


def test_chbrclf_mo_reference(chb_ctx, chb, oracle_lib):
    mo = chb_ctx.mo_vgl(136)
    ref = S.chbrclf_mo_reference()
    abs_err = np.max(np.abs(mo - ref))
    say("mo_vgl vs tests/mo_reference.txt: max abs err", abs_err)
    # tests/test_qmckl_mo.cpp:382-405 checks 1e-14 absolute
    assert abs_err <= 1e-14
    # the file holds 16 fixed decimals: allow its 5e-17 quantisation on top
    # of the relative bar
    _, mag = oracle_lib.mo_vgl(chb, with_mag=True)
    d = np.abs(mo - ref)
    assert np.all(d <= 1e-12 * mag + 1e-16)


def test_chbrclf_ao_value_kat(chb):
    # as tests/test_qmckl_ao.cpp:407-447: the 68 points of walker 0 via set_point
    with Context(0) as c:
        c.set_system(chb, walkers=False)
        pts = DeviceArray.from_numpy(c, chb.elec_coord[0])
        c.call("qmckl_set_point_device", b"N", 68, pts.ptr, 3 * 68)
        ao = c.ao_value(68)
        err = np.max(np.abs(ao[26, 219:223] - AO_VALUE_KAT))
        say("ao_value KAT max abs err", err)
        assert err <= 1e-14


def test_chbrclf_mo_value_matches_mo_vgl(chb):
    # tests/test_qmckl_mo.cpp:279-356: mo_value from scratch, after touch, and cached
    with Context(0) as c:
        c.set_system(chb)
        c.touch()
        mo_value = c.mo_value(136)
        mo = c.mo_vgl(136)
        assert np.max(np.abs(mo[:, 0, :] - mo_value)) <= 1e-12
        c.touch()
        mo_value2 = c.mo_value(136)   # from mo_vgl? no: mo_vgl is stale after touch
        assert np.max(np.abs(mo[:, 0, :] - mo_value2)) <= 1e-12
        mo2 = c.mo_vgl(136)
        mo_value3 = c.mo_value(136)   # cached
        assert np.array_equal(mo_value3, mo_value2)
        assert np.max(np.abs(mo2[:, 0, :] - mo_value3)) <= 1e-12


def test_chbrclf_chain_vs_oracle(chb_ctx, chb, oracle_lib):
    out = device_chain(chb_ctx, 136, want=("en", "shell", "ao", "ao_value", "mo"))
    en = oracle_lib.en_distance(chb)
    assert np.max(np.abs(out["en"] - en) / en) < 1e-15
    sh = oracle_lib.shell_vgl(chb)
    d = np.abs(out["shell"] - sh)
    assert np.all(d <= 1e-12 * np.abs(sh).max(axis=2, keepdims=True) + 1e-300)
    ao, am = oracle_lib.ao_vgl(chb, with_mag=True)
    assert_vgl_close(out["ao"], ao, am, what="chbrclf ao_vgl")
    assert np.array_equal(out["ao_value"], out["ao"][:, 0, :])
    mo, mm = oracle_lib.mo_vgl_from_ao(chb.mo_coef, ao, am)
    assert_vgl_close(out["mo"], mo, mm, what="chbrclf mo_vgl")


@pytest.mark.parametrize("walkers", [1, 3])
def test_n2_like_vs_oracle(oracle_lib, walkers):
    s = S.n2_like(walk_num=walkers)
    with Context(0) as c:
        c.set_system(s)
        out = device_chain(c, s.point_num)
    ao, am = oracle_lib.ao_vgl(s, with_mag=True)
    mo, mm = oracle_lib.mo_vgl_from_ao(s.mo_coef, ao, am)
    assert_vgl_close(out["ao"], ao, am, what=f"n2 ao_vgl w={walkers}")
    assert_vgl_close(out["mo"], mo, mm, what=f"n2 mo_vgl w={walkers}")
    en = oracle_lib.en_distance(s)
    assert np.max(np.abs(out["en"] - en) / en) < 1e-15


def test_chbrclf_sized_full_vs_oracle(oracle_lib):
    s = S.chbrclf_sized()   # 256 walkers, 17,152 points, 145 AOs, 145 MOs
    say("chbrclf_sized", s.summary())
    with Context(0) as c:
        c.set_system(s)
        out = device_chain(c, s.point_num)
    ao, am = oracle_lib.ao_vgl(s, with_mag=True)
    mo, mm = oracle_lib.mo_vgl_from_ao(s.mo_coef, ao, am)
    assert_vgl_close(out["ao"], ao, am, what="chbrclf_sized ao_vgl")
    assert_vgl_close(out["mo"], mo, mm, what="chbrclf_sized mo_vgl")
    en = oracle_lib.en_distance(s)
    assert np.max(np.abs(out["en"] - en) / en) < 1e-15


@pytest.mark.parametrize("P_walk,mo_num", [(1, 1), (1, 17), (5, 333), (13, 64)])
def test_ragged_shapes(oracle_lib, P_walk, mo_num):
    # point counts that are not multiples of the 64-point tile, MO counts that
    # are not multiples of 16 and above the 320 columns of one workgroup
    base = S.n2_like(walk_num=P_walk, seed=2024)
    rng = np.random.default_rng(mo_num)
    s = S.System(**{**base.__dict__})
    s.mo_coef = rng.normal(0, 0.3, size=(mo_num, base.ao_num))
    with Context(0) as c:
        c.set_system(s)
        out = device_chain(c, s.point_num, want=("ao", "mo", "mo_value"))
    ao, am = oracle_lib.ao_vgl(s, with_mag=True)
    mo, mm = oracle_lib.mo_vgl_from_ao(s.mo_coef, ao, am)
    assert_vgl_close(out["mo"], mo, mm, what=f"ragged P={s.point_num} M={mo_num}")
    assert np.max(np.abs(out["mo_value"] - out["mo"][:, 0, :])) <= 1e-12 * (1 + np.abs(mo).max())


@pytest.mark.parametrize("nucl_num", [2, 14])
def test_high_l_and_screened_stages_vs_oracle(oracle_lib, nucl_num):
    """Shells l = 0..6 on nuclei 80 bohr apart along z (84 AOs each: 168 AOs,
    not a multiple of 32 or 64, for the lane-per-AO ao_vgl kernel; 1,176 AOs,
    a multiple of 8 above 1024, for the lane-per-point one), 100 points (not a
    multiple of 64), the first 70 near nucleus 0 and the rest near the last
    nucleus.  The points of the first 64-point tile see every other nucleus
    screened, so its AOs from 84 on (whole 64-AO stages, the last one
    part-filled) are exact zeros there; l = 4..6 take the kernels'
    loop-count paths; the l = 6 shell of nucleus 0 straddles a 64-AO stage
    boundary."""
    rng = np.random.default_rng(77)
    xyz = np.zeros((nucl_num, 3))
    xyz[:, 2] = 80.0 * np.arange(nucl_num)
    bb = S._BasisBuilder()
    for _ in range(nucl_num):
        bb.add_nucleus([(l, int(rng.integers(1, 5))) for l in range(7)], rng, (0.1, 50.0))
    arr = bb.arrays()
    A = int(sum(S.ncart(int(l)) for l in arr["shell_ang_mom"]))
    assert A == 84 * nucl_num
    elec = S._electrons_near(rng, xyz, [0] * 70 + [nucl_num - 1] * 30, 1)
    s = S.System(name="far_chain_l6", charge=np.ones(nucl_num), nucl_coord=np.ascontiguousarray(xyz.T),
                 mo_coef=rng.normal(0.0, 0.3, size=(40, A)), elec_up_num=50, elec_dn_num=50,
                 elec_coord=elec, **arr)
    s.validate()
    with Context(0) as c:
        c.set_system(s)
        out = device_chain(c, s.point_num, want=("ao", "ao_value", "mo"))
        c.touch()
        value_only = c.ao_value(s.point_num)  # ao_vgl is stale: the value-only kernel runs
    ao, am = oracle_lib.ao_vgl(s, with_mag=True)
    assert np.all(ao[:64, :, 84:] == 0.0), "the oracle screens the far nuclei for the first 64 points"
    assert np.all(out["ao"][:64, :, 84:] == 0.0)
    assert np.all(value_only[:64, 84:] == 0.0)
    assert_vgl_close(value_only[:, None, :], ao[:, :1, :], am[:, :1, :], rtol=1e-12,
                     what=f"far chain of {nucl_num}, l<=6 ao_value")
    assert_vgl_close(out["ao"], ao, am, what=f"far chain of {nucl_num}, l<=6 ao_vgl")
    assert np.max(np.abs(out["ao_value"] - out["ao"][:, 0, :])) <= 1e-12 * np.abs(ao).max()
    mo, mm = oracle_lib.mo_vgl_from_ao(s.mo_coef, ao, am)
    assert_vgl_close(out["mo"], mo, mm, what=f"far chain of {nucl_num}, l<=6 mo_vgl")


def test_fp64_point_order_keeps_rows(oracle_lib):
    """FP64 ao_vgl and ao_value over spatially ordered 64-point tiles (the
    default at this size: 4,160 AOs, 5,120 points) against the caller's
    order: every point's rows, and the MOs after them, are bitwise the same,
    the value-only kernel's rows equal the value rows of ao_vgl, and a sample
    of points matches the oracle."""
    s = S.water64(walk_num=8)
    P = s.point_num
    with Context(0) as c:
        c.set_system(s)
        c.call("qmckl_set_mo_basis_tile_config_device", 1, 0)
        ao1, mo1 = c.ao_vgl(P), c.mo_vgl(P)
        c.touch()
        av1 = c.ao_value(P)  # ao_vgl is stale: the value-only kernel runs
        c.call("qmckl_set_mo_basis_tile_config_device", 0, 0)
        c.touch()
        av0 = c.ao_value(P)
        c.touch()
        ao0, mo0 = c.ao_vgl(P), c.mo_vgl(P)
    assert np.array_equal(ao1, ao0)
    assert np.array_equal(mo1, mo0)
    assert np.array_equal(av1, av0)
    assert np.array_equal(av1, ao1[:, 0, :])
    idx = np.sort(np.random.default_rng(3).choice(P, size=64, replace=False))
    coord = np.ascontiguousarray(s.coord_T()[:, idx])
    ao, am = oracle_lib.ao_vgl(s, coord, with_mag=True)
    assert_vgl_close(ao1[idx], ao, am, what="water64 ordered-tile ao_vgl")


def test_walker_count_changes_reallocate(oracle_lib):
    s = S.n2_like(walk_num=2, seed=5)
    with Context(0) as c:
        c.set_system(s)
        for w in (2, 7, 1, 4):
            rng = np.random.default_rng(w)
            coords = rng.normal(0, 1.2, size=(w, s.elec_num, 3))
            c.set_electron_coord(coords)
            sw = s.with_walkers(coords)
            got = c.mo_vgl(sw.point_num)
            ao, am = oracle_lib.ao_vgl(sw, with_mag=True)
            mo, mm = oracle_lib.mo_vgl_from_ao(sw.mo_coef, ao, am)
            assert_vgl_close(got, mo, mm, what=f"walkers={w}")
            assert np.array_equal(c.electron_num(), s.elec_num)


def test_layouts_and_roundtrips(chb):
    with Context(0) as c:
        c.set_system(chb)
        P = chb.point_num
        # electron coordinates back in both layouts
        buf = DeviceArray(c, (3 * P,))
        c.call("qmckl_get_electron_coord_device", b"N", buf.ptr, 3 * P)
        assert np.array_equal(buf.numpy().reshape(P, 3), chb.elec_coord.reshape(P, 3))
        c.call("qmckl_get_electron_coord_device", b"T", buf.ptr, 3 * P)
        assert np.array_equal(buf.numpy().reshape(3, P), chb.coord_T())
        assert c.rc("qmckl_get_point_device", b"N", buf.ptr, 3 * P - 1) == 3
        # nucleus coordinates
        N = chb.nucl_num
        nb = DeviceArray(c, (3 * N,))
        c.call("qmckl_get_nucleus_coord_device", b"N", nb.ptr, 3 * N)
        assert np.array_equal(nb.numpy().reshape(N, 3), chb.nucl_coord.T)
        c.call("qmckl_get_nucleus_charge_device", nb.ptr, N)
        assert np.array_equal(nb.numpy()[:N], chb.charge)
        # AO arrays come back as set
        for name, arr, dt in [("shell_ang_mom", chb.shell_ang_mom, np.int32),
                              ("shell_prim_num", chb.shell_prim_num, np.int64),
                              ("shell_prim_index", chb.shell_prim_index, np.int64),
                              ("nucleus_index", chb.nucleus_index, np.int64),
                              ("nucleus_shell_num", chb.nucleus_shell_num, np.int64),
                              ("shell_factor", chb.shell_factor, np.float64),
                              ("exponent", chb.exponent, np.float64),
                              ("coefficient", chb.coefficient, np.float64),
                              ("prim_factor", chb.prim_factor, np.float64),
                              ("ao_factor", chb.ao_factor, np.float64)]:
            d = DeviceArray(c, arr.shape, dt)
            c.call(f"qmckl_get_ao_basis_{name}_device", d.ptr, arr.size)
            assert np.array_equal(d.numpy(), arr), name
            assert c.rc(f"qmckl_get_ao_basis_{name}_device", d.ptr, arr.size - 1) == 3


def test_points_transposed_layout(chb, oracle_lib):
    with Context(0) as c:
        c.set_system(chb, walkers=False)
        P = chb.point_num
        pts = DeviceArray.from_numpy(c, chb.coord_T())
        c.call("qmckl_set_point_device", b"T", P, pts.ptr, 3 * P)
        mo = c.mo_vgl(P)
    ref = S.chbrclf_mo_reference()
    assert np.max(np.abs(mo - ref)) <= 1e-13


def test_getters_check_size_max(chb_ctx):
    c = chb_ctx
    d = DeviceArray(c, (5 * 136 * 263,))
    assert c.rc("qmckl_get_ao_basis_ao_vgl_device", d.ptr, 5 * 136 * 263 - 1) == 3
    assert c.rc("qmckl_get_mo_basis_mo_vgl_device", d.ptr, 5 * 136 * 224 - 1) == 3
    assert c.rc("qmckl_get_electron_en_distance_device", d.ptr, 136 * 5 - 1) == 3
    assert c.rc("qmckl_get_ao_basis_ao_vgl_inplace_device", d.ptr, 10) == 3
    d.free()


def test_inplace_getters_match(chb_ctx):
    c = chb_ctx
    mo = c.mo_vgl(136)
    ao = c.ao_vgl(136)
    d = DeviceArray(c, (136, 5, 263))
    c.call("qmckl_get_ao_basis_ao_vgl_inplace_device", d.ptr, d.size)
    assert np.array_equal(d.numpy(), ao)
    e = DeviceArray(c, (136, 5, 224))
    c.call("qmckl_get_mo_basis_mo_vgl_inplace_device", e.ptr, e.size)
    assert np.array_equal(e.numpy(), mo)
    # the context's own cache still answers correctly after the touch
    assert np.array_equal(c.mo_vgl(136), mo)
    d.free()
    e.free()


def test_low_level_launchers(chb_ctx, chb, oracle_lib):
    c = chb_ctx
    L = lib()
    P, N, S_, Np, A, M = 136, chb.nucl_num, chb.shell_num, chb.prim_num, chb.ao_num, chb.mo_num
    cn, lmax, rng = oracle_lib.finalize(chb)
    dev = {k: DeviceArray.from_numpy(c, v) for k, v in dict(
        coord=chb.coord_T(), nucl=chb.nucl_coord, nsn=chb.nucleus_shell_num, nidx=chb.nucleus_index,
        rng=rng, spi=chb.shell_prim_index, spn=chb.shell_prim_num, ex=chb.exponent, cn=cn,
        lmax=lmax, sam=chb.shell_ang_mom, af=chb.ao_factor,
        ct=np.ascontiguousarray(chb.mo_coef.T)).items()}
    shell = DeviceArray(c, (P, 5, S_))
    c.check(L.qmckl_compute_ao_basis_shell_gaussian_vgl_device(
        c.handle, Np, S_, P, N, dev["nsn"].ptr, dev["nidx"].ptr, dev["rng"].ptr, dev["spi"].ptr,
        dev["spn"].ptr, dev["coord"].ptr, dev["nucl"].ptr, dev["ex"].ptr, dev["cn"].ptr, shell.ptr))
    sh_ref = oracle_lib.shell_vgl(chb)
    assert np.max(np.abs(shell.numpy() - sh_ref) / (np.abs(sh_ref).max() + 1e-300)) < 1e-13
    ao = DeviceArray(c, (P, 5, A))
    c.check(L.qmckl_compute_ao_vgl_gaussian_device(
        c.handle, A, S_, P, N, dev["coord"].ptr, dev["nucl"].ptr, dev["nidx"].ptr, dev["nsn"].ptr,
        dev["rng"].ptr, dev["lmax"].ptr, dev["sam"].ptr, dev["af"].ptr, shell.ptr, ao.ptr))
    ao_ref, am = oracle_lib.ao_vgl(chb, with_mag=True)
    assert_vgl_close(ao.numpy(), ao_ref, am, what="low-level ao_vgl")
    assert np.max(np.abs(ao.numpy()[26, :, 219] - AO_VGL_KAT)) <= 1e-14
    mo = DeviceArray(c, (P, 5, M))
    c.check(L.qmckl_compute_mo_basis_mo_vgl_device(c.handle, A, M, P, dev["ct"].ptr, ao.ptr, mo.ptr))
    assert np.max(np.abs(mo.numpy() - S.chbrclf_mo_reference())) <= 1e-13
    en = DeviceArray(c, (P, N))
    c.check(L.qmckl_compute_en_distance_device(c.handle, P, N, dev["coord"].ptr, dev["nucl"].ptr, en.ptr))
    en_ref = oracle_lib.en_distance(chb)
    assert np.max(np.abs(en.numpy() - en_ref) / en_ref) < 1e-15
    for d in list(dev.values()) + [shell, ao, mo, en]:
        d.free()


def test_deterministic(chb_ctx):
    c = chb_ctx
    c.touch()
    a = c.mo_vgl(136)
    c.touch()
    b = c.mo_vgl(136)
    assert np.array_equal(a, b)


def test_water64_sampled_points_vs_oracle(oracle_lib):
    """The 64-water system at its full 128 walkers (81,920 points), FP64 plain
    path; every point is independent, so a sample of points is checked
    against the oracle at full size."""
    s = S.water64(walk_num=128)
    say("water64", s.summary())
    with Context(0) as c:
        c.set_system(s)
        mo = c.mo_vgl(s.point_num)
    rng = np.random.default_rng(0)
    idx = np.sort(rng.choice(s.point_num, size=96, replace=False))
    coord = np.ascontiguousarray(s.coord_T()[:, idx])
    ao, am = oracle_lib.ao_vgl(s, coord, with_mag=True)
    ref, mm = oracle_lib.mo_vgl_from_ao(s.mo_coef, ao, am)
    assert_vgl_close(mo[idx], ref, mm, what="water64 sampled mo_vgl")
    assert np.all(np.isfinite(mo))
