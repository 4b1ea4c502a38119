"""Pins the FP64 CPU oracle (CPU only).

Against the reference's own golden vectors (ChBrClF):
  * ao_value[26][219..222], tests/test_qmckl_ao.cpp:440-447 (P = 68), 1e-14 abs;
  * ao_vgl[26][0..4][219], tests/test_qmckl_mo.cpp:250-259 (P = 136), 1e-14 abs;
  * the full mo_vgl[136][5][224] of tests/mo_reference.txt,
    tests/test_qmckl_mo.cpp:382-405, 1e-14 abs;
  * mo_vgl[:,0,:] == mo_value, tests/test_qmckl_mo.cpp:301-356, 1e-12 abs.
Against mathematics, for what those leave open:
  * en_distance: numpy norm and exact rational arithmetic;
  * radial part: a one-primitive s shell in closed form, its screen, r = 0;
  * product-rule assembly: central differences of the value give the
    gradient and the Laplacian (shells up to l = 4);
  * invariances: translation of the whole system; a 90-degree rotation
    permutes p components;
  * the MO contraction: exact rational arithmetic on small inputs;
  * the adjoint: central differences of the oracle forward in nuclear
    coordinates, exponents and MO coefficients;
  * the generators' normalisation: the ChBrClF prim_factor and ao_factor
    arrays and 1D quadrature of contracted shells.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from qmckl_sycl_amd import systems as S

AO_VALUE_KAT = [1.020298798341620e-08, 1.516643537739178e-08,
                -4.686370882518819e-09, 7.514816980753531e-09]
AO_VGL_KAT = [1.020298798341620e-08, -4.928035238010602e-08, -4.691009312035986e-08,
              1.449504046436699e-08, 4.296442111843973e-07]


@pytest.fixture(scope="module")
def chbrclf():
    return S.chbrclf()


def test_ao_value_kat(oracle_lib, chbrclf):
    # test_qmckl_ao.cpp sets the 68 points of walker 0 only (407-419)
    coord = np.ascontiguousarray(chbrclf.elec_coord[0].T)
    ao = oracle_lib.ao_value(chbrclf, coord)
    assert ao.shape == (68, 263)
    assert np.max(np.abs(ao[26, 219:223] - AO_VALUE_KAT)) <= 1e-14


def test_ao_vgl_kat(oracle_lib, chbrclf):
    ao = oracle_lib.ao_vgl(chbrclf)
    assert ao.shape == (136, 5, 263)
    assert np.max(np.abs(ao[26, :, 219] - AO_VGL_KAT)) <= 1e-14


def test_mo_reference(oracle_lib, chbrclf):
    mo = oracle_lib.mo_vgl(chbrclf)
    ref = S.chbrclf_mo_reference()
    assert mo.shape == ref.shape == (136, 5, 224)
    assert np.max(np.abs(mo - ref)) <= 1e-14


def test_mo_value_invariant(oracle_lib, chbrclf):
    ao_v = oracle_lib.ao_value(chbrclf)
    mo_value = ao_v @ chbrclf.mo_coef.T
    mo = oracle_lib.mo_vgl(chbrclf)
    assert np.max(np.abs(mo[:, 0, :] - mo_value)) <= 1e-12
    ao = oracle_lib.ao_vgl(chbrclf)
    assert np.array_equal(ao[:, 0, :], ao_v)


def test_en_distance(oracle_lib, chbrclf):
    d = oracle_lib.en_distance(chbrclf)
    r = chbrclf.elec_coord.reshape(-1, 3)
    R = chbrclf.nucl_coord.T
    ref = np.linalg.norm(r[:, None, :] - R[None, :, :], axis=2)
    assert np.max(np.abs(d - ref) / ref) < 4e-16
    # exact rational check on a few entries (correctly rounded sqrt of an
    # exactly-rounded r^2 is within 1 ulp of the true distance)
    for p, n in [(0, 0), (26, 4), (135, 2)]:
        s2 = sum((Fraction(float(r[p, k])) - Fraction(float(R[n, k]))) ** 2 for k in range(3))
        exact = math.sqrt(float(s2))
        assert abs(d[p, n] - exact) <= 2 * np.spacing(exact)


def test_ee_distance(oracle_lib, chbrclf):
    """ee_distance per walker: numpy brute force, exact rational r^2 with a
    correctly rounded sqrt, symmetry and a zero diagonal."""
    d = oracle_lib.ee_distance(chbrclf)
    W, E = chbrclf.walk_num, chbrclf.elec_num
    assert d.shape == (W, E, E)
    r = chbrclf.elec_coord  # [walk][elec][3]
    ref = np.linalg.norm(r[:, None, :, :] - r[:, :, None, :], axis=3)  # [w][j][i]
    nz = ref > 0
    assert np.max(np.abs(d[nz] - ref[nz]) / ref[nz]) < 4e-16
    assert np.array_equal(d, d.transpose(0, 2, 1))
    assert np.all(np.diagonal(d, axis1=1, axis2=2) == 0.0)
    for w, j, i in [(0, 0, 1), (1, 26, 40), (1, 67, 3)]:
        s2 = sum((Fraction(float(r[w, i, k])) - Fraction(float(r[w, j, k]))) ** 2 for k in range(3))
        exact = math.sqrt(float(s2))
        assert abs(d[w, j, i] - exact) <= 2 * np.spacing(exact)


def _one_shell_system(alpha, coef, l, points, expo_extra=()):
    """1 nucleus at the origin with one contracted shell of angular momentum l."""
    alpha = np.atleast_1d(np.asarray(alpha, dtype=np.float64))
    coef = np.atleast_1d(np.asarray(coef, dtype=np.float64))
    A = S.ncart(l)
    pts = np.asarray(points, dtype=np.float64).reshape(1, -1, 3)
    return S.System(
        name="one_shell", charge=np.array([1.0]), nucl_coord=np.zeros((3, 1)),
        nucleus_index=np.array([0], dtype=np.int64), nucleus_shell_num=np.array([1], dtype=np.int64),
        shell_ang_mom=np.array([l], dtype=np.int32),
        shell_prim_num=np.array([alpha.size], dtype=np.int64),
        shell_prim_index=np.array([0], dtype=np.int64), shell_factor=np.array([1.0]),
        exponent=alpha, coefficient=coef, prim_factor=np.ones_like(alpha),
        ao_factor=np.ones(A), mo_coef=np.eye(A), elec_up_num=pts.shape[1], elec_dn_num=0,
        elec_coord=pts)


def test_single_s_primitive_closed_form(oracle_lib):
    a, c = 0.7, 1.3
    pts = [[0.3, -0.2, 0.5], [1.0, 1.0, 1.0], [0.0, 0.0, 0.0], [2.5, -1.0, 0.4]]
    s = _one_shell_system(a, c, 0, pts)
    ao = oracle_lib.ao_vgl(s)
    for p, r in enumerate(np.asarray(pts)):
        r2 = float(r @ r)
        R = c * math.exp(-a * r2)
        np.testing.assert_allclose(ao[p, 0, 0], R, rtol=1e-15)
        np.testing.assert_allclose(ao[p, 1:4, 0], -2 * a * r * R, rtol=1e-15, atol=1e-300)
        np.testing.assert_allclose(ao[p, 4, 0], (4 * a * a * r2 - 6 * a) * R, rtol=1e-14)
    # at r = 0 the Laplacian is -6 a c
    assert ao[2, 4, 0] == pytest.approx(-6 * a * c, rel=1e-15)


def test_shell_vgl_layout_closed_form(oracle_lib):
    """qo_shell_vgl [P][5][S] against the radial sums in closed form on two
    nuclei with shells of several exponents, normalisation factors and
    angular momenta: R = sum_i c_i f_i g_s exp(-a_i r^2), grad R = -2 sum a_i
    (...) r, lap R = sum (4 a_i^2 r^2 - 6 a_i)(...), zero past either screen
    (src/qmckl_ao_sycl.cpp:53-55, 72-84)."""
    nuc = np.array([[0.0, 0.0, 0.0], [1.4, -0.3, 0.8]])
    shells = [  # (nucleus, l, exponents, coefficients, prim factors, shell factor)
        (0, 0, [5.0, 0.9], [0.4, 0.7], [1.1, 0.9], 1.3),
        (0, 1, [1.7], [1.0], [2.0], 0.5),
        (1, 2, [0.35, 3.1, 12.0], [0.2, -0.5, 0.8], [1.0, 1.2, 0.7], 2.0),
        (1, 0, [40.0], [1.5], [1.0], 1.0),
    ]
    expo = np.concatenate([np.array(sh[2]) for sh in shells])
    coef = np.concatenate([np.array(sh[3]) for sh in shells])
    pf = np.concatenate([np.array(sh[4]) for sh in shells])
    nprim = np.array([len(sh[2]) for sh in shells], dtype=np.int64)
    pidx = np.concatenate([[0], np.cumsum(nprim)[:-1]]).astype(np.int64)
    lvals = np.array([sh[1] for sh in shells], dtype=np.int32)
    A = int(sum(S.ncart(int(l)) for l in lvals))
    rng = np.random.default_rng(3)
    pts = np.concatenate([rng.normal(0.4, 1.2, size=(9, 3)), [[0.0, 0.0, 0.0], [1.4, -0.3, 0.8], [20.0, 0.0, 0.0]]])
    s = S.System(
        name="two_nuclei", charge=np.array([1.0, 1.0]), nucl_coord=np.ascontiguousarray(nuc.T),
        nucleus_index=np.array([0, 2], dtype=np.int64), nucleus_shell_num=np.array([2, 2], dtype=np.int64),
        shell_ang_mom=lvals, shell_prim_num=nprim, shell_prim_index=pidx,
        shell_factor=np.array([sh[5] for sh in shells]), exponent=expo, coefficient=coef, prim_factor=pf,
        ao_factor=np.ones(A), mo_coef=np.eye(A), elec_up_num=len(pts), elec_dn_num=0,
        elec_coord=pts.reshape(1, -1, 3))
    got = oracle_lib.shell_vgl(s)
    assert got.shape == (len(pts), 5, len(shells))
    cut = -math.log(1e-12)
    # nucleus range: the largest 1/alpha of its primitives
    rng_n = [max(1.0 / a for sh in shells if sh[0] == n for a in sh[2]) for n in range(2)]
    want = np.zeros_like(got)
    for p, r in enumerate(pts):
        for k, (n, l, al, c, f, g) in enumerate(shells):
            d = r - nuc[n]
            r2 = float(d @ d)
            if r2 > cut * rng_n[n]:
                continue
            for a, ci, fi in zip(al, c, f):
                if a * r2 > cut:
                    continue
                v = ci * fi * g * math.exp(-a * r2)
                want[p, 0, k] += v
                want[p, 1:4, k] += -2.0 * a * v * d
                want[p, 4, k] += (4.0 * a * a * r2 - 6.0 * a) * v
    scale = np.maximum(np.abs(want), 1e-300)
    assert np.all((want == 0.0) == (got == 0.0)), "screened entries differ"
    assert np.all(np.abs(got - want) <= 1e-13 * scale + 1e-300), np.max(np.abs(got - want) / scale)
    # the far point is past both nuclei's range: all zero; the nuclei
    # themselves have zero gradients
    assert np.all(got[-1] == 0.0)
    assert np.all(got[-3, 1:4, :2] == 0.0) and np.all(got[-2, 1:4, 2:] == 0.0)


def test_primitive_screen_gives_exact_zero(oracle_lib):
    a = 3.0
    r_cut = math.sqrt(27.631021115928547 / a)
    pts = [[r_cut * 1.0001, 0, 0], [r_cut * 0.999, 0, 0]]
    s = _one_shell_system(a, 1.0, 1, pts)
    ao = oracle_lib.ao_vgl(s)
    assert np.all(ao[0] == 0.0)
    assert ao[1, 0, 0] != 0.0


def _fd_system(seed=7, lmax=4):
    rng = np.random.default_rng(seed)
    bb = S._BasisBuilder()
    xyz = np.array([[0.0, 0.0, 0.0], [1.1, -0.4, 0.6]])
    for _ in range(2):
        bb.add_nucleus([(l, int(rng.integers(1, 4))) for l in range(lmax + 1)], rng, (0.2, 3.0))
    arr = bb.arrays()
    A = int(sum(S.ncart(int(l)) for l in arr["shell_ang_mom"]))
    elec = (rng.normal(0.0, 0.8, size=(1, 12, 3)) + 0.3)
    return S.System(name="fd", charge=np.array([1.0, 1.0]), nucl_coord=np.ascontiguousarray(xyz.T),
                    mo_coef=rng.normal(0, 0.3, size=(7, A)), elec_up_num=12, elec_dn_num=0,
                    elec_coord=elec, **arr)


def test_assembly_matches_finite_differences(oracle_lib):
    s = _fd_system()
    coord = s.coord_T()
    ao = oracle_lib.ao_vgl(s, coord)
    h = 1e-4
    lap = np.zeros_like(ao[:, 0, :])
    v0 = ao[:, 0, :]
    for d in range(3):
        cp, cm = coord.copy(), coord.copy()
        cp[d] += h
        cm[d] -= h
        vp = oracle_lib.ao_value(s, cp)
        vm = oracle_lib.ao_value(s, cm)
        grad = (vp - vm) / (2 * h)
        scale = np.abs(ao[:, 1 + d, :]).max(axis=1, keepdims=True)
        assert np.max(np.abs(grad - ao[:, 1 + d, :]) / scale) < 1e-6
        lap += (vp - 2 * v0 + vm) / (h * h)
    scale = np.abs(ao[:, 4, :]).max(axis=1, keepdims=True)
    assert np.max(np.abs(lap - ao[:, 4, :]) / scale) < 1e-5


def test_translation_invariance(oracle_lib, chbrclf):
    shift = np.array([0.37, -1.25, 2.5])
    s2 = S.System(**{**chbrclf.__dict__})
    s2.nucl_coord = chbrclf.nucl_coord + shift[:, None]
    s2.elec_coord = chbrclf.elec_coord + shift
    a1, m1 = oracle_lib.ao_vgl(chbrclf, with_mag=True)
    a2 = oracle_lib.ao_vgl(s2)
    # r - R changes in its last bits under the shift; exp(-a r^2) with
    # a up to 1e7 amplifies that to ~1e-12 relative.  A wrong sign or index
    # would show at O(1).
    assert np.max(np.abs(a1 - a2) / (m1 + 1e-300)) < 1e-10


def test_rotation_permutes_p_components(oracle_lib):
    a = np.array([0.9, 0.3])
    pts = np.array([[0.4, -0.3, 0.8], [1.2, 0.5, -0.7], [-0.6, 0.9, 0.1]])
    s = _one_shell_system(a, [0.6, 0.4], 1, pts)
    rot = pts @ np.array([[0, 1, 0], [-1, 0, 0], [0, 0, 1]], dtype=np.float64)  # (x,y,z)->(-y,x,z)
    s_rot = _one_shell_system(a, [0.6, 0.4], 1, rot)
    v = oracle_lib.ao_vgl(s)[:, 0, :]
    vr = oracle_lib.ao_vgl(s_rot)[:, 0, :]
    # phi_x(Rr) = -y R = -phi_y(r), phi_y(Rr) = phi_x(r), phi_z unchanged
    np.testing.assert_allclose(vr[:, 0], -v[:, 1], rtol=1e-14)
    np.testing.assert_allclose(vr[:, 1], v[:, 0], rtol=1e-14)
    np.testing.assert_allclose(vr[:, 2], v[:, 2], rtol=1e-14)


def test_mo_contraction_exact_rational(oracle_lib):
    rng = np.random.default_rng(3)
    P, A, M = 3, 7, 5
    ao = rng.normal(size=(P, 5, A)) * 10.0 ** rng.integers(-3, 3, size=(P, 5, A))
    C = rng.normal(size=(M, A))
    mo = oracle_lib.mo_vgl_from_ao(C, ao)
    for p in range(P):
        for c in range(5):
            for i in range(M):
                exact = sum(Fraction(float(C[i, k])) * Fraction(float(ao[p, c, k])) for k in range(A))
                mag = sum(abs(float(C[i, k]) * float(ao[p, c, k])) for k in range(A))
                assert abs(mo[p, c, i] - float(exact)) <= 8 * 2.0 ** -53 * mag


def test_mo_magnitude_bounds_value(oracle_lib, chbrclf):
    mo, mag = oracle_lib.mo_vgl(chbrclf, with_mag=True)
    assert np.all(np.abs(mo) <= mag * (1 + 1e-12) + 1e-300)


@pytest.mark.parametrize("system", ["n2_like", "lmax3", "lmax4"])
def test_adjoint_against_finite_differences(oracle_lib, system):
    """Central differences (h = 1e-5) of L = sum w mo_vgl against the
    oracle's adjoint, for every nuclear coordinate, every third primitive
    exponent and (exactly, L being linear in them) the MO coefficients.  The
    l = 3 and l = 4 systems exercise the third derivatives of the monomials
    that the gradient of the Laplacian needs (identically zero for l <= 2)."""
    s = S.n2_like(walk_num=2) if system == "n2_like" else _fd_system(lmax=int(system[-1]))
    rng = np.random.default_rng(11)
    w = rng.normal(size=(s.point_num, 5, s.mo_num))

    def L(sys_):
        return float(np.sum(w * oracle_lib.mo_vgl(sys_)))

    dR, dA, dC = oracle_lib.mo_vgl_adjoint(s, w)
    h = 1e-5
    for n in range(s.nucl_num):
        for d in range(3):
            sp = S.System(**{**s.__dict__})
            sm = S.System(**{**s.__dict__})
            sp.nucl_coord = s.nucl_coord.copy()
            sm.nucl_coord = s.nucl_coord.copy()
            sp.nucl_coord[d, n] += h
            sm.nucl_coord[d, n] -= h
            fd = (L(sp) - L(sm)) / (2 * h)
            assert abs(fd - dR[n, d]) <= 1e-6 * max(1.0, abs(fd)), (n, d, fd, dR[n, d])
    for j in range(0, s.prim_num, 3):
        sp = S.System(**{**s.__dict__})
        sm = S.System(**{**s.__dict__})
        hj = h * s.exponent[j]
        sp.exponent = s.exponent.copy()
        sm.exponent = s.exponent.copy()
        sp.exponent[j] += hj
        sm.exponent[j] -= hj
        fd = (L(sp) - L(sm)) / (2 * hj)
        assert abs(fd - dA[j]) <= 1e-6 * max(1.0, abs(fd)), (j, fd, dA[j])
    ao = oracle_lib.ao_vgl(s)
    dC_ref = np.einsum("pci,pck->ik", w, ao)
    assert np.max(np.abs(dC - dC_ref)) <= 1e-12 * np.max(np.abs(dC_ref))


def test_generator_normalisation_matches_fixture(chbrclf):
    s = chbrclf
    pf = np.empty(s.prim_num)
    for sh in range(s.shell_num):
        l = int(s.shell_ang_mom[sh])
        a, n = s.shell_prim_index[sh], s.shell_prim_num[sh]
        pf[a:a + n] = S.prim_norm(s.exponent[a:a + n], l)
    assert np.max(np.abs(pf / s.prim_factor - 1)) < 1e-14
    k = 0
    for sh in range(s.shell_num):
        l = int(s.shell_ang_mom[sh])
        np.testing.assert_allclose(s.ao_factor[k:k + S.ncart(l)], S.ao_factors(l), rtol=1e-15)
        k += S.ncart(l)
        a, n = s.shell_prim_index[sh], s.shell_prim_num[sh]
        # the fixture's contracted shells are normalised under this convention
        assert abs(S.shell_norm(s.exponent[a:a + n], s.coefficient[a:a + n], l) - 1) < 1e-5


def test_shell_norm_by_quadrature():
    from scipy import integrate
    alpha = np.array([3.0, 0.8, 0.25])
    coef = np.array([0.2, 0.5, 0.6])
    for l in range(4):
        N = S.prim_norm(alpha, l) * coef * S.shell_norm(alpha, coef, l)

        def f1(x, e):
            return sum(N[i] * x ** e * math.exp(-alpha[i] * x * x) for i in range(3))
        # |phi|^2 for the x^l component: separable only per primitive pair
        tot = 0.0
        for i in range(3):
            for j in range(3):
                a = alpha[i] + alpha[j]
                ix = integrate.quad(lambda x: x ** (2 * l) * math.exp(-a * x * x), -np.inf, np.inf)[0]
                iy = integrate.quad(lambda x: math.exp(-a * x * x), -np.inf, np.inf)[0]
                tot += N[i] * N[j] * ix * iy * iy
        assert abs(tot - 1.0) < 1e-9, (l, tot)


def test_synthetic_sizes():
    c1 = S.n2_like()
    assert (c1.nucl_num, c1.ao_num, c1.mo_num, c1.elec_num, c1.point_num) == (2, 30, 30, 14, 14)
    assert c1.shell_num == 12
    c2 = S.chbrclf_sized(walk_num=2)
    assert (c2.nucl_num, c2.ao_num, c2.mo_num, c2.elec_num) == (5, 145, 145, 67)
    assert S.chbrclf_sized().point_num == 17152
    c3 = S.water_cluster(4, walk_num=1)
    assert (c3.nucl_num, c3.ao_num, c3.mo_num, c3.elec_num, c3.shell_num) == (192, 4160, 320, 640, 1408)
    zero = float((c3.mo_coef == 0).mean())
    assert 0.65 < zero < 0.75
    assert c3.prim_num == int(c3.shell_prim_num.sum())
