"""Differentiation form on the GPU: the C-ABI adjoint against the FP64
oracle gradient (itself pinned by central differences in test_oracle.py)
and against central differences of the GPU forward, plus the
torch.autograd.Function harness.

Tolerances: oracle gradient 1e-10 relative to the largest component of each
gradient; finite differences (h = 1e-5, relative for exponents) 1e-6."""
import dataclasses

import numpy as np
import pytest
import torch

from qmckl_sycl_amd import systems as S
from qmckl_sycl_amd.context import Context, DeviceArray

from helpers import say

pytestmark = pytest.mark.gpu


def gpu_adjoint(s, w, point_order=-1):
    with Context(0) as c:
        c.set_system(s)
        c.call("qmckl_set_mo_basis_tile_config_device", point_order, 0)
        wd = DeviceArray.from_numpy(c, w)
        dR = DeviceArray(c, (s.nucl_num, 3))
        dA = DeviceArray(c, (s.prim_num,))
        dC = DeviceArray(c, (s.mo_num, s.ao_num))
        c.call("qmckl_get_mo_basis_mo_vgl_adjoint_device", wd.ptr, wd.size, dR.ptr, dA.ptr, dC.ptr)
        out = dR.numpy(), dA.numpy(), dC.numpy()
        # second call: bitwise reproducible
        c.call("qmckl_get_mo_basis_mo_vgl_adjoint_device", wd.ptr, wd.size, dR.ptr, dA.ptr, dC.ptr)
        again = dR.numpy(), dA.numpy(), dC.numpy()
    for a, b in zip(out, again):
        assert np.array_equal(a, b)
    return out


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("name,point_order", [("n2_like", -1), ("chbrclf", -1), ("water64_2w", -1),
                                              ("n2_like", 1), ("chbrclf", 1), ("water64_2w", 1)])
def test_adjoint_vs_oracle(oracle_lib, name, point_order):
    """point_order 1 visits the points in spatial tiles on these small
    systems too (136 to 1,280 points, the last tile ragged)."""
    if name == "n2_like":
        s = S.n2_like(walk_num=3)
    elif name == "chbrclf":
        s = S.chbrclf()
    else:
        s = S.water64(walk_num=2)
    w = np.random.default_rng(5).normal(size=(s.point_num, 5, s.mo_num))
    gR, gA, gC = gpu_adjoint(s, w, point_order)
    oR, oA, oC = oracle_lib.mo_vgl_adjoint(s, w)
    e = (rel(gR, oR), rel(gA, oA), rel(gC, oC))
    say(f"{name} point_order {point_order}: rel err dR {e[0]:.2e} dalpha {e[1]:.2e} dC {e[2]:.2e}")
    assert max(e) <= 1e-10, e


def test_adjoint_after_points_move(oracle_lib):
    """Two adjoint calls in one context with the points in spatial tiles,
    the electrons moved between them: the blocks of G the screen zeroes are
    not written, so the second call's G keeps the first call's values there,
    and the second gradient must still match the oracle on the new points."""
    s = S.water64(walk_num=2)
    rng = np.random.default_rng(11)
    w = rng.normal(size=(s.point_num, 5, s.mo_num))
    moved = s.with_walkers(s.elec_coord + rng.normal(scale=0.7, size=s.elec_coord.shape))
    with Context(0) as c:
        c.set_system(s)
        c.call("qmckl_set_mo_basis_tile_config_device", 1, 0)
        wd = DeviceArray.from_numpy(c, w)
        dR = DeviceArray(c, (s.nucl_num, 3))
        dA = DeviceArray(c, (s.prim_num,))
        dC = DeviceArray(c, (s.mo_num, s.ao_num))
        c.call("qmckl_get_mo_basis_mo_vgl_adjoint_device", wd.ptr, wd.size, dR.ptr, dA.ptr, dC.ptr)
        c.set_electron_coord(moved.elec_coord)
        c.call("qmckl_get_mo_basis_mo_vgl_adjoint_device", wd.ptr, wd.size, dR.ptr, dA.ptr, dC.ptr)
        gR, gA, gC = dR.numpy(), dA.numpy(), dC.numpy()
    oR, oA, oC = oracle_lib.mo_vgl_adjoint(moved, w)
    e = (rel(gR, oR), rel(gA, oA), rel(gC, oC))
    say(f"moved points: rel err dR {e[0]:.2e} dalpha {e[1]:.2e} dC {e[2]:.2e}")
    assert max(e) <= 1e-10, e


def test_adjoint_water64_32_walkers(oracle_lib):
    """The differentiation workload: 64 waters, 32 walkers (20,480 points),
    w ~ N(0, 1)."""
    s, w = S.water64_diff(walk_num=32)
    gR, gA, gC = gpu_adjoint(s, w)
    oR, oA, oC = oracle_lib.mo_vgl_adjoint(s, w)
    e = (rel(gR, oR), rel(gA, oA), rel(gC, oC))
    say(f"water64_diff: rel err dR {e[0]:.2e} dalpha {e[1]:.2e} dC {e[2]:.2e}")
    assert max(e) <= 1e-10, e


def test_adjoint_vs_finite_differences():
    """Central differences of the GPU forward (fresh contexts per
    perturbation) against the adjoint, n2_like at 2 walkers."""
    s = S.n2_like(walk_num=2)
    w = np.random.default_rng(9).normal(size=(s.point_num, 5, s.mo_num))
    gR, gA, gC = gpu_adjoint(s, w)

    def L(sys_):
        with Context(0) as c:
            c.set_system(sys_)
            return float(np.sum(w * c.mo_vgl(sys_.point_num)))

    h = 1e-5
    worst = 0.0
    for n in range(s.nucl_num):
        for d in range(3):
            nc_p = s.nucl_coord.copy()
            nc_m = s.nucl_coord.copy()
            nc_p[d, n] += h
            nc_m[d, n] -= h
            fd = (L(dataclasses.replace(s, nucl_coord=nc_p)) - L(dataclasses.replace(s, nucl_coord=nc_m))) / (2 * h)
            worst = max(worst, abs(fd - gR[n, d]) / max(1.0, abs(fd)))
    for j in range(0, s.prim_num, 4):
        hj = h * s.exponent[j]
        ep = s.exponent.copy()
        em = s.exponent.copy()
        ep[j] += hj
        em[j] -= hj
        fd = (L(dataclasses.replace(s, exponent=ep)) - L(dataclasses.replace(s, exponent=em))) / (2 * hj)
        worst = max(worst, abs(fd - gA[j]) / max(1.0, abs(fd)))
    for (i, k) in [(0, 0), (3, 7), (29, 29), (11, 17)]:
        cp = s.mo_coef.copy()
        cm = s.mo_coef.copy()
        cp[i, k] += h
        cm[i, k] -= h
        fd = (L(dataclasses.replace(s, mo_coef=cp)) - L(dataclasses.replace(s, mo_coef=cm))) / (2 * h)
        worst = max(worst, abs(fd - gC[i, k]) / max(1.0, abs(fd)))
    say(f"finite differences: worst relative deviation {worst:.2e}")
    assert worst <= 1e-6


def test_autograd_harness():
    from qmckl_sycl_amd.autograd import mo_vgl
    s = S.n2_like(walk_num=2)
    dev = torch.device("cuda", 0)
    R = torch.tensor(s.nucl_coord.T.copy(), dtype=torch.float64, device=dev, requires_grad=True)
    a = torch.tensor(s.exponent, dtype=torch.float64, device=dev, requires_grad=True)
    C = torch.tensor(s.mo_coef, dtype=torch.float64, device=dev, requires_grad=True)
    wt = torch.tensor(np.random.default_rng(2).normal(size=(s.point_num, 5, s.mo_num)), device=dev)
    mo = mo_vgl(R, a, C, s)
    L = (wt * mo).sum()
    gR, ga, gC = torch.autograd.grad(L, (R, a, C))
    oR, oA, oC = gpu_adjoint(s, wt.cpu().numpy())
    assert np.array_equal(gR.cpu().numpy(), oR)
    assert np.array_equal(ga.cpu().numpy(), oA)
    assert np.array_equal(gC.cpu().numpy(), oC)
    # linear in C: a directional derivative along C is exact
    dirC = torch.randn_like(C)
    with torch.no_grad():
        mo2 = mo_vgl(R, a, C + 1e-3 * dirC, s)
    lin = float(((wt * (mo2 - mo)).sum() / 1e-3))
    assert abs(lin - float((gC * dirC).sum())) <= 1e-8 * max(1.0, abs(lin))



def _gpu_forward(sys_, ao=False):
    with Context(0) as c:
        c.set_system(sys_)
        mo = c.mo_vgl(sys_.point_num)
        return (c.ao_vgl(sys_.point_num), mo) if ao else mo


def _fd_of_L(w, sp, sm, step):
    """(L(sp) - L(sm)) / (2 step) with L = sum w mo_vgl, the two forward
    passes subtracted element by element first, so that the round-off of
    summing 32.8 M terms of L does not enter the difference."""
    return float(np.sum(w * (_gpu_forward(sp) - _gpu_forward(sm)))) / (2.0 * step)


def test_adjoint_vs_finite_differences_c5():
    """The differentiation workload itself (64 waters, 32 walkers, 20,480
    points, w ~ N(0, 1)): central differences (h = 1e-5) of the GPU's FP64
    forward pass against the GPU adjoint, each within 1e-6 relative:

    - 3 random directions of all 576 nuclear coordinates and 3 MO
      coefficients, from differences of mo_vgl;
    - the most diffuse primitive exponent of 4 f shells and 2 d shells
      (relative step h alpha), from differences of ao_vgl contracted with
      (w C)[p][c][k] = sum_i w[p][c][i] C[i][k], since L = sum (w C) ao_vgl.
      One exponent moves only its shell's 7-10 AOs, a gradient of order 1
      against sum |w mo_vgl| of order 1e5, and differences of mo_vgl pick
      up the FP64 contraction's round-off (the sums after the changed AOs
      round differently): measured 6.5e-7 and 5.9e-6 relative for exponents
      61 and 2437, against 2.5e-9 and 1.6e-8 from the AO differences.  The
      mo_vgl-level value is printed for the record.  A primitive with a
      point on its screen alpha r^2 = -ln(1e-12) within +-h is passed over
      (exponent 234 here: one point 5.3e-6 from the screen, whose step adds
      3e-6 to the difference quotient)."""
    s, w = S.water64_diff(walk_num=32)
    gR, gA, gC = gpu_adjoint(s, w)
    h = 1e-5
    rng = np.random.default_rng(21)
    worst = 0.0
    for k in range(3):
        u = rng.normal(size=s.nucl_coord.shape)
        u /= np.linalg.norm(u)
        sp = dataclasses.replace(s, nucl_coord=np.ascontiguousarray(s.nucl_coord + h * u))
        sm = dataclasses.replace(s, nucl_coord=np.ascontiguousarray(s.nucl_coord - h * u))
        fd = _fd_of_L(w, sp, sm, h)
        ad = float(np.sum(gR.T * u))
        e = abs(fd - ad) / abs(ad)
        say(f"nuclear direction {k}: fd {fd:.12g} adjoint {ad:.12g} rel {e:.2e}")
        assert e <= 1e-6, (k, fd, ad)
        worst = max(worst, e)
    for (i, k) in ((0, 0), (17, 1000), (319, 4159)):
        cp, cm = s.mo_coef.copy(), s.mo_coef.copy()
        cp[i, k] += h
        cm[i, k] -= h
        fd = _fd_of_L(w, dataclasses.replace(s, mo_coef=cp), dataclasses.replace(s, mo_coef=cm), h)
        e = abs(fd - gC[i, k]) / abs(gC[i, k])
        say(f"coefficient ({i}, {k}): fd {fd:.12g} adjoint {gC[i, k]:.12g} rel {e:.2e}")
        assert e <= 1e-6, (i, k, fd, gC[i, k])
        worst = max(worst, e)
    # The primitive screen alpha r^2 > -ln(1e-12) makes L a step function of
    # alpha where a point sits on it; a primitive is taken only if no point
    # crosses the screen between alpha (1 - h) and alpha (1 + h).
    pts = np.asarray(s.elec_coord).reshape(-1, 3)
    nucleus_of_shell = np.repeat(np.arange(s.nucl_num), s.nucleus_shell_num)
    cut = -np.log(1e-12)
    prims = []
    for l, n in ((3, 4), (2, 2)):
        shells = [sh for sh in range(s.shell_num) if s.shell_ang_mom[sh] == l]
        taken = 0
        for sh in shells[::7]:
            j = int(s.shell_prim_index[sh] + s.shell_prim_num[sh] - 1)
            r2 = np.sum((pts - s.nucl_coord[:, nucleus_of_shell[sh]]) ** 2, axis=1)
            x = s.exponent[j] * r2
            crossing = int(np.sum((x * (1 - h) <= cut) & (x * (1 + h) > cut)))
            if crossing:
                say(f"exponent {j} skipped: {crossing} point(s) cross the primitive screen within +-h")
                continue
            prims.append(j)
            taken += 1
            if taken == n:
                break
    wC = np.matmul(w, s.mo_coef)  # [P][5][A]
    for j in prims:
        hj = h * s.exponent[j]
        ep, em = s.exponent.copy(), s.exponent.copy()
        ep[j] += hj
        em[j] -= hj
        aop, mop = _gpu_forward(dataclasses.replace(s, exponent=ep), ao=True)
        aom, mom = _gpu_forward(dataclasses.replace(s, exponent=em), ao=True)
        dao = aop - aom
        moved = np.nonzero(np.any(dao != 0, axis=(0, 1)))[0]
        # only the AOs of the shell that owns primitive j may move
        assert 0 < len(moved) <= 10, moved
        fd = float(np.sum(dao[:, :, moved] * wC[:, :, moved])) / (2.0 * hj)
        fd_mo = float(np.sum(w * (mop - mom))) / (2.0 * hj)
        e = abs(fd - gA[j]) / abs(gA[j])
        say(f"exponent {j} (alpha {s.exponent[j]:.4g}): fd {fd:.12g} adjoint {gA[j]:.12g} rel {e:.2e}"
            f" (mo_vgl differences {abs(fd_mo - gA[j]) / abs(gA[j]):.2e})")
        assert e <= 1e-6, (j, fd, gA[j])
        worst = max(worst, e)
    say(f"water64_diff finite differences: worst relative difference {worst:.2e}")
