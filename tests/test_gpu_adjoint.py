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


def gpu_adjoint(s, w):
    with Context(0) as c:
        c.set_system(s)
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


@pytest.mark.parametrize("name", ["n2_like", "chbrclf", "water64_2w"])
def test_adjoint_vs_oracle(oracle_lib, name):
    if name == "n2_like":
        s = S.n2_like(walk_num=3)
    elif name == "chbrclf":
        s = S.chbrclf()
    else:
        s = S.water64(walk_num=2)
    w = np.random.default_rng(5).normal(size=(s.point_num, 5, s.mo_num))
    gR, gA, gC = gpu_adjoint(s, w)
    oR, oA, oC = oracle_lib.mo_vgl_adjoint(s, w)
    e = (rel(gR, oR), rel(gA, oA), rel(gC, oC))
    say(f"{name}: rel err dR {e[0]:.2e} dalpha {e[1]:.2e} dC {e[2]:.2e}")
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
