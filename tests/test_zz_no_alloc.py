"""The allocation-free radial launcher against the allocating one, through
the C ABI: same results on two calls in a row, the second with fewer
points and fewer shells, and its argument / table checks."""
import numpy as np
import pytest

from qmckl_sycl_amd import systems as S
from qmckl_sycl_amd.context import Context, DeviceArray, FAILURE, lib
from helpers import say

pytestmark = pytest.mark.gpu


def _radial_args(c, s, oracle_lib, npts, nucl_keep=None):
    """Device arrays of the low-level radial launcher for system s, the
    first npts points and (optionally) only the shells of the first
    nucl_keep nuclei (a smaller basis: fewer shells, same primitives)."""
    cn, _, rng = oracle_lib.finalize(s)
    N = s.nucl_num if nucl_keep is None else nucl_keep
    S_ = int(s.nucleus_index[N - 1] + s.nucleus_shell_num[N - 1])
    arrs = dict(coord=np.ascontiguousarray(s.coord_T()[:, :npts]),
                nucl=np.ascontiguousarray(s.nucl_coord[:, :N]),
                nsn=s.nucleus_shell_num[:N], nidx=s.nucleus_index[:N], rng=rng[:N],
                spi=s.shell_prim_index[:S_], spn=s.shell_prim_num[:S_], ex=s.exponent, cn=cn)
    dev = {k: DeviceArray.from_numpy(c, np.ascontiguousarray(v)) for k, v in arrs.items()}
    return dev, N, S_, s.prim_num


def _call(c, name, dev, P, N, S_, Np, out):
    L = lib()
    return getattr(L, name)(c.handle, Np, S_, P, N, dev["nsn"].ptr, dev["nidx"].ptr, dev["rng"].ptr,
                            dev["spi"].ptr, dev["spn"].ptr, dev["coord"].ptr, dev["nucl"].ptr,
                            dev["ex"].ptr, dev["cn"].ptr, out.ptr)


def test_noalloc_matches_allocating_launcher(oracle_lib):
    s = S.chbrclf()
    with Context(0) as c:
        for P, nk in ((136, None), (40, 3)):  # second call: fewer points and shells
            dev, N, S_, Np = _radial_args(c, s, oracle_lib, P, nk)
            say(f"noalloc radial launcher: P={P} N={N} S={S_}")
            a = DeviceArray(c, (P, 5, S_))
            b = DeviceArray(c, (P, 5, S_))
            c.check(_call(c, "qmckl_compute_ao_basis_shell_gaussian_vgl_device", dev, P, N, S_, Np, a))
            c.check(_call(c, "qmckl_compute_ao_basis_shell_gaussian_vgl_noalloc_device", dev, P, N, S_, Np, b))
            ra, rb = a.numpy(), b.numpy()
            assert np.array_equal(ra, rb)
            assert np.count_nonzero(ra) > 0
            for d in list(dev.values()) + [a, b]:
                d.free()


def test_noalloc_rejects_bad_tables_and_sizes(oracle_lib):
    s = S.chbrclf()
    with Context(0) as c:
        P = 8
        dev, N, S_, Np = _radial_args(c, s, oracle_lib, P)
        out = DeviceArray(c, (P, 5, S_))
        L = lib()
        # beyond the scratch table's 65536 shells: ARG_3, nothing launched
        rc = _call(c, "qmckl_compute_ao_basis_shell_gaussian_vgl_noalloc_device", dev, P, N, 65537, Np, out)
        assert rc == 3
        # a nucleus whose shells run past shell_num: FAILURE before the radial kernel
        bad = s.nucleus_index.copy()
        bad[-1] = S_
        dev["nidx"].free()
        dev["nidx"] = DeviceArray.from_numpy(c, bad)
        rc = _call(c, "qmckl_compute_ao_basis_shell_gaussian_vgl_noalloc_device", dev, P, N, S_, Np, out)
        assert rc == FAILURE
        # primitives past prim_num
        dev["nidx"].free()
        dev["nidx"] = DeviceArray.from_numpy(c, s.nucleus_index)
        badp = s.shell_prim_index.copy()
        badp[5] = Np
        dev["spi"].free()
        dev["spi"] = DeviceArray.from_numpy(c, badp)
        rc = _call(c, "qmckl_compute_ao_basis_shell_gaussian_vgl_noalloc_device", dev, P, N, S_, Np, out)
        assert rc == FAILURE
        # and a good call afterwards still matches the allocating launcher
        dev["spi"].free()
        dev["spi"] = DeviceArray.from_numpy(c, s.shell_prim_index)
        ref = DeviceArray(c, (P, 5, S_))
        c.check(_call(c, "qmckl_compute_ao_basis_shell_gaussian_vgl_device", dev, P, N, S_, Np, ref))
        c.check(_call(c, "qmckl_compute_ao_basis_shell_gaussian_vgl_noalloc_device", dev, P, N, S_, Np, out))
        assert np.array_equal(ref.numpy(), out.numpy())
        del L
        for d in list(dev.values()) + [out, ref]:
            d.free()
