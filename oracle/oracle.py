"""ctypes front end of the FP64 CPU oracle (``libqmckl_oracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke``
and the CPU-baseline leg of ``bench.py`` as the checker.  The product
package never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqmckl_oracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.POINTER(ctypes.c_int64)
_I32 = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.c_int64


def build(force=False):
    import subprocess
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all", "-C", _HERE], check=True)


def use_library(path):
    """Loads the oracle from another build of the same source (the
    -march=native one the CPU baseline times); later calls use it."""
    global _lib, LIB_PATH
    LIB_PATH = path
    _lib = None
    return lib()


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        basis = [_I64, _I64, _I32, _I64, _I64, _D, _D, _D, _D]
        L.qo_en_distance.argtypes = [_i64, _i64, _D, _D, _D]
        L.qo_ee_distance.argtypes = [_i64, _i64, _D, _D]
        L.qo_shell_vgl.argtypes = [_i64, _i64, _i64, _i64, _D, _D] + basis + [_D]
        L.qo_ao_vgl.argtypes = [_i64, _i64, _i64, _i64, _i64, _D, _D] + basis + [_D, _D, _D]
        L.qo_ao_value.argtypes = [_i64, _i64, _i64, _i64, _i64, _D, _D] + basis + [_D, _D]
        L.qo_mo_vgl.argtypes = [_i64, _i64, _i64, _D, _D, _D, _D, _D]
        L.qo_mo_vgl_adjoint.argtypes = [_i64] * 6 + [_D, _D] + basis + [_D, _D, _D, _D, _D, _D]
        L.qo_finalize_ao.argtypes = [_i64, _i64, _i64, _I64, _I64, _I32, _I64, _I64, _D, _D, _D, _D,
                                     _D, _I32, _D]
        L.qo_set_num_threads.argtypes = [ctypes.c_int]
        L.qo_get_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a, dtype=np.float64):
    if a is None:
        return None
    assert a.dtype == dtype and a.flags.c_contiguous, (a.dtype, dtype)
    if dtype == np.float64:
        return a.ctypes.data_as(_D)
    if dtype == np.int64:
        return a.ctypes.data_as(_I64)
    return a.ctypes.data_as(_I32)


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def _basis(sys_):
    arrs = (_c(sys_.nucleus_index, np.int64), _c(sys_.nucleus_shell_num, np.int64),
            _c(sys_.shell_ang_mom, np.int32), _c(sys_.shell_prim_num, np.int64),
            _c(sys_.shell_prim_index, np.int64), _c(sys_.shell_factor), _c(sys_.exponent),
            _c(sys_.coefficient), _c(sys_.prim_factor))
    ptrs = [_p(arrs[0], np.int64), _p(arrs[1], np.int64), _p(arrs[2], np.int32),
            _p(arrs[3], np.int64), _p(arrs[4], np.int64)] + [_p(a) for a in arrs[5:]]
    return arrs, ptrs


def set_num_threads(n):
    lib().qo_set_num_threads(int(n))


def num_threads():
    return int(lib().qo_get_num_threads())


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")


def en_distance(sys_, coord_T=None):
    coord = _c(sys_.coord_T() if coord_T is None else coord_T)
    P = coord.shape[1]
    out = np.empty((P, sys_.nucl_num))
    nc = _c(sys_.nucl_coord)
    _check(lib().qo_en_distance(P, sys_.nucl_num, _p(coord), _p(nc), _p(out)), "en_distance")
    return out


def ee_distance(sys_, coord_T=None):
    """[walk][elec][elec] electron-electron distances of each walker."""
    coord = _c(sys_.coord_T() if coord_T is None else coord_T)
    E = sys_.elec_num
    W = coord.shape[1] // E
    out = np.empty((W, E, E))
    _check(lib().qo_ee_distance(E, W, _p(coord), _p(out)), "ee_distance")
    return out


def finalize(sys_):
    arrs, ptrs = _basis(sys_)
    cn = np.empty(sys_.prim_num)
    lmax = np.empty(sys_.nucl_num, dtype=np.int32)
    rng = np.empty(sys_.nucl_num)
    _check(lib().qo_finalize_ao(sys_.nucl_num, sys_.shell_num, sys_.prim_num, *ptrs,
                                _p(cn), _p(lmax, np.int32), _p(rng)), "finalize")
    return cn, lmax, rng


def shell_vgl(sys_, coord_T=None):
    coord = _c(sys_.coord_T() if coord_T is None else coord_T)
    P = coord.shape[1]
    arrs, ptrs = _basis(sys_)
    nc = _c(sys_.nucl_coord)
    out = np.empty((P, 5, sys_.shell_num))
    _check(lib().qo_shell_vgl(P, sys_.nucl_num, sys_.shell_num, sys_.prim_num, _p(coord), _p(nc),
                              *ptrs, _p(out)), "shell_vgl")
    return out


def ao_vgl(sys_, coord_T=None, with_mag=False):
    coord = _c(sys_.coord_T() if coord_T is None else coord_T)
    P = coord.shape[1]
    arrs, ptrs = _basis(sys_)
    nc = _c(sys_.nucl_coord)
    af = _c(sys_.ao_factor)
    out = np.empty((P, 5, sys_.ao_num))
    mag = np.empty((P, 5, sys_.ao_num)) if with_mag else None
    _check(lib().qo_ao_vgl(P, sys_.nucl_num, sys_.shell_num, sys_.prim_num, sys_.ao_num,
                           _p(coord), _p(nc), *ptrs, _p(af), _p(out), _p(mag)), "ao_vgl")
    return (out, mag) if with_mag else out


def ao_value(sys_, coord_T=None):
    coord = _c(sys_.coord_T() if coord_T is None else coord_T)
    P = coord.shape[1]
    arrs, ptrs = _basis(sys_)
    nc = _c(sys_.nucl_coord)
    af = _c(sys_.ao_factor)
    out = np.empty((P, sys_.ao_num))
    _check(lib().qo_ao_value(P, sys_.nucl_num, sys_.shell_num, sys_.prim_num, sys_.ao_num,
                             _p(coord), _p(nc), *ptrs, _p(af), _p(out)), "ao_value")
    return out


def mo_vgl_from_ao(mo_coef, ao, ao_mag=None):
    mo_coef = _c(mo_coef)
    ao = _c(ao)
    P, _, A = ao.shape
    M = mo_coef.shape[0]
    out = np.empty((P, 5, M))
    mag = np.empty((P, 5, M)) if ao_mag is not None else None
    _check(lib().qo_mo_vgl(P, A, M, _p(mo_coef), _p(ao), _p(out),
                           _p(_c(ao_mag)) if ao_mag is not None else None, _p(mag)), "mo_vgl")
    return (out, mag) if ao_mag is not None else out


def mo_vgl(sys_, coord_T=None, with_mag=False):
    if with_mag:
        ao, am = ao_vgl(sys_, coord_T, with_mag=True)
        return mo_vgl_from_ao(sys_.mo_coef, ao, am)
    return mo_vgl_from_ao(sys_.mo_coef, ao_vgl(sys_, coord_T))


def mo_vgl_adjoint(sys_, w, coord_T=None, want=("nucl", "expo", "coef")):
    coord = _c(sys_.coord_T() if coord_T is None else coord_T)
    P = coord.shape[1]
    w = _c(w)
    assert w.shape == (P, 5, sys_.mo_num)
    arrs, ptrs = _basis(sys_)
    nc = _c(sys_.nucl_coord)
    af = _c(sys_.ao_factor)
    mc = _c(sys_.mo_coef)
    dR = np.empty((sys_.nucl_num, 3)) if "nucl" in want else None
    dA = np.empty(sys_.prim_num) if "expo" in want else None
    dC = np.empty((sys_.mo_num, sys_.ao_num)) if "coef" in want else None
    _check(lib().qo_mo_vgl_adjoint(P, sys_.nucl_num, sys_.shell_num, sys_.prim_num, sys_.ao_num,
                                   sys_.mo_num, _p(coord), _p(nc), *ptrs, _p(af), _p(mc), _p(w),
                                   _p(dR), _p(dA), _p(dC)), "adjoint")
    return dR, dA, dC
