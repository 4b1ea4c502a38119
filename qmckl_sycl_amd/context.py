"""Host-side mirror of the qmckl GPU context/provider API.

``Context`` wraps one ``qmckl_context_device`` and exposes the reference's
entry points under the same names (``include/qmckl_gpu.hpp:473-801``) with
the same argument meaning; a non-zero exit code raises :class:`QmcklError`
carrying the code, the function and the message recorded in the context.
``DeviceArray`` is a device buffer from ``qmckl_malloc_device`` with numpy
upload/download, the way the reference tests move data
(``tests/test_qmckl_mo.cpp:64-76``).
"""
import ctypes

import numpy as np

from . import _lib

SUCCESS = 0
INVALID_ARG = {i: i for i in range(1, 21)}
FAILURE = 101
ERRNO = 102
INVALID_CONTEXT = 103
ALLOCATION_FAILED = 104
DEALLOCATION_FAILED = 105
NOT_PROVIDED = 106
OUT_OF_BOUNDS = 107
ALREADY_SET = 108
INVALID_EXIT_CODE = 109
NO_DEVICE = -1


class QmcklError(RuntimeError):
    def __init__(self, code, function="", message=""):
        self.code = int(code)
        self.function = function
        self.message = message
        super().__init__(f"qmckl error {self.code} ({string_of_error(self.code)}) in "
                         f"{function or '?'}: {message}")


def lib():
    return _lib.load()


def string_of_error(code):
    return lib().qmckl_string_of_error_device(int(code)).decode()


class DeviceArray:
    """Device buffer owned by a context (tracked by its memory table)."""

    def __init__(self, ctx, shape, dtype=np.float64):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self.dtype.itemsize
        self.ptr = lib().qmckl_malloc_device(ctx.handle, max(self.nbytes, 1))
        if not self.ptr:
            ctx.raise_last(ALLOCATION_FAILED, "qmckl_malloc_device")

    @classmethod
    def from_numpy(cls, ctx, arr, dtype=None):
        a = np.ascontiguousarray(arr, dtype=dtype or np.asarray(arr).dtype)
        d = cls(ctx, a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, arr):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        assert a.nbytes <= self.nbytes
        if a.nbytes:
            self.ctx.check(lib().qmckl_memcpy_H2D(self.ctx.handle, self.ptr,
                                                  a.ctypes.data_as(ctypes.c_void_p), a.nbytes),
                           "qmckl_memcpy_H2D")

    def numpy(self, shape=None):
        out = np.empty(self.shape if shape is None else shape, dtype=self.dtype)
        if out.nbytes:
            self.ctx.check(lib().qmckl_memcpy_D2H(self.ctx.handle,
                                                  out.ctypes.data_as(ctypes.c_void_p), self.ptr,
                                                  out.nbytes), "qmckl_memcpy_D2H")
        return out

    def free(self):
        if self.ptr and self.ctx.handle:
            self.ctx.check(lib().qmckl_free_device(self.ctx.handle, self.ptr), "qmckl_free_device")
        self.ptr = None


class Context:
    """One qmckl context on HIP device ``device_id`` (or ``NO_DEVICE``)."""

    def __init__(self, device_id=0):
        self.handle = lib().qmckl_context_create_device(int(device_id))
        if not self.handle:
            raise QmcklError(INVALID_CONTEXT, "qmckl_context_create_device",
                             f"could not create a context on device {device_id}")

    # -- lifetime
    def destroy(self):
        if self.handle:
            lib().qmckl_context_destroy_device(self.handle)
            self.handle = 0

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    # -- errors
    def last_error(self):
        code = ctypes.c_int32(0)
        fn = ctypes.create_string_buffer(256)
        msg = ctypes.create_string_buffer(1024)
        lib().qmckl_get_error_device(self.handle, ctypes.byref(code), fn, msg)
        return int(code.value), fn.value.decode(), msg.value.decode()

    def raise_last(self, code, function):
        c, fn, msg = self.last_error()
        raise QmcklError(code, fn or function, msg)

    def check(self, rc, function=""):
        rc = int(rc)
        if rc != SUCCESS:
            c, fn, msg = self.last_error()
            raise QmcklError(rc, fn if c == rc else function, msg if c == rc else "")
        return rc

    def call(self, name, *args):
        return self.check(getattr(lib(), name)(self.handle, *args), name)

    def rc(self, name, *args):
        """Raw exit code (no exception)."""
        return int(getattr(lib(), name)(self.handle, *args))

    # -- context
    def touch(self):
        self.call("qmckl_context_touch_device")

    def malloc(self, shape, dtype=np.float64):
        return DeviceArray(self, shape, dtype)

    def to_device(self, arr, dtype=None):
        return DeviceArray.from_numpy(self, arr, dtype)

    # -- scalar queries
    def _get_i64(self, name):
        v = ctypes.c_int64(0)
        self.call(name, ctypes.byref(v))
        return int(v.value)

    def nucleus_num(self):
        return self._get_i64("qmckl_get_nucleus_num_device")

    def electron_num(self):
        return self._get_i64("qmckl_get_electron_num_device")

    def ao_num(self):
        return self._get_i64("qmckl_get_ao_basis_ao_num_device")

    def shell_num(self):
        return self._get_i64("qmckl_get_ao_basis_shell_num_device")

    def prim_num(self):
        return self._get_i64("qmckl_get_ao_basis_prim_num_device")

    def mo_num(self):
        return self._get_i64("qmckl_get_mo_basis_mo_num_device")

    def provided(self, what):
        return bool(getattr(lib(), f"qmckl_{what}_provided_device")(self.handle))

    # -- whole-system setup, in the order of tests/test_qmckl_mo.cpp:80-277
    def set_system(self, sys_, walkers=True):
        """Sets electrons, nucleus, AO basis and MO basis of a
        :class:`qmckl_sycl_amd.systems.System` through the device setters."""
        tmp = []

        def dev(a, dt):
            d = DeviceArray.from_numpy(self, a, dt)
            tmp.append(d)
            return d.ptr

        self.call("qmckl_set_electron_num_device", sys_.elec_up_num, sys_.elec_dn_num)
        if walkers:
            self.set_electron_coord(sys_.elec_coord)
        N = sys_.nucl_num
        self.call("qmckl_set_nucleus_num_device", N)
        self.call("qmckl_set_nucleus_coord_device", b"T", dev(sys_.nucl_coord, np.float64), 3 * N)
        self.call("qmckl_set_nucleus_charge_device", dev(sys_.charge, np.float64), N)
        S, Np, A = sys_.shell_num, sys_.prim_num, sys_.ao_num
        self.call("qmckl_set_ao_basis_type_device", b"G")
        self.call("qmckl_set_ao_basis_shell_num_device", S)
        self.call("qmckl_set_ao_basis_prim_num_device", Np)
        self.call("qmckl_set_ao_basis_nucleus_index_device", dev(sys_.nucleus_index, np.int64), N)
        self.call("qmckl_set_ao_basis_nucleus_shell_num_device", dev(sys_.nucleus_shell_num, np.int64), N)
        self.call("qmckl_set_ao_basis_shell_ang_mom_device", dev(sys_.shell_ang_mom, np.int32), S)
        self.call("qmckl_set_ao_basis_shell_factor_device", dev(sys_.shell_factor, np.float64), S)
        self.call("qmckl_set_ao_basis_shell_prim_num_device", dev(sys_.shell_prim_num, np.int64), S)
        self.call("qmckl_set_ao_basis_shell_prim_index_device", dev(sys_.shell_prim_index, np.int64), S)
        self.call("qmckl_set_ao_basis_exponent_device", dev(sys_.exponent, np.float64), Np)
        self.call("qmckl_set_ao_basis_coefficient_device", dev(sys_.coefficient, np.float64), Np)
        self.call("qmckl_set_ao_basis_prim_factor_device", dev(sys_.prim_factor, np.float64), Np)
        self.call("qmckl_set_ao_basis_ao_num_device", A)
        self.call("qmckl_set_ao_basis_ao_factor_device", dev(sys_.ao_factor, np.float64), A)
        self.call("qmckl_set_mo_basis_mo_num_device", sys_.mo_num)
        self.call("qmckl_set_mo_basis_coefficient_device", dev(sys_.mo_coef, np.float64))
        for d in tmp:   # setters deep-copy: inputs may go at once
            d.free()
        return self

    def set_electron_coord(self, elec_coord, transp=b"N"):
        a = np.ascontiguousarray(elec_coord, dtype=np.float64)
        if transp == b"N":
            walk = a.shape[0]
        else:
            walk = a.shape[1] // self.electron_num()
        d = DeviceArray.from_numpy(self, a)
        self.call("qmckl_set_electron_coord_device", transp, walk, d.ptr, a.size)
        d.free()

    # -- computed arrays as numpy (device getters + D2H)
    def _get_array(self, getter, shape):
        d = DeviceArray(self, shape)
        try:
            self.call(getter, d.ptr, d.size)
            return d.numpy()
        finally:
            d.free()

    def en_distance(self, P):
        return self._get_array("qmckl_get_electron_en_distance_device", (P, self.nucleus_num()))

    def shell_vgl(self, P):
        return self._get_array("qmckl_get_ao_basis_shell_vgl_device", (P, 5, self.shell_num()))

    def ao_vgl(self, P):
        return self._get_array("qmckl_get_ao_basis_ao_vgl_device", (P, 5, self.ao_num()))

    def ao_value(self, P):
        return self._get_array("qmckl_get_ao_basis_ao_value_device", (P, self.ao_num()))

    def mo_vgl(self, P):
        return self._get_array("qmckl_get_mo_basis_mo_vgl_device", (P, 5, self.mo_num()))

    def mo_value(self, P):
        return self._get_array("qmckl_get_mo_basis_mo_value_device", (P, self.mo_num()))


def batch_mo_vgl(contexts, outputs):
    """qmckl_get_mo_basis_mo_vgl_batch_device over Context objects and their
    DeviceArray outputs ([P][5][M] each); raises on the returned code."""
    n = len(contexts)
    if n != len(outputs):
        raise ValueError("one output per context")
    handles = (ctypes.c_int64 * n)(*[c.handle for c in contexts])
    ptrs = (ctypes.c_void_p * n)(*[o.ptr for o in outputs])
    sizes = (ctypes.c_int64 * n)(*[o.size for o in outputs])
    rc = lib().qmckl_get_mo_basis_mo_vgl_batch_device(n, handles, ptrs, sizes)
    if rc != 0:
        raise QmcklError(rc, "qmckl_get_mo_basis_mo_vgl_batch_device")
    return rc


class BatchGraph:
    """qmckl_mo_basis_mo_vgl_batch_graph_{create,launch,destroy}_device over
    Context objects and their DeviceArray outputs: the batch recorded once
    into a HIP graph, replayed by launch()."""

    def __init__(self, contexts, outputs):
        n = len(contexts)
        if n != len(outputs):
            raise ValueError("one output per context")
        handles = (ctypes.c_int64 * n)(*[c.handle for c in contexts])
        ptrs = (ctypes.c_void_p * n)(*[o.ptr for o in outputs])
        sizes = (ctypes.c_int64 * n)(*[o.size for o in outputs])
        g = ctypes.c_int64(0)
        rc = lib().qmckl_mo_basis_mo_vgl_batch_graph_create_device(n, handles, ptrs, sizes, ctypes.byref(g))
        if rc != 0:
            raise QmcklError(rc, "qmckl_mo_basis_mo_vgl_batch_graph_create_device")
        self.handle = g.value

    def launch(self):
        """Replays the graph; returns the exit code (0 on success)."""
        return lib().qmckl_mo_basis_mo_vgl_batch_graph_launch_device(self.handle)

    def destroy(self):
        if self.handle:
            lib().qmckl_mo_basis_mo_vgl_batch_graph_destroy_device(self.handle)
            self.handle = 0

