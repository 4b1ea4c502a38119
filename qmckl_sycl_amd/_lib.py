"""ctypes binding of ``libqmckl_gpu_amd.so``.

The prototypes are read from ``include/qmckl_gpu.h`` itself, so the binding
and the C header cannot drift apart.  Pointer arguments are passed as
integers (device addresses) or ctypes pointers (host outputs).

The HIP runtime is process-wide: when PyTorch is importable it is imported
first, so that the library binds to the same ``libamdhip64`` PyTorch uses
(both carry the soname ``libamdhip64.so.7``; loading the library first would
map a second copy when torch is imported later).
"""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
HEADER = os.path.join(REPO, "include", "qmckl_gpu.h")
LIB_PATH = os.path.join(HERE, "libqmckl_gpu_amd.so")


class MemoryInfo(ctypes.Structure):
    _fields_ = [("size", ctypes.c_size_t), ("pointer", ctypes.c_void_p)]


_SCALARS = {
    "qmckl_exit_code_device": ctypes.c_int32,
    "qmckl_context_device": ctypes.c_int64,
    "qmckl_precision_device": ctypes.c_int32,
    "bool": ctypes.c_bool,
    "int64_t": ctypes.c_int64,
    "int32_t": ctypes.c_int32,
    "int": ctypes.c_int,
    "char": ctypes.c_char,
    "size_t": ctypes.c_size_t,
    "double": ctypes.c_double,
    "float": ctypes.c_float,
    "void": None,
    "qmckl_memory_info_struct_device": MemoryInfo,
}


def _ctype(decl):
    decl = re.sub(r"\bconst\b", "", decl).strip()
    ptr = decl.count("*")
    base = decl.replace("*", " ").split()
    # drop the parameter name if present
    if len(base) >= 2 and base[-1] not in _SCALARS:
        base = base[:-1]
    t = " ".join(base)
    if ptr:
        if t == "char" and ptr == 1:
            return ctypes.c_char_p
        return ctypes.c_void_p
    if t not in _SCALARS:
        raise ValueError(f"unknown C type in {decl!r}")
    return _SCALARS[t]


def parse_header(path=HEADER):
    """Returns {name: (restype, [argtypes])} for every prototype in the header."""
    with open(path) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    text = re.sub(r"#[^\n]*", " ", text)
    text = re.sub(r"typedef\s+struct[^;]*?\}[^;]*;", " ", text, flags=re.S)
    text = re.sub(r"typedef[^;]*;", " ", text)
    text = text.replace('extern "C" {', " ").replace("}", " ")
    protos = {}
    for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b(qmckl_\w+)\s*\(([^;]*?)\)\s*;", text, flags=re.S):
        ret, name, args = m.group(1), m.group(2), m.group(3)
        argl = [a.strip() for a in args.split(",") if a.strip() and a.strip() != "void"]
        protos[name] = (_ctype(ret), [_ctype(a) for a in argl])
    return protos


_lib = None


def _prefer_torch_runtime():
    try:
        import torch  # noqa: F401
    except Exception:
        pass


def load(path=LIB_PATH):
    """Loads the HIP library; raises if it has not been built (there is no
    CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -m qmckl_sycl_amd.build` "
            "(the qmckl path has no CPU fallback)")
    _prefer_torch_runtime()
    L = ctypes.CDLL(path)
    for name, (res, args) in parse_header().items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L
