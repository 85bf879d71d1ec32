"""MI355X-native per-walker wavefunction evaluation with the qmckl_context C API.

The product is the C-ABI shared library ``libqmckl_amd.so`` (declared in
``include/qmckl.h``); this package loads it with ctypes and adds a thin Python layer in
the style of the reference's SWIG module (``python/src/qmckl.i``): the ``qmckl_`` prefix
is dropped, array outputs come back as numpy arrays and a non-zero exit code raises
``QmcklError`` (the reference raises ``RuntimeError``; ``QmcklError`` derives from it).

There is no fallback: if the library is missing, importing a function raises.
"""

import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqmckl_amd.so")
HEADER = os.path.join(HERE, "..", "include", "qmckl.h")

_CTYPES = {
    "qmckl_context": ctypes.c_int64,
    "qmckl_exit_code": ctypes.c_int32,
    "int64_t": ctypes.c_int64,
    "uint64_t": ctypes.c_uint64,
    "size_t": ctypes.c_size_t,
    "int32_t": ctypes.c_int32,
    "int": ctypes.c_int,
    "double": ctypes.c_double,
    "float": ctypes.c_float,
    "char": ctypes.c_char,
    "bool": ctypes.c_bool,
    "void": None,
}


class QmcklError(RuntimeError):
    """Raised when a qmckl entry point returns a non-zero exit code."""

    def __init__(self, code, function, message=""):
        self.code = code
        super().__init__(f"{function} returned {code}: {message}")


def parse_header(path=HEADER):
    """Returns {name: (restype, [argtypes])} for every function declared in the header."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"^\s*#[^\n]*", "", text, flags=re.M)
    text = text.replace('extern "C" {', "").replace("}", "")
    protos = {}
    decl = re.compile(r"([A-Za-z_][\w\s\*]*?)\b(qmckl_\w+)\s*\(([^;{]*?)\)\s*;", re.S)
    for m in decl.finditer(text):
        ret, name, params = m.group(1).strip(), m.group(2), m.group(3).strip()
        restype = _ctype_of(ret)
        args = []
        if params and params != "void":
            for p in params.split(","):
                p = " ".join(p.split())
                p = re.sub(r"\b(const)\b", "", p).strip()
                base = p.split()[0]
                is_ptr = "*" in p
                if is_ptr:
                    args.append(ctypes.c_void_p)
                else:
                    args.append(_CTYPES[base])
        protos[name] = (restype, args)
    return protos


def _ctype_of(ret):
    ret = " ".join(ret.replace("const", "").split())
    if ret.endswith("*"):
        return ctypes.c_char_p if ret.startswith("char") else ctypes.c_void_p
    return _CTYPES[ret]


_LIB = None


def lib():
    """The loaded C-ABI library with prototypes from include/qmckl.h."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run `python -m qmckl_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (restype, argtypes) in parse_header().items():
            f = getattr(L, name)
            f.restype = restype
            f.argtypes = argtypes
        _LIB = L
    return _LIB


def ptr(a):
    """Address of a numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data


def last_error(ctx):
    buf = ctypes.create_string_buffer(4096)
    lib().qmckl_last_error(ctx, buf)
    return buf.value.decode(errors="replace")


def check(ctx, rc, fn="qmckl"):
    if rc != 0:
        raise QmcklError(rc, fn, last_error(ctx) if ctx else "")
    return rc


# ---- SWIG-style layer --------------------------------------------------------------

QMCKL_SUCCESS = 0
QMCKL_FAILURE = 101
QMCKL_INVALID_CONTEXT = 103
QMCKL_NOT_PROVIDED = 106
QMCKL_ALREADY_SET = 108
MO_FP64, MO_FP32, MO_BF16X3 = 0, 1, 2

_I64 = np.int64
_I32 = np.int32
_F64 = np.float64


def context_create():
    return lib().qmckl_context_create()


def context_destroy(ctx):
    check(ctx, lib().qmckl_context_destroy(ctx), "qmckl_context_destroy")


def _call(ctx, name, *args):
    rc = getattr(lib(), name)(ctx, *args)
    check(ctx, rc, name)


def _arr(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def set_array(ctx, name, values, dtype=_F64):
    """qmckl_<name>(ctx, values, size) for an array setter."""
    a = _arr(values, dtype)
    _call(ctx, name, ptr(a), a.size)


def set_with_transp(ctx, name, transp, values, *extra):
    a = _arr(values, _F64)
    _call(ctx, name, transp.encode() if isinstance(transp, str) else transp, *extra, ptr(a), a.size)


def get_array(ctx, name, size, dtype=_F64):
    """qmckl_<name>(ctx, out, size) for an array getter; returns the array."""
    out = np.empty(int(size), dtype=dtype)
    _call(ctx, name, ptr(out), out.size)
    return out


def get_scalar(ctx, name, ctype=ctypes.c_int64):
    v = ctype()
    _call(ctx, name, ctypes.byref(v))
    return v.value


def synchronize(ctx):
    _call(ctx, "qmckl_amd_synchronize")


def trexio_read(ctx, file_name):
    """qmckl_trexio_read (org/qmckl_trexio.org:1104): electron numbers, nuclei, AO basis and MO
    coefficients from a TREXIO file of the TEXT back end (a directory of <group>.txt files)."""
    name = os.fsencode(file_name)
    _call(ctx, "qmckl_trexio_read", name, len(name))


# ---- device buffers over the HIP runtime the library itself uses --------------------
# (torch ships its own HIP runtime; mixing two runtimes in one process is not supported,
# so device memory for the qmckl_amd_*_device entry points is managed here.)

_HIP = None


def hip():
    """The HIP runtime libqmckl_amd.so is linked against."""
    global _HIP
    if _HIP is None:
        lib()  # loads libamdhip64 as a dependency of the library
        H = ctypes.CDLL("libamdhip64.so.7")
        H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        H.hipFree.argtypes = [ctypes.c_void_p]
        H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        H.hipDeviceSynchronize.argtypes = []
        H.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        _HIP = H
    return _HIP


def device_count():
    n = ctypes.c_int(0)
    if hip().hipGetDeviceCount(ctypes.byref(n)) != 0:
        return 0
    return n.value


class DeviceBuffer:
    """A device allocation; `ptr` is the device address passed to the C-ABI."""

    H2D, D2H, D2D = 1, 2, 3

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        rc = hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 8))
        if rc != 0:
            raise MemoryError(f"hipMalloc({self.nbytes}) failed with {rc}")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        rc = hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, self.H2D)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D failed with {rc}")

    def to_array(self, shape, dtype=np.float64):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        rc = hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, self.D2H)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H failed with {rc}")
        return out

    def copy_from(self, other, nbytes=None):
        """Device-to-device copy of `nbytes` (default: all of `other`)."""
        n = other.nbytes if nbytes is None else int(nbytes)
        assert n <= self.nbytes and n <= other.nbytes
        rc = hip().hipMemcpy(self.ptr, other.ptr, n, self.D2D)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2D failed with {rc}")

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_synchronize():
    rc = hip().hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed with {rc}")
