"""ctypes bindings to libdvm_hip.so, the C-ABI declared in include/dvm_hip.h.

Device arrays are torch tensors (float64, contiguous, on the context's
device); torch supplies device memory and the stream, the computation is the
library's HIP kernels.  There is no CPU path: if libdvm_hip.so is missing or
no GPU is present, every call raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdvm_hip.so")

F64 = 0
F32C = 1

OK = 0
_STATUS = {0: "ok", 1: "invalid argument", 2: "size mismatch", 3: "non-finite release counts",
           4: "HIP runtime error", 5: "out of memory", 6: "invalid state"}


class DvmError(RuntimeError):
    """A non-zero dvm_status; `.status` holds the code."""

    def __init__(self, status, msg):
        super().__init__(f"{_STATUS.get(status, status)}: {msg}")
        self.status = status


_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_sz = ctypes.c_size_t
_st = ctypes.c_int

# name -> (restype, argtypes); pointers are passed as c_void_p
SIGNATURES = {
    "dvm_ctx_create": (_st, [_int, _vp, ctypes.POINTER(_vp)]),
    "dvm_ctx_destroy": (_st, [_vp]),
    "dvm_ctx_sync": (_st, [_vp]),
    "dvm_ctx_stream": (_vp, [_vp]),
    "dvm_last_error": (ctypes.c_char_p, [_vp]),
    "dvm_status_string": (ctypes.c_char_p, [_int]),
    "dvm_version": (ctypes.c_char_p, []),
    "dvm_prof_enable": (_st, [_vp, _int]),
    "dvm_prof_reset": (_st, [_vp]),
    "dvm_prof_read": (_st, [_vp, ctypes.c_char_p, ctypes.POINTER(_dbl), ctypes.POINTER(_i64)]),
    "dvm_malloc": (_st, [_vp, _sz, ctypes.POINTER(_vp)]),
    "dvm_free": (_st, [_vp, _vp]),
    "dvm_memcpy_h2d": (_st, [_vp, _vp, _vp, _sz]),
    "dvm_memcpy_d2h": (_st, [_vp, _vp, _vp, _sz]),
    "dvm_memcpy_d2d": (_st, [_vp, _vp, _vp, _sz]),
    "dvm_bs_self": (_st, [_vp, _int, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dvm_bs_targets": (_st, [_vp, _int, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _int, _dbl,
                             _vp, _vp]),
    "dvm_bs_self_host": (_st, [_vp, _int, _i64, _vp, _vp]),
    "dvm_bs_targets_host": (_st, [_vp, _int, _i64, _vp, _i64, _vp, _vp, _vp]),
}

_lib = None


def load():
    """Load libdvm_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing; build it with `python -c "
                           f"'import __graft_entry__ as g; g.build()'` or `make -C dvmpp_amd/csrc`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(ctx_handle, status):
    if status != OK:
        msg = load().dvm_last_error(ctx_handle)
        raise DvmError(status, msg.decode() if msg else "")
