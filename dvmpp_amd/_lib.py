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
FMM = 2

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
    "dvm_bs_self_noalloc": (_st, [_vp, _int, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dvm_bs_self_noalloc_scratch_bytes": (_sz, []),
    "dvm_bs_targets": (_st, [_vp, _int, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _int, _dbl,
                             _vp, _vp]),
    "dvm_bs_self_host": (_st, [_vp, _int, _i64, _vp, _vp]),
    "dvm_bs_targets_host": (_st, [_vp, _int, _i64, _vp, _i64, _vp, _vp, _vp]),
    "dvm_set_tuning": (_st, [_vp, _int]),
    "dvm_set_f32c_variant": (_st, [_vp, _int, _int, _int, _int]),
    "dvm_get_f32c_variant": (_st, [_vp, ctypes.POINTER(ctypes.c_int32)]),
}

TUNE_TABLE = 0
TUNE_MEASURE = 1

class SheetParams(ctypes.Structure):
    """dvm_sheet_params."""
    _fields_ = [("rho", _dbl), ("nu", _dbl), ("dt", _dbl), ("ux", _dbl), ("uz", _dbl),
                ("cutoff_exp", _dbl), ("max_npv", _int)]


class RemeshParams(ctypes.Structure):
    """dvm_remesh_params."""
    _fields_ = [("frequency", _int), ("kernel", _int), ("xmin", _dbl), ("xmax", _dbl),
                ("zmin", _dbl), ("zmax", _dbl), ("dx", _dbl), ("dz", _dbl), ("sigma", _dbl),
                ("circ_threshold", _dbl)]


class SimParams(ctypes.Structure):
    """dvm_sim_params."""
    _fields_ = [("sheet", SheetParams), ("scheme", _int), ("crossing", _int), ("seed", _int),
                ("prec", _int), ("rng_mode", _int), ("remesh", RemeshParams),
                ("merge_threshold", _dbl), ("merge_frequency", _int)]


class BatchConfig(ctypes.Structure):
    """dvm_batch_config."""
    _fields_ = [("slots", _int), ("slot_capacity", _i64), ("max_panels", _int), ("scheme", _int),
                ("crossing", _int), ("prec", _int), ("use_graph", _int),
                ("remesh", RemeshParams)]


_dp = ctypes.POINTER(_dbl)
_u32p = ctypes.POINTER(ctypes.c_uint32)

SIGNATURES.update({
    "dvm_sheet_create": (_st, [_vp, _int, _vp, _vp, ctypes.POINTER(SheetParams),
                               ctypes.POINTER(_vp)]),
    "dvm_sheet_destroy": (_st, [_vp]),
    "dvm_sheet_size": (_st, [_vp, ctypes.POINTER(_int)]),
    "dvm_sheet_geometry": (_st, [_vp] + [_vp] * 9),
    "dvm_sheet_influence": (_st, [_vp, _vp]),
    "dvm_sheet_gamma_ptr": (_st, [_vp, ctypes.POINTER(_vp)]),
    "dvm_sheet_get_gamma": (_st, [_vp, _vp]),
    "dvm_sheet_set_gamma": (_st, [_vp, _vp]),
    "dvm_sheet_solve": (_st, [_vp, _int, _i64, _vp, _vp, _vp, _vp, _vp]),
    "dvm_sheet_on_blobs": (_st, [_vp, _i64, _vp, _vp, _vp, _vp]),
    "dvm_sheet_on_blobs_far": (_st, [_vp, _i64, _vp, _vp, _vp, _vp]),
    "dvm_sheet_loads": (_st, [_vp, _dbl, _vp]),
    "dvm_sheet_probe_quirk": (_st, [_vp, _int, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp,
                                    _vp, _vp, _vp]),
    "dvm_sheet_inside": (_st, [_vp, _i64, _vp, _vp, _vp]),
    "dvm_sheet_reflect": (_st, [_vp, _i64, _vp, _vp]),
    "dvm_remesh": (_st, [_vp, ctypes.POINTER(RemeshParams), _i64, _vp, _vp, _vp, _vp, _i64,
                         _vp, _vp, _vp, _vp, ctypes.POINTER(_i64)]),
    "dvm_advect_euler": (_st, [_vp, _i64, _dbl, _dbl, _dbl] + [_vp] * 6),
    "dvm_advect_rk3_stage": (_st, [_vp, _i64, _dbl, _dbl, _dbl, _dbl, _dbl] + [_vp] * 8),
    "dvm_diffuse": (_st, [_vp, _i64, _vp, _vp, _vp]),
    "dvm_sheet_release_counts": (_st, [_vp, ctypes.POINTER(_i64)]),
    "dvm_sheet_release_emit": (_st, [_vp, _i64, _vp, ctypes.c_uint32, _i64] + [_vp] * 9),
    "dvm_sheet_delete_inside": (_st, [_vp, _i64] + [_vp] * 9 + [ctypes.POINTER(_i64)]),
    "dvm_sim_create": (_st, [_vp, _int, _vp, _vp, ctypes.POINTER(SimParams),
                             ctypes.POINTER(_vp)]),
    "dvm_sim_destroy": (_st, [_vp]),
    "dvm_sim_step": (_st, [_vp]),
    "dvm_sim_run": (_st, [_vp, _int]),
    "dvm_sim_substep": (_st, [_vp, _int]),
    "dvm_sim_size": (_st, [_vp, ctypes.POINTER(_i64)]),
    "dvm_sim_step_index": (_st, [_vp, ctypes.POINTER(_int)]),
    "dvm_sim_rng_draws": (_st, [_vp, ctypes.POINTER(ctypes.c_uint64)]),
    "dvm_sim_get_state": (_st, [_vp] + [_vp] * 9),
    "dvm_sim_set_state": (_st, [_vp, _i64] + [_vp] * 9),
    "dvm_sim_get_gamma": (_st, [_vp, _vp]),
    "dvm_sim_set_gamma": (_st, [_vp, _vp]),
    "dvm_sim_loads": (_st, [_vp, _vp]),
    "dvm_sim_blobs": (_st, [_vp] + [ctypes.POINTER(_vp)] * 4),
    "dvm_sim_merged": (_st, [_vp, ctypes.POINTER(_i64)]),
    "dvm_merge_blobs": (_st, [_vp, _dbl, _i64] + [_vp] * 9 + [ctypes.POINTER(_i64)]),
    "dvm_merge_stats": (_st, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "dvm_set_fmm": (_st, [_vp, _int, _dbl]),
    "dvm_fmm_last_level": (_st, [_vp, ctypes.POINTER(_int)]),
    "dvm_sim_snapshot_size": (_st, [_vp, ctypes.POINTER(_sz)]),
    "dvm_sim_snapshot": (_st, [_vp, _vp, _sz]),
    "dvm_sim_restore": (_st, [_vp, _vp, _sz]),
    "dvm_sim_copy_out": (_st, [_vp, _i64] + [_vp] * 5 + [ctypes.POINTER(_i64)]),
    "dvm_sim_probe_quirk": (_st, [_vp, _int, _i64] + [_vp] * 6),
    "dvm_event_create": (_st, [_vp, ctypes.POINTER(_vp)]),
    "dvm_event_record": (_st, [_vp, _vp]),
    "dvm_event_wait": (_st, [_vp]),
    "dvm_event_query": (_st, [_vp, ctypes.POINTER(_int)]),
    "dvm_event_destroy": (_st, [_vp]),
    "dvm_host_alloc": (_st, [_sz, ctypes.POINTER(_vp)]),
    "dvm_host_free": (_st, [_vp]),
    "dvm_batch_footprint": (_st, [ctypes.POINTER(BatchConfig), ctypes.POINTER(_sz)]),
    "dvm_batch_create": (_st, [_vp, ctypes.POINTER(BatchConfig), ctypes.POINTER(_vp)]),
    "dvm_batch_destroy": (_st, [_vp]),
    "dvm_batch_admit": (_st, [_vp, _int, _vp, _vp, ctypes.POINTER(SheetParams), _int, _int,
                              ctypes.POINTER(_int)]),
    "dvm_batch_step": (_st, [_vp, _int]),
    "dvm_batch_poll": (_st, [_vp, _int, _vp, _vp, ctypes.POINTER(_int)]),
    "dvm_batch_slot_info": (_st, [_vp, _int, ctypes.POINTER(_i64), ctypes.POINTER(_int),
                                  ctypes.POINTER(_int), ctypes.POINTER(_int), _vp]),
    "dvm_batch_admit_many": (_st, [_vp, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dvm_batch_set_state": (_st, [_vp, _int, _i64, _vp, _vp, _vp, _vp]),
    "dvm_batch_sort_order": (_st, [_vp, _int, _i64, _vp]),
    "dvm_batch_set_fmm": (_st, [_vp, _i64, _int, _dbl]),
    "dvm_batch_get_state": (_st, [_vp, _int] + [_vp] * 6),
})

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
