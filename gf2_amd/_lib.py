"""ctypes binding of libgf2ldpc.so (the HIP decoder's C ABI, include/gf2ldpc.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C gf2_amd/csrc``).  There is no CPU fallback: when the library is
missing, or when no HIP device is present, every decode call raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgf2ldpc.so")

GF2L_OK = 0
GF2L_ERR_INVALID = -1
GF2L_ERR_HIP = -2
GF2L_ERR_NOMEM = -3
GF2L_ERR_UNSUPPORTED = -4

PATH_AUTO = 0
PATH_LDS = 1
PATH_STREAM = 2
PATH_STATE = 3
PATH_CSR = 4

RULE_MINSUM = 0
RULE_SPA = 1

TUNE_WG_PER_CU = 1
TUNE_XCD_MAP = 2
TUNE_GRAPH_ITERS = 3
TUNE_CHUNK = 4
TUNE_LAYERED_COLUMNS = 5
TUNE_LAYERED_KERNEL = 6
TUNE_LDS_GROUPS = 7
TUNE_HBM_GRAPH = 8
LAYERED_AUTO, LAYERED_GENERAL, LAYERED_STAIRCASE = 0, 1, 2

KERNEL_CHECK = 0
KERNEL_VARIABLE = 1
KERNEL_SYNDROME = 2
KERNEL_LAYERED = 3
PROFILE_ALL = 1
PROFILE_LAYERED = 2

FALLBACK_RESTART = 0
FALLBACK_CONTINUE = 1

_c_void_pp = ctypes.POINTER(ctypes.c_void_p)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)

# name -> (restype, argtypes); every symbol include/gf2ldpc.h declares.
SIGNATURES = {
    "gf2l_create": (ctypes.c_int, [ctypes.c_int, _c_void_pp]),
    "gf2l_destroy": (None, [ctypes.c_void_p]),
    "gf2l_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "gf2l_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "gf2l_stream": (ctypes.c_void_p, [ctypes.c_void_p]),
    "gf2l_set_path": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "gf2l_code_path": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "gf2l_set_check_rule": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "gf2l_device_alloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, _c_void_pp]),
    "gf2l_device_free": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "gf2l_copy_to_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_size_t]),
    "gf2l_copy_to_host": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_size_t]),
    "gf2l_frames_generate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                            ctypes.c_uint64, ctypes.c_size_t, ctypes.c_float,
                                            ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p]),
    "gf2l_frames_generate_zero": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                                 ctypes.c_uint64, ctypes.c_size_t, ctypes.c_float,
                                                 ctypes.c_float, ctypes.c_void_p]),
    "gf2l_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "gf2l_profile_enable": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "gf2l_profile_read": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p]),
    "gf2l_profile_read_kernel": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_void_p]),
    "gf2l_code_qc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                    ctypes.c_uint32, _i32p, _c_void_pp]),
    "gf2l_code_edges": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.c_void_p, ctypes.c_size_t, _c_void_pp]),
    "gf2l_code_info": (ctypes.c_int, [ctypes.c_void_p, _u32p, _u32p, _u32p, _u64p]),
    "gf2l_code_destroy": (None, [ctypes.c_void_p]),
    "gf2l_qc_ira_shifts": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.c_uint32, _u32p, _i32p]),
    "gf2l_dvb_t2_params": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _u32p, _u32p, _u32p, _u32p,
                                          _u32p]),
    "gf2l_dvb_t2_table": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "gf2l_code_dvb_table": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                           ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                           ctypes.c_void_p, ctypes.c_void_p, _c_void_pp]),
    "gf2l_code_dvb_t2": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _c_void_pp]),
    "gf2l_code_gf2data": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t,
                                         _c_void_pp]),
    "gf2l_decode_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_size_t, ctypes.c_uint32, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "gf2l_decode_batch_device_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                        ctypes.c_size_t, ctypes.c_uint32, ctypes.c_void_p,
                                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "gf2l_decode_batch_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_size_t,
                                                ctypes.c_uint32, ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p]),
    "gf2l_decode_trace": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "gf2l_nms_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                       ctypes.c_void_p, ctypes.c_float, _c_void_pp]),
    "gf2l_nms_create_layered": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                               ctypes.c_void_p, ctypes.c_float, _c_void_pp]),
    "gf2l_nms_set_fallback": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "gf2l_nms_layered_columns": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _u32p]),
    "gf2l_nms_destroy": (None, [ctypes.c_void_p]),
    "gf2l_nms_info": (ctypes.c_int, [ctypes.c_void_p, _u32p, _u32p]),
    "gf2l_nms_get_weights": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "gf2l_nms_set_weights": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "gf2l_nms_get_adam": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, _u64p]),
    "gf2l_nms_set_adam": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_uint64]),
    "gf2l_nms_gradient": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                         ctypes.c_void_p]),
    "gf2l_nms_train_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_size_t, ctypes.c_float,
                                           ctypes.c_void_p]),
    "gf2l_nms_decode_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_size_t, ctypes.c_uint32, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p]),
    "gf2l_codeset_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                           _c_void_pp]),
    "gf2l_codeset_destroy": (None, [ctypes.c_void_p]),
    "gf2l_decode_mixed_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_size_t, ctypes.c_void_p, ctypes.c_uint32,
                                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p]),
    "gf2l_decode_mixed_nms_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                    ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                                    ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "gf2_ldpc_decoder_create": (ctypes.c_void_p, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
    "gf2_ldpc_decode": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                         ctypes.c_void_p]),
    "gf2_ldpc_decode_frames": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                                ctypes.c_void_p, ctypes.c_void_p]),
    "gf2_ldpc_decoder_info": (ctypes.c_int, [ctypes.c_void_p, _u32p, _u32p]),
    "gf2_ldpc_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "gf2_ldpc_decoder_destroy": (None, [ctypes.c_void_p]),
}

_lib = None
_lock = threading.Lock()


class Gf2lError(RuntimeError):
    """A libgf2ldpc call returned a negative status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"gf2ldpc error {code}: {message}")
        self.code = code
        self.message = message


def load() -> ctypes.CDLL:
    """Loads libgf2ldpc.so once; raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                    "(the HIP decoder has no CPU fallback)")
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
        return _lib


def check(ctx, rc: int) -> None:
    if rc != GF2L_OK:
        msg = load().gf2l_last_error(ctx).decode() if ctx else ""
        raise Gf2lError(rc, msg)


def ptr(arr, ctype):
    """ctypes pointer to a contiguous numpy array."""
    return arr.ctypes.data_as(ctypes.POINTER(ctype))
