"""ctypes binding of libiirx.so (the C ABI declared in include/iirx.h).

The library is built in-tree by cuda_recursive_filtering_amd.build; there is
no fallback: if it is missing, loading raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libiirx.so")

IIRX_OK = 0
IIRX_ERR_INVALID = 1
IIRX_ERR_GEOMETRY = 2
IIRX_ERR_CONDITIONING = 3
IIRX_ERR_HIP = 4
IIRX_ERR_NO_DEVICE = 5
IIRX_ERR_OOM = 6
IIRX_ERR_TIMEOUT = 7
IIRX_ERR_TOO_LARGE = 8
IIRX_ERR_ACCURACY = 9

SCHEDULES = {"persistent": 0, "ticket": 1}

KERNELS = {"PH": 0, "STCR": 1, "DTCR": 2, "PLR": 3, "BS": 4, "BS_BF16": 5, "PHJ": 6, "PHH": 7}


class Config(ctypes.Structure):
    _fields_ = [
        ("kernel", ctypes.c_int32),
        ("block_size", ctypes.c_int32),
        ("n_blocks", ctypes.c_int32),
        ("n_sections", ctypes.c_int32),
        ("unrolled", ctypes.c_int32),
        ("plr_x", ctypes.c_int32),
        ("plr_x_max", ctypes.c_int32),
        ("schedule", ctypes.c_int32),
    ]


class PlanInfo(ctypes.Structure):
    _fields_ = [
        ("n_samples", ctypes.c_int64),
        ("chunk_samples", ctypes.c_int64),
        ("num_chunks", ctypes.c_int64),
        ("threads_per_block", ctypes.c_int32),
        ("lds_bytes", ctypes.c_int32),
        ("vgprs", ctypes.c_int32),
        ("sgprs", ctypes.c_int32),
        ("scratch_bytes", ctypes.c_int32),
        ("waves_per_simd", ctypes.c_int32),
        ("launches", ctypes.c_int64),
        ("grid", ctypes.c_int64),
        ("schedule", ctypes.c_int32),
        ("halo", ctypes.c_int32),
    ]


BATCH_ALGOS = {"auto": -1, "ph": 0, "blockstate": 1}

TUNE_NO_BLOCKSTATE = 1
TUNE_NO_CR = 2
TUNE_PLR_ALL_ORDERS = 4
TUNE_TICKET = 8


class TuneEntry(ctypes.Structure):
    _fields_ = [
        ("cfg", Config),
        ("status", ctypes.c_int32),
        ("ms", ctypes.c_float),
        ("rel_err", ctypes.c_float),
    ]


class BatchConfig(ctypes.Structure):
    _fields_ = [
        ("n_blocks", ctypes.c_int32),
        ("schedule", ctypes.c_int32),
        ("bs_min_sections", ctypes.c_int32),
        ("bs_tiles", ctypes.c_int32),
    ]


class BatchInfo(ctypes.Structure):
    _fields_ = [
        ("n_channels", ctypes.c_int64),
        ("n_samples", ctypes.c_int64),
        ("num_chunks", ctypes.c_int64),
        ("chunk_samples", ctypes.c_int64),
        ("grid", ctypes.c_int64),
        ("launches", ctypes.c_int64),
        ("vgprs", ctypes.c_int32),
        ("waves_per_simd", ctypes.c_int32),
        ("graph_ready", ctypes.c_int32),
        ("n_launches", ctypes.c_int32),
        ("n_blockstate", ctypes.c_int64),
        ("bs_chunks", ctypes.c_int64),
        ("bs_chunk_samples", ctypes.c_int64),
        ("bs_grid", ctypes.c_int64),
        ("bs_vgprs", ctypes.c_int32),
        ("bs_waves_per_simd", ctypes.c_int32),
    ]


TAPS_MAX_SECTIONS = 8


class FilterTaps(ctypes.Structure):
    """iirx_filter_taps: the reference's generated filter header as data."""
    _fields_ = [
        ("n_sections", ctypes.c_int32),
        ("plr_order", ctypes.c_int32),
        ("gain", ctypes.c_double),
        ("input_fs", ctypes.c_double),
        ("input_freq", ctypes.c_double),
        ("cutoff_wn", ctypes.c_double),
        ("sos", ctypes.c_float * (TAPS_MAX_SECTIONS * 6)),
        ("plr_b", ctypes.c_double * (2 * TAPS_MAX_SECTIONS)),
        ("plr_a", ctypes.c_double * (2 * TAPS_MAX_SECTIONS)),
    ]


# every symbol include/iirx.h declares, with its ctypes signature
_P = ctypes.c_void_p
_SIGS = {
    "iirx_plan_create": (ctypes.c_int, [ctypes.POINTER(Config), _P, ctypes.c_int64, ctypes.POINTER(_P)]),
    "iirx_plan_set_plr_taps": (ctypes.c_int, [_P, _P, _P, ctypes.c_int]),
    "iirx_execute": (ctypes.c_int, [_P, _P, _P, _P]),
    "iirx_plan_info_get": (ctypes.c_int, [_P, ctypes.POINTER(PlanInfo)]),
    "iirx_plan_destroy": (ctypes.c_int, [_P]),
    "iirx_plan_check": (ctypes.c_int, [_P]),
    "iirx_plan_set_spin_limit": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "iirx_config_check": (ctypes.c_int, [ctypes.POINTER(Config), ctypes.c_int64]),
    "iirx_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "iirx_measure": (ctypes.c_int, [_P, _P, ctypes.c_int, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P]),
    "iirx_version": (ctypes.c_char_p, []),
    "iirx_butterworth_sos": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, _P, _P]),
    "iirx_chebyshev_sos": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _P, _P]),
    "iirx_direct_form_taps": (ctypes.c_int, [_P, ctypes.c_int, _P, _P]),
    "iirx_fill_normal": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int64, _P]),
    "iirx_plan_tables": (ctypes.c_int, [ctypes.POINTER(Config), _P, ctypes.c_int64, _P, ctypes.c_int64,
                                        ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32)]),
    "iirx_tune_candidates": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32, _P, ctypes.c_int32,
                                            ctypes.POINTER(ctypes.c_int32)]),
    "iirx_autotune": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32, ctypes.c_int32,
                                     ctypes.POINTER(Config), ctypes.POINTER(ctypes.c_float), _P, ctypes.c_int32,
                                     ctypes.POINTER(ctypes.c_int32)]),
    "iirx_batch_create": (ctypes.c_int, [ctypes.POINTER(BatchConfig), ctypes.c_int64, _P, _P, _P,
                                         ctypes.POINTER(_P)]),
    "iirx_batch_create_ex": (ctypes.c_int, [ctypes.POINTER(BatchConfig), ctypes.c_int64, _P, _P, _P, _P,
                                            ctypes.POINTER(_P)]),
    "iirx_batch_execute": (ctypes.c_int, [_P, _P, _P, _P]),
    "iirx_batch_capture": (ctypes.c_int, [_P, _P, _P, _P]),
    "iirx_batch_replay": (ctypes.c_int, [_P, _P]),
    "iirx_batch_check": (ctypes.c_int, [_P]),
    "iirx_batch_autotune": (ctypes.c_int, [ctypes.POINTER(BatchConfig), ctypes.c_int64, _P, _P, _P, ctypes.c_int32,
                                           ctypes.POINTER(BatchConfig), ctypes.POINTER(ctypes.c_float)]),
    "iirx_batch_info_get": (ctypes.c_int, [_P, ctypes.POINTER(BatchInfo)]),
    "iirx_batch_destroy": (ctypes.c_int, [_P]),
    "iirx_execute_state": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "iirx_copy_ceiling": (ctypes.c_int, [_P, _P, ctypes.c_int64, ctypes.c_int32, _P]),
    "iirx_execute_state_noalloc": (ctypes.c_int, [ctypes.POINTER(Config), _P, ctypes.c_int64, _P, _P, _P, _P, _P]),
    "iirx_normalize_sos": (ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.POINTER(ctypes.c_double)]),
    "iirx_batch_read_tables": (ctypes.c_int, [_P, _P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]),
    "iirx_plan_read_tables": (ctypes.c_int, [_P, _P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]),
    "iirx_filter_taps_design": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.POINTER(FilterTaps)]),
    "iirx_filter_taps_write": (ctypes.c_int, [ctypes.POINTER(FilterTaps), ctypes.c_char_p]),
    "iirx_filter_taps_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(FilterTaps)]),
}

_lib = None


class IirxError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        msg = lib().iirx_status_string(code).decode() if _lib is not None else str(code)
        super().__init__(f"{what}: {msg} (status {code})" if what else f"{msg} (status {code})")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -m cuda_recursive_filtering_amd.build` "
                "(the HIP library is required, there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code, what=""):
    if code != IIRX_OK:
        raise IirxError(code, what)
