"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: imported only by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker.
"""
import ctypes
import os
import subprocess

import numpy as np

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_ORACLE_DIR = os.path.join(_ROOT, "oracle")
_LIB_PATH = os.path.join(_ORACLE_DIR, "liboracle.so")
# the same restatements with double-rounded multiply-adds, the rounding of the
# reference's Python emulators (for bit-exact comparison with their outputs)
_LIB_DR_PATH = os.path.join(_ORACLE_DIR, "liboracle_dr.so")
_libs = {}

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def build():
    subprocess.check_call(["make", "-s", "-C", _ORACLE_DIR])


def _load(path):
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    L.orc_sosfilt_f64.argtypes = [_f64p, ctypes.c_int, _f64p, ctypes.c_long, _f64p]
    L.orc_golden_f64.argtypes = [_f32p, ctypes.c_int, _f32p, ctypes.c_long, _f64p]
    L.orc_ph_f32.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_long, _f32p]
    L.orc_ph_f32_mt.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_long, _f32p,
                                ctypes.c_int]
    L.orc_stcr_f32.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               _f32p, ctypes.c_long, _f32p]
    L.orc_dtcr_f32.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_long, _f32p]
    L.orc_plr_f32.argtypes = [_f64p, _f64p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              _f32p, ctypes.c_long, _f32p]
    L.orc_direct_form_f32.argtypes = [_f64p, _f64p, ctypes.c_int, _f32p, ctypes.c_long, _f32p]
    L.orc_cascade_f32.argtypes = [_f32p, ctypes.c_int, _f32p, ctypes.c_long, _f32p]
    for f in (L.orc_ph_f32, L.orc_ph_f32_mt, L.orc_stcr_f32, L.orc_dtcr_f32, L.orc_plr_f32):
        f.restype = ctypes.c_int
    return L


def lib(double_round=False):
    """The oracle library; double_round=True: the build whose multiply-adds
    round the product and the sum separately, as the emulators do."""
    key = bool(double_round)
    if key not in _libs:
        _libs[key] = _load(_LIB_DR_PATH if key else _LIB_PATH)
    return _libs[key]


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def sosfilt_f64(sos, x):
    sos = _f64(sos)
    x = _f64(x)
    y = np.empty_like(x)
    lib().orc_sosfilt_f64(sos.reshape(-1), sos.shape[0], x, x.size, y)
    return y


def golden(sos_f32, x_f32):
    """float64 filtering of float32 sections and float32 input (ref_generate.py:89)."""
    sos = _f32(sos_f32)
    x = _f32(x_f32)
    y = np.empty(x.size, np.float64)
    lib().orc_golden_f64(sos.reshape(-1), sos.shape[0], x, x.size, y)
    return y


def _check(rc, what):
    if rc != 0:
        raise ValueError(f"oracle {what}: invalid arguments (rc={rc})")


def ph(sos, x, B, N, double_round=False):
    sos = _f32(sos); x = _f32(x); y = np.empty_like(x)
    _check(lib(double_round).orc_ph_f32(sos.reshape(-1), sos.shape[0], B, N, x, x.size, y), "ph")
    return y


def ph_mt(sos, x, B, N, threads):
    """orc_ph_f32 on `threads` host threads (bit-identical output)."""
    sos = _f32(sos); x = _f32(x); y = np.empty_like(x)
    _check(lib().orc_ph_f32_mt(sos.reshape(-1), sos.shape[0], B, N, x, x.size, y, int(threads)), "ph_mt")
    return y


def stcr(sos, x, B, N, unrolled=False, double_round=False):
    sos = _f32(sos); x = _f32(x); y = np.empty_like(x)
    _check(lib(double_round).orc_stcr_f32(sos.reshape(-1), sos.shape[0], B, N, int(unrolled), x, x.size, y), "stcr")
    return y


def dtcr(sos, x, B, N, double_round=False):
    sos = _f32(sos); x = _f32(x); y = np.empty_like(x)
    _check(lib(double_round).orc_dtcr_f32(sos.reshape(-1), sos.shape[0], B, N, x, x.size, y), "dtcr")
    return y


def plr(bt, at, x, block, X, double_round=False):
    bt = _f64(bt); at = _f64(at); x = _f32(x); y = np.empty_like(x)
    _check(lib(double_round).orc_plr_f32(bt, at, bt.size, block, X, x, x.size, y), "plr")
    return y


def direct_form_f32(bt, at, x):
    bt = _f64(bt); at = _f64(at); x = _f32(x); y = np.empty_like(x)
    lib().orc_direct_form_f32(bt, at, bt.size, x, x.size, y)
    return y


def cascade_f32(sos, x, double_round=False):
    sos = _f32(sos); x = _f32(x); y = np.empty_like(x)
    lib(double_round).orc_cascade_f32(sos.reshape(-1), sos.shape[0], x, x.size, y)
    return y


def gate(out, ref, rel_tol=1e-4):
    """Accuracy gate of test_accuracy.cu:178-190: max|out-ref| <= rel_tol*max|ref|.
    Returns (passed, max_abs_err, max_abs_ref, rel)."""
    out = np.asarray(out, np.float64)
    ref = np.asarray(ref, np.float64)
    if not np.all(np.isfinite(out)):
        return False, float("inf"), float(np.max(np.abs(ref))), float("inf")
    err = float(np.max(np.abs(out - ref))) if out.size else 0.0
    mref = float(np.max(np.abs(ref))) if ref.size else 0.0
    rel = err / mref if mref > 0 else err
    return err <= rel_tol * mref, err, mref, rel
