"""MI355X-native parallel cascaded-biquad IIR filtering (host-side mirror).

Python view of the C ABI in include/iirx.h.  It mirrors the reference's
operator interface (Haotian-RA/cuda_recursive_filtering):

* the reference picks KERNEL (PH / STCR / DTCR / PLR), BLOCK_SIZE, N_BLOCKS,
  N_SECTIONS and HANDUNROLLED at compile time (run_accuracy.sh:22-60) and
  uploads the sections with setup_kernel_coefficients(sos)
  (ph_kernels.cuh:151); here `Plan(kernel, sos, n_samples, block_size,
  n_blocks, unrolled)` does both at run time;
* one `Plan.execute(d_in, d_out)` is one grid-wide launch
  (test_accuracy.cu:146);
* sections follow the reference's contract: b0 == 1, the cascade gain is
  applied to the input by the caller (ref_generate.py:61-67).

Device memory and streams come from PyTorch (tensors' data pointers and
torch.cuda streams); the computation is libiirx.so's HIP kernels only.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import IirxError, KERNELS, check, lib

__all__ = [
    "Plan", "Batch", "IirxError", "KERNELS", "butterworth_sos", "normalize_sos", "direct_form_taps", "fill_normal",
    "filter_taps_design", "write_filter_taps", "read_filter_taps",
    "plan_tables", "config_check", "version", "tune_candidates", "autotune",
]


def version():
    return lib().iirx_version().decode()


def _config(kernel, n_sections, block_size=64, n_blocks=64, unrolled=False, plr_x=0, plr_x_max=0,
            schedule="persistent"):
    if isinstance(kernel, str):
        if kernel.upper() not in KERNELS:
            raise ValueError(f"unknown kernel {kernel!r}; expected one of {sorted(KERNELS)}")
        kernel = KERNELS[kernel.upper()]
    c = _lib.Config()
    c.kernel = int(kernel)
    c.block_size = int(block_size)
    c.n_blocks = int(n_blocks)
    c.n_sections = int(n_sections)
    c.unrolled = int(bool(unrolled))
    c.plr_x = int(plr_x)
    c.plr_x_max = int(plr_x_max)
    c.schedule = _lib.SCHEDULES[schedule] if isinstance(schedule, str) else int(schedule)
    return c


def _sos_array(sos):
    s = np.ascontiguousarray(np.asarray(sos, dtype=np.float32).reshape(-1, 6))
    return s


def config_check(kernel, n_sections, n_samples, **kw):
    """Validity of a configuration without a device (gpu_specs::config_is_valid)."""
    c = _config(kernel, n_sections, **kw)
    return lib().iirx_config_check(ctypes.byref(c), int(n_samples))


def butterworth_sos(order, wn):
    """Butterworth lowpass as b0-normalised float32 sections and the float64 gain
    (scipy.signal.butter(order, wn, output='sos') + ref_generate.py:61-67)."""
    ns = order // 2
    sos = np.zeros((ns, 6), np.float32)
    g = ctypes.c_double()
    check(lib().iirx_butterworth_sos(int(order), float(wn), sos.ctypes.data, ctypes.byref(g)), "butterworth")
    return sos, g.value


def chebyshev_sos(kind, order, db, wn):
    """Chebyshev lowpass as b0-normalised float32 sections and the float64 gain:
    kind 1 = scipy.signal.cheby1(order, db, wn, output='sos') (passband ripple
    db), kind 2 = cheby2 (stopband attenuation db), then normalize_sos."""
    ns = order // 2
    sos = np.zeros((ns, 6), np.float32)
    g = ctypes.c_double()
    check(lib().iirx_chebyshev_sos(int(kind), int(order), float(db), float(wn), sos.ctypes.data, ctypes.byref(g)),
          "chebyshev")
    return sos, g.value


def normalize_sos(sos):
    """General sections (rows b0 b1 b2 a0 a1 a2, float64, any scaling; scipy's
    cheby1/ellip/... output) into the filter contract: (float32 rows with
    b0 == a0 == 1, gain) with gain the factor to apply to the input."""
    s = np.ascontiguousarray(np.asarray(sos, np.float64).reshape(-1, 6))
    out = np.zeros(s.shape, np.float32)
    g = ctypes.c_double()
    check(lib().iirx_normalize_sos(s.ctypes.data, s.shape[0], out.ctypes.data, ctypes.byref(g)), "normalize_sos")
    return out, g.value


def direct_form_taps(sos):
    """(b, a): FIR taps past b0 == 1 and feedback taps (recurrence sign), FP64."""
    s = _sos_array(sos)
    ns = s.shape[0]
    b = np.zeros(2 * ns, np.float64)
    a = np.zeros(2 * ns, np.float64)
    check(lib().iirx_direct_form_taps(s.ctypes.data, ns, b.ctypes.data, a.ctypes.data), "direct_form_taps")
    return b, a


def filter_taps_design(n_sections, cutoff_wn=0.2):
    """The reference's test filter as its generated header holds it
    (ref_generate.py:56-81): a dict with n_sections, plr_order, gain,
    input_fs, input_freq, cutoff_wn, sos (float32 [ns, 6]), plr_b, plr_a."""
    t = _lib.FilterTaps()
    check(lib().iirx_filter_taps_design(int(n_sections), float(cutoff_wn), ctypes.byref(t)), "filter_taps_design")
    return _taps_dict(t)


def write_filter_taps(taps, path):
    """Write a taps dict as the reference's build/filter_taps.hpp."""
    ns, k = int(taps["n_sections"]), int(taps["plr_order"])
    t = _lib.FilterTaps(n_sections=ns, plr_order=k, gain=taps["gain"], input_fs=taps["input_fs"],
                        input_freq=taps["input_freq"], cutoff_wn=taps.get("cutoff_wn", 0.0))
    sos = np.asarray(taps["sos"], np.float32).reshape(-1)
    if sos.size != 6 * ns or len(taps["plr_b"]) != k or len(taps["plr_a"]) != k:
        raise ValueError("taps: array sizes disagree with n_sections / plr_order")
    for i, v in enumerate(sos):
        t.sos[i] = v
    for i in range(k):
        t.plr_b[i] = taps["plr_b"][i]
        t.plr_a[i] = taps["plr_a"][i]
    check(lib().iirx_filter_taps_write(ctypes.byref(t), os.fsencode(path)), "filter_taps_write")


def read_filter_taps(path):
    """Read a build/filter_taps.hpp (the reference's or one written here)."""
    t = _lib.FilterTaps()
    check(lib().iirx_filter_taps_read(os.fsencode(path), ctypes.byref(t)), "filter_taps_read")
    return _taps_dict(t)


def _taps_dict(t):
    ns, k = t.n_sections, t.plr_order
    return {
        "n_sections": ns, "plr_order": k, "gain": t.gain, "input_fs": t.input_fs, "input_freq": t.input_freq,
        "cutoff_wn": t.cutoff_wn,
        "sos": np.array(t.sos[:6 * ns], np.float32).reshape(ns, 6),
        "plr_b": np.array(t.plr_b[:k], np.float64), "plr_a": np.array(t.plr_a[:k], np.float64),
    }


def fill_normal(seed, n):
    """N(0,1) float32 samples from the library's documented generator."""
    out = np.empty(int(n), np.float32)
    check(lib().iirx_fill_normal(int(seed) & (2**64 - 1), int(n), out.ctypes.data), "fill_normal")
    return out


def copy_ceiling(d_src, d_dst, n, mode=0, stream=None):
    """One launch of the library's 16-byte streaming copy of n floats (the
    measured HBM ceiling; mode 0: nontemporal loads and stores, mode 1:
    stores through the L2)."""
    check(lib().iirx_copy_ceiling(_ptr(d_src), _ptr(d_dst), int(n), int(mode), _stream_ptr(stream)), "copy_ceiling")


def plan_tables(kernel, sos, n_samples, **kw):
    """Host copy of the FP64-precomputed tables a plan would upload: (tables, stride)."""
    s = _sos_array(sos)
    c = _config(kernel, s.shape[0], **kw)
    size = ctypes.c_int64()
    stride = ctypes.c_int32()
    check(lib().iirx_plan_tables(ctypes.byref(c), s.ctypes.data, int(n_samples), None, 0,
                                 ctypes.byref(size), ctypes.byref(stride)), "plan_tables")
    out = np.zeros(size.value, np.float32)
    check(lib().iirx_plan_tables(ctypes.byref(c), s.ctypes.data, int(n_samples), out.ctypes.data, size.value,
                                 ctypes.byref(size), ctypes.byref(stride)), "plan_tables")
    return out, stride.value


_KERNEL_NAMES = {v: k for k, v in KERNELS.items()}


def _cfg_dict(c):
    return {"kernel": _KERNEL_NAMES.get(c.kernel, c.kernel), "block_size": c.block_size, "n_blocks": c.n_blocks,
            "unrolled": bool(c.unrolled), "schedule": "ticket" if c.schedule == 1 else "persistent"}


def tune_candidates(sos, n_samples, flags=0):
    """Host-only candidate list (the autotuner's search space) with each
    candidate's status: 0 eligible, 2 no instance for the geometry (or too
    many chunks), 3 CR factors ill-conditioned for this filter."""
    s = _sos_array(sos)
    n = ctypes.c_int32()
    check(lib().iirx_tune_candidates(s.ctypes.data, s.shape[0], int(n_samples), int(flags), None, 0,
                                     ctypes.byref(n)), "tune_candidates")
    arr = (_lib.TuneEntry * n.value)()
    check(lib().iirx_tune_candidates(s.ctypes.data, s.shape[0], int(n_samples), int(flags), arr, n.value,
                                     ctypes.byref(n)), "tune_candidates")
    return [dict(_cfg_dict(e.cfg), status=e.status) for e in arr]


def autotune(sos, n_samples, flags=0, reps=10):
    """Time every eligible candidate on the device; returns (best config dict,
    best ms per launch, table of every candidate with its status and ms)."""
    s = _sos_array(sos)
    n = ctypes.c_int32()
    check(lib().iirx_tune_candidates(s.ctypes.data, s.shape[0], int(n_samples), int(flags), None, 0,
                                     ctypes.byref(n)), "tune_candidates")
    arr = (_lib.TuneEntry * n.value)()
    best = _lib.Config()
    ms = ctypes.c_float()
    check(lib().iirx_autotune(s.ctypes.data, s.shape[0], int(n_samples), int(flags), int(reps), ctypes.byref(best),
                              ctypes.byref(ms), arr, n.value, ctypes.byref(n)), "autotune")
    table = [dict(_cfg_dict(e.cfg), status=e.status, ms=e.ms, rel_err=e.rel_err) for e in arr]
    return _cfg_dict(best), ms.value, table


def _ptr(t):
    if isinstance(t, int):
        return t
    if hasattr(t, "data_ptr"):
        if not t.is_cuda:
            raise ValueError("device tensor required")
        if str(t.dtype) != "torch.float32":
            raise ValueError("float32 tensor required")
        if not t.is_contiguous():
            raise ValueError("contiguous tensor required")
        return t.data_ptr()
    raise TypeError(f"expected a device pointer or a torch tensor, got {type(t)}")


def _stream_ptr(stream):
    if stream is None:
        try:
            import torch
            return torch.cuda.current_stream().cuda_stream
        except Exception:
            return 0
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


class Plan:
    """A runtime plan: one filter, one geometry, one length, one device.

    kernel       'PH' | 'STCR' | 'DTCR' | 'PLR' | 'BS'      (KERNEL; BS has no counterpart)
    sos          (n_sections, 6) b0-normalised sections     (FILTER_SOS)
    n_samples    signal length                              (2^N_SAMPLES_LOG2)
    block_size   32 or 64: a chunk is block_size * n_blocks
                 samples, as in the reference; one wave64
                 holds it (64 lanes x n_blocks at 64, 64 lanes
                 x n_blocks / 2 at 32); PLR: threads per
                 block (0/None = per-order default); BS:
                 tiles per chunk, 16, 32 or 64              (BLOCK_SIZE / PLR_BLOCK_SIZE)
    n_blocks     samples per reference lane (8..128, 16..128
                 at block_size 32); BS: tile length         (N_BLOCKS)
    unrolled     literal-index back substitution            (HANDUNROLLED)
    plr_x, plr_x_max   PLR values per thread                (PLR_X / PLR_X_MAX)
    schedule     'persistent' (default) or 'ticket'         (no counterpart)
    """

    def __init__(self, kernel, sos, n_samples, block_size=None, n_blocks=64, unrolled=False, plr_x=0, plr_x_max=0,
                 plr_taps=None, schedule="persistent"):
        s = _sos_array(sos)
        self.kernel = kernel.upper() if isinstance(kernel, str) else kernel
        if block_size is None:
            block_size = 0 if self.kernel in ("PLR", KERNELS["PLR"]) else 64
        self.n_sections = s.shape[0]
        self.n_samples = int(n_samples)
        self._cfg = _config(kernel, self.n_sections, block_size, n_blocks, unrolled, plr_x, plr_x_max, schedule)
        h = ctypes.c_void_p()
        check(lib().iirx_plan_create(ctypes.byref(self._cfg), s.ctypes.data, self.n_samples, ctypes.byref(h)),
              "plan_create")
        self._h = h
        if plr_taps is not None:
            b, a = (np.ascontiguousarray(np.asarray(v, np.float64)) for v in plr_taps)
            check(lib().iirx_plan_set_plr_taps(self._h, b.ctypes.data, a.ctypes.data, b.size), "set_plr_taps")

    @classmethod
    def autotuned(cls, sos, n_samples, flags=0, reps=10):
        """A plan with the configuration the autotuner measures fastest."""
        best, _, _ = autotune(sos, n_samples, flags, reps)
        return cls(best["kernel"], sos, n_samples, block_size=best["block_size"], n_blocks=best["n_blocks"],
                   unrolled=best["unrolled"], schedule=best["schedule"])

    def execute(self, d_in, d_out, stream=None):
        """One grid-wide launch: d_out = filter(d_in).  Tensors or raw device pointers."""
        if self._h is None:
            raise RuntimeError("plan is closed")
        for t in (d_in, d_out):
            if hasattr(t, "numel") and t.numel() != self.n_samples:
                raise ValueError(f"tensor has {t.numel()} samples, plan expects {self.n_samples}")
        check(lib().iirx_execute(self._h, _ptr(d_in), _ptr(d_out), _stream_ptr(stream)), "execute")

    __call__ = execute

    def state_size(self):
        """Floats in a streaming state: 2 + 2 * n_sections (PH, STCR, DTCR, BS) or
        4 * n_sections (PLR: the order-K direct-form history, K = 2 n_sections)."""
        return 4 * self.n_sections if self.kernel in ("PLR", KERNELS["PLR"]) else 2 + 2 * self.n_sections

    def execute_state(self, d_in, d_out, zi=None, zf=None, stream=None):
        """One launch with the filter state carried in and out (a long signal
        filtered in calls of n_samples).  zi / zf: device float32 tensors (or
        pointers) of state_size() values, oldest first: PH, STCR, DTCR and the
        block-state form x[-2], x[-1] of the input, then y_s[-2], y_s[-1] of
        every section's output; PLR x[-K..-1], then y[-K..-1].  None = zero
        state in / no state out."""
        if self._h is None:
            raise RuntimeError("plan is closed")
        for t in (d_in, d_out):
            if hasattr(t, "numel") and t.numel() != self.n_samples:
                raise ValueError(f"tensor has {t.numel()} samples, plan expects {self.n_samples}")
        for t in (zi, zf):
            if t is not None and hasattr(t, "numel") and t.numel() < self.state_size():
                raise ValueError(f"state needs {self.state_size()} values")
        check(lib().iirx_execute_state(self._h, _ptr(d_in), _ptr(d_out), None if zi is None else _ptr(zi),
                                       None if zf is None else _ptr(zf), _stream_ptr(stream)), "execute_state")

    def device_tables(self):
        """The coefficient tables this plan holds on the device, as float32."""
        n = ctypes.c_int64()
        check(lib().iirx_plan_read_tables(self._h, None, 0, ctypes.byref(n)), "plan_read_tables")
        out = np.zeros(n.value, np.float32)
        check(lib().iirx_plan_read_tables(self._h, out.ctypes.data, n.value, ctypes.byref(n)), "plan_read_tables")
        return out

    def info(self):
        i = _lib.PlanInfo()
        check(lib().iirx_plan_info_get(self._h, ctypes.byref(i)), "plan_info")
        return {f: getattr(i, f) for f, _ in _lib.PlanInfo._fields_}

    def check(self):
        """Synchronise and raise IirxError(IIRX_ERR_TIMEOUT) if a lookback wait
        exceeded the spin limit since the last check."""
        check(lib().iirx_plan_check(self._h), "plan_check")

    def set_spin_limit(self, polls):
        check(lib().iirx_plan_set_spin_limit(self._h, int(polls)), "set_spin_limit")

    def measure(self, inputs, d_out, warmup=1000, batches=20, batch=200, stream=None):
        """measure_kernel_time: sorted per-call batch means in ms (P20 = [batches // 5])."""
        ptrs = (ctypes.c_void_p * len(inputs))(*[_ptr(t) for t in inputs])
        out = np.zeros(batches, np.float32)
        check(lib().iirx_measure(self._h, ptrs, len(inputs), _ptr(d_out), int(warmup), int(batches), int(batch),
                                 _stream_ptr(stream), out.ctypes.data), "measure")
        return out

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().iirx_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Batch:
    """Ragged batch: many channels of uneven length, each with its own order
    and sections, filtered by one persistent launch (no counterpart in the
    reference, which filters one channel per binary).

    offsets      n_channels + 1 CSR offsets into the packed input/output
    sections     list of (n_sections_c, 6) arrays, one per channel, or one
                 (total_sections, 6) array with `n_sections` given
    n_blocks     samples per lane (16, 32 or 64): chunks of 64 * n_blocks
    schedule     'persistent' (default) or 'ticket'
    algo         per-channel algorithm: one of 'auto', 'ph', 'blockstate'
                 for every channel, or a sequence of them (or of the
                 BATCH_ALGOS codes), one per channel
    bs_min_sections  'auto' channels with at least this many sections run
                 the block-state form (0: the library's choice, 8 sections =
                 order 16; 9: none)
    bs_tiles     block-state tiles of 64 samples per chunk (16, 32 or 64)
    """

    def __init__(self, offsets, sections, n_sections=None, n_blocks=32, schedule="persistent", algo="auto",
                 bs_min_sections=0, bs_tiles=32):
        off = np.ascontiguousarray(np.asarray(offsets, np.int64))
        if n_sections is None:
            secs = [_sos_array(s) for s in sections]
            ns = np.array([s.shape[0] for s in secs], np.int32)
            sos = np.ascontiguousarray(np.concatenate(secs, axis=0), np.float32)
        else:
            ns = np.ascontiguousarray(np.asarray(n_sections, np.int32))
            sos = np.ascontiguousarray(np.asarray(sections, np.float32).reshape(-1, 6))
        if off.size != ns.size + 1:
            raise ValueError("offsets must have one entry more than there are channels")
        self.offsets = off
        self.n_sections = ns
        self.n_channels = int(ns.size)
        self.n_samples = int(off[-1])
        self._cfg = _lib.BatchConfig()
        self._cfg.n_blocks = int(n_blocks)
        self._cfg.schedule = _lib.SCHEDULES[schedule] if isinstance(schedule, str) else int(schedule)
        self._cfg.bs_min_sections = int(bs_min_sections)
        self._cfg.bs_tiles = int(bs_tiles)
        if isinstance(algo, str):
            algo = [algo] * self.n_channels
        codes = np.ascontiguousarray(
            [_lib.BATCH_ALGOS[a] if isinstance(a, str) else int(a) for a in algo], np.int32)
        if codes.size != self.n_channels:
            raise ValueError("algo needs one entry per channel")
        h = ctypes.c_void_p()
        check(lib().iirx_batch_create_ex(ctypes.byref(self._cfg), self.n_channels, off.ctypes.data, ns.ctypes.data,
                                         sos.ctypes.data, codes.ctypes.data, ctypes.byref(h)), "batch_create")
        self._h = h

    @classmethod
    def autotuned(cls, offsets, sections, n_sections=None, schedule="persistent", reps=10):
        """A batch ('auto' channels) with the n_blocks, bs_tiles and
        bs_min_sections the batch autotuner measures fastest on this device;
        the measured ms per launch is left in `.tuned_ms`."""
        off = np.ascontiguousarray(np.asarray(offsets, np.int64))
        if n_sections is None:
            secs = [_sos_array(s) for s in sections]
            ns = np.array([s.shape[0] for s in secs], np.int32)
            sos = np.ascontiguousarray(np.concatenate(secs, axis=0), np.float32)
        else:
            ns = np.ascontiguousarray(np.asarray(n_sections, np.int32))
            sos = np.ascontiguousarray(np.asarray(sections, np.float32).reshape(-1, 6))
        if off.size != ns.size + 1:
            raise ValueError("offsets must have one entry more than there are channels")
        cfg = _lib.BatchConfig()
        cfg.n_blocks = 32
        cfg.schedule = _lib.SCHEDULES[schedule] if isinstance(schedule, str) else int(schedule)
        cfg.bs_tiles = 32
        best = _lib.BatchConfig()
        ms = ctypes.c_float()
        check(lib().iirx_batch_autotune(ctypes.byref(cfg), int(ns.size), off.ctypes.data, ns.ctypes.data,
                                        sos.ctypes.data, int(reps), ctypes.byref(best), ctypes.byref(ms)),
              "batch_autotune")
        b = cls(off, sos, n_sections=ns, n_blocks=best.n_blocks, schedule=best.schedule,
                bs_min_sections=best.bs_min_sections, bs_tiles=best.bs_tiles)
        b.tuned_ms = ms.value
        return b

    def _check_len(self, *ts):
        for t in ts:
            if hasattr(t, "numel") and t.numel() < self.n_samples:
                raise ValueError(f"tensor has {t.numel()} samples, the batch needs {self.n_samples}")

    def execute(self, d_in, d_out, stream=None):
        """One persistent launch over every channel."""
        self._check_len(d_in, d_out)
        check(lib().iirx_batch_execute(self._h, _ptr(d_in), _ptr(d_out), _stream_ptr(stream)), "batch_execute")

    __call__ = execute

    def capture(self, d_in, d_out, stream):
        """Capture the launch for (d_in, d_out) into a hipGraph on `stream`
        (a torch.cuda.Stream other than the null stream)."""
        self._check_len(d_in, d_out)
        check(lib().iirx_batch_capture(self._h, _ptr(d_in), _ptr(d_out), _stream_ptr(stream)), "batch_capture")

    def replay(self, stream=None):
        check(lib().iirx_batch_replay(self._h, _stream_ptr(stream)), "batch_replay")

    def check(self):
        check(lib().iirx_batch_check(self._h), "batch_check")

    def device_tables(self):
        """The per-channel tables on the device (channel after channel, 16-byte aligned)."""
        n = ctypes.c_int64()
        check(lib().iirx_batch_read_tables(self._h, None, 0, ctypes.byref(n)), "batch_read_tables")
        out = np.zeros(n.value, np.float32)
        check(lib().iirx_batch_read_tables(self._h, out.ctypes.data, n.value, ctypes.byref(n)), "batch_read_tables")
        return out

    def info(self):
        i = _lib.BatchInfo()
        check(lib().iirx_batch_info_get(self._h, ctypes.byref(i)), "batch_info")
        return {f: getattr(i, f) for f, _ in _lib.BatchInfo._fields_}

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().iirx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
