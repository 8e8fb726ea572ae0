"""Synthetic inputs of the reference's workload and of the measured configs.

Conventions of tools/ref_generate.py:56-99: sections normalised to b0 == 1,
the removed cascade gain applied once to the float64 input, then cast to
float32; the golden output is float64 filtering of the float32 quantities.
"""
import numpy as np

import cuda_recursive_filtering_amd as iirx


def reference_workload(order, n, wn=0.2):
    """The reference's own workload: Butterworth lowpass Wn=0.2 (default),
    input f32(gain * cos(2 pi 120 i / 1000)) (ref_generate.py:31-33, 84-86)."""
    sos, g = iirx.butterworth_sos(order, wn)
    i = np.arange(n, dtype=np.float64)
    x = (g * np.cos(2.0 * np.pi * 120.0 * i / 1000.0)).astype(np.float32)
    return sos, g, x


def normal_workload(order, n, seed, wn=0.1, sine=0.0):
    """N(0,1) samples of the library's generator (plus an optional sinusoid at
    0.05 fs of amplitude `sine`), scaled by the cascade gain, float32."""
    sos, g = iirx.butterworth_sos(order, wn)
    x = iirx.fill_normal(seed, n).astype(np.float64)
    if sine:
        x = x + sine * np.sin(2.0 * np.pi * 0.05 * np.arange(n, dtype=np.float64))
    return sos, g, (g * x).astype(np.float32)


def ulp_scaled(a, b, ref):
    """max|a-b| in units of float32 eps * max|ref|."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    m = float(np.max(np.abs(ref)))
    return float(np.max(np.abs(a - b))) / (np.finfo(np.float32).eps * m)


def ragged_channels(n_channels, seed, lo_log2=10, hi_log2=20, orders=(2, 4, 8, 16), wn=(0.02, 0.45)):
    """The ragged-batch workload's channel set: lengths log-uniform in
    [2^lo, 2^hi], order uniform over `orders`, Butterworth Wn uniform in `wn`
    (numpy default_rng(seed)).  Returns (offsets, sections, gains)."""
    rng = np.random.default_rng(seed)
    lens = np.floor(2.0 ** rng.uniform(lo_log2, hi_log2, n_channels)).astype(np.int64)
    ords = rng.choice(np.asarray(orders), n_channels)
    wns = rng.uniform(wn[0], wn[1], n_channels)
    offsets = np.zeros(n_channels + 1, np.int64)
    offsets[1:] = np.cumsum(lens)
    sections, gains = [], []
    for o, w in zip(ords, wns):
        s, g = iirx.butterworth_sos(int(o), float(w))
        sections.append(s)
        gains.append(g)
    return offsets, sections, np.asarray(gains)


def ragged_input(offsets, gains, seed):
    """Packed float32 input: channel c is gain_c * N(0,1) (generator seed + c)."""
    x = np.empty(int(offsets[-1]), np.float32)
    for c in range(len(gains)):
        a, b = int(offsets[c]), int(offsets[c + 1])
        x[a:b] = (gains[c] * iirx.fill_normal(seed + c, b - a).astype(np.float64)).astype(np.float32)
    return x
