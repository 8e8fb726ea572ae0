"""The oracle's float32 restatements against the reference's own float32
emulators (tests/golden/emulator_fixtures.npz, written from
tools/stcr_emulator.py, tools/dtcr_emulator.py and tools/plr_emulator.py by
tests/golden/make_emulator_fixtures.py), and the sequential float32 cascade
against closed forms and the FP64 golden.

Two builds of one oracle source are checked:
  * liboracle_dr.so rounds every multiply-add twice, as the emulators'
    fma(a, b, c) = F32(F32(a) * F32(b) + F32(c)) does (stcr_emulator.py:50,
    dtcr_emulator.py:26).  STCR, DTCR and the direct form must then equal the
    emulators bit for bit: any difference in operation order, index, sign or
    coefficient shows.
  * liboracle.so fuses them (fmaf, what the device executes).  It must agree
    with the emulators within 16 float32 ulps of max|ref|: this is the
    "agreement with the float32 emulators within a stated ULP-scaled
    tolerance" of the parity bar, and the GPU tests compare the HIP kernels
    with this build.
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from workloads import ulp_scaled

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "emulator_fixtures.npz")
ORDERS = (2, 4, 8, 16)
GEOMS = ((32, 32), (32, 64), (32, 128), (64, 64))
FUSED_ULPS = 16.0


@pytest.fixture(scope="module")
def fx():
    return np.load(FIX)


def cos_input(n, gain=None):
    """The emulators' own input: cos(2 pi 120 i / 1000) (stcr_emulator.py:199)."""
    i = np.arange(n, dtype=np.float64)
    x = np.cos(2.0 * np.pi * 120.0 * i / 1000.0)
    return (x if gain is None else gain * x).astype(np.float32)


def gained(x, g):
    """stcr() applies the gain itself: f32(gain * f64(x)) (stcr_emulator.py:116)."""
    return (np.float64(g) * x.astype(np.float64)).astype(np.float32)


@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("B,N", GEOMS)
def test_stcr_restatement_equals_emulator(fx, order, B, N):
    sos, g = fx[f"sos_o{order}"], float(fx[f"gain_o{order}"])
    x = gained(cos_input(3 * B * N), g)
    emu = fx[f"stcr_o{order}_{B}x{N}"]
    got = O.stcr(sos, x, B, N, double_round=True)
    assert np.array_equal(got, emu), np.max(np.abs(got - emu))
    # the unrolled (literal-index) form is the same arithmetic
    assert np.array_equal(O.stcr(sos, x, B, N, unrolled=True, double_round=True), emu)
    ref = O.golden(sos, x)
    assert ulp_scaled(O.stcr(sos, x, B, N), emu, ref) <= FUSED_ULPS


@pytest.mark.parametrize("order", ORDERS)
def test_stcr_hand_unrolled_emulator(fx, order):
    """stcr(..., 'hard32') (stcr_emulator.py:198-201) is bit-identical to the
    loop form in the emulator, and so to the restatement."""
    sos, g = fx[f"sos_o{order}"], float(fx[f"gain_o{order}"])
    x = gained(cos_input(3 * 32 * 32), g)
    emu = fx[f"stcr_hard32_o{order}"]
    assert np.array_equal(emu, fx[f"stcr_o{order}_32x32"])
    assert np.array_equal(O.stcr(sos, x, 32, 32, unrolled=True, double_round=True), emu)


@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("B,N", GEOMS)
def test_dtcr_restatement_equals_emulator(fx, order, B, N):
    sos, g = fx[f"sos_o{order}"], float(fx[f"gain_o{order}"])
    x = cos_input(3 * B * N, g)
    emu = fx[f"dtcr_o{order}_{B}x{N}"]
    got = O.dtcr(sos, x, B, N, double_round=True)
    assert np.array_equal(got, emu), np.max(np.abs(got - emu))
    ref = O.golden(sos, x)
    assert ulp_scaled(O.dtcr(sos, x, B, N), emu, ref) <= FUSED_ULPS


@pytest.mark.parametrize("B", [512, 1024])
@pytest.mark.parametrize("X", [1, 4, 9, 10])
def test_plr_restatement_vs_emulator_part_a(fx, B, X):
    """Part A (plr_emulator.py:29-160, PLR_2 with a partial tail chunk).  The
    restatement follows the kernel's cross-warp stages, which add the two
    factor products to the value one after the other (plr_kernels.cuh:
    656-659); the emulator sums the two products first and adds the sum
    (plr_emulator.py:113-114).  Everything else is the same operation order,
    so the double-rounded build differs from the emulator by at most one
    rounding per cross-warp stage: within 2 ulps of max|ref|, and exactly
    equal where the block has no cross-warp stage to round differently."""
    bt, at = fx["plrB_bt_o2"], fx["plrB_at_o2"]
    x = fx[f"plrA_x_{B}_{X}"]
    emu = fx[f"plrA_{B}_{X}"]
    got = O.plr(bt, at, x, B, X, double_round=True)
    ref = O.golden(fx["sos_o2"], x)
    assert ulp_scaled(got, emu, ref) <= 2.0
    # the first 32 lanes of every segment of the first chunk see no cross-warp
    # stage and no carry: bit for bit
    assert np.array_equal(got[:32], emu[:32])
    assert ulp_scaled(O.plr(bt, at, x, B, X), emu, ref) <= FUSED_ULPS
    assert O.gate(emu, ref)[0] and O.gate(got, ref)[0]


@pytest.mark.parametrize("order", ORDERS)
def test_direct_form_equals_emulator_part_b(fx, order):
    """Part B (plr_emulator.py:185-202): the sequential float32 direct form
    is bit-identical to the emulator's (both unfused), and its gate verdict
    over 2^16 samples is the emulator's: PASS at orders 2 and 4, FAIL at 8
    (1.207e-4 > 1e-4), non-finite at 16."""
    bt, at, g = fx[f"plrB_bt_o{order}"], fx[f"plrB_at_o{order}"], float(fx[f"plrB_gain_o{order}"])
    x = cos_input(1 << 16, g)
    with np.errstate(all="ignore"):
        y = O.direct_form_f32(bt, at, x)
    head = fx[f"plrB_head_o{order}"]
    assert np.array_equal(y[:head.size], head, equal_nan=True)
    finite = bool(np.all(np.isfinite(y)))
    assert finite == bool(fx[f"plrB_finite_o{order}"])
    sos, _ = fx[f"sos_o{order}"], None
    if finite:
        ref = O.golden(sos, x)
        ok, err, mref, rel = O.gate(y, ref)
        assert rel == pytest.approx(float(fx[f"plrB_rel_o{order}"]), rel=1e-9)
        assert ok == (order in (2, 4))
    else:
        assert order == 16


@pytest.mark.parametrize("B,N", [(32, 32), (64, 64)])
def test_cpu_plumbing_workload(fx, B, N):
    """The CPU plumbing workload: order-4 Butterworth at Wn = 0.1, 2^16 N(0,1)
    samples from seed 0 of the library's generator, STCR against the FP64
    golden.  The generator reproduces the fixture's input, the emulator's
    output passes the gate, the restatement (double-rounded) equals it bit
    for bit and the fused restatement is within 16 ulps of it and passes."""
    import cuda_recursive_filtering_amd as iirx
    x = iirx.fill_normal(0, 1 << 16)
    assert np.array_equal(x, fx["cfg1_x"])
    sos, g = iirx.butterworth_sos(4, 0.1)
    assert np.array_equal(sos, fx["cfg1_sos"]) and g == pytest.approx(float(fx["cfg1_gain"]), rel=1e-15)
    xg = gained(x, g)
    ref = O.golden(sos, xg)
    emu = fx[f"cfg1_stcr_{B}x{N}"]
    ok, err, mref, rel = O.gate(emu, ref)
    assert ok and rel == pytest.approx(float(fx[f"cfg1_rel_{B}x{N}"]), rel=1e-6)
    assert np.array_equal(O.stcr(sos, xg, B, N, double_round=True), emu)
    y = O.stcr(sos, xg, B, N)
    assert O.gate(y, ref)[0]
    assert ulp_scaled(y, emu, ref) <= FUSED_ULPS


# ---------------------------------------------------------------------------
# orc_cascade_f32: the sequential float32 cascade (direct form I per section,
# fused), used as the accuracy floor of a plain float32 recurrence
# ---------------------------------------------------------------------------
def test_cascade_f32_first_order_closed_form():
    """b1 = b2 = a2 = 0, a1 = -1/2: y[n] = sum_k 2^-k x[n-k]; for a unit
    impulse every value is a power of two, exact in float32."""
    sos = np.array([[1, 0, 0, 1, -0.5, 0]], np.float32)
    x = np.zeros(40, np.float32)
    x[3] = 1.0
    y = O.cascade_f32(sos, x)
    want = np.zeros(40)
    want[3:] = 0.5 ** np.arange(37)
    assert np.array_equal(y, want.astype(np.float32))


def test_cascade_f32_fir_is_convolution():
    """a1 = a2 = 0: each section is the FIR 1 + b1 z^-1 + b2 z^-2; with small
    integer taps and inputs every value is an exact float32 integer."""
    sos = np.array([[1, 2, 1, 1, 0, 0], [1, -3, 2, 1, 0, 0]], np.float32)
    rng = np.random.default_rng(5)
    x = rng.integers(-8, 9, 200).astype(np.float32)
    want = np.convolve(np.convolve(x.astype(np.float64), [1, 2, 1]), [1, -3, 2])[:200]
    assert np.array_equal(O.cascade_f32(sos, x), want.astype(np.float32))


def test_cascade_f32_resonator_impulse_response():
    """One section with poles r e^{+-i th}: h[n] = r^n sin((n+1) th) / sin th
    (a1 = -2 r cos th, a2 = r^2, scipy sign), within float32 rounding."""
    r, th = 0.97, 0.3
    sos = np.array([[1, 0, 0, 1, -2 * r * np.cos(th), r * r]], np.float32)
    a1, a2 = float(sos[0, 4]), float(sos[0, 5])
    # exact poles of the float32-rounded coefficients
    rr = np.sqrt(a2)
    tt = np.arccos(-a1 / (2 * rr))
    n = np.arange(300)
    x = np.zeros(300, np.float32)
    x[0] = 1.0
    want = rr ** n * np.sin((n + 1) * tt) / np.sin(tt)
    y = O.cascade_f32(sos, x).astype(np.float64)
    assert np.max(np.abs(y - want)) <= 64 * np.finfo(np.float32).eps * np.max(np.abs(want))


@pytest.mark.parametrize("order", [2, 4, 8])
def test_cascade_f32_against_golden_and_lfilter(order):
    """Well-conditioned Butterworth (Wn = 0.2) on N(0,1): the float32 cascade
    is within 2e-6 of max|ref| of the FP64 cascade (the same filter in FP64,
    scipy's lfilter section by section agreeing with the golden to 1e-12).
    A 1e-5 relative error in one coefficient or a dropped section fails
    this bound."""
    from scipy import signal
    import cuda_recursive_filtering_amd as iirx
    sos, g = iirx.butterworth_sos(order, 0.2)
    x = (g * iirx.fill_normal(7, 1 << 15).astype(np.float64)).astype(np.float32)
    ref = O.golden(sos, x)
    u = x.astype(np.float64)
    for s in sos.astype(np.float64):
        u = signal.lfilter(s[:3], s[3:], u)
    assert np.max(np.abs(u - ref)) <= 1e-12 * np.max(np.abs(ref))
    y = O.cascade_f32(sos, x)
    ok, err, mref, rel = O.gate(y, ref)
    assert rel <= 2e-6, rel
    bad = sos.copy()
    bad[-1, 4] *= 1 + 1e-5
    assert O.gate(O.cascade_f32(bad, x), ref)[3] > 2e-6
    assert O.gate(O.cascade_f32(sos[:-1], x), ref)[3] > 2e-6


# ---------------------------------------------------------------------------
# PLR at order 8 (Butterworth Wn = 0.2, the reference's workload): where the
# direct form's error comes from
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("B,X", [(32, 1), (64, 8), (128, 10), (512, 10)])
def test_plr_order8_fails_the_gate_in_float32(fx, B, X):
    """The reference's own Part B emulator fails order 8 (rel 1.207e-4 >
    1e-4, fixture), and the float32 merge tree is far worse at every block
    geometry (128 x 10 is the reference's RTX 3060 geometry, plr_kernels.cuh:
    70-84 with PLR_X_MAX = 10), because the n-nacci factors of this direct
    form reach 284 with alternating signs: each merge adds terms up to 284x
    the result, whose float32 rounding alone is ~1e-4 of the signal per level.
    The same merge in FP64 with the same float32 taps lands at the taps' own
    rounding error (8.3e-5), inside the gate."""
    from scipy import signal
    bt, at, g = fx["plrB_bt_o8"], fx["plrB_at_o8"], float(fx["plrB_gain_o8"])
    sos = fx["sos_o8"]
    n = 1 << 14
    x = cos_input(n, g)
    ref = O.golden(sos, x)
    assert not O.gate(O.direct_form_f32(bt, at, x), ref)[0]
    b32 = bt.astype(np.float32).astype(np.float64)
    a32 = at.astype(np.float32).astype(np.float64)
    y64 = signal.lfilter(np.r_[1.0, b32], np.r_[1.0, -a32], x.astype(np.float64))
    ok64, _, _, rel64 = O.gate(y64, ref)
    assert ok64 and rel64 > 5e-5, rel64   # the float32 taps alone: just inside
    ok, err, mref, rel = O.gate(O.plr(bt, at, x, B, X), ref)
    assert not ok and rel > 1e-3, rel
    # the factors' size: rows of companion-matrix powers, max |.| = 284
    k = 8
    v = np.zeros((k, 64 + k))
    for j in range(k):
        v[j, j] = 1.0
        for i in range(k, 64 + k):
            v[j, i] = np.dot(a32, v[j, i - 1::-1][:k])
    assert 250 < np.max(np.abs(v[:, k:])) < 300
