"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

* float64 golden vs scipy.signal.sosfilt (frozen fixtures and live scipy),
  and vs closed forms: first-order scan, pure FIR, single biquad = lfilter,
  impulse response from the poles, linearity.
* float32 restatements of PH/STCR/DTCR/PLR vs the golden on the reference's
  own workload (3 chunks per geometry, as the emulators run it), against the
  reference's published emulator errors (README.md:117-121) and its gate
  (test_accuracy.cu:178-190).
* planted mistakes in the oracle's inputs must fail the gate.
"""
import os

import numpy as np
import pytest
from scipy import signal

import cuda_recursive_filtering_amd as iirx
import oracle_lib as O
from workloads import reference_workload

FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "scipy_fixtures.npz"), allow_pickle=False)
GEOMS = [(32, 32), (32, 64), (32, 128), (64, 64)]
# README.md:117-121: stcr_emulator.py rel. error per order (same at every geometry up to 10%)
README_STCR = {2: 2.3e-7, 4: 8.3e-7, 8: 3.9e-7, 16: 4.5e-7}


def norm(sos):
    g = float(np.prod(sos[:, 0]))
    s = sos.copy()
    for k in range(s.shape[0]):
        s[k, 0:3] = sos[k, 0:3] / sos[k, 0]
    return s.astype(np.float32), g


@pytest.mark.parametrize("order", [2, 4, 8, 16])
def test_golden_matches_frozen_sosfilt(order):
    x = FIX[f"x_ref_o{order}"]
    y = FIX[f"y_ref_o{order}"]
    sos32, _ = norm(FIX[f"sos_o{order}_w0.2"])
    got = O.golden(sos32, x)
    assert np.max(np.abs(got - y)) <= 1e-13 * np.max(np.abs(y))


@pytest.mark.parametrize("wn", [0.02, 0.1, 0.4])
def test_golden_matches_live_sosfilt_unnormalised(wn):
    sos = signal.butter(8, wn, output="sos")
    x = np.random.default_rng(0).standard_normal(5000)
    ref = signal.sosfilt(sos, x)
    got = O.sosfilt_f64(sos, x)
    assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_golden_first_order_closed_form():
    # b1 = b2 = a2 = 0: y[n] = sum_k c^k x[n-k]
    c = 0.75
    sos = np.array([[1.0, 0.0, 0.0, 1.0, -c, 0.0]])
    x = np.random.default_rng(1).standard_normal(200)
    y = O.sosfilt_f64(sos, x)
    ref = np.array([sum(c ** k * x[n - k] for k in range(n + 1)) for n in range(len(x))])
    assert np.max(np.abs(y - ref)) < 1e-12


def test_golden_pure_fir_is_convolution():
    sos = np.array([[1.0, -0.3, 0.7, 1.0, 0.0, 0.0], [1.0, 2.0, 1.0, 1.0, 0.0, 0.0]])
    x = np.random.default_rng(2).standard_normal(300)
    ref = np.convolve(np.convolve(x, [1.0, -0.3, 0.7]), [1.0, 2.0, 1.0])[: len(x)]
    assert np.max(np.abs(O.sosfilt_f64(sos, x) - ref)) < 1e-12


def test_golden_impulse_response_from_poles():
    # one resonant biquad: h[n] = r^n sin((n+1) w) / sin(w) for 1/(1 - 2r cos w z^-1 + r^2 z^-2)
    r, w = 0.97, 0.3
    sos = np.array([[1.0, 0.0, 0.0, 1.0, -2 * r * np.cos(w), r * r]])
    x = np.zeros(400)
    x[0] = 1.0
    n = np.arange(400)
    ref = r ** n * np.sin((n + 1) * w) / np.sin(w)
    assert np.max(np.abs(O.sosfilt_f64(sos, x) - ref)) < 1e-12


def test_golden_linearity_and_shift():
    sos, _ = norm(FIX["sos_o8_w0.1"])
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal(1000), rng.standard_normal(1000)
    ya, yb = O.sosfilt_f64(sos, a), O.sosfilt_f64(sos, b)
    yab = O.sosfilt_f64(sos, 2.0 * a - 3.0 * b)
    assert np.max(np.abs(yab - (2.0 * ya - 3.0 * yb))) < 1e-9 * np.max(np.abs(yab))
    sh = np.concatenate([np.zeros(7), a])[:1000]
    assert np.max(np.abs(O.sosfilt_f64(sos, sh)[7:] - ya[:993])) < 1e-12 * np.max(np.abs(ya))


@pytest.mark.parametrize("alg", ["ph", "stcr", "dtcr"])
@pytest.mark.parametrize("order", [2, 4, 8, 16])
@pytest.mark.parametrize("B,N", GEOMS)
def test_cascade_restatements_pass_reference_gate(alg, order, B, N):
    """Reference workload, 3 chunks (stcr_emulator.py:210-217): every cascade
    restatement passes the 1e-4 gate, with the error magnitude the reference
    publishes for its float32 emulator (README.md:117-121; within a factor 3,
    since the emulator double-rounds its fma)."""
    sos, g, x = reference_workload(order, 3 * B * N)
    ref = O.golden(sos, x)
    y = {"ph": O.ph, "stcr": O.stcr, "dtcr": O.dtcr}[alg](sos, x, B, N)
    ok, err, mref, rel = O.gate(y, ref)
    assert ok, (alg, order, B, N, rel)
    assert rel < 3.0 * README_STCR[order], rel
    assert rel > 1e-9  # really float32 arithmetic


def test_stcr_unrolled_is_loop():
    # stcr_emulator.py:205-208: the literal-index form is the same operation sequence
    sos, g, x = reference_workload(8, 3 * 32 * 32)
    assert np.array_equal(O.stcr(sos, x, 32, 32, False), O.stcr(sos, x, 32, 32, True))


@pytest.mark.parametrize("alg", ["ph", "stcr", "dtcr"])
def test_planted_mistake_fails_gate(alg):
    """A wrong sign on one coefficient, or a dropped section, must fail the
    gate: the gate discriminates the mistakes a restatement could make."""
    sos, g, x = reference_workload(4, 3 * 32 * 64)
    ref = O.golden(sos, x)
    f = {"ph": O.ph, "stcr": O.stcr, "dtcr": O.dtcr}[alg]
    bad = sos.copy()
    bad[1, 4] = bad[1, 4] * 1.01
    assert not O.gate(f(bad, x, 32, 64), ref)[0]
    assert not O.gate(f(sos[:1].copy(), x, 32, 64), ref)[0]


def test_plr_order2_tree_matches_emulator_checks():
    """plr_emulator.py Part A: PLR_2 at B in {512, 1024}, X in {1,4,9,10}, with
    a partial tail chunk, passes the gate."""
    b, a = FIX["ba_b_o2_w0.2"], FIX["ba_a_o2_w0.2"]
    bt, at = b[1:] / b[0], -a[1:]
    for B in (512, 1024):
        for X in (1, 4, 9, 10):
            n = 3 * B * X + B // 2
            sos, g, x = reference_workload(2, n)
            ref = O.golden(sos, x)
            ok, _, _, rel = O.gate(O.plr(bt, at, x, B, X), ref)
            assert ok and rel < 1e-6, (B, X, rel)


def test_plr_order4_passes_and_is_geometry_invariant():
    b, a = FIX["ba_b_o4_w0.2"], FIX["ba_a_o4_w0.2"]
    bt, at = b[1:] / b[0], -a[1:]
    sos, g, x = reference_workload(4, 3 * 256 * 17 + 100)
    ref = O.golden(sos, x)
    for B, X in ((256, 17), (128, 8), (64, 3)):
        assert O.gate(O.plr(bt, at, x, B, X), ref)[0]


def test_direct_form_part_b_verdicts():
    """plr_emulator.py Part B: sequential float32 direct form at n = 2^16:
    order 2 must pass; order 16 produces non-finite output (the reference's
    published order-16 failure, README.md:39)."""
    n = 1 << 16
    verdicts = {}
    for order in (2, 4, 8, 16):
        b, a = FIX[f"ba_b_o{order}_w0.2"], FIX[f"ba_a_o{order}_w0.2"]
        sos, g, x = reference_workload(order, n)
        ref = O.golden(sos, x)
        with np.errstate(all="ignore"):
            y = O.direct_form_f32(b[1:] / b[0], -a[1:], x)
        verdicts[order] = O.gate(y, ref)
    assert verdicts[2][0] and verdicts[4][0]
    assert not np.isfinite(verdicts[16][1])


def test_plr_order16_float32_taps_are_unstable():
    """Why PLR fails at high order whatever its carry precision: the direct
    form's feedback taps, rounded to float32 as the reference casts
    PLR_FILTER_A (plr_kernels.cuh:423-427), define an unstable filter from
    order 12 on at Wn = 0.1 (a pole outside the unit circle), while the FP64
    taps and the float32 SOS cascade are stable.  Pole radii from numpy's
    companion-matrix roots."""
    radius = {}
    for order in (4, 8, 12, 16):
        sos, _ = iirx.butterworth_sos(order, 0.1)
        a = np.array([1.0])
        for s in sos.astype(np.float64):
            a = np.polymul(a, s[3:])
        a32 = a.astype(np.float32).astype(np.float64)
        radius[order] = (np.abs(np.roots(a)).max(), np.abs(np.roots(a32)).max(),
                         max(np.abs(np.roots(s[3:].astype(np.float64))).max() for s in sos))
    for order, (r64, r32, rsos) in radius.items():
        assert r64 < 1.0 and rsos < 1.0, order
        assert (r32 > 1.0) == (order >= 12), (order, r32)
    assert radius[16][1] > 1.4


@pytest.mark.parametrize("order", [4, 8, 16])
def test_plr_restatement_exact_on_well_conditioned_direct_form(order):
    """The merge-tree rule is exact algebra: on a well-conditioned direct form
    of order k (poles at radius 0.5, so float32 rounding stays small) the
    restatement must equal lfilter of the same float32 taps to float32
    accuracy at every geometry; a dropped or misplaced factor term is O(1)."""
    ang = np.linspace(0.2, 2.8, order // 2)
    poles = 0.5 * np.exp(1j * ang)
    a = np.real(np.poly(np.concatenate([poles, poles.conj()])))
    b = np.real(np.poly(-0.9 * np.ones(order)))
    bt, at = b[1:], -a[1:]
    x = np.random.default_rng(5).standard_normal(4096).astype(np.float32)
    b32 = np.concatenate(([1.0], bt.astype(np.float32).astype(np.float64)))
    a32 = np.concatenate(([1.0], -at.astype(np.float32).astype(np.float64)))
    ref = signal.lfilter(b32, a32, x.astype(np.float64))
    for B, X in ((64, 1), (64, 3), (128, 5), (512, 2)):
        y = O.plr(bt, at, x, B, X)
        assert np.max(np.abs(y - ref)) < 1e-5 * np.max(np.abs(ref)), (B, X)


@pytest.mark.parametrize("order", [2, 8, 16])
@pytest.mark.parametrize("B,N", [(32, 64), (64, 32)])
def test_ph_multithreaded_restatement_is_identical(order, B, N):
    """orc_ph_f32_mt (the all-cores CPU baseline) runs the same float
    operations as orc_ph_f32 with the chunks split over threads: equal bit
    for bit, for any thread count."""
    sos, g = iirx.butterworth_sos(order, 0.1)
    x = (g * iirx.fill_normal(3, 37 * B * N).astype(np.float64)).astype(np.float32)
    want = O.ph(sos, x, B, N)
    for threads in (1, 3, 8):
        assert np.array_equal(O.ph_mt(sos, x, B, N, threads), want)
