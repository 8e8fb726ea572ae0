"""Host-side logic of libiirx.so, no GPU: the library loads and exports every
symbol include/iirx.h declares; filter design, direct-form taps and the
synthetic generator; FP64 table precompute against closed forms; geometry
validation (gpu_specs::config_is_valid, test_accuracy.cu:37-41)."""
import ctypes
import os
import re

import numpy as np
import pytest
from scipy import signal

import cuda_recursive_filtering_amd as iirx
from cuda_recursive_filtering_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = np.load(os.path.join(ROOT, "tests", "golden", "scipy_fixtures.npz"), allow_pickle=False)


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "iirx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(iirx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(_lib._SIGS), "ctypes signatures out of sync with include/iirx.h"


def test_version_and_status_strings():
    assert "gfx950" in iirx.version()
    for code in range(7):
        assert _lib.lib().iirx_status_string(code)


@pytest.mark.parametrize("order", [2, 4, 8, 16])
@pytest.mark.parametrize("wn", [0.02, 0.1, 0.2, 0.4])
def test_butterworth_matches_scipy_bit_exact_float32(order, wn):
    ref = FIX[f"sos_o{order}_w{wn}"]
    g_ref = float(np.prod(ref[:, 0]))
    rn = ref.copy()
    for k in range(rn.shape[0]):
        rn[k, 0:3] = ref[k, 0:3] / ref[k, 0]
    sos, g = iirx.butterworth_sos(order, wn)
    assert np.array_equal(sos, rn.astype(np.float32))
    assert abs(g - g_ref) <= 1e-12 * g_ref


@pytest.mark.parametrize("kind,db", [(1, 0.5), (1, 3.0), (2, 40.0), (2, 80.0)])
@pytest.mark.parametrize("order", [2, 4, 8, 16])
@pytest.mark.parametrize("wn", [0.02, 0.1, 0.3])
def test_chebyshev_matches_scipy_bit_exact_float32(kind, db, order, wn):
    """iirx_chebyshev_sos against frozen scipy 1.15.3 cheby1 / cheby2 designs
    (tests/golden/scipy_fixtures.npz): the float32 b0-normalised rows bit for
    bit, in scipy's section order, the gain to FP64 rounding; and the
    designed filter has the nominal response (type 1: ripple db in the
    passband, unit peak; type 2: at least db attenuation past wn)."""
    ref = FIX[f"cheby{kind}_o{order}_d{db}_w{wn}"]
    want, g_want = iirx.normalize_sos(ref)
    sos, g = iirx.chebyshev_sos(kind, order, db, wn)
    assert np.array_equal(sos, want)
    assert abs(g - g_want) <= 1e-12 * abs(g_want)
    full = sos.astype(np.float64)
    full[0, :3] *= g
    w, h = signal.sosfreqz(full, worN=4096)
    mag = np.abs(h)
    if kind == 1:
        pb = mag[w <= 0.98 * np.pi * wn]
        assert np.max(pb) == pytest.approx(1.0, abs=1e-3)
        assert np.min(pb) >= 10 ** (-db / 20) - 1e-3
    else:
        sb = mag[w >= 1.02 * np.pi * wn]
        assert np.max(sb) <= 10 ** (-db / 20) * 1.01


@pytest.mark.parametrize("order", [2, 4, 8, 16])
def test_direct_form_taps_match_scipy_ba(order):
    b, a = FIX[f"ba_b_o{order}_w0.2"], FIX[f"ba_a_o{order}_w0.2"]
    sos, g = iirx.butterworth_sos(order, 0.2)
    bt, at = iirx.direct_form_taps(sos)
    # float32-rounded sections vs the float64 design: agree to float32 level
    assert np.allclose(bt, b[1:] / b[0], rtol=0, atol=1e-6 * np.max(np.abs(b / b[0])))
    assert np.allclose(at, -a[1:], rtol=0, atol=2e-5 * np.max(np.abs(a)))
    # ref_generate.py:76-79: direct-form b0 equals the cascade gain
    assert abs(b[0] - g) <= 1e-9 * g


def test_fill_normal_deterministic_and_standard():
    a = iirx.fill_normal(1, 1 << 18)
    b = iirx.fill_normal(1, 1 << 18)
    c = iirx.fill_normal(2, 1 << 18)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert abs(a.mean()) < 0.01 and abs(a.std() - 1.0) < 0.01
    assert abs(np.mean(a.astype(np.float64) ** 4) - 3.0) < 0.05
    assert np.array_equal(iirx.fill_normal(1, 1001), a[:1001])


def test_ph_tables_match_closed_form():
    """FP64 tables: C^(t+1) and the zero-input responses against numpy matrix
    powers of the companion form (ph_kernels.cuh:55-144, Appendix A.4)."""
    sos, g = iirx.butterworth_sos(4, 0.1)
    N = 32
    tab, stride = iirx.plan_tables("PH", sos, 64 * N * 4, block_size=64, n_blocks=N)
    assert tab.size == stride * 2
    for s in range(2):
        t = tab[s * stride:(s + 1) * stride]
        a1, a2 = -float(sos[s, 4]), -float(sos[s, 5])
        A = np.array([[0.0, 1.0], [a2, a1]])  # (y[n-2], y[n-1]) -> (y[n-1], y[n])
        C = np.linalg.matrix_power(A, N)
        assert np.allclose(t[4:8], C.reshape(-1), rtol=1e-6, atol=1e-7)
        kcfin = 4 + 4 * 6 + 4 + 4
        for l in (0, 1, 31, 63):
            ref = np.linalg.matrix_power(C, l + 1).reshape(-1)
            assert np.allclose(t[kcfin + 4 * l:kcfin + 4 * l + 4], ref, rtol=1e-5, atol=1e-6 * np.abs(C).max())
        # h1[n] = h0[n+1]: response at n to y[-1] = 1
        kh1 = kcfin + 4 * 64
        zi = signal.lfiltic([1.0], [1.0, -a1, -a2], y=[1.0, 0.0])
        h1 = signal.lfilter([1.0], [1.0, -a1, -a2], np.zeros(N), zi=zi)[0]
        assert np.allclose(t[kh1:kh1 + N], h1, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("wn,N,lo,hi", [(0.1, 64, 1, 1), (0.005, 8, 2, 8), (0.0001, 8, 0, 0)])
def test_lookback_depth(wn, N, lo, hi):
    """Depth of the bounded lookback: smallest d with ||(C^64)^d||_inf <= 2^-36
    (0 = unbounded), checked against numpy matrix powers."""
    sos, g = iirx.butterworth_sos(2, wn)
    tab, stride = iirx.plan_tables("PH", sos, 64 * N * 4, block_size=64, n_blocks=N)
    kM = 4 + 4 * 6
    depth = int(tab[kM + 4])
    assert lo <= depth <= hi, depth
    a1, a2 = -float(sos[0, 4]), -float(sos[0, 5])
    M = np.linalg.matrix_power(np.array([[0.0, 1.0], [a2, a1]]), 64 * N)
    norms = [np.abs(np.linalg.matrix_power(M, d)).sum(axis=1).max() for d in range(1, 65)]
    ref = next((d + 1 for d, v in enumerate(norms) if v <= 2.0 ** -36), 0)
    assert depth == ref


@pytest.mark.parametrize("kernel,B,N,n,ok", [
    ("PH", 64, 64, 1 << 20, True),
    ("PH", 32, 32, 1 << 20, True),       # BLOCK_SIZE 32: the reference's 1024-sample chunk
    ("PH", 32, 8, 1 << 20, False),        # 32 x 8 would need 4 samples per lane of a wave64
    ("PH", 48, 64, 1 << 20, False),       # BLOCK_SIZE not 32/64
    ("PH", 64, 48, 1 << 20, False),       # N_BLOCKS not a power of two
    ("PH", 64, 64, (1 << 20) + 64, True),   # a partial last chunk (the reference needs a multiple)
    ("DTCR", 32, 16, 7, True),            # shorter than one chunk
    ("BS", 16, 128, 1 << 20, True),       # block-state: G = 16 tiles of L = 128
    ("BS", 32, 128, 1 << 20, True),
    ("BS", 64, 128, 1 << 20, False),      # G = 64 at L = 128 does not fit the LDS
    ("BS", 16, 256, 1 << 20, False),      # no L = 256
    ("BS", 16, 128, (1 << 20) + 1024, True),  # not a multiple of the 2048-sample chunk: PH ends it
])
def test_config_check(kernel, B, N, n, ok):
    st = iirx.config_check(kernel, 2, n, block_size=B, n_blocks=N)
    assert (st == 0) == ok, st


def _phj_pair_stride(N, L=64):
    kH = 16 * 5 + 16 + 4 + 16 * 16 + 16 * L
    return kH, (kH + 4 * N + 3) // 4 * 4


@pytest.mark.parametrize("order,wn,N", [(4, 0.1, 32), (8, 0.05, 16), (6, 0.3, 64), (16, 0.1, 8), (4, 0.02, 32)])
def test_ph_joint_tables_reproduce_the_cascade(order, wn, N):
    """PH with sections in pairs: the pair tables (joint transition J and its
    powers, the chunk transition, the responses H of the second section to the
    joint lane state) drive a lane-by-lane FP64 recursion over the whole
    signal -- zero-state passes of both sections per lane, y_t += H U, lane
    end state e + J U -- that must equal the FP64 cascade.  A wrong table
    entry, index or sign fails it; order 6 has an odd last section."""
    sos, g = iirx.butterworth_sos(order, wn)
    ns = sos.shape[0]
    L = 64
    n = L * N * 3
    tab, stride = iirx.plan_tables("PHJ", sos, n, block_size=64, n_blocks=N)
    kH, pstride = _phj_pair_stride(N, L)
    base = ns * stride
    assert tab.size == base + (ns // 2) * pstride
    x = (g * iirx.fill_normal(5, n).astype(np.float64)).astype(np.float32)
    y = x.astype(np.float64)
    sec = 0
    while sec < ns:
        if sec + 1 == ns:
            y = signal.sosfilt(sos[sec:sec + 1].astype(np.float64), y)
            sec += 1
            continue
        t = tab[base + (sec // 2) * pstride: base + (sec // 2 + 1) * pstride].astype(np.float64)
        # the joint one-step matrix in FP64 from the sections: state
        # (y_s[n-2], y_s[n-1], y_t[n-2], y_t[n-1]), zero input
        as1, as2 = -float(sos[sec, 4]), -float(sos[sec, 5])
        bt1, bt2 = float(sos[sec + 1, 1]), float(sos[sec + 1, 2])
        at1, at2 = -float(sos[sec + 1, 4]), -float(sos[sec + 1, 5])
        A4 = np.array([[0.0, 1.0, 0.0, 0.0], [as2, as1, 0.0, 0.0], [0.0, 0.0, 0.0, 1.0],
                       [as2 + bt2, as1 + bt1, at2, at1]])
        J64 = np.linalg.matrix_power(A4, N)

        def close(got, P, tol=2e-6):
            # float32 rounding of each entry, on the scale of the matrix
            # (entries far below it, or below float32's range, round away)
            return np.max(np.abs(got.reshape(4, 4) - P)) <= tol * np.max(np.abs(P)) + 1e-37

        for k in range(5):
            assert close(t[16 * k:16 * k + 16], np.linalg.matrix_power(J64, 2 ** k)), k
        assert close(t[80:96], np.linalg.matrix_power(J64, L), 1e-5)
        for j in (0, 7, 15):
            assert close(t[100 + 16 * j:116 + 16 * j], np.linalg.matrix_power(J64, j + 1), 1e-5), j
        for l in (0, 31, 63):
            assert close(t[356 + 16 * l:372 + 16 * l], np.linalg.matrix_power(J64, l + 1), 1e-5), l
        Href = np.array([np.linalg.matrix_power(A4, k + 1)[3] for k in range(N)]).T  # H[m][n]
        assert np.max(np.abs(t[kH:kH + 4 * N] - Href.reshape(-1))) <= 2e-6 * np.max(np.abs(Href))
        # joined only while every table entry stays within 256
        big = max(np.abs(Href).max(), max(np.abs(np.linalg.matrix_power(J64, l + 1)).max() for l in range(L)))
        assert t[97] == (1.0 if big <= 256.0 else 0.0), big
        J = t[0:16].reshape(4, 4)
        H = t[kH:kH + 4 * N].reshape(4, N)
        b1, b2, a1, a2 = (float(v) for v in tab[sec * stride: sec * stride + 4])
        d1, d2, c1, c2 = (float(v) for v in tab[(sec + 1) * stride: (sec + 1) * stride + 4])
        out = np.empty_like(y)
        U = np.zeros(4)
        xh2 = xh1 = 0.0
        for ls in range(0, n, N):
            xs = y[ls:ls + N]
            x2, x1, y2, y1, v2, v1 = xh2, xh1, 0.0, 0.0, 0.0, 0.0
            zs = np.empty(N)
            for k in range(N):
                yc = xs[k] + b1 * x1 + b2 * x2 + a1 * y1 + a2 * y2
                vc = yc + d1 * y1 + d2 * y2 + c1 * v1 + c2 * v2
                zs[k] = vc
                x2, x1, y2, y1, v2, v1 = x1, xs[k], y1, yc, v1, vc
            out[ls:ls + N] = zs + H.T @ U
            U = np.array([y2, y1, zs[N - 2], zs[N - 1]]) + J @ U
            xh2, xh1 = xs[N - 2], xs[N - 1]
        y = out
        sec += 2
    ref = signal.sosfilt(sos.astype(np.float64), x.astype(np.float64))
    assert np.max(np.abs(y - ref)) <= 1e-5 * np.max(np.abs(ref))


def _bs_tables(sos, L, G=64, kernel="BS"):
    """Dequantised block-state tables (layout of kernels.hpp:bs_layout):
    FP16 hi/lo planes (BS) or BF16 hi/mid/lo planes (BS_BF16)."""
    tab, stride = iirx.plan_tables(kernel, sos, G * L * 4, block_size=G, n_blocks=L)
    MB, KQ = L // 16, L // 32
    split = 3 if kernel == "BS_BF16" else 2
    kT = 0
    kG = kT + split * MB * KQ * 256
    kP = kG + split * KQ * 256
    kQ = kP + 6 * 256
    kPhi = kQ + 3 * 256
    kO = kPhi + 256
    kS = kO + MB * 256
    kMc = kS + 4
    if split == 2:
        halves = tab.view(np.uint32).view(np.float16).astype(np.float64)  # 2 halves per word
    else:  # BF16: the upper 16 bits of a float32
        halves = (tab.view(np.uint32).view(np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)

    def block32(word_off):
        """16 x 32 x32-MFMA A operand: lane l, half j = A[l%16][K row 8(l/16)+j],
        K row -> sample 4g + j (j < 4) or 16 + 4g + j - 4 of the slab."""
        h = halves[2 * word_off: 2 * word_off + 512].reshape(64, 8)
        m = np.zeros((16, 32))
        for l in range(64):
            g = l // 16
            for j in range(8):
                m[l % 16, 4 * g + j if j < 4 else 16 + 4 * g + j - 4] = h[l, j]
        return m

    def blockf(word_off):
        """16 x 16 FP32 A operand, K slab m = states 4m..4m+3: word 4 l + m =
        M[l%16][l/16 + 4m]."""
        w = tab[word_off: word_off + 256].astype(np.float64).reshape(64, 4)
        m = np.zeros((16, 16))
        for l in range(64):
            for k in range(4):
                m[l % 16, l // 16 + 4 * k] = w[l, k]
        return m

    T = np.zeros((L, L))
    for mb in range(MB):
        for q in range(KQ):
            off = kT + (mb * KQ + q) * 256
            T[16 * mb:16 * mb + 16, 32 * q:32 * q + 32] = sum(block32(off + hl * MB * KQ * 256) for hl in range(split))
    T *= 2.0 ** (-tab[kS])
    # Gam's operand row i carries state (i >> 2) + 4 (i & 3) (kernels.hpp: bs_state_of_row)
    Gm = np.zeros((16, L))
    rows = [(i >> 2) + 4 * (i & 3) for i in range(16)]
    for q in range(KQ):
        off = kG + q * 256
        Gm[rows, 32 * q:32 * q + 32] = sum(block32(off + hl * KQ * 256) for hl in range(split))
    Gm *= 2.0 ** (-tab[kS + 1])
    Om = np.vstack([blockf(kO + mb * 256) for mb in range(MB)])
    P0 = blockf(kP)
    Phi = blockf(kPhi)
    Mc = tab[kMc:kMc + 256].reshape(16, 16).astype(np.float64)
    return T, Om, Gm, P0, Phi, Mc, int(tab[kMc + 256])


@pytest.mark.parametrize("order", [2, 8, 16])
@pytest.mark.parametrize("kernel,L,G", [("BS", 32, 16), ("BS", 32, 64), ("BS", 64, 16), ("BS", 64, 64),
                                        ("BS", 128, 16), ("BS", 128, 32), ("BS_BF16", 64, 16),
                                        ("BS_BF16", 64, 32), ("BS_BF16", 128, 16), ("BS_BF16", 128, 32)])
def test_blockstate_tables_reproduce_the_cascade(order, kernel, L, G):
    """The block-state tables (FP64 state space of the cascade, FP16 hi/lo or
    BF16 hi/mid/lo operands) define the same filter: running y_t = T x_t +
    O s_t, s_{t+1} = Phi s_t + Gam x_t in float64 from the dequantised
    tables matches the FP64 cascade to the operand precision."""
    from oracle_lib import golden
    sos, g = iirx.butterworth_sos(order, 0.1)
    T, Om, Gm, P0, Phi, Mc, depth = _bs_tables(sos, L, G, kernel)
    assert np.array_equal(P0, Phi)
    # states 0..order-1 in order: the rows past them are empty
    dead = list(range(order, 16))
    assert not Phi[dead, :].any() and not Phi[:, dead].any()
    assert not Gm[dead, :].any() and not Om[:, dead].any()
    assert np.allclose(np.linalg.matrix_power(Phi, G), Mc, rtol=1e-4, atol=1e-30 + 1e-6 * np.abs(Mc).max())
    n = 64 * L * 3
    x = (g * iirx.fill_normal(4, n).astype(np.float64)).astype(np.float32)
    ref = golden(sos, x)
    y = np.zeros(n)
    s = np.zeros(16)
    for t in range(n // L):
        xt = x[t * L:(t + 1) * L].astype(np.float64)
        y[t * L:(t + 1) * L] = T @ xt + Om @ s
        s = Phi @ s + Gm @ xt
    assert np.max(np.abs(y - ref)) <= 1e-5 * np.max(np.abs(ref))
    assert depth >= 1


@pytest.mark.parametrize("wn", [0.02, 0.1, 0.4])
@pytest.mark.parametrize("order", [2, 4, 8, 16])
@pytest.mark.parametrize("L,G", [(32, 64), (64, 64), (64, 32), (128, 16)])
def test_blockstate_scan_levels(order, wn, L, G):
    """The in-chunk scan levels in the scale slots 2 and 3: with `levels`
    levels run, the dropped lags i >= 2^levels weigh sum ||Phi^i||_inf <=
    2^-36 up to the chunk transition Phi^G, and one level fewer would not
    (levels is the smallest such); every level from `fp32_level` on has
    ||Phi^(2^j)||_inf <= 2^-4 and the one before it does not.  Norms from the
    dequantised float32 Phi in numpy; cases within 0.1% of a threshold are
    not held to minimality."""
    sos, _ = iirx.butterworth_sos(order, wn)
    tab, _ = iirx.plan_tables("BS", sos, G * L * 4, block_size=G, n_blocks=L)
    T, Om, Gm, P0, Phi, Mc, depth = _bs_tables(sos, L, G)
    MB, KQ = L // 16, L // 32
    kS = 2 * MB * KQ * 256 + 2 * KQ * 256 + 6 * 256 + 3 * 256 + 256 + MB * 256
    lv32, nlv = int(tab[kS + 2]), int(tab[kS + 3])
    lvG = {16: 4, 32: 5, 64: 6}[G]
    assert 0 <= nlv <= lvG and 0 <= lv32 <= lvG
    nrm = [0.0] + [np.abs(np.linalg.matrix_power(Phi, i)).sum(axis=1).max() for i in range(1, G + 1)]
    tail = lambda j: sum(nrm[1 << j:G + 1])
    bound = 2.0 ** -36
    assert tail(nlv) <= bound * 1.001 or nlv == lvG
    if nlv > 0 and abs(tail(nlv - 1) / bound - 1) > 1e-3:
        assert tail(nlv - 1) > bound
    pn = [np.abs(np.linalg.matrix_power(Phi, 1 << j)).sum(axis=1).max() for j in range(lvG)]
    for j in range(lv32, lvG):
        assert pn[j] <= 2.0 ** -4 * 1.001
    if lv32 > 0 and abs(pn[lv32 - 1] / 2.0 ** -4 - 1) > 1e-3:
        assert pn[lv32 - 1] > 2.0 ** -4
    # at Wn = 0.1 the long lags are dropped: the filters decay well within a chunk
    if wn == 0.1 and G * L >= 4096:
        assert nlv < lvG


@pytest.mark.parametrize("L,G", [(64, 16), (128, 32)])
@pytest.mark.parametrize("order", [8, 16])
def test_blockstate_bf16_split_holds_24_bits(L, G, order):
    """The BF16 three-way split of the Toeplitz block holds every entry
    h[i - j] (the FP64 impulse response of the cascade) to 24 significant
    bits, |hi + mid + lo - h| <= 2^-23 |h|, across the whole dynamic range:
    BF16 keeps FP32's exponent range.  The FP16 pair, with one power-of-two
    scale for the whole block, loses the entries far below the largest (at
    order 16 the impulse response spans more than 10 decades)."""
    from scipy import signal
    sos, g = iirx.butterworth_sos(order, 0.1)
    with np.errstate(invalid="ignore"):
        T3, _, _, _, _, _, _ = _bs_tables(sos, L, G, "BS_BF16")
        T2, _, _, _, _, _, _ = _bs_tables(sos, L, G, "BS")
    imp = np.zeros(L)
    imp[0] = 1.0
    h = signal.sosfilt(sos.astype(np.float64), imp)
    T = np.zeros((L, L))
    for i in range(L):
        T[i, :i + 1] = h[i::-1]
    nz = T != 0
    e3 = np.abs(T3 - T)[nz] / np.abs(T)[nz]
    e2 = np.abs(T2 - T)[nz] / np.abs(T)[nz]
    assert np.max(e3) <= 2.0 ** -23, np.max(e3)
    if order == 16:
        assert np.max(e2) > 1e3 * np.max(e3)
