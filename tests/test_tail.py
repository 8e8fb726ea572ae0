"""Cascades (PH, STCR, DTCR) and the block-state form at lengths that are
no multiple of the chunk.  The reference requires a multiple
(test_accuracy.cu:37-38: grid = N / (BLOCK_SIZE * N_BLOCKS)); here the last
chunk is partial and runs in a one-workgroup launch of its own after the
others, from the state its predecessor recorded.  Checked against the FP64 cascade with the
reference's gate; guard samples around the buffers check that nothing past
the signal is read into the result or written; the end state is checked
against the FP64 direct-form histories of every section wherever the
signal ends in the chunk; and uneven streaming calls against one call."""
import numpy as np
import pytest
from scipy import signal

import cuda_recursive_filtering_amd as iirx
import oracle_lib as O
from workloads import normal_workload

CASES = [("PH", (64, 32), False), ("PHJ", (64, 32), False), ("PHH", (64, 32), False), ("PHJ", (32, 64), False), ("STCR", (64, 64), False),
         ("STCR", (32, 32), True),
         ("DTCR", (64, 64), False), ("DTCR", (32, 32), True)]


def _run(device, kernel, geom, unrolled, sos, x, schedule="persistent", zi=None, want_zf=False):
    """One call on x with guard samples around input and output: the input
    guards are NaN (read into the result they would show), the output guards
    a sentinel that must survive."""
    import torch
    n, gd = x.size, 64
    xin = torch.full((n + 2 * gd,), float("nan"), dtype=torch.float32, device=device)
    xin[gd:gd + n] = torch.from_numpy(x).to(device)
    out = torch.full((n + 2 * gd,), 12345.0, dtype=torch.float32, device=device)
    ns = sos.shape[0]
    zf = torch.empty(2 + 2 * ns, dtype=torch.float32, device=device) if want_zf else None
    zit = torch.from_numpy(np.asarray(zi, np.float32)).to(device) if zi is not None else None
    with iirx.Plan(kernel, sos, n, block_size=geom[0], n_blocks=geom[1], unrolled=unrolled, schedule=schedule) as p:
        p.execute_state(xin[gd:gd + n], out[gd:gd + n], zi=zit, zf=zf)
        torch.cuda.synchronize()
        p.check()
    o = out.cpu().numpy()
    assert np.all(o[:gd] == 12345.0) and np.all(o[gd + n:] == 12345.0), "write outside the signal"
    return o[gd:gd + n], (zf.cpu().numpy() if want_zf else None)


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,geom,unrolled", CASES)
@pytest.mark.parametrize("r", [1, 2, 3, 517, -1])
def test_cascade_any_length(device, kernel, geom, unrolled, r):
    """n = 37 chunks + r samples (r = -1: one sample short of 38 chunks)."""
    ch = geom[0] * geom[1]
    n = 37 * ch + (r if r > 0 else ch - 1)
    print(f"case tail {kernel} {geom} unrolled={unrolled} n={n}", flush=True)
    sos, g, x = normal_workload(8, n, seed=41 + r % 7, wn=0.1)
    y, _ = _run(device, kernel, geom, unrolled, sos, x)
    ok, err, mref, rel = O.gate(y, O.golden(sos, x))
    assert ok, rel


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,geom,unrolled", CASES)
@pytest.mark.parametrize("schedule", ["persistent", "ticket"])
def test_cascade_shorter_than_one_chunk(device, kernel, geom, unrolled, schedule):
    """The whole signal inside one partial chunk (no main launch), and a
    signal of one chunk plus one sample."""
    ch = geom[0] * geom[1]
    for n in (ch - 5, ch + 1, 3):
        print(f"case tail short {kernel} {geom} {schedule} n={n}", flush=True)
        sos, g, x = normal_workload(4, n, seed=43, wn=0.1)
        y, _ = _run(device, kernel, geom, unrolled, sos, x, schedule=schedule)
        ok, err, mref, rel = O.gate(y, O.golden(sos, x))
        assert ok, (n, rel)


def _histories(sos, x, zi=None):
    """FP64 direct-form histories after x: x[n-2], x[n-1], then y_s[n-2],
    y_s[n-1] of every section's output, from the histories zi before x."""
    ns = sos.shape[0]
    u = x.astype(np.float64)
    h = [0.0, 0.0] if zi is None else [float(zi[0]), float(zi[1])]
    full_u = np.concatenate([h, u])
    z = [full_u[-2], full_u[-1]]
    prev_hist = h
    for s in range(ns):
        b, a = sos[s, :3].astype(np.float64), sos[s, 3:].astype(np.float64)
        yh = [0.0, 0.0] if zi is None else [float(zi[2 + 2 * s]), float(zi[3 + 2 * s])]
        st = signal.lfiltic(b, a, y=[yh[1], yh[0]], x=[prev_hist[1], prev_hist[0]])
        u = signal.lfilter(b, a, u, zi=st)[0]
        full = np.concatenate([yh, u])
        z += [full[-2], full[-1]]
        prev_hist = yh
    return np.array(z)


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,geom,unrolled", CASES)
@pytest.mark.parametrize("r", [1, 2, 37])
def test_cascade_tail_state_out(device, kernel, geom, unrolled, r):
    """zf of a signal that ends r samples into its last chunk: r = 1 takes
    y_s[n-2] from the state before the chunk."""
    ch = geom[0] * geom[1]
    n = 9 * ch + r
    print(f"case tail state {kernel} {geom} n={n}", flush=True)
    sos, g, x = normal_workload(8, n, seed=47, wn=0.1)
    y, zf = _run(device, kernel, geom, unrolled, sos, x, want_zf=True)
    assert zf[0] == x[-2] and zf[1] == x[-1]
    ref = _histories(sos, x)
    u = x.astype(np.float64)
    for s in range(sos.shape[0]):
        u = signal.sosfilt(sos[s:s + 1].astype(np.float64), u)
        scale = np.max(np.abs(u))
        assert abs(zf[2 + 2 * s] - ref[2 + 2 * s]) <= 1e-5 * scale, (s, zf[2 + 2 * s], ref[2 + 2 * s])
        assert abs(zf[3 + 2 * s] - ref[3 + 2 * s]) <= 1e-5 * scale, (s, zf[3 + 2 * s], ref[3 + 2 * s])


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,geom,unrolled", CASES)
def test_cascade_uneven_calls(device, kernel, geom, unrolled):
    """A signal in calls of uneven lengths (one shorter than a chunk, none a
    multiple of it), each call's zf the next call's zi, one call with zf ==
    zi: equal to the whole signal in one call and to the FP64 cascade."""
    import torch
    ch = geom[0] * geom[1]
    lens = [5 * ch + 77, ch // 3, 2 * ch + 1, 7 * ch - 13]
    n = sum(lens)
    print(f"case tail calls {kernel} {geom} lens={lens}", flush=True)
    sos, g, x = normal_workload(16, n, seed=53, wn=0.1)
    ns = sos.shape[0]
    xin = torch.from_numpy(x).to(device)
    out = torch.empty_like(xin)
    za = torch.zeros(2 + 2 * ns, dtype=torch.float32, device=device)
    zb = torch.empty_like(za)
    o = 0
    for i, m in enumerate(lens):
        xc = xin[o:o + m].clone()  # a buffer of its own: 16-byte aligned
        yc = torch.empty_like(xc)
        with iirx.Plan(kernel, sos, m, block_size=geom[0], n_blocks=geom[1], unrolled=unrolled) as p:
            if i == 2:  # the state updated in place
                p.execute_state(xc, yc, zi=za, zf=za)
            else:
                p.execute_state(xc, yc, zi=za, zf=zb)
                za, zb = zb, za
            torch.cuda.synchronize()
            p.check()
        out[o:o + m] = yc
        o += m
    y = out.cpu().numpy()
    ref = O.golden(sos, x)
    ok, err, mref, rel = O.gate(y, ref)
    assert ok, rel
    one, _ = _run(device, kernel, geom, unrolled, sos, x)
    assert O.gate(one, ref)[0]
    # other chunk boundaries, other float32 rounding: agreement well inside the gate
    assert np.max(np.abs(y - one)) <= 1e-5 * mref
    zend = za.cpu().numpy()
    zref = _histories(sos, x)
    assert np.max(np.abs(zend - zref)) <= 1e-5 * max(1.0, float(np.max(np.abs(zref))))


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,geom,unrolled", CASES)
def test_cascade_short_call_state_in_place(device, kernel, geom, unrolled):
    """A call shorter than one chunk (only the last-chunk launch runs) with a
    non-zero state updated in place (zf is zi): every workgroup must read the
    input history x[-2], x[-1] from zi before any writes zf.  Checked against
    scipy's lfilter from lfiltic's state, section by section, and the state
    out against the FP64 histories."""
    import torch
    ch = geom[0] * geom[1]
    sos, g, x0 = normal_workload(8, 3 * ch, seed=59, wn=0.1)
    ns = sos.shape[0]
    # a state: the FP64 histories after a first stretch of signal
    zi = _histories(sos, x0[:ch])
    for n in (ch - 7, 3):
        x = x0[ch:ch + n].copy()
        print(f"case tail in-place state {kernel} {geom} unrolled={unrolled} n={n}", flush=True)
        xin = torch.from_numpy(x).to(device)
        yt = torch.empty_like(xin)
        z = torch.from_numpy(zi.astype(np.float32)).to(device)
        with iirx.Plan(kernel, sos, n, block_size=geom[0], n_blocks=geom[1], unrolled=unrolled) as p:
            p.execute_state(xin, yt, zi=z, zf=z)
            torch.cuda.synchronize()
            p.check()
        y = yt.cpu().numpy()
        zi32 = zi.astype(np.float32).astype(np.float64)
        u = x.astype(np.float64)
        prev = zi32[0:2]
        for s in range(ns):
            b, a = sos[s, :3].astype(np.float64), sos[s, 3:].astype(np.float64)
            yh = zi32[2 + 2 * s: 4 + 2 * s]
            st = signal.lfiltic(b, a, y=[yh[1], yh[0]], x=[prev[1], prev[0]])
            u = signal.lfilter(b, a, u, zi=st)[0]
            prev = yh
        ok, err, mref, rel = O.gate(y, u)
        assert ok, (n, rel)
        zend = z.cpu().numpy()
        zref = _histories(sos, x, zi=zi32)
        assert np.max(np.abs(zend - zref)) <= 1e-5 * max(1.0, float(np.max(np.abs(zref)))), (n, zend, zref)


BS_CASES = [(64, 64), (16, 128), (32, 32)]


@pytest.mark.gpu
@pytest.mark.parametrize("geom", BS_CASES)
@pytest.mark.parametrize("order", [8, 16])
def test_blockstate_any_length(device, geom, order):
    """The block-state form at lengths that are no multiple of its G x L
    chunk: the whole chunks on the matrix cores, the rest through the PH
    last-chunk kernel from the direct-form histories the state-out map gives
    after the whole chunks; signals shorter than one chunk go through PH
    alone."""
    ch = geom[0] * geom[1]
    for n in (7 * ch + 1, 7 * ch + 2, 7 * ch + 777, 8 * ch - 1, 100, ch - 3):
        print(f"case tail BS {geom} order={order} n={n}", flush=True)
        sos, g, x = normal_workload(order, n, seed=59, wn=0.1)
        y, _ = _run(device, "BS", geom, False, sos, x)
        ok, err, mref, rel = O.gate(y, O.golden(sos, x))
        assert ok, (n, rel)


@pytest.mark.gpu
@pytest.mark.parametrize("geom", BS_CASES)
@pytest.mark.parametrize("r", [1, 2, 37])
def test_blockstate_tail_state_out(device, geom, r):
    ch = geom[0] * geom[1]
    n = 5 * ch + r
    print(f"case tail BS state {geom} n={n}", flush=True)
    sos, g, x = normal_workload(16, n, seed=61, wn=0.1)
    y, zf = _run(device, "BS", geom, False, sos, x, want_zf=True)
    assert zf[0] == x[-2] and zf[1] == x[-1]
    ref = _histories(sos, x)
    u = x.astype(np.float64)
    for s in range(sos.shape[0]):
        u = signal.sosfilt(sos[s:s + 1].astype(np.float64), u)
        scale = np.max(np.abs(u))
        assert abs(zf[2 + 2 * s] - ref[2 + 2 * s]) <= 1e-5 * scale, s
        assert abs(zf[3 + 2 * s] - ref[3 + 2 * s]) <= 1e-5 * scale, s


@pytest.mark.gpu
def test_blockstate_uneven_calls(device):
    """Streaming calls of uneven lengths on the block-state form (one shorter
    than a chunk, one updating the state in place) against the FP64 cascade
    and its histories."""
    import torch
    geom = (64, 64)
    ch = geom[0] * geom[1]
    lens = [3 * ch + 77, ch // 3, 2 * ch + 1, 4 * ch - 13]
    n = sum(lens)
    print(f"case tail BS calls lens={lens}", flush=True)
    sos, g, x = normal_workload(16, n, seed=67, wn=0.1)
    ns = sos.shape[0]
    xin = torch.from_numpy(x).to(device)
    out = torch.empty_like(xin)
    za = torch.zeros(2 + 2 * ns, dtype=torch.float32, device=device)
    zb = torch.empty_like(za)
    o = 0
    for i, m in enumerate(lens):
        xc = xin[o:o + m].clone()
        yc = torch.empty_like(xc)
        with iirx.Plan("BS", sos, m, block_size=geom[0], n_blocks=geom[1]) as p:
            if i == 2:
                p.execute_state(xc, yc, zi=za, zf=za)
            else:
                p.execute_state(xc, yc, zi=za, zf=zb)
                za, zb = zb, za
            torch.cuda.synchronize()
            p.check()
        out[o:o + m] = yc
        o += m
    y = out.cpu().numpy()
    ok, err, mref, rel = O.gate(y, O.golden(sos, x))
    assert ok, rel
    zref = _histories(sos, x)
    assert np.max(np.abs(za.cpu().numpy() - zref)) <= 1e-5 * max(1.0, float(np.max(np.abs(zref))))
