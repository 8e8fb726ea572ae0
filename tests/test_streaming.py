"""Streaming state (iirx_execute_state, every kernel): a long
signal filtered in calls, the filter state carried from one call to the
next.  No counterpart in the reference, whose initial-state hooks are always
zero (ph_kernels.cuh:45, 162-165; plr_kernels.cuh:134-142).  Checked by what streaming must preserve:
filtering in calls equals filtering the whole signal (against the FP64
cascade with the reference's gate, and against one call on the whole
signal), a zero state in equals the plain launch bit for bit, and the state
out equals the FP64 direct-form histories of every section."""
import numpy as np
import pytest
from scipy import signal

import cuda_recursive_filtering_amd as iirx
import oracle_lib as O
from workloads import normal_workload


def _stream(device, sos, x, calls, schedule="persistent", kernel="PH", geom=(64, 32), unrolled=False):
    import torch
    n = x.size // calls
    ns = sos.shape[0]
    xin = torch.from_numpy(x).to(device)
    out = torch.empty_like(xin)
    za = torch.zeros(2 + 2 * ns, dtype=torch.float32, device=device)
    zb = torch.empty_like(za)
    with iirx.Plan(kernel, sos, n, block_size=geom[0], n_blocks=geom[1], unrolled=unrolled, schedule=schedule) as p:
        for c in range(calls):
            p.execute_state(xin[c * n:(c + 1) * n], out[c * n:(c + 1) * n], zi=za, zf=zb)
            za, zb = zb, za
        p.check()
    return out.cpu().numpy(), za.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("order", [2, 8, 16])
def test_stream_two_calls_equal_one(device, order):
    import torch
    print(f"case stream two calls order={order}", flush=True)
    sos, g, x = normal_workload(order, 1 << 20, seed=31, wn=0.1)
    y2, _ = _stream(device, sos, x, 2)
    y1, _ = _stream(device, sos, x, 1)
    ref = O.golden(sos, x)
    assert O.gate(y2, ref)[0] and O.gate(y1, ref)[0]
    # the carried state makes the two calls one filter: agreement to float32 rounding
    assert np.max(np.abs(y2 - y1)) <= 64 * np.finfo(np.float32).eps * np.max(np.abs(ref))


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", ["persistent", "ticket"])
def test_stream_many_calls_narrow_filter(device, schedule):
    """16 calls of 2^16 at order 16 and Wn = 0.02 (slow decay: the state
    reaches far into the next call)."""
    print(f"case stream 16 calls narrow {schedule}", flush=True)
    sos, g, x = normal_workload(16, 1 << 20, seed=32, wn=0.02)
    y, _ = _stream(device, sos, x, 16, schedule)
    ok, err, mref, rel = O.gate(y, O.golden(sos, x))
    assert ok, rel


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,geom,unrolled", [("STCR", (32, 64), False), ("STCR", (64, 64), True),
                                                  ("DTCR", (64, 64), False), ("DTCR", (32, 32), False),
                                                  ("PHJ", (64, 32), False), ("PHJ", (64, 16), False),
                                                  ("PHH", (64, 32), False), ("PHH", (32, 32), False)])
@pytest.mark.parametrize("order", [4, 16])
def test_stream_cr_kernels(device, kernel, geom, unrolled, order):
    """STCR and DTCR carry the same direct-form history between chunks, so the
    state streams the same way."""
    print(f"case stream {kernel} {geom} unrolled={unrolled} order={order}", flush=True)
    sos, g, x = normal_workload(order, 1 << 20, seed=35, wn=0.1)
    y, _ = _stream(device, sos, x, 4, kernel=kernel, geom=geom, unrolled=unrolled)
    ok, err, mref, rel = O.gate(y, O.golden(sos, x))
    assert ok, rel
    y1, _ = _stream(device, sos, x, 1, kernel=kernel, geom=geom, unrolled=unrolled)
    assert np.max(np.abs(y - y1)) <= 64 * np.finfo(np.float32).eps * mref


@pytest.mark.gpu
def test_stream_zero_state_is_the_plain_launch(device):
    import torch
    print("case stream zero state", flush=True)
    sos, g, x = normal_workload(8, 1 << 18, seed=33, wn=0.1)
    xin = torch.from_numpy(x).to(device)
    a, b = torch.empty_like(xin), torch.empty_like(xin)
    zi = torch.zeros(2 + 2 * sos.shape[0], dtype=torch.float32, device=device)
    with iirx.Plan("PH", sos, x.size, block_size=64, n_blocks=32) as p:
        p.execute(xin, a)
        p.execute_state(xin, b, zi=zi)
        p.check()
        assert torch.equal(a, b)
        b.zero_()
        p.execute_state(xin, b)  # zi omitted: zero state
        p.check()
    assert torch.equal(a, b)


@pytest.mark.gpu
def test_stream_state_out_is_the_direct_form_history(device):
    import torch
    print("case stream state out", flush=True)
    sos, g, x = normal_workload(8, 1 << 18, seed=34, wn=0.1)
    xin = torch.from_numpy(x).to(device)
    out = torch.empty_like(xin)
    zf = torch.empty(2 + 2 * sos.shape[0], dtype=torch.float32, device=device)
    with iirx.Plan("PH", sos, x.size, block_size=64, n_blocks=32) as p:
        p.execute_state(xin, out, zf=zf)
        p.check()
    z = zf.cpu().numpy()
    assert z[0] == x[-2] and z[1] == x[-1]
    u = x.astype(np.float64)
    for s in range(sos.shape[0]):
        u = signal.sosfilt(sos[s:s + 1].astype(np.float64), u)  # output of section s
        scale = np.max(np.abs(u))
        assert abs(z[2 + 2 * s] - u[-2]) <= 1e-5 * scale and abs(z[3 + 2 * s] - u[-1]) <= 1e-5 * scale, s


def _sos_fp64_with_history(sos, x, zi):
    """FP64 cascade from the DF1 histories zi (x[-2], x[-1], then y_s[-2],
    y_s[-1] per section): lfiltic turns each section's history into
    sosfilt's state."""
    ns = sos.shape[0]
    state = np.zeros((ns, 2))
    u2, u1 = float(zi[0]), float(zi[1])
    for s in range(ns):
        y2, y1 = float(zi[2 + 2 * s]), float(zi[3 + 2 * s])
        state[s] = signal.lfiltic(sos[s, :3].astype(np.float64), sos[s, 3:].astype(np.float64), y=[y1, y2], x=[u1, u2])
        u2, u1 = y2, y1
    return signal.sosfilt(sos.astype(np.float64), x.astype(np.float64), zi=state)[0]


@pytest.mark.gpu
@pytest.mark.parametrize("order", [8, 16])
def test_stream_blockstate_calls(device, order):
    """The block-state form in 4 calls: the histories go in through the
    balanced state map and come out by re-filtering the last tile in FP64;
    the output passes the gate and matches one call on the whole signal."""
    import torch
    print(f"case stream block-state order={order}", flush=True)
    sos, g, x = normal_workload(order, 1 << 20, seed=60 + order, wn=0.1)
    y, _ = _stream(device, sos, x, 4, kernel="BS", geom=(64, 64))
    ref = O.golden(sos, x)
    assert O.gate(y, ref)[0]
    xin = torch.from_numpy(x).to(device)
    whole = torch.empty_like(xin)
    with iirx.Plan("BS", sos, x.size, block_size=64, n_blocks=64) as p:
        p.execute(xin, whole)
        p.check()
    assert np.max(np.abs(y - whole.cpu().numpy())) <= 1e-5 * np.max(np.abs(ref))


@pytest.mark.gpu
def test_stream_blockstate_history_in_and_out(device):
    """A non-zero history in against scipy (lfiltic per section); the
    histories out against the FP64 sections; a zero history is the plain
    launch bit for bit."""
    import torch
    print("case stream block-state history", flush=True)
    sos, g, x = normal_workload(8, 1 << 18, seed=66, wn=0.1)
    ns = sos.shape[0]
    rng = np.random.default_rng(67)
    zi_h = (0.3 * rng.standard_normal(2 + 2 * ns)).astype(np.float32)
    xin = torch.from_numpy(x).to(device)
    out = torch.empty_like(xin)
    plain = torch.empty_like(xin)
    zi = torch.from_numpy(zi_h).to(device)
    zf = torch.empty_like(zi)
    with iirx.Plan("BS", sos, x.size, block_size=64, n_blocks=64) as p:
        p.execute_state(xin, out, zi=zi, zf=zf)
        p.check()
        y = out.cpu().numpy()
        ref = _sos_fp64_with_history(sos, x, zi_h)
        assert np.max(np.abs(y - ref)) <= 1e-5 * np.max(np.abs(ref))
        z = zf.cpu().numpy()
        assert z[0] == x[-2] and z[1] == x[-1]
        u = x.astype(np.float64)
        u2, u1 = float(zi_h[0]), float(zi_h[1])
        for s in range(ns):
            st = signal.lfiltic(sos[s, :3].astype(np.float64), sos[s, 3:].astype(np.float64),
                                y=[float(zi_h[3 + 2 * s]), float(zi_h[2 + 2 * s])], x=[u1, u2])
            u2, u1 = float(zi_h[2 + 2 * s]), float(zi_h[3 + 2 * s])
            u = signal.sosfilt(sos[s:s + 1].astype(np.float64), u, zi=st[None, :])[0]  # output of section s
            scale = np.max(np.abs(u))
            assert abs(z[2 + 2 * s] - u[-2]) <= 1e-5 * scale and abs(z[3 + 2 * s] - u[-1]) <= 1e-5 * scale, s
        p.execute(xin, plain)
        p.execute_state(xin, out, zi=torch.zeros_like(zi))
        p.check()
    assert torch.equal(out, plain)


def _plr_fp64(sos, x, x_hist=None, y_hist=None):
    """FP64 direct form of the cascade with the history x[-K..-1], y[-K..-1]
    (scipy's lfiltic turns a history into lfilter's state)."""
    b, a = iirx.direct_form_taps(sos)
    bb, aa = np.r_[1.0, b], np.r_[1.0, -a]
    if x_hist is None:
        return signal.lfilter(bb, aa, x.astype(np.float64))
    zi = signal.lfiltic(bb, aa, y=np.asarray(y_hist, np.float64)[::-1], x=np.asarray(x_hist, np.float64)[::-1])
    return signal.lfilter(bb, aa, x.astype(np.float64), zi=zi)[0]


@pytest.mark.gpu
@pytest.mark.parametrize("order", [2, 4])
def test_stream_plr_calls_equal_one(device, order):
    """PLR in 3 calls of a length that is no multiple of its chunk (nor of 4:
    the slices are not 16-byte aligned): the output passes the gate against
    the FP64 cascade, as does one call on the whole signal, and the two agree
    to the gate's bound."""
    import torch
    print(f"case stream PLR order={order}", flush=True)
    calls, n = 3, 100003
    sos, g, x = normal_workload(order, calls * n, seed=40 + order, wn=0.1)
    xin = torch.from_numpy(x).to(device)
    out = torch.empty_like(xin)
    whole = torch.empty_like(xin)
    K = 2 * sos.shape[0]
    za = torch.zeros(2 * K, dtype=torch.float32, device=device)
    zb = torch.empty_like(za)
    with iirx.Plan("PLR", sos, n) as p:
        assert p.state_size() == 2 * K
        for c in range(calls):
            p.execute_state(xin[c * n:(c + 1) * n], out[c * n:(c + 1) * n], zi=za, zf=zb)
            za, zb = zb, za
        p.check()
    with iirx.Plan("PLR", sos, calls * n) as p:
        p.execute(xin, whole)
        p.check()
    y, yw = out.cpu().numpy(), whole.cpu().numpy()
    ref = O.golden(sos, x)
    assert O.gate(y, ref)[0] and O.gate(yw, ref)[0]
    # the float32 direct form's own error (about 5e-5 of the peak at order 4)
    # differs between the two geometries; both stay within the gate
    assert np.max(np.abs(y - yw)) <= 1e-4 * np.max(np.abs(ref))


@pytest.mark.gpu
def test_stream_plr_history_in_and_out(device):
    """A non-zero history in (x[-K..-1], y[-K..-1]) against scipy's lfilter
    with lfiltic's state; the history out is the call's last K inputs and
    outputs bit for bit; a call shorter than K takes the rest from the history
    in, also when the state is updated in place (zf = zi)."""
    import torch
    print("case stream PLR history", flush=True)
    sos, g = iirx.butterworth_sos(4, 0.1)
    K = 4
    rng = np.random.default_rng(44)
    x = (g * rng.standard_normal(50000)).astype(np.float32)
    x_hist = (g * rng.standard_normal(K)).astype(np.float32)
    y_hist = (0.5 * rng.standard_normal(K)).astype(np.float32)
    xin = torch.from_numpy(x).to(device)
    out = torch.empty_like(xin)
    zi = torch.from_numpy(np.r_[x_hist, y_hist].astype(np.float32)).to(device)
    zf = torch.empty_like(zi)
    with iirx.Plan("PLR", sos, x.size) as p:
        p.execute_state(xin, out, zi=zi, zf=zf)
        p.check()
    y = out.cpu().numpy()
    ref = _plr_fp64(sos, x, x_hist, y_hist)
    assert np.max(np.abs(y - ref)) <= 1e-4 * np.max(np.abs(ref))
    z = zf.cpu().numpy()
    assert np.array_equal(z[:K], x[-K:]) and np.array_equal(z[K:], y[-K:])
    # a call of 3 < K samples, state updated in place
    xs = torch.from_numpy(x[:3].copy()).to(device)
    ys = torch.empty_like(xs)
    zs = torch.from_numpy(np.r_[x_hist, y_hist].astype(np.float32)).to(device)
    with iirx.Plan("PLR", sos, 3) as p:
        p.execute_state(xs, ys, zi=zs, zf=zs)
        p.check()
    ys, zs = ys.cpu().numpy(), zs.cpu().numpy()
    assert np.max(np.abs(ys - ref[:3])) <= 1e-4 * np.max(np.abs(ref))
    assert np.array_equal(zs[:K], np.r_[x_hist[3:], x[:3]])
    assert np.array_equal(zs[K:], np.r_[y_hist[3:], ys])
