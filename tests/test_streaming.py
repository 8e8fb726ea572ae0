"""Streaming state (iirx_execute_state, PH plans): a long signal filtered in
calls, the state of the cascade carried from one call to the next.  No
counterpart in the reference, whose initial-state hooks are always zero
(ph_kernels.cuh:45, 162-165).  Checked by what streaming must preserve:
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
                                                  ("DTCR", (64, 64), False), ("DTCR", (32, 32), False)])
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


@pytest.mark.gpu
def test_stream_state_rejected_by_plr_and_blockstate(device):
    import torch
    sos, g = iirx.butterworth_sos(4, 0.1)
    xin = torch.zeros(1 << 16, dtype=torch.float32, device=device)
    out = torch.empty_like(xin)
    for kernel, kw in (("PLR", {}), ("BS", {"block_size": 64, "n_blocks": 64})):
        with iirx.Plan(kernel, sos, xin.numel(), **kw) as p:
            with pytest.raises(iirx.IirxError):
                p.execute_state(xin, out)
