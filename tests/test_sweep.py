"""Length / conditioning sweep: order-8 Butterworth at Wn in {0.02, 0.1, 0.4}
(poles near the unit circle at 0.02), lengths 2^16 .. 2^25, input
N(0,1) + 0.5 sin(2 pi 0.05 n) (seed 2).  It exercises lookback chain depth
(deep at Wn = 0.02) and small grids (2^16 samples = 16 chunks of 64 x 64).
Every kernel is gated against the FP64 cascade (test_accuracy.cu:178-190);
STCR/DTCR may instead be rejected as CR-ill-conditioned (status 3)."""
import functools

import numpy as np
import pytest

import cuda_recursive_filtering_amd as iirx
import oracle_lib as O
from workloads import normal_workload

WNS = [0.02, 0.1, 0.4]
GEOM = {"PH": (64, 64), "PHJ": (64, 32), "STCR": (64, 64), "DTCR": (64, 64), "BS": (64, 64)}


@functools.lru_cache(maxsize=8)
def _case(wn, log2n):
    sos, g, x = normal_workload(8, 1 << log2n, seed=2, wn=wn, sine=0.5)
    return sos, x, O.golden(sos, x)


def _run(device, kernel, wn, log2n):
    import torch
    sos, x, ref = _case(wn, log2n)
    B, N = GEOM[kernel]
    print(f"case sweep {kernel} wn={wn} n=2^{log2n}", flush=True)
    try:
        p = iirx.Plan(kernel, sos, x.size, block_size=B, n_blocks=N)
    except iirx.IirxError as e:
        assert kernel in ("STCR", "DTCR") and e.code == 3, e
        return
    with p:
        xin = torch.from_numpy(x).to(device)
        out = torch.empty_like(xin)
        p.execute(xin, out)
        p.check()
        y = out.cpu().numpy()
    ok, err, mref, rel = O.gate(y, ref)
    assert ok, (kernel, wn, log2n, rel)


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", list(GEOM))
@pytest.mark.parametrize("wn", WNS)
@pytest.mark.parametrize("log2n", [16, 18, 20, 22])
def test_sweep(device, kernel, wn, log2n):
    _run(device, kernel, wn, log2n)


@pytest.mark.gpu
@pytest.mark.slow
@pytest.mark.parametrize("kernel", list(GEOM))
@pytest.mark.parametrize("wn", WNS)
@pytest.mark.parametrize("log2n", [17, 19, 21, 23, 24, 25])
def test_sweep_long(device, kernel, wn, log2n):
    _run(device, kernel, wn, log2n)
