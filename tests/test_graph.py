"""A plan's call captured into a HIP graph and replayed.

One filter call is kernel launches only: the launch index of the versioned
lookback flags lives on the device (persistent epoch, or the ticket counter's
epoch), so no memset is recorded and every replay runs a fresh epoch
(ph_kernels.cuh:16-29 keeps the launch index on the host instead).  Replays
alternate with eager work on the same buffers and with eager calls of the
same plan, and every replay must match the eager result to 8 float32 ulps of
its peak.  A chunk's carry may come from its predecessor's full record or
from folding partials, depending on timing; the two agree to rounding, while
a stale carry is off by the signal's own size."""
import numpy as np
import pytest

import cuda_recursive_filtering_amd as iirx


def _close(y, ref):
    import torch
    tol = 8.0 * float(np.finfo(np.float32).eps) * float(ref.abs().max())
    return bool(torch.isfinite(y).all()) and float((y - ref).abs().max()) <= tol


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", ["persistent", "ticket"])
@pytest.mark.parametrize("kernel,B,N,order", [
    ("PH", 64, 32, 4),
    ("PHJ", 64, 32, 8),
    ("PHH", 64, 32, 4),
    ("STCR", 32, 64, 4),
    ("DTCR", 64, 64, 2),
    ("PLR", None, 64, 4),
    ("BS", 64, 64, 8),
])
def test_plan_graph_replay(device, kernel, B, N, order, schedule):
    import torch
    print(f"case graph replay {kernel} {B}x{N} order={order} {schedule}", flush=True)
    n = 1 << 18
    sos, g = iirx.butterworth_sos(order, 0.1)
    xs = [torch.from_numpy((g * iirx.fill_normal(seed, n)).astype(np.float32)) for seed in (11, 12, 13)]
    xin = torch.empty(n, dtype=torch.float32, device=device)
    out = torch.empty_like(xin)
    with iirx.Plan(kernel, sos, n, block_size=B, n_blocks=N, schedule=schedule) as p:
        refs = []
        for x in xs:
            xin.copy_(x)
            p.execute(xin, out)
            torch.cuda.synchronize()
            refs.append(out.clone())
        p.check()
        xin.copy_(xs[0])
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            p.execute(xin, out)
        torch.cuda.synchronize()
        for rep in range(6):
            i = rep % 3
            xin.copy_(xs[i])
            out.fill_(float("nan"))
            torch.cuda.synchronize()
            graph.replay()
            torch.cuda.synchronize()
            p.check()
            assert _close(out, refs[i]), f"replay {rep}"
            if rep % 2 == 1:  # an eager call of the same plan between replays
                out.fill_(float("nan"))
                p.execute(xin, out)
                torch.cuda.synchronize()
                p.check()
                assert _close(out, refs[i]), f"eager after replay {rep}"
