"""Signals past 2^31 samples: every index of the kernels and of the lookback
records must be 64-bit where it has to be.  The input is zero except for
unit impulses at q = 2^31 - 100 and p = n - 3000; by linearity and
time-invariance the output is zero before q and h[k - q] + h[k - p] from q
on, with h the FP64 impulse response of the cascade (the reference's test
lengths stop at 2^25, run_accuracy.sh:37-49)."""
import numpy as np
import pytest
from scipy import signal

import cuda_recursive_filtering_amd as iirx

N_LARGE = (1 << 31) + (1 << 16)


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,B,N,order,tol", [
    ("PH", 64, 32, 8, 1e-5),
    ("PHJ", 64, 32, 8, 1e-5),
    ("PHH", 64, 32, 4, 1e-5),
    ("STCR", 64, 64, 4, 1e-5),
    ("DTCR", 64, 128, 4, 1e-5),
    ("PLR", None, 64, 4, 1e-4),  # float32 direct form: about 5e-5 of the peak at order 4
    ("BS", 64, 64, 8, 1e-5),
])
def test_impulses_past_2_31(device, kernel, B, N, order, tol):
    import torch
    print(f"case large {kernel} {B}x{N} order={order} n={N_LARGE}", flush=True)
    n = N_LARGE
    q, p = (1 << 31) - 100, n - 3000
    sos, g = iirx.butterworth_sos(order, 0.1)
    x = torch.zeros(n, dtype=torch.float32, device=device)
    x[q] = 1.0
    x[p] = 1.0
    y = torch.full((n,), float("nan"), dtype=torch.float32, device=device)
    with iirx.Plan(kernel, sos, n, block_size=B, n_blocks=N) as plan:
        plan.execute(x, y)
        torch.cuda.synchronize()
        plan.check()
    imp = np.zeros(n - q)
    imp[0] = 1.0
    imp[p - q] += 1.0
    ref = signal.sosfilt(sos.astype(np.float64), imp)
    tail = y[q:].cpu().numpy()
    head = y[:q]
    assert bool(torch.isfinite(head).all()) and float(head.abs().max()) == 0.0
    assert np.all(np.isfinite(tail))
    assert np.max(np.abs(tail - ref)) <= tol * np.max(np.abs(ref))
    del x, y
    torch.cuda.empty_cache()


@pytest.mark.gpu
@pytest.mark.parametrize("algo", ["ph", "blockstate"])
def test_batch_channels_past_2_31(device, algo):
    """A ragged batch whose second channel straddles sample 2^31 (CSR offsets
    are 64-bit): an impulse near the end of channel 0 and one at the start of
    channel 1 give each channel its own impulse response, from zero state."""
    import torch
    print(f"case large batch {algo}", flush=True)
    a0 = (1 << 31) - 50
    offsets = np.array([0, a0, a0 + (1 << 16) + 50], dtype=np.int64)
    n = int(offsets[-1])
    sos0, _ = iirx.butterworth_sos(8, 0.1)
    sos1, _ = iirx.butterworth_sos(16, 0.2)
    q0, q1 = a0 - 3000, a0 + 10
    x = torch.zeros(n, dtype=torch.float32, device=device)
    x[q0] = 1.0
    x[q1] = 1.0
    y = torch.full((n,), float("nan"), dtype=torch.float32, device=device)
    with iirx.Batch(offsets, [sos0, sos1], algo=algo) as b:
        b.execute(x, y)
        torch.cuda.synchronize()
        b.check()
    for sos, q, a, e in ((sos0, q0, 0, a0), (sos1, q1, a0, n)):
        imp = np.zeros(e - q)
        imp[0] = 1.0
        ref = signal.sosfilt(sos.astype(np.float64), imp)
        got = y[q:e].cpu().numpy()
        assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))
        before = y[max(a, q - 100000):q]
        assert bool(torch.isfinite(before).all()) and float(before.abs().max()) == 0.0
    del x, y
    torch.cuda.empty_cache()
