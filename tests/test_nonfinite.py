"""A non-finite input sample (NaN or Inf at sample k).  Sequential filtering
keeps every later output non-finite; here the bounded-depth lookback drops a
chunk's carry once it has decayed below 2^-36 of the state, which for a
non-finite carry means dropping it altogether (PHH: a chunk's carry sees only
the 256 samples before it).  The contract pinned here:
outputs before k are those of the finite signal (the block-state form: before
the 32-sample K slab holding k, whose Toeplitz product multiplies the
non-finite value by the zeros above the diagonal), output k is non-finite,
the non-finite stretch ends within a bounded number of chunks, and every
output after it is that of the signal with x[k] = 0."""
import numpy as np
import pytest

import cuda_recursive_filtering_amd as iirx
import oracle_lib as O


@pytest.mark.gpu
@pytest.mark.parametrize("bad", [np.nan, np.inf])
@pytest.mark.parametrize("kernel,B,N,order,early", [
    ("PH", 64, 32, 8, 0),
    ("PHJ", 64, 32, 8, 0),
    ("PHH", 64, 32, 4, 0),
    ("STCR", 64, 64, 4, 0),
    ("DTCR", 64, 64, 4, 0),
    ("PLR", None, 64, 4, 0),
    ("BS", 64, 64, 8, 31),
])
def test_nonfinite_sample(device, kernel, B, N, order, early, bad):
    import torch
    print(f"case non-finite {kernel} {bad}", flush=True)
    n, k = 1 << 20, 300013
    sos, g = iirx.butterworth_sos(order, 0.1)
    x = (g * iirx.fill_normal(7, n)).astype(np.float32)
    xb = x.copy()
    xb[k] = bad
    xin = torch.from_numpy(xb).to(device)
    y = torch.empty_like(xin)
    with iirx.Plan(kernel, sos, n, block_size=B, n_blocks=N) as p:
        p.execute(xin, y)
        torch.cuda.synchronize()
        p.check()
    y = y.cpu().numpy()
    fin = np.isfinite(y)
    first = int(np.argmin(fin))
    last = int(n - 1 - np.argmin(fin[::-1]))
    print(f"  non-finite [{first}, {last}]")
    assert not fin[k]
    assert k - early <= first <= k
    assert (~fin[first:last + 1]).all()  # one contiguous stretch
    assert last - k < 16384
    assert O.gate(y[:first], O.golden(sos, x[:first]))[0]
    x0 = x.copy()
    x0[k] = 0.0
    ref0 = O.golden(sos, x0)
    assert O.gate(y[last + 1:], ref0[last + 1:])[0]
