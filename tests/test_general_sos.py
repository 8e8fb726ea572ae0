"""General sections (iirx_normalize_sos): any design, any section scaling,
odd orders (first-order sections, b2 = a2 = 0).  The reference takes only its
own Butterworth designs (ref_generate.py:50); the filter contract stays the
reference's (b0 == a0 == 1, gain applied to the input), and plans reject rows
outside it."""
import numpy as np
import pytest
from scipy import signal

import cuda_recursive_filtering_amd as iirx
import oracle_lib as O

DESIGNS = {
    "butter3": lambda: signal.butter(3, 0.1, output="sos"),
    "cheby1_6": lambda: signal.cheby1(6, 1.0, 0.2, output="sos"),
    "ellip5": lambda: signal.ellip(5, 0.5, 40.0, 0.3, output="sos"),
    "cheby2_8_highpass": lambda: signal.cheby2(8, 40.0, 0.25, btype="highpass", output="sos"),
}


@pytest.mark.parametrize("name", sorted(DESIGNS))
def test_normalize_sos_is_the_same_filter(name):
    sos = DESIGNS[name]()
    rows, gain = iirx.normalize_sos(sos)
    assert rows.dtype == np.float32 and np.all(rows[:, 0] == 1) and np.all(rows[:, 3] == 1)
    x = np.random.default_rng(7).standard_normal(1 << 14)
    ref = signal.sosfilt(sos, x)
    y = gain * signal.sosfilt(rows.astype(np.float64), x)
    # float32 rounding of the normalised coefficients is the only difference
    assert np.max(np.abs(y - ref)) <= 1e-4 * np.max(np.abs(ref)), name


def test_normalize_sos_rejects_degenerate_rows():
    with pytest.raises(iirx.IirxError):
        iirx.normalize_sos([[0.0, 1.0, 0.0, 1.0, 0.5, 0.0]])  # b0 == 0
    with pytest.raises(iirx.IirxError):
        iirx.normalize_sos([[1.0, 1.0, 0.0, 0.0, 0.5, 0.0]])  # a0 == 0
    with pytest.raises(iirx.IirxError):
        iirx.normalize_sos([[1.0, np.nan, 0.0, 1.0, 0.5, 0.0]])


def test_plans_reject_rows_outside_the_contract():
    rows, _ = iirx.normalize_sos(DESIGNS["cheby1_6"]())
    bad = rows.copy()
    bad[1, 0] = 2.0  # b0 != 1
    with pytest.raises(iirx.IirxError):
        iirx.plan_tables("PH", bad, 1 << 16, block_size=64, n_blocks=32)
    bad = rows.copy()
    bad[0, 3] = 0.5  # a0 != 1
    with pytest.raises(iirx.IirxError):
        iirx.plan_tables("BS", bad, 1 << 16, block_size=64, n_blocks=64)
    iirx.plan_tables("PH", rows, 1 << 16, block_size=64, n_blocks=32)  # in the contract


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(DESIGNS))
@pytest.mark.parametrize("kernel,B,N", [("PH", 64, 32), ("PHJ", 64, 32), ("STCR", 32, 64), ("DTCR", 64, 64),
                                         ("BS", 64, 64)])
def test_general_designs_on_the_gpu(device, name, kernel, B, N):
    import torch
    print(f"case general sos {name} {kernel}", flush=True)
    sos = DESIGNS[name]()
    rows, gain = iirx.normalize_sos(sos)
    n = 1 << 20
    x = (gain * iirx.fill_normal(8, n).astype(np.float64)).astype(np.float32)  # gain on the input
    xin = torch.from_numpy(x).to(device)
    out = torch.empty_like(xin)
    try:
        plan = iirx.Plan(kernel, rows, n, block_size=B, n_blocks=N)
    except iirx.IirxError as e:
        # CR kernels reject ill-conditioned elimination factors instead of computing garbage
        assert kernel in ("STCR", "DTCR") and "ill-conditioned" in str(e), str(e)
        return
    with plan:
        plan.execute(xin, out)
        plan.check()
    y = out.cpu().numpy()
    ok, err, mref, rel = O.gate(y, O.golden(rows, x))
    assert ok, (name, kernel, rel)


@pytest.mark.gpu
def test_first_order_section_in_plr(device):
    """PLR takes the whole odd-order cascade as one direct form (order 3: K = 4
    with a zero tap)."""
    import torch
    rows, gain = iirx.normalize_sos(DESIGNS["butter3"]())
    n = (1 << 18) + 5
    x = (gain * iirx.fill_normal(9, n).astype(np.float64)).astype(np.float32)
    xin = torch.from_numpy(x).to(device)
    out = torch.empty_like(xin)
    with iirx.Plan("PLR", rows, n) as p:
        p.execute(xin, out)
        p.check()
    assert O.gate(out.cpu().numpy(), O.golden(rows, x))[0]
