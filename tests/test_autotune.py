"""Autotuner: the runtime replacement of the reference's geometry sweep and
PLR_X_MAX calibration (run_accuracy.sh:95-151).  Candidate classification is
host logic (CPU tests); the timed search runs on the GPU."""
import collections

import numpy as np
import pytest

import cuda_recursive_filtering_amd as iirx
import oracle_lib as O
from workloads import normal_workload

# a biquad with poles at +-0.9j: a1 = 0 makes the CR elimination divide by
# e_0 = a1 (stcr_kernels.cuh setup), so STCR/DTCR must be excluded
QUADRATURE = np.array([[1.0, 0.5, 0.25, 1.0, 0.0, 0.81]], np.float32)


def _by_kernel(entries):
    out = collections.defaultdict(list)
    for e in entries:
        out[e["kernel"]].append(e)
    return out


def test_candidates_cover_every_algorithm():
    sos, _ = iirx.butterworth_sos(8, 0.1)
    k = _by_kernel(iirx.tune_candidates(sos, 1 << 22))
    assert len(k["PH"]) == 5 and len(k["STCR"]) == 10 and len(k["DTCR"]) == 10 and len(k["BS"]) == 8
    assert len(k["PHJ"]) == 5 and len(k["PHH"]) == 4 and len(k["BS_BF16"]) == 4
    assert "PLR" not in k  # the direct form is offered at orders <= 4 only
    assert all(e["status"] == 0 for es in k.values() for e in es)
    k2 = _by_kernel(iirx.tune_candidates(iirx.butterworth_sos(4, 0.1)[0], 1 << 22))
    assert len(k2["PLR"]) == 1 and k2["PLR"][0]["status"] == 0
    # one section: nothing to pair; halo carries still offered
    k1 = _by_kernel(iirx.tune_candidates(iirx.butterworth_sos(2, 0.1)[0], 1 << 22))
    assert "PHJ" not in k1 and len(k1["PHH"]) == 4


def test_candidates_flags():
    sos, _ = iirx.butterworth_sos(8, 0.1)
    k = _by_kernel(iirx.tune_candidates(sos, 1 << 22, flags=iirx._lib.TUNE_NO_BLOCKSTATE | iirx._lib.TUNE_NO_CR))
    assert set(k) == {"PH", "PHJ", "PHH"}
    k = _by_kernel(iirx.tune_candidates(sos, 1 << 22, flags=iirx._lib.TUNE_PLR_ALL_ORDERS))
    assert "PLR" in k
    k = _by_kernel(iirx.tune_candidates(sos, 1 << 22, flags=iirx._lib.TUNE_TICKET))
    assert all(e["schedule"] == "ticket" for es in k.values() for e in es)


def test_candidates_any_length():
    """Every kernel takes a partial last chunk, so a length that is no
    multiple of any chunk leaves every candidate eligible."""
    sos, _ = iirx.butterworth_sos(4, 0.1)
    n = 64 * 8 * 3 + 64  # no multiple of any block-state chunk (512 samples and up)
    k = _by_kernel(iirx.tune_candidates(sos, n))
    ph = {e["n_blocks"]: e["status"] for e in k["PH"]}
    assert all(ph[N] == 0 for N in (8, 16, 32, 64, 128))
    assert k["BS"] and all(e["status"] == 0 for e in k["STCR"] + k["DTCR"] + k["PLR"] + k["BS"])


def test_candidates_exclude_ill_conditioned_cr():
    k = _by_kernel(iirx.tune_candidates(QUADRATURE, 1 << 20))
    assert all(e["status"] == 3 for e in k["STCR"] + k["DTCR"])
    assert all(e["status"] == 0 for e in k["PH"] + k["BS"])


@pytest.mark.gpu
def test_autotune_picks_the_fastest_valid(device):
    import torch
    print("case autotune order 8", flush=True)
    sos, g, x = normal_workload(8, 1 << 22, seed=7, wn=0.1)
    best, ms, table = iirx.autotune(sos, x.size, reps=5)
    timed = [e for e in table if e["status"] == 0 and e["ms"] > 0]
    assert len(timed) >= 20
    assert ms <= min(e["ms"] for e in timed) + 1e-6
    with iirx.Plan.autotuned(sos, x.size, reps=5) as p:
        xin = torch.from_numpy(x).to(device)
        out = torch.empty_like(xin)
        p.execute(xin, out)
        p.check()
        assert O.gate(out.cpu().numpy(), O.golden(sos, x))[0]


@pytest.mark.gpu
def test_autotune_falls_back_from_cr(device):
    print("case autotune ill-conditioned CR", flush=True)
    best, ms, table = iirx.autotune(QUADRATURE, 1 << 20, reps=3)
    assert best["kernel"] in ("PH", "PHJ", "PHH", "BS")  # any form but the CR ones
    assert all(e["ms"] < 0 for e in table if e["kernel"] in ("STCR", "DTCR"))


@pytest.mark.gpu
def test_autotune_screens_accuracy(device):
    """Every timed candidate carries its error against the FP64 cascade of the
    tuning input; the float32 direct form at order 16 (offered only with
    TUNE_PLR_ALL_ORDERS) fails the screen and is never chosen."""
    print("case autotune accuracy screen order 16", flush=True)
    sos, _ = iirx.butterworth_sos(16, 0.1)
    best, ms, table = iirx.autotune(sos, 1 << 20, flags=iirx._lib.TUNE_PLR_ALL_ORDERS, reps=3)
    timed = [e for e in table if e["ms"] > 0]
    assert timed and all(e["rel_err"] >= 0 for e in timed)
    plr = [e for e in table if e["kernel"] == "PLR"]
    assert plr and all(e["status"] == iirx._lib.IIRX_ERR_ACCURACY for e in plr)
    assert best["kernel"] != "PLR"
    passing = [e for e in timed if e["status"] == 0]
    assert passing and all(e["rel_err"] <= 1e-4 for e in passing)
    assert ms <= min(e["ms"] for e in passing) + 1e-6


def test_accuracy_status_has_a_message():
    assert "accuracy" in iirx._lib.lib().iirx_status_string(iirx._lib.IIRX_ERR_ACCURACY).decode()
