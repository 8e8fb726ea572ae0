"""The C++ driver (cuda_recursive_filtering_amd/harness/iirx_driver.cpp): the
reference's src/test_accuracy.cu and src/test_performance.cu over the C ABI.
It filters the reference workload (ref_generate.py:56-99) and applies the
reference's gate against a golden float32 file, as the reference does with
build/reference.bin; here the golden file is the oracle's FP64 cascade."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from workloads import reference_workload

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cuda_recursive_filtering_amd",
                      "iirx_driver")


def run(*args, timeout=120):
    p = subprocess.run([DRIVER, *map(str, args)], capture_output=True, text=True, timeout=timeout)
    return p.returncode, p.stdout + p.stderr


def test_driver_usage():
    rc, out = run("--help")
    assert rc == 0 and "accuracy|performance" in out


def test_driver_rejects_invalid_configuration():
    """The static_assert of test_accuracy.cu:41 becomes a runtime check, made
    before any device call."""
    rc, out = run("accuracy", "--kernel", "PH", "--n-blocks", "48")
    assert rc == 255 and "configuration rejected" in out
    rc, out = run("accuracy", "--kernel", "XYZ")
    assert rc == 255


def _golden_file(tmp_path, sections, log2n):
    sos, g, x = reference_workload(2 * sections, 1 << log2n)
    ref = O.golden(sos, x).astype(np.float32)
    path = tmp_path / f"reference_{sections}_{log2n}.bin"
    ref.tofile(path)
    return path


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,extra,sections", [
    ("PH", ["--block-size", 64, "--n-blocks", 64], 2),
    ("PH", ["--block-size", 32, "--n-blocks", 64], 8),
    ("STCR", ["--block-size", 32, "--n-blocks", 64, "--unrolled"], 4),
    ("DTCR", ["--block-size", 64, "--n-blocks", 64], 2),
    ("PLR", [], 1),
    ("BS", ["--block-size", 64, "--n-blocks", 64], 8),
])
def test_driver_accuracy(device, tmp_path, kernel, extra, sections):
    print(f"case driver accuracy {kernel} {extra} ns={sections}", flush=True)
    ref = _golden_file(tmp_path, sections, 20)
    rc, out = run("accuracy", "--kernel", kernel, "--sections", sections, "--log2n", 20, "--reference", ref, *extra)
    print(out)
    assert rc == 0 and "PASSED" in out, out


@pytest.mark.gpu
def test_driver_plr_order16_fails(device, tmp_path):
    """The reference reports PLR's order-16 failure (README.md:39); the driver
    reproduces it with exit status 1."""
    print("case driver PLR order 16", flush=True)
    ref = _golden_file(tmp_path, 8, 18)
    rc, out = run("accuracy", "--kernel", "PLR", "--sections", 8, "--log2n", 18, "--reference", ref)
    assert rc == 1 and "FAILED" in out, out


@pytest.mark.gpu
def test_driver_performance(device, tmp_path):
    print("case driver performance", flush=True)
    ref = _golden_file(tmp_path, 1, 22)
    rc, out = run("performance", "--kernel", "PH", "--block-size", 64, "--n-blocks", 32, "--sections", 1,
                  "--log2n", 22, "--reference", ref, "--warmup", 5, "--batches", 5, "--batch", 20)
    print(out)
    assert rc == 0 and "Accuracy: PASSED" in out, out
    gsps = float(re.search(r"Throughput:\s+([0-9.]+) GS/s", out).group(1))
    assert gsps > 1.0
