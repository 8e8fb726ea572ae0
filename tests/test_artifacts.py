"""The reference harness's artefacts: build/filter_taps.hpp written and read
by the library (csrc/artifacts.cpp, the declarations of ref_generate.py:
101-131), and both artefacts driving the C++ driver as they drive the
reference's test_accuracy.cu (:107-109 header check, :164-204 gate)."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import cuda_recursive_filtering_amd as iirx
import ref_artifacts as RA

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cuda_recursive_filtering_amd",
                      "iirx_driver")


def run(*args, timeout=120):
    p = subprocess.run([DRIVER, *map(str, args)], capture_output=True, text=True, timeout=timeout)
    return p.returncode, p.stdout + p.stderr


def _same(t, u):
    assert t["n_sections"] == u["n_sections"] and t["plr_order"] == u["plr_order"]
    for key in ("gain", "input_fs", "input_freq"):
        assert float(t[key]).hex() == float(u[key]).hex(), key
    assert np.array_equal(np.asarray(t["sos"], np.float32).view(np.uint32),
                          np.asarray(u["sos"], np.float32).view(np.uint32))
    for key in ("plr_b", "plr_a"):
        assert np.array_equal(np.asarray(t[key], np.float64).view(np.uint64),
                              np.asarray(u[key], np.float64).view(np.uint64)), key


@pytest.mark.parametrize("ns", [1, 2, 4, 8])
def test_design_is_scipys(ns):
    """The values equal scipy's FP64 design (ref_generate.py:61-81): the
    float32 sections bit for bit, the gain and the direct form to FP64
    rounding (the closed form and scipy's zpk route round differently)."""
    t = iirx.filter_taps_design(ns, 0.2)
    s = RA.design(ns, 0.2)
    assert np.array_equal(t["sos"], s["sos"])
    assert math.isclose(t["gain"], s["gain"], rel_tol=1e-13)
    assert t["input_fs"] == 1000.0 and t["input_freq"] == 120.0 and t["cutoff_wn"] == 0.2
    assert np.allclose(t["plr_b"], s["plr_b"], rtol=0, atol=1e-12 * np.max(np.abs(s["plr_b"])))
    assert np.allclose(t["plr_a"], s["plr_a"], rtol=0, atol=1e-11 * np.max(np.abs(s["plr_a"])))
    # the direct form is the FP64 cascade: the same impulse response
    imp = np.zeros(256)
    imp[0] = 1.0
    sos64 = RA.signal.butter(2 * ns, 0.2, output="sos")
    h_sos = RA.signal.sosfilt(sos64, imp) / s["gain"]
    h_df = RA.signal.lfilter(np.r_[1.0, t["plr_b"]], np.r_[1.0, -t["plr_a"]], imp)
    # (order 16 direct form: lfilter itself loses ~1e-9 of the peak)
    assert np.max(np.abs(h_sos - h_df)) < 1e-8 * np.max(np.abs(h_sos))


@pytest.mark.parametrize("ns", [1, 3, 8])
def test_write_read_roundtrip(tmp_path, ns):
    t = iirx.filter_taps_design(ns, 0.13)
    p = tmp_path / "taps.hpp"
    iirx.write_filter_taps(t, p)
    u = iirx.read_filter_taps(p)
    _same(t, u)
    assert u["cutoff_wn"] == 0.13


@pytest.mark.parametrize("ns", [2, 8])
def test_reads_the_reference_layout(tmp_path, ns):
    """A header in ref_generate.py's text reads back to the values it was
    printed from, bit for bit."""
    s = RA.design(ns)
    p = tmp_path / "filter_taps.hpp"
    p.write_text(RA.header_text(s))
    _same(iirx.read_filter_taps(p), s)


@pytest.mark.parametrize("ns", [1, 2, 4, 8])
def test_writes_the_reference_text(tmp_path, ns):
    """Given scipy's values, the library writes every declaration line as
    ref_generate.py:106-129 prints it ({v:.9e}f, repr)."""
    s = RA.design(ns)
    p = tmp_path / "taps.hpp"
    iirx.write_filter_taps(s, p)
    assert RA.declarations(p.read_text()) == RA.declarations(RA.header_text(s))


def test_double_text_is_repr(tmp_path):
    """Shortest round-trip text in repr's layout, at its layout switches
    (1e-4 / 1e-5, 1e15 / 1e16), for subnormals, -0.0, ties and long values."""
    vals = [1e-4, 1e-5, 1.5e-5, 1e15, 1e16, 1.2345e16, 123.0, 0.1 + 0.2, -0.0, 5e-324, 2.2250738585072014e-308,
            1.7976931348623157e308, 2.0 ** -30, 9007199254740993.0, 1 / 3, -2.5, 100.0, 1e22, 12345678.9]
    k = 16
    t = {"n_sections": 1, "plr_order": k, "gain": vals[0], "input_fs": vals[1], "input_freq": vals[2],
         "sos": np.ones((1, 6), np.float32), "plr_b": np.array(vals[3:3 + k]),
         "plr_a": np.array((vals[3:] + vals[:3])[:k])}
    p = tmp_path / "taps.hpp"
    iirx.write_filter_taps(t, p)
    assert RA.declarations(p.read_text()) == RA.declarations(RA.header_text(t))
    _same(iirx.read_filter_taps(p), t)


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_written_header_compiles(tmp_path):
    """The header builds a program the way the reference's drivers include it,
    and the compiler reads the same values the library wrote."""
    t = iirx.filter_taps_design(4, 0.2)
    iirx.write_filter_taps(t, tmp_path / "filter_taps.hpp")
    (tmp_path / "main.cpp").write_text(r'''
#include <cstdio>
#include <type_traits>
#include "filter_taps.hpp"
static_assert(FILTER_N_SECTIONS == 4, "sections");
static_assert(PLR_FILTER_ORDER == 8, "order");
static_assert(std::is_same<decltype(FILTER_SOS), const float[4][6]>::value, "sos type");
static_assert(std::is_same<decltype(PLR_FILTER_A), const double[8]>::value, "a type");
int main() {
    std::printf("%a %a %a\n", FILTER_GAIN, INPUT_FS, INPUT_FREQ);
    for (int s = 0; s < FILTER_N_SECTIONS; ++s)
        for (int j = 0; j < 6; ++j) std::printf("%a\n", (double)FILTER_SOS[s][j]);
    for (int j = 0; j < PLR_FILTER_ORDER; ++j) std::printf("%a %a\n", PLR_FILTER_B[j], PLR_FILTER_A[j]);
}
''')
    exe = tmp_path / "main"
    subprocess.check_call(["g++", "-std=c++17", "-Wall", "-Werror", "-Wno-unused-variable", "-o", str(exe),
                           str(tmp_path / "main.cpp")])
    out = subprocess.check_output([str(exe)], text=True).split()
    vals = [float.fromhex(v) for v in out]
    assert vals[:3] == [t["gain"], t["input_fs"], t["input_freq"]]
    assert vals[3:27] == [float(v) for v in t["sos"].reshape(-1)]
    assert vals[27::2] == list(t["plr_b"]) and vals[28::2] == list(t["plr_a"])


def test_read_rejects_bad_headers(tmp_path):
    good = RA.header_text(RA.design(2))
    bad = {
        "missing_gain": good.replace("FILTER_GAIN", "FILTER_GAIN_X"),
        "sos_count": good.replace("#define FILTER_N_SECTIONS 2", "#define FILTER_N_SECTIONS 3"),
        "order": good.replace("#define PLR_FILTER_ORDER 4", "#define PLR_FILTER_ORDER 5"),
        "too_many_sections": good.replace("#define FILTER_N_SECTIONS 2", "#define FILTER_N_SECTIONS 9"),
        "garbage": good.replace("{1.000000000e+00f", "{one"),
    }
    for name, text in bad.items():
        p = tmp_path / f"{name}.hpp"
        p.write_text(text)
        with pytest.raises(iirx.IirxError):
            iirx.read_filter_taps(p)
    with pytest.raises(iirx.IirxError):
        iirx.read_filter_taps(tmp_path / "does_not_exist.hpp")


def test_driver_writes_the_header(tmp_path):
    """`iirx_driver taps` writes the reference workload's header without a
    device: the same file as the library's design written directly."""
    p = tmp_path / "a.hpp"
    rc, out = run("taps", "--sections", 4, "--write-taps", p)
    assert rc == 0, out
    q = tmp_path / "b.hpp"
    iirx.write_filter_taps(iirx.filter_taps_design(4, 0.2), q)
    assert p.read_text() == q.read_text()


def test_driver_checks_header_sections(tmp_path):
    """test_accuracy.cu:107-109's static_assert FILTER_N_SECTIONS ==
    N_SECTIONS, made at run time before any device call."""
    hdr, _, _ = RA.write(tmp_path, 4, 10)
    rc, out = run("accuracy", "--kernel", "PH", "--sections", 2, "--taps", hdr)
    assert rc == 255 and "FILTER_N_SECTIONS (4)" in out, out
    rc, out = run("accuracy", "--kernel", "PH", "--taps", tmp_path / "missing.hpp")
    assert rc == 255 and "cannot read" in out, out


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,extra,sections,rc", [
    ("PH", ["--block-size", 64, "--n-blocks", 64], 2, 0),
    ("PH", ["--block-size", 32, "--n-blocks", 64], 8, 0),
    ("STCR", ["--block-size", 32, "--n-blocks", 64], 4, 0),
    ("DTCR", ["--block-size", 64, "--n-blocks", 64], 2, 0),
    ("PLR", [], 1, 0),
    ("PLR", [], 2, 0),
    ("PLR", [], 4, 1),  # direct form in float32 at order 8: gate failed (6.5e-3), exit 1
    ("BS", ["--block-size", 64, "--n-blocks", 64], 8, 0),
])
def test_driver_on_reference_artifacts(device, tmp_path, kernel, extra, sections, rc):
    """The reference's artefacts (header with scipy's design and direct-form
    taps, float32 golden file) drive the driver as they drive
    test_accuracy.cu: N_SECTIONS from the header, PLR on the header's taps."""
    print(f"case driver artefacts {kernel} {extra} ns={sections}", flush=True)
    hdr, ref, _ = RA.write(tmp_path, sections, 20)
    got, out = run("accuracy", "--kernel", kernel, "--taps", hdr, "--reference", ref, "--log2n", 20, *extra)
    print(out)
    assert got == rc and ("PASSED" if rc == 0 else "FAILED") in out and f"N_SECTIONS   = {sections}" in out, out
