"""The reference harness's two artefacts, written the reference's way (test
infrastructure): build/filter_taps.hpp and build/reference.bin as
tools/ref_generate.py:56-131 produces them -- scipy's FP64 Butterworth
design in SOS and direct form, the b0 == 1 normalisation, the float32 input
FILTER_GAIN * cos(2 pi 120 i / 1000), and the golden output (the oracle's
FP64 cascade over the float32 input and sections, ref_generate.py:88-99).

The header text follows ref_generate.py:106-129: floats '{v:.9e}f',
doubles repr().  Only the comment lines differ."""
import numpy as np
from scipy import signal

import oracle_lib as O

INPUT_FS = 1000.0
INPUT_FREQ = 120.0


def design(n_sections, wn=0.2):
    """The values ref_generate.py:61-81 puts in the header (scipy, FP64)."""
    sos = signal.butter(2 * n_sections, wn, btype="lowpass", output="sos")
    gain = float(np.prod(sos[:, 0]))
    norm = sos.copy()
    norm[:, 0:3] = sos[:, 0:3] / sos[:, [0]]
    b, a = signal.butter(2 * n_sections, wn, btype="lowpass", output="ba")
    return {
        "n_sections": n_sections, "plr_order": 2 * n_sections, "gain": gain, "input_fs": INPUT_FS,
        "input_freq": INPUT_FREQ, "cutoff_wn": wn, "sos": norm.astype(np.float32),
        "plr_b": b[1:] / b[0], "plr_a": -a[1:],
    }


def header_text(t):
    """filter_taps.hpp for the values `t` (ref_generate.py:101-129 layout)."""
    ns, k = int(t["n_sections"]), int(t["plr_order"])
    sos = np.asarray(t["sos"], np.float32).reshape(ns, 6)
    rows = ",\n".join("    {" + ", ".join(f"{v:.9e}f" for v in sos[s]) + "}" for s in range(ns))
    lines = [
        "// reference layout (tests/ref_artifacts.py)",
        "#pragma once",
        "",
        f"#define FILTER_N_SECTIONS {ns}",
        f"static const double FILTER_GAIN = {float(t['gain'])!r};",
        f"static const double INPUT_FS    = {float(t['input_fs'])!r};",
        f"static const double INPUT_FREQ  = {float(t['input_freq'])!r};",
        "",
        f"static const float FILTER_SOS[{ns}][6] = {{",
        rows,
        "};",
        "",
        f"#define PLR_FILTER_ORDER {k}",
        f"static const double PLR_FILTER_B[{k}] = {{",
        "    " + ",\n    ".join(repr(float(v)) for v in t["plr_b"]),
        "};",
        f"static const double PLR_FILTER_A[{k}] = {{",
        "    " + ",\n    ".join(repr(float(v)) for v in t["plr_a"]),
        "};",
    ]
    return "\n".join(lines) + "\n"


def declarations(text):
    """The header's lines other than comments and blank lines."""
    return [ln for ln in text.splitlines() if ln.strip() and not ln.lstrip().startswith("//")]


def input_signal(t, n):
    i = np.arange(n, dtype=np.float64)
    return (float(t["gain"]) * np.cos(2.0 * np.pi * float(t["input_freq"]) * i / float(t["input_fs"]))).astype(
        np.float32)


def write(directory, n_sections, log2n, wn=0.2):
    """Write filter_taps.hpp and reference.bin into `directory`; returns
    (header path, reference path, values)."""
    t = design(n_sections, wn)
    hdr = directory / "filter_taps.hpp"
    hdr.write_text(header_text(t))
    x = input_signal(t, 1 << log2n)
    ref = O.golden(t["sos"], x).astype(np.float32)
    binp = directory / "reference.bin"
    ref.tofile(binp)
    return hdr, binp, t
