"""Writes tests/golden/scipy_fixtures.npz: scipy's float64 Butterworth designs
and sosfilt outputs for the reference's conventions, frozen so that the
oracle and the C++ design are pinned against scipy 1.15.3 values.

Imports scipy only (not the reference).  Contents:
  sos_o{order}_w{wn}      butter(order, wn, 'sos') float64       (ref_generate.py:62)
  ba_b_o{..}, ba_a_o{..}  butter(order, wn, 'ba') float64         (ref_generate.py:74)
  x_ref_o{order}          reference workload input, 3 chunks of 32x32:
                          f32(gain * cos(2 pi 120 i / 1000))      (ref_generate.py:84-86)
  y_ref_o{order}          sosfilt(f64(sos_f32), f64(x_ref))      (ref_generate.py:89)
  cheby{k}_o{order}_d{db}_w{wn}  cheby1 (k = 1, ripple db) / cheby2 (k = 2,
                          attenuation db)(order, db, wn, 'sos') float64
Run: python tests/golden/make_fixtures.py
"""
import os

import numpy as np
import scipy
from scipy import signal

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scipy_fixtures.npz")
ORDERS = (2, 4, 8, 16)
WNS = (0.02, 0.1, 0.2, 0.4)


def norm_sos(sos):
    g = float(np.prod(sos[:, 0]))
    s = sos.copy()
    for k in range(s.shape[0]):
        s[k, 0:3] = sos[k, 0:3] / sos[k, 0]
    return s.astype(np.float32), g


def main():
    d = {"scipy_version": np.array(scipy.__version__)}
    for o in ORDERS:
        for wn in WNS:
            key = f"o{o}_w{wn}"
            d["sos_" + key] = signal.butter(o, wn, output="sos")
            b, a = signal.butter(o, wn, output="ba")
            d["ba_b_" + key] = b
            d["ba_a_" + key] = a
        sos = signal.butter(o, 0.2, output="sos")
        sos32, g = norm_sos(sos)
        n = 3 * 32 * 32
        i = np.arange(n, dtype=np.float64)
        x = (g * np.cos(2.0 * np.pi * 120.0 * i / 1000.0)).astype(np.float32)
        d[f"x_ref_o{o}"] = x
        d[f"y_ref_o{o}"] = signal.sosfilt(sos32.astype(np.float64), x.astype(np.float64))
    for k, design, dbs in ((1, signal.cheby1, (0.5, 3.0)), (2, signal.cheby2, (40.0, 80.0))):
        for o in ORDERS:
            for db in dbs:
                for wn in (0.02, 0.1, 0.3):
                    d[f"cheby{k}_o{o}_d{db}_w{wn}"] = design(o, db, wn, output="sos")
    np.savez_compressed(OUT, **d)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
