"""Writes the golden fixtures of tests/golden/ from libraries (no reference code runs):

norminv.json   scipy.special.ndtri at the probabilities of the reference's
               norminv test (src/spectrometer/erf.rs:142-161), with its rmax
               tolerances.
sellmeier.json scipy.optimize.curve_fit of a degree-5 Horner polynomial to
               Sellmeier1 n(l) on numpy.linspace(lo, hi, 100) with p0 = ones,
               the exact call of compound_prism_designer/glasscat.py:114-127,
               for seeded (B, C) draws in the benchmark's ranges.

Run: python tests/golden/make_golden.py
"""
import json
from functools import reduce
from pathlib import Path

import numpy as np
from scipy.optimize import curve_fit
from scipy.special import ndtri

HERE = Path(__file__).resolve().parent

NORMINV_POINTS = [  # (u, rmax) from erf.rs:146-160
    (1e-100, 2e-1), (1e-60, 2e-2), (1e-30, 1e-3), (1e-20, 1e-5), (1e-15, 5e-9), (1e-10, 5e-9),
    (1e-5, 2e-9), (0.1, 1e-9), (0.2, 1e-9), (0.5, 1e-9), (0.7, 1e-9), (0.9, 1e-9), (0.99, 2.04e-9),
    (0.999, 2.5e-9), (0.9999, 4e-9),
]


def sellmeier1(b, c, w):
    w2 = w * w
    return np.sqrt(1 + b[0] * w2 / (w2 - c[0]) + b[1] * w2 / (w2 - c[1]) + b[2] * w2 / (w2 - c[2]))


def main():
    norm = [{"u": u, "z": float(ndtri(u)), "rmax": r} for u, r in NORMINV_POINTS]
    (HERE / "norminv.json").write_text(json.dumps(norm, indent=1))

    rng = np.random.default_rng(20240611)
    cases = []
    for _ in range(8):
        b = rng.uniform(0.3, 1.5, 3)
        c = rng.uniform(0.001, 0.2, 3)
        lo, hi = 0.5, 0.82
        x = np.linspace(lo, hi, 100)
        y = sellmeier1(b, c, x)
        f = lambda x, *coeffs: reduce(lambda a, cc: a * x + cc, coeffs)
        popt, _ = curve_fit(f, x, y, p0=np.ones(6))
        r = y - f(x, *popt)
        xc = np.linspace(lo, hi, 17)
        cases.append({"b": b.tolist(), "c": c.tolist(), "lo": lo, "hi": hi, "coeffs": popt.tolist(),
                      "chisq": float(np.sum(r ** 2)), "check_w": xc.tolist(),
                      "check_n_fit": f(xc, *popt).tolist(), "check_n_sellmeier": sellmeier1(b, c, xc).tolist()})
    (HERE / "sellmeier.json").write_text(json.dumps(cases, indent=1))


if __name__ == "__main__":
    main()
