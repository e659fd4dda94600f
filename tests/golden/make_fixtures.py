"""Writes the known-answer fixtures of tests/golden/.

  python tests/golden/make_fixtures.py

sod_exact_t0.2.json   exact doubled-tube Sod solution (rho, u, p) at t = 0.2 on
                      2001 points of [-1, 1] (sod_exact.py), gamma = 1.4.
reference_values.json numbers held by the reference itself (printed notebook
                      outputs) and textbook star-region values for Sod; each
                      entry names its source.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import sod_exact  # noqa: E402


def main():
    x = np.linspace(-1.0, 1.0, 2001)
    sol = sod_exact.doubled_tube(x, 0.2)
    with open(os.path.join(HERE, "sod_exact_t0.2.json"), "w") as f:
        json.dump({"gamma": 1.4, "t": 0.2, "x": x.round(12).tolist(), "rho": sol[:, 0].tolist(),
                   "u": sol[:, 1].tolist(), "p": sol[:, 2].tolist()}, f)
    ps, us = sod_exact.star_state(1.0, 0.0, 1.0, 0.125, 0.0, 0.1, 1.4)
    print(f"p* = {ps:.6f}, u* = {us:.6f}")


if __name__ == "__main__":
    main()
