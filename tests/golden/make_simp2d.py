"""Writes tests/golden/simp2d_160x40.npz: the oracle's SIMP run on the 2D
point-load cantilever, 160 x 40 unit Q4 cells, E0 = 1, Emin = 1e-9, nu = 0.3
(plane strain), p = 3, volfrac 0.5, density filter rmin = 2.0, OC move 0.2,
rho = 0.5 at the start, 100 iterations, each displacement from the oracle's
banded Cholesky (direct) solve.  The GPU test replays the run and compares
the compliance history and the final design.

Run:  python tests/golden/make_simp2d.py     (about a minute on 8 cores)
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def main():
    import _pkg
    O = _pkg.load_oracle()
    O.build()
    nels, volfrac, rmin, iters = (160, 40), 0.5, 2.0, 100
    r = O.simp(O.Cantilever(nels), volfrac, rmin, iters, p=3.0, xmin=1e-9, solver="direct")
    np.savez_compressed(os.path.join(HERE, "simp2d_160x40.npz"), nels=np.array(nels),
                        volfrac=volfrac, rmin=rmin, iters=iters, obj=r["obj"], x=r["x"],
                        xphys=r["xphys"])
    print(f"obj[0] {r['obj'][0]:.10g} obj[-1] {r['obj'][-1]:.10g}")


if __name__ == "__main__":
    main()
