"""Writes line_operators.npz: outputs of the reference's own compact
operators (Derivatives::calc1stDerivField, calc2ndDerivField and
Filter::filterField, Derivatives.cpp:3-96, Filter.cpp:129-171), run from
the reference's sources as compiled by oracle/Makefile into
oracle/_ref/libref.so, on seeded inputs: x lines (3 x 2 of them) of
N = 8, 16, 33 and 64 points, periodic and non-periodic, first and second
derivatives and the filter at alpha_f = 0.49 and 0.495, line length 1.7.

The vectors are data (inputs and the reference's outputs); the reference
itself is not part of the repository.  Run where oracle/_ref was built:

    python tests/golden/make_line_operators.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import pyoracle as po  # noqa: E402

NS = (8, 16, 33, 64)
ALPHAS = (0.49, 0.495)
LX = 1.7
NY, NZ = 3, 2


def key(kind, bc, n, af):
    return f"k{kind}_b{bc}_n{n}_a{int(round(af * 1000))}"


def main():
    R = po.ref_lib()
    if R is None:
        sys.exit("oracle/_ref/libref.so is not built (make -C oracle ref)")
    dp = ctypes.POINTER(ctypes.c_double)
    out = {}
    for n in NS:
        rng = np.random.default_rng(n)
        f = rng.standard_normal((NZ, NY, n))
        out[f"in_n{n}"] = f
        for bc in (0, 1):
            for kind in (1, 2, 0):
                for af in (ALPHAS if kind == 0 else ALPHAS[:1]):
                    g = np.zeros_like(f)
                    R.ref_field_op(kind, 0, bc, n, NY, NZ, LX, 1.0, 1.0, af,
                                   f.copy().ctypes.data_as(dp), g.ctypes.data_as(dp))
                    out[key(kind, bc, n, af)] = g
    np.savez_compressed(os.path.join(HERE, "line_operators.npz"), **out)


if __name__ == "__main__":
    main()
