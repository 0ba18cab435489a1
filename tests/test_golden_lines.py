"""The compact operators against the reference's own outputs
(tests/golden/line_operators.npz, written by make_line_operators.py from
the reference's Derivatives.cpp / Filter.cpp as compiled into
oracle/_ref): first and second derivatives and the filter on x lines of
8, 16, 33 and 64 points, periodic and non-periodic, alpha_f 0.49 and 0.495.
The oracle on the CPU; the GPU operators (cs3d_ops_apply) in FP64 and FP32
on the GPU."""
import os

import numpy as np
import pytest

import pyoracle as po
from helpers import relerr

HERE = os.path.dirname(os.path.abspath(__file__))
VEC = np.load(os.path.join(HERE, "golden", "line_operators.npz"))
NS = (8, 16, 33, 64)
LX = 1.7
CASES = [(kind, bc, n, af) for n in NS for bc in (0, 1) for kind in (1, 2, 0)
         for af in ((0.49, 0.495) if kind == 0 else (0.49,))]


def key(kind, bc, n, af):
    return f"k{kind}_b{bc}_n{n}_a{int(round(af * 1000))}"


@pytest.mark.parametrize("kind,bc,n,af", CASES)
def test_oracle_matches_reference_vectors(kind, bc, n, af):
    f = VEC[f"in_n{n}"]
    got = po.field_op(kind, 0, bc, f, dh=(1.0 / (n - 1.0)) * LX, alpha_f=af)
    assert relerr(got, VEC[key(kind, bc, n, af)]) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("n", NS)
@pytest.mark.parametrize("bc", [0, 1])
def test_gpu_operators_match_reference_vectors(n, bc):
    import cs3d
    f = VEC[f"in_n{n}"]
    dims = (n, f.shape[1], f.shape[0])
    for prec, tol in ((cs3d.FP64, 1e-12), (cs3d.FP32, 8e-6)):
        der = cs3d.Derivatives(dims, (LX, 1.0, 1.0), bc, 0, prec)
        for kind in (1, 2):
            e = relerr(der.apply(kind, f), VEC[key(kind, bc, n, 0.49)])
            print(f"n={n} bc={bc} prec={prec} kind={kind}: {e:.2e}", flush=True)
            assert e < tol
        for af in (0.49, 0.495):
            flt = cs3d.Filter(af, dims, (LX, 1.0, 1.0), bc, 0, prec)
            e = relerr(flt.apply(0, f), VEC[key(0, bc, n, af)])
            print(f"n={n} bc={bc} prec={prec} filter af={af}: {e:.2e}", flush=True)
            assert e < tol
