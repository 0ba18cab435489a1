"""GPU parity of the standalone compact operators (the Derivatives / Filter
drop-in, cs3d_ops_apply) against the oracle, in every direction, both
layouts, both boundary types, FP64 and FP32, including ragged line
lengths (segments of unequal length) and the smallest legal lines."""
import numpy as np
import pytest

import cs3d
import pyoracle as po
from helpers import relerr

pytestmark = pytest.mark.gpu

DIMS = [(16, 12, 10), (64, 64, 64), (42, 24, 32), (256, 32, 16), (8, 9, 12), (128, 8, 9)]
LENGTHS = (2.0, 1.5, 1.2)


def _to_ref(f, d):
    if d == 0:
        return f
    if d == 1:
        return np.ascontiguousarray(f.transpose(2, 0, 1))
    return np.ascontiguousarray(f.transpose(1, 2, 0))


def _from_ref(f, d):
    if d == 0:
        return f
    if d == 1:
        return f.transpose(1, 2, 0)
    return f.transpose(2, 0, 1)


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("bc", [cs3d.PERIODIC_SOLVE, cs3d.DIRICHLET_SOLVE])
@pytest.mark.parametrize("prec", [cs3d.FP64, cs3d.FP32])
def test_ops_match_oracle(dims, bc, prec):
    rng = np.random.default_rng(sum(dims) + bc)
    Nx, Ny, Nz = dims
    # FP32: about 3x the largest measured error (2.7e-6, second derivatives
    # of random fields)
    tol = 1e-12 if prec == cs3d.FP64 else 8e-6
    for d in range(3):
        n = dims[d]
        dh = (1.0 / (n - 1.0)) * LENGTHS[d]
        der = cs3d.Derivatives(dims, LENGTHS, bc, d, prec)
        flt = cs3d.Filter(0.49, dims, LENGTHS, bc, d, prec)
        f = rng.standard_normal((Nz, Ny, Nx))
        for kind in (1, 2, 0):
            if kind == 0 and n < 8:
                continue
            print(f"ops dims={dims} bc={bc} prec={prec} dir={d} kind={kind}", flush=True)
            ref = po.field_op(kind, d, bc, f, dh=dh, alpha_f=0.49)
            op = flt if kind == 0 else der
            got = op.apply(kind, f)
            got_ref_layout = op.apply(kind, _to_ref(f, d), cs3d.LAYOUT_REFERENCE)
            e1, e2 = relerr(got, ref), relerr(_from_ref(got_ref_layout, d), ref)
            print(f"  errors {e1:.2e} {e2:.2e}", flush=True)
            assert e1 < tol, (d, kind)
            assert e2 < tol, (d, kind)


def test_ops_reject_short_lines():
    with pytest.raises(cs3d.CS3DError):
        cs3d.Filter(0.49, (7, 16, 16), LENGTHS, cs3d.PERIODIC_SOLVE, 0)
