"""GPU parity of the transposed compact operators (cs3d_ops_apply with the
CS3D_*_T kinds: the products the discrete adjoint applies) against the
transpose of the oracle's own line operator, taken as a dense matrix column
by column (po.line_op), in every direction, both layouts, both boundary
types (non-periodic closures included), FP64 and FP32, on uniform and ragged
lines.  The dot-product identity <M f, g> = <f, M^T g> ties each transposed
operator to the forward one the library applies."""
import numpy as np
import pytest

import cs3d
import pyoracle as po
from helpers import relerr

pytestmark = pytest.mark.gpu

# non-periodic transposes need 18 points per line (24 for the filter)
DIMS = [(32, 24, 40), (64, 64, 64), (42, 30, 26), (256, 32, 24), (128, 36, 28)]
LENGTHS = (2.0, 1.5, 1.2)
ALPHA_F = 0.49


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


def dense(kind, bc, n, dh):
    """the oracle's line operator as an n x n matrix (column j = op(e_j))"""
    return np.array([po.line_op(kind, bc, e, dh, ALPHA_F) for e in np.eye(n)]).T


def along(Mat, f, d):
    """Mat applied to every line of an (Nz, Ny, Nx) field along direction d"""
    axis = 2 - d
    return np.moveaxis(np.tensordot(Mat, f, axes=([1], [axis])), 0, axis)


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("bc", [cs3d.PERIODIC_SOLVE, cs3d.DIRICHLET_SOLVE])
@pytest.mark.parametrize("prec", [cs3d.FP64, cs3d.FP32])
def test_transposed_ops_match_oracle_transpose(dims, bc, prec):
    rng = np.random.default_rng(7 * sum(dims) + bc)
    Nx, Ny, Nz = dims
    # FP32: about 3x the largest measured error (3.3e-6)
    tol = 1e-12 if prec == cs3d.FP64 else 1e-5
    for d in range(3):
        n = dims[d]
        dh = (1.0 / (n - 1.0)) * LENGTHS[d]
        der = cs3d.Derivatives(dims, LENGTHS, bc, d, prec)
        flt = cs3d.Filter(ALPHA_F, dims, LENGTHS, bc, d, prec)
        g = rng.standard_normal((Nz, Ny, Nx))
        f = rng.standard_normal((Nz, Ny, Nx))
        for kind in (1, 2, 0):
            print(f"opsT dims={dims} bc={bc} prec={prec} dir={d} kind={kind}", flush=True)
            Mat = dense(kind, bc, n, dh)
            ref = along(Mat.T, g, d)
            op = flt if kind == 0 else der
            got = op.apply(kind + 3, g)
            got_ref_layout = _from_ref(op.apply(kind + 3, _to_ref(g, d), cs3d.LAYOUT_REFERENCE), d)
            e1, e2 = relerr(got, ref), relerr(got_ref_layout, ref)
            # <M f, g> = <f, M^T g> with the library's forward operator
            lhs = float(np.sum(op.apply(kind, f).astype(np.float64) * g))
            rhs = float(np.sum(f * got.astype(np.float64)))
            e3 = abs(lhs - rhs) / (abs(lhs) + np.sqrt(f.size) * np.abs(ref).max())
            print(f"  errors {e1:.2e} {e2:.2e} dot {e3:.2e}", flush=True)
            assert e1 < tol, (d, kind, e1)
            assert e2 < tol, (d, kind, e2)
            assert e3 < (1e-13 if prec == cs3d.FP64 else 1e-5), (d, kind, e3)


def test_transposed_ops_refuse_short_nonperiodic_lines():
    der = cs3d.Derivatives((16, 24, 24), LENGTHS, cs3d.DIRICHLET_SOLVE, 0)
    with pytest.raises(cs3d.CS3DError):
        der.calc1stDerivFieldTransposed(np.zeros((24, 24, 16)))
    flt = cs3d.Filter(ALPHA_F, (20, 24, 24), LENGTHS, cs3d.DIRICHLET_SOLVE, 0)
    with pytest.raises(cs3d.CS3DError):
        flt.filterFieldTransposed(np.zeros((24, 24, 20)))
