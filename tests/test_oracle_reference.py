"""The oracle against the reference's own compiled code (oracle/_ref/libref.so:
Utils.cpp, Derivatives.cpp, Filter.cpp and the header-only IdealGas, SpongeBC,
RK4, TVDRK3 classes, compiled from the reference checkout).  Skipped when the
checker library was not built (no reference checkout on this machine)."""
import ctypes

import numpy as np
import pytest

import pyoracle as po
from helpers import relerr

pytestmark = pytest.mark.skipif(not po.ref_built(),
                                reason="oracle/_ref/libref.so not built (no reference checkout)")


@pytest.fixture(scope="module")
def R():
    """loaded inside the tests, never at collection: a GPU test run maps
    nothing built from the reference"""
    return po.ref_lib()

_dp = ctypes.POINTER(ctypes.c_double)


def _p(a):
    return a.ctypes.data_as(_dp)


@pytest.mark.parametrize("n", [3, 8, 17, 64, 256])
def test_solve_tri_and_cyclic(R, n):
    rng = np.random.default_rng(n)
    a = rng.uniform(-1, 1, n)
    c = rng.uniform(-1, 1, n)
    b = 3.0 + rng.uniform(0, 1, n)
    d = rng.standard_normal(n)
    xr = np.zeros(n)
    R.ref_solve_tri(_p(a.copy()), _p(b.copy()), _p(c.copy()), _p(d.copy()), _p(xr), n)
    xo = po.solve_tri(a, b, c, d)
    assert relerr(xo, xr) < 1e-14
    al, be = 0.7, -0.4
    xr2 = np.zeros(n)
    R.ref_cyclic(_p(a.copy()), _p(b.copy()), _p(c.copy()), al, be, _p(d.copy()), n, _p(xr2))
    xo2 = po.cyclic(a, b, c, al, be, d)
    assert relerr(xo2, xr2) < 1e-14


def _to_ref(f, d):
    if d == 0:
        return f
    if d == 1:  # YZX: [i][k][j]
        return np.ascontiguousarray(f.transpose(2, 0, 1))
    return np.ascontiguousarray(f.transpose(1, 2, 0))  # ZXY: [j][i][k]


@pytest.mark.parametrize("dims", [(16, 12, 10), (8, 9, 11), (33, 20, 16)])
@pytest.mark.parametrize("kind", [1, 2, 0])
@pytest.mark.parametrize("bc", [0, 1])
def test_field_ops_match_reference(R, dims, kind, bc):
    """calc1stDerivField / calc2ndDerivField / filterField in all three
    directions, in the reference's layout contract, against the oracle."""
    Nx, Ny, Nz = dims
    L = (2.0, 1.5, 1.2)
    rng = np.random.default_rng(kind * 10 + bc)
    for d in range(3):
        f = rng.standard_normal((Nz, Ny, Nx))
        fr = _to_ref(f, d)
        outr = np.zeros_like(fr)
        R.ref_field_op(kind, d, bc, Nx, Ny, Nz, L[0], L[1], L[2], 0.49, _p(fr.copy()), _p(outr))
        n = dims[d]
        dh = (1.0 / (n - 1.0)) * L[d]
        # oracle in the reference layout and in XYZ order
        oref = po.field_op(kind, d, bc, fr.reshape(fr.shape), dh=dh, alpha_f=0.49, ref_layout=True) \
            if d == 0 else None
        o = po.field_op(kind, d, bc, f, dh=dh, alpha_f=0.49)
        assert relerr(_to_ref(o, d), outr) < 1e-13
        if oref is not None:
            assert relerr(oref, outr) < 1e-13


def test_domain_spacing(R):
    x, y, z, d = np.zeros(33), np.zeros(20), np.zeros(16), np.zeros(3)
    R.ref_domain(33, 20, 16, 2 * np.pi, 1.5, 1.2, _p(x), _p(y), _p(z), _p(d))
    assert d[0] == (1.0 / 32.0) * (2 * np.pi)
    assert np.array_equal(x, (np.arange(33) / 32.0) * (2 * np.pi))


def test_ideal_gas_relations(R):
    """IdealGas.hpp:49-79 against the oracle's primitives after set_ic."""
    rng = np.random.default_rng(3)
    n = 4096
    rho = 1 + 0.1 * rng.random(n)
    U, V, W = rng.standard_normal((3, n))
    p = 0.7 + 0.1 * rng.random(n)
    rhoE = p / 0.4 + 0.5 * rho * (U * U + V * V + W * W)
    outs = [np.zeros(n) for _ in range(6)]
    mu_ref = 1.7e-3
    R.ref_idealgas(n, mu_ref, _p(rho), _p(rho * U), _p(rho * V), _p(rho * W), _p(rhoE),
                   *[_p(o) for o in outs])
    cfg = po.make_config(16, 16, 16, mu_ref=mu_ref)
    s = po.Solver(cfg)
    sh = (16, 16, 16)
    s.set_ic(rho.reshape(sh), U.reshape(sh), V.reshape(sh), W.reshape(sh), p.reshape(sh))
    # oracle primitives: 3 T, 4 p, 5 mu, 6 Amu, 7 sos ; state rhoE
    assert relerr(s.prim(4).ravel(), outs[0]) < 1e-12
    assert relerr(s.prim(3).ravel(), outs[1]) < 1e-12
    assert relerr(s.prim(5).ravel(), outs[2]) < 1e-12
    assert relerr(s.prim(6).ravel(), outs[3]) < 1e-12
    assert relerr(s.prim(7).ravel(), outs[4]) < 1e-12
    assert relerr(s.field(4).ravel(), outs[5]) < 1e-14


@pytest.mark.parametrize("kinds", [(0, 0, 1, 1, 0, 0), (1, 1, 0, 0, 0, 0), (0, 1, 0, 0, 1, 1)])
def test_sponge_sigma(R, kinds):
    Nx, Ny, Nz = 32, 24, 24
    L = (8.0, 10.0, 4.0)
    sig = np.zeros(Nx * Ny * Nz)
    ki = (ctypes.c_int * 6)(*kinds)
    R.ref_sponge_sigma(Nx, Ny, Nz, L[0], L[1], L[2], ki, _p(sig))
    cfg = po.make_config(Nx, Ny, Nz, Lx=L[0], Ly=L[1], Lz=L[2], bc_type=(0, 1, 0),
                         bc_kind=kinds, sponge_frac=(0.25, 0.125, 0.25))
    s = po.Solver(cfg)
    so = s.sponge_sigma()
    assert so is not None
    assert np.abs(so.ravel() - sig).max() < 1e-14


@pytest.mark.parametrize("rk,stages", [(4, 4), (3, 3)])
def test_rk_updates(R, rk, stages):
    rng = np.random.default_rng(rk)
    N = 1000
    base = [rng.standard_normal(N) for _ in range(20)]
    a_ref = [x.copy() for x in base]
    a_or = [x.copy() for x in base]
    ptrs = (_dp * 20)(*[_p(x) for x in a_ref])
    for s in range(1, stages + 1):
        R.ref_rk_update(rk, s, N, 1, 1, ptrs)
        po.rk_update(rk, s, a_or)
        for i in range(20):
            assert np.array_equal(a_ref[i], a_or[i]), (s, i)
