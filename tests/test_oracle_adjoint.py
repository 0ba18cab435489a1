"""Pins of the numpy adjoint reference (oracle/adjoint.py) on the CPU.

The reference has no adjoint, so the pins come from the oracle's own
forward path and from the mathematics: the numpy forward stage and RK4
step reproduce the C oracle, the dot-product identity holds against
central differences of the forward stage, the kinetic-energy gradient
matches central differences of whole C-oracle runs (dt held fixed), and
the line operators have the symmetries the GPU adjoint relies on
(periodic D1 skew-adjoint, D2 and the filter self-adjoint)."""
import numpy as np
import pytest

import adjoint as oa
import pyoracle as po
from helpers import random_ic, relerr, smooth

DIMS = (10, 8, 12)


def cfg_for(dims, **kw):
    Nx, Ny, Nz = dims
    base = dict(Lx=2 * np.pi * (1 - 1 / Nx), Ly=2 * np.pi * (1 - 1 / Ny),
                Lz=2 * np.pi * (1 - 1 / Nz), alpha_f=0.45, mu_ref=2e-2, Pr=0.71, rk=4,
                filter_step=1)
    base.update(kw)
    return po.make_config(Nx, Ny, Nz, **base)


def oracle_state(cfg, ic, dt):
    o = po.Solver(cfg)
    o.set_ic(*ic)
    o.set_dt(dt)
    return o


def shape_of(dims):
    return (dims[2], dims[1], dims[0])


def test_line_operator_symmetries():
    cfg = cfg_for((12, 9, 16))
    A = oa.Adjoint(cfg)
    for d in range(3):
        D1, D2, F = A.M[1, d], A.M[2, d], A.M[0, d]
        assert np.max(np.abs(D1 + D1.T)) <= 1e-12 * np.max(np.abs(D1))
        assert np.max(np.abs(D2 - D2.T)) <= 1e-12 * np.max(np.abs(D2))
        assert np.max(np.abs(F - F.T)) <= 1e-13
        # the filter passes constants and the derivatives annihilate them
        one = np.ones(D1.shape[0])
        assert np.max(np.abs(F @ one - one)) <= 1e-13
        assert np.max(np.abs(D1 @ one)) <= 1e-11 and np.max(np.abs(D2 @ one)) <= 1e-9


def test_forward_stage_matches_oracle():
    cfg = cfg_for(DIMS)
    ic = random_ic(shape_of(DIMS), seed=2, amp=0.4)
    dt = 0.013
    o = oracle_state(cfg, ic, dt)
    o.pre_step()
    assert o.dt == dt
    o.rhs(1)
    ref = o.conserved(3)
    got = oa.Adjoint(cfg).rhs(o.conserved(1), dt)
    errs = [relerr(g, r) for g, r in zip(got, ref)]
    assert max(errs) <= 1e-12, errs


def test_rk4_step_matches_oracle():
    cfg = cfg_for(DIMS)
    ic = random_ic(shape_of(DIMS), seed=4, amp=0.4)
    dt = 0.02
    o = oracle_state(cfg, ic, dt)
    A = oa.Adjoint(cfg)
    q = o.conserved(1)
    for n in range(1, 4):
        o.step()
        q = A.rk4_step(q, dt, A.filter_order(n))
        errs = [relerr(g, r) for g, r in zip(q, o.conserved(1))]
        assert max(errs) <= 1e-12, (n, errs)


def test_tvdrk3_step_matches_oracle():
    """TVDRK3.hpp:38-96 in numpy against the C oracle's TVDRK3 steps"""
    cfg = cfg_for(DIMS, rk=3)
    ic = random_ic(shape_of(DIMS), seed=6, amp=0.4)
    dt = 0.02
    o = oracle_state(cfg, ic, dt)
    A = oa.Adjoint(cfg)
    q = o.conserved(1)
    for n in range(1, 4):
        o.step()
        q = A.step(q, dt, A.filter_order(n))
        errs = [relerr(g, r) for g, r in zip(q, o.conserved(1))]
        assert max(errs) <= 1e-12, (n, errs)


def test_stage_vjp_dot_product():
    cfg = cfg_for(DIMS)
    shape = shape_of(DIMS)
    ic = random_ic(shape, seed=7, amp=0.4)
    A = oa.Adjoint(cfg)
    q = A.conserved_from_ic(*ic)
    rng = np.random.default_rng(1)
    lam = [rng.standard_normal(shape) for _ in range(5)]
    delta = [s * smooth(rng.standard_normal(shape), 1) for s in (1e-2, 0.1, 0.1, 0.1, 0.1)]
    dt = 0.01
    lhs = sum(np.vdot(a, b) for a, b in zip(A.rhs_vjp(q, lam, dt), delta))
    eps = 1e-5
    kp = A.rhs([a + eps * d for a, d in zip(q, delta)], dt)
    km = A.rhs([a - eps * d for a, d in zip(q, delta)], dt)
    rhs = sum(np.vdot(l_, (a - b) / (2 * eps)) for l_, a, b in zip(lam, kp, km))
    assert abs(lhs - rhs) <= 1e-8 * abs(rhs), (lhs, rhs)


def test_dot_product_catches_a_wrong_seed():
    """the identity fails, by far more than its tolerance, for a reverse
    with one term of the energy equation dropped (the forward intact)"""
    cfg = cfg_for(DIMS)
    shape = shape_of(DIMS)
    A = oa.Adjoint(cfg)
    q = A.conserved_from_ic(*random_ic(shape, seed=7, amp=0.4))
    rng = np.random.default_rng(1)
    lam = [rng.standard_normal(shape) for _ in range(5)]
    delta = [0.1 * smooth(rng.standard_normal(shape), 1) for _ in range(5)]
    dt, eps = 0.01, 1e-5
    kp = A.rhs([a + eps * d for a, d in zip(q, delta)], dt)
    km = A.rhs([a - eps * d for a, d in zip(q, delta)], dt)
    fd = sum(np.vdot(l_, (a - b) / (2 * eps)) for l_, a, b in zip(lam, kp, km))
    saved = oa.RHS_TERMS[4]
    try:
        oa.RHS_TERMS[4] = [t for t in saved if t[1] != ("MuX", "Tx")]
        bad = sum(np.vdot(a, b) for a, b in zip(A.rhs_vjp(q, lam, dt), delta))
    finally:
        oa.RHS_TERMS[4] = saved
    assert abs(bad - fd) > 1e-5 * abs(fd), (bad, fd)


def run_ke(cfg, ic, dt, nsteps):
    o = oracle_state(cfg, ic, dt)
    for _ in range(nsteps):
        o.step()
    rho = o.field(0)
    u = [o.field(i) / rho for i in (1, 2, 3)]
    return float(np.sum(0.5 * rho * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)) / rho.size)


@pytest.mark.parametrize("filter_step,rk", [(1, 4), (2, 4), (1, 3), (2, 3)])
def test_ke_gradient_matches_oracle_finite_differences(filter_step, rk):
    dims = DIMS
    shape = shape_of(dims)
    cfg = cfg_for(dims, filter_step=filter_step, rk=rk)
    ic = random_ic(shape, seed=9, amp=0.5)
    dt, nsteps = 0.02, 3
    A = oa.Adjoint(cfg)
    J, gU, gV, gW = A.ke_gradient(ic, nsteps, dt, filter_step)
    J0 = run_ke(cfg, ic, dt, nsteps)
    assert abs(J - J0) <= 1e-13 * abs(J0)
    rng = np.random.default_rng(12)
    eps = 1e-5
    rho, U, V, W, p = ic
    for _ in range(2):
        d = [smooth(rng.standard_normal(shape), 1) for _ in range(3)]
        Jp = run_ke(cfg, (rho, U + eps * d[0], V + eps * d[1], W + eps * d[2], p), dt, nsteps)
        Jm = run_ke(cfg, (rho, U - eps * d[0], V - eps * d[1], W - eps * d[2], p), dt, nsteps)
        fd = (Jp - Jm) / (2 * eps)
        ad = np.vdot(gU, d[0]) + np.vdot(gV, d[1]) + np.vdot(gW, d[2])
        assert abs(ad - fd) <= 1e-7 * abs(fd), (ad, fd)
