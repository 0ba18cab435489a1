"""Pins of the numpy adjoint reference (oracle/adjoint.py) with sponge faces
and non-periodic lines, on the CPU.

The forward map (stages with the sponge source and the zeroed face planes,
the running averages and the plane reset, the stale primitives at the
first stage after each update) must reproduce the C oracle step for step,
averages included; its reverse must satisfy the dot-product identity
against central differences in each of its inputs (stage state, stale
state, averages; the sponge update's own reverse too), and the kinetic-
energy gradient through several steps must match central differences of
whole C-oracle runs with dt held fixed.  A reverse that sends the stale
primitives' cotangents to the stage state instead fails that last check."""
import numpy as np
import pytest

import adjoint as oa
import pyoracle as po
from helpers import random_ic, relerr, shear_ic, smooth

DIMS = (10, 24, 8)   # x, z periodic; y non-periodic with sponge faces


def cfg_for(dims=DIMS, **kw):
    Nx, Ny, Nz = dims
    base = dict(Lx=2 * np.pi * (1 - 1 / Nx), Ly=6.0, Lz=2 * np.pi * (1 - 1 / Nz),
                alpha_f=0.45, mu_ref=2e-2, Pr=0.71, rk=4, filter_step=1,
                bc_type=(po.PERIODIC_SOLVE, po.DIRICHLET_SOLVE, po.PERIODIC_SOLVE),
                bc_kind=(po.PERIODIC, po.PERIODIC, po.SPONGE, po.SPONGE, po.PERIODIC,
                         po.PERIODIC),
                sponge_frac=(0.25, 0.25, 0.25), sponge_avgT=0.5)
    base.update(kw)
    return po.make_config(Nx, Ny, Nz, **base)


def shape_of(dims):
    return (dims[2], dims[1], dims[0])


def ic_for(dims, seed):
    # a shear layer across the sponges, so that the averages, the plane
    # reset and the stale primitives all see flow
    return shear_ic(dims[0], dims[1], dims[2], 6.0, seed, dU=0.6, amp=0.1)


def oracle_run(cfg, ic, dt, nsteps):
    o = po.Solver(cfg)
    o.set_ic(*ic)
    o.set_dt(dt)
    for _ in range(nsteps):
        o.step()
    return o


@pytest.mark.parametrize("rk,filter_step", [(4, 1), (4, 2), (3, 1)])
def test_forward_with_sponges_matches_oracle(rk, filter_step):
    cfg = cfg_for(rk=rk, filter_step=filter_step)
    ic = ic_for(DIMS, 3)
    dt = 0.02
    A = oa.Adjoint(cfg)
    q0 = A.conserved_from_ic(*ic)
    for n in (1, 2, 4):
        q, avg, _ = A.run(q0, n, dt, filter_step)
        o = oracle_run(cfg, ic, dt, n)
        errs = [relerr(g, r) for g, r in zip(q, o.conserved(1))]
        errs += [relerr(g, r) for g, r in zip(avg, o.sponge_avg())]
        o.close()
        assert max(errs) <= 1e-12, (n, errs)


def test_stage_vjp_dot_product_in_every_input():
    cfg = cfg_for()
    shape = shape_of(DIMS)
    A = oa.Adjoint(cfg)
    rng = np.random.default_rng(4)
    q = A.conserved_from_ic(*ic_for(DIMS, 5))
    qs = [a + 0.01 * smooth(rng.standard_normal(shape), 1) for a in q]
    avg = [a + 0.02 * smooth(rng.standard_normal(shape), 1) for a in q]
    lam = [rng.standard_normal(shape) for _ in range(5)]
    dt, eps = 0.01, 1e-5
    qb, qsb, Ab = A.rhs_vjp_full(q, lam, dt, avg, qs)
    for which in range(3):
        delta = [s * smooth(rng.standard_normal(shape), 1) for s in (1e-2, 0.1, 0.1, 0.1, 0.1)]
        args = [q, qs, avg]

        def shifted(sign):
            a = list(args)
            a[which] = [x + sign * eps * d for x, d in zip(args[which], delta)]
            return A.rhs(a[0], dt, a[2], a[1])

        kp, km = shifted(1), shifted(-1)
        fd = sum(np.vdot(l_, (a - b) / (2 * eps)) for l_, a, b in zip(lam, kp, km))
        ad = sum(np.vdot(a, b) for a, b in zip((qb, qsb, Ab)[which], delta))
        assert abs(ad - fd) <= 1e-8 * abs(fd), (which, ad, fd)


def test_sponge_update_vjp_dot_product():
    cfg = cfg_for()
    shape = shape_of(DIMS)
    A = oa.Adjoint(cfg)
    rng = np.random.default_rng(6)
    qf = A.conserved_from_ic(*ic_for(DIMS, 7))
    avg = [a + 0.02 * smooth(rng.standard_normal(shape), 1) for a in qf]
    lq = [rng.standard_normal(shape) for _ in range(5)]
    lA = [rng.standard_normal(shape) for _ in range(5)]
    dt, eps = 0.02, 1e-6
    lqf, lAo = A.sponge_update_vjp(qf, avg, dt, lq, lA)
    dq = [0.1 * smooth(rng.standard_normal(shape), 1) for _ in range(5)]
    dA = [0.1 * smooth(rng.standard_normal(shape), 1) for _ in range(5)]
    qp, ap = A.sponge_update([x + eps * d for x, d in zip(qf, dq)],
                             [x + eps * d for x, d in zip(avg, dA)], dt)
    qm, am = A.sponge_update([x - eps * d for x, d in zip(qf, dq)],
                             [x - eps * d for x, d in zip(avg, dA)], dt)
    fd = sum(np.vdot(l_, (a - b) / (2 * eps)) for l_, a, b in zip(lq + lA, qp + ap, qm + am))
    ad = sum(np.vdot(a, b) for a, b in zip(lqf + lAo, dq + dA))
    assert abs(ad - fd) <= 1e-8 * abs(fd), (ad, fd)


def run_ke(cfg, ic, dt, nsteps):
    o = oracle_run(cfg, ic, dt, nsteps)
    rho = o.field(0)
    u = [o.field(i) / rho for i in (1, 2, 3)]
    o.close()
    return float(np.sum(0.5 * rho * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)) / rho.size)


def fd_check(A, cfg, ic, dt, nsteps, filter_step):
    """worst relative error of the gradient along two smooth directions"""
    J, gU, gV, gW = A.ke_gradient(ic, nsteps, dt, filter_step)
    J0 = run_ke(cfg, ic, dt, nsteps)
    assert abs(J - J0) <= 1e-13 * abs(J0)
    rng = np.random.default_rng(12)
    eps = 1e-5
    rho, U, V, W, p = ic
    worst = 0.0
    for _ in range(2):
        d = [smooth(rng.standard_normal(U.shape), 1) for _ in range(3)]
        Jp = run_ke(cfg, (rho, U + eps * d[0], V + eps * d[1], W + eps * d[2], p), dt, nsteps)
        Jm = run_ke(cfg, (rho, U - eps * d[0], V - eps * d[1], W - eps * d[2], p), dt, nsteps)
        fd = (Jp - Jm) / (2 * eps)
        ad = np.vdot(gU, d[0]) + np.vdot(gV, d[1]) + np.vdot(gW, d[2])
        worst = max(worst, abs(ad - fd) / abs(fd))
    return worst


@pytest.mark.parametrize("rk,filter_step", [(4, 1), (4, 2), (3, 1)])
def test_ke_gradient_with_sponges_matches_oracle_finite_differences(rk, filter_step):
    cfg = cfg_for(rk=rk, filter_step=filter_step)
    ic = ic_for(DIMS, 9)
    A = oa.Adjoint(cfg)
    err = fd_check(A, cfg, ic, 0.02, 3, filter_step)
    assert err <= 1e-7, err


def test_ignoring_the_stale_primitives_fails_the_dot_product():
    """a plausible mistake: the reverse treats the first stage after a sponge
    update as an ordinary one (the x block's cotangents chained into the
    stage state).  The identity in the stage state's direction then fails
    by far more than its tolerance.  (Over a few whole steps the two routes
    differ only at second order: the averages start as the initial state,
    so the plane values and the averages respond to it alike.)"""
    cfg = cfg_for()
    shape = shape_of(DIMS)
    A = oa.Adjoint(cfg)
    rng = np.random.default_rng(4)
    q = A.conserved_from_ic(*ic_for(DIMS, 5))
    qs = [a + 0.01 * smooth(rng.standard_normal(shape), 1) for a in q]
    avg = [a + 0.02 * smooth(rng.standard_normal(shape), 1) for a in q]
    lam = [rng.standard_normal(shape) for _ in range(5)]
    dt, eps = 0.01, 1e-5
    bad = A.rhs_vjp_full(q, lam, dt, avg, None)[0]
    delta = [s * smooth(rng.standard_normal(shape), 1) for s in (1e-2, 0.1, 0.1, 0.1, 0.1)]
    kp = A.rhs([x + eps * d for x, d in zip(q, delta)], dt, avg, qs)
    km = A.rhs([x - eps * d for x, d in zip(q, delta)], dt, avg, qs)
    fd = sum(np.vdot(l_, (a - b) / (2 * eps)) for l_, a, b in zip(lam, kp, km))
    ad = sum(np.vdot(a, b) for a, b in zip(bad, delta))
    assert abs(ad - fd) > 1e-3 * abs(fd), (ad, fd)
