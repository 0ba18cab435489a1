"""Pins of the numpy adjoint reference (oracle/adjoint.py) with adiabatic and
moving walls, on the CPU: its forward (the wall correction of every stage
state, the first one in place, and the continuity-only wall rows of the
right-hand side) reproduces the C oracle step for step; its reverse
satisfies the dot-product identity for the stage map with wall rows and for
the wall correction itself; the kinetic-energy gradient through several
steps matches central differences of whole C-oracle runs (dt fixed).  A
reverse that leaves out the wall temperature's dependence on its six inward
neighbours fails the correction's identity."""
import numpy as np
import pytest

import adjoint as oa
import pyoracle as po
from helpers import relerr, smooth

ADIA, MOVE = po.ADIABATIC_WALL, po.MOVING_ADIABATIC_WALL


def cfg_for(dims, kinds, bc_type, **kw):
    Nx, Ny, Nz = dims
    L = [2 * np.pi * (1 - 1 / n) if bc_type[d] == po.PERIODIC_SOLVE else 3.0
         for d, n in enumerate(dims)]
    base = dict(Lx=L[0], Ly=L[1], Lz=L[2], alpha_f=0.45, mu_ref=5e-3, Pr=0.71, rk=4,
                filter_step=1, bc_type=bc_type, bc_kind=kinds,
                wall_vel={3: (0.3, 0.0, -0.2), 1: (0.0, 0.25, 0.1)})
    base.update(kw)
    return po.make_config(Nx, Ny, Nz, **base)


CASES = {
    # walls on both y faces (the upper one moving)
    "y": ((10, 16, 8), (0, 0, ADIA, MOVE, 0, 0), (0, 1, 0)),
    # walls on four faces: edges where two faces meet
    "xy": ((16, 14, 6), (ADIA, MOVE, ADIA, MOVE, 0, 0), (1, 1, 0)),
}


def ic_for(dims, seed):
    rng = np.random.default_rng(seed)
    shape = (dims[2], dims[1], dims[0])
    U = 0.2 * smooth(rng.standard_normal(shape))
    V = 0.2 * smooth(rng.standard_normal(shape))
    W = 0.2 * smooth(rng.standard_normal(shape))
    rho = 1.0 + 0.02 * smooth(rng.standard_normal(shape))
    p = 1.0 / 1.4 * (1.0 + 0.05 * smooth(rng.standard_normal(shape)))
    return rho, U, V, W, p


def oracle_run(cfg, ic, dt, nsteps):
    o = po.Solver(cfg)
    o.set_ic(*ic)
    o.set_dt(dt)
    q0 = o.conserved(1)
    for _ in range(nsteps):
        o.step()
    return o, q0


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("rk,filter_step", [(4, 1), (3, 2)])
def test_forward_with_walls_matches_oracle(case, rk, filter_step):
    dims, kinds, bct = CASES[case]
    cfg = cfg_for(dims, kinds, bct, rk=rk, filter_step=filter_step)
    ic = ic_for(dims, 3)
    dt = 0.01
    A = oa.Adjoint(cfg)
    for n in (1, 3):
        o, q0 = oracle_run(cfg, ic, dt, n)
        q, _, _ = A.run(q0, n, dt, filter_step)
        errs = [relerr(g, r) for g, r in zip(q, o.conserved(1))]
        o.close()
        assert max(errs) <= 1e-12, (n, errs)


@pytest.mark.parametrize("case", sorted(CASES))
def test_stage_and_wall_vjp_dot_products(case):
    dims, kinds, bct = CASES[case]
    cfg = cfg_for(dims, kinds, bct)
    A = oa.Adjoint(cfg)
    shape = A.shape
    o, q = oracle_run(cfg, ic_for(dims, 5), 0.01, 0)
    o.close()
    rng = np.random.default_rng(4)
    lam = [rng.standard_normal(shape) for _ in range(5)]
    dt, eps = 0.01, 1e-5
    for name, fwd, rev in (("stage", lambda x: A.rhs(x, dt), lambda l: A.rhs_vjp(q, l, dt)),
                           ("walls", A.wall_pre, lambda l: A.wall_pre_vjp(q, l))):
        delta = [s * smooth(rng.standard_normal(shape), 1) for s in (1e-2, 0.1, 0.1, 0.1, 0.1)]
        kp = fwd([x + eps * d for x, d in zip(q, delta)])
        km = fwd([x - eps * d for x, d in zip(q, delta)])
        fd = sum(np.vdot(l_, (a - b) / (2 * eps)) for l_, a, b in zip(lam, kp, km))
        ad = sum(np.vdot(a, b) for a, b in zip(rev(lam), delta))
        assert abs(ad - fd) <= 1e-8 * abs(fd), (name, ad, fd)


def test_wall_vjp_without_the_neumann_term_fails():
    dims, kinds, bct = CASES["xy"]
    cfg = cfg_for(dims, kinds, bct)
    A = oa.Adjoint(cfg)
    shape = A.shape
    o, q = oracle_run(cfg, ic_for(dims, 5), 0.01, 0)
    o.close()
    rng = np.random.default_rng(4)
    lam = [rng.standard_normal(shape) for _ in range(5)]
    delta = [0.1 * smooth(rng.standard_normal(shape), 1) for _ in range(5)]
    eps = 1e-5
    kp = A.wall_pre([x + eps * d for x, d in zip(q, delta)])
    km = A.wall_pre([x - eps * d for x, d in zip(q, delta)])
    fd = sum(np.vdot(l_, (a - b) / (2 * eps)) for l_, a, b in zip(lam, kp, km))
    saved = oa.Adjoint.CN
    try:
        oa.Adjoint.CN = 0.0 * saved  # the reverse then drops T_w's neighbours
        bad = sum(np.vdot(a, b) for a, b in zip(A.wall_pre_vjp(q, lam), delta))
    finally:
        oa.Adjoint.CN = saved
    assert abs(bad - fd) > 1e-3 * abs(fd), (bad, fd)


def run_ke(cfg, ic, dt, nsteps):
    o, _ = oracle_run(cfg, ic, dt, nsteps)
    rho = o.field(0)
    u = [o.field(i) / rho for i in (1, 2, 3)]
    o.close()
    return float(np.sum(0.5 * rho * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)) / rho.size)


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("rk,filter_step", [(4, 1), (3, 2)])
def test_ke_gradient_with_walls_matches_oracle_finite_differences(case, rk, filter_step):
    dims, kinds, bct = CASES[case]
    cfg = cfg_for(dims, kinds, bct, rk=rk, filter_step=filter_step)
    ic = ic_for(dims, 9)
    dt, nsteps = 0.01, 3
    A = oa.Adjoint(cfg)
    o, q0 = oracle_run(cfg, ic, dt, 0)
    o.close()
    J, gU, gV, gW = A.ke_gradient(ic, nsteps, dt, filter_step, q0=q0)
    assert abs(J - run_ke(cfg, ic, dt, nsteps)) <= 1e-13 * abs(J)
    # directions away from the wall planes (set_initial_conditions
    # overwrites the wall faces' velocities)
    inner = np.ones(A.shape)
    inner[A.rowdir >= 0] = 0.0
    rng = np.random.default_rng(12)
    eps = 1e-5
    rho, U, V, W, p = ic
    for _ in range(2):
        d = [inner * smooth(rng.standard_normal(U.shape), 1) for _ in range(3)]
        Jp = run_ke(cfg, (rho, U + eps * d[0], V + eps * d[1], W + eps * d[2], p), dt, nsteps)
        Jm = run_ke(cfg, (rho, U - eps * d[0], V - eps * d[1], W - eps * d[2], p), dt, nsteps)
        fd = (Jp - Jm) / (2 * eps)
        ad = np.vdot(gU, d[0]) + np.vdot(gV, d[1]) + np.vdot(gW, d[2])
        assert abs(ad - fd) <= 1e-7 * abs(fd), (ad, fd)


# walls on the x faces beside sponges on the y faces: the wall faces cross
# the sponge planes, where the first stage after each sponge update reads
# the stale (pre-reset) temperatures
WS_DIMS = (16, 20, 6)
WS_KINDS = (ADIA, MOVE, po.SPONGE, po.SPONGE, 0, 0)


def ws_cfg(**kw):
    base = dict(sponge_frac=(0.25, 0.2, 0.25), sponge_avgT=0.5,
                wall_vel={1: (0.0, 0.25, 0.1)})
    base.update(kw)
    return cfg_for(WS_DIMS, WS_KINDS, (1, 1, 0), **base)


@pytest.mark.parametrize("rk,filter_step", [(4, 1), (3, 2)])
def test_forward_walls_beside_sponges_matches_oracle(rk, filter_step):
    cfg = ws_cfg(rk=rk, filter_step=filter_step)
    ic = ic_for(WS_DIMS, 3)
    dt = 0.01
    A = oa.Adjoint(cfg)
    for n in (1, 3):
        o, q0 = oracle_run(cfg, ic, dt, n)
        q, Av, _ = A.run(q0, n, dt, filter_step)
        errs = [relerr(g, r) for g, r in zip(q, o.conserved(1))]
        errs += [relerr(g, r) for g, r in zip(Av, o.sponge_avg())]
        o.close()
        assert max(errs) <= 1e-12, (n, errs)


def test_wall_pair_vjp_dot_product():
    """the first stage's wall correction of the pair (state, pre-reset
    state): dot-product identity in both inputs; a reverse that sends the
    wall temperature's cotangent to the state instead of the pre-reset
    state fails it"""
    cfg = ws_cfg()
    A = oa.Adjoint(cfg)
    shape = A.shape
    o, q = oracle_run(cfg, ic_for(WS_DIMS, 5), 0.01, 1)
    qf = o.conserved(1)
    o.close()
    rng = np.random.default_rng(6)
    # a pre-reset state that differs from the state on the sponge planes
    qs = [np.where(A.onplane, x * (1 + 0.05 * rng.standard_normal(shape)), x) for x in qf]
    lam = [rng.standard_normal(shape) for _ in range(5)]
    lqs = [rng.standard_normal(shape) for _ in range(5)]
    eps = 1e-5
    dq = [s * smooth(rng.standard_normal(shape), 1) for s in (1e-2, 0.1, 0.1, 0.1, 0.1)]
    ds = [s * smooth(rng.standard_normal(shape), 1) for s in (1e-2, 0.1, 0.1, 0.1, 0.1)]
    kp = A.wall_pre([x + eps * d for x, d in zip(qf, dq)], qs=[x + eps * d for x, d in zip(qs, ds)])
    km = A.wall_pre([x - eps * d for x, d in zip(qf, dq)], qs=[x - eps * d for x, d in zip(qs, ds)])
    fd = sum(np.vdot(l_, (a - b) / (2 * eps)) for l_, a, b in zip(lam + lqs, kp[0] + kp[1],
                                                                   km[0] + km[1]))
    gq, gs = A.wall_pre_vjp(qf, lam, qs, lqs)
    ad = sum(np.vdot(a, b) for a, b in zip(gq + gs, dq + ds))
    assert abs(ad - fd) <= 1e-8 * abs(fd), (ad, fd)
    bq, _ = A.wall_pre_vjp(qf, lam, qf, lqs)  # temperatures of the state
    bad = sum(np.vdot(a, b) for a, b in zip(bq + gs, dq + ds))
    assert abs(bad - fd) > 1e-6 * abs(fd), (bad, fd)


@pytest.mark.parametrize("rk,filter_step", [(4, 1), (3, 2)])
def test_ke_gradient_walls_beside_sponges_matches_oracle_finite_differences(rk, filter_step):
    cfg = ws_cfg(rk=rk, filter_step=filter_step)
    ic = ic_for(WS_DIMS, 9)
    dt, nsteps = 0.01, 3
    A = oa.Adjoint(cfg)
    o, q0 = oracle_run(cfg, ic, dt, 0)
    o.close()
    J, gU, gV, gW = A.ke_gradient(ic, nsteps, dt, filter_step, q0=q0)
    assert abs(J - run_ke(cfg, ic, dt, nsteps)) <= 1e-13 * abs(J)
    inner = np.ones(A.shape)
    inner[A.rowdir >= 0] = 0.0
    rng = np.random.default_rng(13)
    eps = 1e-5
    rho, U, V, W, p = ic
    for _ in range(2):
        d = [inner * smooth(rng.standard_normal(U.shape), 1) for _ in range(3)]
        Jp = run_ke(cfg, (rho, U + eps * d[0], V + eps * d[1], W + eps * d[2], p), dt, nsteps)
        Jm = run_ke(cfg, (rho, U - eps * d[0], V - eps * d[1], W - eps * d[2], p), dt, nsteps)
        fd = (Jp - Jm) / (2 * eps)
        ad = np.vdot(gU, d[0]) + np.vdot(gV, d[1]) + np.vdot(gW, d[2])
        assert abs(ad - fd) <= 1e-7 * abs(fd), (ad, fd)
