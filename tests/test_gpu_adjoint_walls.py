"""The discrete adjoint with adiabatic and moving walls on the GPU (the
general path): the wall correction of every stage state before its
right-hand side (the first one in place, under the RK combination), and
the continuity-only wall rows.  Checked field by field against the FP64
numpy adjoint of the oracle (oracle/adjoint.py, pinned on the CPU against
the C oracle and finite differences), against central differences of whole
GPU forward runs (directions off the wall planes, which
set_initial_conditions overwrites), and FP32 against FP64; and walls beside
sponge faces, whose first stage after each sponge update corrects the
pre-reset set together with the stage state."""
import numpy as np
import pytest

import adjoint as oa
import cs3d
import pyoracle as po
from helpers import relerr, smooth

pytestmark = pytest.mark.gpu

ADIA, MOVE = po.ADIABATIC_WALL, po.MOVING_ADIABATIC_WALL
WALL_VEL = {3: (0.3, 0.0, -0.2), 1: (0.0, 0.25, 0.1)}
# the transposed operators need 24 points on non-periodic lines
CASES = {
    "y": ((16, 24, 8), (0, 0, ADIA, MOVE, 0, 0), (0, 1, 0)),
    "xy": ((24, 24, 8), (ADIA, MOVE, ADIA, MOVE, 0, 0), (1, 1, 0)),
}


def kw_for(case, **extra):
    dims, kinds, bct = CASES[case]
    L = [2 * np.pi * (1 - 1 / n) if bct[d] == 0 else 3.0 for d, n in enumerate(dims)]
    kw = dict(Lx=L[0], Ly=L[1], Lz=L[2], alpha_f=0.45, mu_ref=5e-3, Pr=0.71, cfl=0.5,
              filter_step=1, rk=4, dt_fixed=0.01, bc_type=bct, bc_kind=kinds,
              wall_vel=[WALL_VEL.get(f, (0.0, 0.0, 0.0)) for f in range(6)])
    kw.update(extra)
    return dims, kw


def oracle_cfg(dims, kw):
    keep = {k: v for k, v in kw.items()
            if k in ("Lx", "Ly", "Lz", "alpha_f", "mu_ref", "Pr", "filter_step", "rk",
                     "bc_type", "bc_kind")}
    return po.make_config(*dims, wall_vel=WALL_VEL, **keep)


def ic_for(dims, seed):
    rng = np.random.default_rng(seed)
    shape = (dims[2], dims[1], dims[0])
    U = 0.2 * smooth(rng.standard_normal(shape))
    V = 0.2 * smooth(rng.standard_normal(shape))
    W = 0.2 * smooth(rng.standard_normal(shape))
    rho = 1.0 + 0.02 * smooth(rng.standard_normal(shape))
    p = 1.0 / 1.4 * (1.0 + 0.05 * smooth(rng.standard_normal(shape)))
    return rho, U, V, W, p


@pytest.mark.parametrize("case", sorted(CASES))
def test_rhs_vjp_with_walls_matches_oracle_adjoint_fp64(case):
    """the product of the map cs3d_rhs evaluates, Q -> dt R(W(Q))"""
    dims, kw = kw_for(case)
    g = cs3d.Solver(cs3d.config(*dims, precision=cs3d.FP64, **kw))
    g.set_initial_conditions(*ic_for(dims, 2))
    g.pre_step()
    q = g.conserved(1)
    shape = (dims[2], dims[1], dims[0])
    rng = np.random.default_rng(8)
    lam = [rng.standard_normal(shape) for _ in range(5)]
    got = g.rhs_vjp(lam, 1)
    A = oa.Adjoint(oracle_cfg(dims, kw))
    ref = A.wall_pre_vjp(q, A.rhs_vjp(A.wall_pre(q), lam, 0.01))
    errs = [relerr(a, b) for a, b in zip(got, ref)]
    print(f"walls vjp ({case}) vs oracle adjoint:", " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 1e-11, errs


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("rk,filter_step,K", [(4, 1, 2), (4, 2, 3), (3, 1, 2), (3, 2, 1)])
def test_ke_gradient_with_walls_matches_oracle_adjoint_fp64(case, rk, filter_step, K):
    dims, kw = kw_for(case, rk=rk, filter_step=filter_step)
    ic = ic_for(dims, 9)
    g = cs3d.Solver(cs3d.config(*dims, precision=cs3d.FP64, **kw))
    g.set_initial_conditions(*ic)
    q0 = g.conserved(1)
    J, *grad = g.ke_gradient(5, K)
    Jo, *gref = oa.Adjoint(oracle_cfg(dims, kw)).ke_gradient(ic, 5, 0.01, filter_step, q0=q0)
    assert abs(J - Jo) <= 1e-13 * abs(Jo)
    errs = [relerr(a, b) for a, b in zip(grad, gref)]
    print(f"walls gradient ({case}) vs oracle adjoint:", " ".join(f"{e:.2e}" for e in errs),
          flush=True)
    assert max(errs) <= 1e-10, errs


def run_ke(dims, kw, ic, nsteps, prec=cs3d.FP64):
    g = cs3d.Solver(cs3d.config(*dims, precision=prec, **kw))
    g.set_initial_conditions(*ic)
    g.step(nsteps)
    return g.kinetic_energy()


@pytest.mark.parametrize("case", sorted(CASES))
def test_ke_gradient_with_walls_matches_finite_differences_fp64(case):
    dims, kw = kw_for(case)
    ic = ic_for(dims, 4)
    g = cs3d.Solver(cs3d.config(*dims, precision=cs3d.FP64, **kw))
    g.set_initial_conditions(*ic)
    J, gU, gV, gW = g.ke_gradient(6, 2)
    shape = (dims[2], dims[1], dims[0])
    inner = np.ones(shape)
    kinds = CASES[case][1]
    for f in range(6):
        if kinds[f]:
            idx = [slice(None)] * 3
            idx[2 - f // 2] = -1 if f % 2 else 0
            inner[tuple(idx)] = 0.0
    rho, U, V, W, p = ic
    rng = np.random.default_rng(31)
    eps = 1e-5
    for trial in range(2):
        d = [inner * smooth(rng.standard_normal(shape), 1) for _ in range(3)]
        Jp = run_ke(dims, kw, (rho, U + eps * d[0], V + eps * d[1], W + eps * d[2], p), 6)
        Jm = run_ke(dims, kw, (rho, U - eps * d[0], V - eps * d[1], W - eps * d[2], p), 6)
        fd = (Jp - Jm) / (2 * eps)
        ad = float(np.vdot(gU, d[0]) + np.vdot(gV, d[1]) + np.vdot(gW, d[2]))
        rel = abs(ad - fd) / abs(fd)
        print(f"walls gradient ({case}) trial {trial}: adjoint {ad:.12e} fd {fd:.12e} "
              f"rel {rel:.2e}", flush=True)
        assert rel <= 1e-4


def test_ke_gradient_with_walls_fp32_tracks_fp64():
    dims, kw = kw_for("xy")
    ic = ic_for(dims, 6)
    out = {}
    for prec in (cs3d.FP64, cs3d.FP32):
        g = cs3d.Solver(cs3d.config(*dims, precision=prec, **kw))
        g.set_initial_conditions(*ic)
        out[prec] = g.ke_gradient(6, 2)
    J64, *g64 = out[cs3d.FP64]
    J32, *g32 = out[cs3d.FP32]
    assert abs(J32 - J64) <= 1e-5 * abs(J64)
    for a, b in zip(g32, g64):
        err = relerr(a, b)
        print(f"walls fp32 gradient rel err {err:.2e}", flush=True)
        assert err <= 1e-3


# walls on the x faces beside sponges on the y faces (z periodic): the wall
# faces cross the sponge planes, where the first stage after each sponge
# update corrects the pre-reset set together with the stage state
WS_DIMS = (24, 24, 8)
WS_KW = dict(Lx=3.0, Ly=3.0, Lz=2 * np.pi * (1 - 1 / 8), alpha_f=0.45, mu_ref=5e-3, Pr=0.71,
             cfl=0.5, filter_step=1, rk=4, dt_fixed=0.01, bc_type=(1, 1, 0),
             bc_kind=(ADIA, MOVE, po.SPONGE, po.SPONGE, 0, 0),
             wall_vel=[WALL_VEL.get(f, (0.0, 0.0, 0.0)) for f in range(6)],
             sponge_frac=(0.25, 0.2, 0.25), sponge_avgT=0.5)


def ws_oracle_cfg(kw):
    keep = {k: v for k, v in kw.items()
            if k in ("Lx", "Ly", "Lz", "alpha_f", "mu_ref", "Pr", "filter_step", "rk",
                     "bc_type", "bc_kind", "sponge_frac", "sponge_avgT")}
    return po.make_config(*WS_DIMS, wall_vel=WALL_VEL, **keep)


@pytest.mark.parametrize("stale", [False, True])
def test_rhs_vjp_walls_beside_sponges_matches_oracle_adjoint_fp64(stale):
    """the stage product at *1: after the initial conditions, and after a
    step (the first stage after a sponge update, whose wall correction acts
    on the pre-reset set too; that set held fixed)"""
    g = cs3d.Solver(cs3d.config(*WS_DIMS, precision=cs3d.FP64, **WS_KW))
    g.set_initial_conditions(*ic_for(WS_DIMS, 2))
    if stale:
        g.step(1)
    g.pre_step()
    q = g.conserved(1)
    Av = g.conserved(5)
    shape = (WS_DIMS[2], WS_DIMS[1], WS_DIMS[0])
    rng = np.random.default_rng(8)
    lam = [rng.standard_normal(shape) for _ in range(5)]
    got = g.rhs_vjp(lam, 1)
    A = oa.Adjoint(ws_oracle_cfg(WS_KW))
    if stale:
        on = A.onplane
        qs = [np.where(on, s, x) for s, x in zip(g.conserved(6), q)]
        q1, qs1 = A.wall_first(q, qs)
        lq1, lqs1, _ = A.rhs_vjp_full(q1, lam, 0.01, Av, qs1)
        lq, lqs = A.wall_first_vjp(q, qs, lq1, lqs1)
        ref = [a + np.where(on, 0.0, b) for a, b in zip(lq, lqs)]
    else:
        ref = A.wall_pre_vjp(q, A.rhs_vjp(A.wall_pre(q), lam, 0.01, Av))
    errs = [relerr(a, b) for a, b in zip(got, ref)]
    print(f"walls+sponges vjp (stale={stale}) vs oracle adjoint:",
          " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 1e-11, errs


@pytest.mark.parametrize("rk,filter_step,K", [(4, 1, 2), (4, 2, 3), (3, 1, 2)])
def test_ke_gradient_walls_beside_sponges_matches_oracle_adjoint_fp64(rk, filter_step, K):
    kw = dict(WS_KW, rk=rk, filter_step=filter_step)
    ic = ic_for(WS_DIMS, 9)
    g = cs3d.Solver(cs3d.config(*WS_DIMS, precision=cs3d.FP64, **kw))
    g.set_initial_conditions(*ic)
    q0 = g.conserved(1)
    J, *grad = g.ke_gradient(5, K)
    Jo, *gref = oa.Adjoint(ws_oracle_cfg(kw)).ke_gradient(ic, 5, 0.01, filter_step, q0=q0)
    assert abs(J - Jo) <= 1e-13 * abs(Jo)
    errs = [relerr(a, b) for a, b in zip(grad, gref)]
    print(f"walls+sponges gradient rk{rk} fs{filter_step} K{K} vs oracle adjoint:",
          " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 1e-10, errs


def test_ke_gradient_walls_beside_sponges_matches_finite_differences():
    """FP64 against central differences of whole GPU runs (directions off
    the wall planes), and FP32 against FP64"""
    ic = ic_for(WS_DIMS, 4)
    shape = (WS_DIMS[2], WS_DIMS[1], WS_DIMS[0])
    out = {}
    for prec in (cs3d.FP64, cs3d.FP32):
        g = cs3d.Solver(cs3d.config(*WS_DIMS, precision=prec, **WS_KW))
        g.set_initial_conditions(*ic)
        out[prec] = g.ke_gradient(6, 2)
    J, gU, gV, gW = out[cs3d.FP64]
    inner = np.ones(shape)
    inner[:, :, 0] = 0.0
    inner[:, :, -1] = 0.0
    rho, U, V, W, p = ic
    rng = np.random.default_rng(32)
    eps = 1e-5
    for trial in range(2):
        d = [inner * smooth(rng.standard_normal(shape), 1) for _ in range(3)]
        Jp = run_ke(WS_DIMS, WS_KW, (rho, U + eps * d[0], V + eps * d[1], W + eps * d[2], p), 6)
        Jm = run_ke(WS_DIMS, WS_KW, (rho, U - eps * d[0], V - eps * d[1], W - eps * d[2], p), 6)
        fd = (Jp - Jm) / (2 * eps)
        ad = float(np.vdot(gU, d[0]) + np.vdot(gV, d[1]) + np.vdot(gW, d[2]))
        rel = abs(ad - fd) / abs(fd)
        print(f"walls+sponges gradient trial {trial}: adjoint {ad:.12e} fd {fd:.12e} "
              f"rel {rel:.2e}", flush=True)
        assert rel <= 1e-4
    J32, *g32 = out[cs3d.FP32]
    assert abs(J32 - J) <= 1e-5 * abs(J)
    for a, b in zip(g32, (gU, gV, gW)):
        err = relerr(a, b)
        print(f"walls+sponges fp32 gradient rel err {err:.2e}", flush=True)
        assert err <= 1e-3
