"""The discrete adjoint's general path on the GPU: non-periodic lines with
their boundary closures and sponge faces (the source, the zeroed face
planes, the running averages and the plane reset, the stale primitives of
the first stage after each sponge update).

Checked field by field against the FP64 numpy adjoint of the oracle
(oracle/adjoint.py: dense transposed operators, pinned on the CPU against
the C oracle and finite differences), against central differences of whole
GPU forward runs, and, on periodic grids where both paths apply, against the
fused five-pass path."""
import os

import numpy as np
import pytest

import adjoint as oa
import cs3d
import pyoracle as po
from helpers import random_ic, relerr, shear_ic, smooth, tgv_ic

pytestmark = pytest.mark.gpu

DIMS = (16, 24, 8)  # x, z periodic; y non-periodic between sponge faces
LY = 6.0


def sponge_kw(**extra):
    Nx, Ny, Nz = DIMS
    kw = dict(Lx=2 * np.pi * (1 - 1 / Nx), Ly=LY, Lz=2 * np.pi * (1 - 1 / Nz), alpha_f=0.45,
              mu_ref=2e-2, Pr=0.71, cfl=0.5, filter_step=1, rk=4, dt_fixed=0.02,
              bc_type=(0, 1, 0), bc_kind=(0, 0, 1, 1, 0, 0), sponge_frac=(0.25, 0.25, 0.25),
              sponge_avgT=0.5)
    kw.update(extra)
    return kw


def oracle_cfg(dims, kw):
    keep = {k: v for k, v in kw.items()
            if k in ("Lx", "Ly", "Lz", "alpha_f", "mu_ref", "Pr", "filter_step", "rk",
                     "bc_type", "bc_kind", "sponge_frac", "sponge_strength", "sponge_avgT",
                     "sponge_epsP")}
    return po.make_config(*dims, **keep)


def ic_for(seed):
    return shear_ic(DIMS[0], DIMS[1], DIMS[2], LY, seed, dU=0.6, amp=0.1)


def planes_mask(shape):
    m = np.zeros(shape, dtype=bool)
    m[:, 0, :] = True
    m[:, -1, :] = True
    return m


@pytest.mark.parametrize("stale", [False, True])
def test_rhs_vjp_with_sponges_matches_oracle_adjoint_fp64(stale):
    """the stage product at *1: right after the initial conditions, and
    after a step (the first stage after a sponge update: the x block's
    cotangents on the sponge planes belong to the pre-reset state, which the
    product holds fixed)"""
    kw = sponge_kw()
    shape = (DIMS[2], DIMS[1], DIMS[0])
    g = cs3d.Solver(cs3d.config(*DIMS, precision=cs3d.FP64, **kw))
    g.set_initial_conditions(*ic_for(2))
    if stale:
        g.step(1)
    g.pre_step()
    q = g.conserved(1)
    A = g.conserved(5)
    rng = np.random.default_rng(8)
    lam = [rng.standard_normal(shape) for _ in range(5)]
    got = g.rhs_vjp(lam, 1)
    adj = oa.Adjoint(oracle_cfg(DIMS, kw))
    if stale:
        on = planes_mask(shape)
        qs = [np.where(on, s, x) for s, x in zip(g.conserved(6), q)]
        qb, qsb, _ = adj.rhs_vjp_full(q, lam, 0.02, A, qs)
        ref = [a + np.where(on, 0.0, b) for a, b in zip(qb, qsb)]
    else:
        ref = adj.rhs_vjp(q, lam, 0.02, A)
    errs = [relerr(a, b) for a, b in zip(got, ref)]
    print(f"sponge vjp (stale={stale}) vs oracle adjoint:", " ".join(f"{e:.2e}" for e in errs),
          flush=True)
    assert max(errs) <= 1e-11, errs


@pytest.mark.parametrize("filter_step,rk,K", [(1, 4, 2), (2, 4, 3), (1, 3, 2), (2, 3, 1)])
def test_ke_gradient_with_sponges_matches_oracle_adjoint_fp64(filter_step, rk, K):
    kw = sponge_kw(filter_step=filter_step, rk=rk)
    ic = ic_for(9)
    g = cs3d.Solver(cs3d.config(*DIMS, precision=cs3d.FP64, **kw))
    g.set_initial_conditions(*ic)
    J, *grad = g.ke_gradient(5, K)
    Jo, *gref = oa.Adjoint(oracle_cfg(DIMS, kw)).ke_gradient(ic, 5, 0.02, filter_step)
    print(f"J {J:.15e} oracle {Jo:.15e}", flush=True)
    assert abs(J - Jo) <= 1e-13 * abs(Jo)
    errs = [relerr(a, b) for a, b in zip(grad, gref)]
    print("sponge gradient vs oracle adjoint:", " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 1e-10, errs
    # the solver is left at the forward pass's end, as after 5 steps
    f = cs3d.Solver(cs3d.config(*DIMS, precision=cs3d.FP64, **kw))
    f.set_initial_conditions(*ic)
    f.step(5)
    for a, b in zip(g.conserved(1) + g.conserved(5), f.conserved(1) + f.conserved(5)):
        assert relerr(a, b) <= 1e-14
    g.step(1)
    f.step(1)
    for a, b in zip(g.conserved(1), f.conserved(1)):
        assert relerr(a, b) <= 1e-14


@pytest.mark.parametrize("rk", [4, 3])
def test_ke_gradient_on_three_point_z_segments_matches_oracle_adjoint_fp64(rk):
    """18-point periodic z lines: the forward z sweep and the adjoint's z
    operators (batched derivatives and transposed terms) run 3-point
    segments; the gradient still equals the oracle adjoint's"""
    dims = (16, 24, 18)
    kw = sponge_kw(rk=rk, Lz=2 * np.pi * (1 - 1 / dims[2]))
    ic = shear_ic(*dims, LY, 11, dU=0.6, amp=0.1)
    g = cs3d.Solver(cs3d.config(*dims, precision=cs3d.FP64, **kw))
    g.set_initial_conditions(*ic)
    J, *grad = g.ke_gradient(4, 2)
    Jo, *gref = oa.Adjoint(oracle_cfg(dims, kw)).ke_gradient(ic, 4, 0.02, 1)
    print(f"z18 rk={rk} J {J:.15e} oracle {Jo:.15e}", flush=True)
    assert abs(J - Jo) <= 1e-13 * abs(Jo)
    errs = [relerr(a, b) for a, b in zip(grad, gref)]
    print(f"z18 rk={rk} gradient vs oracle adjoint:", " ".join(f"{e:.2e}" for e in errs),
          flush=True)
    assert max(errs) <= 1e-10, errs


def run_ke(ic, kw, nsteps, prec=cs3d.FP64):
    g = cs3d.Solver(cs3d.config(*DIMS, precision=prec, **kw))
    g.set_initial_conditions(*ic)
    g.step(nsteps)
    return g.kinetic_energy()


def test_ke_gradient_with_sponges_matches_finite_differences_fp64():
    kw = sponge_kw()
    ic = ic_for(4)
    g = cs3d.Solver(cs3d.config(*DIMS, precision=cs3d.FP64, **kw))
    g.set_initial_conditions(*ic)
    J, gU, gV, gW = g.ke_gradient(8, 3)
    rho, U, V, W, p = ic
    rng = np.random.default_rng(31)
    eps = 1e-5
    for trial in range(3):
        d = [smooth(rng.standard_normal(U.shape), 1) for _ in range(3)]
        Jp = run_ke((rho, U + eps * d[0], V + eps * d[1], W + eps * d[2], p), kw, 8)
        Jm = run_ke((rho, U - eps * d[0], V - eps * d[1], W - eps * d[2], p), kw, 8)
        fd = (Jp - Jm) / (2 * eps)
        ad = float(np.vdot(gU, d[0]) + np.vdot(gV, d[1]) + np.vdot(gW, d[2]))
        rel = abs(ad - fd) / abs(fd)
        print(f"sponge gradient trial {trial}: adjoint {ad:.12e} fd {fd:.12e} rel {rel:.2e}",
              flush=True)
        assert rel <= 1e-4


def test_ke_gradient_with_sponges_fp32_tracks_fp64():
    kw = sponge_kw()
    ic = ic_for(6)
    out = {}
    for prec in (cs3d.FP64, cs3d.FP32):
        g = cs3d.Solver(cs3d.config(*DIMS, precision=prec, **kw))
        g.set_initial_conditions(*ic)
        out[prec] = g.ke_gradient(6, 2)
    J64, *g64 = out[cs3d.FP64]
    J32, *g32 = out[cs3d.FP32]
    assert abs(J32 - J64) <= 1e-5 * abs(J64)
    for a, b in zip(g32, g64):
        err = relerr(a, b)
        print(f"sponge fp32 gradient rel err {err:.2e}", flush=True)
        assert err <= 1e-3


def test_general_path_equals_fused_path_on_periodic_grids():
    """CS3D_ADJ_GENERAL forces the general path; on a periodic grid it must
    give the fused path's stage product and gradient"""
    N = 24
    Nx, Ny, Nz = N, 20, 28
    kw = dict(Lx=2 * np.pi * (1 - 1 / Nx), Ly=2 * np.pi * (1 - 1 / Ny),
              Lz=2 * np.pi * (1 - 1 / Nz), alpha_f=0.45, mu_ref=2e-2, Pr=0.71, cfl=0.5,
              filter_step=1, rk=4, dt_fixed=0.02)
    shape = (Nz, Ny, Nx)
    ic = random_ic(shape, seed=12, amp=0.4)
    lam = [np.random.default_rng(3).standard_normal(shape) for _ in range(5)]
    res = {}
    for general in (False, True):
        if general:
            os.environ["CS3D_ADJ_GENERAL"] = "1"
        try:
            g = cs3d.Solver(cs3d.config(Nx, Ny, Nz, precision=cs3d.FP64, **kw))
            g.set_initial_conditions(*ic)
            g.pre_step()
            v = g.rhs_vjp(lam, 1)
            g.set_initial_conditions(*ic)
            res[general] = (v, g.ke_gradient(4, 2))
            g.close()
        finally:
            os.environ.pop("CS3D_ADJ_GENERAL", None)
    for a, b in zip(res[True][0], res[False][0]):
        assert relerr(a, b) <= 1e-12
    assert abs(res[True][1][0] - res[False][1][0]) <= 1e-15 * abs(res[False][1][0])
    for a, b in zip(res[True][1][1:], res[False][1][1:]):
        e = relerr(a, b)
        print(f"general vs fused gradient {e:.2e}", flush=True)
        assert e <= 1e-12


def test_sponge_gradient_needs_the_initial_averages():
    """with sponges the gradient starts from set_initial_conditions (the
    averages are the initial state); a solver that has stepped refuses"""
    kw = sponge_kw()
    g = cs3d.Solver(cs3d.config(*DIMS, precision=cs3d.FP64, **kw))
    g.set_initial_conditions(*ic_for(1))
    g.step(1)
    with pytest.raises(cs3d.CS3DError):
        g.ke_gradient(2, 1)


def test_shear_layer_512x128x42_fp32_gradient():
    """the shear-layer configuration at its full size (512 x 128 x 42: y
    non-periodic between sponge layers of 15% of the height, z periodic on
    42 points, alpha_f 0.495, the filter every 5 steps): the FP32 adjoint
    through 20 RK4 steps with a checkpoint every 5 steps against FP64
    central differences of whole forward runs (eps 1e-3) along 4 seeded
    Gaussian directions; the dt frozen at the first step's CFL dt.  1e-4
    relative.  Also prints the cost against the forward steps."""
    import time
    import ic as icmod
    dims = (512, 128, 42)
    kw = dict(Lx=172.0, Ly=129.0, Lz=28.2, alpha_f=0.495, mu_ref=0.00375, Pr=0.7, cfl=0.8,
              filter_step=5, rk=4, bc_type=(0, 1, 0), bc_kind=(0, 0, 1, 1, 0, 0),
              sponge_frac=(0.25, 0.15, 0.25))
    icf = icmod.shear_layer_ic(*dims, Ly=kw["Ly"], delta_u=1.2, delta0=1.0, seed=0)
    probe = cs3d.Solver(cs3d.config(*dims, precision=cs3d.FP64, **kw))
    probe.set_initial_conditions(*icf)
    probe.pre_step()
    kw["dt_fixed"] = probe.time()[1]
    probe.close()
    nsteps, K = 20, 5
    f = cs3d.Solver(cs3d.config(*dims, precision=cs3d.FP32, **kw))
    f.set_initial_conditions(*icf)
    f.step(2)
    f.synchronize()
    t0 = time.perf_counter()
    f.step(nsteps)
    f.synchronize()
    fwd = time.perf_counter() - t0
    f.close()
    g = cs3d.Solver(cs3d.config(*dims, precision=cs3d.FP32, **kw))
    g.set_initial_conditions(*icf)
    t0 = time.perf_counter()
    J, gU, gV, gW = g.ke_gradient(nsteps, K)
    tot = time.perf_counter() - t0
    g.close()
    print(f"shear adjoint: J {J:.10e}; forward {fwd * 1e3:.1f} ms, gradient call "
          f"{tot * 1e3:.1f} ms (first call, allocation included): {tot / fwd:.2f}x", flush=True)
    rho, U, V, W, p = icf
    eps = 1e-3
    worst = 0.0
    for seed in range(4):
        rng = np.random.default_rng(seed)
        d = [rng.standard_normal(U.shape) for _ in range(3)]
        Js = []
        for sgn in (1, -1):
            h = cs3d.Solver(cs3d.config(*dims, precision=cs3d.FP64, **kw))
            h.set_initial_conditions(rho, U + sgn * eps * d[0], V + sgn * eps * d[1],
                                     W + sgn * eps * d[2], p)
            h.step(nsteps)
            Js.append(h.kinetic_energy())
            h.close()
        fd = (Js[0] - Js[1]) / (2 * eps)
        ad = float(np.vdot(gU, d[0]) + np.vdot(gV, d[1]) + np.vdot(gW, d[2]))
        rel = abs(ad - fd) / abs(fd)
        worst = max(worst, rel)
        print(f"shear direction {seed}: adjoint {ad:.10e} fd {fd:.10e} rel {rel:.2e}", flush=True)
    assert worst <= 1e-4
