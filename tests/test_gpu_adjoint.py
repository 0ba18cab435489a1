"""The discrete adjoint on the GPU.

The reference has no adjoint, so nothing of the reference pins it; the
mathematics does.  The vector-Jacobian product of the stage map
Q -> dt*RHS(Q) must satisfy the dot-product identity
<lam, J delta> = <J^T lam, delta> against central differences of the
forward stage (itself pinned to the oracle by test_gpu_solver), and the
gradient of the final kinetic energy through RK4 steps with the filter
must match central differences of whole forward runs along random
directions (dt held fixed, so the forward map is the one differentiated)."""
import numpy as np
import pytest

import cs3d
from helpers import random_ic, smooth, tgv_ic

pytestmark = pytest.mark.gpu


def adj_kw(dims, mu_ref=2e-2, **extra):
    Nx, Ny, Nz = dims
    kw = dict(Lx=2 * np.pi * (1 - 1 / Nx), Ly=2 * np.pi * (1 - 1 / Ny),
              Lz=2 * np.pi * (1 - 1 / Nz), alpha_f=0.45, mu_ref=mu_ref, Pr=0.71,
              cfl=0.5, filter_step=1, rk=4)
    kw.update(extra)
    return kw


def rhs_at(g, Q):
    g.set_conserved(Q, 1)
    g.rhs(1)
    return g.conserved(3)


def dot(a, b):
    return float(sum(np.vdot(x, y) for x, y in zip(a, b)))


@pytest.mark.parametrize("dims", [(16, 16, 16), (20, 12, 14)])
def test_rhs_vjp_dot_product_fp64(dims):
    """<J^T lam, delta> against <lam, (k(Q + e delta) - k(Q - e delta)) / 2e>"""
    Nx, Ny, Nz = dims
    shape = (Nz, Ny, Nx)
    ic = random_ic(shape, seed=11, amp=0.4)
    g = cs3d.Solver(cs3d.config(Nx, Ny, Nz, precision=cs3d.FP64, **adj_kw(dims)))
    g.set_initial_conditions(*ic)
    g.pre_step()
    Q = g.conserved(1)
    rng = np.random.default_rng(5)
    lam = [rng.standard_normal(shape) for _ in range(5)]
    scale = [1e-2, 1e-1, 1e-1, 1e-1, 1e-1]
    delta = [s * smooth(rng.standard_normal(shape), 1) for s in scale]
    JTlam = g.rhs_vjp(lam, 1)
    lhs = dot(JTlam, delta)
    eps = 1e-5
    kp = rhs_at(g, [q + eps * d for q, d in zip(Q, delta)])
    km = rhs_at(g, [q - eps * d for q, d in zip(Q, delta)])
    rhs = dot(lam, [(a - b) / (2 * eps) for a, b in zip(kp, km)])
    print(f"dot product {dims}: {lhs:.12e} {rhs:.12e} rel {abs(lhs - rhs) / abs(rhs):.2e}",
          flush=True)
    assert abs(lhs - rhs) <= 1e-7 * abs(rhs)
    # per-component: each conserved cotangent pairs with its own perturbation
    for c in range(5):
        dc = [d if i == c else np.zeros(shape) for i, d in enumerate(delta)]
        kp = rhs_at(g, [q + eps * d for q, d in zip(Q, dc)])
        km = rhs_at(g, [q - eps * d for q, d in zip(Q, dc)])
        r = dot(lam, [(a - b) / (2 * eps) for a, b in zip(kp, km)])
        l_ = float(np.vdot(JTlam[c], delta[c]))
        print(f"  component {c}: {l_:.10e} {r:.10e}", flush=True)
        assert abs(l_ - r) <= 1e-6 * abs(r) + 1e-9 * abs(rhs), c


def run_ke(dims, ic, kw, nsteps, prec=cs3d.FP64):
    Nx, Ny, Nz = dims
    g = cs3d.Solver(cs3d.config(Nx, Ny, Nz, precision=prec, **kw))
    g.set_initial_conditions(*ic)
    g.step(nsteps)
    return g.kinetic_energy()


@pytest.mark.parametrize("nsteps,checkpoint", [(6, 4), (5, 1)])
def test_ke_gradient_matches_finite_differences_fp64(nsteps, checkpoint):
    """dJ/dU0 . dU (and V, W) against central differences of J(U0 + e dU)
    through nsteps RK4 steps with the filter every step; 1e-4 relative"""
    dims = (16, 16, 16)
    N = 16
    kw = adj_kw(dims, dt_fixed=0.02)
    L = kw["Lx"]
    rho, U, V, W, p = tgv_ic(N, L, p0=1.0 / 1.4 * 10)
    rng = np.random.default_rng(3)
    W = 0.2 * smooth(rng.standard_normal(U.shape))
    ic = (rho, U, V, W, p)
    g = cs3d.Solver(cs3d.config(N, N, N, precision=cs3d.FP64, **kw))
    g.set_initial_conditions(*ic)
    J, gU, gV, gW = g.ke_gradient(nsteps, checkpoint)
    J0 = run_ke(dims, ic, kw, nsteps)
    assert abs(J - J0) <= 1e-14 * abs(J0), (J, J0)
    _, _, n = g.time()
    assert n == nsteps
    eps = 1e-5
    for trial in range(3):
        d = [smooth(rng.standard_normal(U.shape), 1) for _ in range(3)]
        Jp = run_ke(dims, (rho, U + eps * d[0], V + eps * d[1], W + eps * d[2], p), kw, nsteps)
        Jm = run_ke(dims, (rho, U - eps * d[0], V - eps * d[1], W - eps * d[2], p), kw, nsteps)
        fd = (Jp - Jm) / (2 * eps)
        ad = float(np.vdot(gU, d[0]) + np.vdot(gV, d[1]) + np.vdot(gW, d[2]))
        print(f"gradient trial {trial}: adjoint {ad:.12e} fd {fd:.12e} "
              f"rel {abs(ad - fd) / abs(fd):.2e}", flush=True)
        assert abs(ad - fd) <= 1e-4 * abs(fd)


def test_ke_gradient_fp32_tracks_fp64():
    """FP32 adjoint gradient against the FP64 one (TGV 32^3, 4 steps)"""
    N = 32
    dims = (N, N, N)
    kw = adj_kw(dims, mu_ref=5e-3, dt_fixed=0.01)
    ic = tgv_ic(N, kw["Lx"], p0=10.0 / 1.4)
    grads = {}
    for prec in (cs3d.FP64, cs3d.FP32):
        g = cs3d.Solver(cs3d.config(N, N, N, precision=prec, **kw))
        g.set_initial_conditions(*ic)
        grads[prec] = g.ke_gradient(4, 2)
    J64, *g64 = grads[cs3d.FP64]
    J32, *g32 = grads[cs3d.FP32]
    assert abs(J32 - J64) <= 1e-5 * abs(J64)
    for a, b in zip(g32, g64):
        err = np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30)
        print(f"fp32 gradient rel err {err:.2e}", flush=True)
        assert err <= 1e-3


def test_adjoint_rejects_unsupported_configurations():
    N = 16
    kw = adj_kw((N, N, N))
    kw["rk"] = 3
    g = cs3d.Solver(cs3d.config(N, N, N, precision=cs3d.FP64, **kw))
    g.set_initial_conditions(*tgv_ic(N, kw["Lx"]))
    with pytest.raises(cs3d.CS3DError):
        g.ke_gradient(2, 1)
    with pytest.raises(cs3d.CS3DError):
        g.ke_gradient(0, 1)
