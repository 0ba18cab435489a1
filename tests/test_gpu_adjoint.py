"""The discrete adjoint on the GPU.

The reference has no adjoint, so nothing of the reference pins it; the
mathematics does.  The vector-Jacobian product of the stage map
Q -> dt*RHS(Q) must satisfy the dot-product identity
<lam, J delta> = <J^T lam, delta> against central differences of the
forward stage (itself pinned to the oracle by test_gpu_solver), and the
gradient of the final kinetic energy through RK4 steps with the filter
must match central differences of whole forward runs along random
directions (dt held fixed, so the forward map is the one differentiated).
Field by field, both agree with the FP64 numpy adjoint of the oracle
(oracle/adjoint.py, dense transposed operators, pinned on the CPU)."""
import numpy as np
import pytest

import adjoint as oa
import cs3d
import pyoracle as po
from helpers import random_ic, relerr, shear_ic, smooth, tgv_ic

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


@pytest.mark.parametrize("nsteps,checkpoint,rk", [(6, 4, 4), (5, 1, 4), (6, 4, 3), (5, 2, 3)])
def test_ke_gradient_matches_finite_differences_fp64(nsteps, checkpoint, rk):
    """dJ/dU0 . dU (and V, W) against central differences of J(U0 + e dU)
    through nsteps RK4 or TVDRK3 steps with the filter every step; 1e-4
    relative"""
    dims = (16, 16, 16)
    N = 16
    kw = adj_kw(dims, dt_fixed=0.02, rk=rk)
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


def test_adjoint_rejects_step_counts_below_one():
    N = 16
    kw = adj_kw((N, N, N))
    g = cs3d.Solver(cs3d.config(N, N, N, precision=cs3d.FP64, **kw))
    g.set_initial_conditions(*tgv_ic(N, kw["Lx"]))
    with pytest.raises(cs3d.CS3DError):
        g.ke_gradient(0, 1)
    g.close()


@pytest.mark.parametrize("rk", [4, 3])
def test_ke_gradient_compensated_fp32_tracks_fp64(rk):
    """the compensated FP32 mode: the replays retrace its forward pass
    (the state's low parts checkpointed) and the products take the FP32
    stage states; the gradient against the FP64 one as for plain FP32, and
    the same J as the mode's own forward run"""
    N = 32
    kw = adj_kw((N, N, N), mu_ref=5e-3, dt_fixed=0.01, rk=rk)
    ic = tgv_ic(N, kw["Lx"], p0=10.0 / 1.4)
    g = cs3d.Solver(cs3d.config(N, N, N, precision=cs3d.FP64, **kw))
    g.set_initial_conditions(*ic)
    J64, *g64 = g.ke_gradient(5, 2)
    g.close()
    g = cs3d.Solver(cs3d.config(N, N, N, precision=cs3d.FP32, compensated=1, **kw))
    g.set_initial_conditions(*ic)
    J, *gc = g.ke_gradient(5, 2)
    g.close()
    g = cs3d.Solver(cs3d.config(N, N, N, precision=cs3d.FP32, compensated=1, **kw))
    g.set_initial_conditions(*ic)
    g.step(5)
    assert J == g.kinetic_energy()
    g.close()
    assert abs(J - J64) <= 1e-6 * abs(J64)
    for a, b in zip(gc, g64):
        err = np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30)
        print(f"compensated rk{rk} gradient rel err {err:.2e}", flush=True)
        assert err <= 1e-3


def oracle_cfg(dims, kw):
    Nx, Ny, Nz = dims
    keep = {k: v for k, v in kw.items() if k in ("Lx", "Ly", "Lz", "alpha_f", "mu_ref", "Pr",
                                                  "filter_step", "rk")}
    return po.make_config(Nx, Ny, Nz, **keep)


@pytest.mark.parametrize("dims", [(20, 12, 14), (18, 10, 14)])
def test_rhs_vjp_matches_oracle_adjoint_fp64(dims):
    """(18, 10, 14): periodic lines of unequal segments in every direction,
    so every pass runs its table variant"""
    Nx, Ny, Nz = dims
    shape = (Nz, Ny, Nx)
    kw = adj_kw(dims, dt_fixed=0.013)
    g = cs3d.Solver(cs3d.config(Nx, Ny, Nz, precision=cs3d.FP64, **kw))
    g.set_initial_conditions(*random_ic(shape, seed=21, amp=0.4))
    g.pre_step()
    q = g.conserved(1)
    rng = np.random.default_rng(8)
    lam = [rng.standard_normal(shape) for _ in range(5)]
    got = g.rhs_vjp(lam, 1)
    ref = oa.Adjoint(oracle_cfg(dims, kw)).rhs_vjp(q, lam, 0.013)
    errs = [relerr(a, b) for a, b in zip(got, ref)]
    print("vjp vs oracle adjoint:", " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 1e-11, errs


@pytest.mark.parametrize("filter_step,rk", [(1, 4), (2, 4), (1, 3), (2, 3)])
def test_ke_gradient_matches_oracle_adjoint_fp64(filter_step, rk):
    dims = (16, 12, 20)
    Nx, Ny, Nz = dims
    shape = (Nz, Ny, Nx)
    kw = adj_kw(dims, dt_fixed=0.02, filter_step=filter_step, rk=rk)
    ic = random_ic(shape, seed=23, amp=0.5)
    g = cs3d.Solver(cs3d.config(Nx, Ny, Nz, precision=cs3d.FP64, **kw))
    g.set_initial_conditions(*ic)
    J, *grad = g.ke_gradient(5, 2)
    Jo, *gref = oa.Adjoint(oracle_cfg(dims, kw)).ke_gradient(ic, 5, 0.02, filter_step)
    assert abs(J - Jo) <= 1e-13 * abs(Jo)
    errs = [relerr(a, b) for a, b in zip(grad, gref)]
    print("gradient vs oracle adjoint:", " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 1e-10, errs


def test_ke_gradient_tgv128_fp32_200_steps():
    """Taylor-Green 128^3, FP32 adjoint through 200 RK4 steps with the
    filter every step, checkpoints every 20 steps, against FP64 central
    differences (eps 1e-3) along 8 seeded Gaussian directions; the dt of
    the run is frozen at the first step's CFL dt.  1e-4 relative."""
    N = 128
    dims = (N, N, N)
    kw = adj_kw(dims, mu_ref=(1 / 1600) / 100 ** 0.76, alpha_f=0.49)
    ic = tgv_ic(N, kw["Lx"])
    probe = cs3d.Solver(cs3d.config(N, N, N, precision=cs3d.FP64, **kw))
    probe.set_initial_conditions(*ic)
    probe.pre_step()
    dt = probe.time()[1]
    probe.close()
    kw["dt_fixed"] = dt
    g = cs3d.Solver(cs3d.config(N, N, N, precision=cs3d.FP32, **kw))
    g.set_initial_conditions(*ic)
    J, gU, gV, gW = g.ke_gradient(200, 20)
    g.close()
    print(f"J = {J:.10e}, dt = {dt:.6e}", flush=True)
    rho, U, V, W, p = ic
    eps = 1e-3
    worst = 0.0
    for seed in range(8):
        rng = np.random.default_rng(seed)
        d = [rng.standard_normal(U.shape) for _ in range(3)]
        Js = []
        for sgn in (1, -1):
            f = cs3d.Solver(cs3d.config(N, N, N, precision=cs3d.FP64, **kw))
            f.set_initial_conditions(rho, U + sgn * eps * d[0], V + sgn * eps * d[1],
                                     W + sgn * eps * d[2], p)
            f.step(200)
            Js.append(f.kinetic_energy())
            f.close()
        fd = (Js[0] - Js[1]) / (2 * eps)
        ad = float(np.vdot(gU, d[0]) + np.vdot(gV, d[1]) + np.vdot(gW, d[2]))
        rel = abs(ad - fd) / abs(fd)
        worst = max(worst, rel)
        print(f"direction {seed}: adjoint {ad:.10e} fd {fd:.10e} rel {rel:.2e}", flush=True)
    assert worst <= 1e-4


def test_autograd_binding_fp64():
    """torch.autograd through the adjoint: backward of J gives the adjoint
    gradient, and a torch-side chain rule on top of it matches finite
    differences of the composite"""
    import torch
    import cs3d_autograd as ag
    N = 16
    dims = (N, N, N)
    kw = adj_kw(dims, dt_fixed=0.02)
    rho, U, V, W, p = tgv_ic(N, kw["Lx"], p0=10.0 / 1.4)
    cfg = cs3d.config(N, N, N, precision=cs3d.FP64, **kw)
    a = torch.tensor(0.7, dtype=torch.float64, requires_grad=True)
    Ut, Vt, Wt = (torch.from_numpy(x.copy()) for x in (U, V, W))
    J = ag.kinetic_energy(a * Ut, a * Vt, Wt, cfg, rho, p, 3, 2)
    J.backward()
    eps = 1e-6
    Jp = ag.kinetic_energy((a.detach() + eps) * Ut, (a.detach() + eps) * Vt, Wt, cfg, rho, p, 3, 2)
    Jm = ag.kinetic_energy((a.detach() - eps) * Ut, (a.detach() - eps) * Vt, Wt, cfg, rho, p, 3, 2)
    fd = float(Jp - Jm) / (2 * eps)
    print(f"autograd dJ/da {a.grad.item():.12e} fd {fd:.12e}", flush=True)
    assert abs(a.grad.item() - fd) <= 1e-6 * abs(fd)


def test_gradient_repeatable_across_pool_reuse_and_release():
    """the state pool is reused by a second call with other sizes, and
    after cs3d_adjoint_release the next call allocates again; all give the
    same gradient for the same initial state"""
    N = 16
    kw = adj_kw((N, N, N), dt_fixed=0.02)
    ic = tgv_ic(N, kw["Lx"], p0=10.0 / 1.4)
    g = cs3d.Solver(cs3d.config(N, N, N, precision=cs3d.FP64, **kw))
    out = []
    for k, release in ((2, False), (3, False), (1, True), (2, False)):
        g.set_initial_conditions(*ic)
        out.append(g.ke_gradient(4, k))
        if release:
            g.adjoint_release()
    for J, *gr in out[1:]:
        assert abs(J - out[0][0]) <= 1e-15 * abs(J)
        for a, b in zip(gr, out[0][1:]):
            assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))
