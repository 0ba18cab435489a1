"""The oracle against what the mathematics fixes: dense solves, the modified
wavenumber / transfer function of the periodic compact operators, polynomial
exactness of the boundary closures, and invariants of the full step."""
import numpy as np
import pytest

import pyoracle as po
from helpers import relerr, random_ic, tgv_ic


def dense(a, b, c, corner_lo=0.0, corner_hi=0.0):
    n = len(b)
    A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    A[n - 1, 0] += corner_lo
    A[0, n - 1] += corner_hi
    return A


@pytest.mark.parametrize("n", [8, 17, 64, 256])
def test_thomas_and_cyclic_dense(n):
    rng = np.random.default_rng(n)
    a, c = rng.uniform(-1, 1, (2, n))
    b = 3 + rng.random(n)
    d = rng.standard_normal(n)
    assert relerr(po.solve_tri(a, b, c, d), np.linalg.solve(dense(a, b, c), d)) < 1e-13
    al, be = 0.8, -0.3
    x = po.cyclic(a, b, c, al, be, d)
    assert relerr(x, np.linalg.solve(dense(a, b, c, al, be), d)) < 1e-13


@pytest.mark.parametrize("n", [16, 64, 256])
@pytest.mark.parametrize("k", [1, 3, 7])
def test_periodic_modified_wavenumber(n, k):
    """D1: k'h = (a sin th + b/2 sin 2th)/(1 + 2 al cos th);
    D2: k''h^2 = (2a(1-cos th) + b/2 (1 - cos 2th))/(1 + 2 al cos th);
    filter: T = sum_n a_n cos(n th) / (1 + 2 af cos th)."""
    if k >= n // 2:
        pytest.skip("mode above Nyquist")
    h = 2 * np.pi / n
    x = np.arange(n) * h
    th = k * h
    f = np.sin(k * x)
    a1, b1, al1 = 14 / 9, 1 / 9, 1 / 3
    kp = (a1 * np.sin(th) + b1 / 2 * np.sin(2 * th)) / (1 + 2 * al1 * np.cos(th)) / h
    assert relerr(po.line_op(1, 0, f, dh=h), kp * np.cos(k * x)) < 1e-12
    a2, b2, al2 = 12 / 11, 3 / 11, 2 / 11
    kpp = (2 * a2 * (1 - np.cos(th)) + b2 / 2 * (1 - np.cos(2 * th))) / (1 + 2 * al2 * np.cos(th)) / h ** 2
    # the second derivative of a low mode cancels to O((kh)^2): roundoff
    # grows like 1/(kh)^2
    assert relerr(po.line_op(2, 0, f, dh=h), -kpp * np.sin(k * x)) < max(1e-12, 1e-14 / th ** 2)
    af = 0.49
    an = [(93 + 70 * af) / 128, (7 + 18 * af) / 16, (-7 + 14 * af) / 32, (1 - 2 * af) / 16,
          (-1 + 2 * af) / 128]
    T = sum(an[m] * np.cos(m * th) for m in range(5)) / (1 + 2 * af * np.cos(th))
    assert relerr(po.line_op(0, 0, f, alpha_f=af), T * f) < 1e-12


def test_filter_transfer_endpoints():
    n = 32
    for af in (0.49, 0.495, 0.3):
        assert np.allclose(po.line_op(0, 0, np.ones(n), alpha_f=af), 1.0, atol=1e-14)
        nyq = (-1.0) ** np.arange(n)
        assert np.abs(po.line_op(0, 0, nyq, alpha_f=af)).max() < 1e-12
        # non-periodic filter keeps constants too (every row sums to its LHS)
        assert np.allclose(po.line_op(0, 1, np.ones(n), alpha_f=af), 1.0, atol=1e-13)


@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4])
def test_nonperiodic_closures_polynomial_exact(deg):
    """Boundary closures of Derivatives.hpp:83-120 are at least 4th order:
    polynomials up to degree 4 are differentiated exactly."""
    n = 24
    h = 0.1
    x = np.arange(n) * h
    f = x ** deg
    d1 = deg * x ** (deg - 1) if deg > 0 else np.zeros(n)
    d2 = deg * (deg - 1) * x ** (deg - 2) if deg > 1 else np.zeros(n)
    assert np.abs(po.line_op(1, 1, f, dh=h) - d1).max() < 1e-9 * max(1, np.abs(d1).max())
    assert np.abs(po.line_op(2, 1, f, dh=h) - d2).max() < 1e-7 * max(1, np.abs(d2).max())


def test_uniform_state_has_zero_rhs():
    N = 16
    cfg = po.make_config(N, N, N, Lx=6.0, Ly=6.0, Lz=6.0, mu_ref=1e-3)
    s = po.Solver(cfg)
    sh = (N, N, N)
    s.set_ic(np.full(sh, 1.2), np.full(sh, 0.3), np.full(sh, -0.1), np.full(sh, 0.2), np.full(sh, 0.8))
    s.pre_step()
    s.rhs(1)
    for v in range(5):
        assert np.abs(s.field(v, 3)).max() < 1e-12


def test_periodic_step_conserves_mass():
    N = 24
    L = 2 * np.pi - 2 * np.pi / N
    cfg = po.make_config(N, N, N, Lx=L, Ly=L, Lz=L, mu_ref=1e-3, alpha_f=0.49)
    s = po.Solver(cfg)
    s.set_ic(*random_ic((N, N, N), 7))
    m0 = [s.field(v).sum() for v in range(5)]
    for _ in range(3):
        s.step()
    m1 = [s.field(v).sum() for v in range(5)]
    # mass of a periodic box is conserved by the flux form of the continuity
    # equation and by the filter (circulant operators: 1^T A^-1 B = 0 and
    # 1^T F = 1^T).  Momentum and energy are not conserved exactly: the
    # viscous terms are in expanded (non-conservative) form
    # (CSolver_AWS.cpp:1311-1312, 1464-1487).
    assert abs(m1[0] - m0[0]) < 1e-13 * abs(m0[0])


def test_filter_directions_commute():
    rng = np.random.default_rng(5)
    f = rng.standard_normal((12, 10, 16))
    xyz = po.field_op(0, 2, 1, po.field_op(0, 1, 0, po.field_op(0, 0, 0, f)))
    zxy = po.field_op(0, 1, 0, po.field_op(0, 0, 0, po.field_op(0, 2, 1, f)))
    assert relerr(xyz, zxy) < 1e-13


def test_tgv_initial_kinetic_energy_closed_form():
    """sum 1/2 rho |u|^2 over the periodic grid is exactly N^3 U0^2/8 for
    uniform rho (the first recorded row of taylorgreen.out is one step later)."""
    N = 16
    L = 2 * np.pi - 2 * np.pi / N
    rho, U, V, W, p = tgv_ic(N, L, isothermal=False)
    ke = 0.5 * (rho * (U * U + V * V + W * W)).sum()
    assert abs(ke - N ** 3 / 8) < 1e-9


def test_turbulence_quantities_tgv_closed_forms():
    """calcTurbulenceQuantities (CSolver_AWS.cpp:2033-2173) pinned on the
    Taylor-Green field at t ~ 0 (one step with dt = 1e-10, so the stage
    derivatives are those of the initial field).  With u = sin x cos y cos z,
    v = -cos x sin y cos z, w = 0, p0 = 100/gamma and rho = p/p0:
    <u.u> = 1/4, <Ux^2 + Vy^2 + Wz^2> = 1/4, div u = 0, <|w|^2> = 3/4,
    2 S:grad u = 2 (Ux^2 + Vy^2 + Uz^2/2 + Vz^2/2) averages to 3/4, the sound
    speed is sqrt(gamma p0) = 10 everywhere, T = 100 so mu = 1/1600, and
    std(rho) = sqrt(4.5)/(16 p0).  The compact derivatives carry the 6th-order
    modified-wavenumber error, ~1e-6 at 32 points per period."""
    N = 32
    L = 2 * np.pi - 2 * np.pi / N
    p0 = 100.0 / 1.4
    o = po.Solver(po.make_config(N, N, N, Lx=L, Ly=L, Lz=L, alpha_f=0.49,
                                 mu_ref=(1 / 1600) / 100 ** 0.76, Pr=0.71, cfl=0.5,
                                 filter_step=1000, rk=4))
    o.set_ic(*tgv_ic(N, L, p0=p0))
    o.set_dt(1e-10)
    o.step()
    t, tau, re3, mt, td, up, kol, rp, dp, vp, ke = o.diag_turb()
    nu = (1 / 1600) / 1.0
    assert t == 0.0
    assert abs(up - np.sqrt(1 / 12)) < 1e-8
    assert abs(tau - np.sqrt(2 * np.pi) / 8 / np.sqrt(1 / 12)) < 1e-7
    assert abs(mt - 0.05) < 1e-8
    assert abs(td - 2 * nu * 3 / 8) < 1e-5 * 2 * nu * 3 / 8
    assert abs(re3 - np.sqrt(1 / 12) / nu) < 1e-5 * np.sqrt(1 / 12) / nu
    assert abs(kol - (nu ** 3 / (2 * nu * 3 / 8)) ** 0.25) < 1e-5 * kol
    assert abs(rp - np.sqrt(4.5) / (16 * p0)) < 1e-6 * rp
    assert dp < 1e-10
    assert abs(vp - np.sqrt(3 / 4)) < 1e-5
    assert abs(ke - 1 / 8) < 1e-3
