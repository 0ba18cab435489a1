"""Adiabatic and moving walls in the oracle (CSolver_AWS.cpp:235-495 for the
initial conditions, :553-826 preStepBCHandling, :1506-1605
postStepBCHandling; calcNeumann, Utils.hpp:52).

The pins, none of which restates the oracle's formulas:
- calcNeumann against the reference's own Utils.hpp compiled into
  oracle/_ref, and its defining property: exact for polynomials of degree
  <= 6 with zero slope at the boundary (a slope is not reproduced);
- the adiabatic wall temperature makes the one-sided 7-point derivative
  vanish (the derivative of the degree-6 interpolant through the wall and
  its six neighbours, from numpy's polyfit);
- a quiescent box with walls on every face stays at rest;
- mirror symmetry: a problem symmetric under y -> Ly - y with walls on both
  y faces stays symmetric (a face that reads its neighbours on the wrong
  side breaks it), and the same problem posed along x, or along z, gives
  the transposed result (the faces of every direction agree);
- plane Couette flow between a no-slip wall and a moving wall settles to
  the linear profile, with the wall's shear work heating the closed gas
  at the rate mu (U_w / L)^2."""
import numpy as np
import pytest

import pyoracle as po

ADIA, MOVE = po.ADIABATIC_WALL, po.MOVING_ADIABATIC_WALL


def test_calc_neumann_matches_reference_and_is_slope_free():
    rng = np.random.default_rng(0)
    ref = po.ref_lib()
    for f in rng.standard_normal((200, 6)):
        v = po.calc_neumann(f)
        if ref is not None:
            # the reference builds with -ffast-math (division by 147 taken as
            # a product), so the last bits may differ
            assert abs(v - ref.ref_calc_neumann(*f)) <= 1e-14 * max(1.0, np.abs(f).max() * 10)
    h = 0.37
    for c in rng.standard_normal((20, 7)):
        c[1] = 0.0  # no linear term
        f = lambda x: np.polyval(c[::-1], x)
        v = po.calc_neumann([f(h * m) for m in range(1, 7)])
        assert abs(v - f(0.0)) <= 1e-12 * np.abs(c).sum()
    # a slope is not reproduced: the boundary value is the zero-slope one
    v = po.calc_neumann([h * m for m in range(1, 7)])
    assert abs(v) > 0.1 * h


def wall_ic(shape, seed, Ly=None):
    """smooth random state with a strong temperature variation"""
    rng = np.random.default_rng(seed)
    Nz, Ny, Nx = shape
    z, y, x = (np.linspace(0, 2 * np.pi, n) for n in shape)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    ph = rng.uniform(0, 2 * np.pi, 8)
    rho = 1.0 + 0.05 * np.cos(X + ph[0]) * np.sin(Y + ph[1]) + 0.03 * np.cos(Z + ph[2])
    U = 0.2 * np.sin(Y + ph[3]) * np.cos(Z)
    V = 0.1 * np.cos(X + ph[4]) * np.sin(Z + ph[5])
    W = 0.15 * np.sin(X + Y + ph[6])
    p = 1.0 / 1.4 * (1.0 + 0.1 * np.cos(Y + ph[7]) * np.cos(X))
    return rho, U, V, W, p


def one_sided_slope(vals):
    """derivative at the first point of the degree-6 interpolant through 7
    equispaced values (unit spacing)"""
    return np.polyfit(np.arange(7.0), vals, 6)[-2]


def test_ic_and_pre_step_set_wall_values():
    dims = (16, 14, 12)
    vel = {1: (0.0, 0.3, -0.2), 2: (0.25, 0.0, 0.1), 5: (-0.1, 0.2, 0.0)}
    kw = dict(bc_type=(1, 1, 1), bc_kind=(ADIA, MOVE, MOVE, ADIA, ADIA, MOVE), wall_vel=vel,
              mu_ref=0.01)
    o = po.Solver(po.make_config(*dims, **kw))
    o.set_ic(*wall_ic((12, 14, 16), 1))
    R = 1.0 / 1.4
    for stage in ("ic", "rhs"):
        if stage == "rhs":
            o.pre_step()
            o.rhs(1)
        q = o.conserved(1)
        T = o.prim(3)
        rho = q[0]
        # the wall temperature has zero one-sided slope on every face
        # (away from edges, where a later face overrides an earlier one)
        assert abs(one_sided_slope(T[5, 6, :7])) < 1e-11
        assert abs(one_sided_slope(T[5, 6, ::-1][:7])) < 1e-11
        assert abs(one_sided_slope(T[5, :7, 7])) < 1e-11
        assert abs(one_sided_slope(T[5, ::-1, 7][:7])) < 1e-11
        assert abs(one_sided_slope(T[:7, 6, 7])) < 1e-11
        assert abs(one_sided_slope(T[::-1, 6, 7][:7])) < 1e-11
        # wall velocity (the last face wins on an edge: Z1 over everything)
        U = [q[1 + m] / rho for m in range(3)]
        for m in range(3):
            assert np.abs(U[m][1:-1, 1:-1, 0]).max() == 0.0
            assert np.allclose(U[m][1:-1, 1:-1, -1], vel[1][m], atol=1e-15, rtol=0)
            assert np.allclose(U[m][1:-1, 0, 1:-1], vel[2][m], atol=1e-15, rtol=0)
            assert np.abs(U[m][1:-1, -1, 1:-1]).max() == 0.0
            assert np.abs(U[m][0, 1:-1, 1:-1]).max() == 0.0
            assert np.allclose(U[m][-1, :, :], vel[5][m], atol=1e-15, rtol=0)
        # p = rho R T on the walls, and the energy carries it
        ke = 0.5 * (q[1] ** 2 + q[2] ** 2 + q[3] ** 2) / rho
        p = 0.4 * (q[4] - ke)
        for sl in (np.s_[:, :, 0], np.s_[:, -1, :], np.s_[-1, :, :]):
            assert np.allclose(p[sl], rho[sl] * R * T[sl], rtol=1e-13, atol=0)
    o.close()


def test_quiescent_box_stays_at_rest():
    dims = (16, 14, 12)
    kw = dict(bc_type=(1, 1, 1), bc_kind=(ADIA, MOVE, MOVE, ADIA, ADIA, MOVE), mu_ref=0.01)
    o = po.Solver(po.make_config(*dims, **kw))
    sh = (12, 14, 16)
    o.set_ic(np.ones(sh), np.zeros(sh), np.zeros(sh), np.zeros(sh), np.full(sh, 1 / 1.4))
    for _ in range(5):
        o.step()
    q = o.conserved(1)
    assert np.abs(q[0] - 1).max() < 1e-12
    assert max(np.abs(q[m]).max() for m in (1, 2, 3)) < 1e-12
    assert np.abs(q[4] - 1 / 0.56).max() < 1e-12
    o.close()


def run_walls(dims, ic, kinds, bct, vel, nsteps=3, **extra):
    kw = dict(bc_type=bct, bc_kind=kinds, wall_vel=vel, mu_ref=0.01, cfl=0.5, **extra)
    o = po.Solver(po.make_config(*dims, **kw))
    o.set_ic(*ic)
    for _ in range(nsteps):
        o.step()
    q = o.conserved(1)
    o.close()
    return q


def symmetric_ic(Nx, Ny, Nz):
    """symmetric under y -> Ly - y (V odd)"""
    z = np.linspace(0, 2 * np.pi, Nz, endpoint=False)
    y = np.linspace(-1, 1, Ny)
    x = np.linspace(0, 2 * np.pi, Nx, endpoint=False)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    rho = 1.0 + 0.05 * np.cos(X) * np.cos(np.pi * Y) + 0.02 * np.sin(Z)
    U = 0.2 * np.cos(0.5 * np.pi * Y) * np.sin(Z)
    V = 0.1 * np.sin(np.pi * Y) * np.cos(X + Z)
    W = 0.1 * np.cos(X) * (1 - Y ** 2)
    p = 1.0 / 1.4 * (1.0 + 0.05 * np.cos(np.pi * Y) * np.sin(X))
    return rho, U, V, W, p


def test_mirror_symmetry_and_direction_agreement():
    Nx, Ny, Nz = 12, 17, 10
    ic = symmetric_ic(Nx, Ny, Nz)
    vy = {2: (0.1, 0.0, -0.05), 3: (0.1, 0.0, -0.05)}
    qy = run_walls((Nx, Ny, Nz), ic, (0, 0, MOVE, MOVE, 0, 0), (0, 1, 0), vy,
                   Lx=2.0, Ly=1.5, Lz=2.5)
    # y -> Ly - y: rho, rhoU, rhoW, rhoE even, rhoV odd
    for v, sgn in zip(range(5), (1, 1, -1, 1, 1)):
        a, b = qy[v], sgn * qy[v][:, ::-1, :]
        assert np.abs(a - b).max() <= 1e-12 * max(np.abs(a).max(), 1e-3), v
    # the same problem along x: the arrays and velocity components with x
    # and y exchanged (the domain lengths too)
    tr = lambda a: np.ascontiguousarray(np.swapaxes(a, 1, 2))
    rho, U, V, W, p = ic
    icx = (tr(rho), tr(V), tr(U), tr(W), tr(p))
    vx = {0: (0.0, 0.1, -0.05), 1: (0.0, 0.1, -0.05)}
    qx = run_walls((Ny, Nx, Nz), icx, (MOVE, MOVE, 0, 0, 0, 0), (1, 0, 0), vx,
                   Lx=1.5, Ly=2.0, Lz=2.5)
    for vy_, vx_ in zip(range(5), (0, 2, 1, 3, 4)):
        a, b = qy[vy_], tr(qx[vx_])
        assert np.abs(a - b).max() <= 1e-12 * max(np.abs(a).max(), 1e-3), vy_
    # and along z: exchange y and z
    tz = lambda a: np.ascontiguousarray(np.swapaxes(a, 0, 1))
    icz = (tz(rho), tz(U), tz(W), tz(V), tz(p))
    vz = {4: (0.1, -0.05, 0.0), 5: (0.1, -0.05, 0.0)}
    qz = run_walls((Nx, Nz, Ny), icz, (0, 0, 0, 0, MOVE, MOVE), (0, 0, 1), vz,
                   Lx=2.0, Ly=2.5, Lz=1.5)
    for vy_, vz_ in zip(range(5), (0, 1, 3, 2, 4)):
        a, b = qy[vy_], tz(qz[vz_])
        assert np.abs(a - b).max() <= 1e-12 * max(np.abs(a).max(), 1e-3), vy_


def test_couette_flow():
    """no-slip adiabatic wall at y = 0, wall moving with U_w at y = Ly:
    the velocity settles to U_w y / Ly (the slowest transient decays like
    exp(-pi^2 nu t / Ly^2)); both walls adiabatic, so the gas heats at the
    rate of the wall's shear work, d<rhoE>/dt = mu(T) (U_w / Ly)^2."""
    Nx, Ny, Nz = 6, 17, 6
    sh = (Nz, Ny, Nx)
    Uw, mu = 0.1, 0.02
    kw = dict(Lx=1.0, Ly=1.0, Lz=1.0, bc_type=(0, 1, 0), bc_kind=(0, 0, ADIA, MOVE, 0, 0),
              wall_vel={3: (Uw, 0.0, 0.0)}, mu_ref=mu, cfl=0.5)
    o = po.Solver(po.make_config(Nx, Ny, Nz, **kw))
    o.set_ic(np.ones(sh), np.zeros(sh), np.zeros(sh), np.zeros(sh), np.full(sh, 1 / 1.4))
    y = np.arange(Ny) / (Ny - 1.0)
    def advance(t_end):
        for _ in range(4000):  # about 2100 steps reach t = 40
            if not o.time < t_end:
                break
            o.step()
        assert o.time >= t_end

    advance(30.0)
    t1, E1 = o.time, o.conserved(1)[4].mean()
    advance(40.0)
    t2, E2 = o.time, o.conserved(1)[4].mean()
    U = o.prim(0)
    u = U.mean(axis=(0, 2))
    assert np.abs(u - Uw * y).max() < 1e-3 * Uw
    assert np.abs(U - u[None, :, None]).max() < 1e-12
    assert np.abs(o.prim(1)).max() < 1e-4 * Uw and np.abs(o.prim(2)).max() < 1e-12
    T = o.prim(3).mean()
    rate = (E2 - E1) / (t2 - t1)
    expect = mu * T ** 0.76 * Uw ** 2
    assert abs(rate - expect) < 0.05 * expect, (rate, expect)
    o.close()


def test_wall_planes_keep_only_the_normal_continuity_term():
    """postStepBCHandling (CSolver_AWS.cpp:1527-1599) keeps, on a wall
    plane, only the wall-normal part of the continuity flux and no momentum
    or energy change.  Uniform flow U_w along walls moving with U_w, density
    varying along x, uniform pressure, no viscosity: rho v = 0 everywhere, so
    the wall planes' density change is exactly zero, while one point inward
    the density is advected, d rho / dt = -U_w d rho / dx."""
    Nx, Ny, Nz = 24, 12, 8
    Lx = 2.0
    Uw = 0.3
    sh = (Nz, Ny, Nx)
    x = np.arange(Nx) * (Lx / (Nx - 1.0))
    kx = 2 * np.pi / (Nx * Lx / (Nx - 1.0))  # one period over the Nx points
    rho = np.broadcast_to(1.0 + 0.1 * np.sin(kx * x), sh).copy()
    vel = {2: (Uw, 0.0, 0.0), 3: (Uw, 0.0, 0.0)}
    kw = dict(Lx=Lx, Ly=1.0, Lz=1.0, bc_type=(0, 1, 0), bc_kind=(0, 0, MOVE, MOVE, 0, 0),
              wall_vel=vel, mu_ref=0.0, cfl=0.5)
    o = po.Solver(po.make_config(Nx, Ny, Nz, **kw))
    o.set_ic(rho, np.full(sh, Uw), np.zeros(sh), np.zeros(sh), np.full(sh, 1 / 1.4))
    o.pre_step()
    o.rhs(1)
    k = o.conserved(3)
    dt = o.dt
    for j in (0, Ny - 1):
        for v in range(5):
            assert np.abs(k[v][:, j, :]).max() == 0.0, (j, v)
    adv = -dt * Uw * 0.1 * kx * np.cos(kx * x)
    assert np.abs(k[0][:, 1, :] - adv[None, :]).max() < 1e-3 * np.abs(adv).max()
    o.close()
