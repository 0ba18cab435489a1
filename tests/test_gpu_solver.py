"""GPU parity of the fused stage pipeline against the oracle: stage
right-hand side, every RK stage, whole steps with the filter round robin,
the CFL dt, the Taylor-Green diagnostics, TVDRK3, and the non-periodic
sponge configuration (including the reference's stale-primitive behaviour
on sponge planes after updateSponge)."""
import numpy as np
import pytest

import cs3d
import pyoracle as po
from helpers import random_ic, relerr, shear_ic, state_errs, tgv_ic

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12      # per field, max-norm relative, after a step
FP64_RHS_TOL = 1e-10  # dt*RHS itself (continuity cancels to O(Mach^2))
# FP32 state vs the FP64 oracle over the first steps, momenta against the
# flow's momentum scale: about 3x the largest measured error (8e-7)
FP32_TOL = 2e-6


def pair(dims, ic, prec, **kw):
    Nx, Ny, Nz = dims
    o = po.Solver(po.make_config(Nx, Ny, Nz, **kw))
    o.set_ic(*ic)
    g = cs3d.Solver(cs3d.config(Nx, Ny, Nz, precision=prec, **kw))
    g.set_initial_conditions(*ic)
    return o, g


def check_fields(tag, got, ref, tol):
    errs = state_errs(got, ref, momentum_scale=tol >= FP32_TOL)
    print(f"{tag}: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) < tol, (tag, errs)


def tgv_kw(N):
    L = 2 * np.pi - 2 * np.pi / N
    return dict(Lx=L, Ly=L, Lz=L, alpha_f=0.49, mu_ref=(1 / 1600) / 100 ** 0.76, Pr=0.71,
                cfl=0.5, filter_step=1, rk=4), L


def run_stages(o, g, prec, nstages, tol, tag):
    o.pre_step()
    g.pre_step()
    for s in range(1, nstages + 1):
        last = s == nstages
        o.sub_step(s, last)
        g.sub_step(s, last)
        which = 1 if last else 2
        check_fields(f"{tag} stage {s}", g.conserved(which), o.conserved(which), tol)
    o.post_step()
    g.post_step()


@pytest.mark.parametrize("prec", [cs3d.FP64, cs3d.FP32])
def test_tgv_rhs_stage1(prec):
    N = 32
    kw, L = tgv_kw(N)
    o, g = pair((N, N, N), tgv_ic(N, L), prec, **kw)
    o.pre_step()
    g.pre_step()
    _, dt, _ = g.time()
    assert abs(dt - o.dt) <= (1e-15 if prec == cs3d.FP64 else 1e-6) * o.dt
    o.rhs(1)
    g.rhs(1)
    check_fields("tgv rhs", g.conserved(3), o.conserved(3),
                 FP64_RHS_TOL if prec == cs3d.FP64 else 2e-3)


@pytest.mark.parametrize("prec", [cs3d.FP64, cs3d.FP32])
def test_tgv_steps(prec):
    N = 32
    kw, L = tgv_kw(N)
    o, g = pair((N, N, N), tgv_ic(N, L), prec, **kw)
    tol = FP64_TOL if prec == cs3d.FP64 else FP32_TOL
    run_stages(o, g, prec, 4, tol, f"tgv prec={prec} step 1")
    for n in range(2, 4):  # filter round robin YZX, ZXY
        o.step()
        g.step(1)
        check_fields(f"tgv prec={prec} step {n}", g.conserved(1), o.conserved(1),
                     tol * (10 if prec == cs3d.FP64 else 1))
    t, dt, ts = g.time()
    assert ts == 3 and abs(t - o.time) < 1e-15 + 1e-6 * o.time


def test_tgv_diagnostics():
    N = 32
    kw, L = tgv_kw(N)
    o, g = pair((N, N, N), tgv_ic(N, L), cs3d.FP64, **kw)
    g.set_diagnostics(True)
    for _ in range(2):
        o.step()
        g.step(1)
        d_o = o.diag_tgv()
        d_g = g.diag_taylor_green()
        print("diag oracle", d_o, "\ndiag gpu   ", d_g, flush=True)
        assert abs(d_g[0] - d_o[0]) <= 1e-15
        for c in range(1, 6):
            assert abs(d_g[c] - d_o[c]) <= 1e-11 * abs(d_o[c]), c
        assert abs(d_g[6] - d_o[6]) <= 1e-8 * abs(d_o[6])


def test_turbulence_diagnostics_fp64():
    """calcTurbulenceQuantities (CSolver_AWS.cpp:2033-2173) on a random
    solenoid-like field with TVDRK3 (the HIT configuration's integrator),
    periodic L = 2 pi (Q9), against the oracle after each of two steps."""
    N = 32
    kw = dict(Lx=2 * np.pi, Ly=2 * np.pi, Lz=2 * np.pi, alpha_f=0.49, mu_ref=2e-3, Pr=0.71,
              cfl=0.5, filter_step=1, rk=3)
    o, g = pair((N, N, N), random_ic((N, N, N), 7), cs3d.FP64, **kw)
    g.set_diagnostics(True)
    for _ in range(2):
        o.step()
        g.step(1)
        d_o = o.diag_turb()
        d_g = g.diag_turbulence()
        print("turb oracle", d_o, "\nturb gpu   ", d_g, flush=True)
        assert abs(d_g[0] - d_o[0]) <= 1e-15
        for c in range(1, 11):
            assert abs(d_g[c] - d_o[c]) <= 1e-10 * abs(d_o[c]), (c, d_g[c], d_o[c])


def test_tvdrk3_noncubic_fp64():
    dims = (32, 16, 24)
    kw = dict(Lx=2 * np.pi, Ly=np.pi, Lz=1.5 * np.pi, alpha_f=0.49, mu_ref=1e-3, Pr=0.71,
              cfl=0.5, filter_step=1, rk=3)
    o, g = pair(dims, random_ic((24, 16, 32), 1), cs3d.FP64, **kw)
    run_stages(o, g, cs3d.FP64, 3, FP64_TOL, "tvdrk3")
    for n in range(2, 5):
        o.step()
        g.step(1)
    check_fields("tvdrk3 step 4", g.conserved(1), o.conserved(1), FP64_TOL)


@pytest.mark.parametrize("prec", [cs3d.FP64, cs3d.FP32])
def test_shear_layer_sponge_nonperiodic(prec):
    """y non-periodic (Thomas with boundary closures), sponges on both y
    faces, filter every second step, several steps so that the sponge update
    and the stale-plane primitives of the first stage come into play."""
    dims = (32, 24, 16)
    kw = dict(Lx=8.0, Ly=10.0, Lz=4.0, alpha_f=0.49, mu_ref=2e-3, Pr=0.7, cfl=0.5,
              filter_step=2, rk=4, bc_type=(0, 1, 0), bc_kind=(0, 0, 1, 1, 0, 0),
              sponge_frac=(0.25, 0.15, 0.25))
    o, g = pair(dims, shear_ic(32, 24, 16, 10.0, 3), prec, **kw)
    tol = FP64_TOL if prec == cs3d.FP64 else FP32_TOL
    run_stages(o, g, prec, 4, tol, f"shear prec={prec} step 1")
    for n in range(2, 5):
        o.pre_step()
        g.pre_step()
        for s in range(1, 5):
            o.sub_step(s, s == 4)
            g.sub_step(s, s == 4)
        o.post_step()
        g.post_step()
        check_fields(f"shear prec={prec} step {n}", g.conserved(1), o.conserved(1),
                     tol * (10 if prec == cs3d.FP64 else 1))


@pytest.mark.parametrize("nz", [18, 42])
@pytest.mark.parametrize("prec", [cs3d.FP64, cs3d.FP32])
def test_z_lines_of_three_point_segments(prec, nz):
    """periodic z lines whose length divides by 3 but not by 4 (42 as in the
    shear layer) run the forward z sweep on uniform 3-point segments (6 and
    14 segments: the short-line carry loop and the unrolled terms), with
    sponges on the y faces for the stale first stage after each update"""
    dims = (32, 24, nz)
    kw = dict(Lx=8.0, Ly=10.0, Lz=0.25 * nz, alpha_f=0.49, mu_ref=2e-3, Pr=0.7, cfl=0.5,
              filter_step=2, rk=4, bc_type=(0, 1, 0), bc_kind=(0, 0, 1, 1, 0, 0),
              sponge_frac=(0.25, 0.15, 0.25))
    o, g = pair(dims, shear_ic(32, 24, nz, 10.0, 5), prec, **kw)
    tol = FP64_TOL if prec == cs3d.FP64 else FP32_TOL
    run_stages(o, g, prec, 4, tol, f"z{nz} prec={prec} step 1")
    for n in range(2, 5):
        o.step()
        g.step(1)
        check_fields(f"z{nz} prec={prec} step {n}", g.conserved(1), o.conserved(1),
                     tol * (10 if prec == cs3d.FP64 else 1))


def test_x_and_z_sponges_fp64():
    """sponges on x and z faces (x lines crossing a sponge plane)"""
    dims = (24, 16, 16)
    kw = dict(Lx=6.0, Ly=4.0, Lz=4.0, alpha_f=0.49, mu_ref=2e-3, Pr=0.7, cfl=0.5,
              filter_step=1, rk=4, bc_type=(1, 0, 1), bc_kind=(1, 1, 0, 0, 1, 1),
              sponge_frac=(0.25, 0.25, 0.25))
    o, g = pair(dims, random_ic((16, 16, 24), 4, amp=0.2), cs3d.FP64, **kw)
    for n in range(1, 4):
        o.step()
        g.step(1)
        check_fields(f"xz sponge step {n}", g.conserved(1), o.conserved(1), 10 * FP64_TOL)


def test_nonfinite_state_is_reported():
    N = 16
    kw, L = tgv_kw(N)
    rho, U, V, W, p = tgv_ic(N, L)
    U = U.copy()
    U[3, 4, 5] = np.nan
    g = cs3d.Solver(cs3d.config(N, N, N, precision=cs3d.FP64, **kw))
    g.set_initial_conditions(rho, U, V, W, p)
    with pytest.raises(cs3d.CS3DError) as e:
        g.step(1)
    assert e.value.code == cs3d.ERR_NONFINITE


def test_fp32_kinetic_energy_tracks_fp64_1000_steps():
    """The accuracy target at reduced size: TGV 64^3, RK4 with the filter
    every step, 1000 steps; FP32 volume-averaged kinetic energy within 1e-5
    relative of the FP64 run."""
    N = 64
    kw, L = tgv_kw(N)
    ic = tgv_ic(N, L)
    ke = {}
    for prec in (cs3d.FP64, cs3d.FP32):
        g = cs3d.Solver(cs3d.config(N, N, N, precision=prec, **kw))
        g.set_initial_conditions(*ic)
        g.step(1000)
        ke[prec] = g.kinetic_energy()
        print("KE", prec, ke[prec], flush=True)
    assert abs(ke[cs3d.FP32] - ke[cs3d.FP64]) <= 1e-5 * abs(ke[cs3d.FP64])


@pytest.mark.parametrize("prec,sponge", [(cs3d.FP64, False), (cs3d.FP32, False),
                                         (cs3d.FP64, True)])
def test_dump_and_restart_continue_the_run(tmp_path, prec, sponge):
    """dump after 3 steps, restart a fresh solver from the files, 3 more
    steps: the same state as 6 uninterrupted steps.  The dump holds full
    FP64 values while the solver stores perturbations of a reference state,
    so the restart reproduces the run to rounding (1e-13 FP64, 2e-6 FP32)."""
    if sponge:
        dims = (24, 16, 16)
        kw = dict(Lx=6.0, Ly=4.0, Lz=4.0, alpha_f=0.49, mu_ref=2e-3, Pr=0.7, cfl=0.5,
                  filter_step=2, rk=4, bc_type=(0, 1, 0), bc_kind=(0, 0, 1, 1, 0, 0),
                  sponge_frac=(0.25, 0.25, 0.25))
        icf = random_ic((16, 16, 24), 6, amp=0.2)
    else:
        N = 24
        kw, L = tgv_kw(N)
        kw["filter_step"] = 2
        dims = (N, N, N)
        icf = tgv_ic(N, L)
    Nx, Ny, Nz = dims
    a = cs3d.Solver(cs3d.config(Nx, Ny, Nz, precision=prec, **kw))
    a.set_initial_conditions(*icf)
    a.step(3)
    n = a.dump(str(tmp_path))
    assert n == 3
    a.step(3)
    b = cs3d.Solver(cs3d.config(Nx, Ny, Nz, precision=prec, **kw))
    b.restart(str(tmp_path), 3)
    b.step(3)
    ta, tb = a.time(), b.time()
    assert ta[2] == tb[2] == 6
    assert abs(ta[0] - tb[0]) <= 1e-13 * abs(ta[0])
    errs = [relerr(x, y) for x, y in zip(b.conserved(1), a.conserved(1))]
    print(f"restart prec={prec} sponge={sponge}: " + " ".join(f"{e:.2e}" for e in errs),
          flush=True)
    assert max(errs) <= (1e-13 if prec == cs3d.FP64 else 2e-6)


@pytest.mark.parametrize("prec", [cs3d.FP64, cs3d.FP32])
def test_shear_layer_profiles_closed_forms(prec):
    """shearLayerInfoCalc's plane statistics on a field whose x-z means and
    moments are exact: s_x = (-1)^i, s_z = (-1)^k (even Nx, Nz);
    U = a(y) + c s_x, V = d s_z, W = e s_x + f s_z, rho = 1 + g(y), p
    uniform.  Q12: the reference's uwprime is u'w' + v'w' = c e + d f and its
    vwprime 0; fixed: c e and d f."""
    Nx, Ny, Nz = 16, 12, 8
    kw = dict(Lx=6.0, Ly=4.0, Lz=4.0, alpha_f=0.49, mu_ref=2e-3, Pr=0.7, cfl=0.5,
              filter_step=1, rk=4)
    y = np.arange(Ny) * (4.0 / (Ny - 1))
    sx = (-1.0) ** np.arange(Nx)
    sz = (-1.0) ** np.arange(Nz)
    Sx = np.broadcast_to(sx[None, None, :], (Nz, Ny, Nx))
    Sz = np.broadcast_to(sz[:, None, None], (Nz, Ny, Nx))
    a = 0.3 * np.tanh(y - 2.0)
    gy = 0.05 * np.cos(y)
    c, d, e, f = 0.1, 0.07, 0.05, 0.03
    U = a[None, :, None] + c * Sx
    V = d * Sz
    W = e * Sx + f * Sz
    rho = np.broadcast_to(1.0 + gy[None, :, None], (Nz, Ny, Nx)).copy()
    p = np.full((Nz, Ny, Nx), 1.0 / 1.4)
    g = cs3d.Solver(cs3d.config(Nx, Ny, Nz, precision=prec, **kw))
    g.set_initial_conditions(rho, U, V, W, p)
    tol = 1e-12 if prec == cs3d.FP64 else 2e-6
    for fix in (False, True):
        o = g.diag_shear_layer(fix_q12=fix)
        assert np.allclose(o["uAvg"], a, atol=tol, rtol=0)
        assert np.allclose(o["rhoAvg"], 1.0 + gy, atol=tol, rtol=0)
        assert np.allclose(o["uprime2"], c * c, atol=tol, rtol=tol)
        assert np.allclose(o["vprime2"], d * d, atol=tol, rtol=tol)
        assert np.allclose(o["wprime2"], e * e + f * f, atol=tol, rtol=tol)
        assert np.allclose(o["uvprime"], 0.0, atol=tol)
        assert np.allclose(o["rhoprime2"], 0.0, atol=tol)
        assert np.allclose(o["uwprime"], c * e + (d * f if not fix else 0.0), atol=tol, rtol=tol)
        assert np.allclose(o["vwprime"], d * f if fix else 0.0, atol=tol, rtol=tol)
        # p is uniform and T varies in y only: both moments vanish to rounding
        assert np.all(o["pprime2"] < 1e-10) and np.all(o["tprime2"] < 1e-10)
