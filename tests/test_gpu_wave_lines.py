"""Lines of 256 points along x: a line is 64 segments of 4 points, one
wavefront.  The x filter, fused into the last stage in both precisions on
such lines, then takes its cross-segment carries from a wavefront scan
instead of the LDS ring.  Against the oracle: FP64 within 1e-12 per field
after every step (TVDRK3 and RK4, with y sponges, with y walls); FP32
within FP32 rounding of the same steps, and against the same grid run with
the filter unfused (truncated LDS carries)."""
import numpy as np
import pytest

import cs3d
import pyoracle as po
from helpers import random_ic, relerr, shear_ic, state_errs

pytestmark = pytest.mark.gpu

DIMS = (256, 8, 8)
KW = dict(Lx=8.0, Ly=0.25, Lz=0.25, alpha_f=0.49, mu_ref=2e-3, Pr=0.7, cfl=0.5,
          filter_step=1, bc_type=(0, 0, 0))


def ic(dims=DIMS, Lx=KW["Lx"]):
    Nx, Ny, Nz = dims
    rho, U, V, W, p = random_ic((Nz, Ny, Nx), 21, amp=0.2)
    # y and z lines of 8 points: keep the state smooth along them
    x = np.arange(Nx) * (Lx / (Nx - 1))
    U = U + 0.3 * np.sin(2 * np.pi * x / Lx)[None, None, :]
    return rho, U, V, W, p


def check(tag, got, ref, tol):
    # FP32 (tol above 1e-9): momenta against the flow's momentum scale
    errs = state_errs(got, ref, momentum_scale=tol > 1e-9)
    print(f"{tag}: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) < tol, (tag, errs)


@pytest.mark.parametrize("rk", [4, 3])
def test_x256_fp64_matches_oracle(rk):
    o = po.Solver(po.make_config(*DIMS, rk=rk, **KW))
    g = cs3d.Solver(cs3d.config(*DIMS, precision=cs3d.FP64, rk=rk, **KW))
    q0 = ic()
    o.set_ic(*q0)
    g.set_initial_conditions(*q0)
    for n in range(1, 4):
        o.step()
        g.step(1)
        check(f"x256 fp64 rk{rk} step {n}", g.conserved(1), o.conserved(1), 1e-12)
    o.close()
    g.close()


@pytest.mark.parametrize("Nx", [256, 512])
def test_x256_fp32_fused_filter_matches_oracle(monkeypatch, Nx):
    """FP32 with the x filter fused into the last stage against the oracle,
    and against the same FP32 run with the filter in separate sweeps.  At
    256 points the fused filter's carries come from the wavefront scan; at
    512 (128 segments, a workgroup of two wavefronts per line) from the LDS
    ring, which shares its LDS with the filter's halo staging."""
    dims = (Nx, 8, 8)
    kw = dict(KW, Lx=KW["Lx"] * Nx / 256)
    o = po.Solver(po.make_config(*dims, rk=4, **kw))
    q0 = ic(dims, kw["Lx"])
    o.set_ic(*q0)
    for _ in range(3):
        o.step()
    ref = o.conserved(1)
    o.close()
    out = {}
    for fused in (True, False):
        if fused:
            monkeypatch.delenv("CS3D_NO_FILTER_FUSION", raising=False)
        else:
            monkeypatch.setenv("CS3D_NO_FILTER_FUSION", "1")
        g = cs3d.Solver(cs3d.config(*dims, precision=cs3d.FP32, rk=4, **kw))
        g.set_initial_conditions(*q0)
        g.step(3)
        out[fused] = g.conserved(1)
        g.close()
    # rho' and rhoE' are perturbations of O(0.01) on an O(1) state
    check(f"x{Nx} fp32 fused vs oracle", out[True], ref, 2e-6)
    check(f"x{Nx} fp32 unfused vs oracle", out[False], ref, 2e-6)
    check(f"x{Nx} fp32 fused vs unfused", out[True], out[False], 2e-6)


@pytest.mark.parametrize("prec", [cs3d.FP64, cs3d.FP32])
def test_x256_sponges_matches_oracle(prec):
    """256-point periodic x lines with sponges on both (non-periodic) y
    faces, filter every second step: the fused x filter with the sponge
    forcing, and the stale-primitive first stage after each sponge update"""
    dims = (256, 24, 8)
    kw = dict(Lx=32.0, Ly=10.0, Lz=2.0, alpha_f=0.49, mu_ref=2e-3, Pr=0.7, cfl=0.5,
              filter_step=2, rk=4, bc_type=(0, 1, 0), bc_kind=(0, 0, 1, 1, 0, 0),
              sponge_frac=(0.25, 0.15, 0.25))
    q0 = shear_ic(*dims, 10.0, 5)
    o = po.Solver(po.make_config(*dims, **kw))
    o.set_ic(*q0)
    g = cs3d.Solver(cs3d.config(*dims, precision=prec, **kw))
    g.set_initial_conditions(*q0)
    tol = 1e-12 if prec == cs3d.FP64 else 2e-6
    for n in range(1, 5):
        o.step()
        g.step(1)
        check(f"x256 sponges prec={prec} step {n}", g.conserved(1), o.conserved(1), tol)
    o.close()
    g.close()


@pytest.mark.parametrize("prec", [cs3d.FP64, cs3d.FP32])
def test_x256_walls_matches_oracle(prec):
    """256-point periodic x lines between an adiabatic and a moving wall in
    y (z periodic), filter every step: the wall rows of the x sweep with the
    fused x filter on 64-segment lines"""
    dims = (256, 16, 12)
    kw = dict(Lx=16.0, Ly=1.5, Lz=1.0, alpha_f=0.49, mu_ref=2e-3, Pr=0.7, cfl=0.5,
              filter_step=1, rk=4, bc_type=(0, 1, 0),
              bc_kind=(0, 0, cs3d.BC_ADIABATIC_WALL, cs3d.BC_MOVING_ADIABATIC_WALL, 0, 0),
              wall_vel={3: (0.1, 0.0, 0.05)})
    q0 = random_ic((12, 16, 256), 13, amp=0.1)
    o = po.Solver(po.make_config(*dims, **kw))
    o.set_ic(*q0)
    g = cs3d.Solver(cs3d.config(*dims, precision=prec, **kw))
    g.set_initial_conditions(*q0)
    tol = 1e-12 if prec == cs3d.FP64 else 2e-6
    for n in range(1, 4):
        o.step()
        g.step(1)
        check(f"x256 walls prec={prec} step {n}", g.conserved(1), o.conserved(1), tol)
    o.close()
    g.close()


@pytest.mark.parametrize("prec", [cs3d.FP64, cs3d.FP32])
def test_x320_three_lines_per_workgroup_matches_oracle(prec):
    """320-point x lines (80 segments): the x sweep takes three lines per
    workgroup (four wavefronts), and 128 lines leave the last workgroup
    with two lines and an idle one; sponges on the y faces put some of the
    lines on sponge planes.  FP64 within 1e-12 of the oracle per field
    after every step, FP32 (with the fused x filter) within 2e-6."""
    dims = (320, 16, 8)
    kw = dict(Lx=40.0, Ly=8.0, Lz=2.0, alpha_f=0.49, mu_ref=2e-3, Pr=0.7, cfl=0.5,
              filter_step=1, rk=4, bc_type=(0, 1, 0), bc_kind=(0, 0, 1, 1, 0, 0),
              sponge_frac=(0.25, 0.2, 0.25))
    q0 = shear_ic(*dims, 8.0, 6)
    o = po.Solver(po.make_config(*dims, **kw))
    o.set_ic(*q0)
    g = cs3d.Solver(cs3d.config(*dims, precision=prec, **kw))
    g.set_initial_conditions(*q0)
    tol = 1e-12 if prec == cs3d.FP64 else 2e-6
    for n in range(1, 4):
        o.step()
        g.step(1)
        check(f"x320 prec={prec} step {n}", g.conserved(1), o.conserved(1), tol)
    o.close()
    g.close()
