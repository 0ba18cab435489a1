"""Adiabatic and moving walls on the GPU against the oracle (FP64 to
1e-12 per field after each step; FP32 tracking FP64): every face a wall
(edges and corners, where the later face of the reference's order wins),
the stage right-hand side alone, TVDRK3, walls beside sponges (the
stale-primitive first stage after a sponge update reads wall neighbours on
sponge planes), and plane Couette flow in FP32."""
import numpy as np
import pytest

import cs3d
import pyoracle as po
from helpers import random_ic, relerr, state_errs

pytestmark = pytest.mark.gpu

ADIA, MOVE = cs3d.BC_ADIABATIC_WALL, cs3d.BC_MOVING_ADIABATIC_WALL
FP64_TOL = 1e-12
VEL = {0: (0.0, 0.2, -0.1), 1: (0.0, -0.15, 0.05), 2: (0.1, 0.0, 0.2),
       3: (-0.2, 0.0, 0.1), 4: (0.05, -0.1, 0.0), 5: (0.15, 0.1, 0.0)}


def pair(dims, ic, prec, **kw):
    Nx, Ny, Nz = dims
    o = po.Solver(po.make_config(Nx, Ny, Nz, **kw))
    o.set_ic(*ic)
    g = cs3d.Solver(cs3d.config(Nx, Ny, Nz, precision=prec, **kw))
    g.set_initial_conditions(*ic)
    return o, g


def check(tag, got, ref, tol):
    # FP32 (tol above 1e-9): momenta against the flow's momentum scale
    errs = state_errs(got, ref, momentum_scale=tol > 1e-9)
    print(f"{tag}: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) < tol, (tag, errs)


@pytest.mark.parametrize("kinds,rk", [((ADIA, MOVE, MOVE, ADIA, ADIA, MOVE), 4),
                                      ((MOVE, ADIA, ADIA, MOVE, MOVE, ADIA), 3),
                                      ((MOVE,) * 6, 4)])
def test_walls_on_every_face_fp64(kinds, rk):
    dims = (24, 20, 16)
    kw = dict(Lx=3.0, Ly=2.5, Lz=2.0, alpha_f=0.49, mu_ref=2e-3, Pr=0.7, cfl=0.5,
              filter_step=1, rk=rk, bc_type=(1, 1, 1), bc_kind=kinds, wall_vel=VEL)
    o, g = pair(dims, random_ic((16, 20, 24), 7, amp=0.2), cs3d.FP64, **kw)
    check(f"walls {kinds} rk{rk} ic", g.conserved(1), o.conserved(1), FP64_TOL)
    for n in range(1, 4):
        o.step()
        g.step(1)
        check(f"walls {kinds} rk{rk} step {n}", g.conserved(1), o.conserved(1), FP64_TOL)
    o.close()
    g.close()


def test_walls_stage_rhs_fp64():
    """the stage right-hand side alone (preSubStep, solveEqnSet,
    postSubStep): the wall-corrected stage state and dt*RHS"""
    dims = (20, 18, 16)
    kw = dict(Lx=2.0, Ly=2.0, Lz=2.0, alpha_f=0.49, mu_ref=2e-3, Pr=0.7, cfl=0.5,
              filter_step=1, rk=4, bc_type=(1, 1, 0), bc_kind=(MOVE, ADIA, ADIA, MOVE, 0, 0),
              wall_vel=VEL)
    o, g = pair(dims, random_ic((16, 18, 20), 8, amp=0.2), cs3d.FP64, **kw)
    o.pre_step()
    g.pre_step()
    o.rhs(1)
    g.rhs(1)
    check("walls rhs state", g.conserved(1), o.conserved(1), FP64_TOL)
    k_o, k_g = o.conserved(3), g.conserved(3)
    # dt*RHS relative to each field's scale (continuity cancels to O(M^2))
    errs = [relerr(k_g[v], k_o[v]) for v in range(5)]
    print("walls rhs k2: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) < 1e-10
    # the wall planes keep only the normal continuity term
    for sl in (np.s_[:, :, 0], np.s_[:, :, -1], np.s_[:, 0, :], np.s_[:, -1, :]):
        for v in range(1, 5):
            assert np.abs(k_g[v][sl]).max() == 0.0
    o.close()
    g.close()


@pytest.mark.parametrize("prec", [cs3d.FP64, cs3d.FP32])
def test_walls_beside_sponges(prec):
    """x walls, y sponges: the wall faces cross the sponge planes, and the
    first stage after each sponge update takes the neighbours' temperatures
    on sponge planes from their pre-update values"""
    dims = (24, 24, 12)
    kw = dict(Lx=3.0, Ly=6.0, Lz=2.0, alpha_f=0.49, mu_ref=2e-3, Pr=0.7, cfl=0.5,
              filter_step=2, rk=4, bc_type=(1, 1, 0), bc_kind=(ADIA, MOVE, 1, 1, 0, 0),
              wall_vel=VEL, sponge_frac=(0.25, 0.2, 0.25))
    o, g = pair(dims, random_ic((12, 24, 24), 9, amp=0.2), prec, **kw)
    tol = FP64_TOL if prec == cs3d.FP64 else 2e-6
    for n in range(1, 5):
        o.step()
        g.step(1)
        check(f"walls+sponge prec={prec} step {n}", g.conserved(1), o.conserved(1),
              tol * (10 if prec == cs3d.FP64 else 1))
    o.close()
    g.close()


def test_couette_fp32_tracks_fp64():
    """plane Couette flow from rest between a no-slip wall and a moving
    wall, 300 steps: FP32 state within 1e-5 of FP64"""
    dims = (16, 33, 16)
    sh = (16, 33, 16)
    kw = dict(Lx=1.0, Ly=1.0, Lz=1.0, alpha_f=0.49, mu_ref=0.005, Pr=0.7, cfl=0.5,
              filter_step=1, rk=4, bc_type=(0, 1, 0), bc_kind=(0, 0, ADIA, MOVE, 0, 0),
              wall_vel={3: (0.1, 0.0, 0.0)})
    ic = (np.ones(sh), np.zeros(sh), np.zeros(sh), np.zeros(sh), np.full(sh, 1 / 1.4))
    out = {}
    for prec in (cs3d.FP64, cs3d.FP32):
        g = cs3d.Solver(cs3d.config(*dims, precision=prec, **kw))
        g.set_initial_conditions(*ic)
        g.step(300)
        out[prec] = g.conserved(1)
        g.close()
    errs = [relerr(out[cs3d.FP32][v], out[cs3d.FP64][v]) for v in (0, 1, 4)]
    print("couette fp32 vs fp64: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) < 1e-5


def test_adjoint_with_walls_on_periodic_lines_matches_oracle_adjoint():
    """walls on the faces of periodic z lines (the reference allows them):
    the gradient through three steps against the numpy adjoint of the
    oracle (oracle/adjoint.py) from the same initial state"""
    import adjoint as oa
    dims = (20, 18, 16)
    kw = dict(Lx=2.0, Ly=2.0, Lz=2.0, alpha_f=0.49, mu_ref=2e-3, Pr=0.7, cfl=0.5,
              filter_step=1, rk=4, bc_type=(0, 0, 0), bc_kind=(0, 0, 0, 0, ADIA, MOVE),
              wall_vel=VEL, dt_fixed=0.005)
    ic = random_ic((16, 18, 20), 3, amp=0.1)
    g = cs3d.Solver(cs3d.config(*dims, precision=cs3d.FP64, **kw))
    g.set_initial_conditions(*ic)
    q0 = g.conserved(1)
    J, *grad = g.ke_gradient(3, 2)
    g.close()
    keep = {k: v for k, v in kw.items() if k not in ("cfl", "dt_fixed")}
    A = oa.Adjoint(po.make_config(*dims, **keep))
    Jo, *gref = A.ke_gradient(ic, 3, 0.005, 1, q0=q0)
    assert abs(J - Jo) <= 1e-13 * abs(Jo)
    errs = [relerr(a, b) for a, b in zip(grad, gref)]
    print("walls on periodic z, gradient vs oracle adjoint: " + " ".join(f"{e:.2e}" for e in errs),
          flush=True)
    assert max(errs) <= 1e-10, errs
