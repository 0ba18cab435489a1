"""The GPU optical-path integrals (AeroOptics::computeAO,
AeroOptics.cpp:45-93) against the oracle, whose restatement is pinned to
the reference's compiled AeroOptics.cpp (tests/test_oracle_aero.py): on
given FP64 density fields through cs3d_aero_optics_field, and on a running
solver's state (set 1) and *2 array (set 4) through cs3d_aero_optics, FP64
and FP32 (the integrals are FP64 over the stored density in both)."""
import numpy as np
import pytest

import cs3d
import pyoracle as po
from helpers import relerr, shear_ic
from test_oracle_aero import GRIDS, smooth_rho

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,lengths", GRIDS + [((256, 64, 32), (2 * np.pi, 5.0, 1.0))])
@pytest.mark.parametrize("nA,max_angle", [(5, 70.0), (3, 45.0)])
def test_field_integrals_match_oracle(dims, lengths, nA, max_angle):
    Nx, Ny, Nz = dims
    rho = smooth_rho((Nz, Ny, Nx), sum(dims))
    got = cs3d.aero_optics_field(rho, lengths, nA, max_angle)
    ref = po.aero_optics(rho, lengths, nA, max_angle)
    err = relerr(got, ref)
    print(f"{dims} {nA} angles: GPU vs oracle {err:.1e}", flush=True)
    assert err < 1e-12


@pytest.mark.parametrize("prec", [cs3d.FP64, cs3d.FP32])
def test_solver_integrals_match_oracle(prec):
    Nx, Ny, Nz = 48, 40, 24
    kw = dict(Lx=8.0, Ly=10.0, Lz=4.0, alpha_f=0.49, mu_ref=2e-3, Pr=0.7, cfl=0.5,
              filter_step=2, rk=4, bc_type=(0, 1, 0), bc_kind=(0, 0, 1, 1, 0, 0),
              sponge_frac=(0.25, 0.15, 0.25))
    g = cs3d.Solver(cs3d.config(Nx, Ny, Nz, precision=prec, **kw))
    g.set_initial_conditions(*shear_ic(Nx, Ny, Nz, kw["Ly"], 4))
    g.step(3)
    lengths = (kw["Lx"], kw["Ly"], kw["Lz"])
    for set_ in (1, 4):
        got = g.aero_optics(5, 70.0, set_)
        ref = po.aero_optics(g.conserved(set_)[0], lengths, 5, 70.0)
        err = relerr(got, ref)
        print(f"prec {prec} set {set_}: GPU vs oracle on the stored density {err:.1e}", flush=True)
        assert err < 1e-12
    g.close()


def test_refuses_bad_arguments():
    rho = np.ones((8, 8, 8))
    with pytest.raises(cs3d.CS3DError):
        cs3d.aero_optics_field(rho, (1.0, 1.0, 1.0), 1, 70.0)
    with pytest.raises(cs3d.CS3DError):
        cs3d.aero_optics_field(rho, (1.0, 1.0, 1.0), 5, 90.0)
