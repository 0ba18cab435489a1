"""cs3d_diag_shear_layer_noalloc, the shear-layer profiles for callers that
cannot allocate (its work arrays in a file-scope device array), against
cs3d_diag_shear_layer on the same states through the C ABI: two solvers in
a row, the second smaller than the first, FP64 and FP32, both Q12 settings.
Both entry points run the same row kernel and reduce the rows over the
planes in the same order, so their outputs agree to the last bit."""
import ctypes

import numpy as np
import pytest

import cs3d
from helpers import shear_ic

pytestmark = pytest.mark.gpu

KW = dict(Lx=8.0, Ly=10.0, Lz=4.0, alpha_f=0.49, mu_ref=2e-3, Pr=0.7, cfl=0.5, filter_step=2,
          rk=4, bc_type=(0, 1, 0), bc_kind=(0, 0, 1, 1, 0, 0), sponge_frac=(0.25, 0.15, 0.25))


def profiles(g, noalloc, fix):
    out = np.full(11 * g.cfg.Ny, np.nan)
    L = cs3d.lib()
    fn = L.cs3d_diag_shear_layer_noalloc if noalloc else L.cs3d_diag_shear_layer
    rc = fn(g.h, fix, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    cs3d.check(rc)
    return out


@pytest.mark.parametrize("prec", [cs3d.FP64, cs3d.FP32])
def test_noalloc_profiles_match_allocating_entry(prec):
    for dims, seed in (((48, 40, 24), 1), ((32, 24, 16), 2)):
        Nx, Ny, Nz = dims
        g = cs3d.Solver(cs3d.config(Nx, Ny, Nz, precision=prec, **KW))
        g.set_initial_conditions(*shear_ic(Nx, Ny, Nz, KW["Ly"], seed))
        g.step(2)
        for fix in (0, 1):
            a = profiles(g, False, fix)
            b = profiles(g, True, fix)
            assert np.isfinite(b).all()
            err = float(np.abs(a - b).max() / np.abs(a).max())
            print(f"prec {prec} dims {dims} fix_q12 {fix}: max rel diff {err:.1e}, "
                  f"identical {np.array_equal(a, b)}", flush=True)
            assert err <= 1e-14
        g.close()


def test_noalloc_profiles_refuse_large_grids():
    Nx, Ny, Nz = 8, 400, 400
    g = cs3d.Solver(cs3d.config(Nx, Ny, Nz, precision=cs3d.FP32, **KW))
    sh = (Nz, Ny, Nx)
    g.set_initial_conditions(np.ones(sh), np.zeros(sh), np.zeros(sh), np.zeros(sh),
                             np.full(sh, 1 / 1.4))
    out = np.zeros(11 * Ny)
    rc = cs3d.lib().cs3d_diag_shear_layer_noalloc(g.h, 0, out.ctypes.data_as(
        ctypes.POINTER(ctypes.c_double)))
    assert rc == cs3d.ERR_ARG and b"131072" in cs3d.lib().cs3d_last_error()
    g.close()
