"""The oracle's optical-path integrals (AeroOptics::computeAO,
AeroOptics.cpp:45-93, with Utils.cpp's getBaseNodeIndex and
linearInterpolation) against the reference's own AeroOptics.cpp and
Utils.cpp compiled into oracle/_ref, and against closed forms: a uniform
density integrates to rho * L_ray; densities linear in y, or in z, or in x on
the vertical ray, are reproduced exactly by the trilinear interpolation and
the trapezoid rule.  The linear-in-z case covers the last z plane, where the
reference reads one plane past its arrays with weight zero."""
import numpy as np
import pytest

import pyoracle as po

GRIDS = [((16, 20, 12), (2 * np.pi - 2 * np.pi / 16, 3.0, 1.5)),
         ((33, 17, 9), (4.0, 2.5, 2.0))]


def ray_lengths(Ly, nA, max_angle):
    a = np.arange(nA) / (nA - 1) * max_angle * np.pi / 180.0
    return 0.6 * Ly / np.cos(a)


def smooth_rho(shape, seed):
    rng = np.random.default_rng(seed)
    Nz, Ny, Nx = shape
    z, y, x = (np.arange(n) / n for n in shape)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    ph = rng.uniform(0, 2 * np.pi, 6)
    return 1.0 + 0.1 * np.sin(2 * np.pi * X + ph[0]) * np.cos(2 * np.pi * Y + ph[1]) + \
        0.05 * np.cos(4 * np.pi * Z + ph[2]) + 0.02 * rng.standard_normal(shape)


@pytest.mark.skipif(not po.ref_built(), reason="oracle/_ref/libref.so not built")
@pytest.mark.parametrize("dims,lengths", GRIDS)
@pytest.mark.parametrize("nA,max_angle", [(5, 70.0), (3, 45.0)])
def test_oracle_matches_reference(dims, lengths, nA, max_angle, tmp_path, monkeypatch):
    Nx, Ny, Nz = dims
    rho = smooth_rho((Nz, Ny, Nx), sum(dims))
    got = po.aero_optics(rho, lengths, nA, max_angle)
    ref = np.zeros_like(got)
    # computeAO also appends its results to files in the working directory
    monkeypatch.chdir(tmp_path)
    po.ref_lib().ref_aero_optics(Nx, Ny, Nz, *lengths, rho.ctypes.data_as(po._dp), nA, max_angle,
                                 ref.ctypes.data_as(po._dp))
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"{dims} {nA} angles to {max_angle}: oracle vs reference {err:.1e}", flush=True)
    assert err < 1e-13


@pytest.mark.parametrize("dims,lengths", GRIDS)
def test_closed_forms(dims, lengths):
    Nx, Ny, Nz = dims
    Lx, Ly, Lz = lengths
    nA, amax = 5, 70.0
    L = ray_lengths(Ly, nA, amax)[:, None, None]
    x = np.arange(Nx) / (Nx - 1.0) * Lx
    y = np.arange(Ny) / (Ny - 1.0) * Ly
    z = np.arange(Nz) / (Nz - 1.0) * Lz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    # uniform density
    opl = po.aero_optics(np.full((Nz, Ny, Nx), 1.3), lengths, nA, amax)
    assert np.abs(opl - 1.3 * L).max() < 1e-13 * L.max()
    # linear in y: the ray's mean density is the value at its midpoint
    opl = po.aero_optics(0.7 + 0.4 * Y, lengths, nA, amax)
    want = (0.7 + 0.4 * 0.5 * Ly) * np.broadcast_to(L, opl.shape)
    assert np.abs(opl - want).max() < 1e-13 * want.max()
    # linear in z: every ray stays on its own z plane, the last one included
    opl = po.aero_optics(0.9 + 0.3 * Z, lengths, nA, amax)
    want = (0.9 + 0.3 * z)[None, :, None] * L
    assert np.abs(opl - want).max() < 1e-13 * want.max()
    # linear in x on the vertical ray (the first angle); the ray positions
    # wrap with fmod(., Lx), so the last column x = Lx samples x = 0
    # (AeroOptics.cpp:68)
    opl = po.aero_optics(1.1 + 0.2 * X, lengths, nA, amax)
    want = (1.1 + 0.2 * np.fmod(x, Lx))[None, :] * L[0]
    assert np.abs(opl[0] - want).max() < 1e-13 * want.max()
