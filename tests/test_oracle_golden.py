"""The oracle against the reference's recorded Taylor-Green output
(taylorgreen_vortex/taylorgreen_run64/taylorgreen.out, first two rows, kept
as tests/golden/tgv64_taylorgreen_out_rows1-2.txt).  That run used
Lx = 2 pi - 2 pi/N, CFL 0.5, filterStep 2, alphaF 0.495, mu_ref 1.4625e-5,
rho0 = 1, p0 = 100 + (cos 2x + cos 2y)(cos 2z + 2)/16 and gamma 1.4, Pr 0.7
(taylorgreen_run64/3DCompact.cpp_N64).  Two RK4 steps pin the whole stage:
derivatives, RHS assembly, RK4, CFL dt, the filter (step 2) and the
diagnostics."""
import os

import numpy as np

import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))


def test_tgv64_recorded_rows():
    rows = np.loadtxt(os.path.join(HERE, "golden", "tgv64_taylorgreen_out_rows1-2.txt"))
    N = 64
    L = 2 * np.pi - 2 * np.pi / N
    cfg = po.make_config(N, N, N, Lx=L, Ly=L, Lz=L, alpha_f=0.495, mu_ref=0.000014625,
                         cfl=0.5, filter_step=2, rk=4, Pr=0.7)
    s = po.Solver(cfg)
    x = np.arange(N) * (L / (N - 1))
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    U = np.sin(X) * np.cos(Y) * np.cos(Z)
    V = -np.cos(X) * np.sin(Y) * np.cos(Z)
    p = 100.0 + (1 / 16) * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
    s.set_ic(np.ones_like(U), U, V, np.zeros_like(U), p)
    for r in range(2):
        s.step()
        d = s.diag_tgv()
        # time column is exact; sums to ~1e-12 (the recorded run used
        # -ffast-math), the tiny dilatation rms to 1e-9
        assert d[0] == rows[r, 0] or abs(d[0] - rows[r, 0]) < 1e-15
        for c in range(1, 6):
            assert abs(d[c] - rows[r, c]) <= 1e-11 * abs(rows[r, c]), (r, c, d[c], rows[r, c])
        assert abs(d[6] - rows[r, 6]) <= 1e-8 * abs(rows[r, 6])
