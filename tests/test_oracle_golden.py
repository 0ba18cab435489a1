"""The oracle against the reference's recorded Taylor-Green output: the
first two rows of taylorgreen_vortex/taylorgreen_run64/taylorgreen.out and
taylorgreen_run128/taylorgreen.out (tests/golden/, written by
tests/golden/extract_reference_rows.py).  Two RK4 steps pin the whole stage:
derivatives, RHS assembly, RK4, CFL dt, the filter and the diagnostics.

The recorded runs (taylorgreen_run64/3DCompact.cpp_N64,
taylorgreen_run128/3DCompact.cpp_N128): Lx = 2 pi - 2 pi/N, alphaF 0.495,
mu_ref 1.4625e-5, rho0 = 1, p0 = 100 + (cos 2x + cos 2y)(cos 2z + 2)/16
(no division by gamma), gamma 1.4, Pr 0.7; 64^3 with CFL 0.5 and
filterStep 2, 128^3 with CFL 1.25 and filterStep 1."""
import os

import numpy as np
import pytest

import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))

RECORDED = {
    64: dict(cfl=0.5, filter_step=2, rows="tgv64_taylorgreen_out_rows1-2.txt"),
    128: dict(cfl=1.25, filter_step=1, rows="tgv128_taylorgreen_out_rows1-2.txt"),
}


def recorded_case(N):
    """(solver keyword arguments, initial fields, recorded rows) of the
    recorded Taylor-Green run at N^3"""
    r = RECORDED[N]
    L = 2 * np.pi - 2 * np.pi / N
    kw = dict(Lx=L, Ly=L, Lz=L, alpha_f=0.495, mu_ref=0.000014625, cfl=r["cfl"],
              filter_step=r["filter_step"], rk=4, Pr=0.7)
    x = np.arange(N) * (L / (N - 1))
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    U = np.sin(X) * np.cos(Y) * np.cos(Z)
    V = -np.cos(X) * np.sin(Y) * np.cos(Z)
    p = 100.0 + (1 / 16) * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
    ic = (np.ones_like(U), U, V, np.zeros_like(U), p)
    rows = np.loadtxt(os.path.join(HERE, "golden", r["rows"]))
    return kw, ic, rows


def check_row(d, row):
    """time exact; sums to 1e-11 (the recorded run used -ffast-math), the
    tiny dilatation rms to 1e-8"""
    assert d[0] == row[0] or abs(d[0] - row[0]) < 1e-15, (d[0], row[0])
    for c in range(1, 6):
        assert abs(d[c] - row[c]) <= 1e-11 * abs(row[c]), (c, d[c], row[c])
    assert abs(d[6] - row[6]) <= 1e-8 * abs(row[6]), (d[6], row[6])


@pytest.mark.parametrize("N", [64, 128])
def test_tgv_recorded_rows(N):
    kw, ic, rows = recorded_case(N)
    s = po.Solver(po.make_config(N, N, N, **kw))
    s.set_ic(*ic)
    for r in range(2):
        s.step()
        check_row(s.diag_tgv(), rows[r])
    s.close()


def recorded_steps(N):
    """{step: recorded row} of the recorded run at N^3 (tests/golden/*_steps.txt)"""
    a = np.loadtxt(os.path.join(HERE, "golden", f"tgv{N}_taylorgreen_out_steps.txt"))
    return {int(r[0]): r[1:] for r in a}


def test_tgv64_recorded_first_twenty_steps():
    """the oracle along the recorded 64^3 run, every step to 10 and at 20"""
    kw, ic, _ = recorded_case(64)
    rows = recorded_steps(64)
    s = po.Solver(po.make_config(64, 64, 64, **kw))
    s.set_ic(*ic)
    for n in range(1, 21):
        s.step()
        if n in rows:
            check_row(s.diag_tgv(), rows[n])
    s.close()
