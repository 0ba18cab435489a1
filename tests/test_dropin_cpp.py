"""The C++ drop-in (include/cs3d.hpp): a Taylor-Green driver written against
the reference's class interface (Domain, BC, TimeStepping, CSolver_AWS, RK4;
3DCompact.cpp:40-110) builds with g++ against the in-tree library, and on the
GPU reproduces the C ABI's own run step for step, including the
taylorgreen.out lines in the reference's format (CSolver_AWS.cpp:2240-2248)."""
import os
import subprocess

import numpy as np
import pytest

import cs3d

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "tgv_dropin")


def _build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], check=True)
    return BIN


def test_dropin_driver_builds():
    assert os.access(_build(), os.X_OK)


def test_dropin_header_is_plain_cxx():
    """The facade needs only the C header: it compiles with -fsyntax-only and
    no HIP include path, in C++11 as the reference's Makefile.in uses."""
    src = '#define CS3D_DROPIN_GLOBAL\n#include "cs3d.hpp"\nint main(){return 0;}\n'
    r = subprocess.run(["g++", "-std=c++11", "-fsyntax-only", "-Wall", "-Werror",
                        "-I" + os.path.join(ROOT, "include"), "-x", "c++", "-"],
                       input=src.encode(), capture_output=True)
    assert r.returncode == 0, r.stderr.decode()


CUSTOM_RK = r"""
#define CS3D_DROPIN_GLOBAL
#include "cs3d.hpp"
// a time integrator of the caller's own: it would update the host mirrors
class MyRK : public AbstractRK {
  public:
    explicit MyRK(AbstractCSolver *c) { cs = c; }
    void updateConservedData() override {}
    void executeSolverLoop() override { cs->setInitialConditions(); }
};
int main()
{
    Domain dom(8, 8, 8, 1.0, 1.0, 1.0);
    BC bc(BC::PERIODIC_SOLVE, BC::PERIODIC, BC::PERIODIC, BC::PERIODIC_SOLVE, BC::PERIODIC,
          BC::PERIODIC, BC::PERIODIC_SOLVE, BC::PERIODIC, BC::PERIODIC);
    TimeStepping ts(TimeStepping::CONST_CFL, 0.5, 1, 1.0, 1, 1, 1);
    CSolver_AWS cs(&dom, &bc, &ts, 0.49, 1e-3, 8, false);
    MyRK rk(&cs);
    rk.executeSolverLoop();
    return 0;
}
"""


def test_dropin_rejects_foreign_integrator(tmp_path):
    """CSolver_AWS keeps its state on the GPU and fuses the RK update into its
    sweeps, so an AbstractRK subclass of the caller's own must not run against
    it silently: setInitialConditions aborts, naming RK4 / TVDRK3, before any
    GPU call."""
    _build()
    src = tmp_path / "custom_rk.cpp"
    src.write_text(CUSTOM_RK)
    exe = tmp_path / "custom_rk"
    lib = os.path.join(ROOT, "3dcompactsolver_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), str(src),
                    "-L" + lib, "-l3dcompact_hip", "-Wl,-rpath," + lib, "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, timeout=60)
    assert r.returncode != 0
    assert b"cs3d::RK4 or cs3d::TVDRK3" in r.stderr


@pytest.mark.gpu
def test_dropin_driver_matches_c_abi(tmp_path):
    exe = BIN if os.access(BIN, os.X_OK) else _build()
    N, steps = 32, 3
    r = subprocess.run([exe, str(N), str(steps), "0", str(tmp_path)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    fields = r.stdout.split()
    assert fields[0] == "steps" and int(fields[1]) == steps
    ke_cpp = float(fields[5])
    rows = np.loadtxt(tmp_path / "taylorgreen.out", ndmin=2)
    assert rows.shape == (steps, 7)

    L = 2 * np.pi - 2 * np.pi / N
    x = np.arange(N) * (L / (N - 1))
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    p0 = 1 / (1.4 * 0.01)
    p = p0 + (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0) / 16.0
    ic = (p / p0, np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z),
          np.zeros_like(X), p)
    g = cs3d.Solver(cs3d.config(N, N, N, Lx=L, Ly=L, Lz=L, alpha_f=0.49,
                                mu_ref=(1 / 1600) / 100 ** 0.76, Pr=0.71, cfl=0.5,
                                filter_step=1, rk=4, precision=cs3d.FP64))
    g.set_initial_conditions(*ic)
    g.set_diagnostics(True)
    for s in range(steps):
        g.step(1)
        ref = g.diag_taylor_green()
        # the driver builds the initial field with libm, the Python side with
        # numpy: last-bit differences, grown over the steps
        assert np.allclose(rows[s][:6], ref[:6], rtol=1e-12, atol=0), (s, rows[s], ref)
        # dilprime = rms(div u) is cancellation-dominated on this field
        assert abs(rows[s][6] - ref[6]) <= 1e-6 * abs(ref[6]), (s, rows[s][6], ref[6])
    assert abs(ke_cpp - g.kinetic_energy()) <= 1e-12 * abs(ke_cpp)


@pytest.mark.gpu
def test_dropin_dump_and_restart(tmp_path):
    """a driver run of 6 steps against one that stops at step 3 (dumping
    every 3 steps) and a second driver that restarts from that dump
    (CSolver_AWS::restartStep) and runs to step 6"""
    exe = BIN if os.access(BIN, os.X_OK) else _build()
    full, part = tmp_path / "full", tmp_path / "part"
    full.mkdir()
    part.mkdir()
    run = lambda *a: subprocess.run([exe, *map(str, a)], capture_output=True, text=True,
                                    timeout=120)
    r_full = run(24, 6, 0, full, 3)
    r_a = run(24, 3, 0, part, 3)
    assert (part / "rho.out.3").exists() and (part / "restart.3").exists()
    r_b = run(24, 6, 0, part, 3, 3)
    for r in (r_full, r_a, r_b):
        assert r.returncode == 0, r.stderr
    f_full, f_b = r_full.stdout.split(), r_b.stdout.split()
    assert int(f_full[1]) == int(f_b[1]) == 6
    assert abs(float(f_full[3]) - float(f_b[3])) <= 1e-13 * float(f_full[3])
    assert abs(float(f_full[5]) - float(f_b[5])) <= 1e-13 * float(f_full[5])
    for name in ("rho", "rhoU", "rhoE"):
        a = np.fromfile(full / f"{name}.out.6")
        b = np.fromfile(part / f"{name}.out.6")
        assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(a)), name


EXTRAS = os.path.join(ROOT, "tests", "cpp", "_build", "dropin_extras")


@pytest.mark.gpu
def test_dropin_lines_aero_optics_and_gradient(tmp_path):
    """The drop-in's per-line Derivatives/Filter methods (Derivatives.hpp:
    170-183, Filter.hpp:111-114) against the oracle's line operators, a
    standalone AeroOptics (AeroOptics.cpp:27-93: the .OPL files) against the
    oracle's integrals, and CSolver_AWS::keGradient against the C ABI's
    gradient on the same run."""
    exe = EXTRAS if os.access(EXTRAS, os.X_OK) else (_build() and EXTRAS)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    import pyoracle as po
    dims, lengths = (16, 12, 10), (3.0, 2.0, 1.5)
    lines = np.fromfile(tmp_path / "lines.bin")
    pos = 0
    for d in range(3):
        n = dims[d]
        dh = (1.0 / (n - 1.0)) * lengths[d]
        for bc in (0, 1):
            f, d1, d2, ff = (lines[pos + i * n:pos + (i + 1) * n] for i in range(4))
            pos += 4 * n
            for kind, got in ((1, d1), (2, d2), (0, ff)):
                ref = po.line_op(kind, bc, f, dh=dh, alpha_f=0.49)
                err = np.abs(got - ref).max() / np.abs(ref).max()
                assert err < 1e-12, (d, bc, kind, err)
    rho = np.fromfile(tmp_path / "rho.bin").reshape(dims[::-1])
    ref = po.aero_optics(rho, lengths, 5, 70.0)
    for ap, name in enumerate(("0.00", "17.50", "35.00", "52.50", "70.00")):
        got = np.fromfile(tmp_path / f"{name}.OPL").reshape(dims[2], dims[0])
        assert np.abs(got - ref[ap]).max() < 1e-12 * np.abs(ref[ap]).max(), name
    # the same gradient through the C ABI
    N = 16
    L = 2 * np.pi - 2 * np.pi / N
    x = np.arange(N) * (L / (N - 1))
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    p0 = 1 / (1.4 * 0.01)
    p = p0 + (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0) / 16.0
    g = cs3d.Solver(cs3d.config(N, N, N, Lx=L, Ly=L, Lz=L, alpha_f=0.49,
                                mu_ref=(1 / 1600) / 100 ** 0.76, Pr=0.71, cfl=0.5,
                                filter_step=1, rk=4, precision=cs3d.FP64))
    g.set_initial_conditions(p / p0, np.sin(X) * np.cos(Y) * np.cos(Z),
                             -np.cos(X) * np.sin(Y) * np.cos(Z), np.zeros_like(X), p)
    J, gU, _, _ = g.ke_gradient(2, 1)
    grad = np.fromfile(tmp_path / "grad.bin")
    # the driver's initial field comes from libm, the Python side's from numpy
    assert abs(grad[0] - J) <= 1e-12 * abs(J)
    assert np.abs(grad[1:] - gU.ravel()).max() <= 1e-10 * np.abs(gU).max()
    g.close()
