"""Closed-form pins of the oracle's sponge path, which no reference output
covers (CSolver_AWS.cpp cannot be compiled here and no recorded run has a
sponge).  Three parts of the reference are checked:

- the sponge source sigma (phi_avg - phi) of every equation
  (CSolver_AWS.hpp:228-230, used at CSolver_AWS.cpp:1262-1265 .. 1452-1455),
  which enters the energy equation twice (:1491 inside engyEuler, :1495),
  and the zeroed right-hand side on sponge face planes (:1606-1670);
- updateSponge's running average with eps = 1/(avgT/dt + 1) and the energy
  target blended with weight epsP toward the sponge pressure 1/gamma at the
  average's own momenta and density (:1937-1953);
- the reset of the face planes of the state to the averages (:1955-2021).

On a uniform state every compact derivative vanishes, so the stage's
dt*RHS is the sponge term alone; with the state held fixed the averages
follow closed forms.  The checks are written in terms the reference's code
does not use (geometric relaxation, the average's pressure), and
test_planted_mistakes_fail shows that each of seven plausible mistakes,
compiled into a private copy of the oracle (OR_PLANT in oracle/oracle.c),
fails at least one of them."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po

GAMMA = 1.4
SP = po.SPONGE
DIMS = (12, 17, 8)  # x and y sponges on both faces, periodic z
SHAPE = DIMS[::-1]
KW = dict(Lx=1.0, Ly=1.3, Lz=0.9, bc_type=(1, 1, 0), bc_kind=(SP, SP, SP, SP, 0, 0),
          mu_ref=0.01, cfl=0.5)


def conserved(rho, u, v, w, p):
    """uniform conserved state as five (Nz, Ny, Nx) fields"""
    ke = 0.5 * rho * (u * u + v * v + w * w)
    vals = (rho, rho * u, rho * v, rho * w, p / (GAMMA - 1.0) + ke)
    return [np.full(SHAPE, x) for x in vals]


def prims(rho, u, v, w, p):
    return [np.full(SHAPE, x) for x in (rho, u, v, w, p)]


def pressure(q):
    return (GAMMA - 1.0) * (q[4] - 0.5 * (q[1] ** 2 + q[2] ** 2 + q[3] ** 2) / q[0])


def planes():
    """sponge face planes (x and y faces)"""
    m = np.zeros(SHAPE, bool)
    m[:, :, 0] = m[:, :, -1] = True
    m[:, 0, :] = m[:, -1, :] = True
    return m


Q = (1.1, 0.2, -0.1, 0.05, 1.05 / GAMMA)   # the state
A = (1.0, 0.1, 0.05, -0.02, 1.0 / GAMMA)   # the running average


def solver(library):
    return po.Solver(po.make_config(*DIMS, **KW), library=library)


def check_rhs(library):
    """uniform state q, uniform average A != q: dt*RHS = dt sigma (A - q)
    for rho .. rhoW, twice that for rhoE, zero on the face planes"""
    o = solver(library)
    o.set_ic(*prims(*Q))
    qa = conserved(*A)
    o.set_sponge_avg(qa)
    dt = 0.01
    o.set_dt(dt)
    o.pre_step()
    o.rhs(1)
    k = o.conserved(3)
    q = o.conserved(1)
    sig = o.sponge_sigma()
    o.close()
    fails = []
    on = planes()
    for v in range(5):
        want = (2.0 if v == 4 else 1.0) * dt * sig * (qa[v] - q[v])
        want[on] = 0.0
        scale = np.abs(want).max()
        assert scale > 0
        if np.abs(k[v] - want).max() > 1e-12 * scale:
            fails.append(f"rhs[{v}]")
    return fails


def check_average_relaxation(library, nsteps=12, dt=0.3):
    """state held fixed away from the face planes: every average relaxes
    geometrically, A_n - q = (1 - eps)^n (A_0 - q), eps = dt / (avgT + dt);
    the energy's blend keeps the average's pressure within epsP of 1/gamma"""
    o = solver(library)
    o.set_ic(*prims(*Q))
    a0 = conserved(*A)
    o.set_sponge_avg(a0)
    o.set_dt(dt)
    o.pre_step()
    q = conserved(*Q)
    eps = dt / (15.0 + dt)
    fails = []
    off = ~planes()
    for n in range(1, nsteps + 1):
        o.post_step()
        a = o.sponge_avg()
        for v in range(4):
            want = q[v] + (1.0 - eps) ** n * (a0[v] - q[v])
            if np.abs(a[v] - want)[off].max() > 1e-13 * np.abs(want).max():
                fails.append(f"avg[{v}] step {n}")
                break
        if n == 1:
            # p(A_1) - 1/gamma = epsP (p_mid - 1/gamma), p_mid the pressure of
            # the eps-update, within |p_q - 1/gamma| + (gamma - 1) * (kinetic
            # energies of A_0 and q) of 1/gamma
            kin = lambda s: 0.5 * s[0] * (s[1] ** 2 + s[2] ** 2 + s[3] ** 2)
            bound = 2 * 0.005 * (abs(Q[4] - 1 / GAMMA) + (GAMMA - 1.0) * (kin(A) + kin(Q)))
            if np.abs(pressure(a) - 1.0 / GAMMA)[off].max() > bound:
                fails.append("avg pressure blend")
    o.close()
    return fails


def check_pressure_blend(library, nsteps=3, dt=0.3):
    """average with the state's density and momenta but another pressure:
    those stay equal to the state's, and the average's pressure follows
    p_{n+1} = epsP [(1 - eps) p_n + eps p_q] + (1 - epsP) / gamma, i.e.
    p_n = p* + a^n (p_0 - p*) with a = epsP (1 - eps)"""
    o = solver(library)
    o.set_ic(*prims(*Q))
    a0 = conserved(Q[0], Q[1], Q[2], Q[3], 0.8 / GAMMA)
    o.set_sponge_avg(a0)
    o.set_dt(dt)
    o.pre_step()
    q = conserved(*Q)
    eps, epsP = dt / (15.0 + dt), 0.005
    a_ = epsP * (1.0 - eps)
    pstar = (epsP * eps * Q[4] + (1.0 - epsP) / GAMMA) / (1.0 - a_)
    fails = []
    off = ~planes()
    for n in range(1, nsteps + 1):
        o.post_step()
        a = o.sponge_avg()
        for v in range(4):
            if np.abs(a[v] - q[v])[off].max() > 1e-15 * np.abs(q[v]).max():
                fails.append(f"blend avg[{v}]")
        want = pstar + a_ ** n * (0.8 / GAMMA - pstar)
        if np.abs(pressure(a) - want)[off].max() > 1e-13:
            fails.append(f"blend pressure step {n}")
    o.close()
    return fails


def check_plane_reset(library, dt=0.3):
    """after updateSponge the state's face planes hold the new averages;
    every other point keeps the state"""
    o = solver(library)
    o.set_ic(*prims(*Q))
    o.set_sponge_avg(conserved(*A))
    o.set_dt(dt)
    o.pre_step()
    o.post_step()
    a = o.sponge_avg()
    s = o.conserved(1)
    q = conserved(*Q)
    o.close()
    on = planes()
    fails = []
    for v in range(5):
        if not (np.array_equal(s[v][on], a[v][on]) and np.array_equal(s[v][~on], q[v][~on])):
            fails.append(f"reset[{v}]")
    return fails


CHECKS = (check_rhs, check_average_relaxation, check_pressure_blend, check_plane_reset)


def all_failures(library):
    out = []
    for c in CHECKS:
        out += c(library)
    return out


def test_oracle_sponge_closed_forms():
    assert all_failures(po.lib()) == []


PLANTS = {
    1: "energy sponge source counted once",
    2: "sponge source sign flipped",
    3: "eps = dt / avgT",
    4: "epsP and 1 - epsP exchanged",
    5: "sponge planes of dt*RHS not zeroed",
    6: "face planes not reset to the averages",
    7: "energy target from the state's momenta, not the average's",
}


def test_planted_mistakes_fail(tmp_path):
    src = os.path.join(os.path.dirname(po.__file__), "oracle.c")
    for k, what in PLANTS.items():
        so = str(tmp_path / f"liboracle_plant{k}.so")
        subprocess.run(["gcc", "-O2", "-std=gnu99", "-fopenmp", "-fPIC", "-shared",
                        f"-DOR_PLANT={k}", "-o", so, src, "-lm"], check=True)
        fails = all_failures(po.load_path(so))
        print(f"plant {k} ({what}): fails {fails}", flush=True)
        assert fails, f"planted mistake {k} ({what}) passes every sponge check"
