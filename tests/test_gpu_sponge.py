"""The sponge path on the GPU against the same closed forms that pin the
oracle's (tests/test_oracle_sponge.py): on a uniform state the stage's
dt*RHS is dt sigma (A - q) (twice for rhoE, CSolver_AWS.cpp:1491,1495) and
zero on the face planes (:1606-1670); updateSponge relaxes the averages
geometrically with eps = 1/(avgT/dt + 1), blends the energy toward the
sponge pressure (:1937-1953) and resets the face planes of the state
(:1955-2021).  Through the C ABI, FP64 and FP32."""
import numpy as np
import pytest

import cs3d
import pyoracle as po
from test_oracle_sponge import A, DIMS, GAMMA, KW, Q, conserved, planes, pressure, prims

pytestmark = pytest.mark.gpu

# FP64: rounding of the stored perturbations; FP32: single rounding of the
# fields (measured errors are a few units of 1e-8)
TOL = {cs3d.FP64: 1e-12, cs3d.FP32: 2e-6}


def gpu(prec, dt):
    g = cs3d.Solver(cs3d.config(*DIMS, precision=prec, dt_fixed=dt, **KW))
    g.set_initial_conditions(*prims(*Q))
    return g


def sigma():
    o = po.Solver(po.make_config(*DIMS, **KW))
    s = o.sponge_sigma()
    o.close()
    return s


def uniform_rhs(prec, dt, avg):
    g = gpu(prec, dt)
    if avg is not None:
        g.set_conserved(avg, 5)
    g.pre_step()
    g.rhs(1)
    k = g.conserved(3)
    g.close()
    return k


@pytest.mark.parametrize("prec", [cs3d.FP64, cs3d.FP32])
def test_sponge_rhs_uniform_state(prec):
    """The sponge term is measured against the same uniform state without
    one (averages equal to the state): on the non-periodic x and y lines the
    one-sided closure rows of a constant field cancel only to the rounding
    of their coefficients, which in FP32 leaves a deterministic dt*RHS of
    about 2e-7 next to the faces (1e-17 in FP64), the same in both runs."""
    dt = 0.01
    qa = conserved(*A)
    k = uniform_rhs(prec, dt, qa)
    k0 = uniform_rhs(prec, dt, None)
    q = conserved(*Q)
    sig = sigma()
    on = planes()
    errs, floor = [], []
    for v in range(5):
        want = (2.0 if v == 4 else 1.0) * dt * sig * (qa[v] - q[v])
        want[on] = 0.0
        scale = np.abs(want).max()
        errs.append(float(np.abs(k[v] - k0[v] - want).max() / scale))
        floor.append(float(np.abs(k0[v]).max() / scale))
    print(f"prec {prec}: sponge dt*RHS errors " + " ".join(f"{e:.1e}" for e in errs) +
          "; uniform-state floor " + " ".join(f"{e:.1e}" for e in floor), flush=True)
    assert max(errs) < TOL[prec]
    assert max(floor) < (1e-12 if prec == cs3d.FP64 else 1e-4)


@pytest.mark.parametrize("prec", [cs3d.FP64, cs3d.FP32])
def test_sponge_average_relaxation_and_reset(prec):
    dt = 0.3
    g = gpu(prec, dt)
    a0 = conserved(*A)
    g.set_conserved(a0, 5)
    q0 = g.conserved(1)
    g.pre_step()
    q = conserved(*Q)
    eps = dt / (15.0 + dt)
    on = planes()
    off = ~on
    tol = TOL[prec]
    worst = 0.0
    for n in range(1, 13):
        g.post_step()
        a = g.conserved(5)
        for v in range(4):
            want = q[v] + (1.0 - eps) ** n * (a0[v] - q[v])
            worst = max(worst, float(np.abs(a[v] - want)[off].max() / np.abs(want).max()))
        if n == 1:
            kin = lambda s: 0.5 * s[0] * (s[1] ** 2 + s[2] ** 2 + s[3] ** 2)
            bound = 2 * 0.005 * (abs(Q[4] - 1 / GAMMA) + (GAMMA - 1.0) * (kin(A) + kin(Q)))
            assert np.abs(pressure(a) - 1.0 / GAMMA)[off].max() < bound + tol
            # the face planes of the state hold the averages, the rest is kept
            s = g.conserved(1)
            for v in range(5):
                assert np.array_equal(s[v][on], a[v][on]), v
                assert np.array_equal(s[v][off], q0[v][off]), v
    print(f"prec {prec}: average relaxation error {worst:.1e}", flush=True)
    g.close()
    assert worst < tol


@pytest.mark.parametrize("prec", [cs3d.FP64, cs3d.FP32])
def test_sponge_pressure_blend(prec):
    dt = 0.3
    g = gpu(prec, dt)
    a0 = conserved(Q[0], Q[1], Q[2], Q[3], 0.8 / GAMMA)
    g.set_conserved(a0, 5)
    g.pre_step()
    eps, epsP = dt / (15.0 + dt), 0.005
    a_ = epsP * (1.0 - eps)
    pstar = (epsP * eps * Q[4] + (1.0 - epsP) / GAMMA) / (1.0 - a_)
    off = ~planes()
    worst = 0.0
    for n in range(1, 4):
        g.post_step()
        a = g.conserved(5)
        want = pstar + a_ ** n * (0.8 / GAMMA - pstar)
        worst = max(worst, float(np.abs(pressure(a) - want)[off].max() / want))
    print(f"prec {prec}: pressure blend error {worst:.1e}", flush=True)
    g.close()
    assert worst < TOL[prec]
