"""The FP32 path filters along x inside the last stage's x sweep (periodic
uniform x lines) instead of in a separate sweep, so the three filter
directions run in another order than the reference's round robin.  Filters
along different axes commute, so both orders give the same state up to
rounding: the fused run stays within FP32 rounding of the unfused one
(CS3D_NO_FILTER_FUSION in the environment at solver creation turns the
fusion off), across the three round-robin orders and a step that does not
filter."""
import os

import numpy as np
import pytest

import cs3d
from helpers import random_ic, relerr, tgv_ic

pytestmark = pytest.mark.gpu


def run(fused, nsteps, filter_step, dims=(64, 64, 64), rk=4, env=None):
    N = 64
    L = 2 * np.pi - 2 * np.pi / N
    kw = dict(Lx=L, Ly=L, Lz=L, alpha_f=0.49, mu_ref=(1 / 1600) / 100 ** 0.76, Pr=0.71,
              cfl=0.5, filter_step=filter_step, rk=rk)
    env = dict(env or {})
    if not fused:
        env["CS3D_NO_FILTER_FUSION"] = "1"
    os.environ.update(env)
    try:
        g = cs3d.Solver(cs3d.config(*dims, precision=cs3d.FP32, **kw))
    finally:
        for k in env:
            os.environ.pop(k, None)
    if dims == (64, 64, 64):
        g.set_initial_conditions(*tgv_ic(N, L))
    else:
        g.set_initial_conditions(*random_ic(dims[::-1], 5, amp=0.2))
    g.step(nsteps)
    q = g.conserved(1)
    g.close()
    return q


@pytest.mark.parametrize("filter_step", [1, 2])
def test_fused_x_filter_matches_separate_sweeps(filter_step):
    a = run(True, 7, filter_step)
    b = run(False, 7, filter_step)
    # the momenta against the flow's momentum scale (rho W is only rounding
    # noise of the Taylor-Green field at these times)
    mscale = max(np.abs(b[1]).max(), np.abs(b[2]).max())
    errs = [relerr(a[0], b[0])] + [float(np.abs(a[v] - b[v]).max() / mscale) for v in (1, 2, 3)] + \
        [relerr(a[4], b[4])]
    print(f"filter_step {filter_step}: fused vs separate " + " ".join(f"{e:.1e}" for e in errs),
          flush=True)
    assert max(errs) < 2e-6


@pytest.mark.parametrize("rk", [4, 3])
def test_fused_x_filter_partial_workgroup(rk):
    """16-point x lines: 16 lines per workgroup, 120 lines, so the last
    workgroup is partial (its idle lanes pass the filter's barriers)"""
    dims = (16, 12, 10)
    a = run(True, 5, 1, dims, rk)
    b = run(False, 5, 1, dims, rk)
    errs = [relerr(a[v], b[v]) for v in range(5)]
    print(f"rk{rk} {dims}: fused vs separate " + " ".join(f"{e:.1e}" for e in errs), flush=True)
    assert max(errs) < 5e-6


@pytest.mark.parametrize("dims,env", [((512, 8, 8), None),
                                      ((64, 16, 16), {"CS3D_SWEEP_X_THREADS": "256"})])
def test_fused_x_filter_multi_wavefront(dims, env):
    """Workgroups of several wavefronts on the fused filter's LDS path:
    512-point x lines (128 segments: one line of two wavefronts per
    workgroup) and 64-point lines (16 segments) in 256-lane workgroups of 16
    lines.  Every wavefront must have
    read its halos from the staging area before any of them writes the
    exchange ring that shares that LDS."""
    a = run(True, 6, 1, dims, env=env)
    b = run(False, 6, 1, dims, env=env)
    errs = [relerr(a[v], b[v]) for v in range(5)]
    print(f"{dims} {env}: fused vs separate " + " ".join(f"{e:.1e}" for e in errs), flush=True)
    assert max(errs) < 2e-6
