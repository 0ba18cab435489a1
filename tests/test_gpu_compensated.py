"""The compensated FP32 mode (cs3d_config.compensated): the state, the RK4
accumulator and TVDRK3's stage state carry FP32 low parts, the stage update
forms hi + lo + c dt R in FP64 and splits it back, and the filter sweeps
subtract their correction from hi + lo (RK4.hpp:45-117, TVDRK3.hpp:38-96,
filterConservedData).  Against the oracle per field over the first steps
(Taylor-Green RK4 with the filter, fused and separate; the shear layer with
sponges), and against FP64 over 500 TVDRK3 steps of decaying turbulence,
where the plain FP32 path's kinetic energy drifts by about 2e-5: an FP64
run from the same fields rounded to FP32 stays within 6e-10 of FP64, so
that drift is the rounding of the state updates, which this mode removes."""
import numpy as np
import pytest

import cs3d
import ic
import pyoracle as po
from helpers import shear_ic, state_errs, tgv_ic
from test_gpu_configs import hit_kw, run

pytestmark = pytest.mark.gpu


def oracle_pair(dims, icf, **kw):
    o = po.Solver(po.make_config(*dims, **kw))
    o.set_ic(*icf)
    g = cs3d.Solver(cs3d.config(*dims, precision=cs3d.FP32, compensated=1, **kw))
    g.set_initial_conditions(*icf)
    return o, g


@pytest.mark.parametrize("fused", [True, False])
def test_tgv_steps_match_oracle(fused, monkeypatch):
    if not fused:
        monkeypatch.setenv("CS3D_NO_FILTER_FUSION", "1")
    N = 32
    L = 2 * np.pi - 2 * np.pi / N
    kw = dict(Lx=L, Ly=L, Lz=L, alpha_f=0.49, mu_ref=(1 / 1600) / 100 ** 0.76, Pr=0.71,
              cfl=0.5, filter_step=1, rk=4)
    o, g = oracle_pair((N, N, N), tgv_ic(N, L), **kw)
    worst = 0.0
    for n in range(1, 5):
        o.step()
        g.step(1)
        errs = state_errs(g.conserved(1), o.conserved(1), momentum_scale=True)
        print(f"compensated tgv fused={fused} step {n}: " + " ".join(f"{e:.1e}" for e in errs),
              flush=True)
        worst = max(worst, max(errs))
    g.close()
    o.close()
    assert worst < 2e-6


def test_shear_layer_sponges_match_oracle():
    dims = (32, 24, 16)
    kw = dict(Lx=8.0, Ly=10.0, Lz=4.0, alpha_f=0.49, mu_ref=2e-3, Pr=0.7, cfl=0.5,
              filter_step=2, rk=4, bc_type=(0, 1, 0), bc_kind=(0, 0, 1, 1, 0, 0),
              sponge_frac=(0.25, 0.15, 0.25))
    o, g = oracle_pair(dims, shear_ic(32, 24, 16, 10.0, 3), **kw)
    for n in range(1, 5):
        o.step()
        g.step(1)
        errs = state_errs(g.conserved(1), o.conserved(1), momentum_scale=True)
        print(f"compensated shear step {n}: " + " ".join(f"{e:.1e}" for e in errs), flush=True)
        assert max(errs) < 2e-6
    g.close()
    o.close()


def test_hit128_tvdrk3_500_steps_tracks_fp64():
    N = 128
    kw = hit_kw()
    icf = ic.hit_initial_conditions(N, k0=8.0, seed=0)
    k64, q64, _ = run((N, N, N), kw, icf, cs3d.FP64, 500)
    k32, _, _ = run((N, N, N), dict(kw, compensated=1), icf, cs3d.FP32, 500)
    rel = abs(k32 - k64) / abs(k64)
    print(f"HIT 500 steps: compensated FP32 KE vs FP64 {rel:.2e}", flush=True)
    assert rel <= 1e-5


def test_compensated_refusals():
    """the compensated mode is an FP32 mode"""
    with pytest.raises(cs3d.CS3DError):
        cs3d.Solver(cs3d.config(16, 16, 16, precision=cs3d.FP64, compensated=1))
