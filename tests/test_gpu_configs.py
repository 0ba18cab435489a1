"""The benchmark configurations at their full sizes.

- Taylor-Green 128^3 and 256^3, RK4 with the filter every step, 1000
  steps: the FP32 path's volume-averaged kinetic energy and the
  enstrophy-based dissipation against the FP64 path (the accuracy target:
  kinetic energy within 1e-5 relative after 1000 steps).
- Decaying isotropic turbulence 128^3 with TVDRK3 from the restated
  generator (ic.py): the FP64 GPU path against the oracle for two steps,
  FP32 against FP64 after two steps (2e-6, as for Taylor-Green) and after
  500 steps, where the turbulent field has amplified the rounding
  difference (measured 2.2e-5 in the kinetic energy; bound 1e-4).
- Temporal shear layer 512 x 128 x 42, y non-periodic with 15% sponges,
  RK4: the FP64 GPU path against the oracle for two steps (the sponge
  update and its stale-plane stage included), then FP32 against FP64 over
  500 steps."""
import numpy as np
import pytest

import cs3d
import ic
import pyoracle as po
from helpers import relerr, tgv_ic

pytestmark = pytest.mark.gpu


def run(dims, kw, icf, prec, nsteps, diag=None):
    Nx, Ny, Nz = dims
    g = cs3d.Solver(cs3d.config(Nx, Ny, Nz, precision=prec, **kw))
    g.set_initial_conditions(*icf)
    out = []
    if diag:
        g.set_diagnostics(True)
    for n in range(nsteps):
        g.step(1)
        if diag and (n + 1) % diag == 0:
            out.append(g.diag_taylor_green())
    ke = g.kinetic_energy()
    q = g.conserved(1)
    g.close()
    return ke, q, out


@pytest.mark.parametrize("N", [128, 256])
def test_tgv_fp32_tracks_fp64_1000_steps(N):
    L = 2 * np.pi - 2 * np.pi / N
    kw = dict(Lx=L, Ly=L, Lz=L, alpha_f=0.49, mu_ref=(1 / 1600) / 100 ** 0.76, Pr=0.71,
              cfl=0.5, filter_step=1, rk=4)
    icf = tgv_ic(N, L)
    res = {p: run((N, N, N), kw, icf, p, 1000, diag=250) for p in (cs3d.FP64, cs3d.FP32)}
    ke64, ke32 = res[cs3d.FP64][0], res[cs3d.FP32][0]
    rel = abs(ke32 - ke64) / abs(ke64)
    print(f"TGV {N}^3 1000 steps: KE fp64 {ke64:.12e} fp32 {ke32:.12e} rel {rel:.2e}", flush=True)
    assert rel <= 1e-5
    for d64, d32 in zip(res[cs3d.FP64][2], res[cs3d.FP32][2]):
        # columns: time, KE sum, dissipation sum, enstrophy sum, ...
        e = abs(d32[2] - d64[2]) / abs(d64[2])
        print(f"  t {d64[0]:.4f}: dissipation rel {e:.2e}", flush=True)
        assert e <= 1e-3


def hit_kw():
    L = 2 * np.pi
    return dict(Lx=L, Ly=L, Lz=L, alpha_f=0.495, mu_ref=0.00033, Pr=0.7, cfl=0.25,
                filter_step=50, rk=3)


def test_hit128_tvdrk3():
    N = 128
    kw = hit_kw()
    icf = ic.hit_initial_conditions(N, k0=8.0, seed=0)
    o = po.Solver(po.make_config(N, N, N, **kw))
    o.set_ic(*icf)
    g = cs3d.Solver(cs3d.config(N, N, N, precision=cs3d.FP64, **kw))
    g.set_initial_conditions(*icf)
    for n in range(2):
        o.step()
        g.step(1)
        errs = [relerr(a, b) for a, b in zip(g.conserved(1), o.conserved(1))]
        print(f"HIT step {n + 1}: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
        assert max(errs) <= 1e-11
    o.close()
    g.close()
    _, q64, _ = run((N, N, N), kw, icf, cs3d.FP64, 2)
    _, q32, _ = run((N, N, N), kw, icf, cs3d.FP32, 2)
    errs = [relerr(a, b) for a, b in zip(q32, q64)]
    print("HIT fp32 vs fp64, 2 steps: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 2e-6
    k64, q64, _ = run((N, N, N), kw, icf, cs3d.FP64, 500)
    k32, q32, _ = run((N, N, N), kw, icf, cs3d.FP32, 500)
    rel = abs(k32 - k64) / abs(k64)
    errs = [relerr(a, b) for a, b in zip(q32, q64)]
    # plain FP32 drifts ~2e-5 here; the compensated FP32 mode's test holds
    # the same run to 1e-6
    print(f"HIT 500 steps: KE rel {rel:.2e}; fields " + " ".join(f"{e:.2e}" for e in errs),
          flush=True)
    assert rel <= 1e-4
    assert max(errs[1:4]) <= 1e-3


def shear_kw():
    return dict(Lx=172.0, Ly=129.0, Lz=28.2, alpha_f=0.495, mu_ref=0.00375, Pr=0.7, cfl=0.8,
                filter_step=5, rk=4, bc_type=(0, 1, 0), bc_kind=(0, 0, 1, 1, 0, 0),
                sponge_frac=(0.25, 0.15, 0.25))


def test_shear_layer_512x128x42():
    dims = (512, 128, 42)
    kw = shear_kw()
    icf = ic.shear_layer_ic(*dims, Ly=kw["Ly"], delta_u=1.2, delta0=1.0, seed=0)
    o = po.Solver(po.make_config(*dims, **kw))
    o.set_ic(*icf)
    g = cs3d.Solver(cs3d.config(*dims, precision=cs3d.FP64, **kw))
    g.set_initial_conditions(*icf)
    for n in range(2):
        o.step()
        g.step(1)
        errs = [relerr(a, b) for a, b in zip(g.conserved(1), o.conserved(1))]
        print(f"shear step {n + 1}: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
        assert max(errs) <= 1e-11
    o.close()
    g.close()
    k64, q64, _ = run(dims, kw, icf, cs3d.FP64, 500)
    k32, q32, _ = run(dims, kw, icf, cs3d.FP32, 500)
    rel = abs(k32 - k64) / abs(k64)
    errs = [relerr(a, b) for a, b in zip(q32, q64)]
    print(f"shear 500 steps: KE rel {rel:.2e}; fields " + " ".join(f"{e:.2e}" for e in errs),
          flush=True)
    assert rel <= 1e-5

