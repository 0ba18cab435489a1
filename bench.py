#!/usr/bin/env python3
"""Benchmark of the compact RHS pipeline on one MI355X.

Workload (one "step" = one RK4 time step = 4 RHS stages + the filter in x, y, z):
Taylor-Green vortex, 256^3 periodic on [0, 2pi)^3, RK4, 8th-order filter
every step (alpha_f 0.49), CFL 0.5, Re 1600, Mach 0.1, gamma 1.4, Pr 0.71,
isothermal density; FP32 bulk arithmetic.  The same workload in FP64 is
timed beside it.  Inputs are resident in HBM before the timed region.

value  = point-stages per second (N^3 * 4 * steps / wall time), FP32.
% roof = value * bytes per point-stage / 8 TB/s, on two bases:
         hbm_roof_frac_fixed_model: the fixed split-sweep model of the
         path, 322 field accesses per point and RK4 step with the filter
         (73 per stage + 30 for the filter), 322 B per point-stage in FP32
         and 644 B in FP64 -- the basis of the 60% target;
         hbm_roof_frac: this build's own accesses (STAGE_ACCESSES: 292 per
         step, the x filter running inside the last x sweep).
roofline = the dominant kernel class, timed with HIP events recorded on the
         solver's own stream around every launch of the timed region.
adjoint = cost of the kinetic-energy gradient through 40 RK4 steps of the
         same workload with a checkpoint every 20 steps: the forward pass,
         the replay of the first segment from its checkpoint, and the
         reverse sweep, against 40 forward steps (reverse_over_forward =
         (replay + reverse) / forward).
fp32_compensated = the same FP32 workload in the compensated mode (state,
         accumulator and stage sums carried as FP32 hi + lo pairs).
shear_layer = the temporal shear layer 512 x 128 x 42 in FP32 (sponges on
         non-periodic y lines, ragged periodic z lines: the non-uniform
         line path), forward step time, and its kinetic-energy gradient
         through 20 steps with a checkpoint every 5 on the adjoint's
         general path (transposed operators, sponge reverse).
cpu_baseline = the oracle (FP64 OpenMP restatement of the reference, built
         with the reference's flags) timed on this host's cores at 256^3
         (1 warm-up + 3 steps), with a 128^3 sample beside it.
hbm_stream_GBps = the library's hand-written 16-byte-vector copy and triad
         kernels over 1 GiB arrays, and the column sweeps' mix (5 arrays
         read, 9 written per element) over 512 MiB arrays, bytes read +
         written per second: the HBM rate plain streams reach on this GPU,
         beside the 8 TB/s peak.

Only one GPU: --gpus N with N > 1 exits with a message.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "3dcompactsolver_amd"))

import cs3d  # noqa: E402

METRIC = "Mpoint-RK-stages/s at 256^3 FP32 vs FP64 (% HBM roof); adjoint/forward ratio"
HBM_PEAK = 8.0e12
# algorithmic field accesses per point of every launch in one RK4 step with
# the filter (no sponge).  The sweeps of a stage run y, z, x: the y sweep
# reads the 5 conserved fields and writes Uy Vy Wy Ty and its 5 partial
# right-hand sides (5 + 9); the z sweep also reads Vy, Wy and writes
# Uz Vz Wz Tz and its own 5 partial sums (7 + 9); the x sweep reads the
# conserved fields, Uy Vy Uz Wz, Wy Ty Vz Tz and both partial sums (23),
# then *1 and *2 from the second stage on (at the first stage *1 is the
# stage state it already holds, and *2 starts from it), and writes *2 and,
# before the last stage, *k.  At the last stage of the step the x filter
# runs inside the sweep (x lines of 256 points: 64 segments, one wavefront
# each; FP32 also at other lengths): it reads only *2 (not *1) and writes the
# filtered new state to *1 instead of *2.
# Each filter sweep reads and writes the 5 conserved fields; the CFL
# reduction rides on the last one: two filter sweeps (y, z).
def stage_accesses(prec):
    return {
        "y_sweep": [14, 14, 14, 14],
        "z_sweep": [16, 16, 16, 16],
        "x_sweep": [23 + 10, 33 + 10, 33 + 10, 28 + 5],
        "filter": [10, 10],
    }


STAGE_ACCESSES = {p: stage_accesses(p) for p in (cs3d.FP32, cs3d.FP64)}
# per launch, averaged over the launches of a step
ACCESSES = {p: {k: sum(v) / len(v) for k, v in a.items()} for p, a in STAGE_ACCESSES.items()}
# 292 in both precisions
ACCESSES_PER_STEP = {p: sum(sum(v) for v in a.values()) for p, a in STAGE_ACCESSES.items()}
# HBM accesses per point actually moved per launch (reads, writes) at 256^3,
# from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs calibrated on a 256 MiB
# copy (profiles/hbm_traffic_tgv256_fp32.md, profiles/hbm_traffic_tgv256_fp64.md)
# (the x sweep's: the mean over its four launches of a step)
MEASURED_TRAFFIC = {
    cs3d.FP32: {"x_sweep": (29.25, 8.75), "y_sweep": (5.00, 9.36), "z_sweep": (7.00, 9.30)},
    cs3d.FP64: {"x_sweep": (29.25, 8.88), "y_sweep": (5.00, 9.75), "z_sweep": (7.00, 9.80)},
}
# bytes per point-stage: ACCESSES_PER_STEP accesses per point over 4 stages
BYTES_PER_POINT_STAGE = {cs3d.FP32: ACCESSES_PER_STEP[cs3d.FP32] * 4 / 4,
                         cs3d.FP64: ACCESSES_PER_STEP[cs3d.FP64] * 8 / 4}
# the fixed split-sweep model: per RK4 stage X 5 + 9, Y 12 + 9, Z with the
# assembly and update 28 + 10 = 73 accesses; the filter 3 sweeps x 5 fields
# x (read + write) = 30 per step: 4 * 73 + 30 = 322 per point and step
FIXED_MODEL_ACCESSES_PER_STEP = 4 * 73 + 30
FIXED_BYTES_PER_POINT_STAGE = {cs3d.FP32: FIXED_MODEL_ACCESSES_PER_STEP * 4 / 4,
                               cs3d.FP64: FIXED_MODEL_ACCESSES_PER_STEP * 8 / 4}


def roof_fracs(value, prec):
    """(fixed-model fraction, own-model fraction) of the 8 TB/s peak"""
    return (round(value * 1e6 * FIXED_BYTES_PER_POINT_STAGE[prec] / HBM_PEAK, 4),
            round(value * 1e6 * BYTES_PER_POINT_STAGE[prec] / HBM_PEAK, 4))


def tgv_ic(N, L, mach=0.1):
    x = np.arange(N) * (L / (N - 1))
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    U = np.sin(X) * np.cos(Y) * np.cos(Z)
    V = -np.cos(X) * np.sin(Y) * np.cos(Z)
    W = np.zeros_like(U)
    p0 = 1.0 / (1.4 * mach ** 2)
    p = p0 + (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0) / 16.0
    rho = p / p0
    return rho, U, V, W, p


def tgv_kw(N):
    L = 2 * np.pi - 2 * np.pi / N
    # mu(T0) = 1/Re with T0 = 100 and mu = mu_ref T^0.76 (IdealGas.hpp:49-51)
    return dict(Lx=L, Ly=L, Lz=L, alpha_f=0.49, mu_ref=(1 / 1600) / 100 ** 0.76, Pr=0.71,
                cfl=0.5, filter_step=1, rk=4), L


def run_gpu(N, prec, steps, warmup, ic, **extra):
    kw, _ = tgv_kw(N)
    g = cs3d.Solver(cs3d.config(N, N, N, precision=prec, **kw, **extra))
    g.set_initial_conditions(*ic)
    g.step(warmup)
    g.synchronize()
    g.profile(True)
    t0 = time.perf_counter()
    g.step(steps)
    g.synchronize()
    t1 = time.perf_counter()
    prof = g.profile_report()
    g.profile(False)
    ke = g.kinetic_energy()
    g.close()
    ms = (t1 - t0) / steps * 1e3
    mpts = N ** 3 * 4 / (ms * 1e-3) / 1e6
    return ms, mpts, prof, ke


def run_adjoint(N, prec, ic, nsteps, checkpoint, fwd_ms):
    """kinetic-energy gradient through nsteps RK4 steps (forward pass,
    checkpointed replay, reverse sweep), gradients left on the device"""
    kw, _ = tgv_kw(N)
    g = cs3d.Solver(cs3d.config(N, N, N, precision=prec, **kw))
    g.set_initial_conditions(*ic)
    g.ke_gradient(nsteps, checkpoint, outputs=False)  # buffer set-up outside the timed call
    g.synchronize()
    t0 = time.perf_counter()
    g.ke_gradient(nsteps, checkpoint, outputs=False)
    g.synchronize()
    t1 = time.perf_counter()
    g.close()
    total = (t1 - t0) * 1e3
    fwd = nsteps * fwd_ms
    return {"steps": nsteps, "checkpoint": checkpoint, "gradient_ms": round(total, 2),
            "forward_ms": round(fwd, 2),
            "reverse_over_forward": round((total - fwd) / fwd, 2),
            "gradient_over_forward": round(total / fwd, 2)}


SHEAR_DIMS = (512, 128, 42)
SHEAR_KW = dict(Lx=172.0, Ly=129.0, Lz=28.2, alpha_f=0.495, mu_ref=0.00375, Pr=0.7, cfl=0.8,
                filter_step=5, rk=4, bc_type=(0, 1, 0), bc_kind=(0, 0, 1, 1, 0, 0),
                sponge_frac=(0.25, 0.15, 0.25))


def run_shear(steps, warmup, adj_steps=20, adj_checkpoint=5):
    """the temporal shear layer 512 x 128 x 42 (y non-periodic between sponge
    layers, z periodic on 42 points: the non-uniform line path) in FP32:
    forward step time, and the kinetic-energy gradient on the adjoint's
    general path (dt frozen at the first step's CFL dt; second call timed)"""
    import ic as icmod
    dims = SHEAR_DIMS
    icf = icmod.shear_layer_ic(*dims, Ly=SHEAR_KW["Ly"], delta_u=1.2, delta0=1.0, seed=0)
    g = cs3d.Solver(cs3d.config(*dims, precision=cs3d.FP32, **SHEAR_KW))
    g.set_initial_conditions(*icf)
    g.pre_step()
    dt = g.time()[1]
    kw = dict(SHEAR_KW, dt_fixed=dt)
    g.close()
    g = cs3d.Solver(cs3d.config(*dims, precision=cs3d.FP32, **kw))
    g.set_initial_conditions(*icf)
    g.step(warmup)
    g.synchronize()
    t0 = time.perf_counter()
    g.step(steps)
    g.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    npts = dims[0] * dims[1] * dims[2]
    out = {"grid": list(dims), "ms_per_step": round(ms, 3),
           "value": round(npts * 4 / (ms * 1e-3) / 1e6, 1), "unit": "Mpoint-RK-stages/s"}
    for _ in range(2):
        g.set_initial_conditions(*icf)
        g.synchronize()
        t0 = time.perf_counter()
        g.ke_gradient(adj_steps, adj_checkpoint, outputs=False)
        g.synchronize()
        tot = (time.perf_counter() - t0) * 1e3
    g.close()
    fwd = adj_steps * ms
    out["adjoint"] = {"steps": adj_steps, "checkpoint": adj_checkpoint,
                      "gradient_ms": round(tot, 2), "forward_ms": round(fwd, 2),
                      "reverse_over_forward": round((tot - fwd) / fwd, 2),
                      "gradient_over_forward": round(tot / fwd, 2)}
    return out


def roofline(prof, N, prec):
    s = 4 if prec == cs3d.FP32 else 8
    acc = ACCESSES[prec]
    best = max((k for k in acc), key=lambda k: prof[k][0])
    tot_ms, cnt = prof[best]
    avg_s = tot_ms / cnt * 1e-3
    alg_bytes = int(acc[best] * s * N ** 3)
    achieved = alg_bytes / avg_s / 1e9
    traffic = None
    if N == 256 and best in MEASURED_TRAFFIC[prec]:
        traffic = int(sum(MEASURED_TRAFFIC[prec][best]) * s * N ** 3)
    return best, {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK / 1e9,
                  "unit": "GB/s", "frac": round(achieved / (HBM_PEAK / 1e9), 4),
                  "traffic": traffic, "kernel": best,
                  "algorithmic_bytes_per_launch": alg_bytes,
                  "avg_launch_ms": round(tot_ms / cnt, 4)}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip() + f" ({os.cpu_count()} logical CPUs)"
    except OSError:
        pass
    return "unknown"


def stream_bandwidth():
    """the library's copy and triad kernels over 1 GiB arrays (the best of
    six launch shapes, 5 timed after 2 warm-up launches each): best and
    mean GB/s"""
    out = {}
    for kind in ("copy", "triad"):
        best, mean = cs3d.stream_probe(kind, 1 << 30, 5)
        out[kind] = {"best": round(best, 1), "mean": round(mean, 1)}
    # the column sweeps' read/write mix (5 read, 9 written; 14 arrays of
    # 512 MiB)
    best, mean = cs3d.stream_probe("mix", 1 << 29, 5)
    out["mix_5r9w"] = {"best": round(best, 1), "mean": round(mean, 1)}
    return out


def cpu_baseline(N, steps):
    """FP64 OpenMP restatement of the reference (liboracle_fast.so, reference
    compiler flags), TGV N^3, one warm-up step then `steps` timed steps.
    The oracle is the CPU baseline here, never the measured GPU path."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    kw, L = tgv_kw(N)
    o = po.Solver(po.make_config(N, N, N, **kw), fast=True)
    o.set_ic(*tgv_ic(N, L))
    o.step()
    t0 = time.perf_counter()
    for _ in range(steps):
        o.step()
    t1 = time.perf_counter()
    o.close()
    mpts = N ** 3 * 4 * steps / (t1 - t0) / 1e6
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(mpts, 3), "unit": "Mpoint-RK-stages/s", "cores": threads,
            "kind": "port", "host": cpu_model(),
            "sample": f"Taylor-Green {N}^3 FP64 RK4 + filter, {steps} steps after 1 warm-up "
                      f"({t1 - t0:.1f} s of CPU time on {threads} OpenMP threads)"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-grid", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--cpu-small-steps", type=int, default=6,
                    help="timed steps of the 128^3 CPU sample beside the main one (0: none)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint cost leg")
    ap.add_argument("--no-compensated", action="store_true",
                    help="skip the compensated FP32 leg")
    ap.add_argument("--no-shear", action="store_true", help="skip the shear-layer leg")
    ap.add_argument("--adjoint-steps", type=int, default=40)
    ap.add_argument("--adjoint-checkpoint", type=int, default=20,
                    help="steps between adjoint checkpoints; 0: all steps (every stage state "
                         "of the run stays in HBM and nothing is recomputed)")
    ap.add_argument("--fp64-only", action="store_true",
                    help="run only the FP64 leg and print its step time (for profiling it)")
    ap.add_argument("--shear-only", action="store_true",
                    help="run only the shear-layer leg (forward and its gradient on the "
                         "adjoint's general path) and print it (for profiling it)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run three 256 MiB device copies (the reference kernel that "
                         "calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        # 256 MiB copies by the library's copy kernel: six launch shapes, 2
        # untimed + 3 timed launches each (every dispatch moves the same bytes)
        best, _ = cs3d.stream_probe("copy", 256 << 20, 3)
        print(f"copy probe: 30 x 256 MiB, best {best:.0f} GB/s", flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "decomposition is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)
    if args.shear_only:
        print(json.dumps({"shear_layer": run_shear(args.steps, args.warmup)}), flush=True)
        return
    N = args.grid
    kw, L = tgv_kw(N)
    ic = tgv_ic(N, L)
    if args.fp64_only:
        ms64, v64, prof64, _ = run_gpu(N, cs3d.FP64, args.steps, args.warmup, ic)
        print(json.dumps({"fp64_ms_per_step": round(ms64, 3), "value": round(v64, 1),
                          "kernel_ms_per_step": {k: round(v[0] / args.steps, 3)
                                                 for k, v in prof64.items()}}), flush=True)
        return
    ms32, v32, prof32, ke32 = run_gpu(N, cs3d.FP32, args.steps, args.warmup, ic)
    best, roof = roofline(prof32, N, cs3d.FP32)
    fp64 = None
    if not args.no_fp64:
        ms64, v64, prof64, ke64 = run_gpu(N, cs3d.FP64, args.steps, args.warmup, ic)
        _, roof64 = roofline(prof64, N, cs3d.FP64)
        fixed64, own64 = roof_fracs(v64, cs3d.FP64)
        fp64 = {"value": round(v64, 1), "ms_per_step": round(ms64, 3),
                "hbm_roof_frac_fixed_model": fixed64,
                "hbm_roof_frac": own64,
                "roofline": roof64,
                "kinetic_energy": ke64,
                "fp32_ke_rel_err": abs(ke32 - ke64) / abs(ke64)}
    # the compensated FP32 mode (off by default in the library): its cost
    cmp = None
    if not args.no_compensated:
        msc, vc, _, kec = run_gpu(N, cs3d.FP32, args.steps, args.warmup, ic, compensated=1)
        cmp = {"value": round(vc, 1), "ms_per_step": round(msc, 3),
               "hbm_roof_frac_fixed_model": roof_fracs(vc, cs3d.FP32)[0]}
        if fp64:
            cmp["ke_rel_err_vs_fp64"] = abs(kec - fp64["kinetic_energy"]) / abs(fp64["kinetic_energy"])
    ck = args.adjoint_checkpoint if args.adjoint_checkpoint > 0 else args.adjoint_steps
    adj = None if args.no_adjoint else run_adjoint(N, cs3d.FP32, ic, args.adjoint_steps, ck, ms32)
    shear = None if args.no_shear else run_shear(args.steps, args.warmup)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_grid, args.cpu_steps)
        if args.cpu_small_steps > 0 and args.cpu_grid != 128:
            small = cpu_baseline(128, args.cpu_small_steps)
            cpu["at_128"] = {k: small[k] for k in ("value", "unit", "sample")}
    stream_bw = stream_bandwidth()
    fixed32, own32 = roof_fracs(v32, cs3d.FP32)
    line = {
        "metric": METRIC,
        "value": round(v32, 1),
        "unit": "Mpoint-RK-stages/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms32, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Taylor-Green initial field)",
        "config": {"workload": f"Taylor-Green vortex {N}^3 periodic, RK4, filter every step "
                               "(alpha_f 0.49), CFL 0.5, Re 1600, Mach 0.1, FP32 bulk",
                   "grid": [N, N, N], "rk": "RK4", "filter_step": 1},
        "hbm_roof_frac_fixed_model": fixed32,
        "hbm_roof_frac": own32,
        "kernel_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in prof32.items()},
        "fp64": fp64,
        "fp32_compensated": cmp,
        "adjoint": adj,
        "shear_layer": shear,
        "roofline": roof,
        "hbm_stream_GBps": stream_bw,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
