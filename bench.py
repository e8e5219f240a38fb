"""Benchmark: cell-updates/s of aither's hot path on one MI355X.

Workload (a "step" is one nonlinear iteration of the whole hot path: ghost
cells, fused face-flux residual + spectral radius + dt + diagonal, LU-SGS
hyperplane sweeps, matrix residual, update and norm reductions):
inviscid Euler, calorically perfect air, one Cartesian block 64^3 on
[0,10]^3, isentropic vortex (strength 5, core radius 1) on a Mach 0.5
freestream, characteristic farfield BCs, Roe flux, MUSCL (kappa = 1/3) with
van Albada limiter, implicit Euler LU-SGS at CFL 5.  The timed region runs
`--steps` iterations from the initial condition (200 = the configuration's
full run); the state is resident in HBM before timing starts.

Prints one JSON line.  Run `python bench.py --help` for options.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLUX_BYTES_PER_CELL = 224      # W 40 + geometry 128 + R 40 + lambda 8 + diag 8 (FP64)
SWEEP_BYTES_PER_CELL = 504     # LU-SGS forward + backward, FP64 operator
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--n", type=int, default=64, help="cells per direction")
    p.add_argument("--precision", choices=["fp64", "fp32"], default="fp64",
                   help="implicit-operator precision (residual is always FP64)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-iters", type=int, default=12)
    p.add_argument("--no-fp32-leg", action="store_true")
    p.add_argument("--sweep", choices=["default", "hyperplane", "tiled", "pencil"], default="default",
                   help="LU-SGS execution order (all give identical results)")
    return p.parse_args()


def traffic_from_profiles():
    """HBM bytes per flux launch from the committed rocprofv3 PMC summary."""
    path = os.path.join(ROOT, "profiles", "flux_hbm_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    except Exception:
        return None


def cpu_baseline(inp, iters):
    """The oracle (single-threaded FP64 C restatement) on the first `iters`
    iterations of the same case."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    c = pyoracle.OracleCase(inp)
    ncell = c.ncells()
    t0 = time.perf_counter()
    for nn in range(iters):
        c.begin_step(nn)
        c.iterate(nn, 0)
    dt = time.perf_counter() - t0
    return {"value": ncell * iters / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"first {iters} iterations of the same {c.dims[0][0]}^3 case, oracle C restatement, "
                      f"1 thread ({dt:.1f} s)"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)
    from aither_amd import Solver, cases

    work = tempfile.mkdtemp(prefix="aither_bench_")
    inp = cases.vortex_case(work, n=args.n, iterations=args.steps)
    s = Solver(inp)
    prec = Solver.FP32 if args.precision == "fp32" else Solver.FP64
    s.set_implicit_precision(prec)
    order = None if args.sweep == "default" else getattr(Solver, args.sweep.upper())
    if order is not None:
        s.set_sweep_ordering(order)
    W0 = s.get_state()
    ncell = s.ncells(0)

    # warmup (graph capture, caches), then restart from the initial condition
    if args.warmup > 0:
        s.run(0, args.warmup)
    s.set_state(W0)
    s.enable_timing(True)
    t0 = time.perf_counter()
    last = s.run(0, args.steps)
    t1 = time.perf_counter()
    wall = t1 - t0
    tm = s.timing()
    value = ncell * args.steps / wall

    flux_avg_s = tm.flux_ms / max(tm.flux_launches, 1) * 1e-3
    sweep_avg_s = tm.sweep_ms / max(tm.sweep_launches, 1) * 1e-3
    flux_gbs = FLUX_BYTES_PER_CELL * ncell / flux_avg_s / 1e9
    sweep_gbs = SWEEP_BYTES_PER_CELL * ncell / sweep_avg_s / 1e9
    out = {
        "metric": "cell-updates/s at <=1e-10 err vs FP64 oracle",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == Solver.FP64 else "f64-residual/f32-implicit",
        "data": "synthetic (seed-free analytic isentropic vortex)",
        "config": {
            "workload": f"euler-vortex-{args.n}^3: 1 block {args.n}^3 cells, Roe, MUSCL kappa=1/3 "
                        "van Albada, implicit Euler LU-SGS CFL 5, characteristic farfield",
            "cells": ncell, "iterations": args.steps, "parallelism": "single GPU",
        },
        "roofline": {
            "kernel": "KResidual (fused face flux + spectral radius + dt + diagonal)",
            "bound": "hbm",
            "achieved": flux_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": flux_gbs / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles(),
            "avg_launch_us": flux_avg_s * 1e6,
            "algorithmic_bytes_per_launch": FLUX_BYTES_PER_CELL * ncell,
        },
        "sweep": {
            "kernels": "LU-SGS forward+backward sweeps of one iteration",
            "avg_per_iteration_us": sweep_avg_s * 1e6,
            "algorithmic_GBs": sweep_gbs,
            "frac_hbm": sweep_gbs / HBM_PEAK_GBS,
        },
        "final_normalized_residual": list(last.normalized[:5]),
    }
    s.close()

    if not args.no_fp32_leg and prec == Solver.FP64:
        # same run with the FP32 implicit operator: throughput and the
        # difference of the final state from the FP64-operator run
        import numpy as np
        ref = None
        s64 = Solver(inp)
        if order is not None:
            s64.set_sweep_ordering(order)
        s64.run(0, args.steps)
        ref = s64.get_state()
        s64.close()
        s32 = Solver(inp)
        s32.set_implicit_precision(Solver.FP32)
        if order is not None:
            s32.set_sweep_ordering(order)
        s32.run(0, min(args.warmup, 5))
        s32.set_state(W0)
        t0 = time.perf_counter()
        s32.run(0, args.steps)
        w32 = time.perf_counter() - t0
        st = s32.get_state()
        s32.close()
        out["fp32_operator"] = {
            "value": ncell * args.steps / w32,
            "state_rel_l2_vs_fp64": float(np.linalg.norm(st - ref) / np.linalg.norm(ref)),
        }

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(inp, args.cpu_iters)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
