"""Benchmark: cell-updates/s of aither's hot path on one MI355X.

Workload (a "step" is one nonlinear iteration of the whole hot path: ghost
cells, fused face-flux residual + spectral radius + dt + diagonal, LU-SGS
forward and backward sweeps, matrix residual, update and norm reductions):
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
# LU-SGS, FP64 operator: forward pass W 40 + R 40 + U^n 40 + geometry 96 +
# V 8 + dt 8 + D^-1 8 + x 40 = 280; backward pass W 40 + geometry 96 + D^-1 8
# + x read 40 + x write 40 = 224; a pass is one forward or backward launch
SWEEP_BYTES_PER_CELL = {"fwd": 280, "bwd": 224}
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch on the 64^3 workload (FP64), from rocprofv3
# FETCH_SIZE / WRITE_SIZE scaled by the same counters on a 256 MiB device
# copy (FETCH_SIZE reads half the copied bytes: factor 2.0; WRITE_SIZE 1.0);
# see profiles/euler-vortex-64_hbm_counters.md
FLUX_TRAFFIC_BYTES = 172.2e6
SWEEP_TRAFFIC_BYTES = 0.5 * (150.7e6 + 132.2e6)   # pencil order, forward and backward


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
    p.add_argument("--hbm-calibration", action="store_true",
                   help="only run a device-to-device copy of known size (to scale HBM counters)")
    return p.parse_args()


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


def hbm_calibration():
    """Copy 256 MiB device to device 5 times with torch's elementwise copy
    kernel: FETCH_SIZE and WRITE_SIZE of that kernel should each read 256 MiB
    per dispatch; their ratio to it scales the counters of the solver's
    kernels.  Runs in its own process (the solver library is not loaded)."""
    import torch
    n = 256 << 20
    a = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "device copy", "bytes_read_per_dispatch": n,
                      "bytes_written_per_dispatch": n, "dispatches": 5}), flush=True)


def main():
    args = parse()
    if args.hbm_calibration:
        hbm_calibration()
        return
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
    pass_avg_s = tm.sweep_ms / max(tm.sweep_passes, 1) * 1e-3
    flux_gbs = FLUX_BYTES_PER_CELL * ncell / flux_avg_s / 1e9
    pass_bytes = 0.5 * (SWEEP_BYTES_PER_CELL["fwd"] + SWEEP_BYTES_PER_CELL["bwd"]) * ncell
    if prec == Solver.FP32:
        pass_bytes -= 0.5 * (4 + 20 + 4 + 20 + 20) * ncell  # D^-1 and x words in FP32
    sweep_gbs = pass_bytes / pass_avg_s / 1e9
    order_name = args.sweep if args.sweep != "default" else "pencil"
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
        # the dominant kernel: one LU-SGS sweep pass (forward or backward)
        "roofline": {
            "kernel": f"LU-SGS sweep pass ({order_name} order), average of forward and backward",
            "bound": "hbm",
            "achieved": sweep_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": sweep_gbs / HBM_PEAK_GBS,
            "traffic": SWEEP_TRAFFIC_BYTES if (order_name == "pencil" and args.n == 64 and
                                               prec == Solver.FP64) else None,
            "avg_launch_us": pass_avg_s * 1e6,
            "algorithmic_bytes_per_launch": pass_bytes,
            "launches_per_pass": tm.sweep_launches / max(tm.sweep_passes, 1),
            "share_of_step": tm.sweep_ms / max(tm.run_ms, 1e-30),
        },
        "roofline_flux": {
            "kernel": "KResidual (fused face flux + spectral radius + dt + diagonal)",
            "bound": "hbm",
            "achieved": flux_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": flux_gbs / HBM_PEAK_GBS,
            "traffic": FLUX_TRAFFIC_BYTES if args.n == 64 else None,
            "avg_launch_us": flux_avg_s * 1e6,
            "algorithmic_bytes_per_launch": FLUX_BYTES_PER_CELL * ncell,
            "share_of_step": tm.flux_ms / max(tm.run_ms, 1e-30),
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
