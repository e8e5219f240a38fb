"""Benchmark: cell-updates/s of aither's hot path on one MI355X.

A "step" is one nonlinear iteration of the whole hot path over every block:
ghost cells (boundary patches and interblock/periodic connections), the
residual (inviscid and viscous face fluxes, spectral radii, dt and the
implicit main diagonal), the implicit sweeps, the matrix residual, the update
and the norm reductions.  The state is resident in HBM before timing starts.

Workloads (--workload):
  * channel (default): batched multiblock laminar Navier-Stokes.  Eight
    point-matched blocks of uneven sizes (48x32x32 to 96x64x32, 1,032,192
    cells) forming a 3D channel: no-slip adiabatic bottom wall, wall-normal
    geometric stretching 1.1 from a first cell of 1e-5 L, spanwise periodic,
    characteristic inflow/outflow/top; Mach 0.3, Re_L = 1e6, Sutherland air
    with k/mu = c_p/0.72; AUSMPW+ with MUSCL (kappa = 1/3, van Albada),
    central viscous flux; BLU-SGS with 5x5 blocks; BDF2 dual time stepping,
    20 nonlinear iterations per physical step, 50 physical steps (1000
    steps here); uniform free stream plus seeded 1e-3 random velocity
    perturbations.
  * vortex: inviscid Euler, one Cartesian block 64^3 on [0,10]^3, isentropic
    vortex on a Mach 0.5 free stream, characteristic farfield, Roe, MUSCL,
    implicit Euler LU-SGS at CFL 5, 200 iterations (the CPU-runnable case).

Prints one JSON line.  Run `python bench.py --help` for options.
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per physical cell (FP64; each array element once, halos
# and re-reads excluded).
# vortex (inviscid, scalar LU-SGS, implicit Euler):
#   residual kernel: W 40 + geometry 128 + R 40 + lambda 8 + diagonal 8 = 224
#   forward pass: W 40 + R 40 + U^n 40 + geometry 96 + V 8 + dt 8 + D^-1 8 +
#   x 40 = 280; backward pass: W 40 + geometry 96 + D^-1 8 + x read 40 + x
#   write 40 = 224
VORTEX_FLUX_BYTES = 224
VORTEX_SWEEP_BYTES = {"fwd": 280, "bwd": 224}
# HBM bytes per launch on the 64^3 vortex (FP64), from rocprofv3 FETCH_SIZE /
# WRITE_SIZE scaled by the same counters on a 256 MiB device copy (factor 2.0
# for FETCH_SIZE, 1.0 for WRITE_SIZE); see profiles/euler-vortex-64_hbm_counters.md
VORTEX_FLUX_TRAFFIC = 85.4e6
VORTEX_SWEEP_TRAFFIC = 0.5 * (263.6e6 + 257.8e6)
# channel (laminar, block LU-SGS, BDF2):
#   residual stage (inviscid and viscous face fluxes, spectral radii, dt,
#   block diagonal and inverse): W 40 + geometry 128 + R 40 + lambda 8 + dt 8
#   + T and mu 16 + D 200 + D^-1 200 + cell velocity gradient 72 = 712
#   forward pass (the off-diagonal blocks formed from the neighbours' state,
#   velocity gradient, viscosity and conductivity): W 40 + R 40 + U^n 40 +
#   U^n-1 40 + geometry 96 + V 8 + dt 8 + D^-1 200 + mu 8 + k 8 + velocity
#   gradient 72 + centre distances 24 + x 40 = 624
#   backward pass (first sweep, no initialisation: x - D^-1 U): W 40 +
#   geometry 96 + D^-1 200 + mu 8 + k 8 + velocity gradient 72 + centre
#   distances 24 + x read 40 + x write 40 = 528
#   the row-parallel sweep reads the right-hand side formed once per relax
#   (rhs 40 in place of R, U^n, U^n-1, V and dt): 528 forward and backward
CHANNEL_FLUX_BYTES = 712
# with the row-parallel sweep (FP64 operator, one sweep) the viscous block
# diagonal also forms the relax's right-hand side and du = 0: U^n 40 +
# U^n-1 40 read, rhs 40 + x 40 written (W, R, V and dt are counted above)
CHANNEL_FLUX_RHS_BYTES = 160
CHANNEL_SWEEP_BYTES = {"fwd": 624, "bwd": 528}
CHANNEL_ROWS_SWEEP_BYTES = {"fwd": 528, "bwd": 528}
# HBM bytes per launch on the full-size channel (refine 1), from rocprofv3
# FETCH_SIZE / WRITE_SIZE scaled by the device-copy calibration (2.0, 1.0);
# see profiles/channel-multiblock-1M_hbm_counters.md
CHANNEL_SWEEP_TRAFFIC = 0.5 * (1338.3e6 + 1119.8e6)   # one pass, forward and backward averaged
CHANNEL_ROWS_SWEEP_TRAFFIC = 0.5 * (1177.5e6 + 1123.9e6)  # the row-parallel sweep
CHANNEL_FLUX_TRAFFIC = 2827.7e6                       # residual stage over all blocks
# what the occupancy and VALU counters show binds each (rocprofv3 SQ_WAVES,
# SQ_INSTS_VALU, VALUBusy, OccupancyPercent on the same case; see
# profiles/channel-multiblock-1M_sq_counters.md)
ROWS_SWEEP_LIMITER = ("dependency latency: one workgroup of 8 waves per CU (22 % occupancy), VALU 20 % busy, HBM about "
                      "28 %; a tile's forming is a chain of memory round trips, and the tile hyperplanes of the largest "
                      "block set the ramps")
SWEEP_LIMITER = ("dependency latency: one workgroup per CU (10.6 % occupancy), VALU 17 % busy, HBM about 20 %; "
                 "the pass is the chain of tile hyperplanes of the largest block")
FLUX_LIMITER = ("mixed: KResidual FP64-VALU issue (VALUBusy 80 %), the viscous block diagonal with the inviscid "
                "block sum and the relax's right-hand side VALU (52 %) and HBM (45 %), the viscous flux the latency "
                "of its gathers (occupancy 22 %, VALU 27 %, HBM 40 %)")


def parse():
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--workload", choices=["channel", "vortex", "serve"], default="channel")
    p.add_argument("--steps", type=int, default=None,
                   help="timed nonlinear iterations (channel: a multiple of 20; default 1000, vortex 200)")
    p.add_argument("--warmup", type=int, default=None, help="untimed nonlinear iterations (default 20 / 10)")
    p.add_argument("--refine", type=float, default=1.0, help="channel: scale every cell count")
    p.add_argument("--n", type=int, default=64, help="vortex: cells per direction")
    p.add_argument("--precision", choices=["fp64", "fp32"], default="fp64",
                   help="vortex: implicit-operator precision (residual is always FP64)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-iters", type=int, default=None,
                   help="oracle iterations timed for cpu_baseline (default 3 channel, 20 vortex)")
    p.add_argument("--no-fp32-leg", action="store_true")
    p.add_argument("--sweep", choices=["default", "hyperplane", "tiled", "pencil", "rows"], default="default",
                   help="scalar LU-SGS execution order (all give identical results)")
    p.add_argument("--requests", type=int, default=8, help="serve: requests (the full stream is 256)")
    p.add_argument("--nmin", type=int, default=24, help="serve: smallest cells per direction")
    p.add_argument("--nmax", type=int, default=40, help="serve: largest cells per direction (full: 96)")
    p.add_argument("--seed", type=int, default=1, help="serve: request and arrival seed")
    p.add_argument("--serve-cfl", type=float, default=20.0)
    p.add_argument("--serve-step", type=int, default=20, help="serve: iterations per batch step / chunk")
    p.add_argument("--serve-load", type=float, default=1.5,
                   help="serve: offered load as a fraction of the per-case path's capacity")
    p.add_argument("--max-iterations", type=int, default=20000, help="serve: iteration cap per request")
    p.add_argument("--max-slots", type=int, default=256, help="serve: cases running at once")
    p.add_argument("--no-serve", action="store_true", help="channel: leave out the serve block")
    p.add_argument("--serve-json", default=None, help="serve: write per-request records to this file")
    p.add_argument("--hbm-calibration", action="store_true",
                   help="only run a device-to-device copy of known size (to scale HBM counters)")
    return p.parse_args()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(inp, iters, what):
    """The oracle (FP64 C restatement) on the first `iters` nonlinear
    iterations of the same case, on every host core: its OpenMP build, whose
    results equal the single-threaded build's bit for bit.  Returns the
    baseline record and the oracle's residual after each iteration and final
    state, which the parity check compares with the GPU."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    c = pyoracle.OracleCase(inp, parallel=True)
    ncell = sum(c.ncells(b) for b in range(c.nb))
    t0 = time.perf_counter()
    done = 0
    nn = 0
    resid = []
    tres = 0.0
    while done < iters:
        c.begin_step(nn)
        for mm in range(c.nonlinear):
            if done == iters:
                break
            c.iterate(nn, mm)
            t1 = time.perf_counter()
            resid.append(np.concatenate([c.residual(b).reshape(-1) for b in range(c.nb)]))
            tres += time.perf_counter() - t1
            done += 1
        nn += 1
    dt = time.perf_counter() - t0 - tres
    state = np.concatenate([c.state(b).reshape(-1) for b in range(c.nb)])
    rec = {"value": ncell * iters / dt, "unit": "cell-updates/s", "cores": c.threads, "kind": "port",
           "cpu": cpu_model(),
           "sample": f"first {iters} nonlinear iterations of the same {what} ({ncell} cells), oracle C "
                     f"restatement, OpenMP on {c.threads} threads ({dt:.1f} s)"}
    return rec, resid, state


def gpu_parity(s, states0, resid_ref, state_ref):
    """The GPU path over the same first iterations from the same initial
    state as the oracle leg: per-iteration residual and final-state relative
    L2 differences (the metric's bound is 1e-12 on the residual and 1e-10 on
    the state)."""
    import numpy as np
    nb = len(states0)
    for b, W in enumerate(states0):
        s.set_state(W, b)
    nl = s.info.nonlinear_iterations
    rr = []
    it = 0
    nn = 0
    while it < len(resid_ref):
        s.begin_step(nn)
        for mm in range(nl):
            if it == len(resid_ref):
                break
            s.iterate(nn, mm)
            R = np.concatenate([s.residual(b).reshape(-1) for b in range(nb)])
            rr.append(float(np.linalg.norm(R - resid_ref[it]) / np.linalg.norm(resid_ref[it])))
            it += 1
        nn += 1
    W = np.concatenate([s.get_state(b).reshape(-1) for b in range(nb)])
    sd = float(np.linalg.norm(W - state_ref) / np.linalg.norm(state_ref))
    ok = max(rr) <= 1e-12 and sd <= 1e-10
    return {"iterations": len(rr), "residual_rel_l2": rr, "state_rel_l2": sd,
            "bound": "residual <= 1e-12, state <= 1e-10", "ok": ok}


def warm_up(s, warm):
    """`warm` untimed nonlinear iterations: whole physical steps replayed
    from the captured graph, the rest launched eagerly; the graph is captured
    first, so no capture or instantiation falls in the timed region."""
    s.run(0, 0)
    sub = s.info.nonlinear_iterations
    nphys, rest = divmod(warm, sub)
    if nphys:
        s.run(0, nphys)
    if rest:
        s.begin_step(nphys)
        for mm in range(rest):
            s.iterate(nphys, mm)


def run_iterations(s, steps):
    """exactly `steps` nonlinear iterations from outer step 0: whole physical
    steps from the graph, a partial last step eagerly"""
    sub = s.info.nonlinear_iterations
    nphys, rest = divmod(steps, sub)
    last = s.run(0, nphys) if nphys else None
    if rest:
        s.begin_step(nphys)
        for mm in range(rest):
            last = s.iterate(nphys, mm)
    return last


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


def roofline(name, nbytes, avg_s, share, traffic=None, **extra):
    gbs = nbytes / avg_s / 1e9
    out = {"kernel": name, "bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": gbs / HBM_PEAK_GBS, "traffic": traffic, "avg_launch_us": avg_s * 1e6,
           "algorithmic_bytes_per_launch": nbytes, "share_of_step": share}
    out.update(extra)
    return out


def timed_run(s, states0, steps, warm, prof_phys):
    """`warm` untimed iterations (graph capture, caches), restart from the
    initial condition, then exactly `steps` nonlinear iterations (whole
    physical steps replaying the captured graph), timed by the wall clock with
    the device synchronised on both sides (ath_run and ath_iterate return
    after a stream synchronisation); then, from the initial condition again,
    `prof_phys` physical steps launched eagerly with a HIP event pair around
    every residual stage and sweep pass (the per-kernel breakdown)"""
    warm_up(s, warm)
    for b, W in enumerate(states0):
        s.set_state(W, b)
    s.enable_timing(False)
    t0 = time.perf_counter()
    last = run_iterations(s, steps)
    wall = time.perf_counter() - t0
    for b, W in enumerate(states0):
        s.set_state(W, b)
    s.enable_timing(True)
    s.run(0, prof_phys)
    tm = s.timing()
    s.enable_timing(False)
    return last, wall, tm


def bench_channel(args, Solver, cases):
    steps = args.steps if args.steps is not None else 1000
    warm = args.warmup if args.warmup is not None else 20
    work = tempfile.mkdtemp(prefix="aither_bench_")
    inp, meta = cases.channel_case(work, refine=args.refine, steps=max(1, -(-steps // 20)), subits=20)
    sub = meta["subits"]
    nphys = -(-steps // sub)
    order = None if args.sweep == "default" else getattr(Solver, args.sweep.upper())
    s = Solver(inp)
    if order is not None:
        s.set_sweep_ordering(order)
    nb = s.info.num_blocks
    states0 = [s.get_state(b) for b in range(nb)]
    ncell = sum(s.ncells(b) for b in range(nb))
    last, wall, tm = timed_run(s, states0, steps, warm, prof_phys=min(nphys, 2))
    value = ncell * steps / wall
    # one residual stage (every kernel launched once over all blocks) and two
    # passes (a forward and a backward sweep over all blocks) per iteration
    flux_avg_s = tm.flux_ms / max(tm.flux_launches, 1) * 1e-3
    pass_avg_s = tm.sweep_ms / max(tm.sweep_passes, 1) * 1e-3
    flux_share = tm.flux_ms / max(tm.run_ms, 1e-30)
    sweep_share = tm.sweep_ms / max(tm.run_ms, 1e-30)
    rows = s.sweep_ordering() == Solver.ROWS
    per_cell = CHANNEL_ROWS_SWEEP_BYTES if rows else CHANNEL_SWEEP_BYTES
    pass_bytes = 0.5 * (per_cell["fwd"] + per_cell["bwd"]) * ncell
    full = args.refine == 1.0
    kname = "KBlockSweepRows<double, FWD>, two lanes per cell" if rows else "KBlockSweepTiles<double, FWD>"
    sweep = roofline(f"block LU-SGS sweep pass over all eight blocks ({kname}, one launch), average of forward "
                     "and backward", pass_bytes, pass_avg_s, sweep_share,
                     traffic=(CHANNEL_ROWS_SWEEP_TRAFFIC if rows else CHANNEL_SWEEP_TRAFFIC) if full else None,
                     launches_per_pass=tm.sweep_launches / max(tm.sweep_passes, 1),
                     limiter=ROWS_SWEEP_LIMITER if rows else SWEEP_LIMITER)
    rhs_in_stage = rows  # the FP64 operator (this run) with the row-parallel sweep
    flux = roofline("residual stage over all eight blocks (inviscid + viscous fluxes, spectral radii, dt, block "
                    "diagonal and inverse" + (", the relax's right-hand side" if rhs_in_stage else "") +
                    "; one launch per kernel)",
                    (CHANNEL_FLUX_BYTES + (CHANNEL_FLUX_RHS_BYTES if rhs_in_stage else 0)) * ncell, flux_avg_s,
                    flux_share, traffic=CHANNEL_FLUX_TRAFFIC if full and rhs_in_stage else None,
                    limiter=FLUX_LIMITER)
    dominant, other = (sweep, flux) if sweep_share >= flux_share else (flux, sweep)
    out = {
        "metric": "cell-updates/s at <=1e-10 err vs FP64 oracle",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": wall / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (generated stretched channel grid, uniform free stream plus seeded 1e-3 random "
                "velocity perturbations, seed 0)",
        "config": {
            "workload": "multiblock laminar channel: 8 blocks 48x32x32..96x64x32, no-slip wall, stretching 1.1 "
                        "from 1e-5 L, spanwise periodic, Mach 0.3, Re_L 1e6, Sutherland air (k/mu = cp/0.72), "
                        "AUSMPW+ MUSCL, central viscous flux, BLU-SGS 5x5, BDF2 20 subiterations per step",
            "cells": ncell, "blocks": nb, "physical_steps": nphys, "subiterations": sub,
            "parallelism": "single GPU",
        },
        "roofline": dominant,
        "roofline_other": other,
        "final_normalized_residual": list(last.normalized[:5]),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"], resid_ref, state_ref = cpu_baseline(inp, args.cpu_iters or 3, "multiblock channel")
        out["parity"] = gpu_parity(s, states0, resid_ref, state_ref)
    s.close()
    if not args.no_fp32_leg:
        # the same timed run with the FP32 implicit operator (residual, dt
        # and spectral radius stay FP64), and its converged-state error
        s32 = Solver(inp)
        s32.set_implicit_precision(Solver.FP32)
        if order is not None:
            s32.set_sweep_ordering(order)
        _, w32, tm32 = timed_run(s32, states0, steps, warm, prof_phys=min(nphys, 2))
        W32 = np.concatenate([s32.get_state(b).reshape(-1) for b in range(nb)])
        s32.close()
        s64 = Solver(inp)
        run_iterations(s64, steps)
        W64 = np.concatenate([s64.get_state(b).reshape(-1) for b in range(nb)])
        s64.close()
        out["fp32_operator"] = {
            "value": ncell * steps / w32, "ms_per_step": w32 / steps * 1e3,
            "sweep_pass_ms": tm32.sweep_ms / max(tm32.sweep_passes, 1),
            "state_rel_l2_vs_fp64_after_timed_steps": float(np.linalg.norm(W32 - W64) / np.linalg.norm(W64)),
            "converged": fp32_converged(Solver, cases, work),
        }
    return out


def fp32_converged(Solver, cases, work, drop=1e-11):
    """The FP32 implicit operator's converged-state error and extra
    iterations: a steady 10^3 bump channel with BLU-SGS (5x5 blocks) run to
    a `drop` residual drop with each operator precision and in the oracle
    (its OpenMP build)."""
    from aither_amd.serve import drop_of
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    path, meta = cases.bump_case(os.path.join(work, "fp32conv"), seed=5, n=10, iterations=1, cfl=20.0,
                                 limiter="minmod")
    text = open(path).read().replace("matrixSolver: lusgs", "matrixSolver: blusgs")
    open(path, "w").write(text)

    def gpu(prec):
        s = Solver(path)
        s.set_implicit_precision(prec)
        it = 0
        while it < 20000:
            n = s.run(it, 10)
            it += 10
            if drop_of(list(n.l2[:5]), list(n.normalized[:5])) <= drop:
                break
        W = s.get_state().reshape(-1)
        s.close()
        return W, it

    o = pyoracle.OracleCase(path, parallel=True)
    it = 0
    while it < 20000:
        o.begin_step(it)
        r = o.iterate(it, 0)
        it += 1
        if it % 10 == 0 and drop_of(list(r["l2"]), list(r["normalized"])) <= drop:
            break
    Wo = o.state().reshape(-1)
    W64, i64 = gpu(Solver.FP64)
    W32, i32 = gpu(Solver.FP32)
    err = lambda W: float(np.linalg.norm(W - Wo) / np.linalg.norm(Wo))
    return {"case": f"steady 10^3 bump channel (Mach {meta['mach']:.3f}, h/l {meta['h_over_l']:.3f}), Roe + "
                    f"MUSCL minmod, BLU-SGS, CFL 20, to a {drop:g} residual drop",
            "iterations_oracle": it, "iterations_fp64": i64, "iterations_fp32": i32,
            "extra_iterations_fp32": i32 - i64,
            "converged_state_rel_l2_fp64_vs_oracle": err(W64),
            "converged_state_rel_l2_fp32_vs_oracle": err(W32)}


def serve_requests(cases, work, n, nmin, nmax, seed, cfl):
    """n seeded bump-channel requests (sizes uniform in nmin^3..nmax^3, bump
    height/length uniform in [0.02, 0.10], Mach uniform in [0.2, 0.8]; Roe,
    MUSCL kappa = 1/3 with the minmod limiter, LU-SGS, constant CFL)"""
    reqs = []
    rng = np.random.default_rng(seed)
    for q in range(n):
        size = int(rng.integers(nmin, nmax + 1))
        path, meta = cases.bump_case(os.path.join(work, f"r{q}"), seed=seed * 100000 + q, n=size,
                                     iterations=1, cfl=cfl, limiter="minmod")
        reqs.append({"path": path, "cells": size ** 3, "n": size, "mach": meta["mach"],
                     "h_over_l": meta["h_over_l"]})
    return reqs


def bench_serve(args, Solver, Server, cases_mod):
    """C4: steady Euler bump channels as a Poisson request stream, served by
    continuous batching, with the plain per-case GPU path beside it."""
    from aither_amd import serve
    work = tempfile.mkdtemp(prefix="aither_serve_")

    def progress(msg):
        print(f"[serve] {msg}", file=sys.stderr, flush=True)

    reqs = serve_requests(cases_mod, work, args.requests, args.nmin, args.nmax, args.seed, args.serve_cfl)
    drop = 1e-8
    step = args.serve_step
    # plain per-case path, back to back: service times
    t0 = time.perf_counter()
    service = serve.per_case_service(Solver, reqs, step_iters=step, drop=drop, max_iterations=args.max_iterations,
                                     precision=Solver.FP32, progress=progress)
    plain_wall = time.perf_counter() - t0
    total_service = sum(r["service_s"] for r in service)
    rate = args.serve_load * len(reqs) / total_service
    arrivals = serve.poisson_arrivals(len(reqs), rate, args.seed + 1)
    plain = serve.per_case_queue(service, arrivals)
    srv = Server(max_slots=args.max_slots, precision=Solver.FP32)
    batched = serve.serve(srv, reqs, arrivals, step_iters=step, drop=drop, max_iterations=args.max_iterations,
                          progress=progress)
    srv.close()
    # the same result alone and in the mix: a served case's iteration count
    # equals its per-case count (same arithmetic, same stopping rule)
    same_iters = sum(a["iterations"] == b["iterations"] for a, b in zip(batched["records"], service))

    def summary(o):
        return {k: o[k] for k in ("requests", "converged", "failed", "limit", "p50_s", "p99_s", "max_s",
                                  "cell_updates", "makespan_s", "busy_s", "cell_updates_per_s",
                                  "busy_cell_updates_per_s")}

    out = {
        "workload": f"{len(reqs)} steady Euler bump channels, {args.nmin}^3..{args.nmax}^3 cells, Mach 0.2-0.8, "
                    f"bump h/l 0.02-0.10, Roe + MUSCL (minmod), LU-SGS FP32 operator, CFL {args.serve_cfl:g}, "
                    f"to a 1e-8 residual drop (cap {args.max_iterations} iterations), Poisson arrivals at "
                    f"{rate:.3g}/s (offered load {args.serve_load:g} of the per-case path)",
        "step_iterations": step,
        "serve": summary(batched),
        "per_case": summary(plain),
        "per_case_back_to_back_s": plain_wall,
        "iterations_equal_per_case": same_iters,
        "batch_steps": batched["batch_steps"], "peak_running": batched["peak_running"],
    }
    return out, batched, plain


def serve_line(args, Solver, Server, cases):
    """the serve workload as the headline: aggregate cell-updates/s of the
    batched server over the makespan of the request stream"""
    block, batched, plain = bench_serve(args, Solver, Server, cases)
    out = {
        "metric": "cell-updates/s at <=1e-10 err vs FP64 oracle; serve p50/p99 time-to-converge",
        "value": batched["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": batched["batch_steps"] * args.serve_step,
        "warmup": 0,
        "ms_per_step": batched["makespan_s"] / max(1, batched["batch_steps"] * args.serve_step) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64-residual/f32-implicit",
        "data": "synthetic (seeded bump channels and Poisson arrivals)",
        "config": {"workload": block["workload"], "parallelism": "single GPU"},
        "serve": block,
    }
    if args.serve_json:
        with open(args.serve_json, "w") as f:
            json.dump({"serve": batched["records"], "per_case": plain["records"]}, f)
    return out


def bench_vortex(args, Solver, cases):
    steps = args.steps if args.steps is not None else 200
    warm = args.warmup if args.warmup is not None else 10
    work = tempfile.mkdtemp(prefix="aither_bench_")
    inp = cases.vortex_case(work, n=args.n, iterations=steps)
    s = Solver(inp)
    prec = Solver.FP32 if args.precision == "fp32" else Solver.FP64
    s.set_implicit_precision(prec)
    order = None if args.sweep == "default" else getattr(Solver, args.sweep.upper())
    if order is not None:
        s.set_sweep_ordering(order)
    W0 = s.get_state()
    ncell = s.ncells(0)
    last, wall, tm = timed_run(s, [W0], steps, warm, prof_phys=min(steps, 50))
    cpu = None
    if not args.no_cpu_baseline:
        cpu, resid_ref, state_ref = cpu_baseline(inp, args.cpu_iters or 20, f"{args.n}^3 vortex")
        parity = gpu_parity(s, [W0], resid_ref, state_ref)
    value = ncell * steps / wall
    flux_avg_s = tm.flux_ms / max(tm.flux_launches, 1) * 1e-3
    pass_avg_s = tm.sweep_ms / max(tm.sweep_passes, 1) * 1e-3
    pass_bytes = 0.5 * (VORTEX_SWEEP_BYTES["fwd"] + VORTEX_SWEEP_BYTES["bwd"]) * ncell
    if prec == Solver.FP32:
        pass_bytes -= 0.5 * (4 + 20 + 4 + 20 + 20) * ncell  # D^-1 and x words in FP32
    order_name = args.sweep if args.sweep != "default" else "pencil"
    out = {
        "metric": "cell-updates/s at <=1e-10 err vs FP64 oracle",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": wall / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == Solver.FP64 else "f64-residual/f32-implicit",
        "data": "synthetic (seed-free analytic isentropic vortex)",
        "config": {
            "workload": f"euler-vortex-{args.n}^3: 1 block {args.n}^3 cells, Roe, MUSCL kappa=1/3 "
                        "van Albada, implicit Euler LU-SGS CFL 5, characteristic farfield",
            "cells": ncell, "iterations": steps, "parallelism": "single GPU",
        },
        "roofline": roofline(
            f"LU-SGS sweep pass ({order_name} order), average of forward and backward", pass_bytes, pass_avg_s,
            tm.sweep_ms / max(tm.run_ms, 1e-30),
            traffic=VORTEX_SWEEP_TRAFFIC if (order_name == "pencil" and args.n == 64 and prec == Solver.FP64)
            else None, launches_per_pass=tm.sweep_launches / max(tm.sweep_passes, 1)),
        "roofline_other": roofline(
            "KResidual (fused face flux + spectral radius + dt + diagonal)", VORTEX_FLUX_BYTES * ncell,
            flux_avg_s, tm.flux_ms / max(tm.run_ms, 1e-30),
            traffic=VORTEX_FLUX_TRAFFIC if args.n == 64 else None),
        "final_normalized_residual": list(last.normalized[:5]),
    }
    s.close()
    if not args.no_fp32_leg and prec == Solver.FP64:
        # same run with the FP32 implicit operator: throughput and the
        # difference of the final state from the FP64-operator run
        import numpy as np
        s64 = Solver(inp)
        if order is not None:
            s64.set_sweep_ordering(order)
        s64.run(0, steps)
        ref = s64.get_state()
        s64.close()
        s32 = Solver(inp)
        s32.set_implicit_precision(Solver.FP32)
        if order is not None:
            s32.set_sweep_ordering(order)
        _, w32, _ = timed_run(s32, [W0], steps, min(warm, 5), prof_phys=1)
        st = s32.get_state()
        s32.close()
        out["fp32_operator"] = {
            "value": ncell * steps / w32,
            "state_rel_l2_vs_fp64": float(np.linalg.norm(st - ref) / np.linalg.norm(ref)),
        }
    if cpu is not None:
        out["cpu_baseline"] = cpu
        out["parity"] = parity
    return out


def main():
    args = parse()
    if args.hbm_calibration:
        hbm_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)
    from aither_amd import Server, Solver, cases
    if args.workload == "serve":
        out = serve_line(args, Solver, Server, cases)
    elif args.workload == "channel":
        out = bench_channel(args, Solver, cases)
        if not args.no_serve:
            out["serve"], _, _ = bench_serve(args, Solver, Server, cases)
    else:
        out = bench_vortex(args, Solver, cases)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
