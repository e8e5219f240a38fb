"""Request loop for steady solves on one GPU: continuous batching over an
``ath_server``, and the plain per-case path it is compared with.

The reference has no serving layer (it runs one case per process,
src/main.cpp:231-302); this module is the host-side scheduler around the
C ABI's server entry points (``ath_admit``, ``ath_step_batch``,
``ath_retire``):

* ``serve``: requests arrive at given times; every arrived request is
  admitted as soon as the server accepts it (a free slot and room in device
  memory; otherwise it waits in FIFO order), and all admitted cases advance
  together, ``step_iters`` nonlinear iterations per batch step.  After each
  step every finished case (converged, failed or out of iterations) is read
  back and retired, which frees its memory and slot for the queue.
* ``per_case``: the same requests one at a time, each in its own context
  (``ath_open`` / ``ath_run`` in chunks of ``step_iters`` iterations / close),
  FIFO.  Its service times are measured back to back; its latencies under an
  arrival stream follow from them by the single-server FIFO recurrence
  (start = max(arrival, previous finish)).

Time-to-converge is arrival to the end of the batch step (or chunk) after
which the case is reported finished.
"""
import collections
import math
import time

import numpy as np

RUNNING, CONVERGED, FAILED, LIMIT = 0, 1, 2, 3
STATE_NAMES = {RUNNING: "running", CONVERGED: "converged", FAILED: "failed", LIMIT: "limit"}


def poisson_arrivals(n, rate, seed):
    """n arrival times (s) of a Poisson stream of `rate` requests/s"""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n)) if n else np.zeros(0)


def drop_of(l2, normalized):
    """the residual drop of one iteration's norms: the largest normalised L2
    over the equations whose first residual carries a signal (first L2 at
    least 1e-10 of the largest; the first L2 is recovered from l2 and the
    normalised value)"""
    eps = 1.0e-30
    first = [(a + eps) / b - eps if b > 0 else 0.0 for a, b in zip(l2, normalized)]
    big = max(first)
    return max(nz for nz, f in zip(normalized, first) if f >= 1e-10 * big)


def percentile(values, q):
    """nearest-rank percentile (q in [0, 100]) of a non-empty list"""
    v = sorted(values)
    k = max(0, min(len(v) - 1, int(math.ceil(q / 100.0 * len(v))) - 1))
    return v[k]


def serve(server, requests, arrivals, step_iters=20, drop=1e-8, max_iterations=20000, clock=time.perf_counter,
          sleep=time.sleep, keep_state=None, progress=None):
    """Run the request loop.  `requests`: list of dicts with "path" and
    "cells"; `arrivals`: arrival time of each (s from the start).  Returns a
    dict of per-request records and totals.  `keep_state(index, server,
    slot)` is called for a finished request before it is retired."""
    order = sorted(range(len(requests)), key=lambda i: arrivals[i])
    pending = collections.deque(order)
    queue = collections.deque()
    slots = {}  # slot -> request index
    rec = [None] * len(requests)
    t0 = clock()
    steps = 0
    busy = 0.0
    peak_running = 0
    while pending or queue or slots:
        now = clock() - t0
        while pending and arrivals[pending[0]] <= now:
            queue.append(pending.popleft())
        while queue:
            i = queue[0]
            slot = server.admit(requests[i]["path"], drop, max_iterations)
            if slot is None:
                if not slots:
                    raise RuntimeError(f"request {i} does not fit on an empty server")
                break
            queue.popleft()
            slots[slot] = i
            rec[i] = {"arrival": float(arrivals[i]), "admitted": clock() - t0, "cells": requests[i]["cells"]}
        if not slots:
            if pending:
                sleep(max(0.0, arrivals[pending[0]] - (clock() - t0)))
            continue
        peak_running = max(peak_running, len(slots))
        ts = clock()
        server.step(step_iters)
        busy += clock() - ts
        steps += 1
        done_t = clock() - t0
        for slot in list(slots):
            info = server.info(slot)
            if info.state == RUNNING:
                continue
            i = slots.pop(slot)
            r = rec[i]
            r.update(finished=done_t, latency=done_t - r["arrival"], iterations=info.iterations,
                     state=STATE_NAMES[info.state], error=info.error, drop=info.drop)
            if keep_state is not None:
                keep_state(i, server, slot)
            server.retire(slot)
            if progress is not None:
                progress(f"request {i} {r['state']} after {r['iterations']} iterations, "
                         f"latency {r['latency']:.2f} s, {len(slots)} running, {len(queue) + len(pending)} waiting")
    makespan = clock() - t0
    return _totals(rec, makespan, busy, extra={"batch_steps": steps, "peak_running": peak_running})


def _totals(rec, makespan, busy, extra=None):
    lat = [r["latency"] for r in rec]
    upd = sum(r["cells"] * r["iterations"] for r in rec)
    out = {
        "requests": len(rec),
        "converged": sum(r["state"] == "converged" for r in rec),
        "failed": sum(r["state"] == "failed" for r in rec),
        "limit": sum(r["state"] == "limit" for r in rec),
        "p50_s": percentile(lat, 50), "p99_s": percentile(lat, 99), "max_s": max(lat),
        "cell_updates": upd, "makespan_s": makespan, "busy_s": busy,
        "cell_updates_per_s": upd / makespan, "busy_cell_updates_per_s": upd / busy if busy > 0 else None,
        "records": rec,
    }
    if extra:
        out.update(extra)
    return out


def per_case_service(solver_cls, requests, step_iters=20, drop=1e-8, max_iterations=20000, precision=1,
                     clock=time.perf_counter, progress=None):
    """Service time of each request on the plain per-case GPU path, run back
    to back: open the case, run chunks of `step_iters` iterations (graph
    replay) until the residual drop or the iteration limit, close."""
    out = []
    for req in requests:
        t = clock()
        s = solver_cls(req["path"])
        s.set_implicit_precision(precision)
        it = 0
        state = "limit"
        d = 1.0
        while it < max_iterations:
            n = s.run(it, step_iters)
            it += step_iters
            l2 = list(n.l2[:5])
            nz = list(n.normalized[:5])
            if not all(math.isfinite(v) for v in l2):
                state = "failed"
                break
            d = drop_of(l2, nz)
            if d <= drop:
                state = "converged"
                break
        s.close()
        out.append({"service_s": clock() - t, "iterations": it, "state": state, "drop": d, "cells": req["cells"]})
        if progress is not None:
            progress(f"per-case request {len(out) - 1} {state} after {it} iterations in {out[-1]['service_s']:.2f} s")
    return out


def per_case_queue(service, arrivals):
    """Latencies of the per-case path under an arrival stream: one case at a
    time, FIFO, from the measured service times."""
    order = sorted(range(len(service)), key=lambda i: arrivals[i])
    rec = [None] * len(service)
    free_at = 0.0
    for i in order:
        start = max(float(arrivals[i]), free_at)
        free_at = start + service[i]["service_s"]
        rec[i] = dict(service[i], arrival=float(arrivals[i]), finished=free_at, latency=free_at - arrivals[i])
    busy = sum(r["service_s"] for r in service)
    return _totals(rec, max(r["finished"] for r in rec), busy)
