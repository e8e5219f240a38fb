"""The request loop's scheduling logic on the CPU, with a stand-in server
whose cases need a known number of iterations and whose batch steps advance
a simulated clock."""
import math

import numpy as np
import pytest

from aither_amd import serve


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(0.0, dt)


class Info:
    def __init__(self, state, iterations, drop):
        self.state, self.iterations, self.drop, self.error = state, iterations, drop, 0


class FakeServer:
    """cases finish after `need` iterations; at most `slots` run at once; a
    batch step costs `per_cell` seconds per running cell per iteration"""

    def __init__(self, clock, slots, per_cell=1e-6):
        self.clock, self.slots, self.per_cell = clock, slots, per_cell
        self.cases = {}
        self.next = 0
        self.admit_log = []

    def admit(self, path, drop, max_iterations):
        running = sum(1 for c in self.cases.values() if c["state"] == serve.RUNNING)
        if running >= self.slots:
            return None
        need, cells = (int(x) for x in path.split(":"))
        sid = self.next
        self.next += 1
        self.cases[sid] = dict(need=need, cells=cells, it=0, state=serve.RUNNING, max=max_iterations)
        self.admit_log.append((path, self.clock()))
        return sid

    def step(self, iters):
        run = [c for c in self.cases.values() if c["state"] == serve.RUNNING]
        self.clock.t += iters * self.per_cell * sum(c["cells"] for c in run)
        for c in run:
            c["it"] += iters
            if c["it"] >= c["need"]:
                c["state"] = serve.CONVERGED
            elif c["it"] >= c["max"]:
                c["state"] = serve.LIMIT
        return sum(1 for c in self.cases.values() if c["state"] == serve.RUNNING)

    def info(self, sid):
        c = self.cases[sid]
        return Info(c["state"], c["it"], 1e-9 if c["state"] == serve.CONVERGED else 1.0)

    def retire(self, sid):
        del self.cases[sid]


def make_requests(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        need = int(rng.integers(1, 20)) * 10
        cells = int(rng.integers(1000, 50000))
        out.append({"path": f"{need}:{cells}", "cells": cells, "need": need})
    return out


@pytest.mark.parametrize("slots", [1, 3, 64])
def test_every_request_finishes_fifo_after_arrival(slots):
    clk = FakeClock()
    srv = FakeServer(clk, slots)
    reqs = make_requests(40, 1)
    arr = serve.poisson_arrivals(40, 20.0, 2)
    out = serve.serve(srv, reqs, arr, step_iters=10, clock=clk, sleep=clk.sleep)
    assert out["requests"] == 40 and out["converged"] == 40
    for r, q in zip(out["records"], reqs):
        assert r["admitted"] >= r["arrival"] - 1e-12
        assert r["finished"] > r["admitted"]
        assert r["iterations"] == q["need"]
    # FIFO: admission order is arrival order
    adm = [p for p, _ in srv.admit_log]
    order = sorted(range(40), key=lambda i: arr[i])
    assert adm == [reqs[i]["path"] for i in order]
    assert out["peak_running"] <= slots
    assert out["cell_updates"] == sum(q["cells"] * q["need"] for q in reqs)
    assert not srv.cases


def test_one_slot_equals_fifo_queue():
    """With one slot the loop is the per-case FIFO queue with service times
    of whole steps."""
    clk = FakeClock()
    srv = FakeServer(clk, 1)
    reqs = make_requests(15, 3)
    arr = serve.poisson_arrivals(15, 2.0, 4)
    out = serve.serve(srv, reqs, arr, step_iters=10, clock=clk, sleep=clk.sleep)
    service = [{"service_s": q["need"] * q["cells"] * 1e-6, "iterations": q["need"], "state": "converged",
                "drop": 0.0, "cells": q["cells"]} for q in reqs]
    ref = serve.per_case_queue(service, arr)
    for a, b in zip(out["records"], ref["records"]):
        assert math.isclose(a["finished"], b["finished"], rel_tol=1e-9, abs_tol=1e-9)


def test_iteration_limit_is_reported():
    clk = FakeClock()
    srv = FakeServer(clk, 4)
    reqs = [{"path": "1000:100", "cells": 100}, {"path": "20:100", "cells": 100}]
    out = serve.serve(srv, reqs, [0.0, 0.0], step_iters=10, max_iterations=50, clock=clk, sleep=clk.sleep)
    assert [r["state"] for r in out["records"]] == ["limit", "converged"]
    assert out["records"][0]["iterations"] == 50 and out["limit"] == 1


def test_per_case_queue_recurrence():
    service = [{"service_s": s, "iterations": 1, "state": "converged", "drop": 0.0, "cells": 1}
               for s in (2.0, 1.0, 3.0)]
    out = serve.per_case_queue(service, [0.0, 0.5, 10.0])
    fin = [r["finished"] for r in out["records"]]
    assert fin == [2.0, 3.0, 13.0]
    assert [r["latency"] for r in out["records"]] == [2.0, 2.5, 3.0]
    assert out["busy_s"] == 6.0 and out["makespan_s"] == 13.0


def test_percentile_and_drop():
    assert serve.percentile([5, 1, 4, 2, 3], 50) == 3
    assert serve.percentile(list(range(1, 101)), 99) == 99
    assert serve.percentile([7.0], 99) == 7.0
    # an equation whose first residual is round-off does not count
    l2 = [1e-9, 2e-9, 3e-9, 1e-25, 4e-9]
    first = [1.0, 1.0, 1.0, 1e-20, 1.0]
    nz = [(a + 1e-30) / (f + 1e-30) for a, f in zip(l2, first)]
    assert math.isclose(serve.drop_of(l2, nz), 4e-9, rel_tol=1e-9)
    assert serve.drop_of([1.0] * 5, [0.5] * 5) == 0.5


def test_poisson_arrivals_rate():
    a = serve.poisson_arrivals(20000, 50.0, 7)
    assert np.all(np.diff(a) > 0)
    assert abs(len(a) / a[-1] - 50.0) < 2.0
