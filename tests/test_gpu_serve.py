"""The server (continuous batching of steady solves) on the GPU, through the
C ABI.

* Each case's result is the same alone and in a mix of cases admitted at
  different times, bit for bit (FP32 implicit operator, as served).
* With the FP64 operator every case of a mix matches the oracle over its
  first iterations (state within 1e-12 relative).
* A case driven to a nonphysical state fails alone (ATH_ERR_NONPHYSICAL for
  that slot) while the others run on and end as they would alone.
* Admission follows the memory budget: a case that does not fit waits and
  is admitted once a retire frees room.
* A small case converges to a 1e-8 residual drop.
"""
import os
import sys

import numpy as np
import pytest

from aither_amd import Server, Solver, cases
import pyoracle

pytestmark = pytest.mark.gpu

SIZES = [6, 9, 7, 12, 8, 10, 5, 11]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def bump(tmp, seed, n, cfl=10.0):
    path, meta = cases.bump_case(str(tmp), seed=seed, n=n, iterations=10, cfl=cfl, limiter="minmod")
    return path, meta


def run_alone(path, iters, step, precision=1):
    srv = Server(max_slots=4, precision=precision)
    slot = srv.admit(path, 1e-30, iters)
    while srv.step(step):
        pass
    info = srv.info(slot)
    W = srv.state(slot)
    srv.close()
    return W, info


def test_serve_alone_equals_mix(tmp_path):
    paths = [bump(tmp_path / f"c{q}", 100 + q, n)[0] for q, n in enumerate(SIZES)]
    iters, step = 15, 5
    say("serve: 8 cases alone")
    alone = [run_alone(p, iters, step) for p in paths]
    say("serve: 8 cases in a mix, admitted in two waves")
    srv = Server(max_slots=16, precision=1)
    slots = [srv.admit(p, 1e-30, iters) for p in paths[:4]]
    assert srv.step(step) == 4
    slots += [srv.admit(p, 1e-30, iters) for p in paths[4:]]
    while srv.step(step):
        pass
    for q, slot in enumerate(slots):
        info = srv.info(slot)
        assert info.state == Server.LIMIT and info.iterations == iters, (q, info.state, info.iterations)
        W = srv.state(slot)
        assert np.array_equal(W, alone[q][0]), (q, rel(W, alone[q][0]))
        assert info.last.l2_sq[:5] == alone[q][1].last.l2_sq[:5]
    srv.close()


def test_serve_mix_matches_oracle_fp64(tmp_path):
    paths = [bump(tmp_path / f"c{q}", 200 + q, n)[0] for q, n in enumerate(SIZES[:4])]
    srv = Server(max_slots=8, precision=0)
    slots = [srv.admit(p, 1e-30, 3) for p in paths]
    say("serve: FP64 mix of 4 against the oracle")
    assert srv.step(3) == 0
    for q, (p, slot) in enumerate(zip(paths, slots)):
        o = pyoracle.OracleCase(p)
        for nn in range(3):
            o.begin_step(nn)
            no = o.iterate(nn, 0)
        d = rel(srv.state(slot), o.state())
        say(f"  case {q}: state rel diff {d:.3e}")
        assert d <= 1e-12, (q, d)
        info = srv.info(slot)
        assert abs(info.last.l2[0] - no["l2"][0]) <= 1e-12 * no["l2"][0]
    srv.close()


def test_serve_quarantines_a_failing_case(tmp_path):
    paths = [bump(tmp_path / f"c{q}", 300 + q, n)[0] for q, n in enumerate(SIZES[:4])]
    iters, step = 10, 5
    alone = [run_alone(p, iters, step) for p in paths]
    srv = Server(max_slots=8, precision=1)
    slots = [srv.admit(p, 1e-30, iters) for p in paths]
    bad = slots[2]
    W = srv.state(bad)
    W[len(W) // 2, 4] = -1.0  # a negative pressure in one cell
    srv.set_state(bad, W)
    say("serve: one case with a negative pressure among 4")
    running = srv.step(step)
    assert running == 3
    info = srv.info(bad)
    assert info.state == Server.FAILED and info.error == -5, (info.state, info.error)
    while srv.step(step):
        pass
    for q, slot in enumerate(slots):
        if slot == bad:
            continue
        info = srv.info(slot)
        assert info.state == Server.LIMIT and info.iterations == iters
        assert np.array_equal(srv.state(slot), alone[q][0]), q
    srv.retire(bad)
    srv.close()


def test_serve_admission_by_memory_budget(tmp_path):
    paths = [bump(tmp_path / f"c{q}", 400 + q, 10)[0] for q in range(3)]
    est = Server.estimate(paths[0], 1)
    # room for two such cases, not three
    srv = Server(hbm_budget=2.5 * est, max_slots=8, precision=1)
    a = srv.admit(paths[0], 1e-30, 5)
    b = srv.admit(paths[1], 1e-30, 5)
    assert a is not None and b is not None
    assert srv.admit(paths[2], 1e-30, 5) is None
    used, free = srv.memory()
    assert used == 2 * est and free > 0
    # the estimate bounds what a case holds
    assert 0 < srv.info(a).device_bytes <= est
    srv.step(5)
    srv.retire(a)
    c = srv.admit(paths[2], 1e-30, 5)
    assert c is not None
    while srv.step(5):
        pass
    assert srv.info(c).iterations == 5
    srv.close()


def test_serve_slot_limit(tmp_path):
    paths = [bump(tmp_path / f"c{q}", 500 + q, 6)[0] for q in range(3)]
    srv = Server(max_slots=2, precision=1)
    assert srv.admit(paths[0], 1e-30, 5) is not None
    assert srv.admit(paths[1], 1e-30, 5) is not None
    assert srv.admit(paths[2], 1e-30, 5) is None
    srv.close()


def test_serve_converges_small_case(tmp_path):
    path, meta = bump(tmp_path / "c", 600, 8, cfl=20.0)
    srv = Server(max_slots=2, precision=1)
    slot = srv.admit(path, 1e-8, 20000)
    say("serve: one 8^3 case to a 1e-8 residual drop", meta)
    while srv.step(50):
        pass
    info = srv.info(slot)
    say(f"  state {info.state} after {info.iterations} iterations, drop {info.drop:.3e}")
    assert info.state == Server.CONVERGED and info.drop <= 1e-8
    srv.close()
