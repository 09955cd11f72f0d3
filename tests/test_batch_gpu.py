"""Ensembles, serving, checkpoints and the autotuner on the GPU.

The state of an instance does not depend on which launch its tiles ride in, on the
tile shape, or on what else is on the GPU: every cell is computed by the same code
from the same inputs.  So an instance stepped in an ensemble or inside the server
must equal, bit for bit, the same case stepped alone; a restored checkpoint must
continue exactly as the original; the autotuner must leave the state untouched.
"""
import os

import numpy as np
import pytest

from oracle_lib import Oracle
from parity_util import compare_states, worst
from swehr_amd import _abi
from swehr_amd import solver as sv
from swehr_amd.inputs import config_c4_member, config_c5_request, read_case

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = ("depth", "uh", "vh", "topo", "vinf", "totm", "dep", "maxvel", "maxdepth")


def _golden(name):
    return read_case(os.path.join(GOLD, name, "inputs"), name=name)


def _mix(tend=90.0):
    rng = np.random.default_rng(21)
    cases = [_golden("transmissive"), _golden("closed_basin"), _golden("inflow_mix")]
    for k in range(3):
        c = config_c4_member(rng, k, tend=tend, scale_max=3.0)
        cases.append(c)
    return cases


def _same(a, b, case):
    for f in FIELDS:
        np.testing.assert_array_equal(a.field(f), b.field(f), err_msg="%s %s" % (case.name, f))
    for k in range(case.nclass):
        np.testing.assert_array_equal(a.field("ch", k + 1), b.field("ch", k + 1))
        np.testing.assert_array_equal(a.field("m", k + 1), b.field("m", k + 1))
    sa, fa = a.stage()
    sb, fb = b.stage()
    np.testing.assert_array_equal(sa, sb)
    np.testing.assert_array_equal(fa, fb)
    qa, qsa = a.discharge()
    qb, qsb = b.discharge()
    np.testing.assert_array_equal(qa, qb)
    np.testing.assert_array_equal(qsa, qsb)


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32], ids=["fp64", "fp32"])
def test_ensemble_members_equal_alone_runs(precision):
    cases = _mix()
    print("ensemble of", [(c.name, c.nx, c.ny, c.nclass) for c in cases], flush=True)
    with sv.Batch(cases, precision=precision) as b:
        left = b.advance()
        assert left == 0
        for i, c in enumerate(cases):
            m = b.member(i)
            mi = m.info()
            with sv.Solver(c, precision=precision) as s:
                s.run()
                si = s.info()
                assert si["steps"] == mi["steps"] and si["t"] == mi["t"], c.name
                _same(m, s, c)
        # per-instance dt: the members took different numbers of steps
        assert len({b.member(i).info()["steps"] for i in range(len(cases))}) > 1


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32], ids=["fp64", "fp32"])
def test_ensemble_mixed_class_counts_equal_alone_runs(precision):
    """Members with many size classes beside members with few.  The launch takes the
    variant for many classes for every member (FP64: two LDS chunks; FP32: 4 waves per
    SIMD with the own cell's classes in registers); each member equals its own run
    alone, which takes the variant its S selects.  In FP32 this also checks that the
    unbiased roundings draw the same random bits in both."""
    from parity_util import many_classes_case

    cases = [many_classes_case(10, 37, 53, seed=1), _golden("transmissive"), many_classes_case(16, 21, 70, seed=2),
             _golden("closed_basin")]
    with sv.Batch(cases, precision=precision) as b:
        b.advance(300)
        for i, c in enumerate(cases):
            with sv.Solver(c, precision=precision) as s:
                s.run(300)
                assert s.info()["steps"] == b.member(i).info()["steps"], c.name
                _same(b.member(i), s, c)


def test_batch_profile_times_each_launch_and_advances():
    """The ensemble diagnostic: positive durations of the three launches, and every
    unfinished member advanced by exactly the profiled steps."""
    cases = [_golden("transmissive"), _golden("closed_basin")]
    with sv.Batch(cases, precision=_abi.FP32) as b:
        b.advance(10)
        before = [b.member(i).info()["steps"] for i in range(len(cases))]
        ms = b.profile(5)
        assert ms.shape == (3,) and all(x > 0 for x in ms)
        after = [b.member(i).info()["steps"] for i in range(len(cases))]
        assert after == [x + 5 for x in before]


def test_ensemble_fp64_matches_oracle():
    cases = [_golden("transmissive"), _golden("closed_basin")]
    with sv.Batch(cases, precision=_abi.FP64) as b:
        b.advance()
        for i, c in enumerate(cases):
            o = Oracle(c)
            o.run()
            errs = compare_states(b.member(i).state(), o.state(), c.nclass)
            k, e = worst(errs)
            print(c.name, "worst %s %.3e" % (k, e))
            assert e <= 1e-12, (c.name, k, e)


def test_ensemble_step_limit_and_resume():
    cases = _mix()
    with sv.Batch(cases, precision=_abi.FP64) as b:
        b.advance(37)
        steps = [b.member(i).info()["steps"] for i in range(len(cases))]
        assert all(s == 37 or b.member(i).info()["done"] for i, s in enumerate(steps))
        b.advance()
        c = cases[4]
        with sv.Solver(c) as s:
            s.run()
            _same(b.member(4), s, c)


def test_server_continuous_batching_results_equal_alone():
    """Requests admitted as slots and memory free up; each result equals the same case
    run alone.  Finished requests hold their device memory until released, so the
    client checks and releases each one as it completes.  The server's work counter
    never falls, never lags the finished requests' cells x classes x steps, and equals
    their sum once all have finished."""
    rng = np.random.default_rng(31)
    reqs = [config_c5_request(rng, k, sizes=(32, 48, 64), weights=(1, 1, 1), tmin=30.0, tmax=90.0) for k in range(6)]
    sizes = [sv.instance_bytes_for(c, _abi.FP32) for c in reqs]
    budget = int(2.5 * max(sizes))
    checked = 0
    with sv.Server(mem_bytes=budget, max_slots=4, precision=_abi.FP32) as srv:
        ids = {srv.submit(c): c for c in reqs}
        max_running = 0
        done_work, last_work = 0, srv.work()
        assert last_work == 0
        for _ in range(10000):
            running, queued, finished = srv.pump(16)
            used, _ = srv.stats()
            assert used <= budget
            max_running = max(max_running, running)
            while True:
                d = srv.next_done()
                if d is None:
                    break
                rid, info = d
                c = ids[rid]
                done_work += c.nx * c.ny * c.nclass * info["steps"]
                r = srv.result(rid)
                with sv.Solver(c, precision=_abi.FP32) as s:
                    s.run()
                    assert s.info()["steps"] == info["steps"]
                    _same(r, s, c)
                srv.release(rid)
                checked += 1
            w = srv.work()
            assert last_work <= w and done_work <= w
            last_work = w
            if checked == len(reqs):
                break
        assert checked == len(reqs)
        assert srv.work() == done_work
        assert 2 <= max_running <= min(4, budget // min(sizes))  # slots and memory budget
        assert max_running < len(reqs)


def test_server_rejects_request_larger_than_budget():
    rng = np.random.default_rng(32)
    c = config_c5_request(rng, 0, sizes=(64,), weights=(1,), tmin=30, tmax=40)
    with sv.Server(mem_bytes=64 << 10, max_slots=2, precision=_abi.FP32) as srv:
        with pytest.raises(sv.SwehrError) as ei:
            srv.submit(c)
        assert ei.value.code == _abi.ERR_BUSY


def test_snapshot_restore_continues_identically():
    c = _golden("transmissive")
    with sv.Solver(c) as s:
        s.run(100)
        snap = s.snapshot()
        s.run(200)
        a = {f: s.field(f) for f in FIELDS}
        s.restore(snap)
        assert s.info()["steps"] == 100
        s.run(200)
        for f in FIELDS:
            np.testing.assert_array_equal(s.field(f), a[f])


def test_autotune_keeps_state_and_picks_a_fitting_tile():
    c = _golden("closed_basin")
    with sv.Solver(c) as s, sv.Solver(c) as ref:
        s.run(50)
        best, ms = s.autotune(5)
        assert 1 <= best <= len(ms) and ms[best - 1] > 0
        assert ms[best - 1] == min(m for m in ms if m > 0)
        s.run(150)
        ref.run(200)
        for f in FIELDS:
            np.testing.assert_array_equal(s.field(f), ref.field(f))


def test_server_sized_by_free_device_memory():
    rng = np.random.default_rng(33)
    c = config_c5_request(rng, 0, sizes=(64,), weights=(1,), tmin=30, tmax=40)
    with sv.Server(mem_bytes=0, max_slots=2, precision=_abi.FP32) as srv:
        used, budget = srv.stats()
        assert used == 0 and budget > 8 * 2 ** 30  # most of the GPU's HBM is free
        rid = srv.submit(c)
        for _ in range(1000):
            srv.pump(16)
            d = srv.next_done()
            if d is not None:
                assert d[0] == rid
                break
        srv.release(rid)


def _drain(srv, ids, reqs_by_id, max_pumps=20000):
    """Pump until every request finished; check each against its alone run and release it,
    and the server's work counter against the requests' cells x classes x steps (each
    step counted once, across evictions and re-admissions).  Returns the largest number
    of requests running at once."""
    done, max_running, work = 0, 0, 0
    for _ in range(max_pumps):
        running, queued, finished = srv.pump(16)
        max_running = max(max_running, running)
        while True:
            d = srv.next_done()
            if d is None:
                break
            rid, info = d
            c = reqs_by_id[rid]
            work += c.nx * c.ny * c.nclass * info["steps"]
            r = srv.result(rid)
            with sv.Solver(c, precision=srv.precision) as s:
                s.run()
                assert s.info()["steps"] == info["steps"], c.name
                _same(r, s, c)
            srv.release(rid)
            done += 1
        if done == len(ids):
            assert srv.work() == work
            return max_running
    raise AssertionError("requests did not finish")


def test_server_evicted_request_resumes_identically():
    """A running request checkpointed to host memory and requeued frees its device memory
    and, re-admitted, finishes exactly as if it had never stopped."""
    rng = np.random.default_rng(34)
    reqs = [config_c5_request(rng, k, sizes=(48, 64), weights=(1, 1), tmin=60.0, tmax=90.0) for k in range(2)]
    with sv.Server(mem_bytes=1 << 30, max_slots=4, precision=_abi.FP32) as srv:
        ids = {srv.submit(c): c for c in reqs}
        srv.pump(32)
        used0, _ = srv.stats()
        victim = min(ids)
        srv.evict(victim)
        used1, _ = srv.stats()
        assert used1 < used0
        assert srv.evictions() == 1
        with pytest.raises(sv.SwehrError):
            srv.evict(victim)  # no longer running
        _drain(srv, ids, ids)


def test_server_time_slicing_round_robins_requests_that_do_not_fit_together():
    """A budget that holds one request at a time and a quantum of 32 steps: the requests
    take turns (evicted and restored many times), and each result equals its alone run."""
    rng = np.random.default_rng(35)
    reqs = [config_c5_request(rng, k, sizes=(48,), weights=(1,), tmin=60.0, tmax=90.0) for k in range(3)]
    need = max(sv.instance_bytes_for(c, _abi.FP32) for c in reqs)
    with sv.Server(mem_bytes=int(1.5 * need), max_slots=4, precision=_abi.FP32) as srv:
        srv.set_quantum(32)
        ids = {srv.submit(c): c for c in reqs}
        max_running = _drain(srv, ids, ids)
        assert max_running == 1
        assert srv.evictions() >= 2



@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32], ids=["fp64", "fp32"])
def test_persistent_ensemble_equals_graph_ensemble(precision):
    """The persistent mode (one cooperative launch per 16 steps, grid barriers between
    the stages) runs the per-stage kernels' device functions in the same order: every
    member ends bit-identical to the graph-launched ensemble's (and so to its run
    alone), with the same per-instance step counts."""
    cases = _mix()
    print("persistent ensemble of", [(c.name, c.nx, c.ny, c.nclass) for c in cases], flush=True)
    with sv.Batch(cases, precision=precision, tile=1) as a, sv.Batch(cases, precision=precision, tile=1) as b:
        b.set_persistent(True)
        assert a.advance() == 0
        assert b.advance() == 0
        for i, c in enumerate(cases):
            assert a.member(i).info()["steps"] == b.member(i).info()["steps"], c.name
            _same(a.member(i), b.member(i), c)
