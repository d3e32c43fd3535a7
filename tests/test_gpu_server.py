"""Serving: jobs of mixed size and length submitted over time, admitted
against a memory budget, run by continuous batching; every finished job equals
the same job run alone (bit for bit), and memory bookkeeping holds."""
import time

import numpy as np
import pytest

from new_arch_amd import scmf
from new_arch_amd import workloads as W
from new_arch_amd.abi import (CANCEL_CALLER, CANCEL_TIMEOUT, JOB_CANCELLED, JOB_DONE, JOB_QUEUED,
                              JOB_RELEASED, make_opts)

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-server] {msg}", flush=True)


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32),
                          np.ascontiguousarray(b).view(np.uint32))


def jobs(n, seed=11):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        spec, _ = W.served_job(rng, n_range=(12, 40), grid_range=(12, 28))
        sweeps = int(rng.integers(3, 30))
        out.append((spec, make_opts(seed=100 + k, system_id=k, omega_every_step=bool(k % 2),
                                    density_bits=16), sweeps))
    return out


def alone(spec, opts, sweeps):
    g = scmf.System(spec, opts)
    g.grid_density()
    g.update_omega()
    for step in range(1, sweeps + 1):
        g.step(step)
    return g


def test_jobs_equal_alone_under_continuous_batching():
    banner("jobs vs alone")
    js = jobs(10)
    sv = scmf.Server(steps_per_tick=7)
    ids = []
    for k, (spec, opts, sweeps) in enumerate(js):
        ids.append(sv.submit(spec, opts, sweeps))
        if k % 3 == 2:                   # arrivals while others are running
            sv.tick()
    for _ in range(200):
        q, r = sv.tick()
        if q == 0 and r == 0:
            break
    for jid, (spec, opts, sweeps) in zip(ids, js):
        info = sv.info(jid)
        assert info.state == JOB_DONE and info.sweeps_done == sweeps
        assert info.t_done >= info.t_admit >= info.t_submit
        got = sv.job(jid)
        ref = alone(spec, opts, sweeps)
        assert same(got.positions(), ref.positions()), f"job {jid}"
        np.testing.assert_array_equal(got.density(), ref.density())
        assert same(got.omega(), ref.omega())
        assert got.stats().accepts == ref.stats().accepts
        ref.close()
        sv.release(jid)
        assert sv.info(jid).state == JOB_RELEASED
    st = sv.stats()
    banner(f"membership changes {st.rebuilds}, graph captures {st.captures}")
    assert 1 <= st.captures < st.rebuilds   # most changes only upload the graph's slot table
    sv.close()


def test_admission_against_budget():
    banner("admission")
    js = jobs(6, seed=5)
    def est(spec):                        # the server's admission estimate
        s = scmf.inspect_spec(spec)
        return (s.device_bytes_16 + 12 * s.n_beads + 16 * s.n_chains + 64 * s.n_cells
                + (1 << 20))
    need = max(est(spec) for spec, _, _ in js)
    sv = scmf.Server(hbm_budget=int(need * 2.5), steps_per_tick=5)
    ids = [sv.submit(spec, opts, sweeps) for spec, opts, sweeps in js]
    seen_queued = False
    for _ in range(3000):          # admission runs on the server's worker thread
        q, r = sv.tick()
        assert q + r <= len(js)
        seen_queued |= q > 0 and r >= 1
        if r >= 1:
            break
        time.sleep(0.001)
    assert r >= 1
    for _ in range(3000):
        q, r = sv.tick()
        seen_queued |= q > 0 and r >= 1
        for jid in ids:                   # finished jobs give their memory back
            if sv.info(jid).state == JOB_DONE:
                sv.release(jid)
        if q == 0 and r == 0:
            break
    assert all(sv.info(j).state == JOB_RELEASED for j in ids)
    assert seen_queued
    spec, opts, _ = js[0]
    tiny = scmf.Server(hbm_budget=1024)
    with pytest.raises(scmf.ScmfError):
        tiny.submit(spec, opts, 5)
    tiny.close()
    sv.close()


def test_cancel_and_timeout_leave_the_others_unchanged():
    """A running job cancelled mid-run, a job that passes its wall-clock limit
    and a queued job cancelled before admission all end CANCELLED with their
    memory returned; every other job still equals its run alone."""
    banner("cancel and timeout")
    js = jobs(6, seed=21)
    long_spec, _ = W.served_job(np.random.default_rng(4), n_range=(24, 24), grid_range=(16, 16))
    sv = scmf.Server(steps_per_tick=5)
    ids = [sv.submit(spec, opts, sweeps) for spec, opts, sweeps in js]
    victim = sv.submit(long_spec, make_opts(seed=900, system_id=90, density_bits=16), 100_000)
    late = sv.submit(long_spec, make_opts(seed=901, system_id=91, density_bits=16), 100_000)
    sv.set_timeout(late, 0.5)
    for _ in range(200):
        sv.tick()
        if sv.info(victim).state == 1 and sv.info(victim).sweeps_done >= 10:
            break
    assert sv.info(victim).sweeps_done >= 10
    sv.cancel(victim)
    for _ in range(5000):
        q, r = sv.tick()
        if q == 0 and r == 0:
            break
    assert sv.info(victim).state == JOB_CANCELLED and sv.info(victim).reason == CANCEL_CALLER
    assert sv.info(late).state == JOB_CANCELLED and sv.info(late).reason == CANCEL_TIMEOUT
    for jid, (spec, opts, sweeps) in zip(ids, js):
        assert sv.info(jid).state == JOB_DONE
        got = sv.job(jid)
        ref = alone(spec, opts, sweeps)
        assert same(got.positions(), ref.positions()), f"job {jid}"
        np.testing.assert_array_equal(got.density(), ref.density())
        assert same(got.omega(), ref.omega())
        ref.close()
        sv.release(jid)
    # a queued job cancelled before admission is never admitted: the budget
    # holds the first two jobs, so the third waits at the head of the queue
    # and the fourth behind it
    b = [sv.info(j).device_bytes for j in ids[:3]]
    tight = scmf.Server(hbm_budget=b[0] + b[1] + b[2] // 2, steps_per_tick=5)
    first = [tight.submit(spec, opts, 200) for spec, opts, _ in js[:3]]
    waiting = tight.submit(js[3][0], js[3][1], 5)
    tight.tick()
    tight.cancel(waiting)
    assert tight.info(waiting).state == JOB_CANCELLED
    done = set()
    for _ in range(5000):
        q, r = tight.tick()
        for j in first:                   # finished jobs give their memory back
            if j not in done and tight.info(j).state == JOB_DONE:
                done.add(j)
                tight.release(j)
        if q == 0 and r == 0:
            break
    assert done == set(first)
    assert tight.info(waiting).t_admit == 0.0
    assert tight.stats().cancelled == 1
    st = sv.stats()
    assert st.cancelled == 2 and st.finished == len(js)
    for _ in range(500):              # the worker frees released memory
        if sv.stats().used_bytes == 0 and tight.stats().used_bytes == 0:
            break
        sv.tick()
        tight.tick()
        time.sleep(0.01)
    assert sv.stats().used_bytes == 0 and tight.stats().used_bytes == 0
    assert 0 < st.busy_s <= st.span_s * 1.001
    tight.close()
    sv.close()
