"""The batched and served workloads at their own shapes.

* The 64-system batched mix (N in {16, 24, 32, 48, 64}, grids 16^3 to 64^3,
  diblock melts and homopolymer blends): every member of the batch equals the
  same system stepped alone, for each member's own kernel shape and for one
  shape forced on all, and one member of every chain length equals the oracle's
  trajectory (positions and counts bit for bit, omega the oracle's fp64 value
  rounded once, accept counts).
* Served jobs drawn from the full job ranges (N 16-128, grids 16^3-128^3, AB
  diblock or ABA triblock, chiN 10-50): each equals the same job run alone, and
  the longest-chain job equals the oracle.
Sweep counts are cut to a few per system so the oracle finishes in seconds.
"""
import numpy as np
import pytest

from new_arch_amd import scmf
from new_arch_amd import workloads as W
from new_arch_amd.abi import JOB_DONE, make_opts
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-mixes] {msg}", flush=True)


def same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32),
                          np.ascontiguousarray(b).view(np.uint32))


def mix_opts(seed, sid):
    return make_opts(seed=seed, system_id=sid, omega_every_step=True, density_bits=16)


def start(spec, opts):
    g = scmf.System(spec, opts)
    g.grid_density()
    g.update_omega()
    return g


def oracle_run(spec, opts, steps):
    """The test.cpp:19-43 sequence on the oracle: init, grid, omega, then
    steps of sweep, grid, omega (omega after every sweep)."""
    o = orc.Oracle(spec, opts)
    pos = o.init_coords()
    counts = o.grid(pos)
    om = o.omega(counts).astype(np.float32)
    acc = 0
    for s in range(steps):
        acc += o.sweep(pos, om, s)
        counts = o.grid(pos)
        om64 = o.omega(counts)
        om = om64.astype(np.float32)
    return pos, counts, om, acc


def check_against(g, ref, what):
    pos, counts, om, acc = ref
    assert same_bits(g.positions(), pos), f"{what}: positions"
    np.testing.assert_array_equal(g.density(), counts, err_msg=f"{what}: counts")
    assert same_bits(g.omega(), om), f"{what}: omega"
    assert g.stats().accepts == acc, f"{what}: accepts"


@pytest.mark.parametrize("lanes", [0, 8])
def test_batched_mix_members_equal_alone_and_oracle(lanes):
    steps = 3
    mix = W.batched_mix(n_systems=64, seed=0)
    Ns = sorted({spec.polymers[0].length for spec, _, _ in mix})
    grids = sorted({spec.n[0] for spec, _, _ in mix})
    banner(f"batched mix lanes={lanes}: {len(mix)} systems, N {Ns}, grids {grids}, "
           f"{sum(s.n_beads for s, _, _ in mix)} beads")
    assert Ns == [16, 24, 32, 48, 64] and max(grids) == 64
    mem = [start(spec, mix_opts(seed, sid)) for spec, seed, sid in mix]
    b = scmf.Batch(mem)
    if lanes:
        b.set_lanes(lanes)
    b.step(steps)
    b.sync()
    banner("  alone")
    for i, (spec, seed, sid) in enumerate(mix):
        a = start(spec, mix_opts(seed, sid))
        for k in range(1, steps + 1):
            a.step(k)
        m = mem[i]
        assert same_bits(a.positions(), m.positions()), f"member {i} positions"
        np.testing.assert_array_equal(a.density(), m.density(), err_msg=f"member {i} counts")
        assert same_bits(a.omega(), m.omega()), f"member {i} omega"
        assert a.stats().accepts == m.stats().accepts, f"member {i} accepts"
        a.close()
    if lanes == 0:
        for N in Ns:
            i = next(k for k, (spec, _, _) in enumerate(mix) if spec.polymers[0].length == N)
            spec, seed, sid = mix[i]
            banner(f"  oracle member {i}: N={N}, grid {spec.n[0]}^3, {spec.n_beads} beads")
            check_against(mem[i], oracle_run(spec, mix_opts(seed, sid), steps), f"member {i}")
    b.close()
    for g in mem:
        g.close()


def test_served_jobs_full_ranges_equal_alone_and_oracle():
    rng = np.random.default_rng(2)
    jobs = []
    for k in range(14):
        spec, _ = W.served_job(rng)
        jobs.append((spec, mix_opts(500 + k, k), int(rng.integers(2, 7))))
    # the ranges' ends are in the draw: the longest chains and the largest grid
    spec, _ = W.served_job(np.random.default_rng(0), n_range=(128, 128), grid_range=(128, 128))
    jobs.append((spec, mix_opts(600, 14), 3))
    spec, _ = W.served_job(np.random.default_rng(1), n_range=(16, 16), grid_range=(16, 16))
    jobs.append((spec, mix_opts(601, 15), 5))
    Ns = [s.polymers[0].length for s, _, _ in jobs]
    grids = [s.n[0] for s, _, _ in jobs]
    banner(f"served jobs: N {min(Ns)}-{max(Ns)}, grids {min(grids)}-{max(grids)}, "
           f"{sum(s.n_beads for s, _, _ in jobs)} beads")
    sv = scmf.Server(steps_per_tick=4)
    ids = []
    for k, (spec, opts, sweeps) in enumerate(jobs):
        ids.append(sv.submit(spec, opts, sweeps))
        if k % 4 == 3:
            sv.tick()
    for _ in range(100):
        q, r = sv.tick()
        if q == 0 and r == 0:
            break
    for jid, (spec, opts, sweeps) in zip(ids, jobs):
        info = sv.info(jid)
        assert info.state == JOB_DONE and info.sweeps_done == sweeps, jid
        a = start(spec, opts)
        for k in range(1, sweeps + 1):
            a.step(k)
        got = sv.job(jid)
        assert same_bits(got.positions(), a.positions()), f"job {jid} positions"
        np.testing.assert_array_equal(got.density(), a.density(), err_msg=f"job {jid}")
        assert same_bits(got.omega(), a.omega()), f"job {jid} omega"
        assert got.stats().accepts == a.stats().accepts
        a.close()
    k = int(np.argmax(Ns))
    spec, opts, sweeps = jobs[k]
    banner(f"  oracle job {ids[k]}: N={spec.polymers[0].length}, grid {spec.n[0]}^3")
    check_against(sv.job(ids[k]), oracle_run(spec, opts, sweeps), f"job {ids[k]}")
    for jid in ids:
        sv.release(jid)
    sv.close()
