"""Throughput benchmark of the SCMF Monte Carlo step on one MI355X.

A step is one pass of the hot path over the whole system: a smart-MC sweep
(every chain makes N sequential moves), NGP density gridding and the omega
update.  The workload is the single large system: AB diblock melt, N=64,
fA=0.3, 4,194,304 chains (268,435,456 beads), 256^3 grid, chiN=25, kappaN=70,
uint16 counts, fp32 coordinates; synthetic Gaussian-walk initial state.

Prints one JSON line.  `value` is bead-moves per second over the timed steps
(positions, counts and fields resident in HBM throughout); the timed window
starts at a spatial ordering (every 100 steps) and holds ceil(steps / 100) of
them.  `roofline` is for the dominant kernel (the sweep: move schedules +
moves), from HIP events on the engine's stream, against the scope table's
algorithmic bytes, with the counter-based VALU issue figure beside it;
`passes` gives the gridding and omega passes.  `small` is the 262,144-bead
trial-move melt's whole run (10,000 steps), `largest` the 1.0e9-bead blend,
`batched` and `served` the many-system forms.  `cpu_baseline` is the CPU
oracle (the restated reference move, gridding and omega logic, OpenMP) built
for and timed on this host.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def spec_pow2(spec):
    """Every axis has cells of a power-of-two length (the engine's fast binning)."""
    return all(math.frexp(L / n)[0] == 0.5 for L, n in zip(spec.L, spec.n))


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_info():
    """The host the CPU baseline runs on: logical CPUs visible to this process,
    the CPU model (lscpu, else /proc/cpuinfo) and the OpenMP threads used."""
    import subprocess
    info = {"nproc": len(os.sched_getaffinity(0)), "cpu_count": os.cpu_count()}
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core"):
                info[k.strip()] = v.strip()
        model = info.get("Model name")
    except (OSError, subprocess.SubprocessError):
        pass
    if model is None:
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        info["Model name"] = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    return info


def native_oracle():
    """Compile the oracle for this host (-O3 -march=native -fopenmp, FP
    contraction off, as the checker is built) into a temporary directory and
    point the oracle binding at it; fall back to the prebuilt portable library
    (-march=x86-64-v3) if the compiler is unavailable.  Returns the build used."""
    import subprocess
    import tempfile
    from oracle import oracle as orc

    src = os.path.join(ROOT, "oracle", "scmf_oracle.cpp")
    out = os.path.join(tempfile.mkdtemp(prefix="scmf_oracle_"), "liboracle_native.so")
    cmd = ["g++", "-O3", "-march=native", "-std=c++17", "-fPIC", "-shared", "-fopenmp",
           "-ffp-contract=off", "-fno-fast-math", "-o", out, src]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
        if r.returncode == 0 and orc._lib is None:
            orc.LIB_PATH = out
            return "g++ -O3 -march=native -fopenmp -ffp-contract=off (built on this host)"
        log(f"native oracle build failed: {r.stderr[-300:]}")
    except (OSError, subprocess.SubprocessError) as e:
        log(f"native oracle build failed: {e}")
    return "g++ -O3 -march=x86-64-v3 -fopenmp -ffp-contract=off (prebuilt)"


def cpu_step_figures(spec, opts, sweep_target_s=12.0, max_sweeps=20):
    """The oracle's step (sweep + gridding + omega) on one named system: the
    gridding and the omega update of the whole system, then the sweep with the
    omega just computed, over every chain if that fits the time budget, else
    over the first chains (a bounded sample) scaled to the whole system."""
    from oracle import oracle as orc

    o = orc.Oracle(spec, opts)
    pos = o.init_coords()
    t = time.perf_counter()
    counts = o.grid(pos)
    t_grid = time.perf_counter() - t
    t = time.perf_counter()
    om = o.omega(counts).astype(np.float32)
    t_omega = time.perf_counter() - t
    N = max(p.length for p in spec.polymers)
    probe = min(spec.n_chains, 4096)
    t = time.perf_counter()
    o.sweep(pos, om, 0, 0, probe)
    per_chain = (time.perf_counter() - t) / probe
    chains = int(max(probe, min(spec.n_chains, sweep_target_s / per_chain)))
    sweeps = int(max(1, min(max_sweeps, sweep_target_s / (chains * per_chain))))
    t = time.perf_counter()
    for k in range(sweeps):
        o.sweep(pos, om, 1 + k, 0, chains)
    dt = time.perf_counter() - t
    beads_swept = sum(p.count * p.length for p in spec.polymers) * chains / spec.n_chains
    sweep_rate = beads_swept * sweeps / dt
    B = spec.n_beads
    step_s = B / sweep_rate + t_grid + t_omega
    return {"bead_moves_per_s": B / step_s, "sweep_bead_moves_per_s": sweep_rate,
            "grid_s": round(t_grid, 4), "omega_s": round(t_omega, 5),
            "sweep_chains": chains, "sweeps": sweeps, "sweep_s": round(dt, 2),
            "step_s": round(step_s, 4)}


def cpu_full_run(spec, opts, steps):
    """The oracle's whole step loop (test.cpp:37-43 with omega every sweep) for
    `steps` steps on the whole system: bead-moves/s of the full run."""
    from oracle import oracle as orc

    o = orc.Oracle(spec, opts)
    pos = o.init_coords()
    om = o.omega(o.grid(pos)).astype(np.float32)
    t = time.perf_counter()
    for k in range(steps):
        o.sweep(pos, om, k)
        om = o.omega(o.grid(pos)).astype(np.float32)
    dt = time.perf_counter() - t
    return {"bead_moves_per_s": spec.n_beads * steps / dt, "steps": steps,
            "seconds": round(dt, 2)}


def cpu_baseline(spec, opts, small_steps=10000, with_largest=True):
    """The CPU oracle (the restated reference move, gridding and omega logic,
    OpenMP over chains / private histograms / cells) on this host: the headline
    system's step from a bounded sweep sample plus whole-system gridding and
    omega; the small melt's whole run; the largest system's step."""
    from new_arch_amd import workloads as W
    from new_arch_amd.abi import make_opts
    from oracle import oracle as orc

    build = native_oracle()
    hi = host_info()
    log(f"cpu baseline: {build}; {orc.threads()} threads; {hi.get('Model name')}")
    head = cpu_step_figures(spec, opts)
    log(f"cpu headline: {head}")
    small = W.small_diblock_melt()
    sopts = make_opts(seed=0, move="trial", trial_delta=0.1, omega_every_step=True,
                      density_bits=16)
    small_run = cpu_full_run(small, sopts, small_steps)
    log(f"cpu small melt, {small_steps} steps: {small_run}")
    largest = None
    if with_largest:
        lg = W.largest_blend()
        largest = cpu_step_figures(lg, make_opts(seed=0, move="smart", omega_every_step=True,
                                                 density_bits=32), sweep_target_s=8.0,
                                   max_sweeps=1)
        log(f"cpu largest: {largest}")
    N = spec.polymers[0].length
    return {"value": head["bead_moves_per_s"], "unit": "bead-moves/s", "cores": orc.threads(),
            "kind": "port", "build": build, "host": hi,
            "sample": f"one step of the headline system = the oracle's smart-MC sweep of the first "
                      f"{head['sweep_chains']} chains x {head['sweeps']} sweeps ({head['sweep_s']} s, "
                      f"against the omega of the system's own counts) scaled to {spec.n_beads} "
                      f"beads, plus its gridding ({head['grid_s']} s) and omega update "
                      f"({head['omega_s']} s) of the whole system",
            "headline": head,
            "small_diblock_melt_full_run": small_run,
            "largest_blend_step": largest}


def batched_vs_plain(steps, warmup, n_systems=64):
    """The batched form against the same job mix run one system at a time with
    the plain per-system kernels: aggregate bead-moves/s of each."""
    from new_arch_amd import scmf
    from new_arch_amd import workloads as W
    from new_arch_amd.abi import make_opts

    mix = W.batched_mix(n_systems=n_systems, seed=0)

    def members():
        out = []
        for spec, seed, sid in mix:
            g = scmf.System(spec, make_opts(seed=seed, system_id=sid, omega_every_step=True,
                                            density_bits=16))
            g.grid_density()
            g.update_omega()
            out.append(g)
        for g in out:
            g.sync()
        return out

    beads = sum(spec.n_beads for spec, _, _ in mix)
    plain = members()
    for g in plain:                      # one system at a time, plain kernels
        for k in range(warmup):
            g.step(k + 1)
        g.sync()
    t0 = time.perf_counter()
    for g in plain:
        for k in range(steps):
            g.step(warmup + k + 1)
        g.sync()
    t_plain = time.perf_counter() - t0
    for g in plain:
        g.close()
    batch_members = members()
    b = scmf.Batch(batch_members)
    best = None
    for lanes in (1, 2, 4, 8, 16):       # autotune the shared chain-to-wavefront shape
        b.set_lanes(lanes)
        b.step(1)
        b.sync()
        t = time.perf_counter()
        b.step(3)
        b.sync()
        dt = time.perf_counter() - t
        if best is None or dt < best[1]:
            best = (lanes, dt)
    b.set_lanes(best[0])
    b.step(max(0, warmup - 4))
    b.sync()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    t_batch = time.perf_counter() - t0
    b.close()
    for g in batch_members:
        g.close()
    plain_v = beads * steps / t_plain
    batch_v = beads * steps / t_batch
    log(f"batched mix: {n_systems} systems, {beads} beads; plain {plain_v:.3e}, "
        f"batched {batch_v:.3e} bead-moves/s ({batch_v / plain_v:.1f}x), lanes {best[0]}")
    return {"workload": "batched_mix", "systems": n_systems, "beads": beads, "steps": steps,
            "plain_bead_moves_per_s": plain_v, "batched_bead_moves_per_s": batch_v,
            "batched_over_plain": batch_v / plain_v, "lanes_per_chain": best[0],
            "plain_ms_per_step_all_systems": t_plain * 1e3 / steps,
            "batched_ms_per_step": t_batch * 1e3 / steps}


def plain_jobs(jobs, seed_base=10_000):
    """The same jobs run one at a time with the plain per-system kernels
    (create, grid, omega, then scmf_step for every sweep): bead-moves/s."""
    from new_arch_amd import scmf
    from new_arch_amd.abi import make_opts

    moves = 0
    t0 = time.perf_counter()
    for k, (spec, sweeps) in enumerate(jobs):
        g = scmf.System(spec, make_opts(seed=seed_base + k, system_id=k, omega_every_step=True,
                                        density_bits=16))
        g.grid_density()
        g.update_omega()
        for step in range(1, sweeps + 1):
            g.step(step)
        g.sync()
        g.close()
        moves += spec.n_beads * sweeps
    dt = time.perf_counter() - t0
    return {"jobs": len(jobs), "moves": int(moves), "seconds": round(dt, 2),
            "bead_moves_per_s": moves / dt}


def served(n_jobs, rate, seed=0, steps_per_tick=25, plain_sample=24):
    """Poisson stream of jobs into the continuous-batching server, driven in
    wall-clock time: p50/p99 latency (finish - arrival), aggregate
    bead-moves/s over the run and the GPU's busy fraction (device time of the
    server's ticks over their span); then the first `plain_sample` jobs of the
    same mix run one at a time with the plain kernels."""
    from new_arch_amd import scmf
    from new_arch_amd import workloads as W
    from new_arch_amd.abi import JOB_DONE, make_opts

    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
    js = [W.served_job(rng) for _ in range(n_jobs)]
    sv = scmf.Server(steps_per_tick=steps_per_tick)
    t0 = sv.clock()
    submitted, live, lat, moves = 0, {}, [], 0
    t_first = None
    next_note = 30.0
    while len(lat) < n_jobs:
        now = sv.clock() - t0
        if now >= next_note:  # progress of a long run
            log(f"served: {len(lat)} of {n_jobs} jobs done, {submitted} submitted, {now:.0f} s")
            next_note += 30.0
        while submitted < n_jobs and arrivals[submitted] <= now:
            spec, sweeps = js[submitted]
            jid = sv.submit(spec, make_opts(seed=10_000 + submitted, system_id=submitted,
                                            omega_every_step=True, density_bits=16), sweeps)
            live[jid] = (spec.n_beads, sweeps)
            t_first = t_first if t_first is not None else sv.info(jid).t_submit
            submitted += 1
        q, r = sv.tick()
        for jid in [j for j in live if sv.info(j).state == JOB_DONE]:
            info = sv.info(jid)
            lat.append(info.t_done - info.t_submit)
            beads, sweeps = live.pop(jid)
            moves += beads * sweeps
            t_last = info.t_done
            sv.release(jid)
        if q == 0 and r == 0 and submitted < n_jobs:
            time.sleep(max(0.0, min(0.01, arrivals[submitted] - (sv.clock() - t0))))
    st = sv.stats()
    sv.close()
    lat = np.array(lat)
    agg = moves / (t_last - t_first)
    busy = st.busy_s / st.span_s if st.span_s > 0 else None
    log(f"served: {n_jobs} jobs at {rate}/s; p50 {np.percentile(lat, 50):.3f} s, "
        f"p99 {np.percentile(lat, 99):.3f} s, {agg:.3e} bead-moves/s, GPU busy {busy}")
    plain = plain_jobs(js[:plain_sample]) if plain_sample else None
    if plain:
        log(f"served mix one at a time (first {plain_sample} jobs): {plain}")
    return {"workload": "served_jobs", "jobs": n_jobs, "arrival_rate_per_s": rate,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()), "aggregate_bead_moves_per_s": agg,
            "moves": int(moves), "span_s": t_last - t_first,
            "gpu_busy_fraction": busy, "ticks": int(st.ticks),
            "membership_changes": int(st.rebuilds), "graph_captures": int(st.captures),
            "steps_per_tick": steps_per_tick,
            "plain_one_at_a_time": plain,
            "served_over_plain": agg / plain["bead_moves_per_s"] if plain else None,
            "job_mix": "AB diblock or ABA triblock melts, N 16-128, grid 16^3-128^3 (Re/8 "
                       "cells), chiN 10-50, 500-5000 sweeps, 128 beads/Re^3, uint16"}


def profile_json(workload, kind):
    """Per-kernel figures of a committed rocprofv3 run of this workload
    (profiles/<workload>_<kind>.json), or {} if there is none."""
    path = os.path.join(ROOT, "profiles", f"{workload}_{kind}.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def issue_figures(workload, kernels, B):
    """VALU issue of the sweep from the committed counter run: SIMD issue
    cycles 4 (SQ_INSTS_VALU - SQ_ACTIVE_INST_VALU2) / SIMDs against the
    elapsed GRBM_GUI_ACTIVE cycles per XCD (the model the calibration kernels
    of libscmf_calib.so confirm within 94-96% for VALU-saturated kernels)."""
    prof = profile_json(workload, "valu").get("per_launch", {})
    cal = profile_json("valu", "calibration")
    if not all(k in prof for k in kernels):
        return None
    insts = sum(prof[k]["SQ_INSTS_VALU"] for k in kernels)
    dual = sum(prof[k]["SQ_ACTIVE_INST_VALU2"] for k in kernels)
    cycles = sum(prof[k]["GRBM_GUI_ACTIVE"] for k in kernels) / 8.0
    simds = 1024
    busy = 4.0 * (insts - dual) / simds / cycles
    sat = cal.get("saturated_issue_fraction")
    return {"bound": "valu-issue", "valu_wave_instructions": insts,
            "dual_issued_quad_cycles": dual, "elapsed_cycles": cycles,
            "issue_busy": round(busy, 4),
            "of_saturated": round(busy / sat, 4) if sat else None,
            "saturated_issue_fraction": sat,
            "wave_instructions_per_move": round(insts / B, 3),
            "source": f"profiles/{workload}_valu.json, profiles/valu_calibration.json"}


# the engine's automatic spatial-ordering period (systems of 2^23 beads or more)
ORDER_PERIOD = 100
LDS_PER_CU = 160 * 1024   # MI355X LDS per compute unit


def sweep_run_lanes(spec, lanes, pre):
    """(chain length, lanes a chain) of each sweep launch, as the engine makes
    them: consecutive architectures of one length share a launch, and at 8
    lanes a chain with pre-made schedules a run of chains longer than 64 beads
    takes 16 lanes when LDS then keeps more waves resident (at most 20 a CU)."""
    lengths = []
    for p in spec.polymers:
        if p.count > 0 and (not lengths or lengths[-1] != p.length):
            lengths.append(p.length)
    if spec.n_chains < 65536:
        lengths = [max(lengths)]

    def resident(L, n):
        G = 64 // L
        return min(20, LDS_PER_CU // (3 * n * G * 4 + G * (3 * n + 4)))
    return [(n, 16 if lanes == 8 and pre and n > 64 and resident(16, n) > resident(8, n) else lanes)
            for n in lengths]


def measure_system(workload, steps, warmup, with_host_copy=False, autotune=True):
    """Create the named system, autotune the sweep shape, run `warmup` steps
    (plus the untimed steps up to the next spatial ordering, so the timed
    window starts with one and holds ceil(steps / ORDER_PERIOD) of them), then time
    `steps` steps (sweep + gridding + omega every sweep)."""
    from new_arch_amd import scmf
    from new_arch_amd import workloads as W
    from new_arch_amd.abi import make_opts

    spec = W.NAMED[workload]()
    move = "trial" if workload == "small_diblock_melt" else "smart"
    bits = 32 if workload == "largest_blend" else 16
    opts = make_opts(seed=0, move=move, trial_delta=0.1, omega_every_step=True, density_bits=bits)
    log(f"creating {workload}: {spec.n_beads} beads, {spec.n_chains} chains, grid {spec.n}")
    t = time.perf_counter()
    g = scmf.System(spec, opts)
    g.grid_density()
    g.update_omega()
    g.sync()
    log(f"created in {time.perf_counter() - t:.1f} s")
    lanes, tuned_ms = g.autotune(2) if autotune else (0, 0.0)
    log(f"autotune: {lanes} lanes per chain, {tuned_ms:.3f} ms per sweep")
    period = ORDER_PERIOD if spec.n_beads >= (1 << 23) else 0   # automatic ordering period
    step_no = 0
    for _ in range(warmup):
        step_no += 1
        g.step(step_no)
    if period:
        while g.stats().sweeps % period:
            step_no += 1
            g.step(step_no)
    g.sync()
    log(f"warmup {step_no} steps done")
    # the throughput window runs without timing events (each is a record on
    # the stream, noticeable in the small melt's 0.07 ms steps)
    g.enable_timing(False)
    g.reset_stats()
    g.sync()
    st0 = g.stats()
    t0 = time.perf_counter()
    for i in range(steps):
        g.step(step_no + i + 1)
    g.sync()
    t1 = time.perf_counter()
    step_no += steps
    st = g.stats()
    # per-phase device times (sweep, gridding, omega) from HIP events around
    # the kernels of the next steps (at most 25)
    g.enable_timing(True)
    g.reset_timing()
    for i in range(min(steps, 25)):
        g.step(step_no + i + 1)
    g.sync()
    step_no += min(steps, 25)
    tm = g.timing()
    g.enable_timing(False)
    from_cells = st.cell_grids - st0.cell_grids > 0
    ordering = None
    if period:
        # one ordering and the gridding after it (the cells move with the
        # positions, so it reads cells), timed on their own; then the
        # gridding of the scope table, binning positions (12 B a bead), after
        # the positions went out and back in (which leaves no cells to read)
        g.enable_timing(False)
        g.sync()
        t = time.perf_counter()
        g.sort_chains()
        g.sync()
        t_sort = time.perf_counter() - t
        g.enable_timing(True)
        g.reset_timing()
        cg0 = g.stats().cell_grids
        g.grid_density()
        g.sync()
        t_after = g.timing().grid_ms
        after_from_cells = g.stats().cell_grids > cg0
        g.enable_timing(False)
        pos = g.positions()
        g.set_positions(pos)
        del pos
        g.enable_timing(True)
        g.reset_timing()
        g.grid_density()
        g.sync()
        t_bin = g.timing().grid_ms
        ordering = {"period_steps": period, "in_timed_steps": int(st.sorts - st0.sorts),
                    "ordering_ms": round(t_sort * 1e3, 3),
                    "gridding_after_ordering_ms": round(t_after, 4),
                    "gridding_after_ordering_reads_cells": bool(after_from_cells),
                    "gridding_from_positions_ms": round(t_bin, 4)}
    host = None
    if with_host_copy:
        # The rate a caller would see if it copied the positions out and back
        # in every step (not the headline: the engine keeps them resident).
        t = time.perf_counter()
        pos = g.positions()
        t_out = time.perf_counter() - t
        t = time.perf_counter()
        g.set_positions(pos)
        g.sync()
        t_in = time.perf_counter() - t
        del pos
        step_s = (t1 - t0) / steps
        host = {"copy_out_ms": round(t_out * 1e3, 2), "copy_in_ms": round(t_in * 1e3, 2),
                "position_bytes": 12 * spec.n_beads,
                "bead_moves_per_s_with_copies": spec.n_beads / (step_s + t_out + t_in)}
    analysis = None
    if with_host_copy:
        # The composition structure factor of the final state on the device
        # (weights, in-place D2Z FFT, shell sums), as a run would take it every
        # check interval; synchronous, shell sums copied to the host.
        coef = [1.0, -1.0] + [0.0] * (spec.n_types - 2)
        dq = 2.0 * math.pi / spec.L[0]
        n_bins = spec.n[0]
        for _ in range(2):
            g.structure_factor(coef, dq, n_bins)
        reps = 10
        t = time.perf_counter()
        for _ in range(reps):
            sums, modes = g.structure_factor(coef, dq, n_bins)
        analysis = {"structure_factor_ms": round((time.perf_counter() - t) * 1e3 / reps, 3),
                    "grid": list(spec.n), "shells": n_bins, "modes": int(modes.sum()),
                    "S_first_shell": float(sums[1] / max(modes[1], 1))}
        log(f"structure factor: {analysis['structure_factor_ms']} ms per call")
        # A density_<step>.dat dump of this state: synchronous, and queued
        # (the call's own cost, then the time until the file is written).
        import shutil
        import tempfile
        d = tempfile.mkdtemp(prefix="scmf_dump_")
        try:
            t = time.perf_counter()
            g.write_density(1, d)
            t_sync = time.perf_counter() - t
            size = os.path.getsize(os.path.join(d, "density_1.dat"))
            g.write_density_async(3, d)  # first call allocates the snapshot slots
            g.output_wait()
            t = time.perf_counter()
            g.write_density_async(2, d)
            t_call = time.perf_counter() - t
            g.output_wait()
            t_done = time.perf_counter() - t
        finally:
            shutil.rmtree(d, ignore_errors=True)
        analysis["density_dump"] = {"bytes": size, "sync_ms": round(t_sync * 1e3, 1),
                                    "async_call_ms": round(t_call * 1e3, 3),
                                    "async_written_ms": round(t_done * 1e3, 1)}
        log(f"density dump: {analysis['density_dump']}")
    g.close()
    log(f"{workload}: timed {steps} steps in {t1 - t0:.3f} s; sweep {tm.sweep_ms:.3f} ms, "
        f"grid {tm.grid_ms:.3f} ms, omega {tm.omega_ms:.3f} ms; acceptance "
        f"{st.accepts / max(st.moves, 1):.3f}; grid fallback {st.grid_fallback}; ordering {ordering}")

    B, TC = spec.n_beads, spec.n_types * spec.n_cells
    s = 2 if bits == 16 else 4
    nmax = max(p.length for p in spec.polymers)
    pre = nmax <= 255 and lanes >= 2 and spec.n_chains >= 65536
    # algorithmic bytes of the scope table: positions read + written, omega read once
    sweep_bytes = 24 * B + 4 * TC
    # bytes the implementation moves beyond them: the bead cells the sweep stores
    # for the next gridding, the move-schedule records written and read back
    impl = {"cells_written": 4 * B if from_cells else 0,
            "schedule_records": 2 * (3 * nmax + 3) * spec.n_chains if pre else 0}
    grid_bytes_cells = 4 * B + s * TC   # the pass as run: cells read, counts written
    omega_bytes = s * TC + 4 * TC       # counts read, fp32 omega written
    tiled = spec.n_beads >= (1 << 23)
    sweep_kernel = {-1: "k_sweep_global", 1: "k_sweep_lds"}.get(lanes, f"k_sweep_sched<{lanes}>")
    if lanes > 1:
        sweep_kernel = "+".join(f"k_sweep_sched<{L}>"
                                for L in sorted({L for _, L in sweep_run_lanes(spec, lanes, pre)}))

    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    # of this workload (calibration and commands in the profile summary).
    measured = profile_json(workload, "traffic").get("bytes_per_launch", {})
    byte_t = "unsigned char" if nmax <= 255 else "unsigned short"
    run_lanes = sweep_run_lanes(spec, lanes, pre)
    # the schedule kernel keeps round counts in registers for chains of <= 128 beads
    # and rounds of 2, 4, 8 or 16 moves (else one k_schedule_lds launch at the
    # sweep's lanes over every chain)
    in_regs = all(n <= 128 for n, _ in run_lanes)
    if not in_regs:
        run_lanes = [(n, lanes) for n, _ in run_lanes]
    sched_keys = ([f"k_schedule_reg<{c}, {w}>" for c, w in
                   sorted({(L.bit_length() - 1, 1 if n <= 64 else 2) for n, L in run_lanes})]
                  if in_regs else ["k_schedule_lds"])
    sweep_keys = ([f"k_sweep_sched<{L}, {byte_t}>" for L in sorted({L for _, L in run_lanes})]
                  if lanes > 1 else [sweep_kernel]) + (sched_keys if pre else [])
    omega_key = (f"k_omega4<{spec.n_types}, {bits}>" if (spec.n_cells % 4 == 0)
                 else "k_omega")
    pair = "true" if spec.n_types == 2 else "false"
    # the step's gridding (its counts cleared by the step's sweep)
    grid_keys = [f"k_grid_tiles<{bits}, false, true, {pair}>" if from_cells
                 else f"k_grid_tiles<{bits}, {'true' if spec_pow2(spec) else 'false'}, false, {pair}>",
                 f"k_tiles_sum_pair<{bits}>" if spec.n_types == 2 else f"k_tiles_sum<{bits}>"]

    def traffic_of(keys):
        if all(k in measured for k in keys):
            return sum(measured[k] for k in keys)
        return None

    def roof(nbytes, ms, traffic):
        gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        out = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
               "algorithmic_bytes": nbytes, "ms": round(ms, 4)}
        if traffic is not None:
            out["traffic_source"] = f"profiles/{workload}_kernel_stats.md"
        return out

    passes = {"grid": dict(roof(grid_bytes_cells, tm.grid_ms,
                                traffic_of(grid_keys) if tiled else None),
                           kernel=("k_grid_tiles" + ("(cells)+" if from_cells else "+")
                                   + ("k_tiles_sum_pair" if spec.n_types == 2 else "k_tiles_sum"))
                           if tiled else "k_grid",
                           bytes_note="cells read (4 B a bead) + counts written" if from_cells
                           else "positions read (12 B a bead) + counts written"),
              "omega": dict(roof(omega_bytes, tm.omega_ms, traffic_of([omega_key])),
                            kernel=omega_key)}
    if not from_cells:
        passes["grid"]["algorithmic_bytes"] = 12 * B + s * TC
        passes["grid"] = dict(roof(12 * B + s * TC, tm.grid_ms, passes["grid"]["traffic"]),
                              kernel=passes["grid"]["kernel"])
    if ordering:
        # the gridding of the scope table: binning positions, 12 B a bead
        passes["grid_from_positions"] = dict(
            roof(12 * B + s * TC, ordering["gridding_from_positions_ms"], None),
            kernel=f"k_zero_counts+k_grid_tiles(positions)+"
                   f"{'k_tiles_sum_pair' if spec.n_types == 2 else 'k_tiles_sum'}",
            when="after any change of positions other than a sweep or an ordering")
    dtNxi, nref = spec.dtNxi, spec.nref
    A = dtNxi / nref
    roofline = dict(roof(sweep_bytes, tm.sweep_ms, traffic_of(sweep_keys)),
                    kernel=("+".join(sched_keys) + "+" if pre else "") + sweep_kernel,
                    implementation_bytes=impl)
    issue = issue_figures(workload, sweep_keys, B)
    if issue:
        roofline["issue"] = issue
    return {
        "_spec": spec, "_opts": opts,
        "workload": workload,
        "bead_moves_per_s": B * steps / (t1 - t0),
        "ms_per_step": (t1 - t0) * 1e3 / steps,
        "steps": steps,
        "dtype": "f32 coords / f64 energy+binning / " + ("u16" if bits == 16 else "u32") + " counts",
        "config": {"workload": workload, "beads": B, "chains": spec.n_chains,
                   "N": [p.length for p in spec.polymers], "grid": list(spec.n),
                   "box_Re": list(spec.L), "xn": spec.xn, "move": move,
                   "dtNxi": dtNxi, "Nref": nref,
                   "A_Re2": A if move == "smart" else None,
                   "R_Re": math.sqrt(2 * A) if move == "smart" else None,
                   "trial_delta_Re": 0.1 if move == "trial" else None,
                   "step_parameter": "smart-MC drift A = dtNxi / Nref and noise R = sqrt(2 A) "
                                     "(init_config.cu:804-809); the configuration's step 0.05 "
                                     "is read as dtNxi" if move == "smart" else
                                     "uniform displacement of +-0.1 Re per component",
                   "density_bits": bits, "omega": "every sweep", "parallelism": "1 GPU",
                   "lanes_per_chain": lanes},
        "roofline": roofline,
        "passes": passes,
        "acceptance": st.accepts / max(st.moves, 1),
        "grid_fallback_beads": int(st.grid_fallback),
        "orderings_timed": int(st.sorts - st0.sorts),
        "ordering": ordering,
        "pcie_inclusive": host,
        "analysis": analysis,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=ORDER_PERIOD)  # one spatial ordering period
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="large_diblock_melt")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-largest", action="store_true")
    ap.add_argument("--no-small", action="store_true")
    ap.add_argument("--small-steps", type=int, default=10000)  # the small melt's whole run
    ap.add_argument("--served-jobs", type=int, default=300)
    ap.add_argument("--served-rate", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs one system per GPU and measures a single GPU; "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    m = measure_system(args.workload, args.steps, args.warmup, with_host_copy=True)
    spec, opts = m.pop("_spec"), m.pop("_opts")
    out = {
        "metric": "bead-moves/s/GPU",
        "value": m["bead_moves_per_s"],
        "unit": "bead-moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": m["dtype"],
        "data": "synthetic (seeded Gaussian-walk chains, uniform chain starts)",
        "config": m["config"],
        "roofline": m["roofline"],
        "passes": m["passes"],
        "acceptance": m["acceptance"],
        "orderings_timed": m["orderings_timed"],
        "ordering": m["ordering"],
        "pcie_inclusive": m["pcie_inclusive"],
        "analysis": m["analysis"],
    }
    if not args.no_small and args.workload != "small_diblock_melt":
        log("small system")
        sm = measure_system("small_diblock_melt", steps=args.small_steps, warmup=100)
        sm.pop("_spec"), sm.pop("_opts")
        out["small"] = sm
    if not args.no_largest and args.workload != "largest_blend":
        log("largest system")
        lg = measure_system("largest_blend", steps=ORDER_PERIOD, warmup=1)
        lg.pop("_spec"), lg.pop("_opts")
        out["largest"] = lg
    if not args.no_batched:
        log("batched vs plain")
        out["batched"] = batched_vs_plain(steps=50, warmup=5)
    if not args.no_served:
        log("served")
        out["served"] = served(args.served_jobs, args.served_rate)
    if not args.no_cpu_baseline:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(spec, opts, with_largest=not args.no_largest)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
