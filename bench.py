"""Throughput benchmark of the SCMF Monte Carlo step on one MI355X.

A step is one pass of the hot path over the whole system: a smart-MC sweep
(every chain makes N sequential moves), NGP density gridding and the omega
update.  The workload is the single large system: AB diblock melt, N=64,
fA=0.3, 4,194,304 chains (268,435,456 beads), 256^3 grid, chiN=25, kappaN=70,
uint16 counts, fp32 coordinates; synthetic Gaussian-walk initial state.

Prints one JSON line.  `value` is bead-moves per second over the timed steps
(positions, counts and fields resident in HBM throughout).  `roofline` is
for the dominant kernel (the sweep), from HIP events on the engine's stream;
`passes` gives the same for the gridding and omega passes.  `cpu_baseline` is
the CPU oracle (the restated reference move logic, OpenMP over chains) timed
on a bounded sample of the same workload on this host.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def spec_pow2(spec):
    """Every axis has cells of a power-of-two length (the engine's fast binning)."""
    return all(math.frexp(L / n)[0] == 0.5 for L, n in zip(spec.L, spec.n))


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(spec, opts, target_s=15.0):
    """The oracle's step on all host threads: its sweep timed on the first
    chains of the workload, its gridding and omega update on the whole system."""
    import numpy as np

    from new_arch_amd.abi import SystemSpec
    from oracle import oracle as orc

    N = spec.polymers[0].length

    def sub(chains):
        p = spec.polymers[0]
        return SystemSpec(n_types=spec.n_types, xn=spec.xn, L=spec.L, n=spec.n, dtNxi=spec.dtNxi,
                          nref=spec.nref, polymers=[type(p)(length=N, count=chains,
                                                            monotype=p.monotype)])

    om = np.zeros(spec.n_types * spec.n_cells, np.float32)
    chains = 4096
    o = orc.Oracle(sub(chains), opts)
    pos = o.init_coords()
    t = time.perf_counter()
    o.sweep(pos, om, 0)
    dt = time.perf_counter() - t
    per_chain_sweep = dt / chains
    chains = int(max(4096, min(spec.n_chains, target_s / (2 * per_chain_sweep))))
    sweeps = int(max(2, min(20, target_s / (chains * per_chain_sweep))))
    o = orc.Oracle(sub(chains), opts)
    pos = o.init_coords()
    t = time.perf_counter()
    for s in range(sweeps):
        o.sweep(pos, om, s)
    dt = time.perf_counter() - t
    moves = chains * N * sweeps
    sweep_rate = moves / dt
    # The gridding and the omega update of the whole system, once each.
    o = orc.Oracle(spec, opts)
    pos = o.init_coords()
    t = time.perf_counter()
    counts = o.grid(pos)
    t_grid = time.perf_counter() - t
    del pos
    t = time.perf_counter()
    o.omega(counts)
    t_omega = time.perf_counter() - t
    B = spec.n_beads
    step_s = B / sweep_rate + t_grid + t_omega
    return {"value": B / step_s, "unit": "bead-moves/s", "cores": orc.threads(), "kind": "port",
            "sweep_bead_moves_per_s": sweep_rate, "grid_s": round(t_grid, 3),
            "omega_s": round(t_omega, 4),
            "sample": f"one step = sweep + gridding + omega: the oracle's smart-MC sweep of the "
                      f"first {chains} chains ({chains * N} beads) x {sweeps} sweeps ({dt:.1f} s) "
                      f"scaled to {B} beads, plus its gridding ({t_grid:.2f} s) and omega update "
                      f"({t_omega:.3f} s) of the whole system"}


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
    for lanes in (1, 2, 4, 8):           # autotune the shared chain-to-wavefront shape
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


def served(n_jobs, rate, seed=0, steps_per_tick=25):
    """Poisson stream of jobs into the continuous-batching server, driven in
    wall-clock time: p50/p99 latency (finish - arrival) and aggregate
    bead-moves/s over the run."""
    import numpy as np

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
    sv.close()
    lat = np.array(lat)
    agg = moves / (t_last - t_first)
    log(f"served: {n_jobs} jobs at {rate}/s; p50 {np.percentile(lat, 50):.3f} s, "
        f"p99 {np.percentile(lat, 99):.3f} s, {agg:.3e} bead-moves/s")
    return {"workload": "served_jobs", "jobs": n_jobs, "arrival_rate_per_s": rate,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()), "aggregate_bead_moves_per_s": agg,
            "moves": int(moves), "span_s": t_last - t_first,
            "job_mix": "AB diblock or ABA triblock melts, N 16-128, grid 16^3-128^3 (Re/8 "
                       "cells), chiN 10-50, 500-5000 sweeps, 128 beads/Re^3, uint16"}


def measure_system(workload, steps, warmup, with_host_copy=False):
    """Create the named system, autotune the sweep shape, run `warmup` steps,
    then time `steps` steps (sweep + gridding + omega every sweep)."""
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
    lanes, tuned_ms = g.autotune(2)
    log(f"autotune: {lanes} lanes per chain, {tuned_ms:.3f} ms per sweep")
    for i in range(warmup):
        g.step(i + 1)
    g.sync()
    log(f"warmup {warmup} steps done")
    g.enable_timing(True)
    g.reset_timing()
    g.reset_stats()
    g.sync()
    cell_grids0 = g.stats().cell_grids
    sorts0 = g.stats().sorts
    t0 = time.perf_counter()
    for i in range(steps):
        g.step(warmup + i + 1)
    g.sync()
    t1 = time.perf_counter()
    tm = g.timing()
    st = g.stats()
    # every timed gridding read the bead cells the sweep stored
    from_cells = st.cell_grids - cell_grids0 == steps
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
        f"{st.accepts / max(st.moves, 1):.3f}; grid fallback {st.grid_fallback}")

    B, TC = spec.n_beads, spec.n_types * spec.n_cells
    s = 2 if bits == 16 else 4
    # positions read + written, omega read once (+ the bead cells written
    # for the gridding); the gridding reads the cells (4 B) or the positions
    # (12 B) and writes the counts
    sweep_bytes = 24 * B + 4 * TC + (4 * B if from_cells else 0)
    # chains of at most 255 beads: k_schedule writes each chain's schedule
    # record (2 nmax + 3 bytes) and the sweep reads it back
    nmax = max(p.length for p in spec.polymers)
    pre = nmax <= 255 and lanes >= 2
    if pre:
        sweep_bytes += 2 * (2 * nmax + 3) * spec.n_chains
    grid_bytes = (4 if from_cells else 12) * B + s * TC
    omega_bytes = s * TC + 4 * TC       # counts read, fp32 omega written
    tiled = spec.n_beads >= (1 << 23)
    sweep_kernel = {-1: "k_sweep_global", 1: "k_sweep_lds"}.get(lanes, f"k_sweep_sched<{lanes}>")

    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    # of this workload (calibration and commands in the profile summary).
    measured = {}
    tpath = os.path.join(ROOT, "profiles", f"{workload}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            measured = json.load(f).get("bytes_per_launch", {})
    byte_t = "unsigned char" if spec.polymers and max(p.length for p in spec.polymers) <= 255 else "unsigned short"
    sweep_keys = ([f"k_sweep_sched<{lanes}, {byte_t}>"] if lanes > 1 else [sweep_kernel]) + \
        (["k_schedule"] if pre else [])
    omega_key = (f"k_omega4<{spec.n_types}, {bits}>" if (spec.n_cells % 4 == 0)
                 else "k_omega")
    pair = "true" if spec.n_types == 2 else "false"
    grid_keys = [f"k_grid_tiles<{bits}, false, true, {pair}>" if from_cells
                 else f"k_grid_tiles<{bits}, {'true' if spec_pow2(spec) else 'false'}, false, {pair}>",
                 f"k_tiles_sum_pair<{bits}>" if spec.n_types == 2 else f"k_tiles_sum<{bits}>",
                 "k_zero_counts"]

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
                   "density_bits": bits, "omega": "every sweep", "parallelism": "1 GPU",
                   "lanes_per_chain": lanes},
        "roofline": dict(roof(sweep_bytes, tm.sweep_ms, traffic_of(sweep_keys)),
                         kernel=("k_schedule+" if pre else "") + sweep_kernel),
        "passes": {"grid": dict(roof(grid_bytes, tm.grid_ms,
                                     traffic_of(grid_keys) if tiled else None),
                                kernel=("k_zero_counts+k_grid_tiles" + ("(cells)+" if from_cells else "+")
                                        + ("k_tiles_sum_pair" if spec.n_types == 2 else "k_tiles_sum"))
                                if tiled else "k_zero_counts+k_grid"),
                   "omega": dict(roof(omega_bytes, tm.omega_ms, traffic_of([omega_key])),
                                 kernel=omega_key)},
        "acceptance": st.accepts / max(st.moves, 1),
        "grid_fallback_beads": int(st.grid_fallback),
        # spatial orderings inside the timed steps (one every 25 sweeps)
        "orderings_timed": int(st.sorts - sorts0),
        "pcie_inclusive": host,
        "analysis": analysis,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=25)  # one spatial ordering period
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="large_diblock_melt")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-largest", action="store_true")
    ap.add_argument("--served-jobs", type=int, default=200)
    ap.add_argument("--served-rate", type=float, default=6.0)
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
        "pcie_inclusive": m["pcie_inclusive"],
        "analysis": m["analysis"],
    }
    if not args.no_largest and args.workload != "largest_blend":
        log("largest system")
        lg = measure_system("largest_blend", steps=5, warmup=1)
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
        out["cpu_baseline"] = cpu_baseline(spec, opts)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
