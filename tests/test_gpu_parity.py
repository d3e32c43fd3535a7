"""Parity of the HIP engine (through its C ABI) with the CPU oracle.

Bar: bit-exact positions, counts and accept decisions under the same injected
Philox streams; omega bit-exact against the oracle's fp64 value rounded once
to fp32 (hence within 2^-24 relative of the fp64 reference value, inside the
1e-6 relative tolerance).  Sizes are small enough for the oracle to finish in
seconds; the full-size systems are checked through size-independent
invariants.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from new_arch_amd import scmf
from new_arch_amd import workloads as W
from new_arch_amd.abi import PolymerType, SystemSpec, make_opts
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SPECS = {
    "tiny": W.tiny,
    "mixed": W.mixed_arch,
    "shipped": W.reference_shipped,
}


def banner(msg):
    print(f"[gpu-parity] {msg}", flush=True)


def pair(spec, **kw):
    o = make_opts(**kw)
    return scmf.System(spec, o), orc.Oracle(spec, o)


def same_bits(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape and a.dtype == b.dtype
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("name", list(SPECS))
def test_init_bit_exact(name):
    banner(f"init {name}")
    spec = SPECS[name]()
    g, o = pair(spec, seed=12345)
    assert same_bits(g.positions(), o.init_coords())
    g.close()


@pytest.mark.parametrize("bits", [16, 32])
@pytest.mark.parametrize("name", list(SPECS))
def test_grid_bit_exact(name, bits):
    banner(f"grid {name} u{bits}")
    spec = SPECS[name]()
    g, o = pair(spec, seed=7, density_bits=bits)
    pos = o.init_coords()
    g.grid_density()
    np.testing.assert_array_equal(g.density(), o.grid(pos))
    # unwrapped, far-away and negative coordinates, cell edges
    rng = np.random.default_rng(3)
    L = np.array(spec.L)
    far = (rng.uniform(-40, 40, (spec.n_beads, 3)) * L).astype(np.float32)
    edges = (rng.integers(-3 * spec.n[0], 3 * spec.n[0], (spec.n_beads // 4, 3))
             * (L / np.array(spec.n))).astype(np.float32)
    far[: edges.shape[0]] = edges
    g.set_positions(far)
    g.grid_density()
    np.testing.assert_array_equal(g.density(), o.grid(far))
    assert g.stats().overflow == 0
    g.close()


@pytest.mark.parametrize("name", list(SPECS))
def test_omega_bit_exact(name):
    banner(f"omega {name}")
    spec = SPECS[name]()
    g, o = pair(spec, seed=9)
    g.grid_density()
    g.update_omega()
    counts = g.density()
    ref64 = o.omega(counts)
    got = g.omega()
    assert same_bits(got, ref64.astype(np.float32))
    rel = np.abs(got.astype(np.float64) - ref64) / np.maximum(np.abs(ref64), 1e-300)
    assert rel[ref64 != 0].max() <= 2.0 ** -24
    g.close()


@pytest.mark.parametrize("step", [240000, 999000])
def test_omega_reference_dumps_on_gpu(golden, step):
    banner(f"omega dump {step}")
    f = np.load(os.path.join(golden, "density_fixtures.npz"))
    grid = [int(v) for v in f[f"grid_{step}"]]
    kappa, chi = f[f"kappa_chi_{step}"]
    poly = PolymerType(length=32, count=12800, monotype=[0] * 4 + [1] * 24 + [2] * 4)
    spec = SystemSpec(n_types=3, xn=W.xn_matrix(3, kappa, chi), L=[v / 6.0 for v in grid],
                      n=grid, dtNxi=0.17, nref=32, polymers=[poly])
    g = scmf.System(spec, make_opts(seed=1, density_bits=16, init_coords=False))
    prev = np.ascontiguousarray(f[f"prev_{step}"].astype(np.uint32).T)
    g.set_density(prev)
    g.update_omega()
    got = g.omega().reshape(3, -1).T.astype(np.float64)
    printed = f[f"omega_{step}"].astype(np.float64)
    assert np.all(np.abs(got - printed) <= 6e-6 * np.abs(printed) + 2e-6)
    g.close()


@pytest.mark.parametrize("lanes", [0, -1, 1, 2, 4, 8, 16])
@pytest.mark.parametrize("move", ["smart", "trial"])
@pytest.mark.parametrize("name", list(SPECS))
def test_sweep_bit_exact(name, move, lanes):
    banner(f"sweep {name} {move} lanes={lanes}")
    spec = SPECS[name]()
    g, o = pair(spec, seed=2024, move=move, trial_delta=0.1, lanes_per_chain=lanes)
    pos = o.init_coords()
    g.grid_density()
    g.update_omega()
    om = g.omega()
    acc_ref = 0
    for s in range(3):
        g.mc_sweep(1)
        acc_ref += o.sweep(pos, om, s)
        assert same_bits(g.positions(), pos), f"sweep {s}"
    st = g.stats()
    assert st.accepts == acc_ref and st.moves == 3 * spec.n_beads
    assert 0 < st.accepts < st.moves
    g.close()


@pytest.mark.parametrize("cadence", [0, 1])
def test_step_loop_bit_exact(cadence):
    """The test.cpp:37-43 loop: sweep, grid, omega unless step%10==0."""
    banner(f"step loop cadence={cadence}")
    spec = W.tiny(n_chains=256, N=24, grid=(10, 10, 10), L=(2.5, 2.5, 2.5), chiN=30.0)
    g, o = pair(spec, seed=77, omega_every_step=bool(cadence), density_bits=16)
    pos = o.init_coords()
    g.grid_density()
    g.update_omega()
    om = o.omega(o.grid(pos)).astype(np.float32)
    for step in range(1, 26):
        g.step(step)
        o.sweep(pos, om, step - 1)
        counts = o.grid(pos)
        if cadence or step % 10:
            om = o.omega(counts).astype(np.float32)
    assert same_bits(g.positions(), pos)
    np.testing.assert_array_equal(g.density(), counts)
    assert same_bits(g.omega(), om)
    g.close()


def test_small_diblock_melt_two_steps():
    """The 262,144-bead trial-move melt, omega every sweep, uint16 counts."""
    banner("small diblock melt")
    spec = W.small_diblock_melt()
    g, o = pair(spec, seed=0, move="trial", trial_delta=0.1, omega_every_step=True,
                density_bits=16)
    pos = o.init_coords()
    g.grid_density()
    g.update_omega()
    om = o.omega(o.grid(pos)).astype(np.float32)
    for step in range(1, 3):
        g.step(step)
        o.sweep(pos, om, step - 1)
        om = o.omega(o.grid(pos)).astype(np.float32)
    assert same_bits(g.positions(), pos)
    assert same_bits(g.omega(), om)
    g.close()


def test_autotune_leaves_trajectory_unchanged():
    banner("autotune")
    spec = W.mixed_arch(scale=4)
    g, o = pair(spec, seed=8)
    g.grid_density()
    g.update_omega()
    lanes, ms = g.autotune(2)
    assert lanes in (-1, 1, 2, 4, 8, 16) and ms > 0
    assert g.stats().accepts == 0 and g.stats().sweeps == 0
    pos = o.init_coords()
    om = g.omega()
    g.mc_sweep(2)
    o.sweep(pos, om, 0)
    o.sweep(pos, om, 1)
    assert same_bits(g.positions(), pos)
    g.close()


def tiled_spec(grid, T=2, N=16):
    L = [g * 0.125 for g in grid]
    poly = [PolymerType(length=N, count=600, monotype=[t % T for t in range(N)])]
    if T == 3:
        poly.append(PolymerType(length=N // 2, count=400, monotype=[2] * (N // 2)))
    return SystemSpec(n_types=T, xn=W.xn_matrix(T, 40.0, 15.0), L=L, n=list(grid), dtNxi=0.1,
                      nref=N, polymers=poly)


@pytest.mark.parametrize("grid,N", [((32, 32, 32), 16), ((32, 32, 32), 24), ((30, 32, 28), 16)])
@pytest.mark.parametrize("cadence", [0, 1])
@pytest.mark.parametrize("bits", [16, 32])
def test_step_loop_on_ordered_two_type_system(bits, cadence, grid, N):
    """scmf_step on a spatially ordered two-type system (the sweep clearing the
    counts, the tiled gridding advancing the stream word, orderings every 3
    sweeps): positions, counts and omega equal the oracle's after every step
    of both cadences (omega kept on steps 10 and 20 with the reference
    cadence).  The gridding from cells takes computed window offsets on the
    power-of-two grids (with each lane's types read once for N = 16, whose
    length divides the 4096-bead pass, and per pass for N = 24) and the
    axis tables on the 30 x 32 x 28 grid."""
    banner(f"ordered step loop u{bits} cadence={cadence} grid {grid} N={N}")
    spec = tiled_spec(grid, T=2, N=N)
    g, o = pair(spec, seed=41, density_bits=bits, sort_every=3, omega_every_step=bool(cadence))
    pos = o.init_coords()
    g.grid_density()
    g.update_omega()
    om = o.omega(o.grid(pos)).astype(np.float32)
    assert same_bits(g.omega(), om)
    for step in range(1, 26):
        g.step(step)
        o.sweep(pos, om, step - 1)
        counts = o.grid(pos)
        if cadence or step % 10:
            om = o.omega(counts).astype(np.float32)
        assert same_bits(g.positions(), pos), f"step {step} positions"
        np.testing.assert_array_equal(g.density(), counts, err_msg=f"step {step} counts")
        assert same_bits(g.omega(), om), f"step {step} omega"
    assert g.stats().sorts >= 8
    g.close()


@pytest.mark.parametrize("bits", [16, 32])
@pytest.mark.parametrize("grid,T", [((27, 31, 40), 2), ((32, 32, 32), 3), ((64, 24, 33), 2),
                                    ((48, 40, 10), 2), ((40, 36, 30), 1)])
def test_tiled_gridding_bit_exact(grid, T, bits):
    """Spatially ordered chains, windows with partial tiles, outliers on the
    global-atomic fallback; counts equal the oracle's at every stage.  (A slab
    of 10 cells is thinner than two halos: one tile and no halo along z.)"""
    banner(f"tiled grid {grid} T={T} u{bits}")
    spec = tiled_spec(grid, T)
    g, o = pair(spec, seed=5, density_bits=bits, sort_every=3)
    pos = o.init_coords()
    g.grid_density()                               # orders the chains first
    np.testing.assert_array_equal(g.density(), o.grid(pos))
    st = g.stats()
    assert st.sorts == 1
    assert same_bits(g.positions(), pos)            # storage order is invisible
    # outliers: a few chains spread over far periodic images
    rng = np.random.default_rng(1)
    far = pos.copy()
    idx = rng.choice(spec.n_beads, spec.n_beads // 20, replace=False)
    far[idx] += (rng.integers(-3, 4, (idx.size, 3)) * np.array(spec.L)
                 + rng.uniform(-2, 2, (idx.size, 3))).astype(np.float32)
    g.set_positions(far)
    g.reset_stats()
    g.grid_density()
    np.testing.assert_array_equal(g.density(), o.grid(far))
    assert g.stats().grid_fallback > 0
    # steps with reordering every 3 sweeps
    g.update_omega()
    om = g.omega()
    for step in range(1, 8):
        g.step(step)
        o.sweep(far, om, step - 1)
        counts = o.grid(far)
        om = o.omega(counts).astype(np.float32) if step % 10 else om
    np.testing.assert_array_equal(g.density(), counts)
    assert same_bits(g.positions(), far)
    assert g.stats().sorts >= 3
    g.close()


def test_sort_is_invisible():
    banner("sort invisible")
    spec = W.small_diblock_melt()
    g, o = pair(spec, seed=2, sort_every=-1)
    pos = o.init_coords()
    re2 = g.stats().re2
    g.sort_chains()
    assert same_bits(g.positions(), pos)
    assert g.stats().re2 == re2
    g.grid_density()
    g.update_omega()
    om = g.omega()
    g.mc_sweep(1)
    o.sweep(pos, om, 0)
    assert same_bits(g.positions(), pos)
    g.close()


def test_set_rng_replays_a_sweep():
    banner("set_rng replay")
    spec = W.tiny()
    g = scmf.System(spec, make_opts(seed=4))
    g.grid_density()
    g.update_omega()
    start = g.positions()
    g.mc_sweep(1)
    once = g.positions()
    g.set_positions(start)
    g.set_rng(4, 0)
    g.mc_sweep(1)
    assert same_bits(g.positions(), once)
    g.set_positions(start)
    g.set_rng(5, 0)
    g.mc_sweep(1)
    assert not same_bits(g.positions(), once)
    g.close()


def full_size_check(spec, bits, type_sums, sweep_steps=1):
    """Init, gridding, omega and whole sweeps of a full-size system against
    the oracle on the same positions: counts and positions bit-exact, omega
    equal to the oracle's fp64 value rounded once; then size-independent
    invariants (bead sums per type, idempotent gridding, <Re^2> of the walk)."""
    g = scmf.System(spec, make_opts(seed=0, density_bits=bits))
    try:
        o = orc.Oracle(spec, g.opts)
        g.grid_density()
        c = g.density()
        cc = c.reshape(spec.n_types, -1)
        assert cc.sum(dtype=np.int64) == spec.n_beads
        for t, n in enumerate(type_sums):
            assert cc[t].sum(dtype=np.int64) == n
        assert g.stats().overflow == 0
        g.grid_density()                          # idempotent
        assert np.array_equal(g.density(), c)
        assert abs(g.stats().re2 - 1.0) < 0.01    # Gaussian walks: <Re^2> = 1
        pos = g.positions()
        banner("  oracle gridding")
        assert np.array_equal(o.grid(pos), c)
        g.update_omega()
        om = g.omega()
        banner("  oracle omega")
        assert same_bits(om, o.omega(c).astype(np.float32))
        for k in range(sweep_steps):
            banner(f"  oracle sweep {k}")
            acc = o.sweep(pos, om, k)
            g.reset_stats()
            g.mc_sweep(1)
            assert g.stats().accepts == acc
            assert same_bits(g.positions(), pos)
            g.grid_density()
            c = g.density()
            assert np.array_equal(o.grid(pos), c)
            g.update_omega()
            om = g.omega()
    finally:
        g.close()


def test_large_melt_full_size():
    """268M beads, 256^3 grid, uint16 counts, smart moves: one whole step
    bit-exact against the oracle."""
    banner("large diblock melt, full size")
    spec = W.large_diblock_melt()
    full_size_check(spec, 16, [19 * 4_194_304, 45 * 4_194_304])


def test_largest_blend_full_size():
    """1.0e9 beads (A/B homopolymers + AB diblocks), 512x512x256 grid, uint32
    counts, smart moves: one whole step bit-exact against the oracle."""
    banner("largest blend, full size")
    spec = W.largest_blend()
    na = 10_937_500 * 64 + 781_250 * 64
    nb = 3_125_000 * 64 + 781_250 * 64
    assert na + nb == spec.n_beads == 1_000_000_000
    full_size_check(spec, 32, [na, nb])


@pytest.mark.parametrize("sorted_", [False, True])
def test_re2_matches_oracle(sorted_):
    """<Re^2> over chains of three architectures (first to last bead, test.cpp:45-67
    generalised) after a few sweeps, from chains in reference order and from
    spatially reordered slots: equal to the oracle's up to fp64 summation order."""
    banner(f"re2 sorted={sorted_}")
    spec = W.mixed_arch(scale=8)
    g, o = pair(spec, seed=404, sort_every=2 if sorted_ else -1)
    pos = o.init_coords()
    g.grid_density()
    g.update_omega()
    om = g.omega()
    for s in range(4):
        g.mc_sweep(1)
        o.sweep(pos, om, s)
        if sorted_:
            g.sort_chains()
    assert same_bits(g.positions(), pos)
    assert (g.stats().sorts > 0) == sorted_
    want = o.re2(pos)
    got = g.stats().re2
    assert abs(got - want) <= 1e-12 * want, (got, want)
    # a chain's end-to-end vector uses its own first and last bead: moving the
    # last bead of one star (13 beads) changes the mean by exactly its share
    p2 = pos.copy()
    base = spec.polymers[0].count * 24 + spec.polymers[1].count * 20
    p2[base + 12] += np.float32(0.5)
    g.set_positions(p2)
    assert abs(g.stats().re2 - o.re2(p2)) <= 1e-12 * want
    assert g.stats().re2 != got
    g.close()


def test_copy_errors():
    banner("copy errors")
    g = scmf.System(W.tiny(), make_opts(seed=1))
    with pytest.raises(scmf.ScmfError):
        g.set_positions(np.zeros((3, 3), np.float32))
    with pytest.raises(scmf.ScmfError):
        g.set_omega(np.zeros(5, np.float32))
    g.close()
    g16 = scmf.System(W.tiny(), make_opts(seed=1, density_bits=16))
    bad = np.zeros(2 * 512, np.uint32)
    bad[3] = 70000
    with pytest.raises(scmf.ScmfError):
        g16.set_density(bad)
    g16.close()


def test_files_and_driver(golden, tmp_path):
    banner("files + driver")
    spec = W.reference_shipped()
    g = scmf.System(spec, make_opts(seed=3))
    g.grid_density()
    g.update_omega()
    g.write_density(10, str(tmp_path))
    g.write_coord(str(tmp_path / "coord.dat"))
    lines = (tmp_path / "density_10.dat").read_text().splitlines()
    assert lines[0] == "Nx=24, Ny=24, Nz=24" and lines[1] == "dx=0.166667, dy=0.166667, dz=0.166667"
    assert len(lines) == 2 + 24 ** 3
    xyz = np.loadtxt(tmp_path / "coord.dat", dtype=np.float64)
    np.testing.assert_allclose(xyz, g.positions(), rtol=1e-5, atol=1e-5)
    g.close()
    # the C++ driver on the reference's own input files
    inp = os.path.join(golden, "reference_inputs")
    for f in ("configuration.dat", "polymer.dat"):
        (tmp_path / f).write_text(open(os.path.join(inp, f)).read())
    exe = os.path.join(ROOT, "new_arch_amd", "lib", "scmf_run")
    r = subprocess.run([exe, "-Steps_N=20", "-Check_T=10", "-seed=5", "-vtk=1"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "step 10 end to end dis" in r.stdout and "time =" in r.stdout
    assert (tmp_path / "density_20.dat").exists()
    # -vtk=1: density_20.vtk is vtk.c's conversion of density_20.dat
    conv = tmp_path / "conv.vtk"
    assert scmf.lib().scmf_density_to_vtk(str(tmp_path / "density_20.dat").encode(),
                                          str(conv).encode()) == 0
    a = (tmp_path / "density_20.vtk").read_text().split("\n")
    b = conv.read_text().split("\n")
    assert a[1] == "# comment: converted from file density_20.dat"
    assert a[:1] + a[2:] == b[:1] + b[2:]
