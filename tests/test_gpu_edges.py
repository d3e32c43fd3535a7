"""Edge cases of the HIP engine against the oracle, bit for bit: chains longer
than 255 beads (16-bit schedule indices), one-, two- and three-bead chains,
architectures with no chains, one and eight monomer types, tiled gridding
with cells that are not a power of two or architectures longer than its type
table, and a cell holding more beads than a 16-bit counter (the window
normalisation of the tiled gridding)."""
import numpy as np
import pytest

from new_arch_amd import scmf
from new_arch_amd import workloads as W
from new_arch_amd.abi import PolymerType, SystemSpec, make_opts
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-edges] {msg}", flush=True)


def same_bits(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape and a.dtype == b.dtype
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def run_pair(spec, sweeps=2, **kw):
    """init, grid, omega, then sweeps with a grid + omega after each: every
    stage bit-exact against the oracle."""
    opts = make_opts(**kw)
    g = scmf.System(spec, opts)
    try:
        o = orc.Oracle(spec, opts)
        pos = o.init_coords()
        assert same_bits(g.positions(), pos)
        g.grid_density()
        c = g.density()
        assert np.array_equal(c, o.grid(pos))
        g.update_omega()
        om = g.omega()
        assert same_bits(om, o.omega(c).astype(np.float32))
        for k in range(sweeps):
            g.reset_stats()
            acc = o.sweep(pos, om, k)
            g.mc_sweep(1)
            assert g.stats().accepts == acc
            assert same_bits(g.positions(), pos)
            g.grid_density()
            c = g.density()
            assert np.array_equal(c, o.grid(pos))
            g.update_omega()
            om = g.omega()
            assert same_bits(om, o.omega(c).astype(np.float32))
    finally:
        g.close()


def long_chain_spec():
    return SystemSpec(n_types=2, xn=W.xn_matrix(2, 50.0, 20.0), L=[3.0, 3.0, 3.0], n=[12, 12, 12],
                      dtNxi=0.17, nref=300,
                      polymers=[W.diblock(300, 0.4, 40), W.homopolymer(20, 100, 1)])


@pytest.mark.parametrize("lanes", [0, -1, 4, 8])
def test_chains_longer_than_255(lanes):
    banner(f"N=300, lanes {lanes}")
    run_pair(long_chain_spec(), seed=11, lanes_per_chain=lanes, density_bits=32)


def test_long_chains_reject_shapes_that_do_not_fit():
    with pytest.raises(scmf.ScmfError):
        scmf.System(long_chain_spec(), make_opts(seed=1, lanes_per_chain=1))


def short_chain_spec():
    return SystemSpec(n_types=2, xn=W.xn_matrix(2, 40.0, 25.0), L=[2.0, 2.0, 2.0], n=[8, 8, 8],
                      dtNxi=0.3, nref=16,
                      polymers=[W.homopolymer(1, 200, 0), W.diblock(2, 0.5, 150),
                                PolymerType(length=3, count=100, monotype=[1, 0, 1]),
                                W.homopolymer(7, 0, 1),  # an architecture with no chains
                                W.diblock(16, 0.5, 50)])


@pytest.mark.parametrize("move", ["smart", "trial"])
@pytest.mark.parametrize("lanes", [0, -1, 1, 8])
def test_one_two_three_bead_chains(move, lanes):
    banner(f"short chains {move} lanes {lanes}")
    run_pair(short_chain_spec(), seed=12, move=move, trial_delta=0.15, lanes_per_chain=lanes)


@pytest.mark.parametrize("T", [1, 8])
def test_one_and_eight_monomer_types(T):
    banner(f"T={T}")
    poly = PolymerType(length=24, count=300, monotype=[i % T for i in range(24)])
    spec = SystemSpec(n_types=T, xn=W.xn_matrix(T, 45.0, 18.0), L=[2.5, 2.0, 3.0],
                      n=[10, 8, 12], dtNxi=0.2, nref=24, polymers=[poly])
    run_pair(spec, seed=13, density_bits=16)


@pytest.mark.parametrize("bits", [16, 32])
def test_tiled_gridding_general_cells(bits):
    """Cells of 0.1 Re (not a power of two): the tiled gridding takes the
    general binning path; uneven tiles (two per axis, one cell longer) with
    halos on every axis; ordering every 2 sweeps."""
    banner(f"tiled, 0.1 Re cells, u{bits}")
    grid = (31, 27, 35)
    spec = SystemSpec(n_types=2, xn=W.xn_matrix(2, 40.0, 15.0), L=[g * 0.1 for g in grid],
                      n=list(grid), dtNxi=0.1, nref=16,
                      polymers=[PolymerType(length=8, count=3000,
                                            monotype=[t % 2 for t in range(8)])])
    run_pair(spec, sweeps=4, seed=14, density_bits=bits, sort_every=2)


@pytest.mark.parametrize("bits,crowd", [(16, 0), (32, 0), (32, 1)])
def test_crowded_cell(bits, crowd):
    """70,000 single beads of one type in one cell: the tile holds more than
    65,535 beads, so its window is normalised while binning (for two types
    both counters of a cell share a word: type 1 is the high half).  uint32
    counts stay exact; uint16 counts raise the overflow flag."""
    banner(f"crowded cell u{bits} type {crowd}")
    many, few = W.homopolymer(1, 70_000, crowd), W.homopolymer(1, 5_000, 1 - crowd)
    spec = SystemSpec(n_types=2, xn=W.xn_matrix(2, 40.0, 15.0), L=[2.0, 2.0, 2.0],
                      n=[16, 16, 16], dtNxi=0.1, nref=16,
                      polymers=[many, few] if crowd == 0 else [few, many])
    g = scmf.System(spec, make_opts(seed=15, density_bits=bits, sort_every=1))
    try:
        o = orc.Oracle(spec, g.opts)
        pos = o.init_coords()
        first = 0 if crowd == 0 else 5_000
        pos[first:first + 70_000] = np.array([0.51, 1.26, 0.77], np.float32)  # one cell
        g.set_positions(pos)
        g.grid_density()
        want = o.grid(pos)
        assert want.max() == 70_000
        if bits == 32:
            assert np.array_equal(g.density(), want)
            assert g.stats().overflow == 0
        else:
            assert g.stats().overflow == 1
        assert g.stats().sorts >= 1
    finally:
        g.close()


@pytest.mark.parametrize("bits", [16, 32])
def test_crowded_tile_without_wrap(bits):
    """40,000 beads of type 0 and 30,000 of type 1 in one cell (both halves
    of one window word): a tile of more than 65,535 beads whose halves pass
    0x8000 but never wrap, so the window check passes without a second
    binning.  Both count widths exact, no overflow."""
    banner(f"crowded tile without wrap u{bits}")
    a, b = W.homopolymer(1, 40_000, 0), W.homopolymer(1, 30_000, 1)
    spec = SystemSpec(n_types=2, xn=W.xn_matrix(2, 40.0, 15.0), L=[2.0, 2.0, 2.0],
                      n=[16, 16, 16], dtNxi=0.1, nref=16, polymers=[a, b])
    g = scmf.System(spec, make_opts(seed=16, density_bits=bits, sort_every=1))
    try:
        o = orc.Oracle(spec, g.opts)
        pos = o.init_coords()
        pos[:] = np.array([0.51, 1.26, 0.77], np.float32)
        g.set_positions(pos)
        g.grid_density()
        want = o.grid(pos)
        assert want.max() == 40_000
        assert np.array_equal(g.density(), want)
        assert g.stats().overflow == 0
        assert g.stats().sorts >= 1
    finally:
        g.close()


@pytest.mark.parametrize("half", [0, 1])
@pytest.mark.parametrize("n", [65_535, 65_536])
def test_atomic_gridding_uint16_boundary(n, half):
    """The atomic gridding (one atomic per distinct count word of a wave, both
    uint16 halves of the word at once): n type-0 beads in one cell and 3,000 in
    the cell beside it along x, whose count is the other half of the same
    32-bit word (cells x = 4 and 5: the low and the high half), beside 2,000
    type-1 beads spread out.  65,535 in one half is exact with no overflow;
    65,536 raises the flag, in either half."""
    banner(f"atomic gridding boundary n={n} half {half}")
    spec = SystemSpec(n_types=2, xn=W.xn_matrix(2, 40.0, 15.0), L=[2.0, 2.0, 2.0],
                      n=[16, 16, 16], dtNxi=0.1, nref=16,
                      polymers=[W.homopolymer(1, n + 3_000, 0), W.homopolymer(1, 2_000, 1)])
    g = scmf.System(spec, make_opts(seed=16, density_bits=16, sort_every=-1))
    try:
        o = orc.Oracle(spec, g.opts)
        pos = o.init_coords()
        low = np.array([0.51, 1.26, 0.77], np.float32)    # x cell 4
        high = np.array([0.70, 1.26, 0.77], np.float32)   # x cell 5, same word
        pos[:n] = high if half else low
        pos[n:n + 3_000] = low if half else high
        g.set_positions(pos)
        g.grid_density()
        want = o.grid(pos)
        c = 4 + 10 * 16 + 6 * 256
        assert want[c + half] == n and want[c + 1 - half] == 3_000
        if n <= 65_535:
            assert g.stats().overflow == 0
            np.testing.assert_array_equal(g.density(), want)
        else:
            assert g.stats().overflow == 1
        assert g.stats().sorts == 0
    finally:
        g.close()


@pytest.mark.parametrize("bits", [16, 32])
def test_tiled_gridding_architecture_longer_than_type_table(bits):
    """An architecture longer than the tiled gridding's LDS type table (4096
    monomers) reads its types from global memory, a short one beside it from
    the table; both in the same tiles, two monomer types (paired counters)."""
    banner(f"tiled, N=4500 beside N=32, u{bits}")
    N = 4500
    spec = SystemSpec(n_types=2, xn=W.xn_matrix(2, 40.0, 15.0), L=[3.0, 3.0, 3.0],
                      n=[24, 24, 24], dtNxi=0.1, nref=N,
                      polymers=[PolymerType(length=N, count=24, monotype=[(i // 9) % 2 for i in range(N)]),
                                W.diblock(32, 0.5, 2000)])
    # (the initial walk stops at 4096 beads a chain: positions from a random walk here)
    rng = np.random.default_rng(16)
    chains = [(N, 24), (32, 2000)]
    pos = []
    for n, count in chains:
        start = rng.uniform(0.0, 3.0, (count, 1, 3))
        steps = rng.normal(0.0, (3.0 * (N - 1)) ** -0.5, (count, n - 1, 3))
        walk = np.concatenate([start, start + np.cumsum(steps, axis=1)], axis=1)
        pos.append(walk.reshape(-1, 3))
    pos = np.ascontiguousarray(np.concatenate(pos).astype(np.float32))
    g = scmf.System(spec, make_opts(seed=16, density_bits=bits, sort_every=1, init_coords=False))
    try:
        o = orc.Oracle(spec, g.opts)
        g.set_positions(pos)
        g.grid_density()
        c = g.density()
        assert np.array_equal(c, o.grid(pos))
        assert g.stats().sorts >= 1                 # the tiled path
        g.update_omega()
        om = g.omega()
        acc = o.sweep(pos, om, 0)
        g.reset_stats()
        g.mc_sweep(1)
        assert g.stats().accepts == acc
        assert same_bits(g.positions(), pos)
        g.grid_density()
        assert np.array_equal(g.density(), o.grid(pos))
    finally:
        g.close()


@pytest.mark.parametrize("L,lanes", [(4.0, 0), (3.3, 8)])
def test_gridding_follows_every_position_change(tmp_path, L, lanes):
    """A tiled system grids from the bead cells its last sweep stored, also
    after a spatial ordering (the cells move with the positions); every other
    change of positions (set_positions, checkpoint restore, autotune) makes it
    bin positions again, and so does an ordering of positions whose cells
    are stale.  Counts equal the oracle's after each, and cell_grids counts
    the griddings that read cells.  L = 4 gives power-of-two cells, L = 3.3
    the general binning in the sweep."""
    banner(f"cells follow positions, L={L}")
    spec = W.tiny(n_chains=3000, N=32, grid=(32, 32, 32), L=(L, L, L))
    g = scmf.System(spec, make_opts(seed=17, density_bits=32, sort_every=1,
                                    lanes_per_chain=lanes))
    try:
        o = orc.Oracle(spec, g.opts)

        def check(cell_grids):
            g.grid_density()
            assert np.array_equal(g.density(), o.grid(g.positions()))
            assert g.stats().cell_grids == cell_grids

        check(0)                                   # initial coordinates
        g.update_omega()
        g.step(1)                                  # order, sweep, grid from cells
        assert np.array_equal(g.density(), o.grid(g.positions()))
        assert g.stats().cell_grids == 1
        g.mc_sweep(1)
        check(2)
        pos = g.positions()
        pos[::7] += np.float32(0.37)               # some beads leave their cells
        g.set_positions(pos)
        g.sort_chains()                            # stale cells: not carried
        check(2)
        g.mc_sweep(1)
        g.sort_chains()                            # slots permuted, cells with them
        check(3)
        g.mc_sweep(1)
        ck = tmp_path / "c.ckpt"
        g.checkpoint(ck, 3)
        g.mc_sweep(2)
        assert g.restore(ck) == 3
        check(3)
        g.mc_sweep(1)
        tuned, _ = g.autotune(1)                   # tries every shape, restores positions
        check(3)
        g.mc_sweep(1)
        check(4 if tuned >= 2 else 3)              # only the scheduled shapes store cells
    finally:
        g.close()


@pytest.mark.parametrize("N,lanes", [(96, 2), (96, 8), (96, 16), (128, 2), (128, 8), (160, 2),
                                     (160, 8)])
def test_schedule_kernel_long_chains(N, lanes):
    """70,000 diblocks of N beads, rounds of 2, 8 or 16 moves.  Chains of 65-128
    beads take the register-count schedule kernel with two 64-bit words per
    bit-plane (rounds of 2 moves reach past round 64 there); at 8 lanes a
    chain they are swept at 16 lanes (rounds of 16 moves, five bit-planes),
    which keeps more waves resident; longer chains take the kernel that counts
    each round's moves in LDS (k_schedule_lds, records staged in LDS).  Sweeps
    bit-exact against the oracle."""
    banner(f"schedule kernel, N={N}, lanes {lanes}")
    spec = W.tiny(n_chains=70_000, N=N, grid=(32, 32, 32), L=(16.0, 16.0, 16.0))
    run_pair(spec, sweeps=2, seed=31, density_bits=32, lanes_per_chain=lanes)


@pytest.mark.parametrize("long_first", [False, True])
def test_blend_runs_at_their_own_lanes(long_first):
    """A blend of 64-bead homopolymers and 128-bead diblocks (70,000 chains) at
    8 lanes a chain: the 64-bead run sweeps at 8 lanes, the 128-bead run at 16
    (rounds of 16 moves, its own schedule launch), in either order of the runs.
    Sweeps bit-exact against the oracle."""
    banner(f"blend runs, 128-bead run first: {long_first}")
    homo = [W.homopolymer(64, 35_000, 0), W.homopolymer(64, 25_000, 1)]
    dib = [W.diblock(128, 0.5, 10_000)]
    spec = SystemSpec(n_types=2, xn=W.xn_matrix(2, 50.0, 15.0), L=[16.0, 16.0, 16.0],
                      n=[32, 32, 32], dtNxi=0.05, nref=64,
                      polymers=dib + homo if long_first else homo + dib)
    run_pair(spec, sweeps=2, seed=37, density_bits=32, lanes_per_chain=8, move="smart")


@pytest.mark.parametrize("lanes", [0, 4, 8, 16])
def test_schedule_kernel_with_branched_chains(lanes):
    """66,000 chains (homopolymers, diblocks and 4-arm stars): enough chains
    that the move schedules come from the schedule kernels (register round
    counts: chains of at most 64 beads), whose general neighbour scan the
    stars take.  Sweeps bit-exact against the oracle."""
    banner(f"schedule kernel, stars, lanes {lanes}")
    run_pair(W.mixed_arch(scale=550), sweeps=2, seed=16, density_bits=32, lanes_per_chain=lanes)
