"""Pins of the CPU oracle against what the reference and mathematics fix.

* NGP binning (device_function.cu:313-360): exact rational arithmetic away from
  cell boundaries, hand-derived answers on them, periodicity, invariants.
* omega update (device_function.cu:373-412): the reference's own density dumps
  (density_240000/300000/999000.dat) — their omega columns are reproduced from
  the integer counts they decode to.
* initial walk (device_function.cu:140-207): closed-form bond variance 1/(3(N-1)).
* move (device_function.cu:444-799): relaxation of collapsed chains to the ideal
  Gaussian chain (<Re^2> = 1, <b^2> per component = 1/(2 harmonic_normb)), and
  Boltzmann occupancy of a two-level field for unbonded beads; both fail if an
  energy term has the wrong sign or the wrong cell.
"""
import os
from fractions import Fraction

import numpy as np
import pytest

from new_arch_amd import workloads as W
from new_arch_amd.abi import PolymerType, SystemSpec, make_opts
from oracle import oracle as orc


# ---------------------------------------------------------------- binning
def exact_cell(r, L, d, n):
    rf, Lf, df = Fraction(float(r)), Fraction(float(L)), Fraction(float(d))
    m = (rf / Lf).__floor__()
    t = rf - m * Lf
    q = t / df
    return int(q.__floor__()) % n, float(q - q.__floor__()), float(rf / Lf - (rf / Lf).__floor__())


@pytest.mark.parametrize("L,n", [(4.0, 24), (4.0, 32), (32.0, 256), (1.0, 6), (3.0, 12)])
def test_binning_matches_exact_arithmetic(L, n):
    spec = W.tiny(grid=(n, n, n), L=(L, L, L))
    o = orc.Oracle(spec)
    d = o.derived.dx
    rng = np.random.default_rng(n)
    r = rng.uniform(-3 * L, 3 * L, 3000).astype(np.float32).astype(np.float64)
    pts = np.stack([r, np.zeros_like(r), np.zeros_like(r)], 1)
    got = o.bin_many(pts, np.zeros(r.size, np.uint32))
    checked = 0
    for ri, g in zip(r, got):
        xi, fq, fs = exact_cell(ri, L, d, n)
        if min(fq, 1 - fq) < 1e-9 or min(fs, 1 - fs) < 1e-12:
            continue
        assert int(g) == xi, (ri, g, xi)
        checked += 1
    assert checked > 2900


def test_binning_boundaries_and_layout():
    spec = W.tiny(grid=(24, 24, 24), L=(4.0, 4.0, 4.0))
    o = orc.Oracle(spec)
    C = 24 ** 3
    b = lambda x, y, z, t=0: orc.lib().orc_bin_index(o.desc, x, y, z, t)  # noqa: E731
    assert b(0.0, 0.0, 0.0) == 0
    assert b(4.0, 0.0, 0.0) == 0                       # r = L folds to 0
    assert b(-4.0, 0.0, 0.0) == 0                      # r < 0 branch, t = L: n % n
    assert b(-1e-7, 0.0, 0.0) == 23                    # just below 0 -> last cell
    # largest float32 below L: last cell
    assert b(float(np.nextafter(np.float32(4.0), np.float32(0))), 0.0, 0.0) == 23
    # largest double below L: dx = fl(4/24) is below 1/6, the quotient rounds to
    # 24 and the reference's "% n" wraps it to cell 0
    assert b(np.nextafter(4.0, 0), 0.0, 0.0) == 0
    assert b(0.5, 0.0, 0.0) == 3                       # 0.5 / (1/6) = 3 exactly
    assert b(0.0, 0.5, 0.0) == 3 * 24                  # y stride nx
    assert b(0.0, 0.0, 0.5) == 3 * 24 * 24             # z stride nx*ny
    assert b(0.1, 0.2, 0.3, 1) == C + 0 + 1 * 24 + 1 * 576
    # periodic images on exactly representable coordinates
    rng = np.random.default_rng(5)
    for _ in range(200):
        x, y, z = (rng.integers(0, 4096, 3) / 1024.0)
        k = rng.integers(-5, 6, 3)
        assert b(x + 4 * k[0], y + 4 * k[1], z + 4 * k[2]) == b(x, y, z)


def test_grid_invariants():
    spec = W.mixed_arch()
    o = orc.Oracle(spec, seed=3)
    pos = o.init_coords()
    counts = o.grid(pos)
    assert counts.sum() == spec.n_beads
    per_type = counts.reshape(spec.n_types, -1).sum(1)
    want = np.bincount(spec.bead_types(), minlength=spec.n_types)
    np.testing.assert_array_equal(per_type, want)
    # order independence: permuting chains inside the arrays changes nothing
    perm = np.random.default_rng(0).permutation(spec.polymers[0].count)
    p2 = pos.copy()
    N0 = spec.polymers[0].length
    blk = p2[: spec.polymers[0].count * N0].reshape(-1, N0, 3)
    p2[: spec.polymers[0].count * N0] = blk[perm].reshape(-1, 3)
    np.testing.assert_array_equal(o.grid(p2), counts)


# ---------------------------------------------------------------- omega
def dump_spec(grid, L, kappa, chi):
    """A 3-type, 409,600-bead system with the reference dumps' field scaling."""
    poly = PolymerType(length=32, count=12800, monotype=[0] * 4 + [1] * 24 + [2] * 4)
    return SystemSpec(n_types=3, xn=W.xn_matrix(3, kappa, chi), L=list(L), n=list(grid),
                      dtNxi=0.17, nref=32, polymers=[poly])


@pytest.mark.parametrize("step", [240000, 300000, 999000])
def test_omega_reproduces_reference_dumps(golden, step):
    f = np.load(os.path.join(golden, "density_fixtures.npz"))
    grid = tuple(int(v) for v in f[f"grid_{step}"])
    kappa, chi = f[f"kappa_chi_{step}"]
    L = [g / 6.0 for g in grid]  # dx = 1/6 Re in every dump
    o = orc.Oracle(dump_spec(grid, L, kappa, chi))
    prev = f[f"prev_{step}"].astype(np.uint32)         # [cell][type]
    om = o.omega(np.ascontiguousarray(prev.T))          # [type][cell]
    printed = f[f"omega_{step}"].astype(np.float64)     # [cell][type], %g values
    got = om.reshape(3, -1).T
    assert np.all(np.abs(got - printed) <= 6e-6 * np.abs(printed) + 2e-6)
    # and a dropped interaction term would not: it moves omega by O(chiN/N)
    assert np.abs(got - printed).max() < 1e-4


def test_omega_closed_forms():
    spec = W.tiny(grid=(4, 4, 4), L=(1.0, 1.0, 1.0), kappaN=50.0, chiN=20.0, n_chains=64, N=16)
    o = orc.Oracle(spec)
    C = 64
    # uniform melt at the mean density, half A half B: omega = 0 exactly
    per = spec.n_beads // C // 2
    counts = np.full(2 * C, per, np.uint32)
    np.testing.assert_array_equal(o.omega(counts), 0.0)
    # pure A cell at twice the mean density:
    # omega_A = kappa/N (2 - 1) - chi/2N (2 - 0);  omega_B = kappa/N (2 - 1) - chi/2N (0 - 2)
    counts = np.zeros(2 * C, np.uint32)
    counts[:C] = 2 * per * 2
    om = o.omega(counts).reshape(2, C)
    np.testing.assert_allclose(om[0], 50 / 16 - 20 / 32 * 2, rtol=1e-14)
    np.testing.assert_allclose(om[1], 50 / 16 + 20 / 32 * 2, rtol=1e-14)


# ---------------------------------------------------------------- initial walk
def test_init_walk_statistics():
    spec = W.tiny(n_chains=20000, N=16, grid=(8, 8, 8), L=(2.0, 2.0, 2.0))
    o = orc.Oracle(spec, seed=11)
    pos = o.init_coords().reshape(-1, 16, 3).astype(np.float64)
    bonds = np.diff(pos, axis=1).reshape(-1, 3)
    var = bonds.var(axis=0)
    np.testing.assert_allclose(var, 1 / (3 * 15), rtol=0.02)
    assert abs(bonds.mean()) < 0.002
    starts = pos[:, 0, :]
    assert starts.min() > 0 and starts.max() <= 2.0
    np.testing.assert_allclose(starts.mean(0), 1.0, atol=0.02)
    re2 = o.re2(pos.astype(np.float32).reshape(-1, 3))
    assert abs(re2 - 1.0) < 0.03


def test_init_walk_branched():
    spec = W.mixed_arch()
    o = orc.Oracle(spec, seed=1)
    pos = o.init_coords()
    assert np.isfinite(pos).all()
    # every bonded pair of the star is at a Gaussian distance, never a copy
    star = spec.polymers[2]
    base = spec.polymers[0].count * 24 + spec.polymers[1].count * 20
    p = pos[base: base + star.count * star.length].reshape(star.count, star.length, 3)
    i, j = np.nonzero(np.triu(star.bond_matrix()))
    d = np.linalg.norm(p[:, i] - p[:, j], axis=2)
    assert (d > 0).all() and d.mean() < 1.0


# ---------------------------------------------------------------- moves
def run_relax(move, N, seed, sweeps=300, chains=3000):
    spec = SystemSpec(n_types=2, xn=W.xn_matrix(2, 0.0, 0.0), L=[2.0, 2.0, 2.0], n=[8, 8, 8],
                      dtNxi=0.17, nref=N, polymers=[W.diblock(N, 0.5, chains)])
    o = orc.Oracle(spec, make_opts(seed=seed, move=move, trial_delta=0.15))
    pos = o.init_coords()
    # collapse every chain onto its first bead: <Re^2> starts at 0
    p = pos.reshape(chains, N, 3)
    p[:] = p[:, :1, :]
    omega = np.zeros(2 * spec.n_cells, np.float32)
    re2, b2, acc = [], [], 0
    for s in range(sweeps):
        acc += o.sweep(pos, omega, s)
        if s >= sweeps // 2:
            re2.append(o.re2(pos))
            b2.append((np.diff(pos.reshape(chains, N, 3).astype(np.float64), axis=1) ** 2).mean())
    return np.mean(re2), np.mean(b2), acc / (sweeps * spec.n_beads), spec


@pytest.mark.parametrize("move", ["smart", "trial"])
def test_moves_relax_to_ideal_chain(move):
    N = 8
    re2, b2, acc, spec = run_relax(move, N, seed=21)
    harm = 3 * (N - 1) / 2
    assert abs(re2 - 1.0) < 0.06, re2
    assert abs(b2 - 1 / (2 * harm)) < 0.006, b2
    assert 0.05 < acc < 0.99


def test_field_boltzmann_occupancy():
    """Unbonded beads in omega = +1 (x < L/2) / 0 (x >= L/2): P(left) = 1/(1+e)."""
    beads = 20000
    spec = SystemSpec(n_types=1, xn=[10.0], L=[2.0, 2.0, 2.0], n=[2, 1, 1], dtNxi=0.17, nref=2,
                      polymers=[PolymerType(length=1, count=beads, monotype=[0],
                                            bonds=np.zeros((1, 1), np.uint8))])
    o = orc.Oracle(spec, make_opts(seed=5, move="trial", trial_delta=0.4))
    pos = o.init_coords()
    omega = np.array([1.0, 0.0], np.float32)  # cell 0 = x in [0,1), cell 1 = [1,2)
    frac = []
    for s in range(120):
        o.sweep(pos, omega, s)
        if s >= 40:
            x = np.mod(pos[:, 0].astype(np.float64), 2.0)
            frac.append((x < 1.0).mean())
    want = np.exp(-1.0) / (1 + np.exp(-1.0))
    assert abs(np.mean(frac) - want) < 0.01, (np.mean(frac), want)
