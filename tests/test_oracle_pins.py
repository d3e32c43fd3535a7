"""Pins of the CPU oracle against what the reference and mathematics fix.

* NGP binning (device_function.cu:313-360): exact rational arithmetic away from
  cell boundaries, hand-derived answers on them, periodicity, invariants.
* omega update (device_function.cu:373-412): the reference's own density dumps
  (density_240000/300000/999000.dat) — their omega columns are reproduced from
  the integer counts they decode to.
* derived constants (init_config.cu:480-484, 804-809) with N != Nref.
* initial walk (device_function.cu:140-207): closed-form bond variance
  1/(3(N-1)) from the chain's own length, also with N != Nref.
* move (device_function.cu:444-799): single moves of a 3-bead chain computed by
  hand from the physics (bond energy, force-biased Gaussian proposal,
  Metropolis-Hastings ratio) on both sides of the accept threshold, for smart
  and trial moves, with the fp64 new cell across a cell edge (:577) and the
  bead pick (:749-751); relaxation of collapsed chains to the ideal Gaussian
  chain (<b^2> per component = 1/(2 harmonic_normb), harm from Nref, N = Nref
  and N != Nref); dimer detailed balance; Boltzmann occupancy of a two-level
  field for unbonded beads.
Each plausible slip of the restatement (A without /Nref, R = sqrt(A), harm or
the initial walk from the wrong length, a flipped or dropped proposal-ratio
term, a flipped bond energy, the new cell from the fp32 position) fails at
least one of these tests.
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
def run_relax(move, N, seed, sweeps=300, chains=3000, nref=None):
    spec = SystemSpec(n_types=2, xn=W.xn_matrix(2, 0.0, 0.0), L=[2.0, 2.0, 2.0], n=[8, 8, 8],
                      dtNxi=0.17, nref=N if nref is None else nref,
                      polymers=[W.diblock(N, 0.5, chains)])
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


# ---------------------------------------------------------------- single moves by hand
# One chain of N = 3 beads with Nref = 4 (so N != Nref): harm = 3 (Nref - 1) / 2 = 4.5,
# A = dtNxi / Nref = 0.125 and R = sqrt(2 A) = 0.5 (init_config.cu:480-484, 804-809).
# Cells of 1/8 in a box of 2 (16 cells per axis).  The expected values below come from
# the physics of the move, not from the oracle's formulas: the bond energy
# U = harm * sum |r_i - r_j|^2 over bonds, the force F = -grad U, the force-biased
# proposal r' = r + A F(r) + sqrt(2 A) xi (a Gaussian of variance 2A per component
# around the drift), and the Metropolis-Hastings test against
# exp(-(dU_bond + d_omega)) q(r'->r) / q(r->r').
HAND_NREF, HAND_N, HAND_DTNXI = 4, 3, 0.5
HAND_L, HAND_CELLS = 2.0, 16


def hand_spec(move_types=(0, 1, 0)):
    return SystemSpec(n_types=2, xn=W.xn_matrix(2, 30.0, 10.0), L=[HAND_L] * 3,
                      n=[HAND_CELLS] * 3, dtNxi=HAND_DTNXI, nref=HAND_NREF,
                      polymers=[PolymerType(length=HAND_N, count=1, monotype=list(move_types))])


def hand_cell(r):
    """NGP cell of an exactly known point: floor(r / (1/8)) mod 16 per axis,
    index x + 16 y + 256 z (device_function.cuh:19)."""
    i = [int(np.floor(c * HAND_CELLS / HAND_L)) % HAND_CELLS for c in r]
    return i[0] + HAND_CELLS * i[1] + HAND_CELLS * HAND_CELLS * i[2]


def hand_bond_energy(pos, b, r, harm):
    return sum(harm * float(np.sum((r - pos[j]) ** 2)) for j in (b - 1, b + 1)
               if 0 <= j < len(pos))


def hand_force(pos, b, r, harm):
    f = np.zeros(3)
    for j in (b - 1, b + 1):
        if 0 <= j < len(pos):
            f += 2.0 * harm * (pos[j] - r)
    return f


def hand_smart(pos, b, xi, omega_t):
    """(d, dE) of a force-biased move of bead b with normals xi."""
    harm = 3.0 * (HAND_NREF - 1) / 2.0
    A = HAND_DTNXI / HAND_NREF
    r = pos[b]
    F = hand_force(pos, b, r, harm)
    d = A * F + np.sqrt(2.0 * A) * np.asarray(xi, np.float64)
    rn = r + d
    Fn = hand_force(pos, b, rn, harm)
    dU = hand_bond_energy(pos, b, rn, harm) - hand_bond_energy(pos, b, r, harm)
    dW = float(omega_t[hand_cell(rn)]) - float(omega_t[hand_cell(r)])
    log_fwd = -float(np.sum((d - A * F) ** 2)) / (4.0 * A)     # log q(r -> r')
    log_bwd = -float(np.sum((-d - A * Fn) ** 2)) / (4.0 * A)   # log q(r' -> r)
    return d, dU + dW - (log_bwd - log_fwd)


def hand_trial(pos, b, u, delta, omega_t):
    harm = 3.0 * (HAND_NREF - 1) / 2.0
    d = delta * (2.0 * np.asarray(u, np.float64) - 1.0)
    r = pos[b]
    dU = hand_bond_energy(pos, b, r + d, harm) - hand_bond_energy(pos, b, r, harm)
    dW = float(omega_t[hand_cell(r + d)]) - float(omega_t[hand_cell(r)])
    return d, dU + dW


HAND_POS = np.array([[0.5, 0.5, 0.5], [0.75, 0.625, 0.5], [1.0, 0.5, 0.625]], np.float32)


def hand_omega(spec, seed=4):
    C = spec.n_cells
    om = np.random.default_rng(seed).uniform(-0.5, 0.5, 2 * C).astype(np.float32)
    return om


def run_injected(o, pos, omega, picks, draws, accs):
    p = pos.copy()
    acc = o.sweep_injected(p, omega, np.asarray(picks, np.float64),
                           np.asarray(draws, np.float32).ravel(), np.asarray(accs, np.float32))
    return p, acc


def test_derived_constants_use_nref():
    """harm = 3 (Nref - 1) / 2, A = dtNxi / Nref, R = sqrt(2 A) with Nref != N
    (init_config.cu:480-484, 804-809)."""
    o = orc.Oracle(hand_spec())
    assert o.derived.harmonic_normb == 4.5
    assert o.derived.A == 0.125 and o.derived.R == 0.5
    assert o.derived.inverse_refbeads == 0.25


@pytest.mark.parametrize("side", ["accept", "reject"])
def test_smart_move_by_hand(side):
    spec = hand_spec()
    o = orc.Oracle(spec, make_opts(move="smart"))
    omega = hand_omega(spec)
    C = spec.n_cells
    pos64 = HAND_POS.astype(np.float64)
    xi = [0.25, -0.5, 0.125]
    d, dE = hand_smart(pos64, 1, xi, omega[C:])      # bead 1 is of type 1
    # d = A F + R xi = 0.125 (0, -2.25, 1.125) + 0.5 xi, exactly
    np.testing.assert_array_equal(d, [0.125, -0.53125, 0.203125])
    # make the field term push dE well above 0 so that exp(-dE) < 1 decides
    om = omega.copy()
    om[C + hand_cell(pos64[1] + d)] = om[C + hand_cell(pos64[1])] + 1.5
    d, dE = hand_smart(pos64, 1, xi, om[C:])
    assert 0.2 < dE < 5.0, dE
    p = np.exp(-dE)
    u1 = p * (1 - 1e-5) if side == "accept" else p * (1 + 1e-5)
    # moves 2 and 3 displace the end beads by ~3 Re: rejected for any uniform
    big = [6.0, 6.0, 6.0]
    for b in (0, 2):
        assert hand_smart(pos64, b, big, om[C * (b == 1):])[1] > 40
    picks = [0.5, 0.1, 0.9]          # floor(U N): beads 1, 0, 2
    out, acc = run_injected(o, HAND_POS, om, picks, [xi, big, big], [u1, 0.5, 0.5])
    want = HAND_POS.copy()
    if side == "accept":
        want[1] = (pos64[1] + d).astype(np.float32)
    assert acc == (1 if side == "accept" else 0)
    np.testing.assert_array_equal(out, want)


def test_smart_move_drift_and_noise_separately():
    """With xi = 0 the bead moves by the drift A F alone; with a symmetric
    neighbourhood (F = 0) by the noise R xi alone: pins A and R separately."""
    spec = hand_spec()
    o = orc.Oracle(spec, make_opts(move="smart"))
    omega = np.zeros(2 * spec.n_cells, np.float32)
    pos = HAND_POS.copy()
    # end bead 0: F = 2 harm (r1 - r0) = 9 (0.25, 0.125, 0) -> A F = (0.28125, 0.140625, 0)
    out, _ = run_injected(o, pos, omega, [0.0, 0.5, 0.5], [[0, 0, 0], [0, 0, 0], [0, 0, 0]],
                          [1e-30, 1.0, 1.0])
    np.testing.assert_array_equal(out[0], np.float32([0.5 + 0.28125, 0.5 + 0.140625, 0.5]))
    # symmetric middle bead: r1 at the midpoint of r0 and r2, F = 0: d = R xi = xi / 2
    sym = np.array([[0.5, 0.5, 0.5], [0.75, 0.75, 0.75], [1.0, 1.0, 1.0]], np.float32)
    out, _ = run_injected(o, sym, omega, [0.5, 0.0, 0.0], [[0.25, -0.125, 0.0625], [0, 0, 0],
                                                           [0, 0, 0]], [1e-30, 1.0, 1.0])
    np.testing.assert_array_equal(out[1], np.float32([0.875, 0.6875, 0.78125]))


@pytest.mark.parametrize("side", ["accept", "reject"])
def test_trial_move_by_hand(side):
    spec = hand_spec()
    delta = 0.25
    o = orc.Oracle(spec, make_opts(move="trial", trial_delta=delta))
    omega = hand_omega(spec, seed=9)
    C = spec.n_cells
    pos64 = HAND_POS.astype(np.float64)
    u = [0.75, 0.125, 0.5]           # d = delta (2u - 1) = (0.125, -0.1875, 0)
    d, dE = hand_trial(pos64, 2, u, delta, omega[:C])   # bead 2 is of type 0
    np.testing.assert_array_equal(d, [0.125, -0.1875, 0.0])
    om = omega.copy()
    om[hand_cell(pos64[2] + d)] = om[hand_cell(pos64[2])] + 0.75
    d, dE = hand_trial(pos64, 2, u, delta, om[:C])
    assert 0.2 < dE < 5.0, dE
    p = np.exp(-dE)
    u1 = p * (1 - 1e-5) if side == "accept" else p * (1 + 1e-5)
    # moves 2 and 3 draw u = 1/2: zero displacement, dE = 0, always accepted, no change
    out, acc = run_injected(o, HAND_POS, om, [1.0, 0.0, 0.5], [u, [0.5] * 3, [0.5] * 3],
                            [u1, 0.5, 0.5])
    want = HAND_POS.copy()
    if side == "accept":
        want[2] = (pos64[2] + d).astype(np.float32)
    assert acc == (3 if side == "accept" else 2)
    np.testing.assert_array_equal(out, want)


def test_new_cell_from_fp64_sum_across_a_cell_edge():
    """calc_delta_energy_arr takes the new cell from the fp64 sum old + d
    (device_function.cu:577), not from the fp32 position that is stored: with
    x_old = 1 - 2^-24 and d = 3 2^-26 the fp64 sum 1 - 2^-26 lies in cell 7
    while the stored fp32 x rounds to 1.0, in cell 8."""
    spec = hand_spec()
    delta = 2.0 ** -20
    o = orc.Oracle(spec, make_opts(move="trial", trial_delta=delta))
    C = spec.n_cells
    x0 = np.float32(1.0) - np.float32(2.0 ** -24)
    pos = np.array([[0.5, 0.5, 0.5], [0.75, 0.5, 0.5], [x0, 0.5, 0.5]], np.float32)
    u = np.float32(0.5 + 3 * 2.0 ** -7)      # 2u - 1 = 3 2^-6
    assert np.float32(delta) * (np.float32(2) * u - np.float32(1)) == np.float32(3 * 2.0 ** -26)
    assert np.float32(x0 + np.float32(3 * 2.0 ** -26)) == np.float32(1.0)
    cell7 = hand_cell([0.99, 0.5, 0.5])
    cell8 = hand_cell([1.0, 0.5, 0.5])
    assert cell8 == cell7 + 1
    omega = np.zeros(2 * C, np.float32)
    omega[cell8] = 50.0                       # cell 8 costs exp(-50): never accepted
    out, acc = run_injected(o, pos, omega, [1.0, 0.5, 0.5], [[u, 0.5, 0.5], [0.5] * 3,
                                                             [0.5] * 3], [0.5, 0.5, 0.5])
    assert acc == 3
    assert out[2, 0] == np.float32(1.0)       # accepted: its cell was taken as 7
    # the old cell is the stored position's: from x = 1.0 (cell 8, omega 50) a
    # move back into cell 7 gains 50 and is accepted whatever the uniform
    pos2 = out.copy()
    u_back = np.float32(0.5 - 2.0 ** -6)      # d = -2^-25
    out2, acc2 = run_injected(o, pos2, omega, [1.0, 0.5, 0.5], [[u_back, 0.5, 0.5], [0.5] * 3,
                                                               [0.5] * 3], [1.0, 0.5, 0.5])
    assert out2[2, 0] == np.float32(1.0) - np.float32(2.0 ** -25)


def test_pick_is_floor_of_uniform_times_length():
    """ibead = (int)(U N), and U = 1 picks the last bead (device_function.cu:749-751)."""
    spec = hand_spec()
    o = orc.Oracle(spec, make_opts(move="trial", trial_delta=0.25))
    omega = np.zeros(2 * spec.n_cells, np.float32)
    shift = [0.75, 0.5, 0.5]                  # d = (0.125, 0, 0)
    still = [0.5, 0.5, 0.5]
    for pick, bead in [(0.0, 0), (0.3333, 0), (0.3334, 1), (0.6666, 1), (0.6667, 2), (1.0, 2)]:
        out, _ = run_injected(o, HAND_POS, omega, [pick, 0.5, 0.5], [shift, still, still],
                              [1e-30, 1.0, 1.0])
        moved = np.nonzero(np.any(out != HAND_POS, axis=1))[0].tolist()
        assert moved == [bead], (pick, moved)


# ---------------------------------------------------------------- N != Nref closed forms
def test_init_walk_uses_chain_length_not_nref():
    """initialize_coord scales the walk by 1/sqrt(3 (N - 1)) with N the chain's
    own length (device_function.cu:169-171), whatever Nref is: bond variance
    1/45 for N = 16 (not 1/93 from Nref = 32), <Re^2> = 1."""
    spec = SystemSpec(n_types=2, xn=W.xn_matrix(2, 50.0, 20.0), L=[2.0] * 3, n=[8] * 3,
                      dtNxi=0.17, nref=32, polymers=[W.diblock(16, 0.5, 20000)])
    o = orc.Oracle(spec, seed=12)
    pos = o.init_coords()
    bonds = np.diff(pos.reshape(-1, 16, 3).astype(np.float64), axis=1).reshape(-1, 3)
    np.testing.assert_allclose(bonds.var(axis=0), 1 / 45, rtol=0.02)
    assert abs(o.re2(pos) - 1.0) < 0.03


@pytest.mark.parametrize("move", ["smart", "trial"])
def test_relaxation_with_nref_not_chain_length(move):
    """The bonds equilibrate to variance 1/(2 harm) per component with harm from
    Nref (init_config.cu:480-484): for N = 8, Nref = 16, 1/45 per component
    and <Re^2> = 3 (N - 1)/45 = 7/15 (harm from N would give 1/21 and 1)."""
    re2, b2, acc, _ = run_relax(move, 8, seed=23, nref=16)
    assert abs(b2 - 1 / 45) < 0.001, b2
    assert abs(re2 - 7 / 15) < 0.03, re2
    assert 0.05 < acc < 0.99


@pytest.mark.parametrize("move", ["smart", "trial"])
def test_dimer_detailed_balance(move):
    """Two-bead chains in omega = 0: the bond vector's stationary density is
    exp(-harm |b|^2), so 2 harm |b|^2 follows chi-square with 3 degrees of
    freedom.  Ten bins of equal chi-square probability each hold 0.1 of the
    samples.  The force-biased move reaches it only with the proposal ratio
    q(r'->r)/q(r->r') in its acceptance: the step below (A 2 harm ~ 0.5)
    makes the drift large enough that a dropped or sign-flipped ratio term
    moves the bins far outside the tolerance."""
    from scipy.stats import chi2

    nref, chains = 8, 20000
    harm = 3 * (nref - 1) / 2
    spec = SystemSpec(n_types=1, xn=[0.0], L=[2.0] * 3, n=[8] * 3, dtNxi=0.2, nref=nref,
                      polymers=[PolymerType(length=2, count=chains, monotype=[0, 0])])
    o = orc.Oracle(spec, make_opts(seed=31, move=move, trial_delta=0.2))
    pos = o.init_coords()
    omega = np.zeros(spec.n_cells, np.float32)
    edges = chi2.ppf(np.linspace(0, 1, 11), 3)
    hist = np.zeros(10)
    for s in range(400):
        o.sweep(pos, omega, s)
        if s >= 100 and s % 5 == 0:
            b = np.diff(pos.reshape(chains, 2, 3).astype(np.float64), axis=1).reshape(-1, 3)
            hist += np.histogram(2 * harm * (b * b).sum(1), edges)[0]
    frac = hist / hist.sum()
    assert np.abs(frac - 0.1).max() < 0.006, frac
