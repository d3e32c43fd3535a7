"""Synthetic systems: the named benchmark configurations and small test systems.

Lengths are in units of the reference end-to-end distance Re; grids use cells
of Re/8 unless a workload says otherwise.  The smart-MC step follows the
reference's derived constant A = (dt*N/xi) / Nref (init_config.cu:804-809);
"step 0.05" in the large-system description is read as dt*N/xi = 0.05.
"""
from typing import List

import numpy as np

from .abi import PolymerType, SystemSpec


def xn_matrix(n_types: int, kappaN: float, chiN: float) -> List[float]:
    """Flat xn[t + T*s]: kappaN on the diagonal, chiN off it (struct.h:169-175)."""
    m = np.full((n_types, n_types), float(chiN))
    np.fill_diagonal(m, float(kappaN))
    return [float(v) for v in m.ravel()]


def diblock(N: int, fA: float, count: int, a_type: int = 0, b_type: int = 1) -> PolymerType:
    na = int(round(N * fA))
    return PolymerType(length=N, count=count, monotype=[a_type] * na + [b_type] * (N - na))


def homopolymer(N: int, count: int, t: int) -> PolymerType:
    return PolymerType(length=N, count=count, monotype=[t] * N)


def small_diblock_melt() -> SystemSpec:
    """Symmetric diblock melt, N=32, 8192 chains (262,144 beads), 32^3 cells of
    Re/8, chiN=20, kappaN=50; run with trial moves of +-0.1 Re and omega after
    every sweep."""
    return SystemSpec(n_types=2, xn=xn_matrix(2, 50.0, 20.0), L=[4.0, 4.0, 4.0], n=[32, 32, 32],
                      dtNxi=0.17, nref=32, polymers=[diblock(32, 0.5, 8192)])


def large_diblock_melt() -> SystemSpec:
    """Cylinder-forming diblock melt, N=64, fA=0.3, 4,194,304 chains (268M beads),
    256^3 cells of Re/8, chiN=25, kappaN=70, smart-MC moves with dt*N/xi = 0.05."""
    return SystemSpec(n_types=2, xn=xn_matrix(2, 70.0, 25.0), L=[32.0, 32.0, 32.0],
                      n=[256, 256, 256], dtNxi=0.05, nref=64,
                      polymers=[diblock(64, 0.3, 4_194_304)])


def largest_blend() -> SystemSpec:
    """A/B homopolymer blend + AB diblock compatibiliser, 70/20/10 by beads,
    N_A = N_B = 64, N_AB = 128, 1.0e9 beads, 512x512x256 cells of Re/8,
    chiN=15, kappaN=50, uint32 counts, smart-MC moves."""
    return SystemSpec(n_types=2, xn=xn_matrix(2, 50.0, 15.0), L=[64.0, 64.0, 32.0],
                      n=[512, 512, 256], dtNxi=0.05, nref=64,
                      polymers=[homopolymer(64, 10_937_500, 0), homopolymer(64, 3_125_000, 1),
                                diblock(128, 0.5, 781_250)])


def reference_shipped() -> SystemSpec:
    """The reference's shipped input (configuration.dat + polymer.dat): 12,800
    AB diblocks of N=32 (16 type-1 beads then 16 type-0 beads), 4^3 Re box,
    24^3 grid, kappaN=50, chiN=30, dt*N/xi=0.17."""
    return SystemSpec(n_types=2, xn=xn_matrix(2, 50.0, 30.0), L=[4.0, 4.0, 4.0], n=[24, 24, 24],
                      dtNxi=0.17, nref=32,
                      polymers=[PolymerType(length=32, count=12800, monotype=[1] * 16 + [0] * 16)])


def tiny(n_chains=64, N=16, grid=(8, 8, 8), L=(2.0, 2.0, 2.0), chiN=20.0, kappaN=50.0,
         dtNxi=0.17, fA=0.5) -> SystemSpec:
    return SystemSpec(n_types=2, xn=xn_matrix(2, kappaN, chiN), L=list(L), n=list(grid),
                      dtNxi=dtNxi, nref=N, polymers=[diblock(N, fA, n_chains)])


def mixed_arch(scale=1) -> SystemSpec:
    """Three architectures of different length, a branched one among them, for
    indexing tests: 3 monomer types, homopolymer + diblock + 4-arm star."""
    star_n = 13
    star = np.zeros((star_n, star_n), np.uint8)
    # centre bead 0; four arms of three beads: 0-1-2-3, 0-4-5-6, 0-7-8-9, 0-10-11-12
    for arm in range(4):
        prev = 0
        for k in range(3):
            b = 1 + arm * 3 + k
            star[prev, b] = star[b, prev] = 1
            prev = b
    return SystemSpec(
        n_types=3, xn=xn_matrix(3, 40.0, 12.0), L=[3.0, 2.5, 2.0], n=[12, 10, 9],
        dtNxi=0.2, nref=24,
        polymers=[homopolymer(24, 40 * scale, 2), diblock(20, 0.25, 50 * scale),
                  PolymerType(length=star_n, count=30 * scale,
                              monotype=[0] + [1, 1, 2] * 4, bonds=star)])


def batched_mix(n_systems=64, seed=0, sizes=(16, 24, 32, 48, 64), lengths=(16, 24, 32, 48, 64),
                density_per_re3=128.0):
    """Independent systems of uneven size for batched runs: AB diblock melts or
    A/B homopolymer blends, N from `lengths`, grid edge from `sizes` (cells of
    Re/8), chiN uniform in [10, 40], kappaN = 50, fA uniform in [0.2, 0.8]
    (A fraction of each diblock, or of the chains in a blend), chain counts
    giving ~`density_per_re3` beads per Re^3.  Returns (spec, seed, system_id)
    triples; the streams of system i are keyed by (1000 seed + i, i)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_systems):
        N = int(rng.choice(lengths))
        g = int(rng.choice(sizes))
        chi = float(rng.uniform(10.0, 40.0))
        fA = float(rng.uniform(0.2, 0.8))
        L = g / 8.0
        chains = max(1, int(round(density_per_re3 * L ** 3 / N)))
        if rng.uniform() < 0.5:
            polys = [diblock(N, fA, chains)]
        else:
            na = max(1, int(round(fA * chains)))
            polys = [homopolymer(N, na, 0)] + ([homopolymer(N, chains - na, 1)]
                                               if chains > na else [])
        spec = SystemSpec(n_types=2, xn=xn_matrix(2, 50.0, chi), L=[L, L, L], n=[g, g, g],
                          dtNxi=0.17, nref=N, polymers=polys)
        out.append((spec, seed * 1000 + i, i))
    return out


def triblock(N: int, fA: float, count: int) -> PolymerType:
    """ABA triblock: A ends of fA/2 each around a B middle block."""
    na = int(round(N * fA / 2))
    return PolymerType(length=N, count=count, monotype=[0] * na + [1] * (N - 2 * na) + [0] * na)


def served_job(rng, density_per_re3=128.0, n_range=(16, 128), grid_range=(16, 128),
               chi_range=(10.0, 50.0), sweep_range=(500, 5000)):
    """One serving job: an AB diblock or ABA triblock melt with N, grid edge
    (cells of Re/8), chiN and sweep count drawn uniformly from the ranges,
    kappaN = 50, fA uniform in [0.2, 0.8], ~`density_per_re3` beads per Re^3.
    Returns (spec, sweeps)."""
    N = int(rng.integers(n_range[0], n_range[1] + 1))
    g = int(rng.integers(grid_range[0], grid_range[1] + 1))
    chi = float(rng.uniform(*chi_range))
    fA = float(rng.uniform(0.2, 0.8))
    L = g / 8.0
    chains = max(1, int(round(density_per_re3 * L ** 3 / N)))
    poly = diblock(N, fA, chains) if rng.uniform() < 0.5 else triblock(N, fA, chains)
    spec = SystemSpec(n_types=2, xn=xn_matrix(2, 50.0, chi), L=[L, L, L], n=[g, g, g],
                      dtNxi=0.17, nref=N, polymers=[poly])
    return spec, int(rng.integers(sweep_range[0], sweep_range[1] + 1))


NAMED = {
    "small_diblock_melt": small_diblock_melt,
    "large_diblock_melt": large_diblock_melt,
    "largest_blend": largest_blend,
    "reference_shipped": reference_shipped,
}
