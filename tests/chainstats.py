"""Chain and field statistics for the seed-ensemble comparisons.

Observables per run (sampled over the second half of the run):
  * acceptance rate of the moves;
  * mean-squared end-to-end distance <Re^2> (test.cpp:45-67, any N);
  * structure factor of the A-B composition, S(k) = |sum_c (n_A - n_B) e^{-ik.c}|^2 / n_beads,
    on the lattice shells |k|^2 = 1, 2, 3, 4 (units of 2 pi / L); the gated
    figure is its mean over all wave vectors of those shells (the low-k band).

Two runs of the same algorithm with independent random numbers must agree on
the three gated observables (acceptance, <Re^2>, low-k S) within two standard
errors of the difference of the 8-seed means.  The per-shell S values are
reported beside them but not gated: a dozen separate 2-sigma gates would
fail by chance about 40% of the time for identical algorithms.
"""
import numpy as np

from new_arch_amd import workloads as W
from new_arch_amd.abi import make_opts

N_SEEDS = 8
SHELLS = (1, 2, 3, 4)
NAMES = ("acceptance", "re2", "S(low k)") + tuple(f"S(k^2={s})" for s in SHELLS)
GATED = 3


def melt(grid=16):
    """Symmetric AB diblock melt, N=32, 4096 beads per Re^3, cells of Re/8,
    chiN=20, kappaN=50 (the small benchmark melt in a smaller box)."""
    L = grid / 8.0
    chains = int(round(4096 * L ** 3 / 32))
    return W.tiny(n_chains=chains, N=32, grid=(grid,) * 3, L=(L,) * 3, chiN=20.0, kappaN=50.0,
                  dtNxi=0.17, fA=0.5)


def opts(seed, move):
    return make_opts(seed=seed, move=move, trial_delta=0.1, omega_every_step=1)


def structure_factor(counts, grid):
    c = np.asarray(counts, np.float64).reshape(2, grid, grid, grid)
    f = np.fft.fftn(c[0] - c[1])
    s = (f * f.conj()).real / c.sum()
    k = np.fft.fftfreq(grid, 1.0 / grid)
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    k2 = np.rint(kx ** 2 + ky ** 2 + kz ** 2).astype(int)
    band = (k2 >= SHELLS[0]) & (k2 <= SHELLS[-1])
    return np.array([s[band].mean()] + [s[k2 == sh].mean() for sh in SHELLS])


def structure_factor_from_shells(sums, modes):
    """The vector of structure_factor from the shell sums and mode counts of
    the engine's scmf_structure_factor (integer |k|^2 shells)."""
    sh = np.asarray(SHELLS)
    band = sums[sh[0]:sh[-1] + 1].sum() / modes[sh[0]:sh[-1] + 1].sum()
    return np.array([band] + [sums[k] / modes[k] for k in sh])


class Sampler:
    """Accumulates the observables of one run."""

    def __init__(self):
        self.accepts = 0
        self.moves = 0
        self.re2 = []
        self.sq = []

    def add_sweep(self, accepts, moves):
        self.accepts += accepts
        self.moves += moves

    def add_sample(self, re2, counts, grid):
        self.re2.append(re2)
        self.sq.append(structure_factor(counts, grid))

    def add_sample_sq(self, re2, sq):
        self.re2.append(re2)
        self.sq.append(np.asarray(sq, np.float64))

    def result(self):
        return np.array([self.accepts / self.moves, np.mean(self.re2), *np.mean(self.sq, axis=0)])


def oracle_run_injected(orc, spec, seed, move, sweeps, grid, every):
    """The oracle's step loop (sweep, grid, omega every sweep) driven by
    numpy's PCG64 generator: 53-bit uniform picks, ziggurat fp32 normals
    (uniforms in [0,1) for trial moves) and fp32 accept uniforms in (0,1].
    Initial coordinates from the oracle's own walk with this seed."""
    o = orc.Oracle(spec, opts(seed, move))
    pos = o.init_coords()
    n = o.n_beads
    om = o.omega(o.grid(pos)).astype(np.float32)
    rng = np.random.Generator(np.random.PCG64(seed))
    smp = Sampler()
    for sw in range(sweeps):
        pick = rng.random(n)
        if move == "smart":
            gauss = rng.standard_normal(3 * n, dtype=np.float32)
        else:
            gauss = rng.random(3 * n, dtype=np.float32)
        acc_u = (np.float32(1.0) - rng.random(n, dtype=np.float32)).astype(np.float32)
        a = o.sweep_injected(pos, om, pick, gauss, acc_u)
        counts = o.grid(pos)
        om = o.omega(counts).astype(np.float32)
        if sw >= sweeps // 2:
            smp.add_sweep(a, n)
            if (sw + 1) % every == 0:
                smp.add_sample(o.re2(pos), counts, grid)
    return smp.result()


def oracle_run_streams(orc, spec, seed, move, sweeps, grid, every):
    """The same loop on the Philox streams the engine consumes."""
    o = orc.Oracle(spec, opts(seed, move))
    pos = o.init_coords()
    n = o.n_beads
    om = o.omega(o.grid(pos)).astype(np.float32)
    smp = Sampler()
    for sw in range(sweeps):
        a = o.sweep(pos, om, sw)
        counts = o.grid(pos)
        om = o.omega(counts).astype(np.float32)
        if sw >= sweeps // 2:
            smp.add_sweep(a, n)
            if (sw + 1) % every == 0:
                smp.add_sample(o.re2(pos), counts, grid)
    return smp.result()


def compare(a, b):
    """Per observable: (mean a, mean b, (mean a - mean b) / standard error of
    the difference)."""
    a = np.asarray(a)
    b = np.asarray(b)
    se = np.sqrt(a.var(0, ddof=1) / len(a) + b.var(0, ddof=1) / len(b))
    return a.mean(0), b.mean(0), (a.mean(0) - b.mean(0)) / se


def report(tag, a, b):
    ma, mb, z = compare(a, b)
    lines = [f"{tag}: {nm:12s} {x:10.6g} {y:10.6g}  z={zz:+.2f}{'' if i < GATED else ' (reported)'}"
             for i, (nm, x, y, zz) in enumerate(zip(NAMES, ma, mb, z))]
    print("\n".join(lines), flush=True)
    return z[:GATED]
