"""Chain statistics of the HIP engine within 2 sigma over 8 seeds.

The engine (Philox streams, seeds 1-8) runs the reference step loop on a
110,592-bead symmetric diblock melt (N=32, chiN=20, kappaN=50, 24^3 cells,
omega every sweep) for 200 sweeps; the oracle runs the same loop with moves
driven by an unrelated generator (numpy PCG64, seeds 101-108).  Acceptance
rate, <Re^2> (the engine's scmf_stats) and the low-k composition structure
factor (the engine's scmf_structure_factor on the device, numpy's FFT on
the oracle side), sampled over the second half, agree within two standard errors of the difference of the
means.  (With identical streams the engine equals the oracle bit for bit;
that is tested elsewhere.)
"""
import numpy as np
import pytest

import chainstats as CS
from new_arch_amd import scmf
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GRID, SWEEPS, EVERY = 24, 200, 20


def engine_run(spec, seed, move):
    g = scmf.System(spec, CS.opts(seed, move))
    try:
        g.grid_density()  # test.cpp:21-23
        g.update_omega()
        smp = CS.Sampler()
        for sw in range(SWEEPS):
            if sw == SWEEPS // 2:
                g.reset_stats()
            g.step(sw)
            if sw >= SWEEPS // 2 and (sw + 1) % EVERY == 0:
                sums, modes = g.structure_factor((1.0, -1.0), 2.0 * np.pi / spec.L[0],
                                                 CS.SHELLS[-1] + 1, by_q2=True)
                sq = CS.structure_factor_from_shells(sums, modes)
                np.testing.assert_allclose(sq, CS.structure_factor(g.density(), GRID), rtol=1e-9)
                smp.add_sample_sq(g.stats().re2, sq)
        st = g.stats()
        smp.add_sweep(st.accepts, st.moves)
        assert st.moves == (SWEEPS - SWEEPS // 2) * g.n_beads
        return smp.result()
    finally:
        g.close()


@pytest.mark.parametrize("move", ["smart", "trial"])
def test_engine_statistics_within_two_sigma(move):
    print(f"[gpu-stats] {move}: engine seeds 1-8", flush=True)
    spec = CS.melt(GRID)
    a = [engine_run(spec, s, move) for s in range(1, CS.N_SEEDS + 1)]
    print(f"[gpu-stats] {move}: oracle (PCG64) seeds 101-108", flush=True)
    b = [CS.oracle_run_injected(orc, spec, 100 + s, move, SWEEPS, GRID, EVERY)
         for s in range(1, CS.N_SEEDS + 1)]
    z = CS.report(f"engine-vs-oracle(pcg64) {move}", a, b)
    assert np.all(np.abs(z) < 2.0), z
