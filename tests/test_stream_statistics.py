"""The injected Philox streams against an independent generator (CPU).

The reference's own random numbers (cuRAND) cannot be reproduced, so the
engine and the oracle share counter-based Philox streams with their own
normal and uniform transforms.  This pins that choice statistically: the
oracle's move algorithm driven by those streams and by numpy's PCG64 (an
unrelated generator and normal method) gives the same acceptance rate,
<Re^2> and low-k composition structure factor within 2 sigma over 8 seeds
each.
"""
import numpy as np
import pytest

import chainstats as CS
from oracle import oracle as orc

SWEEPS, EVERY = 100, 10


@pytest.mark.parametrize("move", ["smart", "trial"])
def test_streams_match_independent_generator(move):
    spec = CS.melt(grid=16)
    a = [CS.oracle_run_streams(orc, spec, s, move, SWEEPS, 16, EVERY)
         for s in range(1, CS.N_SEEDS + 1)]
    b = [CS.oracle_run_injected(orc, spec, 100 + s, move, SWEEPS, 16, EVERY)
         for s in range(1, CS.N_SEEDS + 1)]
    z = CS.report(f"streams-vs-pcg64 {move}", a, b)
    assert np.all(np.abs(z) < 2.0), z


def test_injected_sweep_matches_streams_when_fed_the_same_numbers():
    """The injected entry is the same move: fed the numbers the Philox streams
    would give, it reproduces the streamed sweep bit for bit."""
    spec = CS.melt(grid=8)
    o = orc.Oracle(spec, CS.opts(3, "smart"))
    pos = o.init_coords()
    om = o.omega(o.grid(pos)).astype(np.float32)
    n = o.n_beads
    # Rebuild the draws of sweep 0 from raw Philox blocks: move i of chain c
    # picks with words (0, 1) or (2, 3) of the pick block of index i // 2 (53
    # bits), and takes its accept uniform from the low bytes of words 0, 1, 2
    # of its displacement block (index i).
    N = spec.polymers[0].length
    pick = np.empty(n)
    gauss = np.empty((n, 3), np.float32)
    acc_u = np.empty(n, np.float32)
    for c in range(spec.polymers[0].count):
        gauss[c * N:(c + 1) * N] = orc.normals_f(3, 0, c, 0, N)
        for i in range(N):
            w = [int(v) for v in orc.philox((c, 0, ((i >> 1) << 2) | 0, 0), (3, 0))]
            a, b = (w[2], w[3]) if i & 1 else (w[0], w[1])
            m = ((a >> 5) << 26) | (b >> 6)
            pick[c * N + i] = (m + 1) * 2.0 ** -53
            d = [int(v) for v in orc.philox((c, 0, (i << 2) | 1, 0), (3, 0))]
            bits = (d[0] & 0xFF) | ((d[1] & 0xFF) << 8) | ((d[2] & 0xFF) << 16)
            acc_u[c * N + i] = np.float32((bits + 1) * 2.0 ** -24)
    p1 = pos.copy()
    a1 = o.sweep(p1, om, 0)
    p2 = pos.copy()
    a2 = o.sweep_injected(p2, om, pick, gauss.ravel(), acc_u)
    assert a1 == a2
    assert np.array_equal(p1.view(np.uint32), p2.view(np.uint32))

