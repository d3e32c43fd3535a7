"""Pins of the shared arithmetic primitives (random streams, exp, log, sincos).

These primitives replace cuRAND and CUDA's exp (third-party code of the
reference, not available here); they are pinned by ROCm's rocRAND Philox
header (fixture made by tests/golden/make_fixtures.py), the published
Random123 known-answer vectors, libm accuracy bounds and distribution moments.
"""
import os

import numpy as np
import pytest

from oracle import oracle as orc


def ulp_err(got, want):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    return np.abs(got - want) / np.spacing(np.abs(want))


def test_philox_matches_rocrand(golden):
    f = np.load(os.path.join(golden, "philox_rocrand.npz"))
    for ctr, key, out in zip(f["ctr"], f["key"], f["out"]):
        np.testing.assert_array_equal(orc.philox(ctr, key), out)


def test_philox_random123_kat():
    # Known-answer vectors published with Random123 (kat_vectors, philox4x32 R=10).
    assert list(orc.philox([0] * 4, [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(orc.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [
        0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert list(orc.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                           [0xA4093822, 0x299F31D0])) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_exp_accuracy():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-50, 50, 200000), rng.uniform(-1, 1, 100000),
                        rng.uniform(-745, 709, 100000), [0.0, -0.0, 1e-300, -1e-300, 709.7, -700]])
    y = orc.exp_many(x)
    want = np.exp(x)
    normal = want > 2.3e-308
    assert ulp_err(y[normal], want[normal]).max() <= 2.0
    assert orc.exp_many(np.array([0.0]))[0] == 1.0
    assert np.isinf(orc.exp_many(np.array([710.0]))[0])
    assert orc.exp_many(np.array([-746.0]))[0] == 0.0


def test_exp_monotone_around_zero():
    # The accept test compares exp(-dE) with 1 and with U; exp must not dip
    # below 1 for positive arguments or exceed 1 for negative ones.
    x = np.concatenate([np.linspace(-1e-6, 1e-6, 20001), np.array([-1e-300, 1e-300])])
    y = orc.exp_many(x)
    assert np.all(y[x > 0] >= 1.0) and np.all(y[x < 0] <= 1.0)


def test_log_accuracy():
    rng = np.random.default_rng(2)
    u = ((rng.integers(0, 2**32, 300000).astype(np.float64) + 0.5) * 2.0**-32)
    x = np.concatenate([u, rng.uniform(1e-10, 1e10, 100000), [1.0, 2.0, 0.5, 1e-300, 1e300]])
    y = orc.log_many(x)
    want = np.log(x)
    nz = np.abs(want) > 1e-300
    # absolute error relative to ulp of the result, plus exact log(1) == 0
    assert ulp_err(y[nz], want[nz]).max() <= 2.0
    assert orc.log_many(np.array([1.0]))[0] == 0.0


def test_sincos_accuracy():
    rng = np.random.default_rng(3)
    u = rng.integers(0, 2**32, 300000).astype(np.float64) * 2.0**-32
    u = np.concatenate([u, [0.0, 0.25, 0.5, 0.75, 1 - 2.0**-32, 0.125]])
    s, c = orc.sincos_many(u)
    # u - round(u) is exact, so the reference angle carries one rounding only
    th = 2 * np.pi * (u - np.round(u))
    assert np.abs(s - np.sin(th)).max() < 5e-16
    assert np.abs(c - np.cos(th)).max() < 5e-16
    assert s[-6] == 0.0 and c[-6] == 1.0


def test_logf_accuracy():
    # every uniform the displacement normals feed it, 2^-24 k for k = 1..2^24
    k = np.arange(1, 2**24 + 1, dtype=np.uint64)
    x = (k.astype(np.float64) * 2.0**-24).astype(np.float32)
    y = orc.logf_many(x).astype(np.float64)
    want = np.log(x.astype(np.float64))
    # absolute error against the exact log, in float ulps of the result
    err = np.abs(y - want) / np.spacing(np.maximum(np.abs(want), 2.0**-24).astype(np.float32))
    assert err.max() <= 1.0
    assert orc.logf_many(np.array([1.0], np.float32))[0] == 0.0


def test_sincosf_accuracy():
    k = np.arange(0, 2**24, 61, dtype=np.uint32)
    s, c = orc.sincosf_turn24_many(k)
    th = 2 * np.pi * (k.astype(np.float64) * 2.0**-24)
    assert np.abs(s - np.sin(th)).max() < 2.5e-7
    assert np.abs(c - np.cos(th)).max() < 2.5e-7
    s0, c0 = orc.sincosf_turn24_many(np.array([0, 1 << 22, 1 << 23, 3 << 22], np.uint32))
    np.testing.assert_array_equal(s0, [0, 1, 0, -1])
    np.testing.assert_array_equal(c0, [1, 0, -1, 0])


def test_normals_f_moments():
    n = orc.normals_f(seed=7, system=0, chain=3, sweep=11, n=400000).astype(np.float64).ravel()
    assert abs(n.mean()) < 4 / np.sqrt(n.size)
    assert abs(n.var() - 1) < 4 * np.sqrt(2 / n.size)
    assert abs((n ** 4).mean() - 3) < 0.05
    assert abs((np.abs(n) > 3).mean() - 0.0026998) < 0.0004
    a = orc.normals_f(seed=9, system=2, chain=5, sweep=0, n=200000)
    assert np.abs(np.corrcoef(a.T.astype(np.float64)) - np.eye(3)).max() < 0.01


def test_normals_moments():
    n = orc.normals(seed=7, system=0, chain=3, sweep=11, n=400000).ravel()
    assert abs(n.mean()) < 4 * 1 / np.sqrt(n.size)
    assert abs(n.var() - 1) < 4 * np.sqrt(2 / n.size)
    # fourth moment of a standard normal is 3
    assert abs((n ** 4).mean() - 3) < 0.05
    # tail mass beyond 3 sigma: 0.0027
    assert abs((np.abs(n) > 3).mean() - 0.0026998) < 0.0004


def test_normals_independent_components():
    n = orc.normals(seed=9, system=2, chain=5, sweep=0, n=200000)
    cc = np.corrcoef(n.T)
    assert np.abs(cc - np.eye(3)).max() < 0.01


@pytest.mark.parametrize("key", [(1, 0), (123456789, 42)])
def test_streams_differ_by_slot(key):
    a = orc.philox([5, 0, (3 << 2) | 0, key[1]], [key[0], 0])
    b = orc.philox([5, 0, (3 << 2) | 1, key[1]], [key[0], 0])
    assert not np.array_equal(a, b)


def test_pick_forms_agree():
    """The engine's pick (the 53-bit uniform as one exact fma of its halves,
    times n 2^-53) equals the defining (int)(u n) with the last bead for
    u = 1, on random blocks and on the words at the ends of both halves,
    for chain lengths up to past 255."""
    rng = np.random.default_rng(17)
    w = rng.integers(0, 2 ** 32, (400_000, 4), dtype=np.uint64).astype(np.uint32)
    edge = np.array([0, 1, 31, 32, 63, 64, 2 ** 31, 2 ** 32 - 65, 2 ** 32 - 64, 2 ** 32 - 1],
                    np.uint32)
    a, b = np.meshgrid(edge, edge)
    e = np.stack([a.ravel(), b.ravel(), a.ravel(), b.ravel()], 1).astype(np.uint32)
    words = np.concatenate([w, e])
    for n in (1, 2, 3, 7, 31, 32, 63, 64, 65, 96, 127, 128, 160, 255, 300, 4500, 1 << 20):
        assert orc.pick_disagreements(words, n) == 0, n
    # a uniform of exactly 1 (both halves all ones) picks the last bead
    top = np.array([[2 ** 32 - 1] * 4], np.uint32)
    assert orc.pick_disagreements(top, 64) == 0
