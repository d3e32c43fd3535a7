"""The device builds of the shared primitives equal the host builds the
oracle uses, bit for bit.  The device code differs (exp's Taylor steps take
their coefficients from scalar registers; the Box-Muller radius uses a
correctly rounded sqrt for its fixed argument range), so the results are
pinned directly, beyond the trajectories of the parity tests:
  * Philox4x32-10 blocks (the device rounds XOR three words in one
    instruction) against the oracle's host build and a numpy restatement;
  * normal3f over every 24-bit uniform of both radii, so the fp32 log and
    the sqrt see each input they can ever get;
  * exp over wide, narrow, near-overflow, gradual-underflow and special
    arguments."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "new_arch_amd", "lib", "libscmf_check.so")


def check_lib():
    if not os.path.exists(LIB):
        raise ImportError(f"{LIB} is missing: build it with new_arch_amd/build.py")
    L = C.CDLL(LIB)
    L.scmf_check_normal3f.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
    L.scmf_check_exp.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
    L.scmf_check_philox.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
    return L


def philox_numpy(w):
    """Philox4x32-10 (Salmon et al. 2011) of counter w[:, 0:4] and key
    w[:, 4:6], vectorised: ten rounds of the two 32x32->64 products, the XOR
    with the other counter words and the key, and the Weyl key schedule."""
    c = [w[:, j].astype(np.uint64) for j in range(4)]
    k0, k1 = w[:, 4].astype(np.uint64), w[:, 5].astype(np.uint64)
    m32 = np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0 = np.uint64(0xD2511F53) * c[0]
        p1 = np.uint64(0xCD9E8D57) * c[2]
        c = [(p1 >> np.uint64(32)) ^ c[1] ^ k0, p1 & m32, (p0 >> np.uint64(32)) ^ c[3] ^ k1, p0 & m32]
        k0 = (k0 + np.uint64(0x9E3779B9)) & m32
        k1 = (k1 + np.uint64(0xBB67AE85)) & m32
    return np.stack(c, 1).astype(np.uint32)


def test_philox_device_equals_host():
    print("[gpu-prims] philox", flush=True)
    rng = np.random.default_rng(4)
    n = 1 << 20
    w = rng.integers(0, 2 ** 32, (n, 6), dtype=np.uint64).astype(np.uint32)
    w[:4] = 0
    w[4:8] = np.uint32(0xFFFFFFFF)
    dev = np.empty((n, 4), np.uint32)
    assert check_lib().scmf_check_philox(w.ctypes.data, n, dev.ctypes.data) == 0
    want = philox_numpy(w)
    bad = np.nonzero((dev != want).any(1))[0]
    assert bad.size == 0, f"{bad.size} mismatches, first block {bad[0]}: {w[bad[0]]}"
    for i in list(range(8)) + [12345, n - 1]:
        assert np.array_equal(dev[i], orc.philox(w[i, :4], w[i, 4:]))


def test_normal3f_every_radius_uniform():
    print("[gpu-prims] normal3f over 2^24 uniforms", flush=True)
    n = 1 << 24
    k = np.arange(n, dtype=np.uint32)
    rng = np.random.default_rng(5)
    w = np.empty((n, 4), np.uint32)
    w[:, 0] = (k << 8) | rng.integers(0, 256, n, dtype=np.uint32)
    w[:, 1] = rng.integers(0, 2 ** 32, n, dtype=np.uint32)
    w[:, 2] = (((k * np.uint32(2654435761)) & np.uint32(0xFFFFFF)) << 8) | \
        rng.integers(0, 256, n, dtype=np.uint32)   # every value again, in another order
    w[:, 3] = rng.integers(0, 2 ** 32, n, dtype=np.uint32)
    dev = np.empty((n, 3), np.float32)
    assert check_lib().scmf_check_normal3f(w.ctypes.data, n, dev.ctypes.data) == 0
    host = orc.normal3f_words(w)
    bad = np.nonzero(dev.view(np.uint32) != host.view(np.uint32))[0]
    assert bad.size == 0, f"{bad.size} mismatches, first block {bad[0]}: {w[bad[0]]}"


def test_exp_device_equals_host():
    print("[gpu-prims] exp", flush=True)
    rng = np.random.default_rng(6)
    x = np.concatenate([
        rng.uniform(-50, 50, 2_000_000), rng.uniform(-1, 1, 1_000_000),
        rng.uniform(-1e-6, 1e-6, 200_000), rng.uniform(700, 710, 200_000),
        rng.uniform(-746, -700, 500_000),           # gradual underflow
        np.array([0.0, -0.0, 1e-300, -1e-300, 709.782712893384, 709.7827128933841,
                  -745.2, -745.1999999999999, np.inf, -np.inf, np.nan, 1.0, -1.0,
                  np.log(2.0), -np.log(2.0)]),
    ])
    dev = np.empty_like(x)
    assert check_lib().scmf_check_exp(x.ctypes.data, x.size, dev.ctypes.data) == 0
    host = orc.exp_many(x)
    nan = np.isnan(host)
    assert np.array_equal(np.isnan(dev), nan)
    bad = np.nonzero(dev[~nan].view(np.uint64) != host[~nan].view(np.uint64))[0]
    assert bad.size == 0, f"{bad.size} mismatches, first x = {x[~nan][bad[0]]!r}"
