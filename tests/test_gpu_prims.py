"""The device builds of the shared primitives equal the host builds the
oracle uses, bit for bit.  The device code differs (exp's Taylor steps take
their coefficients from scalar registers; the Box-Muller radius uses a
correctly rounded sqrt for its fixed argument range), so the results are
pinned directly, beyond the trajectories of the parity tests:
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
    return L


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
