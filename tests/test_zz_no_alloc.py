"""scmf_copy_out_noalloc, the copy-out for callers that cannot allocate device
memory, against scmf_copy_out: the same positions, counts and omega bit for
bit on two calls in a row (a larger system, then a smaller one), uint16 and
uint32 counts, reordered slots; the device's free memory is unchanged by the
calls; a system beyond the fixed staging array is refused with SCMF_E_ARG."""
import ctypes as C

import numpy as np
import pytest

from new_arch_amd import scmf
from new_arch_amd import workloads as W
from new_arch_amd.abi import DENSITY, OMEGA, POS, SCMF_E_ARG, make_opts

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-noalloc] {msg}", flush=True)


def device_free():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def same_bits(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def stepped(spec, **kw):
    g = scmf.System(spec, make_opts(**kw))
    g.grid_density()
    g.update_omega()
    for k in range(1, 4):
        g.step(k)
    g.sync()
    return g


def check(g):
    # the reference copies first: they run the same gather and widening
    # kernels (the runtime loads a kernel's code on its first launch) and make
    # scmf_copy_out's own staging buffer
    ref = {POS: g.positions(), DENSITY: g.density(), OMEGA: g.omega()}
    free0 = device_free()
    got = {w: g.copy_out_noalloc(w) for w in (POS, DENSITY, OMEGA)}
    assert device_free() == free0, "the call allocated device memory"
    assert same_bits(got[POS], ref[POS])
    assert np.array_equal(got[DENSITY], ref[DENSITY])
    assert same_bits(got[OMEGA], ref[OMEGA])
    assert got[DENSITY].sum(dtype=np.int64) == g.n_beads


@pytest.mark.parametrize("bits", [16, 32])
def test_noalloc_copy_out_equals_copy_out(bits):
    banner(f"u{bits}: larger system, then a smaller one")
    big = stepped(W.tiny(n_chains=30_000, N=32, grid=(24, 24, 24), L=(3.0, 3.0, 3.0)),
                  seed=5, density_bits=bits, sort_every=2)     # 960,000 beads, reordered
    assert big.stats().sorts > 0
    check(big)
    small = stepped(W.mixed_arch(), seed=6, density_bits=bits)   # 2,350 beads, 3 types
    check(small)
    check(big)      # and back: the array holds the last call's contents, fully overwritten
    big.close()
    small.close()


def test_noalloc_refuses_systems_beyond_the_array():
    banner("beyond the array")
    spec = W.tiny(n_chains=50_000, N=32, grid=(16, 16, 16), L=(2.0, 2.0, 2.0))  # 1.6M beads
    g = scmf.System(spec, make_opts(seed=7, density_bits=16))
    out = np.empty((spec.n_beads, 3), np.float32)
    rc = scmf.lib().scmf_copy_out_noalloc(g._h, POS, out.ctypes.data, out.nbytes)
    assert rc == SCMF_E_ARG
    assert b"at most 1398101 beads" in scmf.lib().scmf_last_error()
    # counts and omega of the same system still fit
    g.grid_density()
    g.update_omega()
    assert np.array_equal(g.copy_out_noalloc(DENSITY), g.density())
    assert same_bits(g.copy_out_noalloc(OMEGA), g.omega())
    g.close()
