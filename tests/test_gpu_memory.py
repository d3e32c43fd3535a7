"""The device-memory estimate of scmf_inspect_desc (what the job server admits
against) covers what a system really takes over a run: creation, the first
spatial ordering, the cells and the schedule records."""
import ctypes as C

import pytest

from new_arch_amd import scmf
from new_arch_amd import workloads as W
from new_arch_amd.abi import make_opts

pytestmark = pytest.mark.gpu


def device_free():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


# 9.6M beads, 300,000 chains: ordered automatically, schedule records in use.
# "wide halo": cells of 0.094 Re put the halo at 8 cells, so the only tiling
# that fits LDS has 32^3 windows around 16^3 tiles (8 window cells per
# interior cell, the most the tile chooser allows).
GEOMETRIES = {"melt": dict(grid=(64, 64, 64), L=(8.0, 8.0, 8.0)),
              "wide halo": dict(grid=(96, 96, 96), L=(9.0, 9.0, 9.0))}


@pytest.mark.parametrize("bits,geometry", [(16, "melt"), (32, "melt"), (16, "wide halo")])
def test_estimate_covers_a_run(bits, geometry):
    spec = W.tiny(n_chains=300_000, N=32, **GEOMETRIES[geometry])
    est = scmf.inspect_spec(spec)
    want = est.device_bytes_16 if bits == 16 else est.device_bytes_32
    print(f"[gpu-memory] u{bits} {geometry}: estimate {want / 2**20:.1f} MiB", flush=True)
    # warm the runtime first (code objects, queues, kernel scratch), so the
    # difference below is the system's own memory
    w = scmf.System(W.tiny(n_chains=70_000, N=32, grid=(16, 16, 16), L=(2.0, 2.0, 2.0)),
                    make_opts(seed=1, density_bits=bits, sort_every=1))
    w.grid_density()
    w.update_omega()
    w.step(1)
    w.sync()
    w.close()
    before = device_free()
    g = scmf.System(spec, make_opts(seed=2, density_bits=bits))
    try:
        g.grid_density()
        g.update_omega()
        for k in range(1, 4):
            g.step(k)
        g.sync()
        assert g.stats().sorts >= 1 and g.stats().cell_grids >= 1
        used = before - device_free()
    finally:
        g.close()
    print(f"[gpu-memory] u{bits} {geometry}: used {used / 2**20:.1f} MiB", flush=True)
    assert used <= want                # the estimate includes allocation granularity
    assert used >= 0.7 * want          # and is not wildly above the use
