"""Asynchronous density output: each queued file holds the state at the
time of the call (the steps that follow do not leak into it), is byte-equal
to the synchronous writer's, survives more calls than snapshot slots, is
finished by close, and write errors come back from the wait."""
import os

import pytest

from new_arch_amd import scmf
from new_arch_amd import workloads as W
from new_arch_amd.abi import make_opts

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bits", [16, 32])
def test_async_files_hold_the_state_of_the_call(tmp_path, bits):
    print(f"[gpu-output] u{bits}", flush=True)
    sync_dir, async_dir = tmp_path / "sync", tmp_path / "async"
    sync_dir.mkdir()
    async_dir.mkdir()
    g = scmf.System(W.tiny(n_chains=800, N=16, grid=(12, 10, 8), L=(2.4, 2.0, 1.6)),
                    make_opts(seed=41, density_bits=bits))
    try:
        g.grid_density()
        g.update_omega()
        steps = list(range(1, 8))
        for k in steps:
            g.write_density_async(k, async_dir)   # queued first ...
            g.write_density(k, sync_dir)          # ... the same state, written now
            g.step(k)                             # changes counts and omega at once
        g.output_wait()
        for k in steps:
            name = f"density_{k}.dat"
            a = (async_dir / name).read_bytes()
            assert a == (sync_dir / name).read_bytes(), name
        assert len({(async_dir / f"density_{k}.dat").read_bytes() for k in steps}) == len(steps)
        assert not [p for p in os.listdir(async_dir) if p.endswith(".tmp")]
    finally:
        g.close()


def test_async_errors_and_close(tmp_path):
    g = scmf.System(W.tiny(n_chains=100, N=16), make_opts(seed=42))
    g.grid_density()
    g.update_omega()
    g.write_density_async(1, tmp_path / "missing_dir")
    with pytest.raises(scmf.ScmfError):
        g.output_wait()
    g.output_wait()                               # the error was reported once
    g.write_density_async(2, tmp_path)
    g.write_density_async(3, tmp_path)
    g.write_density_async(4, tmp_path)
    g.close()                                     # finishes the queue
    for k in (2, 3, 4):
        text = (tmp_path / f"density_{k}.dat").read_text()
        assert text.startswith("Nx=8, Ny=8, Nz=8\n")
        assert text.count("\n") == 2 + 8 * 8 * 8


def test_async_vtk_equals_converter(tmp_path):
    """write_outputs_async with a VTK path: the VTK file equals converting the
    density file written with it (and that file equals the synchronous one)."""
    g = scmf.System(W.tiny(n_chains=900, N=16, grid=(12, 10, 8), L=(2.4, 2.0, 1.6)),
                    make_opts(seed=43, density_bits=16))
    try:
        g.grid_density()
        g.update_omega()
        g.write_outputs_async(5, tmp_path, vtk_path=tmp_path / "a.vtk")
        sync_dir = tmp_path / "sync"
        sync_dir.mkdir()
        g.write_density(5, sync_dir)
        g.step(1)
        g.output_wait()
        dens = tmp_path / "density_5.dat"
        assert dens.read_bytes() == (sync_dir / "density_5.dat").read_bytes()
        assert scmf.lib().scmf_density_to_vtk(str(dens).encode(), str(tmp_path / "c.vtk").encode()) == 0
        assert (tmp_path / "a.vtk").read_bytes() == (tmp_path / "c.vtk").read_bytes()
    finally:
        g.close()
