"""Exact checkpoint/restart through the C ABI: a run restored from a
checkpoint continues bit for bit (positions, counts, omega, stream position),
including the reference omega cadence that depends on the step number, and
ends where the oracle's uninterrupted run of the same steps ends (the
reference's restart path is init_config.cu:930-987 with io.cpp:95-118, which
loses digits to %g; this one does not); a checkpoint of another system or a
damaged file is refused."""
import os

import numpy as np
import pytest

from new_arch_amd import scmf
from new_arch_amd import workloads as W
from new_arch_amd.abi import make_opts
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32),
                          np.ascontiguousarray(b).view(np.uint32))


def steps(g, first, last):
    for k in range(first, last + 1):
        g.step(k)
    g.sync()


@pytest.mark.parametrize("bits,sort_every", [(16, 0), (32, 3)])
def test_restored_run_continues_bit_for_bit(tmp_path, bits, sort_every):
    print(f"[gpu-ckpt] u{bits} sort_every={sort_every}", flush=True)
    spec = W.tiny(n_chains=400, N=20, grid=(10, 10, 10), L=(2.5, 2.5, 2.5), chiN=25.0)
    opts = dict(density_bits=bits, sort_every=sort_every, omega_every_step=0)  # reference cadence
    a = scmf.System(spec, make_opts(seed=21, **opts))
    a.grid_density()
    a.update_omega()
    steps(a, 1, 8)
    ck = tmp_path / "run.ckpt"
    a.checkpoint(ck, 8)
    steps(a, 9, 14)                      # crosses step 10, where omega is not updated
    b = scmf.System(spec, make_opts(seed=99, **opts))   # another seed: its state is overwritten
    assert b.restore(ck) == 8
    steps(b, 9, 14)
    assert same_bits(a.positions(), b.positions())
    assert np.array_equal(a.density(), b.density())
    assert same_bits(a.omega(), b.omega())
    assert a.stats().sweeps == b.stats().sweeps
    # and both equal the oracle's uninterrupted run of the same 14 steps
    # (test.cpp:37-43 with the reference cadence: no omega update at step 10)
    o = orc.Oracle(spec, make_opts(seed=21, **opts))
    pos = o.init_coords()
    counts = o.grid(pos)
    om = o.omega(counts).astype(np.float32)
    acc = 0
    for step in range(1, 15):
        acc += o.sweep(pos, om, step - 1)
        counts = o.grid(pos)
        if step % 10:
            om = o.omega(counts).astype(np.float32)
    assert same_bits(b.positions(), pos)
    assert np.array_equal(b.density(), counts)
    assert same_bits(b.omega(), om)
    assert a.stats().accepts == acc
    a.close()
    b.close()


def test_checkpoint_refuses_other_systems_and_damage(tmp_path):
    spec = W.tiny(n_chains=100, N=16)
    g = scmf.System(spec, make_opts(seed=3, density_bits=32))
    g.grid_density()
    g.update_omega()
    ck = tmp_path / "a.ckpt"
    g.checkpoint(ck, 0)
    other = scmf.System(W.tiny(n_chains=101, N=16), make_opts(seed=3, density_bits=32))
    with pytest.raises(scmf.ScmfError):
        other.restore(ck)
    other16 = scmf.System(spec, make_opts(seed=3, density_bits=16))
    with pytest.raises(scmf.ScmfError):
        other16.restore(ck)
    cut = tmp_path / "cut.ckpt"
    data = ck.read_bytes()
    cut.write_bytes(data[: len(data) - 7])
    with pytest.raises(scmf.ScmfError):
        g.restore(cut)
    junk = tmp_path / "junk.ckpt"
    junk.write_bytes(b"not a checkpoint" * 10)
    with pytest.raises(scmf.ScmfError):
        g.restore(junk)
    with pytest.raises(scmf.ScmfError):
        g.restore(tmp_path / "missing.ckpt")
    assert not os.path.exists(str(ck) + ".tmp")
    for s in (g, other, other16):
        s.close()


def test_driver_restart_matches_straight_run(golden, tmp_path):
    """scmf_run: 20 steps straight, or 12 steps + -checkpoint then -restart for
    8 more, give the same density_20.dat."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "new_arch_amd", "lib", "scmf_run")
    inp = os.path.join(golden, "reference_inputs")
    runs = {}
    for name in ("straight", "split"):
        d = tmp_path / name
        d.mkdir()
        for f in ("configuration.dat", "polymer.dat"):
            (d / f).write_text(open(os.path.join(inp, f)).read())
        runs[name] = d
    common = ["-Check_T=4", "-seed=5"]
    r = subprocess.run([exe, "-Steps_N=20", *common], cwd=runs["straight"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe, "-Steps_N=12", "-checkpoint=run.ckpt", *common], cwd=runs["split"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe, "-Steps_N=8", "-restart=run.ckpt", "-seed=77", "-Check_T=4"],
                       cwd=runs["split"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "restarted from run.ckpt at step 12" in r.stdout
    a = (runs["straight"] / "density_20.dat").read_text()
    b = (runs["split"] / "density_20.dat").read_text()
    assert a == b
