"""Host-side formats and the C ABI, without a GPU.

* configuration.dat / polymer.dat readers against the reference's shipped inputs
* density_<step>.dat writer against the reference's density_240000.dat
* VTK converter against the reference's out.vtk (made by vtk.c from that dump)
* every symbol include/scmf.h declares is exported; argument errors come back
  as status codes with a message instead of ending the process
"""
import ctypes as C
import gzip
import os
import re

import numpy as np
import pytest

from new_arch_amd import scmf
from new_arch_amd import workloads as W
from new_arch_amd.abi import Summary, make_opts

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def inspect(cfg, poly):
    s = Summary()
    rc = scmf.lib().scmf_inspect_files(str(cfg).encode(), str(poly).encode(), C.byref(s))
    return rc, s


def test_header_symbols_exported():
    hdr = open(os.path.join(ROOT, "include", "scmf.h")).read()
    names = set(re.findall(r"\b(scmf_[a-z_0-9]+)\s*\(", hdr))
    assert len(names) >= 25
    L = scmf.lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_read_reference_inputs(golden):
    d = os.path.join(golden, "reference_inputs")
    rc, s = inspect(os.path.join(d, "configuration.dat"), os.path.join(d, "polymer.dat"))
    assert rc == 0, scmf.lib().scmf_last_error()
    assert s.n_mono_types == 2 and s.n_poly_types == 1 and s.reference_Nbeads == 32
    assert list(s.n) == [24, 24, 24] and list(s.L) == [4.0, 4.0, 4.0]
    assert s.dtNxi == 0.17
    assert list(s.xn[:4]) == [50.0, 30.0, 30.0, 50.0]
    assert s.n_chains == 12800 and s.n_beads == 409600 and s.n_cells == 13824


def test_read_errors(tmp_path, golden):
    d = os.path.join(golden, "reference_inputs")
    cfg = os.path.join(d, "configuration.dat")
    rc, _ = inspect(cfg, tmp_path / "nope.dat")
    assert rc == -2 and b"cannot open" in scmf.lib().scmf_last_error()
    poly = open(os.path.join(d, "polymer.dat")).read()
    bad = tmp_path / "polymer.dat"
    bad.write_text(poly.replace("Number of total polymer:12800", "Number of total polymer:12801"))
    rc, _ = inspect(cfg, bad)
    assert rc == -2 and b"add up" in scmf.lib().scmf_last_error()
    bad.write_text(poly[: len(poly) // 2])
    rc, _ = inspect(cfg, bad)
    assert rc == -2
    badcfg = tmp_path / "configuration.dat"
    badcfg.write_text(open(cfg).read().replace("N = 24 24 24", "N = 24 x 24"))
    rc, _ = inspect(badcfg, os.path.join(d, "polymer.dat"))
    assert rc == -2 and b"ny" in scmf.lib().scmf_last_error()


def test_density_writer_matches_reference_dump(golden, tmp_path):
    f = np.load(os.path.join(golden, "density_fixtures.npz"))
    counts = np.ascontiguousarray(f["counts_240000"].astype(np.uint32).T)   # [type][cell]
    omega = np.ascontiguousarray(f["omega_240000"].T.astype(np.float32))
    n = (C.c_int32 * 3)(24, 24, 24)
    L = (C.c_double * 3)(4.0, 4.0, 4.0)
    out = tmp_path / "density_240000.dat"
    rc = scmf.lib().scmf_write_density_host(str(out).encode(), n, L, 3, counts.ctypes.data,
                                            omega.ctypes.data, C.c_double(13824 / 409600))
    assert rc == 0
    ref = gzip.open(os.path.join(golden, "density_240000.dat.gz"), "rt").read().splitlines()
    got = out.read_text().splitlines()
    assert got[:2] == ref[:2]
    assert len(got) == len(ref)
    for a, b in zip(got[2:], ref[2:]):
        ta, tb = a.split(), b.split()
        assert ta[:3] == tb[:3]                       # density columns: identical text
        np.testing.assert_allclose([float(v) for v in ta[3:]], [float(v) for v in tb[3:]],
                                   rtol=1e-5, atol=1e-6)
        assert a.endswith(" ") and b.endswith(" ")


def test_density_file_over_several_format_blocks(tmp_path):
    """153,600 cells: the writer formats blocks of 65,536 rows on several
    threads; rows must come out in cell order, each "%g " per value."""
    nx, ny, nz, T = 64, 48, 50, 2
    cells = nx * ny * nz
    rng = np.random.default_rng(3)
    counts = rng.integers(0, 300, size=T * cells).astype(np.uint32)
    omega = (rng.standard_normal(T * cells) * 3).astype(np.float32)
    n = (C.c_int32 * 3)(nx, ny, nz)
    L = (C.c_double * 3)(6.4, 4.8, 5.0)
    out = tmp_path / "density_7.dat"
    fs = 0.0625
    rc = scmf.lib().scmf_write_density_host(str(out).encode(), n, L, T, counts.ctypes.data,
                                            omega.ctypes.data, C.c_double(fs))
    assert rc == 0
    got = out.read_text().split("\n")
    assert got[0] == f"Nx={nx}, Ny={ny}, Nz={nz}"
    assert len(got) == 2 + cells + 1 and got[-1] == ""
    for c in (0, 1, 65535, 65536, 65537, 131071, 131072, cells - 1):
        vals = [counts[c + t * cells] * fs for t in range(T)] + \
               [float(omega[c + t * cells]) for t in range(T)]
        assert got[2 + c] == "".join("%g " % v for v in vals), c
    assert not os.path.exists(str(out) + ".tmp")


@pytest.mark.parametrize("T", [1, 2, 3])
def test_vtk_from_counts_equals_converter(tmp_path, T):
    """The VTK file written straight from counts and omega (the background
    writer's path) is byte-equal to converting the density file written from
    them: every value read back through its "%g" text."""
    nx, ny, nz = 40, 30, 60                   # 72,000 cells: more than one block
    cells = nx * ny * nz
    rng = np.random.default_rng(5 + T)
    counts = rng.integers(0, 40, size=T * cells).astype(np.uint32)
    counts[:7] = [0, 65535, 123456, 1, 2, 3, 99999]
    omega = (rng.standard_normal(T * cells) * 7).astype(np.float32)
    n = (C.c_int32 * 3)(nx, ny, nz)
    L = (C.c_double * 3)(5.0, 3.3, 7.77)
    fs = 13824 / 409600
    dens = str(tmp_path / "density_12.dat")
    lib = scmf.lib()
    assert lib.scmf_write_density_host(dens.encode(), n, L, T, counts.ctypes.data,
                                       omega.ctypes.data, C.c_double(fs)) == 0
    conv = tmp_path / "conv.vtk"
    assert lib.scmf_density_to_vtk(dens.encode(), str(conv).encode()) == 0
    direct = tmp_path / "direct.vtk"
    assert lib.scmf_write_vtk_host(str(direct).encode(), dens.encode(), n, L, T, counts.ctypes.data,
                                   omega.ctypes.data, C.c_double(fs)) == 0
    assert direct.read_bytes() == conv.read_bytes()
    assert not os.path.exists(str(direct) + ".tmp")


def test_density_file_names():
    L = scmf.lib()
    # the name rule of io.cpp:16-19 is exercised through the host writer path
    assert L.scmf_version()


def test_vtk_matches_reference(golden, tmp_path):
    src = tmp_path / "density_240000.dat"
    src.write_bytes(gzip.open(os.path.join(golden, "density_240000.dat.gz")).read())
    out = tmp_path / "out.vtk"
    scmf.density_to_vtk(src, out)
    ref = gzip.open(os.path.join(golden, "out.vtk.gz"), "rt").read().splitlines()
    got = out.read_text().splitlines()
    assert len(got) == len(ref)
    assert got[1].startswith("# comment: converted from file ")
    assert got[0] == ref[0] and got[2:] == ref[2:]


def test_create_rejects_bad_descriptions():
    spec = W.tiny()
    d, keep = spec.to_desc()
    h = C.c_void_p()
    o = make_opts()
    L = scmf.lib()
    d.n[1] = 0
    assert L.scmf_create(C.byref(h), C.byref(d), C.byref(o)) == -1
    assert b"grid" in L.scmf_last_error()
    d.n[1] = 8
    d.reference_Nbeads = 1
    assert L.scmf_create(C.byref(h), C.byref(d), C.byref(o)) == -1
    d.reference_Nbeads = 16
    o.density_bits = 8
    rc = L.scmf_create(C.byref(h), C.byref(d), C.byref(o))
    assert rc in (-1, -6)
    assert not h.value


def test_null_handles_are_errors():
    L = scmf.lib()
    assert L.scmf_mc_sweep(None, 1) == -1
    assert L.scmf_grid_density(None) == -1
    assert L.scmf_update_omega(None) == -1
    assert L.scmf_step(None, 1) == -1
    L.scmf_destroy(None)
