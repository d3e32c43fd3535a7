"""Regenerate the committed fixtures from the reference's own data files.

Run in a container that has the reference checkout at /root/reference (it is
not needed afterwards):

    python tests/golden/make_fixtures.py

Writes, next to this script:
  reference_inputs/configuration.dat, polymer.dat   the reference's input files
  density_240000.dat.gz, out.vtk.gz                 a density dump and the VTK the
                                                    reference's vtk.c made from it
  density_fixtures.npz   per dump: the integer counts the density columns encode,
                         the integer counts the omega columns decode to (the
                         previous step's; omega is not refreshed on steps that are
                         multiples of 10), and the printed omega values
  philox_rocrand.npz     Philox4x32-10 outputs computed by ROCm's rocRAND header
"""
import gzip
import os
import shutil
import subprocess
import tempfile

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

# Dumps to decode: (step, kappaN, chiN).  The 24^3 dumps decode with the
# shipped kappaN=50, chiN=30; the 6x60x60 dump with kappaN=120, chiN=80.
DUMPS = [(240000, 50.0, 30.0), (300000, 50.0, 30.0), (999000, 120.0, 80.0)]
NREF = 32
N_BEADS = 409600


def read_dump(path):
    with open(path) as f:
        h1 = f.readline()
        h2 = f.readline()
        rows = np.loadtxt(f)
    nx, ny, nz = [int(t.split("=")[1]) for t in h1.strip().split(", ")]
    return (nx, ny, nz), h2.strip(), rows


def omega_matrix(T, kappa, chi, scale):
    """omega = M n + c for the update of device_function.cu:373-412."""
    M = np.zeros((T, T))
    c = np.full(T, -kappa / NREF)
    for t in range(T):
        M[t, :] += kappa * scale / NREF
        for s in range(T):
            if s != t:
                M[t, t] += (-chi / 2) * scale / NREF
                M[t, s] -= (-chi / 2) * scale / NREF
    return M, c


def decode_dumps():
    out = {}
    for step, kappa, chi in DUMPS:
        grid, h2, rows = read_dump(os.path.join(REF, f"density_{step}.dat"))
        C = grid[0] * grid[1] * grid[2]
        T = rows.shape[1] // 2
        assert rows.shape == (C, 2 * T)
        scale = C / N_BEADS
        dens, om = rows[:, :T], rows[:, T:]
        counts = dens / scale
        assert np.abs(counts - np.rint(counts)).max() < 5e-3, step
        M, c = omega_matrix(T, kappa, chi, scale)
        prev = np.linalg.solve(M, (om - c).T).T
        dev = np.abs(prev - np.rint(prev)).max()
        assert dev < 5e-3, (step, dev)
        out[f"grid_{step}"] = np.array(grid, np.int32)
        out[f"header2_{step}"] = np.array(h2)
        out[f"kappa_chi_{step}"] = np.array([kappa, chi])
        out[f"counts_{step}"] = np.rint(counts).astype(np.uint16)
        out[f"prev_{step}"] = np.rint(prev).astype(np.uint16)
        # float32 keeps every printed 6-significant-digit value recoverable
        out[f"omega_{step}"] = om.astype(np.float32)
    np.savez_compressed(os.path.join(HERE, "density_fixtures.npz"), **out)


PHILOX_SRC = r"""
#include <rocrand/rocrand_philox4x32_10.h>
#include <stdio.h>
// ten_rounds is protected; expose it to compute raw blocks.
struct Open : rocrand_device::philox4x32_10_engine {
  Open() : rocrand_device::philox4x32_10_engine(0ull, 0ull, 0ull) {}
  uint4 block(uint4 c, uint2 k) { return ten_rounds(c, k); }
};
int main() {
  unsigned s = 12345u;
  Open eng;
  for (int i = 0; i < 256; ++i) {
    unsigned v[6];
    for (int k = 0; k < 6; ++k) { s = s * 1664525u + 1013904223u; v[k] = s; }
    if (i == 0) for (int k = 0; k < 6; ++k) v[k] = 0u;
    if (i == 1) for (int k = 0; k < 6; ++k) v[k] = 0xffffffffu;
    uint4 c = {v[0], v[1], v[2], v[3]};
    uint2 key = {v[4], v[5]};
    uint4 r = eng.block(c, key);
    printf("%u %u %u %u %u %u %u %u %u %u\n", v[0], v[1], v[2], v[3], v[4], v[5], r.x, r.y, r.z, r.w);
  }
  return 0;
}
"""


def philox_vectors():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "p.cpp")
        exe = os.path.join(d, "p")
        with open(src, "w") as f:
            f.write(PHILOX_SRC)
        subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O1", "-o", exe, src], check=True)
        txt = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    a = np.array([[int(t) for t in line.split()] for line in txt.strip().splitlines()],
                 dtype=np.uint64).astype(np.uint32)
    np.savez_compressed(os.path.join(HERE, "philox_rocrand.npz"), ctr=a[:, :4], key=a[:, 4:6],
                        out=a[:, 6:])


def copy_inputs():
    d = os.path.join(HERE, "reference_inputs")
    os.makedirs(d, exist_ok=True)
    for name in ("configuration.dat", "polymer.dat"):
        shutil.copyfile(os.path.join(REF, name), os.path.join(d, name))
    for name in ("density_240000.dat", "out.vtk"):
        with open(os.path.join(REF, name), "rb") as f, \
                open(os.path.join(HERE, name + ".gz"), "wb") as raw, \
                gzip.GzipFile(fileobj=raw, mode="wb", compresslevel=9, mtime=0) as g:
            g.write(f.read())


if __name__ == "__main__":
    copy_inputs()
    decode_dumps()
    philox_vectors()
    print("fixtures written to", HERE)
