"""Structure factor S(q) on the device (hipFFT) against numpy's full complex
FFT of the same counts, shell sums within a tolerance scaled by the total
power, mode counts exact; closed forms for a checkerboard along x and a
uniform density."""
import math

import numpy as np
import pytest

from new_arch_amd import scmf
from new_arch_amd import workloads as W
from new_arch_amd.abi import PolymerType, SystemSpec, make_opts

pytestmark = pytest.mark.gpu

TWO_PI = 6.283185307179586


def sq_numpy(counts, spec, n_beads, coef, dq, by_q2, n_bins):
    """Shell sums of |FFT(w)|^2 / n_beads over the full spectrum, w = coef . n."""
    nx, ny, nz = spec.n
    c = np.asarray(counts, np.float64).reshape(spec.n_types, nz, ny, nx)
    w = np.tensordot(np.asarray(coef, np.float64), c, axes=1)
    F = np.fft.fftn(w)
    S = (F.real ** 2 + F.imag ** 2) / n_beads
    kz, ky, kx = np.meshgrid(np.fft.fftfreq(nz, 1.0 / nz), np.fft.fftfreq(ny, 1.0 / ny),
                             np.fft.fftfreq(nx, 1.0 / nx), indexing="ij")
    qx = (TWO_PI / spec.L[0]) * kx
    qy = (TWO_PI / spec.L[1]) * ky
    qz = (TWO_PI / spec.L[2]) * kz
    q2 = qx * qx + qy * qy + qz * qz
    inv = 1.0 / dq
    u = q2 * inv * inv if by_q2 else np.sqrt(q2) * inv
    b = np.rint(u).astype(np.int64).ravel()
    keep = b < n_bins
    sums = np.bincount(b[keep], weights=S.ravel()[keep], minlength=n_bins)
    modes = np.bincount(b[keep], minlength=n_bins)
    return sums, modes, S.sum()


def check(g, spec, coef, dq, by_q2, n_bins):
    sums, modes = g.structure_factor(coef, dq, n_bins, by_q2=by_q2)
    want, want_modes, total = sq_numpy(g.density(), spec, g.n_beads, coef, dq, by_q2, n_bins)
    assert np.array_equal(modes, want_modes)
    np.testing.assert_allclose(sums, want, rtol=1e-9, atol=1e-11 * total)
    return sums, modes


CASES = {
    # cubic box, integer |k|^2 shells (dq = 2 pi / L), A-B contrast
    "cubic_k2": (W.tiny(n_chains=600, N=16, grid=(16, 16, 16), L=(2.0, 2.0, 2.0)),
                 (1.0, -1.0), TWO_PI / 2.0, True, 200, 16),
    # odd, non-cubic grid, |q| shells, three types
    "odd_noncubic": (SystemSpec(n_types=3, xn=W.xn_matrix(3, 30.0, 12.0), L=[1.5, 1.2, 0.9],
                                n=[15, 12, 9], dtNxi=0.2, nref=12,
                                polymers=[PolymerType(length=12, count=400,
                                                      monotype=[t % 3 for t in range(12)])]),
                     (0.5, -1.25, 2.0), 2.0, False, 64, 32),
    # uint16 counts, even grid with an odd axis, shells cut short by n_bins
    "u16_cut": (W.tiny(n_chains=500, N=20, grid=(12, 10, 14), L=(1.8, 1.5, 2.1)),
                (1.0, 1.0), 3.0, False, 5, 16),
}


@pytest.mark.parametrize("case", list(CASES))
def test_structure_factor_matches_numpy(case):
    spec, coef, dq, by_q2, n_bins, bits = CASES[case]
    print(f"[gpu-sq] {case}", flush=True)
    g = scmf.System(spec, make_opts(seed=31, density_bits=bits))
    try:
        g.grid_density()
        g.update_omega()
        check(g, spec, coef, dq, by_q2, n_bins)
        g.mc_sweep(3)
        g.grid_density()
        sums, modes = check(g, spec, coef, dq, by_q2, n_bins)
        if n_bins >= 64:
            assert modes.sum() == g.n_cells      # every mode lands in a shell
        again, _ = g.structure_factor(coef, dq, n_bins, by_q2=by_q2)
        np.testing.assert_allclose(again, sums, rtol=1e-12, atol=1e-14 * sums.sum())
    finally:
        g.close()


def test_structure_factor_closed_forms():
    """w = (-1)^x puts all power into the single mode q = (pi nx / Lx, 0, 0):
    S = n_cells^2 / n_beads there and 0 elsewhere; a uniform w = 1 puts it all
    at q = 0."""
    print("[gpu-sq] closed forms", flush=True)
    nx, ny, nz = 8, 6, 10
    spec = W.tiny(n_chains=300, N=8, grid=(nx, ny, nz), L=(1.6, 1.2, 2.0))
    g = scmf.System(spec, make_opts(seed=5, density_bits=32))
    try:
        C = nx * ny * nz
        x = np.arange(C) % nx
        counts = np.zeros((2, C), np.uint32)
        counts[0, x % 2 == 0] = 1
        counts[1, x % 2 == 1] = 1
        g.set_density(counts.ravel())
        dq = TWO_PI / spec.L[0]
        sums, modes = g.structure_factor((1.0, -1.0), dq, 64)
        peak = C * C / g.n_beads
        assert modes.sum() == C
        assert math.isclose(sums[nx // 2], peak, rel_tol=1e-12)
        rest = np.delete(sums, nx // 2)
        assert np.all(np.abs(rest) < 1e-12 * peak)
        one = g.structure_factor((1.0, 1.0), dq, 64)[0]
        assert math.isclose(one[0], peak, rel_tol=1e-12)
        assert np.all(np.abs(one[1:]) < 1e-12 * peak)
    finally:
        g.close()


def test_structure_factor_rejects_bad_arguments():
    g = scmf.System(W.tiny(n_chains=50, N=8), make_opts(seed=1))
    try:
        g.grid_density()
        for dq, n_bins in ((0.0, 8), (-1.0, 8), (float("nan"), 8), (1.0, 0), (1.0, 1 << 21)):
            with pytest.raises(scmf.ScmfError):
                g.structure_factor((1.0, -1.0), dq, n_bins)
        with pytest.raises(scmf.ScmfError):
            g.structure_factor((1.0, float("inf")), 1.0, 8)
        with pytest.raises(ValueError):
            g.structure_factor((1.0,), 1.0, 8)
    finally:
        g.close()
