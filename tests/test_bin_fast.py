"""The engine's fast NGP binning equals the reference computation on every
input class (random, unwrapped, cell and box edges to the ulp, fp32 edges,
old+displacement sums, huge and non-finite coordinates; for power-of-two
boxes also negative cell edges plus tiny steps and tiny negative fp32
coordinates, where the reference's k L + r rounds).

The harness tests/bin_selftest.cpp is compiled here with g++ and FP
contraction off, exactly as the engine's arithmetic must behave on the GPU.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("bin") / "bin_selftest")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                    "-march=x86-64-v3", "-o", exe, os.path.join(HERE, "bin_selftest.cpp")],
                   check=True)
    return exe


@pytest.mark.parametrize("L,n,pow2", [(4.0, 24, 0), (4.0, 32, 1), (32.0, 256, 1), (64.0, 512, 1),
                                      (1.0, 6, 0), (3.0, 12, 1), (2.5, 10, 1), (0.7, 7, 0),
                                      (100.0, 1000, 0), (2.0, 16, 1), (0.5, 64, 1),
                                      (1024.0, 64, 1), (3.0, 24, 1), (6.0, 3, 1), (1.0, 8, 1),
                                      (8.0, 4, 1)])
def test_fast_binning_equals_reference(harness, L, n, pow2):
    r = subprocess.run([harness, str(L), str(n), "2000000", str(n * 7 + 1), str(pow2)],
                       capture_output=True, text=True)
    checked, bad = map(int, r.stdout.split())
    assert bad == 0, r.stderr
    assert checked >= 2000000
