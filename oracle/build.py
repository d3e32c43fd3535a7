"""Build recipe for the CPU oracle (test infrastructure; see scmf_oracle.cpp).

g++ with FP contraction off so every fp32/fp64 operation rounds once, exactly
as the HIP kernels (also built contraction-off) do.  -march=x86-64-v3 keeps
the library loadable on any AVX2 host; OpenMP parallelises over chains.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build():
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-std=c++17", "-fPIC", "-shared", "-fopenmp",
           "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wno-unknown-pragmas",
           "-o", tmp, os.path.join(HERE, "scmf_oracle.cpp")]
    print("+ " + " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
