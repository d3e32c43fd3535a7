"""Build recipe for the native parts: libscmf.so + scmf_run (HIP, gfx950),
libscmf_calib.so (kernels of known VALU instruction count, for measurement)
and libscmf_check.so (the device builds of the shared arithmetic primitives
over given inputs, for the tests).

Every HIP translation unit is compiled with FP contraction off (results must
match the CPU oracle bit for bit) and with packed-FP32 instructions disabled.
The oracle's own library is built by oracle/build.py.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libscmf.so")
RUNNER = os.path.join(LIBDIR, "scmf_run")
CALIB = os.path.join(LIBDIR, "libscmf_calib.so")  # VALU-rate calibration kernels (measurement)
CHECK = os.path.join(LIBDIR, "libscmf_check.so")  # device builds of the primitives (tests)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    print("+ " + " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", *COMMON, "-shared", "-o", tmp,
          os.path.join(CSRC, "scmf_engine.hip"), os.path.join(CSRC, "scmf_host.cpp"),
          os.path.join(CSRC, "scmf_server.cpp"), "-L/opt/rocm/lib", "-lhipfft",
          "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    tmp = CALIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", *COMMON, "-shared", "-o", tmp,
          os.path.join(CSRC, "scmf_calib.hip")])
    os.replace(tmp, CALIB)
    tmp = CHECK + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", *COMMON, "-shared", "-o", tmp,
          os.path.join(CSRC, "scmf_check.hip")])
    os.replace(tmp, CHECK)
    tmp = RUNNER + ".tmp"
    _run([HIPCC, *COMMON, "-o", tmp, os.path.join(CSRC, "scmf_main.cpp"),
          f"-L{LIBDIR}", "-lscmf", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, RUNNER)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
