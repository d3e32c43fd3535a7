"""VALU issue-rate calibration: kernels with a known instruction stream.

`python3 -m new_arch_amd.calib` times each kernel of libscmf_calib.so with HIP
events and prints one JSON line: per kind, the VALU instructions one wave
issues (from the loop body hipcc emits for gfx950, csrc/scmf_calib.hip), the
waves launched and the measured time.  Run under rocprofv3 with
SQ_INSTS_VALU / SQ_ACTIVE_INST_VALU / SQ_ACTIVE_INST_VALU2 / GRBM_GUI_ACTIVE,
the same run gives the factor between those counters and SIMD issue cycles,
which the engine's counter figures are read against.
"""
import ctypes as C
import json
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libscmf_calib.so")

# VALU instructions per loop iteration of each kind (csrc/scmf_calib.hip)
KINDS = {0: ("v_fma_f32", 8), 1: ("v_fma_f64", 8), 2: ("v_mad_u64_u32 + v_xor_b32", 16),
         3: ("v_fma_f64 + v_lshrrev_b32 + v_xor_b32", 12)}


def run(blocks=8192, iters=4096, reps=5):
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with new_arch_amd/build.py")
    lib = C.CDLL(LIB_PATH)
    lib.scmf_calib_valu.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float)]
    out = {"blocks": blocks, "lanes_per_block": 256, "waves": blocks * 4, "iters": iters,
           "reps": reps, "kinds": {}}
    for kind, (name, per_iter) in KINDS.items():
        ms = C.c_float()
        rc = lib.scmf_calib_valu(kind, blocks, iters, reps, C.byref(ms))
        if rc != 0:
            raise RuntimeError(f"calibration kernel {kind} failed: {rc}")
        out["kinds"][str(kind)] = {"ops": name, "valu_per_wave": per_iter * iters,
                                   "ms": float(ms.value)}
    return out


if __name__ == "__main__":
    print(json.dumps(run()), flush=True)
