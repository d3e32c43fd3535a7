"""MI355X-native single-chain-in-mean-field Monte Carlo engine.

The compute path is libscmf.so (hand-written HIP for gfx950 behind the C ABI of
include/scmf.h); `new_arch_amd.scmf` binds it, `new_arch_amd.abi` mirrors the
ABI's structs, `new_arch_amd.workloads` defines the synthetic systems.
"""
from .abi import PolymerType, SystemSpec, make_opts  # noqa: F401

__all__ = ["PolymerType", "SystemSpec", "make_opts"]
