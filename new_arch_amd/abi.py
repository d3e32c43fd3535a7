"""ctypes mirror of include/scmf.h and the Python-side system description.

This module loads nothing; new_arch_amd.scmf binds the HIP library.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

SCMF_OK = 0
SCMF_E_ARG = -1
SCMF_E_IO = -2
SCMF_E_HIP = -3
SCMF_E_NOMEM = -4
SCMF_E_OVERFLOW = -5
SCMF_E_NODEV = -6
SCMF_E_BUDGET = -7

MOVE_TRIAL = 0
MOVE_SMART = 1

POS = 0
DENSITY = 1
OMEGA = 2


class Desc(C.Structure):
    _fields_ = [
        ("n_mono_types", C.c_int32),
        ("xn", C.POINTER(C.c_double)),
        ("L", C.c_double * 3),
        ("n", C.c_int32 * 3),
        ("dtNxi", C.c_double),
        ("reference_Nbeads", C.c_int32),
        ("n_poly_types", C.c_int32),
        ("poly_length", C.POINTER(C.c_int32)),
        ("n_polymers", C.POINTER(C.c_int64)),
        ("monotype", C.POINTER(C.c_int32)),
        ("connection", C.POINTER(C.c_uint8)),
    ]


class Opts(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("system_id", C.c_uint32),
        ("move_type", C.c_int32),
        ("trial_delta", C.c_float),
        ("omega_every_step", C.c_int32),
        ("density_bits", C.c_int32),
        ("init_coords", C.c_int32),
        ("lanes_per_chain", C.c_int32),
        ("sort_every", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("moves", C.c_uint64),
        ("accepts", C.c_uint64),
        ("sweeps", C.c_uint64),
        ("re2", C.c_double),
        ("overflow", C.c_int32),
        ("grid_fallback", C.c_uint64),
        ("sorts", C.c_uint64),
        ("cell_grids", C.c_uint64),
    ]


class Timing(C.Structure):
    _fields_ = [
        ("sweep_ms", C.c_double),
        ("grid_ms", C.c_double),
        ("omega_ms", C.c_double),
        ("sweep_n", C.c_uint64),
        ("grid_n", C.c_uint64),
        ("omega_n", C.c_uint64),
    ]


class Summary(C.Structure):
    _fields_ = [
        ("n_mono_types", C.c_int32),
        ("n_poly_types", C.c_int32),
        ("reference_Nbeads", C.c_int32),
        ("n", C.c_int32 * 3),
        ("L", C.c_double * 3),
        ("dtNxi", C.c_double),
        ("n_chains", C.c_int64),
        ("n_beads", C.c_int64),
        ("n_cells", C.c_int64),
        ("xn", C.c_double * 64),
        ("device_bytes_16", C.c_uint64),
        ("device_bytes_32", C.c_uint64),
    ]


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_jobs", C.c_int32),
                ("steps_per_tick", C.c_int32), ("job_timeout_s", C.c_double)]


JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_RELEASED, JOB_FAILED, JOB_CANCELLED = range(6)
CANCEL_NONE, CANCEL_CALLER, CANCEL_TIMEOUT = range(3)


class ServerStats(C.Structure):
    _fields_ = [("ticks", C.c_uint64), ("rebuilds", C.c_uint64), ("admitted", C.c_uint64),
                ("finished", C.c_uint64), ("cancelled", C.c_uint64), ("failed", C.c_uint64),
                ("busy_s", C.c_double), ("span_s", C.c_double), ("used_bytes", C.c_uint64),
                ("captures", C.c_uint64)]


class BatchStats(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("captures", C.c_uint64), ("table_uploads", C.c_uint64),
                ("members", C.c_int32), ("member_capacity", C.c_int32),
                ("blocks_launched", C.c_uint64), ("blocks_used", C.c_uint64)]


class JobInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("reason", C.c_int32), ("sweeps_done", C.c_int64),
                ("sweeps_target", C.c_int64), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double), ("device_bytes", C.c_uint64), ("beads", C.c_int64)]


def linear_bonds(n: int) -> np.ndarray:
    """Bond matrix of a linear chain, as polymer.dat stores it (row i, column j)."""
    m = np.zeros((n, n), dtype=np.uint8)
    idx = np.arange(n - 1)
    m[idx, idx + 1] = 1
    m[idx + 1, idx] = 1
    return m


@dataclass
class PolymerType:
    length: int
    count: int
    monotype: List[int]
    bonds: Optional[np.ndarray] = None  # N x N; None = linear chain

    def bond_matrix(self) -> np.ndarray:
        return linear_bonds(self.length) if self.bonds is None else np.asarray(self.bonds, np.uint8)


@dataclass
class SystemSpec:
    """The content of configuration.dat + polymer.dat."""
    n_types: int
    xn: List[float]              # flat, xn[t + T*s]
    L: List[float]
    n: List[int]
    dtNxi: float
    nref: int
    polymers: List[PolymerType] = field(default_factory=list)

    @property
    def n_chains(self) -> int:
        return sum(p.count for p in self.polymers)

    @property
    def n_beads(self) -> int:
        return sum(p.count * p.length for p in self.polymers)

    @property
    def n_cells(self) -> int:
        return int(self.n[0]) * int(self.n[1]) * int(self.n[2])

    def bead_types(self) -> np.ndarray:
        """Monomer type of every bead in reference order."""
        return np.concatenate([np.tile(np.asarray(p.monotype, np.int32), p.count)
                               for p in self.polymers])

    def chain_first_last(self):
        firsts, lasts = [], []
        base = 0
        for p in self.polymers:
            f = base + np.arange(p.count, dtype=np.int64) * p.length
            firsts.append(f)
            lasts.append(f + p.length - 1)
            base += p.count * p.length
        return np.concatenate(firsts), np.concatenate(lasts)

    def to_desc(self):
        """Return (Desc, keepalive) — keep `keepalive` referenced while Desc is used."""
        xn = np.ascontiguousarray(self.xn, dtype=np.float64)
        lens = np.array([p.length for p in self.polymers], np.int32)
        counts = np.array([p.count for p in self.polymers], np.int64)
        mono = np.ascontiguousarray(np.concatenate([np.asarray(p.monotype, np.int32)
                                                    for p in self.polymers]), np.int32)
        conn = np.ascontiguousarray(np.concatenate([p.bond_matrix().ravel()
                                                    for p in self.polymers]), np.uint8)
        d = Desc()
        d.n_mono_types = self.n_types
        d.xn = xn.ctypes.data_as(C.POINTER(C.c_double))
        for i in range(3):
            d.L[i] = float(self.L[i])
            d.n[i] = int(self.n[i])
        d.dtNxi = float(self.dtNxi)
        d.reference_Nbeads = int(self.nref)
        d.n_poly_types = len(self.polymers)
        d.poly_length = lens.ctypes.data_as(C.POINTER(C.c_int32))
        d.n_polymers = counts.ctypes.data_as(C.POINTER(C.c_int64))
        d.monotype = mono.ctypes.data_as(C.POINTER(C.c_int32))
        d.connection = conn.ctypes.data_as(C.POINTER(C.c_uint8))
        return d, (xn, lens, counts, mono, conn)


def make_opts(seed=1, system_id=0, move="smart", trial_delta=0.1, omega_every_step=False,
              density_bits=32, init_coords=True, lanes_per_chain=0, sort_every=0) -> Opts:
    o = Opts()
    o.seed = int(seed)
    o.system_id = int(system_id)
    o.move_type = MOVE_SMART if move == "smart" else MOVE_TRIAL
    o.trial_delta = float(trial_delta)
    o.omega_every_step = 1 if omega_every_step else 0
    o.density_bits = int(density_bits)
    o.init_coords = 1 if init_coords else 0
    o.lanes_per_chain = int(lanes_per_chain)
    o.sort_every = int(sort_every)
    return o
