"""ctypes binding of liboracle.so — the CPU restatement of the reference path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker and the CPU baseline.  The engine
never uses it.  Positions are float [bead][xyz] in reference order.
"""
import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_HERE)
if _ROOT not in sys.path:
    sys.path.insert(0, _ROOT)

from new_arch_amd.abi import Desc, Opts, SystemSpec, make_opts  # noqa: E402

LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class Derived(C.Structure):
    _fields_ = [("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("inverse_refbeads", C.c_double), ("harmonic_normb", C.c_double),
                ("field_scaling", C.c_double), ("A", C.c_double), ("R", C.c_double),
                ("n_cells", C.c_int64), ("n_chains", C.c_int64), ("n_beads", C.c_int64)]


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} missing: build it with oracle/build.py")
    L = C.CDLL(LIB_PATH)
    pD, pO, vp = C.POINTER(Desc), C.POINTER(Opts), C.c_void_p
    L.orc_derive.argtypes = [pD, C.POINTER(Derived)]
    L.orc_bin_index.argtypes = [pD, C.c_double, C.c_double, C.c_double, C.c_uint]
    L.orc_bin_index.restype = C.c_uint
    L.orc_bin_many.argtypes = [pD, vp, C.c_int64, vp, vp]
    L.orc_bin_many.restype = None
    L.orc_grid.argtypes = [pD, vp, vp]
    L.orc_omega.argtypes = [pD, vp, vp]
    L.orc_init_coords.argtypes = [pD, pO, vp]
    L.orc_sweep.argtypes = [pD, pO, vp, vp, C.c_uint32, C.c_int64, C.c_int64,
                            C.POINTER(C.c_uint64)]
    L.orc_sweep_injected.argtypes = [pD, pO, vp, vp, vp, vp, vp, C.POINTER(C.c_uint64)]
    L.orc_re2.argtypes = [pD, vp]
    L.orc_re2.restype = C.c_double
    L.orc_philox.argtypes = [C.c_uint32] * 6 + [vp]
    L.orc_philox.restype = None
    L.orc_pick_disagreements.argtypes = [vp, C.c_int64, C.c_int32]
    L.orc_pick_disagreements.restype = C.c_int64
    for n in ("orc_exp_many", "orc_log_many"):
        getattr(L, n).argtypes = [vp, vp, C.c_int64]
        getattr(L, n).restype = None
    L.orc_sincos_many.argtypes = [vp, vp, vp, C.c_int64]
    L.orc_sincos_many.restype = None
    L.orc_normals.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int64, vp]
    L.orc_normals.restype = None
    L.orc_logf_many.argtypes = [vp, vp, C.c_int64]
    L.orc_logf_many.restype = None
    L.orc_sincosf_turn24_many.argtypes = [vp, vp, vp, C.c_int64]
    L.orc_sincosf_turn24_many.restype = None
    L.orc_normals_f.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int64, vp]
    L.orc_normals_f.restype = None
    L.orc_normal3f_words.argtypes = [vp, C.c_int64, vp]
    L.orc_normal3f_words.restype = None
    L.orc_threads.restype = C.c_int
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data


class Oracle:
    """CPU reference for one system description."""

    def __init__(self, spec: SystemSpec, opts: Opts = None, **kw):
        self.spec = spec
        self.desc, self._keep = spec.to_desc()
        self.opts = opts if opts is not None else make_opts(**kw)
        d = Derived()
        if lib().orc_derive(C.byref(self.desc), C.byref(d)) != 0:
            raise ValueError("oracle rejected the description")
        self.derived = d
        self.n_beads = d.n_beads
        self.n_cells = d.n_cells
        self.n_types = spec.n_types

    def init_coords(self) -> np.ndarray:
        pos = np.empty((self.n_beads, 3), np.float32)
        lib().orc_init_coords(C.byref(self.desc), C.byref(self.opts), _ptr(pos))
        return pos

    def grid(self, pos) -> np.ndarray:
        pos = np.ascontiguousarray(pos, np.float32)
        out = np.empty(self.n_types * self.n_cells, np.uint32)
        lib().orc_grid(C.byref(self.desc), _ptr(pos), _ptr(out))
        return out

    def omega(self, counts) -> np.ndarray:
        counts = np.ascontiguousarray(counts, np.uint32)
        out = np.empty(self.n_types * self.n_cells, np.float64)
        lib().orc_omega(C.byref(self.desc), _ptr(counts), _ptr(out))
        return out

    def sweep(self, pos, omega32, sweep, chain_begin=0, chain_end=-1):
        """In place on pos (float32 [beads][3]); returns accepted moves."""
        assert pos.dtype == np.float32 and pos.flags.c_contiguous
        om = np.ascontiguousarray(omega32, np.float32)
        acc = C.c_uint64()
        lib().orc_sweep(C.byref(self.desc), C.byref(self.opts), _ptr(pos), _ptr(om),
                        int(sweep), int(chain_begin), int(chain_end), C.byref(acc))
        return acc.value

    def sweep_injected(self, pos, omega32, pick, gauss, acc_u):
        """One sweep in place on pos, driven by injected random numbers: move
        i of the chain whose first bead is b takes pick[b+i] in [0,1] (1 picks the last bead),
        gauss[b+i] (three normals; three uniforms in [0,1) for trial moves)
        and acc_u[b+i] in (0,1].  Returns accepted moves."""
        assert pos.dtype == np.float32 and pos.flags.c_contiguous
        n = self.n_beads
        pick = np.ascontiguousarray(pick, np.float64)
        gauss = np.ascontiguousarray(gauss, np.float32)
        acc_u = np.ascontiguousarray(acc_u, np.float32)
        assert pick.size == n and gauss.size == 3 * n and acc_u.size == n
        om = np.ascontiguousarray(omega32, np.float32)
        acc = C.c_uint64()
        lib().orc_sweep_injected(C.byref(self.desc), C.byref(self.opts), _ptr(pos), _ptr(om),
                                 _ptr(pick), _ptr(gauss), _ptr(acc_u), C.byref(acc))
        return acc.value

    def re2(self, pos) -> float:
        pos = np.ascontiguousarray(pos, np.float32)
        return lib().orc_re2(C.byref(self.desc), _ptr(pos))

    def bin_many(self, r, types) -> np.ndarray:
        r = np.ascontiguousarray(r, np.float64)
        t = np.ascontiguousarray(types, np.uint32)
        out = np.empty(t.size, np.uint32)
        lib().orc_bin_many(C.byref(self.desc), _ptr(r), t.size, _ptr(t), _ptr(out))
        return out


def pick_disagreements(words, n):
    """Picks (moves 0 and 1 of each 4-word block) on which the engine's
    pick_of and the defining pick_bead(pick_u01(.)) differ, for chains of n beads."""
    w = np.ascontiguousarray(words, np.uint32).reshape(-1, 4)
    return int(lib().orc_pick_disagreements(_ptr(w), w.shape[0], int(n)))


def philox(ctr, key):
    out = np.empty(4, np.uint32)
    lib().orc_philox(*[int(c) for c in ctr], *[int(k) for k in key], _ptr(out))
    return out


def exp_many(x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    lib().orc_exp_many(_ptr(x), _ptr(y), x.size)
    return y


def log_many(x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    lib().orc_log_many(_ptr(x), _ptr(y), x.size)
    return y


def sincos_many(u):
    u = np.ascontiguousarray(u, np.float64)
    s = np.empty_like(u)
    c = np.empty_like(u)
    lib().orc_sincos_many(_ptr(u), _ptr(s), _ptr(c), u.size)
    return s, c


def normals(seed, system, chain, sweep, n):
    out = np.empty((n, 3), np.float64)
    lib().orc_normals(int(seed), int(system), int(chain), int(sweep), int(n), _ptr(out))
    return out


def logf_many(x):
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    lib().orc_logf_many(_ptr(x), _ptr(y), x.size)
    return y


def sincosf_turn24_many(k):
    k = np.ascontiguousarray(k, np.uint32)
    s = np.empty(k.size, np.float32)
    c = np.empty(k.size, np.float32)
    lib().orc_sincosf_turn24_many(_ptr(k), _ptr(s), _ptr(c), k.size)
    return s, c


def normals_f(seed, system, chain, sweep, n):
    out = np.empty((n, 3), np.float32)
    lib().orc_normals_f(int(seed), int(system), int(chain), int(sweep), int(n), _ptr(out))
    return out


def normal3f_words(words):
    """normal3f of each 4-word Philox block in `words` (uint32, shape [n, 4])."""
    w = np.ascontiguousarray(words, np.uint32)
    out = np.empty((w.shape[0], 3), np.float32)
    lib().orc_normal3f_words(_ptr(w), w.shape[0], _ptr(out))
    return out


def threads() -> int:
    return lib().orc_threads()
