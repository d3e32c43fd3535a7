"""Python binding of libscmf.so (the HIP engine) through its C ABI.

The engine must be built (new_arch_amd/build.py); importing this module on a
machine without the library raises instead of falling back to anything else.
"""
import ctypes as C
import os

import numpy as np

from .abi import (DENSITY, OMEGA, POS, SCMF_OK, BatchStats, Desc, JobInfo, Opts, ServerOpts,
                  ServerStats, Stats, Summary, SystemSpec, Timing, make_opts)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libscmf.so")

_lib = None


class ScmfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"scmf error {code}: {msg}")
        self.code = code


def lib():
    """Load libscmf.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with new_arch_amd/build.py")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.scmf_create.argtypes = [C.POINTER(vp), C.POINTER(Desc), C.POINTER(Opts)]
    L.scmf_create_from_files.argtypes = [C.POINTER(vp), C.c_char_p, C.c_char_p, C.POINTER(Opts)]
    L.scmf_destroy.argtypes = [vp]
    L.scmf_destroy.restype = None
    for name in ("scmf_mc_sweep", "scmf_step", "scmf_enable_timing"):
        getattr(L, name).argtypes = [vp, C.c_int32]
    for name in ("scmf_grid_density", "scmf_update_omega", "scmf_reset_stats",
                 "scmf_reset_timing", "scmf_sync", "scmf_sort_chains"):
        getattr(L, name).argtypes = [vp]
    L.scmf_copy_out.argtypes = [vp, C.c_int32, vp, C.c_size_t]
    L.scmf_copy_out_noalloc.argtypes = [vp, C.c_int32, vp, C.c_size_t]
    L.scmf_copy_in_pos.argtypes = [vp, vp, C.c_size_t]
    L.scmf_copy_in_omega.argtypes = [vp, vp, C.c_size_t]
    L.scmf_copy_in_density.argtypes = [vp, vp, C.c_size_t]
    L.scmf_set_rng.argtypes = [vp, C.c_uint64, C.c_uint64]
    L.scmf_stats.argtypes = [vp, C.POINTER(Stats)]
    L.scmf_timing.argtypes = [vp, C.POINTER(Timing)]
    L.scmf_sizes.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                             C.POINTER(C.c_int64), C.POINTER(C.c_int32)]
    L.scmf_write_density.argtypes = [vp, C.c_char_p, C.c_int32]
    L.scmf_write_coord.argtypes = [vp, C.c_char_p]
    L.scmf_write_density_async.argtypes = [vp, C.c_char_p, C.c_int32]
    L.scmf_write_outputs_async.argtypes = [vp, C.c_char_p, C.c_int32, C.c_char_p]
    L.scmf_write_vtk_host.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_int32),
                                      C.POINTER(C.c_double), C.c_int32, vp, vp, C.c_double]
    L.scmf_output_wait.argtypes = [vp]
    L.scmf_checkpoint_write.argtypes = [vp, C.c_char_p, C.c_int64]
    L.scmf_checkpoint_read.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int64)]
    L.scmf_structure_factor.argtypes = [vp, vp, C.c_double, C.c_int32, C.c_int32, vp, vp]
    L.scmf_density_to_vtk.argtypes = [C.c_char_p, C.c_char_p]
    L.scmf_inspect_files.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(Summary)]
    L.scmf_inspect_desc.argtypes = [C.POINTER(Desc), C.POINTER(Summary)]
    L.scmf_write_density_host.argtypes = [C.c_char_p, C.POINTER(C.c_int32),
                                          C.POINTER(C.c_double), C.c_int32, vp, vp, C.c_double]
    L.scmf_autotune.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_float)]
    L.scmf_batch_create.argtypes = [C.POINTER(vp), C.POINTER(vp), C.c_int32]
    L.scmf_batch_step.argtypes = [vp, C.c_int32]
    L.scmf_batch_sync.argtypes = [vp]
    L.scmf_batch_set_lanes.argtypes = [vp, C.c_int32]
    L.scmf_batch_update.argtypes = [vp, C.POINTER(vp), C.c_int32, vp]
    L.scmf_batch_destroy.argtypes = [vp]
    L.scmf_batch_stats.argtypes = [vp, C.POINTER(BatchStats)]
    L.scmf_batch_destroy.restype = None
    L.scmf_server_create.argtypes = [C.POINTER(vp), C.POINTER(ServerOpts)]
    L.scmf_server_submit.argtypes = [vp, C.POINTER(Desc), C.POINTER(Opts), C.c_int32,
                                     C.POINTER(C.c_uint64)]
    L.scmf_server_tick.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.scmf_server_info.argtypes = [vp, C.c_uint64, C.POINTER(JobInfo)]
    L.scmf_server_job_system.argtypes = [vp, C.c_uint64, C.POINTER(vp)]
    L.scmf_server_release.argtypes = [vp, C.c_uint64]
    L.scmf_server_cancel.argtypes = [vp, C.c_uint64]
    L.scmf_server_set_timeout.argtypes = [vp, C.c_uint64, C.c_double]
    L.scmf_server_stats.argtypes = [vp, C.POINTER(ServerStats)]
    L.scmf_server_clock.argtypes = [vp]
    L.scmf_server_clock.restype = C.c_double
    L.scmf_server_last_error.restype = C.c_char_p
    L.scmf_server_destroy.argtypes = [vp]
    L.scmf_server_destroy.restype = None
    L.scmf_last_error.restype = C.c_char_p
    L.scmf_version.restype = C.c_char_p
    _lib = L
    return L


def _check(rc):
    if rc != SCMF_OK:
        raise ScmfError(rc, lib().scmf_last_error().decode())


def inspect_files(configuration_dat, polymer_dat) -> Summary:
    s = Summary()
    _check(lib().scmf_inspect_files(str(configuration_dat).encode(), str(polymer_dat).encode(),
                                    C.byref(s)))
    return s


def inspect_spec(spec: SystemSpec) -> Summary:
    d, keep = spec.to_desc()
    s = Summary()
    _check(lib().scmf_inspect_desc(C.byref(d), C.byref(s)))
    return s


def density_to_vtk(density_path, vtk_path):
    _check(lib().scmf_density_to_vtk(str(density_path).encode(), str(vtk_path).encode()))


class System:
    """One SCMF system on the current GPU, driven through the C ABI."""

    def __init__(self, spec: SystemSpec = None, opts: Opts = None, *, files=None, **kw):
        L = lib()
        self._h = C.c_void_p()
        if opts is None:
            opts = make_opts(**kw)
        self.opts = opts
        if files is not None:
            cfg, poly = files
            _check(L.scmf_create_from_files(C.byref(self._h), str(cfg).encode(),
                                            str(poly).encode(), C.byref(opts)))
        else:
            d, self._keep = spec.to_desc()
            _check(L.scmf_create(C.byref(self._h), C.byref(d), C.byref(opts)))
        self.spec = spec
        nb, nc, ncell, nt = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        _check(L.scmf_sizes(self._h, C.byref(nb), C.byref(nc), C.byref(ncell), C.byref(nt)))
        self.n_beads, self.n_chains, self.n_cells, self.n_types = (nb.value, nc.value,
                                                                   ncell.value, nt.value)

    # -- lifetime
    def close(self):
        if self._h:
            lib().scmf_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- the reference call sequence
    def mc_sweep(self, n=1):
        _check(lib().scmf_mc_sweep(self._h, int(n)))

    def grid_density(self):
        _check(lib().scmf_grid_density(self._h))

    def update_omega(self):
        _check(lib().scmf_update_omega(self._h))

    def step(self, step):
        _check(lib().scmf_step(self._h, int(step)))

    def sync(self):
        _check(lib().scmf_sync(self._h))

    def sort_chains(self):
        _check(lib().scmf_sort_chains(self._h))

    # -- data movement
    def positions(self) -> np.ndarray:
        out = np.empty((self.n_beads, 3), np.float32)
        _check(lib().scmf_copy_out(self._h, POS, out.ctypes.data, out.nbytes))
        return out

    def density(self) -> np.ndarray:
        out = np.empty(self.n_types * self.n_cells, np.uint32)
        _check(lib().scmf_copy_out(self._h, DENSITY, out.ctypes.data, out.nbytes))
        return out

    def omega(self) -> np.ndarray:
        out = np.empty(self.n_types * self.n_cells, np.float32)
        _check(lib().scmf_copy_out(self._h, OMEGA, out.ctypes.data, out.nbytes))
        return out

    def copy_out_noalloc(self, what):
        """scmf_copy_out_noalloc of positions (POS), counts (DENSITY) or omega (OMEGA)."""
        n = 3 * self.n_beads if what == POS else self.n_types * self.n_cells
        out = np.empty(n, np.float32 if what in (POS, OMEGA) else np.uint32)
        _check(lib().scmf_copy_out_noalloc(self._h, what, out.ctypes.data, out.nbytes))
        return out.reshape(-1, 3) if what == POS else out

    def set_positions(self, xyz):
        a = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        _check(lib().scmf_copy_in_pos(self._h, a.ctypes.data, a.shape[0]))

    def set_omega(self, om):
        a = np.ascontiguousarray(om, np.float32).ravel()
        _check(lib().scmf_copy_in_omega(self._h, a.ctypes.data, a.size))

    def set_density(self, counts):
        a = np.ascontiguousarray(counts, np.uint32).ravel()
        _check(lib().scmf_copy_in_density(self._h, a.ctypes.data, a.size))

    def set_rng(self, seed, next_sweep):
        _check(lib().scmf_set_rng(self._h, int(seed), int(next_sweep)))

    def autotune(self, sweeps_per_variant=2):
        """Pick the fastest sweep kernel shape; returns (lanes_per_chain, ms/sweep)."""
        lanes, ms = C.c_int32(), C.c_float()
        _check(lib().scmf_autotune(self._h, int(sweeps_per_variant), C.byref(lanes),
                                   C.byref(ms)))
        return lanes.value, ms.value

    # -- diagnostics
    def stats(self) -> Stats:
        s = Stats()
        _check(lib().scmf_stats(self._h, C.byref(s)))
        return s

    def reset_stats(self):
        _check(lib().scmf_reset_stats(self._h))

    def enable_timing(self, on=True):
        _check(lib().scmf_enable_timing(self._h, 1 if on else 0))

    def timing(self) -> Timing:
        t = Timing()
        _check(lib().scmf_timing(self._h, C.byref(t)))
        return t

    def reset_timing(self):
        _check(lib().scmf_reset_timing(self._h))

    # -- analysis
    def structure_factor(self, coef, dq, n_bins, by_q2=False):
        """Shell sums of S(q) = |F[sum_t coef[t] n_t](q)|^2 / n_beads and the
        number of modes per shell (shell = nearest integer to |q|/dq, or to
        q^2/dq^2 with by_q2).  Returns (sums, n_modes)."""
        cf = np.ascontiguousarray(coef, np.float64)
        if cf.shape != (self.n_types,):
            raise ValueError(f"need {self.n_types} coefficients")
        sums = np.empty(int(n_bins), np.float64)
        modes = np.empty(int(n_bins), np.int64)
        _check(lib().scmf_structure_factor(self._h, cf.ctypes.data, float(dq), int(bool(by_q2)),
                                           int(n_bins), sums.ctypes.data, modes.ctypes.data))
        return sums, modes

    # -- files
    def write_density(self, step, directory=""):
        _check(lib().scmf_write_density(self._h, str(directory).encode(), int(step)))

    def write_density_async(self, step, directory=""):
        """Queue density_<step>.dat from a device snapshot of the current
        state; the file is written by a background thread."""
        _check(lib().scmf_write_density_async(self._h, str(directory).encode(), int(step)))

    def write_outputs_async(self, step, directory="", vtk_path=None):
        """write_density_async, plus the VTK file of that density file at
        vtk_path when given."""
        _check(lib().scmf_write_outputs_async(self._h, str(directory).encode(), int(step),
                                              None if vtk_path is None else str(vtk_path).encode()))

    def output_wait(self):
        _check(lib().scmf_output_wait(self._h))

    def checkpoint(self, path, step):
        _check(lib().scmf_checkpoint_write(self._h, str(path).encode(), int(step)))

    def restore(self, path) -> int:
        step = C.c_int64()
        _check(lib().scmf_checkpoint_read(self._h, str(path).encode(), C.byref(step)))
        return step.value

    def write_coord(self, path):
        _check(lib().scmf_write_coord(self._h, str(path).encode()))


class Batch:
    """Many systems advanced together by captured one-launch-per-phase steps."""

    def __init__(self, members):
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m._h.value for m in self.members])
        self._h = C.c_void_p()
        _check(lib().scmf_batch_create(C.byref(self._h), arr, len(self.members)))

    def step(self, n=1):
        _check(lib().scmf_batch_step(self._h, int(n)))

    def sync(self):
        _check(lib().scmf_batch_sync(self._h))

    def set_lanes(self, lanes_per_chain):
        _check(lib().scmf_batch_set_lanes(self._h, int(lanes_per_chain)))

    def update(self, members, stops=None):
        """New members, each stopping at its sweep count in `stops` (None or -1: never)."""
        self.members = list(members)
        arr = (C.c_void_p * max(1, len(self.members)))(*[m._h.value for m in self.members])
        st = None
        if stops is not None:
            st = np.ascontiguousarray([-1 if v is None else int(v) for v in stops], np.int64)
        _check(lib().scmf_batch_update(self._h, arr, len(self.members),
                                       st.ctypes.data if st is not None else None))

    def stats(self) -> BatchStats:
        st = BatchStats()
        _check(lib().scmf_batch_stats(self._h, C.byref(st)))
        return st

    def close(self):
        if self._h:
            lib().scmf_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _scheck(rc):
    if rc != SCMF_OK:
        raise ScmfError(rc, lib().scmf_server_last_error().decode())


class JobView(System):
    """A finished job's resident state, borrowed from the server (not owned)."""

    def __init__(self, handle, spec):
        self._h = C.c_void_p(handle)
        self.spec = spec
        nb, nc, ncell, nt = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        _check(lib().scmf_sizes(self._h, C.byref(nb), C.byref(nc), C.byref(ncell), C.byref(nt)))
        self.n_beads, self.n_chains, self.n_cells, self.n_types = (nb.value, nc.value,
                                                                   ncell.value, nt.value)

    def close(self):
        self._h = C.c_void_p()


class Server:
    """Continuous-batching job server (see include/scmf.h)."""

    def __init__(self, hbm_budget=0, max_jobs=0, steps_per_tick=0, job_timeout_s=0.0):
        o = ServerOpts(int(hbm_budget), int(max_jobs), int(steps_per_tick), float(job_timeout_s))
        self._h = C.c_void_p()
        _scheck(lib().scmf_server_create(C.byref(self._h), C.byref(o)))
        self.specs = {}

    def submit(self, spec: SystemSpec, opts: Opts, sweeps: int) -> int:
        d, keep = spec.to_desc()
        jid = C.c_uint64()
        _scheck(lib().scmf_server_submit(self._h, C.byref(d), C.byref(opts), int(sweeps),
                                         C.byref(jid)))
        self.specs[jid.value] = spec
        return jid.value

    def tick(self):
        q, r = C.c_int32(), C.c_int32()
        _scheck(lib().scmf_server_tick(self._h, C.byref(q), C.byref(r)))
        return q.value, r.value

    def info(self, jid) -> JobInfo:
        i = JobInfo()
        _scheck(lib().scmf_server_info(self._h, int(jid), C.byref(i)))
        return i

    def job(self, jid) -> JobView:
        h = C.c_void_p()
        _scheck(lib().scmf_server_job_system(self._h, int(jid), C.byref(h)))
        return JobView(h.value, self.specs.get(jid))

    def release(self, jid):
        _scheck(lib().scmf_server_release(self._h, int(jid)))

    def cancel(self, jid):
        _scheck(lib().scmf_server_cancel(self._h, int(jid)))

    def set_timeout(self, jid, seconds):
        _scheck(lib().scmf_server_set_timeout(self._h, int(jid), float(seconds)))

    def stats(self) -> ServerStats:
        st = ServerStats()
        _scheck(lib().scmf_server_stats(self._h, C.byref(st)))
        return st

    def clock(self) -> float:
        return lib().scmf_server_clock(self._h)

    def close(self):
        if self._h:
            lib().scmf_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
