/* scmf.h — C ABI of the MI355X single-chain-in-mean-field Monte Carlo engine.
 *
 * The drop-in boundary for the reference's host "operator API" (a set of
 * extern C++ functions taking Phase, GPU_info and MPI_info pointers).  Every entry point
 * below names the reference function it replaces.  Conventions:
 *   - plain C types only; the opaque handle owns all device memory;
 *   - every call returns SCMF_OK (0) or a negative SCMF_E* code and never
 *     exits the process; scmf_last_error() gives the message of the last
 *     failure on the calling thread;
 *   - work is queued on the handle's own HIP stream; calls that hand data to
 *     the host (scmf_copy_out, scmf_stats, the writers) synchronise it;
 *   - one handle is used by one host thread at a time.
 * Positions cross the boundary as the reference stores them: float
 * [bead][xyz], chains contiguous, polymer types in polymer.dat order
 * (struct.h:203, init_config.cu:610-629).
 */
#ifndef SCMF_H
#define SCMF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCMF_OK 0
#define SCMF_E_ARG -1      /* invalid argument or inconsistent description */
#define SCMF_E_IO -2       /* file missing or malformed */
#define SCMF_E_HIP -3      /* HIP runtime failure */
#define SCMF_E_NOMEM -4    /* device allocation failed */
#define SCMF_E_OVERFLOW -5 /* a uint16 density count exceeded 65535 */
#define SCMF_E_NODEV -6    /* no usable GPU */
#define SCMF_E_BUDGET -7   /* admission refused: HBM budget exhausted */

/* Move kinds.  SMART is the reference's force-biased move (struct.h:45
 * move_type_arg_SMART, device_function.cu:637-717).  TRIAL is the uniform
 * local displacement the reference only names as an enum (struct.h:45). */
#define SCMF_MOVE_TRIAL 0
#define SCMF_MOVE_SMART 1

/* What scmf_copy_out / scmf_copy_in move. */
#define SCMF_POS 0      /* float [beads][3], reference order */
#define SCMF_DENSITY 1  /* uint32 [types][cells] (widened from uint16 if needed) */
#define SCMF_OMEGA 2    /* float  [types][cells] */

/* System description: the content of configuration.dat + polymer.dat
 * (init_config.cu:210-451, 487-514). */
typedef struct scmf_desc {
  int32_t n_mono_types;      /* T: "Monomer type number" */
  const double* xn;          /* [T*T], xn[t + T*s]: diagonal kappaN, off-diagonal chiN */
  double L[3];               /* box edge lengths in Re: "D0 = ; LY = ; LZ = ;" */
  int32_t n[3];              /* grid cells per edge: "N = nx ny nz" */
  double dtNxi;              /* "dt*N/xi"; smart-MC step A = dtNxi / reference_Nbeads */
  int32_t reference_Nbeads;  /* "Polymer length unit" (Nref) */
  int32_t n_poly_types;      /* P: "Number of polymer type" */
  const int32_t* poly_length;  /* [P] beads per chain of each type */
  const int64_t* n_polymers;   /* [P] chains of each type */
  const int32_t* monotype;     /* [sum N_p] monomer type of every bead, types concatenated */
  const uint8_t* connection;   /* [sum N_p*N_p] row-major bond matrices, types concatenated:
                                  connection[off_p + i*N_p + j] != 0 <=> beads i and j bonded */
} scmf_desc;

typedef struct scmf_opts {
  uint64_t seed;            /* Philox key of every random stream of this system */
  uint32_t system_id;       /* fourth counter word: separates systems sharing a seed */
  int32_t move_type;        /* SCMF_MOVE_SMART (reference) or SCMF_MOVE_TRIAL */
  float trial_delta;        /* TRIAL: each component uniform in [-delta, delta) Re */
  int32_t omega_every_step; /* 0: reference cadence, omega skipped when step%10==0
                               (test.cpp:42-43); 1: omega after every sweep */
  int32_t density_bits;     /* 16 (uint16 counts, overflow-checked) or 32 */
  int32_t init_coords;      /* 1: build the initial Gaussian walks on the device
                               (initialize_coord, device_function.cu:140-207);
                               0: the caller supplies positions (scmf_copy_in_pos) */
  int32_t lanes_per_chain;  /* sweep kernel shape: 1, 2, 4, 8 or 16 lanes share a chain
                               staged in LDS; -1 = one lane per chain in global
                               memory; 0 = pick automatically.  At 8, systems of
                               65,536 chains or more sweep their runs of chains
                               longer than 64 beads at 16 when that keeps more
                               waves resident.  Results do not depend on it. */
  int32_t sort_every;       /* reorder chains spatially every k sweeps (enables the
                               tiled gridding); 0 = automatic (every 100 sweeps for
                               systems of 2^23 beads or more, else never), -1 =
                               never.  Results do not depend on it. */
} scmf_opts;

typedef struct scmf_stats_t {
  uint64_t moves;        /* attempted moves since creation / last reset */
  uint64_t accepts;      /* accepted moves */
  uint64_t sweeps;       /* sweeps done (also the sweep word of the next sweep's stream) */
  double re2;            /* mean squared end-to-end distance, first to last bead (Re^2) */
  int32_t overflow;      /* 1 if a uint16 count saturated since the last reset */
  uint64_t grid_fallback; /* beads the tiled gridding counted with global atomics */
  uint64_t sorts;         /* spatial reorderings done */
  uint64_t cell_grids;    /* griddings that read the bead cells stored by the sweep
                             instead of binning positions (tiled systems) */
} scmf_stats_t;

typedef struct scmf_timing_t {
  /* Mean device time per launch over the launches since the last reset,
   * measured with HIP events on the handle's stream. */
  double sweep_ms, grid_ms, omega_ms;
  uint64_t sweep_n, grid_n, omega_n;
} scmf_timing_t;

typedef struct scmf_summary {
  int32_t n_mono_types, n_poly_types, reference_Nbeads;
  int32_t n[3];
  double L[3];
  double dtNxi;
  int64_t n_chains, n_beads, n_cells;
  double xn[64];             /* first min(T*T, 64) entries of xn[t + T*s] */
  uint64_t device_bytes_16;  /* device memory a system needs with uint16 counts
                                (ordering buffers, the tiled gridding's window images
                                among them, included from 2^23 beads, where ordering
                                is automatic; not the optional analysis and output
                                buffers) */
  uint64_t device_bytes_32;  /* ... with uint32 counts */
} scmf_summary;

typedef struct scmf_sys scmf_sys;

/* init_all_config (init_config.cu:992-1011): validate, derive constants
 * (init_config.cu:480-484, 662-672, 784-809), allocate device state and,
 * when opts->init_coords, build the initial configuration. */
int scmf_create(scmf_sys** out, const scmf_desc* desc, const scmf_opts* opts);

/* init_all_config reading configuration.dat and polymer.dat from the given
 * paths (Read_polymer_config init_config.cu:210-451, initialize_values
 * init_config.cu:487-514). */
int scmf_create_from_files(scmf_sys** out, const char* configuration_dat, const char* polymer_dat,
                           const scmf_opts* opts);

void scmf_destroy(scmf_sys* sys);

/* smc_move (smc.cu:3-72 -> mc_polymer_move_arr, device_function.cu:718-799):
 * n_sweeps sweeps, each giving every chain N sequential moves against the
 * current omega field. */
int scmf_mc_sweep(scmf_sys* sys, int32_t n_sweeps);

/* density_reset_mpi / reset_density_fields (mesh_mpi.cpp:4-27, mesh.cu:161-190
 * -> init_array + coord_to_density, device_function.cu:471-508): zero and
 * re-grid the NGP density counts.  The multi-GPU reduction collapses to
 * nothing on one device. */
int scmf_grid_density(scmf_sys* sys);

/* update_omega_fields_scmf0 (mesh.cu:217-248 -> omega_field_update,
 * device_function.cu:373-412). */
int scmf_update_omega(scmf_sys* sys);

/* One iteration of the step loop body of test.cpp:37-43: sweep, grid, then
 * omega unless the reference cadence skips it for this step number. */
int scmf_step(scmf_sys* sys, int32_t step);

/* Replace the host's direct reads of managed memory (io.cpp:36,44,107,
 * test.cpp:51-58): copy a device array to host memory. */
int scmf_copy_out(scmf_sys* sys, int32_t what, void* host, size_t bytes);

/* No counterpart in the reference.  scmf_copy_out for callers that cannot
   allocate device memory (code under HIP graph capture, serving loops, hosts
   that forbid allocation): the same copies, staged through a fixed 16 MiB
   device array instead of buffers made on demand (scmf_copy_out allocates a
   position staging buffer on first use and a widening buffer on every
   uint16 copy).  Positions of at most 1,398,101 beads and uint16 counts of
   at most 4,194,304 types x cells; larger systems are refused with
   SCMF_E_ARG (use scmf_copy_out).  uint32 counts and omega are copied
   directly.  Calls in one process take a lock, since every call on a GPU
   shares that GPU's array (which keeps the last call's contents). */
int scmf_copy_out_noalloc(scmf_sys* sys, int32_t what, void* host, size_t bytes);

/* coord.dat restart (init_config.cu:930-987) and parity injection. */
int scmf_copy_in_pos(scmf_sys* sys, const float* xyz, size_t n_beads);
int scmf_copy_in_omega(scmf_sys* sys, const float* omega, size_t n_values);
/* Density counts [types][cells] (uint32; must fit the system's count width). */
int scmf_copy_in_density(scmf_sys* sys, const uint32_t* counts, size_t n_values);

/* Reorder the chains spatially now (storage order only: results, copies and
 * files are unaffected).  No counterpart in the reference. */
int scmf_sort_chains(scmf_sys* sys);

/* Re-key the random streams: seed, and the sweep counter of the next sweep. */
int scmf_set_rng(scmf_sys* sys, uint64_t seed, uint64_t next_sweep);

int scmf_stats(scmf_sys* sys, scmf_stats_t* out);
int scmf_reset_stats(scmf_sys* sys);
/* Per-kernel device timing with HIP events (off by default; the reference
 * times one whole step with cudaEvents, test.cpp:82-95). */
int scmf_enable_timing(scmf_sys* sys, int32_t on);
int scmf_timing(scmf_sys* sys, scmf_timing_t* out);
int scmf_reset_timing(scmf_sys* sys);
int scmf_sync(scmf_sys* sys);

/* Autotuner for the sweep kernel shape (lanes per chain, LDS tile): times
 * every variant that fits for a few sweeps on the system's own state, keeps
 * the fastest, and restores positions, counters and the stream position, so
 * tuning leaves the trajectory unchanged.  No counterpart in the reference,
 * whose launch shape is the largest divisor of the chain count up to 1024
 * (init_config.cu:8-64). */
int scmf_autotune(scmf_sys* sys, int32_t sweeps_per_variant, int32_t* lanes_per_chain,
                  float* ms_per_sweep);

/* Sizes of the system (for host buffers). */
int scmf_sizes(scmf_sys* sys, int64_t* n_beads, int64_t* n_chains, int64_t* n_cells,
               int32_t* n_types);

/* io_out_density (io.cpp:4-56): density_<step>.dat in directory dir ("" = cwd). */
int scmf_write_density(scmf_sys* sys, const char* dir, int32_t step);
/* io_out_density without the stall: returns once a device snapshot of the
   counts and omega is queued on the system's stream (two device-to-device
   copies); a host thread copies it out and writes the same file as
   scmf_write_density while the caller keeps stepping.  At most two snapshots
   are in flight: a third call waits until the oldest file is written.  A
   write error is reported by the next call or by scmf_output_wait.  No
   counterpart in the reference, whose dumps are synchronous. */
int scmf_write_density_async(scmf_sys* sys, const char* dir, int32_t step);
/* The same, and when vtk_path is not null also the VTK file that
   scmf_density_to_vtk(<that density file>, vtk_path) would make of it
   (vtk.c:42-76), written by the background thread from the same snapshot. */
int scmf_write_outputs_async(scmf_sys* sys, const char* dir, int32_t step, const char* vtk_path);
/* Wait until every queued file is written; returns the first error since the
   last report.  scmf_destroy also finishes the queue. */
int scmf_output_wait(scmf_sys* sys);
/* io_out_coord (io.cpp:95-118): "%g %g %g" per bead. */
int scmf_write_coord(scmf_sys* sys, const char* path);

/* Exact checkpoint, the lossless counterpart of coord.dat restarts
   (io.cpp:95-118 with init_config.cu:930-987, which lose digits to %g):
   positions, counts, omega, seed, system id and the stream sweep counter, so
   a run restored from it continues bit for bit.  `step` is stored for the
   caller (the omega cadence of scmf_step depends on it). */
int scmf_checkpoint_write(scmf_sys* sys, const char* path, int64_t step);
/* Restore into a system created from the same description and counter
   width; returns the stored step. */
int scmf_checkpoint_read(scmf_sys* sys, const char* path, int64_t* step);
/* Static structure factor of the current counts, an analysis the reference
   names only as a link dependency (-lcufft, Makefile:265):
   w(c) = sum_t coef[t] * n_t(c), F(q) = sum_c w(c) exp(-i q.r_c) over the
   grid's wave vectors q = 2*pi*(kx/Lx, ky/Ly, kz/Lz), S(q) = |F(q)|^2 / n_beads.
   Each mode goes to shell b = nearest integer to |q|/dq (by_q2 = 0) or to
   q^2/dq^2 (by_q2 = 1; with dq = 2*pi/L on a cubic box the shells are the
   integer |k|^2); shells b >= n_bins are dropped.  On return s_sum[b] is the
   sum of S over the modes of shell b (q = 0 is in shell 0) and n_modes[b]
   their number (every mode of the full spectrum counts once).  Sums are fp64
   atomics: reproducible to rounding, not bit for bit.  Synchronous. */
#define SCMF_SQ_MAX_BINS (1 << 20)
int scmf_structure_factor(scmf_sys* sys, const double* coef, double dq, int32_t by_q2,
                          int32_t n_bins, double* s_sum, int64_t* n_modes);
/* Host-only helpers (no device needed). */
/* Parse configuration.dat + polymer.dat and report sizes and memory needs. */
int scmf_inspect_files(const char* configuration_dat, const char* polymer_dat, scmf_summary* out);
/* Sizes and memory needs of a description. */
int scmf_inspect_desc(const scmf_desc* desc, scmf_summary* out);
/* io_out_density's format from host arrays: counts uint32 [T][cells], omega float [T][cells]. */
int scmf_write_density_host(const char* path, const int32_t n[3], const double L[3], int32_t T,
                            const uint32_t* counts, const float* omega, double field_scaling);
/* vtk.c:4-81: density_*.dat -> STRUCTURED_POINTS VTK of phi_0 - phi_1. */
int scmf_density_to_vtk(const char* density_path, const char* vtk_path);
/* The VTK file scmf_density_to_vtk(density_name, vtk_path) makes of the file
   scmf_write_density_host writes from the same counts and omega, made from
   them directly (no density file is read). */
int scmf_write_vtk_host(const char* vtk_path, const char* density_name, const int32_t n[3],
                        const double L[3], int32_t T, const uint32_t* counts, const float* omega,
                        double field_scaling);

/* ---- batches: many independent systems advanced together -------------------
 * No counterpart in the reference (one system per process).  Each step runs
 * one launch per phase across all members (sweep, clear, gridding, omega with
 * every member's own cadence) from a captured HIP graph.  Every member ends
 * up exactly where the same number of scmf_step calls on it alone would
 * leave it.  Call scmf_batch_sync before using a member's handle again. */
typedef struct scmf_batch scmf_batch;
int scmf_batch_create(scmf_batch** out, scmf_sys* const* members, int32_t n_members);
/* Replace the members and give each a stop: from sweep count stop_sweeps[i]
   on (its stream sweep word, scmf_stats().sweeps) member i does no more work
   in the batch's steps, so members with different sweep targets share every
   step; a negative stop (or stop_sweeps null) means none.  Takes effect for
   the steps queued after it, without waiting for those queued before; members
   new to the batch must have no work of their own pending. */
int scmf_batch_update(scmf_batch* batch, scmf_sys* const* members, int32_t n_members,
                      const int64_t* stop_sweeps);
int scmf_batch_step(scmf_batch* batch, int32_t n_steps);
int scmf_batch_sync(scmf_batch* batch);
/* Force one sweep-kernel shape for every member (0 = each member's own). */
int scmf_batch_set_lanes(scmf_batch* batch, int32_t lanes_per_chain);
void scmf_batch_destroy(scmf_batch* batch);

/* The step graph is captured with room (members, workgroups, chain length);
   a membership that fits is installed by uploading the graph's slot table in
   stream order, with no capture.  No counterpart in the reference. */
typedef struct scmf_batch_stats_t {
  uint64_t steps;            /* steps enqueued */
  uint64_t captures;         /* step graphs captured */
  uint64_t table_uploads;    /* memberships installed into a captured graph */
  int32_t members, member_capacity;
  uint64_t blocks_launched;  /* workgroups per step, all phases */
  uint64_t blocks_used;      /* of those, the current members' */
} scmf_batch_stats_t;
int scmf_batch_stats(scmf_batch* batch, scmf_batch_stats_t* out);

/* ---- serving: jobs submitted at any time, run by continuous batching -------
 * No counterpart in the reference.  A job is a system (description, options)
 * and a sweep count.  A worker thread admits queued jobs in arrival order
 * while the device-memory budget and the slot count allow, creating and
 * initialising each on its own stream while the running jobs keep stepping.
 * scmf_server_tick brings admitted jobs into one batch, enqueues
 * steps_per_tick steps of it (sweep, grid, omega per step, as scmf_step;
 * each job stops on the device at its own sweep count, so a tick is never cut
 * short for the job closest to its end) and returns without waiting for them
 * while fewer than two ticks are queued; a job is DONE when the tick holding
 * its last sweep has run, and its state stays resident until released.  A
 * finished job's result equals running it alone. */
#define SCMF_JOB_QUEUED 0
#define SCMF_JOB_RUNNING 1
#define SCMF_JOB_DONE 2
#define SCMF_JOB_RELEASED 3
#define SCMF_JOB_FAILED 4
#define SCMF_JOB_CANCELLED 5

/* Why a job was cancelled (scmf_job_info.reason). */
#define SCMF_CANCEL_NONE 0
#define SCMF_CANCEL_CALLER 1   /* scmf_server_cancel */
#define SCMF_CANCEL_TIMEOUT 2  /* its wall-clock limit passed before it finished */

typedef struct scmf_server_opts {
  uint64_t hbm_budget;    /* device bytes the jobs may hold (0: 200 GiB) */
  int32_t max_jobs;       /* admitted jobs at most (0: 4096) */
  int32_t steps_per_tick; /* sweeps per tick (0: 50) */
  double job_timeout_s;   /* default wall-clock limit of a job from its submission
                             (0: none); a job not DONE by then is cancelled */
} scmf_server_opts;

typedef struct scmf_job_info {
  int32_t state;
  int32_t reason;                   /* SCMF_CANCEL_* of a CANCELLED job */
  int64_t sweeps_done, sweeps_target;  /* sweeps_done counts the sweeps enqueued */
  double t_submit, t_admit, t_done; /* seconds on the server clock; t_done is when
                                       the device finished the job's last step */
  uint64_t device_bytes;            /* admission estimate */
  int64_t beads;
} scmf_job_info;

typedef struct scmf_server_stats_t {
  uint64_t ticks;        /* ticks enqueued */
  uint64_t rebuilds;     /* changes of the batch's membership */
  uint64_t admitted, finished, cancelled, failed;
  double busy_s;         /* device time of the finished ticks */
  double span_s;         /* from the start of the first tick to the end of the last
                            finished one; busy_s / span_s is the GPU's busy fraction */
  uint64_t used_bytes;   /* admission estimates of the jobs holding memory */
  uint64_t captures;     /* step graphs captured; the other membership changes only
                            uploaded the graph's slot table */
} scmf_server_stats_t;

typedef struct scmf_server scmf_server;
int scmf_server_create(scmf_server** out, const scmf_server_opts* opts);
int scmf_server_submit(scmf_server* server, const scmf_desc* desc, const scmf_opts* opts,
                       int32_t n_sweeps, uint64_t* job_id);
int scmf_server_tick(scmf_server* server, int32_t* n_queued, int32_t* n_running);
int scmf_server_info(scmf_server* server, uint64_t job_id, scmf_job_info* out);
/* Handle of a finished job, for scmf_copy_out / scmf_stats / the writers. */
int scmf_server_job_system(scmf_server* server, uint64_t job_id, scmf_sys** out);
/* Free a finished, failed or cancelled job's memory (once no queued step uses it). */
int scmf_server_release(scmf_server* server, uint64_t job_id);
/* Cancel a queued or running job: a queued one is never admitted; a running one
   leaves the batch at the next tick and its memory is freed once the steps
   queued with it have run.  Its state becomes CANCELLED (reason CALLER). */
int scmf_server_cancel(scmf_server* server, uint64_t job_id);
/* Wall-clock limit of one job from its submission (0: none), replacing the
   server's default. */
int scmf_server_set_timeout(scmf_server* server, uint64_t job_id, double seconds);
int scmf_server_stats(scmf_server* server, scmf_server_stats_t* out);
double scmf_server_clock(scmf_server* server);
const char* scmf_server_last_error(void);
/* Waits for the queued steps, then frees every job. */
void scmf_server_destroy(scmf_server* server);

const char* scmf_last_error(void);
const char* scmf_version(void);

#ifdef __cplusplus
}
#endif
#endif /* SCMF_H */
