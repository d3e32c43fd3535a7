// scmf_main.cpp — the host driver: main.cpp + test_program (test.cpp:9-95)
// on top of the C ABI.
//
// Flags keep the reference's names and defaults (init.cpp:8-45, parsed the
// way the CUDA-samples helper does: "-Name=value" or "--Name=value"):
//   -Check_T=1000  -Start_time=0  -Steps_N=100  -read_file=0  -GPU_N=1
// and add:
//   -seed=S            Philox key (default: time(NULL), like init_config.cu:864)
//   -move=smart|trial  -trial_delta=D  -omega_every_step=0|1  -density_bits=16|32
//   -config=PATH -polymer=PATH -outdir=DIR
//   -checkpoint=PATH   exact checkpoint after the last step
//   -restart=PATH      continue from such a checkpoint (instead of the
//                      initial gridding; the run starts after its step)
//   -vtk=1             also write density_<step>.vtk (vtk.c's conversion of
//                      that density file) beside every dump
// density_<step>.dat files are written by a background thread from device
// snapshots while the steps go on; the driver waits for them after the loop.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <string>
#include <vector>

#include "../../include/scmf.h"

namespace {

const char* flag_value(int argc, char** argv, const char* name) {
  const size_t n = strlen(name);
  for (int i = 1; i < argc; ++i) {
    const char* a = argv[i];
    while (*a == '-') ++a;
    if (strncmp(a, name, n) == 0 && a[n] == '=') return a + n + 1;
  }
  return nullptr;
}

long long flag_int(int argc, char** argv, const char* name, long long dflt) {
  const char* v = flag_value(argc, argv, name);
  if (!v) return dflt;
  const long long x = atoll(v);
  return x == 0 ? dflt : x;  // the helper returns 0 for absent flags; 0 means "default"
}

int die(const char* what, int rc) {
  fprintf(stderr, "scmf_run: %s failed (%d): %s\n", what, rc, scmf_last_error());
  return 1;
}

// coord.dat reader for -read_file=1 (init_config.cu:951-977).
bool read_coord(const char* path, std::vector<float>& xyz, int64_t n_beads) {
  FILE* f = fopen(path, "r");
  if (!f) return false;
  xyz.resize(3 * (size_t)n_beads);
  for (int64_t i = 0; i < n_beads; ++i) {
    double a, b, c;
    if (fscanf(f, "%lg %lg %lg", &a, &b, &c) != 3) {
      fclose(f);
      return false;
    }
    xyz[3 * i] = (float)a;
    xyz[3 * i + 1] = (float)b;
    xyz[3 * i + 2] = (float)c;
  }
  fclose(f);
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  const int gpu_n = (int)flag_int(argc, argv, "GPU_N", 1);
  if (gpu_n != 1) {
    fprintf(stderr, "scmf_run: this engine runs one system on one GPU (GPU_N=%d requested)\n", gpu_n);
    return 2;
  }
  const int tcheck = (int)flag_int(argc, argv, "Check_T", 1000);
  int start_time = (int)flag_int(argc, argv, "Start_time", 0);
  const int n_steps = (int)flag_int(argc, argv, "Steps_N", 100);
  const int read_file = (int)flag_int(argc, argv, "read_file", 0);
  const char* cfg = flag_value(argc, argv, "config");
  const char* poly = flag_value(argc, argv, "polymer");
  const char* outdir = flag_value(argc, argv, "outdir");
  const char* mv = flag_value(argc, argv, "move");
  const char* td = flag_value(argc, argv, "trial_delta");
  const char* sd = flag_value(argc, argv, "seed");
  const char* restart = flag_value(argc, argv, "restart");        // exact checkpoint to continue
  const char* checkpoint = flag_value(argc, argv, "checkpoint");  // written after the last step
  const bool vtk = flag_int(argc, argv, "vtk", 0) != 0;            // density_<step>.vtk beside each dump

  scmf_opts o{};
  o.seed = sd ? strtoull(sd, nullptr, 10) : (uint64_t)time(nullptr);
  o.system_id = 0;
  o.move_type = (mv && strcmp(mv, "trial") == 0) ? SCMF_MOVE_TRIAL : SCMF_MOVE_SMART;
  o.trial_delta = td ? (float)atof(td) : 0.1f;
  o.omega_every_step = (int)flag_int(argc, argv, "omega_every_step", 0);
  o.density_bits = (int)flag_int(argc, argv, "density_bits", 32);
  o.init_coords = 1;
  printf("seed = %llu\n", (unsigned long long)o.seed);

  scmf_sys* sys = nullptr;
  int rc = scmf_create_from_files(&sys, cfg ? cfg : "configuration.dat", poly ? poly : "polymer.dat", &o);
  if (rc) return die("init_all_config", rc);
  int64_t n_beads = 0;
  scmf_sizes(sys, &n_beads, nullptr, nullptr, nullptr);
  if (read_file == 1) {
    std::vector<float> xyz;
    if (read_coord("coord.dat", xyz, n_beads)) {
      if ((rc = scmf_copy_in_pos(sys, xyz.data(), (size_t)n_beads))) return die("read coord.dat", rc);
    } else {
      printf("coord.dat did not exist or was short; keeping the generated configuration.\n");
    }
  }
  if (restart) {
    // positions, counts, omega and the stream position of the checkpoint;
    // the run continues after its step
    int64_t step0 = 0;
    if ((rc = scmf_checkpoint_read(sys, restart, &step0))) return die("restart", rc);
    start_time = (int)step0;
    printf("restarted from %s at step %d\n", restart, start_time);
  } else {
    // test.cpp:21-23
    if ((rc = scmf_grid_density(sys))) return die("density_reset_mpi", rc);
    if ((rc = scmf_update_omega(sys))) return die("update_omega_fields_scmf0", rc);
  }
  // test.cpp:37-73
  for (int step = start_time + 1; step <= n_steps + start_time; step++) {
    if ((rc = scmf_step(sys, step))) return die("step", rc);
    if (step % tcheck == 0) {
      scmf_stats_t st;
      if ((rc = scmf_stats(sys, &st))) return die("stats", rc);
      printf("step %d end to end dis %g\n", step, st.re2);
      if (st.overflow) fprintf(stderr, "scmf_run: warning: uint16 density count overflow\n");
      std::string vtk_path;
      if (vtk) {
        vtk_path = std::string(outdir ? outdir : ".") + "/density_" + std::to_string(step) + ".vtk";
      }
      if ((rc = scmf_write_outputs_async(sys, outdir ? outdir : "", step,
                                         vtk ? vtk_path.c_str() : nullptr)))
        return die("io_out_density", rc);
    }
  }
  if ((rc = scmf_output_wait(sys))) return die("io_out_density", rc);
  std::string coord = outdir ? std::string(outdir) + "/coord.dat" : std::string("coord.dat");
  if ((rc = scmf_write_coord(sys, coord.c_str()))) return die("io_out_coord", rc);
  if (checkpoint && (rc = scmf_checkpoint_write(sys, checkpoint, n_steps + start_time)))
    return die("checkpoint", rc);
  // test.cpp:82-95: one timed step.
  scmf_sync(sys);
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  (void)hipEventRecord(a, 0);
  if ((rc = scmf_mc_sweep(sys, 1)) || (rc = scmf_grid_density(sys)) || (rc = scmf_update_omega(sys)))
    return die("timed step", rc);
  scmf_sync(sys);
  (void)hipEventRecord(b, 0);
  (void)hipEventSynchronize(b);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  printf("time =%g\n", ms);
  scmf_stats_t st;
  if (scmf_stats(sys, &st) == 0 && st.moves)
    printf("acceptance %g over %llu moves\n", (double)st.accepts / (double)st.moves,
           (unsigned long long)st.moves);
  scmf_destroy(sys);
  return 0;
}
