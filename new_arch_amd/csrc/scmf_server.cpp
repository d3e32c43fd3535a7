// scmf_server.cpp — a job server above the C ABI: simulation jobs arrive at
// any time, are admitted against a device-memory budget as slots and memory
// free up, run together (continuous batching through scmf_batch) with their
// state resident on the device, and stay resident after they finish until
// their results are fetched and released.
//
// A tick admits what fits, (re)builds the batch when its membership changed,
// and advances every running job by the same number of sweeps, cut short so
// that no job passes its own sweep count: each job stops exactly where the
// same number of scmf_step calls on it alone would leave it.
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/scmf.h"

namespace {

// Owning copy of a job description (the caller's arrays may go away).
struct OwnedDesc {
  std::vector<double> xn;
  std::vector<int32_t> len;
  std::vector<int64_t> count;
  std::vector<int32_t> mono;
  std::vector<uint8_t> conn;
  scmf_desc d{};

  void take(const scmf_desc* src) {
    d = *src;
    const int T = src->n_mono_types, P = src->n_poly_types;
    xn.assign(src->xn, src->xn + (size_t)T * T);
    len.assign(src->poly_length, src->poly_length + P);
    count.assign(src->n_polymers, src->n_polymers + P);
    size_t ms = 0, cs = 0;
    for (int32_t n : len) {
      ms += (size_t)n;
      cs += (size_t)n * n;
    }
    mono.assign(src->monotype, src->monotype + ms);
    conn.assign(src->connection, src->connection + cs);
    d.xn = xn.data();
    d.poly_length = len.data();
    d.n_polymers = count.data();
    d.monotype = mono.data();
    d.connection = conn.data();
  }
};

struct Job {
  uint64_t id = 0;
  int32_t state = 0;
  OwnedDesc desc;
  scmf_opts opts{};
  int64_t target = 0, done = 0;
  uint64_t bytes = 0;
  double t_submit = 0, t_admit = 0, t_done = 0;
  scmf_sys* sys = nullptr;
  std::string error;
};

thread_local std::string g_server_err;

int sfail(int code, const std::string& m) {
  g_server_err = m;
  return code;
}

}  // namespace

struct scmf_server {
  scmf_server_opts o{};
  std::chrono::steady_clock::time_point t0;
  std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
  std::deque<uint64_t> queue;
  std::vector<uint64_t> running;
  scmf_batch* batch = nullptr;
  bool dirty = false;
  uint64_t next_id = 1;
  uint64_t used = 0;
  uint64_t admitted = 0, finished = 0;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

// Device bytes a job will hold: the engine's own estimate plus the double
// buffer and window scratch of the spatial ordering.
uint64_t job_bytes(const scmf_desc* d, const scmf_opts* o, uint64_t* est) {
  scmf_summary s;
  if (scmf_inspect_desc(d, &s) != SCMF_OK) return 0;
  const uint64_t base = o->density_bits == 16 ? s.device_bytes_16 : s.device_bytes_32;
  *est = base + 12ull * (uint64_t)s.n_beads + 16ull * (uint64_t)s.n_chains +
         4ull * (uint64_t)s.n_cells * (uint64_t)s.n_mono_types * 8ull + (1ull << 20);
  return *est;
}

int rebuild(scmf_server* sv) {
  if (sv->batch) {
    scmf_batch_destroy(sv->batch);
    sv->batch = nullptr;
  }
  sv->dirty = false;
  if (sv->running.empty()) return SCMF_OK;
  std::vector<scmf_sys*> m;
  for (uint64_t id : sv->running) m.push_back(sv->jobs[id]->sys);
  int rc = scmf_batch_create(&sv->batch, m.data(), (int32_t)m.size());
  if (rc) return sfail(rc, std::string("batch: ") + scmf_last_error());
  return SCMF_OK;
}

}  // namespace

extern "C" {

int scmf_server_create(scmf_server** out, const scmf_server_opts* opts) {
  if (!out) return sfail(SCMF_E_ARG, "out is null");
  scmf_server* sv = new scmf_server();
  if (opts) sv->o = *opts;
  if (sv->o.max_jobs <= 0) sv->o.max_jobs = 4096;
  if (sv->o.steps_per_tick <= 0) sv->o.steps_per_tick = 50;
  if (sv->o.hbm_budget == 0) sv->o.hbm_budget = 200ull << 30;
  sv->t0 = std::chrono::steady_clock::now();
  *out = sv;
  return SCMF_OK;
}

void scmf_server_destroy(scmf_server* sv) {
  if (!sv) return;
  if (sv->batch) scmf_batch_destroy(sv->batch);
  for (auto& kv : sv->jobs)
    if (kv.second->sys) scmf_destroy(kv.second->sys);
  delete sv;
}

double scmf_server_clock(scmf_server* sv) { return sv ? sv->now() : 0.0; }

int scmf_server_submit(scmf_server* sv, const scmf_desc* desc, const scmf_opts* opts,
                       int32_t n_sweeps, uint64_t* job_id) {
  if (!sv || !desc || !opts || !job_id) return sfail(SCMF_E_ARG, "null argument");
  if (n_sweeps <= 0) return sfail(SCMF_E_ARG, "a job needs a positive sweep count");
  uint64_t est = 0;
  if (!job_bytes(desc, opts, &est)) return sfail(SCMF_E_ARG, scmf_last_error());
  if (est > sv->o.hbm_budget)
    return sfail(SCMF_E_BUDGET, "the job alone needs more device memory than the budget");
  std::unique_ptr<Job> j(new Job());
  j->id = sv->next_id++;
  j->desc.take(desc);
  j->opts = *opts;
  j->opts.init_coords = 1;
  j->target = n_sweeps;
  j->bytes = est;
  j->t_submit = sv->now();
  j->state = SCMF_JOB_QUEUED;
  *job_id = j->id;
  sv->queue.push_back(j->id);
  sv->jobs[j->id] = std::move(j);
  return SCMF_OK;
}

int scmf_server_tick(scmf_server* sv, int32_t* n_queued, int32_t* n_running) {
  if (!sv) return sfail(SCMF_E_ARG, "null server");
  // admit in arrival order while memory and slots allow
  while (!sv->queue.empty() && (int32_t)sv->running.size() < sv->o.max_jobs) {
    Job& j = *sv->jobs[sv->queue.front()];
    if (sv->used + j.bytes > sv->o.hbm_budget) break;
    sv->queue.pop_front();
    int rc = scmf_create(&j.sys, &j.desc.d, &j.opts);
    if (!rc) rc = scmf_grid_density(j.sys);   // test.cpp:21-23
    if (!rc) rc = scmf_update_omega(j.sys);
    if (rc) {
      j.state = SCMF_JOB_FAILED;
      j.error = scmf_last_error();
      if (j.sys) scmf_destroy(j.sys);
      j.sys = nullptr;
      continue;
    }
    j.state = SCMF_JOB_RUNNING;
    j.t_admit = sv->now();
    sv->used += j.bytes;
    sv->running.push_back(j.id);
    sv->admitted += 1;
    sv->dirty = true;
  }
  if (sv->dirty) {
    int rc = rebuild(sv);
    if (rc) return rc;
  }
  if (!sv->running.empty()) {
    int64_t k = sv->o.steps_per_tick;
    for (uint64_t id : sv->running) k = std::min(k, sv->jobs[id]->target - sv->jobs[id]->done);
    int rc = scmf_batch_step(sv->batch, (int32_t)k);
    if (!rc) rc = scmf_batch_sync(sv->batch);
    if (rc) return sfail(rc, std::string("batch step: ") + scmf_last_error());
    const double t = sv->now();
    std::vector<uint64_t> still;
    for (uint64_t id : sv->running) {
      Job& j = *sv->jobs[id];
      j.done += k;
      if (j.done >= j.target) {
        j.state = SCMF_JOB_DONE;
        j.t_done = t;
        sv->finished += 1;
        sv->dirty = true;
      } else {
        still.push_back(id);
      }
    }
    sv->running.swap(still);
  }
  if (n_queued) *n_queued = (int32_t)sv->queue.size();
  if (n_running) *n_running = (int32_t)sv->running.size();
  return SCMF_OK;
}

int scmf_server_info(scmf_server* sv, uint64_t job_id, scmf_job_info* out) {
  if (!sv || !out) return sfail(SCMF_E_ARG, "null argument");
  auto it = sv->jobs.find(job_id);
  if (it == sv->jobs.end()) return sfail(SCMF_E_ARG, "unknown job");
  const Job& j = *it->second;
  memset(out, 0, sizeof *out);
  out->state = j.state;
  out->sweeps_done = j.done;
  out->sweeps_target = j.target;
  out->t_submit = j.t_submit;
  out->t_admit = j.t_admit;
  out->t_done = j.t_done;
  out->device_bytes = j.bytes;
  if (j.sys) {
    int64_t nb = 0;
    scmf_sizes(j.sys, &nb, nullptr, nullptr, nullptr);
    out->beads = nb;
  }
  return SCMF_OK;
}

int scmf_server_job_system(scmf_server* sv, uint64_t job_id, scmf_sys** out) {
  if (!sv || !out) return sfail(SCMF_E_ARG, "null argument");
  auto it = sv->jobs.find(job_id);
  if (it == sv->jobs.end() || !it->second->sys) return sfail(SCMF_E_ARG, "job has no resident state");
  if (it->second->state != SCMF_JOB_DONE) return sfail(SCMF_E_ARG, "job is not finished");
  *out = it->second->sys;
  return SCMF_OK;
}

int scmf_server_release(scmf_server* sv, uint64_t job_id) {
  if (!sv) return sfail(SCMF_E_ARG, "null server");
  auto it = sv->jobs.find(job_id);
  if (it == sv->jobs.end()) return sfail(SCMF_E_ARG, "unknown job");
  Job& j = *it->second;
  if (j.state == SCMF_JOB_RUNNING) return sfail(SCMF_E_ARG, "job is still running");
  if (j.sys) {
    scmf_destroy(j.sys);
    j.sys = nullptr;
    sv->used -= j.bytes;
  }
  if (j.state == SCMF_JOB_QUEUED) {
    sv->queue.erase(std::remove(sv->queue.begin(), sv->queue.end(), job_id), sv->queue.end());
  }
  j.state = SCMF_JOB_RELEASED;
  return SCMF_OK;
}

const char* scmf_server_last_error(void) { return g_server_err.c_str(); }

}  // extern "C"
