// scmf_server.cpp — a job server above the C ABI: simulation jobs arrive at
// any time, are admitted against a device-memory budget as slots and memory
// free up, run together (continuous batching through scmf_batch) with their
// state resident on the device, and stay resident after they finish until
// their results are fetched and released.
//
// Three things keep the GPU busy:
//   * admission runs on a worker thread: creating a job's system, its initial
//     configuration, gridding and omega happen on the system's own stream
//     while the running jobs keep stepping; the worker also frees the memory
//     of jobs that are released or cancelled, so no free stalls the ticks;
//   * every tick enqueues the same number of steps for the whole batch and
//     each job stops on the device at its own sweep count (scmf_batch_update's
//     stops): no tick is cut short for the job closest to its end, and a job
//     ends exactly where the same number of scmf_step calls on it alone would
//     leave it;
//   * ticks are not waited for: up to two stay queued while the host handles
//     arrivals, completions and membership changes; a job's completion time is
//     read from the event after the tick holding its last sweep.
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/scmf.h"
#include "scmf_batch_tickets.h"

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
  int32_t state = SCMF_JOB_QUEUED;
  int32_t reason = SCMF_CANCEL_NONE;
  OwnedDesc desc;
  scmf_opts opts{};
  int64_t target = 0;
  int64_t enq = 0;            // sweeps enqueued
  uint64_t bytes = 0;
  double t_submit = 0, t_admit = 0, t_done = 0;
  double timeout = 0;         // seconds from submission, 0: none
  scmf_sys* sys = nullptr;
  std::string error;
  bool cancel = false;        // asked for; a running job leaves at the next tick
  bool finishing = false;     // its last sweep is enqueued
  bool parked = false;        // out of the batch, memory freed once last_ticket has run
  uint64_t last_ticket = 0;   // the last ticket whose steps include the job
};

thread_local std::string g_server_err;

int sfail(int code, const std::string& m) {
  g_server_err = m;
  return code;
}

struct InFlight {
  uint64_t ticket;
  std::vector<uint64_t> finishing;  // jobs whose last sweep is in these steps
};

}  // namespace

struct scmf_server {
  scmf_server_opts o{};
  std::chrono::steady_clock::time_point t0;
  // shared with the worker (under mu)
  std::mutex mu;
  std::condition_variable cv;
  std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
  std::deque<uint64_t> queue;                  // submitted, not yet taken by the worker
  std::vector<uint64_t> ready;                 // created and initialised, not yet in the batch
  std::vector<std::pair<scmf_sys*, uint64_t>> doomed;  // systems to free and their bytes
  uint64_t used = 0;                           // admission estimates holding memory
  int32_t slots = 0;                           // jobs from admission until they leave the batch
  int32_t creating = 0;
  bool quit = false;
  uint64_t admitted = 0, failed = 0, cancelled = 0, finished = 0;
  std::thread worker;
  // main thread only
  scmf_batch* batch = nullptr;
  std::vector<uint64_t> running;               // the batch's members, in order
  std::vector<uint64_t> parked;                // out of the batch, waiting for their last ticket
  std::deque<InFlight> inflight;
  uint64_t completed_ticket = 0;
  bool dirty = false;
  uint64_t next_id = 1;
  uint64_t ticks = 0, rebuilds = 0;
  uint64_t captures = 0;                       // step graphs the batch captured (under mu)
  double busy_s = 0, first_start = -1, last_end = 0;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

// Device bytes a job will hold: the engine's own estimate plus the double
// buffer and window images of the spatial ordering.
uint64_t job_bytes(const scmf_desc* d, const scmf_opts* o, uint64_t* est) {
  scmf_summary s;
  if (scmf_inspect_desc(d, &s) != SCMF_OK) return 0;
  const uint64_t base = o->density_bits == 16 ? s.device_bytes_16 : s.device_bytes_32;
  *est = base + 12ull * (uint64_t)s.n_beads + 16ull * (uint64_t)s.n_chains +
         4ull * (uint64_t)s.n_cells * (uint64_t)s.n_mono_types * 8ull + (1ull << 20);
  return *est;
}

// The worker may take the job at the head of the queue (under mu).
bool admissible(scmf_server* sv) {
  if (sv->queue.empty() || sv->slots >= sv->o.max_jobs) return false;
  const Job& j = *sv->jobs[sv->queue.front()];
  return sv->used + j.bytes <= sv->o.hbm_budget;
}

void worker_loop(scmf_server* sv) {
  std::unique_lock<std::mutex> lk(sv->mu);
  for (;;) {
    sv->cv.wait(lk, [&] { return sv->quit || !sv->doomed.empty() || admissible(sv); });
    if (!sv->doomed.empty()) {
      const auto d = sv->doomed.back();
      sv->doomed.pop_back();
      lk.unlock();
      {
        std::lock_guard<std::mutex> cl(scmf_batch_tickets::capture_mutex());
        scmf_destroy(d.first);
      }
      lk.lock();
      sv->used -= d.second;
      sv->cv.notify_all();
      continue;
    }
    if (sv->quit) return;
    const uint64_t id = sv->queue.front();
    sv->queue.pop_front();
    Job& j = *sv->jobs[id];
    sv->used += j.bytes;
    sv->slots += 1;
    sv->creating += 1;
    lk.unlock();
    // test.cpp:19-23: create, initial configuration, gridding, omega
    scmf_sys* sys = nullptr;
    int rc;
    std::string err;
    {
      std::lock_guard<std::mutex> cl(scmf_batch_tickets::capture_mutex());
      rc = scmf_create(&sys, &j.desc.d, &j.opts);
      if (!rc) rc = scmf_grid_density(sys);
      if (!rc) rc = scmf_update_omega(sys);
      if (!rc) rc = scmf_sync(sys);
      if (rc) err = scmf_last_error();
      if (rc && sys) {
        scmf_destroy(sys);
        sys = nullptr;
      }
    }
    lk.lock();
    sv->creating -= 1;
    if (rc || j.cancel) {
      if (sys) sv->doomed.push_back({sys, 0});
      sv->used -= j.bytes;
      sv->slots -= 1;
      if (rc) {
        j.state = SCMF_JOB_FAILED;
        j.error = err;
        sv->failed += 1;
      } else {
        j.state = SCMF_JOB_CANCELLED;
        sv->cancelled += 1;
      }
    } else {
      j.sys = sys;
      j.t_admit = sv->now();
      sv->ready.push_back(id);
      sv->admitted += 1;
    }
    sv->cv.notify_all();
  }
}

// Free the memory of parked jobs whose last ticket has run (their systems go
// to the worker).
void release_parked(scmf_server* sv) {
  std::vector<uint64_t> keep;
  std::lock_guard<std::mutex> lk(sv->mu);
  for (uint64_t id : sv->parked) {
    Job& j = *sv->jobs[id];
    if (j.last_ticket <= sv->completed_ticket) {
      if (j.sys) sv->doomed.push_back({j.sys, j.bytes});
      j.sys = nullptr;
      j.parked = false;
    } else {
      keep.push_back(id);
    }
  }
  sv->parked.swap(keep);
  sv->cv.notify_all();
}

// Take the finished ticks off the queue front (waiting for the first if
// `wait`): mark their finishing jobs DONE at the ticks' device end time.
int reap(scmf_server* sv, bool wait) {
  while (!sv->inflight.empty()) {
    InFlight& f = sv->inflight.front();
    int done = 0;
    double t_end = 0, busy = 0;
    int rc = scmf_batch_tickets::poll(sv->batch, f.ticket, wait, &done, &t_end, &busy);
    if (rc) return sfail(rc, std::string("tick: ") + scmf_last_error());
    if (!done) break;
    wait = false;
    const double off =
        std::chrono::duration<double>(scmf_batch_tickets::origin(sv->batch) - sv->t0).count();
    const double end = off + t_end;
    if (sv->first_start < 0) sv->first_start = end - busy;
    sv->last_end = end;
    sv->busy_s += busy;
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      for (uint64_t id : f.finishing) {
        Job& j = *sv->jobs[id];
        if (j.state == SCMF_JOB_RUNNING) {
          j.state = SCMF_JOB_DONE;
          j.t_done = end;
          sv->finished += 1;
        }
      }
    }
    sv->completed_ticket = f.ticket;
    sv->inflight.pop_front();
  }
  release_parked(sv);
  return SCMF_OK;
}

// Take a job out of the batch at the next rebuild; its memory is freed once
// the steps queued with it have run.
void leave(scmf_server* sv, Job& j) {
  sv->running.erase(std::remove(sv->running.begin(), sv->running.end(), j.id), sv->running.end());
  j.parked = true;
  sv->parked.push_back(j.id);
  sv->dirty = true;
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->slots -= 1;
  sv->cv.notify_all();
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
  if (!(sv->o.job_timeout_s >= 0)) sv->o.job_timeout_s = 0;
  sv->t0 = std::chrono::steady_clock::now();
  sv->worker = std::thread(worker_loop, sv);
  *out = sv;
  return SCMF_OK;
}

void scmf_server_destroy(scmf_server* sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->quit = true;
    sv->cv.notify_all();
  }
  if (sv->worker.joinable()) sv->worker.join();
  if (sv->batch) {
    (void)scmf_batch_sync(sv->batch);
    scmf_batch_destroy(sv->batch);
  }
  for (auto& d : sv->doomed) scmf_destroy(d.first);
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
  j->timeout = sv->o.job_timeout_s;
  j->t_submit = sv->now();
  *job_id = j->id;
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->queue.push_back(j->id);
  sv->jobs[j->id] = std::move(j);
  sv->cv.notify_all();
  return SCMF_OK;
}

int scmf_server_tick(scmf_server* sv, int32_t* n_queued, int32_t* n_running) {
  if (!sv) return sfail(SCMF_E_ARG, "null server");
  int rc = SCMF_OK;
  if (sv->batch && (rc = reap(sv, false))) return rc;
  const double now = sv->now();
  auto expired = [&](const Job& j) { return j.timeout > 0 && now - j.t_submit > j.timeout; };
  // cancellations and timeouts of running jobs
  for (uint64_t id : std::vector<uint64_t>(sv->running)) {
    Job& j = *sv->jobs[id];
    if (j.cancel || expired(j)) {
      {
        std::lock_guard<std::mutex> lk(sv->mu);
        j.state = SCMF_JOB_CANCELLED;
        if (!j.cancel) j.reason = SCMF_CANCEL_TIMEOUT;
        sv->cancelled += 1;
      }
      leave(sv, j);
    }
  }
  // queued and admitted jobs: timeouts, then the admitted ones join the batch
  std::vector<uint64_t> joining;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    std::deque<uint64_t> q;
    for (uint64_t id : sv->queue) {
      Job& j = *sv->jobs[id];
      if (expired(j)) {
        j.state = SCMF_JOB_CANCELLED;
        j.reason = SCMF_CANCEL_TIMEOUT;
        sv->cancelled += 1;
      } else {
        q.push_back(id);
      }
    }
    sv->queue.swap(q);
    for (uint64_t id : sv->ready) {
      Job& j = *sv->jobs[id];
      if (j.cancel || expired(j)) {
        j.state = SCMF_JOB_CANCELLED;
        if (!j.cancel) j.reason = SCMF_CANCEL_TIMEOUT;
        sv->cancelled += 1;
        sv->doomed.push_back({j.sys, j.bytes});
        j.sys = nullptr;
        sv->slots -= 1;
      } else {
        j.state = SCMF_JOB_RUNNING;
        joining.push_back(id);
      }
    }
    sv->ready.clear();
    sv->cv.notify_all();
  }
  for (uint64_t id : joining) sv->running.push_back(id);
  if (!joining.empty()) sv->dirty = true;
  // new membership: every running job with its stop at its own sweep count
  if (sv->dirty) {
    std::vector<scmf_sys*> m;
    std::vector<int64_t> stops;
    for (uint64_t id : sv->running) {
      m.push_back(sv->jobs[id]->sys);
      stops.push_back(sv->jobs[id]->target);
    }
    if (!sv->batch && !m.empty()) {
      rc = scmf_batch_create(&sv->batch, m.data(), (int32_t)m.size());
      if (!rc) rc = scmf_batch_set_lanes(sv->batch, 8);
      if (!rc) rc = scmf_batch_update(sv->batch, m.data(), (int32_t)m.size(), stops.data());
    } else if (sv->batch) {
      rc = scmf_batch_update(sv->batch, m.data(), (int32_t)m.size(), stops.data());
    }
    if (rc) return sfail(rc, std::string("batch: ") + scmf_last_error());
    sv->dirty = false;
    sv->rebuilds += 1;
    scmf_batch_stats_t bs;
    if (scmf_batch_stats(sv->batch, &bs) == SCMF_OK) {
      std::lock_guard<std::mutex> lk(sv->mu);
      sv->captures = bs.captures;
    }
  }
  // one tick of the whole batch
  bool enqueued = false;
  if (!sv->running.empty()) {
    const int32_t K = sv->o.steps_per_tick;
    uint64_t ticket = 0;
    rc = scmf_batch_tickets::enqueue(sv->batch, K, &ticket);
    if (rc) return sfail(rc, std::string("batch step: ") + scmf_last_error());
    enqueued = true;
    sv->ticks += 1;
    InFlight f{ticket, {}};
    for (uint64_t id : sv->running) {
      Job& j = *sv->jobs[id];
      j.last_ticket = ticket;
      j.enq += std::min<int64_t>(K, j.target - j.enq);
      if (j.enq >= j.target) f.finishing.push_back(id);
    }
    for (uint64_t id : f.finishing) {
      Job& j = *sv->jobs[id];
      j.finishing = true;
      leave(sv, j);
      j.parked = false;  // a finished job keeps its memory until released
      sv->parked.erase(std::remove(sv->parked.begin(), sv->parked.end(), id), sv->parked.end());
    }
    sv->inflight.push_back(std::move(f));
  }
  // at most two ticks queued; with nothing new to run, wait for the oldest
  while (sv->batch && sv->inflight.size() > 2)
    if ((rc = reap(sv, true))) return rc;
  if (sv->batch && !enqueued && !sv->inflight.empty())
    if ((rc = reap(sv, true))) return rc;
  std::unique_lock<std::mutex> lk(sv->mu);
  if (!enqueued && sv->inflight.empty() && sv->ready.empty() &&
      (sv->creating > 0 || !sv->queue.empty()))
    sv->cv.wait_for(lk, std::chrono::milliseconds(2), [&] { return !sv->ready.empty(); });
  int32_t finishing = 0;
  for (const InFlight& f : sv->inflight) finishing += (int32_t)f.finishing.size();
  if (n_queued) *n_queued = (int32_t)(sv->queue.size() + sv->ready.size()) + sv->creating;
  if (n_running) *n_running = (int32_t)sv->running.size() + finishing;
  return SCMF_OK;
}

int scmf_server_info(scmf_server* sv, uint64_t job_id, scmf_job_info* out) {
  if (!sv || !out) return sfail(SCMF_E_ARG, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->jobs.find(job_id);
  if (it == sv->jobs.end()) return sfail(SCMF_E_ARG, "unknown job");
  const Job& j = *it->second;
  memset(out, 0, sizeof *out);
  out->state = j.state;
  out->reason = j.reason;
  out->sweeps_done = j.state == SCMF_JOB_DONE ? j.target : j.enq;
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
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->jobs.find(job_id);
  if (it == sv->jobs.end() || !it->second->sys) return sfail(SCMF_E_ARG, "job has no resident state");
  if (it->second->state != SCMF_JOB_DONE) return sfail(SCMF_E_ARG, "job is not finished");
  *out = it->second->sys;
  return SCMF_OK;
}

int scmf_server_release(scmf_server* sv, uint64_t job_id) {
  if (!sv) return sfail(SCMF_E_ARG, "null server");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->jobs.find(job_id);
  if (it == sv->jobs.end()) return sfail(SCMF_E_ARG, "unknown job");
  Job& j = *it->second;
  if (j.state == SCMF_JOB_RUNNING) return sfail(SCMF_E_ARG, "job is still running");
  if (j.state == SCMF_JOB_QUEUED) {
    if (std::find(sv->queue.begin(), sv->queue.end(), job_id) == sv->queue.end())
      return sfail(SCMF_E_ARG, "job is being admitted; cancel it instead");
    sv->queue.erase(std::remove(sv->queue.begin(), sv->queue.end(), job_id), sv->queue.end());
  }
  if (j.sys && !j.parked) {
    // a finished job: its last ticket has run (it is DONE), and the batch drops
    // it before the next tick
    sv->doomed.push_back({j.sys, j.bytes});
    j.sys = nullptr;
    sv->cv.notify_all();
  }
  j.state = SCMF_JOB_RELEASED;
  return SCMF_OK;
}

int scmf_server_cancel(scmf_server* sv, uint64_t job_id) {
  if (!sv) return sfail(SCMF_E_ARG, "null server");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->jobs.find(job_id);
  if (it == sv->jobs.end()) return sfail(SCMF_E_ARG, "unknown job");
  Job& j = *it->second;
  if (j.state != SCMF_JOB_QUEUED && j.state != SCMF_JOB_RUNNING)
    return sfail(SCMF_E_ARG, "only a queued or running job can be cancelled");
  if (j.finishing) return sfail(SCMF_E_ARG, "the job's last sweep is already queued");
  j.cancel = true;
  j.reason = SCMF_CANCEL_CALLER;
  auto q = std::find(sv->queue.begin(), sv->queue.end(), job_id);
  if (q != sv->queue.end()) {  // not admitted yet
    sv->queue.erase(q);
    j.state = SCMF_JOB_CANCELLED;
    sv->cancelled += 1;
    sv->cv.notify_all();
  }
  return SCMF_OK;
}

int scmf_server_set_timeout(scmf_server* sv, uint64_t job_id, double seconds) {
  if (!sv) return sfail(SCMF_E_ARG, "null server");
  if (!(seconds >= 0)) return sfail(SCMF_E_ARG, "the limit must be non-negative");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->jobs.find(job_id);
  if (it == sv->jobs.end()) return sfail(SCMF_E_ARG, "unknown job");
  it->second->timeout = seconds;
  return SCMF_OK;
}

int scmf_server_stats(scmf_server* sv, scmf_server_stats_t* out) {
  if (!sv || !out) return sfail(SCMF_E_ARG, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  memset(out, 0, sizeof *out);
  out->ticks = sv->ticks;
  out->rebuilds = sv->rebuilds;
  out->admitted = sv->admitted;
  out->finished = sv->finished;
  out->cancelled = sv->cancelled;
  out->failed = sv->failed;
  out->busy_s = sv->busy_s;
  out->span_s = sv->first_start >= 0 ? sv->last_end - sv->first_start : 0.0;
  out->used_bytes = sv->used;
  out->captures = sv->captures;
  return SCMF_OK;
}

const char* scmf_server_last_error(void) { return g_server_err.c_str(); }

}  // extern "C"
