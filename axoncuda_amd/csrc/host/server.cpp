// server.cpp - continuous-batching partition service on one GPU (built on the C ABI).
//
// Jobs arrive at any time (hgp_server_submit) and run concurrently: every worker thread owns
// a context with its own HIP stream, so the kernels of different jobs overlap on the device
// and one job's host phases (initial partitioning, launch gaps, synchronisations) are
// covered by the others' device work.  A job is admitted when a worker is free and its
// estimated HBM footprint fits next to the jobs already running; the first queued job that
// fits goes next, so small jobs slip past a large one that must wait for memory.  Each job's
// level hierarchy stays resident on the device for its whole run, and the results do not
// depend on what else runs: every stage is deterministic and uses only the job's own data.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "hgp.h"

namespace {

using clock_type = std::chrono::steady_clock;

struct Job {
  uint64_t id = 0;
  const hgp_hgraph* hg = nullptr;
  hgp_params prm{};
  uint32_t n = 0;
  uint64_t footprint = 0;
  std::vector<uint32_t> part;
  hgp_run_stats st{};
  hgp_status status = HGP_OK;
  std::string error;
  bool done = false;
  clock_type::time_point t_submit, t_start, t_end;
};

}  // namespace

struct hgp_server {
  uint64_t budget = 0;
  std::vector<hgp_ctx*> ctxs;
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Job*> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
  uint64_t next_id = 1;
  uint64_t in_flight = 0;
  bool stop = false;
};

namespace {

// Peak device bytes of one pipeline run, from the level-0 sizes.  Measured peaks with every
// level resident (-smh 0), in bytes per level-0 pin: random hypergraphs (|e| = 2 + Geom(1/2))
// 71 at 10^4 pins, 118 at 2 x 10^5, 192 at 3 x 10^6 and 232 at 2 x 10^7 (k = 8: twelve levels
// that shrink slowly, then the device initial partitioner on a coarsest level of millions of
// pins); layered SNNs 61-139; the 3.2M-node netlist 166.  The bound keeps about 10% above the
// largest of them.
uint64_t footprint(uint32_t n, uint32_t e, uint64_t p) {
  return 256ull * p + 128ull * n + 64ull * e + (64ull << 20);
}

void worker_loop(hgp_server* s, hgp_ctx* ctx) {
  for (;;) {
    Job* job = nullptr;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] {
        if (s->stop) return true;
        for (Job* j : s->queue)
          if (s->in_flight == 0 || s->in_flight + j->footprint <= s->budget) return true;
        return false;
      });
      if (s->stop && s->queue.empty()) return;
      for (auto it = s->queue.begin(); it != s->queue.end(); ++it) {
        if (s->in_flight == 0 || s->in_flight + (*it)->footprint <= s->budget) {
          job = *it;
          s->queue.erase(it);
          break;
        }
      }
      if (!job) continue;
      s->in_flight += job->footprint;
    }
    job->t_start = clock_type::now();
    job->part.assign(job->n, 0);
    job->status = hgp_partition(ctx, &job->prm, job->hg, job->part.data(), &job->st);
    if (job->status != HGP_OK) job->error = hgp_ctx_last_error(ctx);
    // The worker's context keeps the job's freed blocks for reuse; admission counts only
    // running jobs, so give them back before the job's reservation is released.
    (void)hgp_ctx_trim(ctx);
    job->t_end = clock_type::now();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->in_flight -= job->footprint;
      job->done = true;
    }
    s->cv_done.notify_all();
    s->cv_work.notify_all();
  }
}

}  // namespace

extern "C" {

uint64_t hgp_job_footprint(const hgp_hgraph* hg, const hgp_params* params) {
  (void)params;
  if (!hg) return 0;
  uint32_t n = 0, e = 0;
  uint64_t p = 0;
  hgp_hgraph_sizes(hg, &n, &e, &p);
  return footprint(n, e, p);
}

hgp_status hgp_server_create(int device, uint64_t hbm_budget, uint32_t workers, hgp_server** out) {
  if (!out || workers == 0) return HGP_E_INVALID;
  *out = nullptr;
  auto* s = new hgp_server();
  for (uint32_t i = 0; i < workers; ++i) {
    hgp_ctx* c = nullptr;
    const hgp_status st = hgp_ctx_create(device, 0, &c);
    if (st != HGP_OK) {
      for (hgp_ctx* x : s->ctxs) hgp_ctx_destroy(x);
      delete s;
      return st;
    }
    // the workers already run side by side: more streams than the process's hardware queues
    // would queue one worker's kernels behind another's, so the workers keep to one stream
    // each; and they sort without grid barriers, so no number of workers can hold the device's
    // slots with workgroups that wait for each other
    hgp_tuning t;
    if (hgp_tuning_get(c, &t) == HGP_OK) {
      t.side_streams = 0;
      t.coop_max_n = 0;
      (void)hgp_tuning_set(c, &t);
    }
    s->ctxs.push_back(c);
  }
  if (hbm_budget == 0) {
    size_t free_b = 0, total_b = 0;
    if (hipSetDevice(device) != hipSuccess || hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
      (void)hipGetLastError();
      free_b = 64ull << 30;
    }
    hbm_budget = free_b / 10 * 9;
  }
  s->budget = hbm_budget;
  for (hgp_ctx* c : s->ctxs) s->workers.emplace_back(worker_loop, s, c);
  *out = s;
  return HGP_OK;
}

hgp_status hgp_server_submit(hgp_server* s, const hgp_hgraph* hg, const hgp_params* params, uint64_t* job_id) {
  if (!s || !hg || !params || !job_id) return HGP_E_INVALID;
  uint32_t n = 0, e = 0;
  uint64_t p = 0;
  hgp_hgraph_sizes(hg, &n, &e, &p);
  auto job = std::make_unique<Job>();
  job->hg = hg;
  job->prm = *params;
  job->n = n;
  job->footprint = hgp_job_footprint(hg, params);
  job->t_submit = clock_type::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    job->id = s->next_id++;
    *job_id = job->id;
    s->queue.push_back(job.get());
    s->jobs.emplace(job->id, std::move(job));
  }
  s->cv_work.notify_all();
  return HGP_OK;
}

hgp_status hgp_server_wait(hgp_server* s, uint64_t job_id, uint32_t* h_partitions_out, hgp_run_stats* stats,
                           double* latency_ms, double* queue_ms) {
  if (!s) return HGP_E_INVALID;
  std::unique_ptr<Job> job;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return HGP_E_INVALID;
    Job* j = it->second.get();
    s->cv_done.wait(lk, [&] { return j->done; });
    job = std::move(it->second);
    s->jobs.erase(it);
  }
  if (job->status == HGP_OK && h_partitions_out)
    std::memcpy(h_partitions_out, job->part.data(), (size_t)job->n * sizeof(uint32_t));
  if (stats) *stats = job->st;
  using ms = std::chrono::duration<double, std::milli>;
  if (latency_ms) *latency_ms = ms(job->t_end - job->t_submit).count();
  if (queue_ms) *queue_ms = ms(job->t_start - job->t_submit).count();
  return job->status;
}

void hgp_server_destroy(hgp_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& t : s->workers) t.join();
  for (hgp_ctx* c : s->ctxs) hgp_ctx_destroy(c);
  delete s;
}

}  // extern "C"
