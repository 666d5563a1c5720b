// Serving loop over the engine's jobs: a request queue with continuous
// batching.  Requests (plain IS, AMIS, NPMC, DM-PMC descriptors) are copied at
// submission; a dispatcher admits them in arrival order whenever an execution
// slot is free and their device-memory estimate fits the budget; each admitted
// job is created (its proposal slots, mixture sums and samples stay resident
// on the device), run on its own stream by a worker thread, and held until the
// caller takes it.  Jobs in flight overlap on the device; each job's results
// depend only on its descriptor (its own counter-RNG key, its own buffers and
// a fixed reduction order), so they are the same alone and in the mix.
// The reference has no serving layer; this is the engine's request loop for
// importance_sample! calls (src/methods/importance_sampling.jl:912-930).
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/issampler.h"
#include "host.h"

using namespace issampler::host;

namespace {

int64_t now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// A request's descriptor with every array it points into owned by the request.
struct OwnedDesc {
  is_job_desc d{};
  std::vector<int64_t> round_size;
  std::vector<double> location, scale, bank_locations, bank_scales, bank_masses, target_params;
  std::vector<int32_t> group_offsets, group_members, transform_kinds;
  std::vector<double> transform_bounds;
  std::vector<is_transform_block> transform_blocks;

  int copy_from(const is_job_desc& s) {
    d = s;
    const int T = (s.method == IS_PLAIN || s.method == IS_MIS) ? 1 : s.rounds;
    const int dim = s.dim;
    if (T < 1 || !s.round_size || dim < 1) return fail(IS_EINVAL, "descriptor needs rounds, round sizes and dim");
    round_size.assign(s.round_size, s.round_size + T);
    const int scale_len = s.scale_kind == IS_SPHERICAL ? 1 : s.scale_kind == IS_DIAGONAL ? dim : dim * dim;
    if (s.method == IS_DMPMC || s.method == IS_MIS) {
      const int K = s.bank_size;
      if (K < 1 || !s.bank_locations || !s.bank_scales || !s.bank_masses)
        return fail(IS_EINVAL, "a proposal bank is required");
      bank_locations.assign(s.bank_locations, s.bank_locations + (size_t)K * dim);
      bank_scales.assign(s.bank_scales, s.bank_scales + (size_t)K * scale_len);
      bank_masses.assign(s.bank_masses, s.bank_masses + K);
      if (s.method == IS_MIS && s.mis_scheme == IS_MIS_PARTIAL) {
        const int G = s.mis_group_count;
        if (G < 1 || !s.mis_group_offsets || !s.mis_group_members) return fail(IS_EINVAL, "partial-MIS groups cannot be empty");
        group_offsets.assign(s.mis_group_offsets, s.mis_group_offsets + G + 1);
        const int nm = group_offsets[G];
        if (nm < 0) return fail(IS_EINVAL, "invalid partial-MIS group offsets");
        group_members.assign(s.mis_group_members, s.mis_group_members + nm);
      }
    } else {
      if (!s.location || !s.scale) return fail(IS_EINVAL, "proposal location and scale are required");
      location.assign(s.location, s.location + dim);
      scale.assign(s.scale, s.scale + scale_len);
    }
    if (s.target_params && s.target_params_len > 0)
      target_params.assign(s.target_params, s.target_params + s.target_params_len);
    d.round_size = round_size.data();
    d.location = location.empty() ? nullptr : location.data();
    d.scale = scale.empty() ? nullptr : scale.data();
    d.bank_locations = bank_locations.empty() ? nullptr : bank_locations.data();
    d.bank_scales = bank_scales.empty() ? nullptr : bank_scales.data();
    d.bank_masses = bank_masses.empty() ? nullptr : bank_masses.data();
    d.target_params = target_params.empty() ? nullptr : target_params.data();
    if (s.transform_kinds) {
      if (!s.transform_bounds) return fail(IS_EINVAL, "transforms need bounds (2 per coordinate)");
      transform_kinds.assign(s.transform_kinds, s.transform_kinds + dim);
      transform_bounds.assign(s.transform_bounds, s.transform_bounds + 2 * dim);
      d.transform_kinds = transform_kinds.data();
      d.transform_bounds = transform_bounds.data();
    }
    if (s.transform_block_count > 0) {
      if (!s.transform_blocks) return fail(IS_EINVAL, "a transform layout needs its blocks");
      transform_blocks.assign(s.transform_blocks, s.transform_blocks + s.transform_block_count);
      d.transform_blocks = transform_blocks.data();
    }
    d.mis_group_offsets = group_offsets.empty() ? nullptr : group_offsets.data();
    d.mis_group_members = group_members.empty() ? nullptr : group_members.data();
    return IS_OK;
  }
};

struct Request {
  OwnedDesc desc;
  int64_t bytes = 0;
  int32_t state = IS_REQ_QUEUED;
  int rc = IS_OK;
  std::string error;
  is_job* job = nullptr;
  int64_t submit_ns = 0, admit_ns = 0, done_ns = 0;
};

}  // namespace

struct is_server {
  is_ctx* ctx = nullptr;
  is_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_dispatch, cv_work, cv_done;
  std::map<uint64_t, std::unique_ptr<Request>> reqs;
  std::deque<uint64_t> pending;   // submitted, not admitted (arrival order)
  std::deque<uint64_t> admitted;  // admitted, waiting for a worker
  uint64_t next_ticket = 1;
  int64_t bytes_in_use = 0;       // admitted and not yet taken
  int active = 0;                 // admitted and not finished
  bool stopping = false;
  is_server_stats stats{};
  std::thread dispatcher;
  std::vector<std::thread> workers;
};

namespace {

void dispatch_loop(is_server* s) {
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_dispatch.wait(lk, [&] {
      if (s->stopping) return true;
      if (s->pending.empty() || s->active >= s->cfg.max_active) return false;
      const Request& r = *s->reqs[s->pending.front()];
      return s->bytes_in_use + r.bytes <= s->cfg.memory_budget;
    });
    if (s->stopping) return;
    // admit in arrival order while a slot and the memory budget allow
    while (!s->pending.empty() && s->active < s->cfg.max_active) {
      const uint64_t t = s->pending.front();
      Request& r = *s->reqs[t];
      if (s->bytes_in_use + r.bytes > s->cfg.memory_budget) break;
      s->pending.pop_front();
      r.state = IS_REQ_RUNNING;
      r.admit_ns = now_ns();
      s->bytes_in_use += r.bytes;
      s->active += 1;
      s->stats.peak_active = std::max<int64_t>(s->stats.peak_active, s->active);
      s->stats.peak_bytes = std::max<int64_t>(s->stats.peak_bytes, s->bytes_in_use);
      s->admitted.push_back(t);
      s->cv_work.notify_one();
    }
  }
}

void worker_loop(is_server* s) {
  for (;;) {
    uint64_t t;
    Request* r;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] { return s->stopping || !s->admitted.empty(); });
      if (s->stopping && s->admitted.empty()) return;
      t = s->admitted.front();
      s->admitted.pop_front();
      r = s->reqs[t].get();
    }
    // create the job (device buffers), run it on its own stream
    is_job* job = nullptr;
    int rc = is_job_create(s->ctx, &r->desc.d, &job);
    if (rc == IS_OK) {
      (void)is_job_set_graph(job, s->cfg.use_graph);
      rc = is_job_run(job);
    }
    std::string err = rc == IS_OK ? std::string() : std::string(is_last_error());
    {
      std::lock_guard<std::mutex> lk(s->mu);
      r->rc = rc;
      r->error = err;
      r->job = job;
      r->done_ns = now_ns();
      r->state = rc == IS_OK ? IS_REQ_DONE : IS_REQ_FAILED;
      s->active -= 1;
      if (rc == IS_OK)
        s->stats.completed += 1;
      else
        s->stats.failed += 1;
      if (!job) s->bytes_in_use -= r->bytes;  // nothing stays resident
      if (!job) r->bytes = 0;
    }
    s->cv_dispatch.notify_one();
    s->cv_done.notify_all();
  }
}

}  // namespace

extern "C" {

int is_server_create(is_ctx* ctx, const is_server_config* cfg, is_server** out) {
  if (!ctx || !cfg || !out) return fail(IS_EINVAL, "null argument");
  if (cfg->max_active < 1 || cfg->max_active > 64) return fail(IS_EINVAL, "max_active must be in [1, 64]");
  if (cfg->memory_budget <= 0) return fail(IS_EINVAL, "memory budget must be positive");
  auto s = new is_server();
  s->ctx = ctx;
  s->cfg = *cfg;
  s->dispatcher = std::thread(dispatch_loop, s);
  for (int i = 0; i < cfg->max_active; ++i) s->workers.emplace_back(worker_loop, s);
  *out = s;
  return IS_OK;
}

int is_server_submit(is_server* s, const is_job_desc* desc, uint64_t* ticket) {
  if (!s || !desc || !ticket) return fail(IS_EINVAL, "null argument");
  auto r = std::make_unique<Request>();
  int rc = r->desc.copy_from(*desc);
  if (rc) return rc;
  r->bytes = is_job_device_bytes(&r->desc.d);
  if (r->bytes < 0) return fail(IS_EINVAL, "invalid job descriptor");
  if (r->bytes > s->cfg.memory_budget)
    return fail(IS_ENOMEM, "job needs %lld bytes, more than the %lld-byte budget", (long long)r->bytes,
                (long long)s->cfg.memory_budget);
  r->submit_ns = now_ns();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) return fail(IS_ESTATE, "server is shutting down");
    const uint64_t t = s->next_ticket++;
    s->reqs[t] = std::move(r);
    s->pending.push_back(t);
    s->stats.submitted += 1;
    *ticket = t;
  }
  s->cv_dispatch.notify_one();
  return IS_OK;
}

int is_server_poll(is_server* s, uint64_t ticket, int32_t* state) {
  if (!s || !state) return fail(IS_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return fail(IS_EINVAL, "unknown ticket");
  *state = it->second->state;
  return IS_OK;
}

int is_server_wait(is_server* s, uint64_t ticket, int32_t timeout_ms, int32_t* state) {
  if (!s) return fail(IS_EINVAL, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return fail(IS_EINVAL, "unknown ticket");
  Request* r = it->second.get();
  auto done = [&] { return r->state == IS_REQ_DONE || r->state == IS_REQ_FAILED; };
  if (timeout_ms < 0)
    s->cv_done.wait(lk, done);
  else
    s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done);
  if (state) *state = r->state;
  return IS_OK;
}

int is_server_take(is_server* s, uint64_t ticket, is_job** job, is_server_timing* timing) {
  if (!s || !job) return fail(IS_EINVAL, "null argument");
  std::unique_ptr<Request> r;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(IS_EINVAL, "unknown ticket");
    if (it->second->state != IS_REQ_DONE && it->second->state != IS_REQ_FAILED)
      return fail(IS_ESTATE, "request has not finished");
    r = std::move(it->second);
    s->reqs.erase(it);
    s->bytes_in_use -= r->bytes;  // the caller now owns the job's device memory
  }
  s->cv_dispatch.notify_one();
  if (timing) {
    timing->submit_ns = r->submit_ns;
    timing->admit_ns = r->admit_ns;
    timing->done_ns = r->done_ns;
    timing->device_bytes = r->bytes;
  }
  *job = r->job;
  if (r->rc != IS_OK) {
    const int rc = r->rc;
    if (!r->job) return fail(rc, "%s", r->error.c_str());
    g_last_error = r->error;
    return rc;  // the failed job is handed over for is_job_failure
  }
  return IS_OK;
}

int is_server_get_stats(is_server* s, is_server_stats* out) {
  if (!s || !out) return fail(IS_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  out->active = s->active;
  out->queued = (int64_t)s->pending.size();
  out->bytes_in_use = s->bytes_in_use;
  return IS_OK;
}

int is_server_destroy(is_server* s) {
  if (!s) return IS_OK;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    // finish what was admitted; drop what is still queued
    for (uint64_t t : s->pending) s->reqs.erase(t);
    s->pending.clear();
    s->cv_done.wait(lk, [&] { return s->active == 0; });
    s->stopping = true;
  }
  s->cv_dispatch.notify_all();
  s->cv_work.notify_all();
  if (s->dispatcher.joinable()) s->dispatcher.join();
  for (auto& w : s->workers)
    if (w.joinable()) w.join();
  for (auto& kv : s->reqs)
    if (kv.second->job) (void)is_job_destroy(kv.second->job);
  delete s;
  return IS_OK;
}

}  // extern "C"
