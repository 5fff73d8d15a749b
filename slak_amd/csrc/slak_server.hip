// Serving: a request loop that refines subdivision requests with continuous
// batching. Requests are admitted while their projected device footprint fits
// the memory budget; a step refines every admitted request to its target
// level in the batched level kernels (one launch per kernel per level over
// the union of the requests still active at that level), and several steps
// run at once, one per slot, so requests that arrive while others are on the
// GPU start beside them. Intermediate levels and their structure stay
// resident on the device; a finished request's refined mesh is handed to the
// caller. No counterpart in the reference, which refines one mesh per
// blocking call.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <thread>

#include "slak_c.h"
#include "slak_plan.hpp"

namespace slak {
int guarded_call(const std::function<void()>& fn);
}

namespace {

using Clock = std::chrono::steady_clock;

double us_since(Clock::time_point t0, Clock::time_point t) {
  return std::chrono::duration<double, std::micro>(t - t0).count();
}

struct Request {
  uint64_t id = 0;
  slak_mesh in{};
  int target = 0;
  int levels = 0;
  uint64_t footprint = 0;
  Clock::time_point t_submit, t_admit;
  // result: written by the worker under the server mutex together with
  // `finished`, read by the waiter under the same mutex
  slak_mesh cur{};
  Clock::time_point t_done;
  int status = SLAK_OK;
  std::string error;
  bool finished = false;
  int steps = 0;
};

// Closed-form sizes of a request's levels, with the edge count bounded by the
// corner count: n_f' = nnz, nnz' = 4 nnz, n_v' = n_v + n_f + n_e.
struct ChainBound {
  double peak_bytes = 0;   // device bytes while the chain runs
  double last_corners = 0; // corners of the input of the last level
  double last_verts = 0;   // vertices of the input of the last level
};

ChainBound chain_bound(const slak_mesh& m, int levels) {
  ChainBound cb;
  double nv = (double)m.nverts, nf = (double)m.nfaces, nnz = (double)m.nnz;
  double held = 16 * nv + 4 * nnz;  // the caller's input (gathered once)
  double peak = 0;
  for (int l = 0; l < std::max(levels, 1); ++l) {
    const double ne = nnz;  // bound
    const double nv2 = nv + nf + ne, nnz2 = 4 * nnz;
    cb.last_corners = nnz;
    cb.last_verts = nv;
    // this level's input structure (M^T rows with in-neighbours, (edge,
    // twin) per corner, boundary flags) and its refined positions and ids
    // (with the LAK form at the last level)
    const double level = 28 * nnz + 8 * nv + 16 * nv2 + 4 * nnz2 * 2 + 4 * nnz;
    peak = std::max(peak, held + level);
    held += 16 * nv2 + 4 * nnz2;  // a level stays resident while the next one runs
    nv = nv2;
    nf = nnz;
    nnz = nnz2;
  }
  cb.peak_bytes = peak;
  return cb;
}

}  // namespace

// Continuous batching over `slots` concurrent steps. A step takes the pending
// requests that fit (FIFO, by projected device footprint and max_batch) and
// refines every one of them to its target level in one multi-mesh plan: one
// launch per kernel per level over the union of the requests still active at
// that level, the first level's analysis once per request, every later
// level's structure in closed form from the level before it, the last level
// from the quad mesh two levels back. Each slot has its own worker thread
// and stream, so a step starts as soon as requests are pending and a slot is
// free: requests that arrive while another step is on the GPU run beside it.
struct slak_server {
  slak_server_config cfg{};
  int nslots = 2;
  std::vector<hipStream_t> streams;
  std::mutex mu;
  std::condition_variable cv, done_cv;
  std::deque<std::shared_ptr<Request>> pending;
  std::map<uint64_t, std::shared_ptr<Request>> all;
  uint64_t next_id = 1;
  uint64_t inflight_bytes = 0;
  int running = 0;  // steps on the GPU
  bool stop = false;
  std::vector<std::thread> workers;
  Clock::time_point t0 = Clock::now();
  // counters (under mu)
  uint64_t steps = 0, completed = 0, faces_done = 0, active = 0;
  double busy_us = 0;
  int max_batch_seen = 0;
  // corners (and vertices) one plan's union may hold at any level: the int32
  // index range, or less with SLAK_SERVE_STEP_CORNERS
  double step_limit = 2147483647.0;

  void loop(int slot);
  bool admissible(const Request& r) const {
    return inflight_bytes == 0 || inflight_bytes + r.footprint <= cfg.memory_budget_bytes;
  }
  void step(std::vector<std::shared_ptr<Request>>& batch, hipStream_t st);
  void run_plan(std::vector<std::shared_ptr<Request>>& batch, hipStream_t st);
  void publish(Request& r, const slak_mesh& m, int status, const std::string& err, Clock::time_point t);
};

// Sets a request's result and marks it finished, under the mutex the waiters
// read it with.
void slak_server::publish(Request& r, const slak_mesh& m, int status, const std::string& err, Clock::time_point t) {
  std::lock_guard<std::mutex> lk(mu);
  r.cur = m;
  r.status = status;
  r.error = err;
  r.t_done = t;
  r.steps += 1;
  r.finished = true;
  inflight_bytes -= r.footprint;
  --active;
  ++completed;
  if (status == SLAK_OK) faces_done += m.nfaces;
}

// A plan indexes the union of its meshes with int32 at every level: the step
// runs the batch as consecutive sub-batches whose deepest level fits.
void slak_server::step(std::vector<std::shared_ptr<Request>>& batch, hipStream_t st) {
  const double lim = step_limit;
  size_t i = 0;
  while (i < batch.size()) {
    size_t j = i;
    double corners = 0, verts = 0;
    while (j < batch.size()) {
      const ChainBound b = chain_bound(batch[j]->in, batch[j]->levels);
      const double c = 4.0 * b.last_corners, v = b.last_verts + 2.0 * b.last_corners;
      if (j > i && (corners + c > lim || verts + v > lim)) break;
      corners += c;
      verts += v;
      ++j;
    }
    std::vector<std::shared_ptr<Request>> sub(batch.begin() + i, batch.begin() + j);
    run_plan(sub, st);
    i = j;
  }
}

void slak_server::run_plan(std::vector<std::shared_ptr<Request>>& batch, hipStream_t st) {
  using namespace slak;
  while (!batch.empty()) {
    std::vector<MeshJob> jobs;
    jobs.reserve(batch.size());
    for (auto& r : batch) {
      MeshJob J;
      J.in = r->in;
      J.target = r->target;
      jobs.push_back(J);
    }
    Plan plan(cfg.scheme, std::move(jobs), OutputMode::kSeparate, st);
    try {
      plan.prepare();
    } catch (const Error& e) {
      // fail the requests whose meshes were rejected and run the rest again
      std::vector<std::shared_ptr<Request>> rest;
      bool named = false;
      for (size_t i = 0; i < batch.size(); ++i)
        if (plan.jobs()[i].bad_flags) named = true;
      for (size_t i = 0; i < batch.size(); ++i) {
        const int bad = plan.jobs()[i].bad_flags;
        if (bad || !named)
          publish(*batch[i], slak_mesh{}, e.code,
                  bad ? std::string("invalid mesh:") + flags_message(bad) : std::string(e.what()), Clock::now());
        else
          rest.push_back(batch[i]);
      }
      batch.swap(rest);
      continue;
    }
    plan.launch(false);
    plan.finish();
    plan.release_outputs();
    const auto t_end = Clock::now();
    for (size_t i = 0; i < batch.size(); ++i) {
      MeshJob& J = plan.jobs()[i];
      slak_mesh m{};
      const int L = J.levels;
      m.type = SLAK_QUAD;
      m.nfaces = (uint64_t)J.nf[L];
      m.nnz = (uint64_t)J.nnz[L];
      m.nverts = (uint64_t)J.nv[L];
      m.max_face_order = 4;
      m.ids = J.out_ids;
      m.verts = reinterpret_cast<float*>(J.out_verts);
      m.ptr = J.out_ptr;
      m.vals = J.out_vals;
      m.is_reduced = cfg.scheme == SLAK_SCHEME_LAK ? 0 : 1;
      publish(*batch[i], m, SLAK_OK, std::string(), t_end);
    }
    return;
  }
}

void slak_server::loop(int slot) {
  (void)hipSetDevice(cfg.device);
  hipStream_t st = streams[slot];
  for (;;) {
    std::vector<std::shared_ptr<Request>> batch;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] { return (stop && pending.empty()) || (!pending.empty() && admissible(*pending.front())); });
      if (pending.empty()) return;  // stopping, nothing left
      // admission by projected device footprint, FIFO
      const auto now = Clock::now();
      while (!pending.empty() && (int)batch.size() < cfg.max_batch && admissible(*pending.front())) {
        auto r = pending.front();
        pending.pop_front();
        r->t_admit = now;
        inflight_bytes += r->footprint;
        batch.push_back(r);
      }
      active += batch.size();
      ++running;
      ++steps;
      max_batch_seen = std::max(max_batch_seen, (int)batch.size());
    }
    const auto t_step = Clock::now();
    try {
      step(batch, st);
    } catch (const std::exception& e) {
      (void)hipStreamSynchronize(st);
      for (auto& r : batch) {
        bool done;
        {
          std::lock_guard<std::mutex> lk(mu);
          done = r->finished;
        }
        if (!done) publish(*r, slak_mesh{}, SLAK_ERR_INTERNAL, e.what(), Clock::now());
      }
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      busy_us += us_since(t_step, Clock::now());
      --running;
    }
    done_cv.notify_all();
    cv.notify_all();  // memory freed: the other slots may admit again
  }
}

extern "C" {

int slak_server_start(const slak_server_config* cfg, slak_server** out) {
  return slak::guarded_call([&] {
    if (!cfg || !out) throw slak::Error(SLAK_ERR_INVALID_ARGUMENT, "null argument");
    if (cfg->scheme != SLAK_SCHEME_LAK && cfg->scheme != SLAK_SCHEME_SLAK)
      throw slak::Error(SLAK_ERR_INVALID_ARGUMENT, "unknown scheme");
    if (cfg->slots < 0 || cfg->slots > 8) throw slak::Error(SLAK_ERR_INVALID_ARGUMENT, "slots must be in 0..8");
    auto s = std::make_unique<slak_server>();
    s->cfg = *cfg;
    if (const char* e = std::getenv("SLAK_SERVE_STEP_CORNERS")) {
      const double x = std::atof(e);
      if (x > 0 && x < s->step_limit) s->step_limit = x;
    }
    if (s->cfg.max_batch <= 0) s->cfg.max_batch = 1024;
    if (s->cfg.memory_budget_bytes == 0) s->cfg.memory_budget_bytes = (uint64_t)64 << 30;
    s->nslots = s->cfg.slots > 0 ? s->cfg.slots : 2;
    SLAK_HIP_CHECK(hipSetDevice(s->cfg.device));
    slak_server* raw = s.get();
    auto cleanup = [raw] {
      for (hipStream_t st : raw->streams) (void)hipStreamDestroy(st);
      raw->streams.clear();
    };
    try {
      for (int k = 0; k < s->nslots; ++k) {
        hipStream_t st = nullptr;
        SLAK_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        s->streams.push_back(st);
      }
    } catch (...) {
      cleanup();
      throw;
    }
    for (int k = 0; k < s->nslots; ++k) s->workers.emplace_back([raw, k] { raw->loop(k); });
    *out = s.release();
  });
}

int slak_server_submit(slak_server* s, const slak_mesh* mesh, int32_t target_level, uint64_t* request_id) {
  return slak::guarded_call([&] {
    if (!s || !mesh || !request_id) throw slak::Error(SLAK_ERR_INVALID_ARGUMENT, "null argument");
    if (target_level < 0 || target_level > 12) throw slak::Error(SLAK_ERR_INVALID_ARGUMENT, "target level out of range");
    auto r = std::make_shared<Request>();
    r->in = *mesh;
    r->target = target_level;
    r->levels = mesh->type != SLAK_QUAD ? std::max(1, (int)target_level) : (int)target_level;
    r->footprint = (uint64_t)chain_bound(*mesh, r->levels).peak_bytes;
    r->t_submit = Clock::now();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->stop) throw slak::Error(SLAK_ERR_REJECTED, "server stopped");
      if (r->footprint > s->cfg.memory_budget_bytes)
        throw slak::Error(SLAK_ERR_REJECTED, "request exceeds the server's memory budget");
      r->id = s->next_id++;
      s->all[r->id] = r;
      if (r->levels == 0) {  // nothing to refine: done on arrival
        r->finished = true;
        r->t_admit = r->t_done = r->t_submit;
      } else {
        s->pending.push_back(r);
      }
      *request_id = r->id;
    }
    s->cv.notify_one();
  });
}

static void fill_result(const Request& r, slak_mesh* result, slak_request_stats* st, Clock::time_point t0) {
  if (result && r.status == SLAK_OK && r.levels > 0) *result = r.cur;  // finished: read under the mutex
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->status = r.status;
    st->levels = r.levels;
    st->steps = r.steps;
    st->submit_us = us_since(t0, r.t_submit);
    st->admit_us = us_since(t0, r.t_admit);
    st->done_us = us_since(t0, r.t_done);
    st->latency_us = us_since(r.t_submit, r.t_done);
    st->out_faces = r.levels > 0 ? r.cur.nfaces : 0;
  }
}

int slak_server_wait(slak_server* s, uint64_t request_id, int32_t timeout_ms, slak_mesh* result,
                     slak_request_stats* st) {
  return slak::guarded_call([&] {
    if (!s) throw slak::Error(SLAK_ERR_INVALID_ARGUMENT, "null server");
    std::shared_ptr<Request> r;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      auto it = s->all.find(request_id);
      if (it == s->all.end()) throw slak::Error(SLAK_ERR_INVALID_ARGUMENT, "unknown request id");
      r = it->second;
      auto pred = [&] { return r->finished; };
      if (timeout_ms < 0)
        s->done_cv.wait(lk, pred);
      else if (!s->done_cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred))
        throw slak::Error(SLAK_ERR_REJECTED, "timeout");
      s->all.erase(it);  // the result now belongs to the caller
    }
    fill_result(*r, result, st, s->t0);
    if (r->status != SLAK_OK) throw slak::Error(r->status, r->error);
  });
}

int slak_server_get_stats(slak_server* s, slak_server_stats* out) {
  return slak::guarded_call([&] {
    if (!s || !out) throw slak::Error(SLAK_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(out, 0, sizeof(*out));
    out->steps = s->steps;
    out->completed = s->completed;
    out->pending = s->pending.size();
    out->active = s->active;
    out->inflight_bytes = s->inflight_bytes;
    out->faces_done = s->faces_done;
    out->busy_us = s->busy_us;
    out->max_batch = s->max_batch_seen;
  });
}

int slak_server_stop(slak_server* s) {
  return slak::guarded_call([&] {
    if (!s) return;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
    }
    s->cv.notify_all();
    for (auto& w : s->workers)
      if (w.joinable()) w.join();
    // results nobody collected go back to the library
    for (hipStream_t st : s->streams) (void)hipStreamSynchronize(st);
    for (auto& kv : s->all) {
      Request& r = *kv.second;
      if (r.finished && r.status == SLAK_OK && r.levels > 0)
        for (void* p : {(void*)r.cur.ids, (void*)r.cur.verts, (void*)r.cur.ptr, (void*)r.cur.vals})
          if (p) slak::pool().release(p);
    }
    for (hipStream_t st : s->streams) (void)hipStreamDestroy(st);
    delete s;
  });
}

}  // extern "C"
