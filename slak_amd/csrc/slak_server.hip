// Serving: a request loop that refines subdivision requests with continuous
// batching. Requests are admitted while their projected device footprint fits
// the memory budget; every step advances each admitted request by one level,
// all of them together in the batched level kernels (one launch per kernel
// over the union), so requests that arrive while others are on the GPU join
// at the next step. Intermediate levels stay resident on the device; a
// finished request's refined mesh is handed to the caller. No counterpart in
// the reference, which refines one mesh per blocking call.
#include <hip/hip_runtime.h>

#include <atomic>
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
  int done_levels = 0;
  uint64_t footprint = 0;
  // current resident mesh (after done_levels > 0)
  slak_mesh cur{};
  int64_t cur_ne = -1;  // edges of cur (closed form)
  Clock::time_point t_submit, t_admit, t_done;
  int status = SLAK_OK;
  std::string error;
  bool finished = false;
  int steps = 0;
};

// Upper bound of the device bytes a request needs at its largest level:
// input + output positions and ids plus the level scratch (edge count bounded
// by the corner count).
uint64_t footprint(const slak_mesh& m, int levels) {
  double nv = (double)m.nverts, nf = (double)m.nfaces, nnz = (double)m.nnz;
  double peak = 0;
  for (int l = 0; l < std::max(levels, 1); ++l) {
    double ne = nnz;  // bound
    double nv2 = nv + nf + ne, nnz2 = 4 * nnz;
    double b = 16 * nv + 4 * nnz                  // input
               + 16 * nv2 + 4 * nnz2 * 2 + 4 * nnz  // output ids, LAK ptr/vals
               + 16 * nv + 28 * nnz;                // scratch: M^T, rows, (edge, twin), flags
    peak = std::max(peak, b);
    nv = nv2;
    nf = nnz;
    nnz = nnz2;
  }
  return (uint64_t)peak;
}

}  // namespace

struct slak_server {
  slak_server_config cfg{};
  hipStream_t stream = nullptr;
  std::mutex mu;
  std::condition_variable cv, done_cv;
  std::deque<std::shared_ptr<Request>> pending;
  std::vector<std::shared_ptr<Request>> active;
  std::map<uint64_t, std::shared_ptr<Request>> all;
  uint64_t next_id = 1;
  uint64_t inflight_bytes = 0;
  bool stop = false;
  std::thread worker;
  Clock::time_point t0 = Clock::now();
  // counters
  uint64_t steps = 0, completed = 0, faces_done = 0;
  double busy_us = 0;
  int max_batch_seen = 0;
  // refined corners (and vertices) one level plan of a step may hold: the
  // int32 index range, or less with SLAK_SERVE_STEP_CORNERS
  double step_limit = 2147483647.0;

  void loop();
  void step(std::vector<std::shared_ptr<Request>>& batch);
  void step_plan(std::vector<std::shared_ptr<Request>>& batch);
};

// A level plan indexes its union of meshes with int32: the step runs the
// batch as consecutive sub-batches whose refined corners and vertices fit.
void slak_server::step(std::vector<std::shared_ptr<Request>>& batch) {
  const double lim = step_limit;
  size_t i = 0;
  while (i < batch.size()) {
    size_t j = i;
    double corners = 0, verts = 0;
    while (j < batch.size()) {
      const Request& r = *batch[j];
      const slak_mesh& m = r.done_levels == 0 ? r.in : r.cur;
      const double c = 4.0 * (double)m.nnz, v = (double)m.nverts + (double)m.nfaces + (double)m.nnz;
      if (j > i && (corners + c > lim || verts + v > lim)) break;
      corners += c;
      verts += v;
      ++j;
    }
    std::vector<std::shared_ptr<Request>> sub(batch.begin() + i, batch.begin() + j);
    step_plan(sub);
    i = j;
  }
}

void slak_server::step_plan(std::vector<std::shared_ptr<Request>>& batch) {
  using namespace slak;
  std::vector<MeshJob> jobs;
  jobs.reserve(batch.size());
  for (auto& r : batch) {
    MeshJob J;
    if (r->done_levels == 0) {
      J.in = r->in;
    } else {
      J.in = r->cur;
      J.trusted = true;
      J.known_ne = r->cur_ne;
    }
    J.target = 1;
    J.lak_form = r->done_levels + 1 == r->levels;
    jobs.push_back(J);
  }
  auto t_step = Clock::now();
  Plan plan(cfg.scheme, std::move(jobs), OutputMode::kSeparate, stream);
  try {
    plan.prepare();
  } catch (const Error& e) {
    // fail the requests whose meshes were rejected, keep the rest for the next step
    bool named = false;
    for (size_t i = 0; i < batch.size(); ++i)
      if (plan.jobs()[i].bad_flags) {
        batch[i]->status = e.code;
        batch[i]->error = std::string("invalid mesh:") + flags_message(plan.jobs()[i].bad_flags);
        batch[i]->finished = true;
        named = true;
      }
    if (!named)
      for (auto& r : batch) {
        r->status = e.code;
        r->error = e.what();
        r->finished = true;
      }
    for (auto& r : batch)
      if (r->finished) r->t_done = Clock::now();
    return;
  }
  plan.launch(false);
  plan.finish();
  plan.release_outputs();
  auto t_end = Clock::now();
  busy_us += us_since(t_step, t_end);
  for (size_t i = 0; i < batch.size(); ++i) {
    Request& r = *batch[i];
    MeshJob& J = plan.jobs()[i];
    // the previous intermediate level is no longer needed
    if (r.done_levels > 0) {
      for (void* p : {(void*)r.cur.ids, (void*)r.cur.verts, (void*)r.cur.ptr, (void*)r.cur.vals})
        if (p) pool().release(p);
    }
    slak_mesh m{};
    m.type = SLAK_QUAD;
    m.nfaces = (uint64_t)J.nf[1];
    m.nnz = (uint64_t)J.nnz[1];
    m.nverts = (uint64_t)J.nv[1];
    m.max_face_order = 4;
    m.ids = J.out_ids;
    m.verts = reinterpret_cast<float*>(J.out_verts);
    m.ptr = J.out_ptr;
    m.vals = J.out_vals;
    m.is_reduced = cfg.scheme == SLAK_SCHEME_LAK ? 0 : 1;
    r.cur = m;
    r.cur_ne = J.ne[1];
    r.done_levels += 1;
    r.steps += 1;
    if (r.done_levels >= r.levels) {
      r.finished = true;
      r.t_done = t_end;
      faces_done += m.nfaces;
    }
  }
}

void slak_server::loop() {
  (void)hipSetDevice(cfg.device);
  for (;;) {
    std::vector<std::shared_ptr<Request>> batch;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] { return stop || !pending.empty() || !active.empty(); });
      if (stop && pending.empty() && active.empty()) return;
      // admission by projected device footprint, FIFO
      while (!pending.empty() && (int)active.size() < cfg.max_batch) {
        auto& r = pending.front();
        if (!active.empty() && inflight_bytes + r->footprint > cfg.memory_budget_bytes) break;
        r->t_admit = Clock::now();
        inflight_bytes += r->footprint;
        active.push_back(r);
        pending.pop_front();
      }
      batch = active;
    }
    max_batch_seen = std::max(max_batch_seen, (int)batch.size());
    try {
      step(batch);
    } catch (const std::exception& e) {
      for (auto& r : batch) {
        r->status = SLAK_ERR_INTERNAL;
        r->error = e.what();
        r->finished = true;
        r->t_done = Clock::now();
      }
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      ++steps;
      std::vector<std::shared_ptr<Request>> still;
      for (auto& r : active) {
        if (r->finished) {
          inflight_bytes -= r->footprint;
          ++completed;
        } else {
          still.push_back(r);
        }
      }
      active.swap(still);
    }
    done_cv.notify_all();
  }
}

extern "C" {

int slak_server_start(const slak_server_config* cfg, slak_server** out) {
  return slak::guarded_call([&] {
    if (!cfg || !out) throw slak::Error(SLAK_ERR_INVALID_ARGUMENT, "null argument");
    if (cfg->scheme != SLAK_SCHEME_LAK && cfg->scheme != SLAK_SCHEME_SLAK)
      throw slak::Error(SLAK_ERR_INVALID_ARGUMENT, "unknown scheme");
    auto s = std::make_unique<slak_server>();
    s->cfg = *cfg;
    if (const char* e = std::getenv("SLAK_SERVE_STEP_CORNERS")) {
      const double x = std::atof(e);
      if (x > 0 && x < s->step_limit) s->step_limit = x;
    }
    if (s->cfg.max_batch <= 0) s->cfg.max_batch = 1024;
    if (s->cfg.memory_budget_bytes == 0) s->cfg.memory_budget_bytes = (uint64_t)64 << 30;
    SLAK_HIP_CHECK(hipSetDevice(s->cfg.device));
    SLAK_HIP_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    slak_server* raw = s.get();
    s->worker = std::thread([raw] { raw->loop(); });
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
    r->footprint = footprint(*mesh, r->levels);
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
  if (result && r.status == SLAK_OK && r.levels > 0) *result = r.cur;
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
    out->active = s->active.size();
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
    if (s->worker.joinable()) s->worker.join();
    // results nobody collected go back to the library
    (void)hipStreamSynchronize(s->stream);
    for (auto& kv : s->all) {
      Request& r = *kv.second;
      if (r.levels > 0 && r.done_levels > 0)
        for (void* p : {(void*)r.cur.ids, (void*)r.cur.verts, (void*)r.cur.ptr, (void*)r.cur.vals})
          if (p) slak::pool().release(p);
    }
    (void)hipStreamDestroy(s->stream);
    delete s;
  });
}

}  // extern "C"
