// Host runtime of libslak.so: device pool, level orchestration and the C-ABI.
//
// Level chain (subdivideIteratively, LAKCatmullClarkImpl.cu:512-571 and
// CatmullClarkImpl.cu:457-515). Sizes of every level after the first follow
// from the first level's edge count in closed form (a manifold quad level
// with n_ext boundary edges refines to n_ext' = 2 n_ext, n_e' = 2 nnz + n_ext),
// so a call needs two blocking round trips (input validation, first-level
// edge count) plus the final synchronization; levels 2.. are launched back
// to back on one stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <functional>
#include <memory>

#include "slak_c.h"
#include "slak_internal.hpp"
#include "slak_plan.hpp"

namespace slak {

namespace {
thread_local std::string t_last_error;

constexpr size_t kAlign = 256;
constexpr size_t kLargeGranule = size_t(2) << 20;

size_t round_bytes(size_t b) {
  if (b == 0) b = 1;
  if (b >= (size_t(8) << 20)) return (b + kLargeGranule - 1) / kLargeGranule * kLargeGranule;
  return (b + kAlign - 1) / kAlign * kAlign;
}

}  // namespace

// ---------------------------------------------------------------------------
// Pool

void* Pool::alloc(size_t bytes) {
  size_t want = round_bytes(bytes);
  {
    std::lock_guard<std::mutex> g(mu_);
    auto it = free_.lower_bound(want);
    if (it != free_.end() && it->first <= want + want / 4 + kLargeGranule) {
      void* p = it->second;
      size_t sz = it->first;
      free_.erase(it);
      used_[p] = sz;
      cached_ -= sz;
      in_use_ += sz;
      peak_ = std::max(peak_, in_use_);
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    trim();
    e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error(SLAK_ERR_OUT_OF_MEMORY, "device allocation of " + std::to_string(want) + " bytes failed");
    }
  }
  std::lock_guard<std::mutex> g(mu_);
  used_[p] = want;
  in_use_ += want;
  peak_ = std::max(peak_, in_use_);
  return p;
}

bool Pool::release(void* p) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = used_.find(p);
  if (it == used_.end()) return false;
  size_t sz = it->second;
  used_.erase(it);
  in_use_ -= sz;
  free_.emplace(sz, p);
  cached_ += sz;
  return true;
}

bool Pool::owns(void* p) {
  std::lock_guard<std::mutex> g(mu_);
  return used_.count(p) != 0;
}

void Pool::trim() {
  std::multimap<size_t, void*> victims;
  {
    std::lock_guard<std::mutex> g(mu_);
    victims.swap(free_);
    cached_ = 0;
  }
  if (!victims.empty()) (void)hipDeviceSynchronize();
  for (auto& kv : victims) (void)hipFree(kv.second);
}

Pool& pool() {
  static std::mutex mu;
  static std::vector<std::unique_ptr<Pool>> pools;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> g(mu);
  if ((int)pools.size() <= dev) pools.resize(dev + 1);
  if (!pools[dev]) pools[dev].reset(new Pool());
  return *pools[dev];
}

void* Arena::get(size_t bytes) {
  const size_t b = (bytes + kAlign - 1) / kAlign * kAlign;
  if (b > 0 && b <= bump_left_) {
    void* p = bump_;
    bump_ += b;
    bump_left_ -= b;
    return p;
  }
  void* p = pool_.alloc(bytes);
  blocks_.push_back(p);
  bytes_ += bytes;
  return p;
}

void Arena::reserve(size_t bytes) {
  if (bytes == 0) return;
  bytes = (bytes + kAlign - 1) / kAlign * kAlign;
  void* p = pool_.alloc(bytes);
  blocks_.push_back(p);
  bytes_ += bytes;
  bump_ = static_cast<char*>(p);
  bump_left_ = bytes;
}

void Arena::keep(void* p) {
  auto it = std::find(blocks_.begin(), blocks_.end(), p);
  if (it != blocks_.end()) blocks_.erase(it);
}

void Arena::release_all() {
  for (void* p : blocks_) pool_.release(p);
  blocks_.clear();
  bump_ = nullptr;
  bump_left_ = 0;
}

// ---------------------------------------------------------------------------
// Profiler

Profiler& Profiler::get() {
  static Profiler p;
  return p;
}

hipEvent_t Profiler::event() {
  if (!spare_.empty()) {
    hipEvent_t e = spare_.back();
    spare_.pop_back();
    return e;
  }
  hipEvent_t e;
  SLAK_HIP_CHECK(hipEventCreate(&e));
  return e;
}

void Profiler::begin(hipStream_t st, const char* name, double bytes) {
  std::lock_guard<std::mutex> g(mu_);
  Rec r{name, bytes, event(), event()};
  SLAK_HIP_CHECK(hipEventRecord(r.a, st));
  recs_.push_back(r);
}

void Profiler::end(hipStream_t st) {
  std::lock_guard<std::mutex> g(mu_);
  if (!recs_.empty()) SLAK_HIP_CHECK(hipEventRecord(recs_.back().b, st));
}

void Profiler::amend_last(const char* name, double bytes) {
  std::lock_guard<std::mutex> g(mu_);
  for (auto it = recs_.rbegin(); it != recs_.rend(); ++it)
    if (std::strcmp(it->name, name) == 0) {
      it->bytes += bytes;
      return;
    }
}

void Profiler::reset() {
  std::lock_guard<std::mutex> g(mu_);
  for (auto& r : recs_) {
    (void)hipEventSynchronize(r.b);
    spare_.push_back(r.a);
    spare_.push_back(r.b);
  }
  recs_.clear();
}

std::string Profiler::report() {
  std::lock_guard<std::mutex> g(mu_);
  struct Agg {
    int64_t n = 0;
    double ms = 0, bytes = 0;
  };
  std::vector<std::pair<std::string, Agg>> agg;
  for (auto& r : recs_) {
    SLAK_HIP_CHECK(hipEventSynchronize(r.b));
    float ms = 0;
    SLAK_HIP_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
    auto it = std::find_if(agg.begin(), agg.end(), [&](auto& kv) { return kv.first == r.name; });
    if (it == agg.end()) {
      agg.push_back({r.name, Agg{}});
      it = agg.end() - 1;
    }
    it->second.n += 1;
    it->second.ms += ms;
    it->second.bytes += r.bytes;
  }
  std::string js = "{\"stages\": [";
  for (size_t i = 0; i < agg.size(); ++i) {
    char buf[256];
    snprintf(buf, sizeof(buf), "%s{\"name\": \"%s\", \"launches\": %lld, \"total_ms\": %.6f, \"bytes\": %.0f}",
             i ? ", " : "", agg[i].first.c_str(), (long long)agg[i].second.n, agg[i].second.ms, agg[i].second.bytes);
    js += buf;
  }
  js += "]}";
  return js;
}

// ---------------------------------------------------------------------------
// single-mesh entry points on top of Plan

void subdivide_one(int scheme, const slak_mesh& cmesh, slak_mesh& rmesh, int target_level, slak_stats* stats,
                   hipStream_t st) {
  auto t0 = std::chrono::steady_clock::now();
  Pool& P = pool();
  P.reset_peak();
  MeshJob job;
  job.in = cmesh;
  job.target = target_level;
  std::vector<MeshJob> jobs{job};
  Plan plan(scheme, std::move(jobs), OutputMode::kSeparate, st);
  plan.prepare();
  MeshJob& J = plan.jobs()[0];
  if (J.levels > 0) {
    plan.launch(false);
    plan.finish();
    plan.release_outputs();
    const int L = J.levels;
    rmesh.type = SLAK_QUAD;
    rmesh.nfaces = (uint64_t)J.nf[L];
    rmesh.nnz = (uint64_t)J.nnz[L];
    rmesh.nverts = (uint64_t)J.nv[L];
    rmesh.ids = J.out_ids;
    rmesh.verts = reinterpret_cast<float*>(J.out_verts);
    if (scheme == kLAK) {  // LAKCatmullClarkImpl.cu:533-534,549-550,97
      rmesh.is_reduced = 0;
      rmesh.max_face_order = 4;
      rmesh.ptr = J.out_ptr;
      rmesh.vals = J.out_vals;
    } else {  // CatmullClarkImpl.cu:124-125,478,495
      rmesh.is_reduced = 1;
      rmesh.ptr = nullptr;
      rmesh.vals = nullptr;
    }
  }
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    stats->peak_bytes = P.peak();
    stats->levels = J.levels;
    stats->host_syncs = plan.syncs();
    if (J.levels > 0) {
      stats->out_faces = (uint64_t)J.nf[J.levels];
      stats->out_verts = (uint64_t)J.nv[J.levels];
    }
  }
}

void spgemm_operator(const slak_mesh& cmesh, int32_t** F_ptr, int32_t** F_ids, int32_t** F_vals, int32_t** D_ptr,
                     int32_t** D_ids, int32_t** D_vals, int32_t** E_ptr, int32_t** E_ids, int32_t** E_vals,
                     uint64_t* nedges, uint64_t* nextern, hipStream_t st) {
  if (!F_ptr || !F_ids || !F_vals || !D_ptr || !D_ids || !D_vals || !E_ptr || !E_ids || !E_vals)
    throw Error(SLAK_ERR_INVALID_ARGUMENT, "null output pointer");
  MeshJob job;
  job.in = cmesh;
  job.target = 1;
  if (cmesh.type == SLAK_QUAD) job.target = 1;
  Plan plan(kLAK, {job}, OutputMode::kPooled, st);
  plan.prepare();
  LevelPlan& L = plan.level(0);
  MeshJob& J = plan.jobs()[0];
  const int64_t nv = L.nv, nnz = L.nnz, ne = J.ne[0];
  Pool& P = pool();
  std::vector<void*> out;
  auto take = [&](int64_t n) {
    void* p = P.alloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    out.push_back(p);
    return static_cast<int32_t*>(p);
  };
  try {
    int32_t* fptr = take(nv + 1);
    int32_t* dptr = take(nv + 1);
    int32_t* eptr = take(nv + 1);
    SLAK_HIP_CHECK(hipMemcpyAsync(fptr, L.tptr, (nv + 1) * 4, hipMemcpyDeviceToDevice, st));
    SLAK_HIP_CHECK(hipMemcpyAsync(dptr, L.tptr, (nv + 1) * 4, hipMemcpyDeviceToDevice, st));
    SLAK_HIP_CHECK(hipMemcpyAsync(eptr, L.eptr, (nv + 1) * 4, hipMemcpyDeviceToDevice, st));
    LevelView V{};
    V.ids = L.ids;
    V.fptr = L.fptr;
    V.cface = L.cface;
    V.P = L.P;
    V.nv = (int32_t)nv;
    V.nf = (int32_t)L.nf;
    V.nnz = nnz;
    V.tptr = L.tptr;
    V.tids = L.tids;
    V.eptr = L.eptr;
    V.flags = plan.flags();
    V.seg = L.table();
    V.ne_cap = ne;
    V.F_ids = take(nnz);
    V.F_vals = take(nnz);
    V.D_ids = take(nnz);
    V.D_vals = take(nnz);
    V.E_ids = take(ne);
    V.E_vals = take(ne);
    launch_spgemm_fde(L.quad, V, st);
    plan.finish();
    *F_ptr = fptr;
    *F_ids = V.F_ids;
    *F_vals = V.F_vals;
    *D_ptr = dptr;
    *D_ids = V.D_ids;
    *D_vals = V.D_vals;
    *E_ptr = eptr;
    *E_ids = V.E_ids;
    *E_vals = V.E_vals;
    if (nedges) *nedges = (uint64_t)ne;
    if (nextern) *nextern = (uint64_t)(2 * ne - nnz);
  } catch (...) {
    (void)hipStreamSynchronize(st);
    for (void* p : out) P.release(p);
    throw;
  }
}

}  // namespace slak

// ---------------------------------------------------------------------------
// C-ABI

using slak::Error;

template <class Fn>
static int guarded(Fn&& fn);

namespace slak {
int guarded_call(const std::function<void()>& fn) { return guarded(fn); }
}  // namespace slak

template <class Fn>
static int guarded(Fn&& fn) {
  try {
    fn();
    slak::t_last_error.clear();
    return SLAK_OK;
  } catch (const Error& e) {
    slak::t_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    slak::t_last_error = "host out of memory";
    return SLAK_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    slak::t_last_error = e.what();
    return SLAK_ERR_INTERNAL;
  }
}

extern "C" {

int slak_subdivide(int32_t scheme, const slak_mesh* cmesh, slak_mesh* rmesh, int32_t target_level, slak_stats* stats,
                   void* stream) {
  return guarded([&] {
    if (!cmesh || !rmesh) throw Error(SLAK_ERR_INVALID_ARGUMENT, "null mesh");
    if (scheme != SLAK_SCHEME_LAK && scheme != SLAK_SCHEME_SLAK) throw Error(SLAK_ERR_INVALID_ARGUMENT, "unknown scheme");
    if (target_level < 0 || target_level > 12) throw Error(SLAK_ERR_INVALID_ARGUMENT, "target_level out of range");
    slak_mesh in = *cmesh;  // rmesh may alias cmesh
    slak::subdivide_one(scheme, in, *rmesh, target_level, stats, static_cast<hipStream_t>(stream));
  });
}

int slak_free(void* p) {
  return guarded([&] {
    if (!p) return;
    // Every library call returns with its own work complete, so a returned
    // buffer only needs the caller's work on it to be finished: the legacy
    // default stream and, through it, the blocking streams. Non-blocking
    // streams (a server's steps) are not waited for.
    SLAK_HIP_CHECK(hipStreamSynchronize(nullptr));
    if (!slak::pool().release(p)) SLAK_HIP_CHECK(hipFree(p));
  });
}

const char* slak_last_error(void) { return slak::t_last_error.c_str(); }

int32_t slak_version(void) { return 100; }

int slak_malloc(void** p, uint64_t bytes) {
  return guarded([&] {
    if (!p) throw Error(SLAK_ERR_INVALID_ARGUMENT, "null out pointer");
    *p = slak::pool().alloc(bytes);
  });
}

int slak_memcpy_h2d(void* dst, const void* src, uint64_t bytes) {
  return guarded([&] { SLAK_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}
int slak_memcpy_d2h(void* dst, const void* src, uint64_t bytes) {
  return guarded([&] { SLAK_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}
int slak_memcpy_d2d(void* dst, const void* src, uint64_t bytes) {
  return guarded([&] { SLAK_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice)); });
}
int slak_device_synchronize(void) {
  return guarded([&] { SLAK_HIP_CHECK(hipDeviceSynchronize()); });
}
int slak_pool_trim(void) {
  return guarded([&] { slak::pool().trim(); });
}
int slak_pool_stats(uint64_t* in_use, uint64_t* cached, uint64_t* peak) {
  return guarded([&] {
    slak::Pool& P = slak::pool();
    if (in_use) *in_use = P.in_use();
    if (cached) *cached = P.cached();
    if (peak) *peak = P.peak();
  });
}

int slak_spgemm_mapped(const slak_mesh* cmesh, int32_t** F_ptr, int32_t** F_ids, int32_t** F_vals, int32_t** D_ptr,
                       int32_t** D_ids, int32_t** D_vals, int32_t** E_ptr, int32_t** E_ids, int32_t** E_vals,
                       uint64_t* nedges, uint64_t* nextern, void* stream) {
  return guarded([&] {
    if (!cmesh) throw Error(SLAK_ERR_INVALID_ARGUMENT, "null mesh");
    slak::spgemm_operator(*cmesh, F_ptr, F_ids, F_vals, D_ptr, D_ids, D_vals, E_ptr, E_ids, E_vals, nedges, nextern,
                          static_cast<hipStream_t>(stream));
  });
}

int slak_profile_enable(int32_t on) {
  return guarded([&] {
    slak::Profiler::get().reset();
    slak::Profiler::get().set_enabled(on != 0);
  });
}

int slak_profile_report(char* buf, uint64_t len) {
  return guarded([&] {
    std::string r = slak::Profiler::get().report();
    if (!buf || len == 0) throw Error(SLAK_ERR_INVALID_ARGUMENT, "null buffer");
    if (r.size() + 1 > len) throw Error(SLAK_ERR_INVALID_ARGUMENT, "buffer too small: need " + std::to_string(r.size() + 1));
    std::memcpy(buf, r.c_str(), r.size() + 1);
  });
}

int slak_copy_probe(uint64_t bytes, int32_t iters, double* gbps) {
  return guarded([&] {
    if (!gbps || iters <= 0 || bytes < 16) throw Error(SLAK_ERR_INVALID_ARGUMENT, "bad argument");
    slak::Arena arena(slak::pool());
    const int64_t n = (int64_t)(bytes / 16);
    float4* a = arena.get_n<float4>(n);
    float4* b = arena.get_n<float4>(n);
    SLAK_HIP_CHECK(hipMemsetAsync(a, 0, n * 16, nullptr));
    slak::launch_copy_probe(a, b, n, nullptr);  // warm-up
    hipEvent_t e0, e1;
    SLAK_HIP_CHECK(hipEventCreate(&e0));
    SLAK_HIP_CHECK(hipEventCreate(&e1));
    SLAK_HIP_CHECK(hipEventRecord(e0, nullptr));
    for (int i = 0; i < iters; ++i) slak::launch_copy_probe(a, b, n, nullptr);
    SLAK_HIP_CHECK(hipEventRecord(e1, nullptr));
    SLAK_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    SLAK_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *gbps = 2.0 * 16.0 * (double)n * iters / (ms * 1e-3) / 1e9;
  });
}

int slak_digest(const void* p, uint64_t bytes, uint64_t* digest, void* stream) {
  return guarded([&] {
    if (!digest || (!p && bytes)) throw Error(SLAK_ERR_INVALID_ARGUMENT, "null pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    slak::Arena arena(slak::pool());
    uint64_t* d = arena.get_n<uint64_t>(1);
    slak::launch_digest(p, (int64_t)bytes, d, st);
    SLAK_HIP_CHECK(hipMemcpyAsync(digest, d, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    SLAK_HIP_CHECK(hipStreamSynchronize(st));
  });
}

// per-face reorder of (ids, vals) and the validity read-back; throws on an
// invalid mesh (ids and vals may already be reordered in place)
static void sort_and_check(bool reduce, int32_t* ptr, int32_t stride, slak_mesh* m, int32_t* vals, int* flags,
                           hipStream_t st) {
  slak::launch_sort_faces(reduce, ptr, stride, (int32_t)m->nfaces, (int64_t)m->nnz, m->ids, vals, flags, st);
  int fl = 0;
  SLAK_HIP_CHECK(hipMemcpyAsync(&fl, flags, sizeof(int), hipMemcpyDeviceToHost, st));
  SLAK_HIP_CHECK(hipStreamSynchronize(st));
  if (fl) throw Error(SLAK_ERR_INVALID_MESH, "invalid mesh: ptr does not describe faces of the mesh type's order");
}

static void convert_homogeneous(slak_mesh* m, bool reduce, hipStream_t st) {
  if (!m) throw Error(SLAK_ERR_INVALID_ARGUMENT, "null mesh");
  if (m->type == SLAK_POLY) throw Error(SLAK_ERR_INVALID_ARGUMENT, "mesh type not homogeneous");
  if (m->type != SLAK_TRI && m->type != SLAK_QUAD) throw Error(SLAK_ERR_INVALID_ARGUMENT, "unknown mesh type");
  if (reduce == (m->is_reduced != 0)) return;  // already in the requested form
  const int32_t stride = m->type == SLAK_QUAD ? 4 : 3;
  if (m->nnz > 0 && !m->ids) throw Error(SLAK_ERR_INVALID_ARGUMENT, "mesh without ids");
  if ((int64_t)m->nnz != (int64_t)stride * (int64_t)m->nfaces)
    throw Error(SLAK_ERR_INVALID_ARGUMENT, "homogeneous mesh with nnz != face order * nfaces");
  if (m->nnz > 0x7fffffffull) throw Error(SLAK_ERR_INVALID_ARGUMENT, "nnz too large for int32 indices");
  slak::Arena arena(slak::pool());
  int* flags = arena.get_n<int>(1);
  slak::launch_fill_i32(flags, 1, 0, st);
  if (reduce) {
    if (!m->vals) throw Error(SLAK_ERR_INVALID_ARGUMENT, "unreduced mesh without vals");
    sort_and_check(reduce, m->ptr, stride, m, m->vals, flags, st);
  } else {
    // ptr and vals of a reduced mesh may be absent: stride offsets and 1-based
    // corner positions (the reference's unreduce fills vals with (c - 1) mod
    // order, i.e. -1, 0, 1, ..., which no LAK kernel accepts as positions)
    // new buffers stay the library's until the conversion has succeeded
    int32_t* ptr = nullptr;
    int32_t* vals = nullptr;
    struct Owned {
      int32_t*& ptr;
      int32_t*& vals;
      hipStream_t st;
      ~Owned() {
        if (!ptr && !vals) return;
        (void)hipStreamSynchronize(st);
        if (ptr) slak::pool().release(ptr);
        if (vals) slak::pool().release(vals);
      }
    } owned{ptr, vals, st};
    if (!m->ptr) ptr = static_cast<int32_t*>(slak::pool().alloc(((size_t)m->nfaces + 1) * sizeof(int32_t)));
    if (!m->vals) vals = static_cast<int32_t*>(slak::pool().alloc(std::max<size_t>(m->nnz, 1) * sizeof(int32_t)));
    slak::launch_fill_homogeneous(ptr, (int64_t)m->nfaces + 1, vals, (int64_t)m->nnz, stride, st);
    int32_t* use_ptr = ptr ? ptr : m->ptr;
    int32_t* use_vals = vals ? vals : m->vals;
    sort_and_check(reduce, use_ptr, stride, m, use_vals, flags, st);
    if (ptr) m->ptr = ptr;
    if (vals) m->vals = vals;
    ptr = vals = nullptr;  // handed to the caller's mesh
  }
  m->is_reduced = reduce ? 1 : 0;
}

int slak_mesh_reduce(slak_mesh* mesh, void* stream) {
  return guarded([&] { convert_homogeneous(mesh, true, static_cast<hipStream_t>(stream)); });
}

int slak_mesh_unreduce(slak_mesh* mesh, void* stream) {
  return guarded([&] { convert_homogeneous(mesh, false, static_cast<hipStream_t>(stream)); });
}

int slak_scan_exclusive(const int32_t* in, int32_t* out, uint64_t n, void* stream) {
  return guarded([&] {
    if ((!in && n) || !out) throw Error(SLAK_ERR_INVALID_ARGUMENT, "null pointer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    slak::Arena arena(slak::pool());
    int* flags = arena.get_n<int>(1);
    slak::launch_fill_i32(flags, 1, 0, st);
    int64_t* partial = arena.get_n<int64_t>(slak::scan_partials_len((int64_t)n));
    slak::launch_scan(in, out, (int64_t)n, partial, flags, st);
    int fl = 0;
    SLAK_HIP_CHECK(hipMemcpyAsync(&fl, flags, sizeof(int), hipMemcpyDeviceToHost, st));
    SLAK_HIP_CHECK(hipStreamSynchronize(st));
    if (fl) throw Error(SLAK_ERR_INVALID_ARGUMENT, "scan total exceeds int32");
  });
}

int slak_scan_exclusive_noalloc(const int32_t* in, int32_t* out, uint64_t n, void* stream) {
  return guarded([&] {
    if ((!in && n) || !out) throw Error(SLAK_ERR_INVALID_ARGUMENT, "null pointer");
    const int64_t max_n = slak::noalloc_scan_max_items();
    if (n > (uint64_t)max_n)
      throw Error(SLAK_ERR_INVALID_ARGUMENT, "scan of " + std::to_string(n) + " items exceeds the scratch capacity of " +
                                                 std::to_string(max_n) + " items");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const slak::NoAllocScratch s = slak::noalloc_scratch();
    slak::launch_fill_i32(s.flags, 1, 0, st);  // the flags word starts at zero (a kernel, not a memset)
    slak::launch_scan(in, out, (int64_t)n, s.partial, s.flags, st);
    int fl = 0;
    SLAK_HIP_CHECK(hipMemcpyAsync(&fl, s.flags, sizeof(int), hipMemcpyDeviceToHost, st));
    SLAK_HIP_CHECK(hipStreamSynchronize(st));
    if (fl) throw Error(SLAK_ERR_INVALID_ARGUMENT, "scan total exceeds int32");
  });
}

}  // extern "C"
