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
#include <memory>

#include "slak_c.h"
#include "slak_internal.hpp"

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

struct Pinned {
  int64_t* p = nullptr;
  Pinned() {
    if (hipHostMalloc((void**)&p, 64 * sizeof(int64_t)) != hipSuccess) p = nullptr;
  }
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
};
int64_t* pinned_scratch() {
  thread_local Pinned pin;
  if (!pin.p) throw Error(SLAK_ERR_HIP, "hipHostMalloc of the readback buffer failed");
  return pin.p;
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
  void* p = pool_.alloc(bytes);
  blocks_.push_back(p);
  return p;
}

void Arena::keep(void* p) {
  auto it = std::find(blocks_.begin(), blocks_.end(), p);
  if (it != blocks_.end()) blocks_.erase(it);
}

void Arena::release_all() {
  for (void* p : blocks_) pool_.release(p);
  blocks_.clear();
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
// level pipeline

namespace {

std::string flags_message(int f) {
  std::string m;
  if (f & kFlagBadPtr) m += " ptr is not a valid CSR offset array (ptr[0]=0, non-decreasing, ptr[nf]=nnz);";
  if (f & kFlagBadOrder) m += " face with fewer than three vertices;";
  if (f & kFlagBadId) m += " vertex id out of range;";
  if (f & kFlagBadVals) m += " vals is not a permutation of 1..face order;";
  if (f & kFlagBadQuad) m += " QUAD mesh with a face that is not a quad at ptr = 4i;";
  if (f & kFlagDupHalfEdge) m += " non-manifold: a directed edge occurs twice (inconsistent orientation or >2 faces on an edge);";
  if (f & kFlagDegenerate) m += " degenerate face (repeated vertex along a face);";
  if (f & kFlagOverflow) m += " index overflow (counts exceed int32);";
  return m;
}

struct LevelIn {
  bool quad;
  int64_t nv, nf, nnz;
  const int32_t* ids;   // position ordered
  const int32_t* fptr;  // poly only
  const int32_t* cface; // poly only
  const float4* P;
};

struct LevelOut {
  int64_t nv, nf, nnz, ne;
  int32_t* ids;
  int32_t* ptr;
  int32_t* vals;
  float4* verts;
};

struct Chain {
  int scheme;
  hipStream_t st;
  Arena& arena;
  int* flags;           // device word
  int64_t* partial;     // scan partials
  int64_t partial_len;
  int syncs = 0;

  int32_t* scan(const int32_t* in, int64_t n) {
    int32_t* out = arena.get_n<int32_t>(n + 1);
    int64_t need = scan_partials_len(n);
    if (need > partial_len) {
      partial = arena.get_n<int64_t>(need);
      partial_len = need;
    }
    {
      ProfScope ps(st, "scan", 8.0 * n + 4);
      launch_scan(in, out, n, partial, flags, st);
    }
    return out;
  }

  // reads `count` int32 values (device) plus the flag word with one round trip
  void readback(const int32_t* const* src, int count, int64_t* dst, int& fl) {
    int64_t* pin = pinned_scratch();
    int32_t* p32 = reinterpret_cast<int32_t*>(pin);
    SLAK_HIP_CHECK(hipMemcpyAsync(p32, flags, sizeof(int), hipMemcpyDeviceToHost, st));
    for (int i = 0; i < count; ++i)
      SLAK_HIP_CHECK(hipMemcpyAsync(p32 + 1 + i, src[i], sizeof(int32_t), hipMemcpyDeviceToHost, st));
    SLAK_HIP_CHECK(hipStreamSynchronize(st));
    ++syncs;
    fl = p32[0];
    for (int i = 0; i < count; ++i) dst[i] = p32[1 + i];
  }
};

// Symbolic half of a level: M^T, owned-edge counts and their offsets.
struct Analysis {
  int32_t* tptr;
  int32_t* tids;
  int32_t* eptr;
};

Analysis analyse(Chain& C, const LevelIn& in, bool count_done, int32_t* cnt) {
  Analysis a;
  const double poly_b = in.quad ? 0.0 : 4.0 * (in.nf + 1) + 4.0 * in.nnz;  // fptr + cface
  if (!count_done) {
    {
      ProfScope ps(C.st, "fill", 4.0 * in.nv);
      launch_fill_i32(cnt, in.nv, 0, C.st);
    }
    ProfScope ps(C.st, in.quad ? "count_quad" : "count_poly", 4.0 * in.nnz + 8.0 * in.nv + (in.quad ? 0.0 : 4.0 * (in.nf + 1)));
    launch_faces(in.quad, false, true, in.ids, in.fptr, in.P, (int32_t)in.nf, (int32_t)in.nv, nullptr, cnt, C.st);
  }
  a.tptr = C.scan(cnt, in.nv);
  a.tids = C.arena.get_n<int32_t>(in.nnz);
  {
    ProfScope ps(C.st, in.quad ? "transpose_quad" : "transpose_poly", 8.0 * in.nnz + 12.0 * in.nv + (in.quad ? 0.0 : 4.0 * (in.nf + 1)));
    launch_transpose(in.quad, in.ids, in.fptr, (int32_t)in.nf, a.tptr, cnt, a.tids, C.st);
  }
  int32_t* ecnt = C.arena.get_n<int32_t>(in.nv);
  LevelView L{};
  L.ids = in.ids;
  L.fptr = in.fptr;
  L.cface = in.cface;
  L.P = in.P;
  L.nv = (int32_t)in.nv;
  L.nf = (int32_t)in.nf;
  L.nnz = in.nnz;
  L.tptr = a.tptr;
  L.tids = a.tids;
  L.ecnt = ecnt;
  L.flags = C.flags;
  {
    ProfScope ps(C.st, in.quad ? "spgemm_symbolic_quad" : "spgemm_symbolic_poly", 8.0 * in.nnz + 8.0 * in.nv + poly_b);
    launch_symbolic(in.quad, C.scheme, L, C.st);
  }
  a.eptr = C.scan(ecnt, in.nv);
  return a;
}

// Numeric half: face points, edge numbering + edge/vertex points, topology.
void finish(Chain& C, const LevelIn& in, const Analysis& a, LevelOut& out, bool fp_done, bool final_level,
            int32_t* E_ids = nullptr, int32_t* E_vals = nullptr) {
  const double poly_b = in.quad ? 0.0 : 4.0 * (in.nf + 1) + 4.0 * in.nnz;
  if (!fp_done) {
    ProfScope ps(C.st, in.quad ? "facepoints_quad" : "facepoints_poly", 4.0 * in.nnz + 16.0 * in.nv + 16.0 * in.nf + (in.quad ? 0.0 : 4.0 * (in.nf + 1)));
    launch_faces(in.quad, true, false, in.ids, in.fptr, in.P, (int32_t)in.nf, (int32_t)in.nv, out.verts, nullptr, C.st);
  }
  int32_t* eout = C.arena.get_n<int32_t>(in.nnz);
  // every corner's edge id is written by the numeric pass for a valid mesh;
  // zero-filling keeps refined ids in range even if one were not
  {
    ProfScope ps(C.st, "fill", 4.0 * in.nnz);
    launch_fill_i32(eout, in.nnz, 0, C.st);
  }
  LevelView L{};
  L.ids = in.ids;
  L.fptr = in.fptr;
  L.cface = in.cface;
  L.P = in.P;
  L.nv = (int32_t)in.nv;
  L.nf = (int32_t)in.nf;
  L.nnz = in.nnz;
  L.tptr = a.tptr;
  L.tids = a.tids;
  L.eptr = a.eptr;
  L.eout = eout;
  L.E_ids = E_ids;
  L.E_vals = E_vals;
  L.out = out.verts;
  L.flags = C.flags;
  L.ne_cap = out.ne;
  {
    // reads tptr, tids, ids, eptr, P, fp; writes vertex points, edge points, eout
    ProfScope ps(C.st, in.quad ? "spgemm_numeric_quad" : "spgemm_numeric_poly", 12.0 * in.nv + 8.0 * in.nnz + 16.0 * in.nv + 16.0 * in.nf +
                                             16.0 * in.nv + 16.0 * out.ne + 4.0 * in.nnz + poly_b);
    launch_numeric(in.quad, C.scheme, L, C.st);
  }
  bool lak_form = final_level && C.scheme == kLAK;
  {
    ProfScope ps(C.st, in.quad ? "topology_quad" : "topology_poly", 8.0 * in.nnz + 16.0 * in.nnz + (lak_form ? 20.0 * in.nnz : 0.0) +
                                       (in.quad ? 0.0 : 4.0 * (in.nf + 1)));
    launch_topology(in.quad, in.ids, in.fptr, eout, (int32_t)in.nv, (int32_t)in.nf, in.nnz, out.ids,
                    lak_form ? out.ptr : nullptr, lak_form ? out.vals : nullptr, C.st);
  }
}

void check_int32(int64_t x, const char* what) {
  if (x < 0 || x > 0x7fffffffll)
    throw Error(SLAK_ERR_INVALID_ARGUMENT, std::string("mesh too large for int32 indices: ") + what);
}

// Validates the caller's mesh, brings it into position order and returns the
// first level's input.
LevelIn first_level_input(Chain& C, const slak_mesh& m) {
  if (m.type != SLAK_TRI && m.type != SLAK_QUAD && m.type != SLAK_POLY)
    throw Error(SLAK_ERR_INVALID_ARGUMENT, "unknown mesh type");
  LevelIn in{};
  in.quad = m.type == SLAK_QUAD;
  in.nv = (int64_t)m.nverts;
  in.nf = (int64_t)m.nfaces;
  in.nnz = (int64_t)m.nnz;
  check_int32(in.nv, "nverts");
  check_int32(in.nf, "nfaces");
  check_int32(in.nnz, "nnz");
  check_int32(4 * in.nnz, "refined nnz (4 * nnz)");
  if (!m.ids || !m.verts) throw Error(SLAK_ERR_INVALID_ARGUMENT, "mesh ids/verts must not be null");
  int stride = 0;
  if (!m.ptr) {
    if (m.type == SLAK_POLY) throw Error(SLAK_ERR_INVALID_ARGUMENT, "POLY mesh without ptr");
    stride = m.type == SLAK_QUAD ? 4 : 3;
    if (in.nnz != stride * in.nf) throw Error(SLAK_ERR_INVALID_ARGUMENT, "nnz != face order * nfaces");
  }
  int32_t* ids_pos = C.arena.get_n<int32_t>(in.nnz);
  int32_t* cface = in.quad ? nullptr : C.arena.get_n<int32_t>(in.nnz);
  int32_t* fptr = nullptr;
  if (!in.quad) {
    if (m.ptr) {
      fptr = m.ptr;
    } else {
      // reduced homogeneous TRI mesh: materialise ptr = 3i
      fptr = C.arena.get_n<int32_t>(in.nf + 1);
      std::vector<int32_t> h(in.nf + 1);
      for (int64_t i = 0; i <= in.nf; ++i) h[i] = (int32_t)(3 * i);
      SLAK_HIP_CHECK(hipMemcpyAsync(fptr, h.data(), h.size() * 4, hipMemcpyHostToDevice, C.st));
      SLAK_HIP_CHECK(hipStreamSynchronize(C.st));
    }
  }
  launch_fill_i32(C.flags, 1, 0, C.st);
  // SLAK never reads vals (CatmullClarkImpl.cu); LAK orders corners by them.
  const int32_t* vals = C.scheme == kLAK ? m.vals : nullptr;
  {
    ProfScope ps(C.st, "validate", 4.0 * (in.nf + 1) + (vals ? 12.0 : 8.0) * in.nnz + (cface ? 4.0 * in.nnz : 0.0));
    launch_validate(m.ptr, m.ids, vals, stride, (int32_t)in.nf, (int32_t)in.nv, in.nnz, in.quad, ids_pos, cface,
                    C.flags, C.st);
  }
  int fl = 0;
  C.readback(nullptr, 0, nullptr, fl);
  if (fl) throw Error(SLAK_ERR_INVALID_MESH, "invalid mesh:" + flags_message(fl));
  in.ids = ids_pos;
  in.fptr = fptr;
  in.cface = cface;
  in.P = reinterpret_cast<const float4*>(m.verts);
  return in;
}

void alloc_out(Chain& C, LevelOut& o, bool final_level, bool lak_form) {
  o.verts = C.arena.get_n<float4>(o.nv);
  o.ids = C.arena.get_n<int32_t>(4 * o.nf);
  o.ptr = (final_level && lak_form) ? C.arena.get_n<int32_t>(o.nf + 1) : nullptr;
  o.vals = (final_level && lak_form) ? C.arena.get_n<int32_t>(4 * o.nf) : nullptr;
}

}  // namespace

// Runs the whole chain; returns the number of levels and fills rmesh.
int subdivide_chain(int scheme, const slak_mesh& cmesh, slak_mesh& rmesh, int target_level, slak_stats* stats,
                    hipStream_t st) {
  auto t0 = std::chrono::steady_clock::now();
  Pool& P = pool();
  P.reset_peak();
  Arena arena(P);
  int levels = cmesh.type != SLAK_QUAD ? std::max(1, target_level) : target_level;
  if (levels <= 0) {
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return 0;
  }
  Chain C{scheme, st, arena, arena.get_n<int>(1), nullptr, 0};
  LevelIn in = first_level_input(C, cmesh);
  const bool lak_form = scheme == kLAK;

  // level 1: analyse, then learn the edge count
  int32_t* cnt = arena.get_n<int32_t>(in.nv);
  Analysis a = analyse(C, in, false, cnt);
  int64_t ne = 0;
  int fl = 0;
  const int32_t* src[1] = {a.eptr + in.nv};
  C.readback(src, 1, &ne, fl);
  if (fl) throw Error(SLAK_ERR_INVALID_MESH, "invalid mesh:" + flags_message(fl));
  int64_t next = 2 * ne - in.nnz;  // boundary edges
  if (next < 0) throw Error(SLAK_ERR_INTERNAL, "inconsistent edge count");

  // closed-form sizes of all later levels
  std::vector<int64_t> lv_nv(levels + 1), lv_nf(levels + 1), lv_nnz(levels + 1), lv_ne(levels + 1);
  lv_nv[0] = in.nv;
  lv_nf[0] = in.nf;
  lv_nnz[0] = in.nnz;
  lv_ne[0] = ne;
  for (int l = 1; l <= levels; ++l) {
    lv_nv[l] = lv_nv[l - 1] + lv_nf[l - 1] + lv_ne[l - 1];
    lv_nf[l] = lv_nnz[l - 1];
    lv_nnz[l] = 4 * lv_nnz[l - 1];
    int64_t nx = next << l;
    lv_ne[l] = (lv_nnz[l] - nx) / 2 + nx;
    check_int32(lv_nv[l], "refined vertex count");
    check_int32(lv_nnz[l], "refined nnz");
  }

  LevelOut out{};
  for (int l = 0; l < levels; ++l) {
    bool final_level = l == levels - 1;
    out = LevelOut{};
    out.nv = lv_nv[l + 1];
    out.nf = lv_nf[l + 1];
    out.nnz = lv_nnz[l + 1];
    out.ne = lv_ne[l + 1];
    alloc_out(C, out, final_level, lak_form);
    if (l == 0) {
      finish(C, in, a, out, false, final_level);
    } else {
      // fused face points + valence count, then the symbolic half
      {
        ProfScope ps(C.st, "fill", 4.0 * in.nv);
        launch_fill_i32(cnt, in.nv, 0, C.st);
      }
      {
        ProfScope ps(C.st, "facepoints+count", 4.0 * in.nnz + 16.0 * in.nv + 16.0 * in.nf + 8.0 * in.nv);
        launch_faces(true, true, true, in.ids, nullptr, in.P, (int32_t)in.nf, (int32_t)in.nv, out.verts, cnt, C.st);
      }
      Analysis al = analyse(C, in, true, cnt);
      finish(C, in, al, out, true, final_level);
    }
    if (!final_level) {
      LevelIn nx{};
      nx.quad = true;
      nx.nv = out.nv;
      nx.nf = out.nf;
      nx.nnz = out.nnz;
      nx.ids = out.ids;
      nx.P = out.verts;
      in = nx;
      cnt = arena.get_n<int32_t>(in.nv);
    }
  }
  // final synchronization + flag check
  fl = 0;
  C.readback(nullptr, 0, nullptr, fl);
  if (fl) throw Error(SLAK_ERR_INTERNAL, "refinement produced invalid topology:" + flags_message(fl));

  arena.keep(out.verts);
  arena.keep(out.ids);
  if (out.ptr) arena.keep(out.ptr);
  if (out.vals) arena.keep(out.vals);
  rmesh.type = SLAK_QUAD;
  rmesh.nfaces = (uint64_t)out.nf;
  rmesh.nnz = (uint64_t)out.nnz;
  rmesh.nverts = (uint64_t)out.nv;
  rmesh.ids = out.ids;
  rmesh.verts = reinterpret_cast<float*>(out.verts);
  if (lak_form) {
    rmesh.is_reduced = 0;
    rmesh.max_face_order = 4;
    rmesh.ptr = out.ptr;
    rmesh.vals = out.vals;
  } else {
    rmesh.is_reduced = 1;
    rmesh.ptr = nullptr;
    rmesh.vals = nullptr;
  }
  if (stats) {
    stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    stats->peak_bytes = P.peak();
    stats->levels = levels;
    stats->host_syncs = C.syncs;
    stats->out_faces = (uint64_t)out.nf;
    stats->out_verts = (uint64_t)out.nv;
  }
  return levels;
}

void spgemm_operator(const slak_mesh& cmesh, int32_t** F_ptr, int32_t** F_ids, int32_t** F_vals, int32_t** D_ptr,
                     int32_t** D_ids, int32_t** D_vals, int32_t** E_ptr, int32_t** E_ids, int32_t** E_vals,
                     uint64_t* nedges, uint64_t* nextern, hipStream_t st) {
  if (!F_ptr || !F_ids || !F_vals || !D_ptr || !D_ids || !D_vals || !E_ptr || !E_ids || !E_vals)
    throw Error(SLAK_ERR_INVALID_ARGUMENT, "null output pointer");
  Pool& P = pool();
  Arena arena(P);
  Chain C{kLAK, st, arena, arena.get_n<int>(1), nullptr, 0};
  LevelIn in = first_level_input(C, cmesh);
  int32_t* cnt = arena.get_n<int32_t>(in.nv);
  Analysis a = analyse(C, in, false, cnt);
  int64_t ne = 0;
  int fl = 0;
  const int32_t* src[1] = {a.eptr + in.nv};
  C.readback(src, 1, &ne, fl);
  if (fl) throw Error(SLAK_ERR_INVALID_MESH, "invalid mesh:" + flags_message(fl));
  int32_t* dptr = arena.get_n<int32_t>(in.nv + 1);
  SLAK_HIP_CHECK(hipMemcpyAsync(dptr, a.tptr, (in.nv + 1) * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
  LevelView L{};
  L.ids = in.ids;
  L.fptr = in.fptr;
  L.cface = in.cface;
  L.P = in.P;
  L.nv = (int32_t)in.nv;
  L.nf = (int32_t)in.nf;
  L.nnz = in.nnz;
  L.tptr = a.tptr;
  L.tids = a.tids;
  L.eptr = a.eptr;
  L.flags = C.flags;
  L.ne_cap = ne;
  L.F_ids = arena.get_n<int32_t>(in.nnz);
  L.F_vals = arena.get_n<int32_t>(in.nnz);
  L.D_ids = arena.get_n<int32_t>(in.nnz);
  L.D_vals = arena.get_n<int32_t>(in.nnz);
  L.E_ids = arena.get_n<int32_t>(ne);
  L.E_vals = arena.get_n<int32_t>(ne);
  launch_spgemm_fde(in.quad, L, st);
  C.readback(nullptr, 0, nullptr, fl);
  if (fl) throw Error(SLAK_ERR_INTERNAL, "spgemm failed:" + flags_message(fl));
  int32_t* keep[] = {a.tptr, L.F_ids, L.F_vals, dptr, L.D_ids, L.D_vals, a.eptr, L.E_ids, L.E_vals};
  for (int32_t* k : keep) arena.keep(k);
  *F_ptr = a.tptr;
  *F_ids = L.F_ids;
  *F_vals = L.F_vals;
  *D_ptr = dptr;
  *D_ids = L.D_ids;
  *D_vals = L.D_vals;
  *E_ptr = a.eptr;
  *E_ids = L.E_ids;
  *E_vals = L.E_vals;
  if (nedges) *nedges = (uint64_t)ne;
  if (nextern) *nextern = (uint64_t)(2 * ne - in.nnz);
}

}  // namespace slak

// ---------------------------------------------------------------------------
// C-ABI

using slak::Error;

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
    slak::subdivide_chain(scheme, in, *rmesh, target_level, stats, static_cast<hipStream_t>(stream));
  });
}

int slak_free(void* p) {
  return guarded([&] {
    if (!p) return;
    SLAK_HIP_CHECK(hipDeviceSynchronize());
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

}  // extern "C"
