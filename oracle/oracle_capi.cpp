// C entry points of the CPU oracle, loaded by the tests (ctypes) and by the
// CPU-baseline leg of the benchmark. TEST INFRASTRUCTURE ONLY: the shipped
// library never links this.
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <string>

#include "slak_oracle.hpp"

using namespace slak_oracle;

namespace {

thread_local std::string g_err;

struct Handle {
  int precision = 64;
  int levels = 0;
  Mesh<double> out64;
  Mesh<float> out32;
  std::vector<LevelDebug> dbg;
};

template <class Real>
Mesh<Real> make_mesh(int type, int64_t nf, int64_t nv, int64_t nnz, const int32_t* ptr, const int32_t* ids,
                     const int32_t* vals, const float* verts4) {
  Mesh<Real> m;
  m.type = type;
  m.nf = nf;
  m.nv = nv;
  m.ids.assign(ids, ids + nnz);
  if (ptr) {
    m.ptr.assign(ptr, ptr + nf + 1);
  } else {  // reduced homogeneous mesh: implicit stride
    int k = type == QUAD ? 4 : 3;
    m.ptr.resize(nf + 1);
    for (int64_t f = 0; f <= nf; ++f) m.ptr[f] = (int32_t)(k * f);
  }
  if (vals) m.vals.assign(vals, vals + nnz);
  m.P.resize(4 * nv);
  for (int64_t i = 0; i < 4 * nv; ++i) m.P[i] = (Real)verts4[i];
  return m;
}

const std::vector<int32_t>* find_array(const LevelDebug& d, const std::string& n) {
  static const std::map<std::string, int> names = {
      {"order", 0},  {"Mt_ptr", 1},  {"Mt_ids", 2},  {"Mt_vals", 3}, {"F_ptr", 4},   {"F_ids", 5},
      {"F_vals", 6}, {"D_ptr", 7},   {"D_ids", 8},   {"D_vals", 9},  {"E_ptr", 10},  {"E_ids", 11},
      {"E_vals", 12}, {"int0", 13},  {"int1", 14},   {"intid", 15},  {"f0", 16},     {"f1", 17},
      {"ext0", 18},  {"ext1", 19},   {"extid", 20},  {"S_ptr", 21},  {"S_ids", 22},  {"S_epids", 23}};
  auto it = names.find(n);
  if (it == names.end()) return nullptr;
  const std::vector<int32_t>* t[] = {&d.order, &d.Mt.ptr, &d.Mt.ids, &d.Mt.vals, &d.F.ptr, &d.F.ids,
                                     &d.F.vals, &d.D.ptr, &d.D.ids, &d.D.vals,   &d.E.ptr, &d.E.ids,
                                     &d.E.vals, &d.int0,  &d.int1,  &d.intid,    &d.f0,    &d.f1,
                                     &d.ext0,   &d.ext1,  &d.extid, &d.S_ptr,    &d.S_ids, &d.S_epids};
  return t[it->second];
}

}  // namespace

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

void* oracle_run(int scheme, int precision, int nthreads, int keep_debug, int type, int64_t nf, int64_t nv,
                 int64_t nnz, const int32_t* ptr, const int32_t* ids, const int32_t* vals, const float* verts4,
                 int target_level) {
  try {
    auto h = std::make_unique<Handle>();
    h->precision = precision;
    Options opt;
    opt.nthreads = nthreads;
    std::vector<LevelDebug>* dbg = keep_debug ? &h->dbg : nullptr;
    if (precision == 64) {
      auto m = make_mesh<double>(type, nf, nv, nnz, ptr, ids, vals, verts4);
      h->levels = subdivide<double>(scheme, m, h->out64, target_level, opt, dbg);
    } else {
      auto m = make_mesh<float>(type, nf, nv, nnz, ptr, ids, vals, verts4);
      h->levels = subdivide<float>(scheme, m, h->out32, target_level, opt, dbg);
    }
    return h.release();
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

int64_t oracle_info(void* hv, const char* key) {
  auto* h = static_cast<Handle*>(hv);
  std::string k(key);
  bool d = h->precision == 64;
  if (k == "levels") return h->levels;
  if (k == "nf") return d ? h->out64.nf : h->out32.nf;
  if (k == "nv") return d ? h->out64.nv : h->out32.nv;
  if (k == "nnz") return d ? h->out64.nnz() : h->out32.nnz();
  if (k == "has_ptr") return d ? !h->out64.ptr.empty() : !h->out32.ptr.empty();
  if (k == "has_vals") return d ? !h->out64.vals.empty() : !h->out32.vals.empty();
  if (k == "ndebug") return (int64_t)h->dbg.size();
  return -1;
}

// Output arrays: "ptr", "ids", "vals" (int32), "P" (double, 4 per vertex).
// Level intermediates (keep_debug): level >= 0 and a name from find_array,
// plus "nedges"/"nextern" through oracle_level_scalar.
int64_t oracle_array_len(void* hv, int level, const char* name) {
  auto* h = static_cast<Handle*>(hv);
  std::string n(name);
  bool d = h->precision == 64;
  if (level < 0) {
    if (n == "ptr") return d ? h->out64.ptr.size() : h->out32.ptr.size();
    if (n == "ids") return d ? h->out64.ids.size() : h->out32.ids.size();
    if (n == "vals") return d ? h->out64.vals.size() : h->out32.vals.size();
    if (n == "P") return d ? h->out64.P.size() : h->out32.P.size();
    return -1;
  }
  if (level >= (int)h->dbg.size()) return -1;
  auto* a = find_array(h->dbg[level], n);
  return a ? (int64_t)a->size() : -1;
}

int oracle_array_copy(void* hv, int level, const char* name, void* dst) {
  auto* h = static_cast<Handle*>(hv);
  std::string n(name);
  bool d = h->precision == 64;
  if (level < 0) {
    const std::vector<int32_t>* a = nullptr;
    if (n == "ptr") a = d ? &h->out64.ptr : &h->out32.ptr;
    if (n == "ids") a = d ? &h->out64.ids : &h->out32.ids;
    if (n == "vals") a = d ? &h->out64.vals : &h->out32.vals;
    if (a) {
      if (!a->empty()) std::memcpy(dst, a->data(), a->size() * sizeof(int32_t));
      return 0;
    }
    if (n == "P") {
      double* o = static_cast<double*>(dst);
      if (d)
        std::memcpy(o, h->out64.P.data(), h->out64.P.size() * sizeof(double));
      else
        for (size_t i = 0; i < h->out32.P.size(); ++i) o[i] = h->out32.P[i];
      return 0;
    }
    return -1;
  }
  if (level >= (int)h->dbg.size()) return -1;
  auto* a = find_array(h->dbg[level], n);
  if (!a) return -1;
  if (!a->empty()) std::memcpy(dst, a->data(), a->size() * sizeof(int32_t));
  return 0;
}

int64_t oracle_level_scalar(void* hv, int level, const char* name) {
  auto* h = static_cast<Handle*>(hv);
  if (level < 0 || level >= (int)h->dbg.size()) return -1;
  std::string n(name);
  if (n == "nedges") return h->dbg[level].nedges;
  if (n == "nextern") return h->dbg[level].nextern;
  return -1;
}

void oracle_release(void* hv) { delete static_cast<Handle*>(hv); }

// CPU baseline: the FP32 LAK algorithm on `nthreads` threads, `warmup` untimed
// runs then the median of `reps` timed runs (std::chrono::steady_clock) of the
// whole level chain. Returns milliseconds (negative on error); the final face
// count is written to *out_faces.
double oracle_time_ms(int scheme, int nthreads, int type, int64_t nf, int64_t nv, int64_t nnz, const int32_t* ptr,
                      const int32_t* ids, const int32_t* vals, const float* verts4, int target_level, int warmup,
                      int reps, int64_t* out_faces) {
  try {
    auto m = make_mesh<float>(type, nf, nv, nnz, ptr, ids, vals, verts4);
    Options opt;
    opt.nthreads = nthreads;
    std::vector<double> t;
    for (int r = 0; r < warmup + reps; ++r) {
      Mesh<float> out;
      auto t0 = std::chrono::steady_clock::now();
      subdivide<float>(scheme, m, out, target_level, opt, nullptr);
      auto t1 = std::chrono::steady_clock::now();
      if (out_faces) *out_faces = out.nf;
      if (r >= warmup) t.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
    }
    std::sort(t.begin(), t.end());
    return t.empty() ? 0.0 : t[t.size() / 2];
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1.0;
  }
}

// CPU baseline of a batch: the FP32 LAK algorithm on every mesh of a list,
// each at its own level, the meshes spread over `nthreads` OpenMP threads
// (dynamic schedule, one single-threaded chain per mesh). Median of `reps`
// timed passes over the whole list after `warmup` untimed ones; the total
// final face count goes to *out_faces.
double oracle_time_batch_ms(int scheme, int nthreads, int nmeshes, const int* types, const int64_t* nf,
                            const int64_t* nv, const int64_t* nnz, const int32_t* const* ptr,
                            const int32_t* const* ids, const int32_t* const* vals, const float* const* verts4,
                            const int* levels, int warmup, int reps, int64_t* out_faces) {
  try {
    std::vector<Mesh<float>> in(nmeshes);
    for (int i = 0; i < nmeshes; ++i)
      in[i] = make_mesh<float>(types[i], nf[i], nv[i], nnz[i], ptr[i], ids[i], vals ? vals[i] : nullptr, verts4[i]);
    Options opt;
    opt.nthreads = 1;
    std::vector<double> t;
    std::vector<int64_t> faces(nmeshes, 0);
    for (int r = 0; r < warmup + reps; ++r) {
      std::string err;
      auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
      for (int i = 0; i < nmeshes; ++i) {
        try {
          Mesh<float> out;
          subdivide<float>(scheme, in[i], out, levels[i], opt, nullptr);
          faces[i] = out.nf;
        } catch (const std::exception& e) {
#pragma omp critical
          err = e.what();
        }
      }
      auto t1 = std::chrono::steady_clock::now();
      if (!err.empty()) throw std::runtime_error(err);
      if (r >= warmup) t.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
    }
    if (out_faces) {
      int64_t s = 0;
      for (int64_t f : faces) s += f;
      *out_faces = s;
    }
    std::sort(t.begin(), t.end());
    return t.empty() ? 0.0 : t[t.size() / 2];
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1.0;
  }
}

int oracle_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
