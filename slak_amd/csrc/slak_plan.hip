// Level plans over a union of meshes (see slak_plan.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "slak_plan.hpp"

namespace slak {

std::string flags_message(int f) {
  std::string m;
  if (f & kFlagBadPtr) m += " ptr is not a valid CSR offset array (ptr[0]=0, non-decreasing, ptr[nf]=nnz);";
  if (f & kFlagBadOrder) m += " face with fewer than three vertices;";
  if (f & kFlagBadId) m += " vertex id out of range;";
  if (f & kFlagBadVals) m += " vals is not a permutation of 1..face order;";
  if (f & kFlagBadQuad) m += " QUAD mesh with a face that is not a quad at ptr = 4i;";
  if (f & kFlagDupHalfEdge)
    m += " non-manifold: a directed edge occurs twice (inconsistent orientation or >2 faces on an edge);";
  if (f & kFlagDegenerate) m += " degenerate face (repeated vertex along a face);";
  if (f & kFlagOverflow) m += " index overflow (counts exceed int32);";
  return m;
}

namespace {

int64_t* pinned_buffer(size_t n) {
  struct Pinned {
    int64_t* p = nullptr;
    size_t n = 0;
    ~Pinned() {
      if (p) (void)hipHostFree(p);
    }
  };
  thread_local Pinned pin;
  if (pin.n < n) {
    if (pin.p) (void)hipHostFree(pin.p);
    pin.p = nullptr;
    pin.n = 0;
    SLAK_HIP_CHECK(hipHostMalloc((void**)&pin.p, n * sizeof(int64_t)));
    pin.n = n;
  }
  return pin.p;
}

// Pinned host blocks for table uploads, by power-of-two size class (4 KiB
// and up): a pageable source makes hipMemcpyAsync stage the bytes itself
// before it returns, which delays every launch queued after it.
std::mutex g_pin_mu;
std::vector<std::pair<size_t, char*>> g_pin_free;

char* pinned_acquire(size_t bytes, size_t& cap) {
  cap = 4096;
  while (cap < bytes) cap <<= 1;
  {
    std::lock_guard<std::mutex> g(g_pin_mu);
    for (size_t i = 0; i < g_pin_free.size(); ++i)
      if (g_pin_free[i].first == cap) {
        char* p = g_pin_free[i].second;
        g_pin_free[i] = g_pin_free.back();
        g_pin_free.pop_back();
        return p;
      }
  }
  char* p = nullptr;
  SLAK_HIP_CHECK(hipHostMalloc((void**)&p, cap));
  return p;
}

void pinned_release(char* p, size_t cap) {
  if (!p) return;
  std::lock_guard<std::mutex> g(g_pin_mu);
  if (g_pin_free.size() < 64) {
    g_pin_free.emplace_back(cap, p);
    return;
  }
  (void)hipHostFree(p);
}

void check_i32(int64_t x, const std::string& what) {
  if (x < 0 || x > 0x7fffffffll)
    throw Error(SLAK_ERR_INVALID_ARGUMENT, "too large for int32 indices: " + what + " = " + std::to_string(x));
}

}  // namespace

// Register capacity of the SpGEMM row kernels for a level with the given
// row-length statistics. Rows above the capacity take the in-memory path
// (correct for any length, slower); capacity 4 keeps the row state in half
// the registers, which doubles occupancy. Pick 4 when every row fits or the
// long rows are rare enough that the waves they stall cost less than the
// occupancy gained; the threshold is a tuned constant (SLAK_CAP_THRESHOLD
// overrides it, SLAK_ROW_CAP forces a capacity).
int autotune_cap(bool quad, int64_t nv, int64_t long_rows, int max_row) {
  const char* ef = std::getenv("SLAK_ROW_CAP");
  const char* et = std::getenv("SLAK_CAP_THRESHOLD");
  const int forced = ef ? std::atoi(ef) : 0;
  const double thresh = et ? std::atof(et) : 0.02;
  if (forced == 4 || forced == 8) return forced;
  if (max_row <= 4) return 4;
  if (nv <= 0) return 8;
  double frac = (double)long_rows / (double)nv;
  (void)quad;
  return frac < thresh ? 4 : 8;
}

// Rows per block and LDS row tile of the first level's row kernels
// (k_symbolic, the numbering of k_number_faces). A block of R rows needs
// about R m + 3 sqrt(R) sigma tile entries, with m the mean row length
// (corners over vertices) and sigma the spread of row lengths from the
// histogram (hist: kRowHistBins counts, the last bin "longer"; without one,
// sigma = 1): R is the largest of 256, 128, 64 whose need fits kRowTileMax,
// and the tile that need rounded up to 256 entries (at least 1024). A block
// whose rows exceed the tile reads them from global memory. With rows
// averaging more than 5 the numbering takes half as many rows per block (at
// least 64). Medians of interleaved runs on a 1024 x 1024 triangle torus
// (rows of 6): numbering 89.5 -> 72.5 us, while the sizing pass is fastest at
// R = 256 (62.8 against 74.7 us at 128). With rows of about 4 the numbering
// gains little or loses at R/2 (quad 1024 x 1024 torus 38.6 -> 35.0 us, the
// 256 x 128 mixed torus 11.8 -> 13.5 us, the 4096-mesh batch 385 -> 450 us).
// SLAK_ROWS_PER_BLOCK forces the sizing pass's R, SLAK_ROWS_NUMBER the
// numbering's.
RowTiling autotune_rows(int64_t nv, int64_t nnz, const int64_t* hist) {
  double m = nv > 0 ? (double)nnz / (double)nv : 4.0;
  double sigma = 1.0;
  if (hist) {
    double n = 0, sq = 0;
    for (int b = 0; b < kRowHistBins; ++b) {
      n += (double)hist[b];
      sq += (double)hist[b] * (b - m) * (b - m);
    }
    if (n > 0) sigma = std::sqrt(sq / n);
  }
  auto need = [&](int R) { return R * m + 3.0 * std::sqrt((double)R) * sigma; };
  int R = 64;
  for (int r : {256, 128})
    if (need(r) <= (double)kRowTileMax) {
      R = r;
      break;
    }
  if (const char* e = std::getenv("SLAK_ROWS_PER_BLOCK")) {
    const int forced = std::atoi(e);
    if (forced == 64 || forced == 128 || forced == 256) R = forced;
  }
  const double want = std::ceil(need(R) / 256.0) * 256.0;
  const int tile = (int)std::min<double>(kRowTileMax, std::max(1024.0, want));
  int Rn = m > 5.0 ? std::max(64, R / 2) : R;
  if (const char* e = std::getenv("SLAK_ROWS_NUMBER")) {
    const int forced = std::atoi(e);
    if (forced == 64 || forced == 128 || forced == 256) Rn = forced;
  }
  return RowTiling{R, Rn, tile};
}

Plan::Plan(int scheme, std::vector<MeshJob> jobs, OutputMode mode, hipStream_t st)
    : scheme_(scheme), jobs_(std::move(jobs)), mode_(mode), st_(st), arena_(pool()) {
  init_device_options();
}

Plan::~Plan() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  if (owned_st_) {
    (void)hipStreamSynchronize(owned_st_);
    (void)hipStreamDestroy(owned_st_);
    (void)hipEventDestroy(order_ev_);
  }
  bool held = false;
  for (auto& j : jobs_) held |= j.out_ids != nullptr;
  if (!outputs_released_ && held) {
    (void)hipStreamSynchronize(st_);
    if (mode_ == OutputMode::kSeparate)
      for (auto& j : jobs_)
        for (void* p : {(void*)j.out_verts, (void*)j.out_ids, (void*)j.out_ptr, (void*)j.out_vals})
          if (p) pool().release(p);
  }
  if (!arena_.empty() || staging_first_.p || staging_all_.p) (void)hipStreamSynchronize(st_);
  pinned_release(staging_first_.p, staging_first_.cap);
  pinned_release(staging_all_.p, staging_all_.cap);
}

// Reads the flags word and the count ints after it in the state block in one
// copy (count = 0: flags only) and waits for it.
int Plan::readback(const int32_t* src, int count, int64_t* dst, const std::function<void()>& overlap) {
  (void)src;
  int64_t* pin = pinned_buffer((size_t)count / 2 + 2);
  int32_t* p32 = reinterpret_cast<int32_t*>(pin);
  SLAK_HIP_CHECK(hipMemcpyAsync(p32, d_flags_, (1 + count) * sizeof(int32_t), hipMemcpyDeviceToHost, st_));
  if (overlap) {
    struct Ev {
      hipEvent_t e = nullptr;
      ~Ev() {
        if (e) (void)hipEventDestroy(e);
      }
    };
    thread_local Ev ev;
    if (!ev.e) SLAK_HIP_CHECK(hipEventCreateWithFlags(&ev.e, hipEventDisableTiming));
    SLAK_HIP_CHECK(hipEventRecord(ev.e, st_));
    overlap();
    SLAK_HIP_CHECK(hipEventSynchronize(ev.e));
  } else {
    SLAK_HIP_CHECK(hipStreamSynchronize(st_));
  }
  ++syncs_;
  for (int i = 0; i < count; ++i) dst[i] = p32[1 + i];
  return p32[0];
}

// Uploads the segment tables (SegDesc, vertex, face and corner bases) of levels
// [l0, l1) in one copy from a plan-owned staging buffer (one per call site,
// so a pending copy never sees its source rewritten).
void Plan::upload_tables(int l0, int l1, Staging& staging, int64_t state_ints) {
  auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t total = 0;
  std::vector<size_t> offs;
  // k_final tables of several meshes: FinalSeg[n], corner and vertex item bases
  auto table_bytes = [&](const FinalTable& T) -> size_t {
    return final_table(T) ? align(T.segs.size() * sizeof(FinalSeg)) + align(T.units.size() * sizeof(int4)) : 0;
  };
  auto fin_bytes = [&](const LevelPlan& L) -> size_t { return table_bytes(L.fin) + table_bytes(L.pos); };
  for (int l = l0; l < l1; ++l) {
    const size_t n = levels_[l].segs.size();
    offs.push_back(total);
    total = align(total + n * sizeof(SegDesc)) + align(3 * (n + 1) * sizeof(int64_t)) + fin_bytes(levels_[l]);
  }
  const size_t state_off = total;
  total += align((size_t)state_ints * sizeof(int32_t));
  if (total == 0) return;
  if (staging.p) SLAK_HIP_CHECK(hipStreamSynchronize(st_));  // reused: an earlier copy may still read it
  if (staging.cap < total) {
    pinned_release(staging.p, staging.cap);
    staging.p = pinned_acquire(total, staging.cap);
  }
  std::memset(staging.p, 0, total);
  char* d = arena_.get_n<char>((int64_t)total);
  if (state_ints > 0) {  // the plan's state block, zeros
    d_state_ = reinterpret_cast<int32_t*>(d + state_off);
    d_flags_ = d_state_;
  }
  for (int l = l0; l < l1; ++l) {
    LevelPlan& L = levels_[l];
    const size_t n = L.segs.size();
    size_t o = offs[l - l0];
    std::memcpy(staging.p + o, L.segs.data(), n * sizeof(SegDesc));
    L.d_segs = reinterpret_cast<SegDesc*>(d + o);
    o = align(o + n * sizeof(SegDesc));
    std::memcpy(staging.p + o, L.vbase.data(), (n + 1) * sizeof(int64_t));
    L.d_vbase = reinterpret_cast<int64_t*>(d + o);
    std::memcpy(staging.p + o + (n + 1) * sizeof(int64_t), L.fbase.data(), (n + 1) * sizeof(int64_t));
    L.d_fbase = reinterpret_cast<int64_t*>(d + o) + (n + 1);
    std::memcpy(staging.p + o + 2 * (n + 1) * sizeof(int64_t), L.cbase.data(), (n + 1) * sizeof(int64_t));
    L.d_cbase = reinterpret_cast<int64_t*>(d + o) + 2 * (n + 1);
    if (!L.d_mflags) {
      if (l == 0 && d_state_)  // the first level's mesh flags live in the state block
        L.d_mflags = d_state_ + kStateBounds + (n + 1);
      else
        L.d_mflags = arena_.get_n<int>((int64_t)std::max<size_t>(n, 1));
    }
    o = align(o + 3 * (n + 1) * sizeof(int64_t));
    for (FinalTable* T : {&L.fin, &L.pos}) {
      if (!final_table(*T)) continue;
      std::memcpy(staging.p + o, T->segs.data(), T->segs.size() * sizeof(FinalSeg));
      T->d_segs = reinterpret_cast<FinalSeg*>(d + o);
      o = align(o + T->segs.size() * sizeof(FinalSeg));
      std::memcpy(staging.p + o, T->units.data(), T->units.size() * sizeof(int4));
      T->d_units = reinterpret_cast<int4*>(d + o);
      o = align(o + T->units.size() * sizeof(int4));
    }
  }
  SLAK_HIP_CHECK(hipMemcpyAsync(d, staging.p, total, hipMemcpyHostToDevice, st_));
}

std::string Plan::describe_failure(int fl) const {
  std::string m;
  for (size_t j = 0; j < jobs_.size(); ++j)
    if (jobs_[j].bad_flags) {
      m = "mesh " + std::to_string(j) + ":" + flags_message(jobs_[j].bad_flags);
      break;
    }
  if (m.empty()) m = flags_message(fl);
  return "invalid mesh: " + m;
}

// First level: per-job bases, the caller meshes gathered into the union.
void Plan::plan_first_level() {
  levels_.clear();
  levels_.emplace_back();
  LevelPlan& L = levels_[0];
  bool all_quad = true;
  for (size_t j = 0; j < jobs_.size(); ++j) {
    MeshJob& J = jobs_[j];
    if (J.levels <= 0) continue;
    all_quad &= J.in.type == SLAK_QUAD;
  }
  L.quad = all_quad;
  int64_t vb = 0, fb = 0, cb = 0;
  for (size_t j = 0; j < jobs_.size(); ++j) {
    MeshJob& J = jobs_[j];
    if (J.levels <= 0) continue;
    SegDesc S{};
    S.vb = vb;
    S.fb = fb;
    S.cb = cb;
    S.nv = (int32_t)J.in.nverts;
    S.nf = (int32_t)J.in.nfaces;
    S.nnz = (int64_t)J.in.nnz;
    S.src_ptr = J.in.ptr;
    S.src_ids = J.in.ids;
    S.src_vals = J.in.vals;
    S.src_verts = reinterpret_cast<const float4*>(J.in.verts);
    S.src_stride = J.in.ptr ? 0 : (J.in.type == SLAK_QUAD ? 4 : 3);
    S.require_quad = J.in.type == SLAK_QUAD;
    L.jobs.push_back((int)j);
    L.segs.push_back(S);
    L.vbase.push_back(vb);
    L.fbase.push_back(fb);
    L.cbase.push_back(cb);
    vb += S.nv;
    fb += S.nf;
    cb += S.nnz;
  }
  L.vbase.push_back(vb);
  L.fbase.push_back(fb);
  L.cbase.push_back(cb);
  L.nv = vb;
  L.nf = fb;
  L.nnz = cb;
  check_i32(L.nv, "vertices of the first level");
  check_i32(L.nf, "faces of the first level");
  check_i32(L.nnz, "corners of the first level");
  check_i32(4 * L.nnz, "refined corners of the first level");
  const RowTiling rt = autotune_rows(L.nv, L.nnz, nullptr);
  L.rows = rt.rows;
  L.rows_num = rt.rows_num;
  L.tile = rt.tile;

  // one arena block for the first level's buffers and the analysis state
  // that follows (the table upload, the scan partials, a single mesh's
  // boundary flags); anything beyond it comes from the pool as before
  {
    auto line = [](int64_t n, size_t elem) { return ((size_t)std::max<int64_t>(n, 1) * elem + 255) & ~(size_t)255; };
    size_t est = line(L.nnz, 4) + line(L.nv, 4) + line(L.nv + 1, 4) + line(L.nnz, 4) + 2 * line(L.nnz, 8) +
                 line(L.nv, 4) + line(L.nv + 1, 4) + line(L.nnz, 4);
    if (!L.quad) est += line(L.nf + 1, 4) + line(L.nnz, 4);
    if (L.segs.size() != 1) est += line(L.nv, 16);
    if (scheme_ == kSLAK) est += line(L.nnz, 4);
    est += (size_t)256 << 10;
    arena_.reserve(est);
  }
  int32_t* ids_pos = arena_.get_n<int32_t>(L.nnz);
  L.ids = ids_pos;
  if (!L.quad) {
    L.fptr = arena_.get_n<int32_t>(L.nf + 1);
    L.cface = arena_.get_n<int32_t>(L.nnz);
  }
  if (L.segs.size() == 1) {
    L.P = L.segs[0].src_verts;
  } else {
    P1_ = arena_.get_n<float4>(L.nv);
    L.P = P1_;
  }
  d_cnt_ = arena_.get_n<int32_t>(L.nv);
  L.cnt = d_cnt_;
  L.tptr = arena_.get_n<int32_t>(L.nv + 1);
  L.tids = arena_.get_n<int32_t>(L.nnz);
  L.trow = arena_.get_n<int2>(L.nnz);
  L.ci = arena_.get_n<int2>(L.nnz);
  if (scheme_ == kSLAK) L.mpos = arena_.get_n<int32_t>(L.nnz);
  L.ecnt = arena_.get_n<int32_t>(L.nv);
  L.eptr = arena_.get_n<int32_t>(L.nv + 1);
}

// First-level analysis: M^T by counting sort, sorted rows with in-neighbours
// and owned-edge counts (later levels receive their structure from the level
// before).
void Plan::record_analysis(LevelPlan& L) {
  const double poly_b = L.quad ? 0.0 : 4.0 * (L.nf + 1) + 4.0 * L.nnz;
  {
    ProfScope ps(st_, L.quad ? "count_quad" : "count_poly", 4.0 * L.nnz + 8.0 * L.nv + (L.quad ? 0.0 : 4.0 * L.nf));
    launch_faces(L.quad, false, true, L.ids, L.fptr, L.P, (int32_t)L.nf, L.table(), L.cnt, d_flags_, st_);
  }
  {
    ProfScope ps(st_, "scan", 8.0 * L.nv + 4);
    launch_scan(L.cnt, L.tptr, L.nv, d_partial_, d_flags_, st_);
  }
  {
    ProfScope ps(st_, L.quad ? "transpose_quad" : "transpose_poly",
                 8.0 * L.nnz + 12.0 * L.nv + (L.quad ? 0.0 : 4.0 * L.nf));
    launch_transpose(L.quad, L.ids, L.fptr, (int32_t)L.nf, L.tptr, L.cnt, L.tids, d_flags_, st_);
  }
  LevelView V{};
  V.ids = L.ids;
  V.fptr = L.fptr;
  V.cface = L.cface;
  V.P = L.P;
  V.nv = (int32_t)L.nv;
  V.nf = (int32_t)L.nf;
  V.nnz = L.nnz;
  V.tptr = L.tptr;
  V.tids = L.tids;
  V.trow = L.trow;
  V.mpos = L.mpos;
  V.ecnt = L.ecnt;
  V.hist_rows = d_state_ + kStateRowHist;
  V.rows_per_block = L.rows;
  V.tile_entries = L.tile;
  V.flags = d_flags_;
  V.seg = L.table();
  {
    ProfScope ps(st_, L.quad ? "spgemm_symbolic_quad" : "spgemm_symbolic_poly",
                 8.0 * L.nnz + 8.0 * L.nv + 8.0 * L.nnz + poly_b);
    launch_symbolic(L.quad, scheme_, L.cap, V, st_);
  }
  {
    // a single mesh's edge total goes straight to its read-back slot
    int32_t* bounds = d_state_ + kStateBounds;
    ProfScope ps(st_, "scan", 8.0 * L.nv + 4);
    launch_scan(L.ecnt, L.eptr, L.nv, d_partial_, d_flags_, st_, L.segs.size() == 1 ? bounds + 1 : nullptr);
  }
}

void Plan::prepare() {
  for (auto& J : jobs_) {
    if (J.in.type != SLAK_TRI && J.in.type != SLAK_QUAD && J.in.type != SLAK_POLY)
      throw Error(SLAK_ERR_INVALID_ARGUMENT, "unknown mesh type");
    if (J.target < 0 || J.target > 12) throw Error(SLAK_ERR_INVALID_ARGUMENT, "target level out of range");
    J.levels = J.in.type != SLAK_QUAD ? std::max(1, J.target) : J.target;
    if (J.levels <= 0) continue;
    if (!J.in.ids || !J.in.verts) throw Error(SLAK_ERR_INVALID_ARGUMENT, "mesh ids/verts must not be null");
    if (!J.in.ptr) {
      if (J.in.type == SLAK_POLY) throw Error(SLAK_ERR_INVALID_ARGUMENT, "POLY mesh without ptr");
      int k = J.in.type == SLAK_QUAD ? 4 : 3;
      if ((int64_t)J.in.nnz != k * (int64_t)J.in.nfaces)
        throw Error(SLAK_ERR_INVALID_ARGUMENT, "nnz != face order * nfaces");
    }
    check_i32((int64_t)J.in.nverts, "nverts");
    check_i32((int64_t)J.in.nfaces, "nfaces");
    check_i32((int64_t)J.in.nnz, "nnz");
  }
  bool any = false;
  for (auto& J : jobs_) any |= J.levels > 0;
  if (!any) {
    d_flags_ = arena_.get_n<int>(1);
    launch_fill_i32(d_flags_, 1, 0, st_);
    prepared_ = true;
    return;
  }
  plan_first_level();
  LevelPlan& L1 = levels_[0];
  const int nseg = (int)L1.segs.size();
  // a single mesh numbers its edges before the read-back: the first upload
  // gives it nnz >= ne as the edge-id bound and a flag buffer of that size
  const bool early = nseg == 1 && jobs_.size() == 1 && jobs_[L1.jobs[0]].known_ne < 0;
  early_p1_ = nullptr;
  early_faces_ = false;
  if (early) {
    L1.segs[0].ne = L1.segs[0].nnz;
    L1.bnd = arena_.get_n<int32_t>(L1.nnz);
    // when the first level feeds another, its positions go to a plan buffer
    // (sized for nnz >= ne edge points) and the early numbering also forms
    // the face points
    const MeshJob& J = jobs_[L1.jobs[0]];
    if (J.levels >= 2) {
      early_p1_cap_ = (int64_t)J.in.nverts + (int64_t)J.in.nfaces + (int64_t)J.in.nnz + 8;  // + line padding
      early_p1_ = arena_.get_n<float4>(early_p1_cap_);
      L1.segs[0].dst_verts = early_p1_;
    }
  }
  n_state_ = kStateBounds + (nseg + 1) + nseg;
  upload_tables(0, 1, staging_first_, n_state_);
  int64_t need = scan_partials_len(L1.nv);
  d_partial_ = arena_.get_n<int64_t>(need);
  partial_len_ = need;

  // 1. validate + gather (the state block arrived zeroed with the tables)
  {
    ProfScope ps(st_, "gather", 4.0 * (L1.nf + 1) + 12.0 * L1.nnz + (L1.quad ? 0.0 : 8.0 * L1.nnz + 4.0 * L1.nf));
    launch_gather(L1.table(), (int32_t)L1.nf, scheme_ == kLAK, const_cast<int32_t*>(L1.ids),
                  const_cast<int32_t*>(L1.fptr), const_cast<int32_t*>(L1.cface), d_flags_, d_cnt_, (int32_t)L1.nv,
                  d_state_ + kStateFaceHist, st_);
  }
  if (P1_) launch_gather_verts(L1.table(), (int32_t)L1.nv, P1_, st_);
  auto fail_if = [&](int fl) {
    if (!fl) return;
    std::vector<int> mf(nseg, 0);
    SLAK_HIP_CHECK(hipMemcpy(mf.data(), L1.d_mflags, nseg * sizeof(int), hipMemcpyDeviceToHost));
    for (int k = 0; k < nseg; ++k) jobs_[L1.jobs[k]].bad_flags = mf[k];
    throw Error(SLAK_ERR_INVALID_MESH, describe_failure(fl));
  };
  bool all_trusted = true, all_known = true;
  for (int k = 0; k < nseg; ++k) {
    all_trusted &= jobs_[L1.jobs[k]].trusted;
    all_known &= jobs_[L1.jobs[k]].known_ne >= 0;
  }
  // 2. first-level analysis: M^T, owned-edge counts, per-mesh edge totals
  // (its kernels do nothing on an input the gather flagged; the flags come
  // back with the edge counts in one round trip)
  record_analysis(L1);
  int32_t* d_bounds = d_state_ + kStateBounds;
  if (nseg > 1) launch_seg_bounds(L1.d_vbase, nseg, L1.eptr, d_bounds, st_);
  std::vector<int64_t> hb(2 * kRowHistBins + nseg + 1, 0);
  bool have_hist = false;
  if (!all_known) {
    std::function<void()> overlap;
    if (early) overlap = [this, &L1] { launch_number(L1, early_p1_ != nullptr); };
    // state ints 1 .. kStateBounds + nseg: flags, the histogram copies, the edge bounds
    std::vector<int64_t> sb((size_t)kStateBounds + nseg, 0);
    fail_if(readback(nullptr, kStateBounds + nseg, sb.data(), overlap));
    for (int c = 0; c < kHistCopies; ++c)
      for (int b = 0; b < kRowHistBins; ++b) {
        hb[b] += sb[kStateRowHist - 1 + c * kHistStride + b];
        hb[kRowHistBins + b] += sb[kStateFaceHist - 1 + c * kHistStride + b];
      }
    for (int k = 0; k <= nseg; ++k) hb[2 * kRowHistBins + k] = sb[kStateBounds - 1 + k];
    // the early numbering was recorded before its edge count was known: its
    // boundary-flag writes (4 bytes per edge) join its algorithmic bytes now
    if (early && Profiler::get().enabled())
      Profiler::get().amend_last(L1.quad ? "number_facepoints_quad" : "number_facepoints_poly",
                                 4.0 * (double)hb[2 * kRowHistBins + 1]);
    number_early_ = early;
    early_faces_ = early && early_p1_ != nullptr;
    have_hist = true;
  } else {
    if (!all_trusted) fail_if(readback(nullptr, 0, nullptr));
    for (int k = 0; k < nseg; ++k) hb[2 * kRowHistBins + k + 1] = hb[2 * kRowHistBins + k] + jobs_[L1.jobs[k]].known_ne;
  }
  const int64_t* bounds = hb.data() + 2 * kRowHistBins;
  // row-length statistics of every level, predicted from the first level:
  // original vertices keep their valence, face points take the face order,
  // edge points have 4 (2 on the boundary), later face points 4.
  int64_t vlong = 0, flong = 0;
  int vmax = 0, fmax = 0;
  if (have_hist) {
    const int64_t* hv = hb.data();
    const int64_t* hf = hv + kRowHistBins;
    for (int b = 0; b < kRowHistBins; ++b) {
      if (b > 4) {
        vlong += hv[b];
        flong += hf[b];
      }
      if (hv[b]) vmax = b;
      if (hf[b]) fmax = b;
    }
  }

  // 3. closed-form sizes of every later level
  int Lmax = 0;
  for (int k = 0; k < nseg; ++k) {
    MeshJob& J = jobs_[L1.jobs[k]];
    int lv = J.levels;
    Lmax = std::max(Lmax, lv);
    J.nv.assign(lv + 1, 0);
    J.nf.assign(lv + 1, 0);
    J.nnz.assign(lv + 1, 0);
    J.ne.assign(lv + 1, 0);
    J.nv[0] = (int64_t)J.in.nverts;
    J.nf[0] = (int64_t)J.in.nfaces;
    J.nnz[0] = (int64_t)J.in.nnz;
    J.ne[0] = bounds[k + 1] - bounds[k];
    int64_t next = 2 * J.ne[0] - J.nnz[0];  // boundary edges
    if (next < 0) throw Error(SLAK_ERR_INTERNAL, "inconsistent edge count");
    for (int l = 1; l <= lv; ++l) {
      J.nv[l] = J.nv[l - 1] + J.nf[l - 1] + J.ne[l - 1];
      J.nf[l] = J.nnz[l - 1];
      J.nnz[l] = 4 * J.nnz[l - 1];
      int64_t nx = next << l;
      J.ne[l] = (J.nnz[l] - nx) / 2 + nx;
      check_i32(J.nv[l], "refined vertices");
      check_i32(J.nnz[l], "refined corners");
    }
  }

  // 4. level unions, destinations and outputs
  const bool lak_scheme = scheme_ == kLAK;
  auto out_alloc = [&](size_t bytes) -> void* {
    if (mode_ == OutputMode::kSeparate) return pool().alloc(bytes);
    return nullptr;  // carved from pooled blocks below
  };
  // pooled outputs: one block per kind
  // each mesh's pooled arrays start on a 128-byte line: the int4 and 64-byte
  // stores of the final kernels then never straddle lines (a pointer array
  // of nf + 1 ints would otherwise shift every later mesh by 4 bytes)
  auto line_v = [](int64_t n) { return (n + 7) & ~(int64_t)7; };     // float4
  auto line_i = [](int64_t n) { return (n + 31) & ~(int64_t)31; };   // int32
  int64_t tot_v = 0, tot_i = 0, tot_p = 0;
  for (auto& J : jobs_)
    if (J.levels > 0) {
      tot_v += line_v(J.nv[J.levels]);
      tot_i += line_i(4 * J.nnz[J.levels - 1]);
      tot_p += line_i(J.nf[J.levels] + 1);
    }
  float4* pool_v = nullptr;
  int32_t *pool_i = nullptr, *pool_p = nullptr, *pool_a = nullptr;
  if (mode_ == OutputMode::kPooled) {
    pool_v = arena_.get_n<float4>(tot_v);
    pool_i = arena_.get_n<int32_t>(tot_i);
    if (lak_scheme) {
      pool_p = arena_.get_n<int32_t>(tot_p);
      pool_a = arena_.get_n<int32_t>(tot_i);
    }
  }
  int64_t ov = 0, oi = 0, op = 0;
  for (auto& J : jobs_) {
    if (J.levels <= 0) continue;
    int64_t nvo = J.nv[J.levels], nio = 4 * J.nnz[J.levels - 1], npo = J.nf[J.levels] + 1;
    const bool lak_form = lak_scheme && J.lak_form;
    if (mode_ == OutputMode::kSeparate) {
      J.out_verts = static_cast<float4*>(out_alloc(nvo * sizeof(float4)));
      J.out_ids = static_cast<int32_t*>(out_alloc(nio * sizeof(int32_t)));
      if (lak_form) {
        J.out_ptr = static_cast<int32_t*>(out_alloc(npo * sizeof(int32_t)));
        J.out_vals = static_cast<int32_t*>(out_alloc(nio * sizeof(int32_t)));
      }
    } else {
      J.out_verts = pool_v + ov;
      J.out_ids = pool_i + oi;
      if (lak_form) {
        J.out_ptr = pool_p + op;
        J.out_vals = pool_a + oi;
      }
    }
    ov += line_v(nvo);
    oi += line_i(nio);
    op += line_i(npo);
  }

  {
    const char* e = std::getenv("SLAK_FINAL_TABLE");  // one mesh through the table path too
    force_final_table_ = e && std::atoi(e) != 0;
  }
  // k_final computes the last level of every mesh refined two or more levels
  // whose mesh two levels before the output is a quad mesh of a quad level
  // (levels >= 2 are; the first is when every input is a quad mesh)
  {
    const char* e = std::getenv("SLAK_FUSED_FINAL");
    const bool allow = !e || std::atoi(e) != 0;
    fused_final_ = false;
    for (int j : levels_[0].jobs) {
      MeshJob& J = jobs_[j];
      J.fused = allow && J.levels >= 2 && (J.levels >= 3 || levels_[0].quad);
      J.p1 = nullptr;
      fused_final_ |= J.fused;
    }
  }
  // one arena block for the level buffers that follow (positions and ids
  // handed on, the positions for k_final, the alternating structure, the
  // boundary flags and their scan, the tables), from the closed-form sizes
  // of what the level loop below allocates; anything beyond it comes from
  // the pool as before
  {
    auto line = [](int64_t n, size_t elem) { return ((size_t)std::max<int64_t>(n, 1) * elem + 255) & ~(size_t)255; };
    size_t est = 0;
    int64_t mnv = 0, mnnz = 0, mne = 0;
    for (int l = 1; l <= Lmax; ++l) {
      int64_t nv = 0, pad = 0, ids = 0, snv = 0, snnz = 0, ne = 0;
      for (int j : levels_[0].jobs) {
        const MeshJob& J = jobs_[j];
        const bool to_final = J.fused && J.levels == l + 1;
        if (J.levels > l) {  // positions handed on (to the union or to k_final)
          nv += J.nv[l];
          pad += 8;
          if (!to_final) ids += J.nnz[l];
        }
        if (J.levels > l && !to_final) {  // index l's structure: level l + 1 runs it
          snv += J.nv[l];
          snnz += J.nnz[l];
        }
        if (J.levels >= l) ne += J.ne[l - 1];
      }
      est += line(nv + pad, 16) + 256 + line(ids, 4);
      mnv = std::max(mnv, snv);
      mnnz = std::max(mnnz, snnz);
      mne = std::max(mne, ne);
    }
    est += 2 * (line(mnv + 1, 4) + 2 * line(mnnz, 8) + line(mnnz, 4));
    est += 2 * line(mne + 1, 4) + line(scan_partials_len(mne), 8);
    est += (size_t)1 << 20;
    if (Lmax >= 1) arena_.reserve(est);
  }
  // level l (1-based) consumes sizes index l-1 and produces index l
  int64_t max_nv = 0, max_nnz = 0;
  levels_.resize(Lmax);
  for (int l = 1; l <= Lmax; ++l) {
    LevelPlan& L = levels_[l - 1];
    if (l > 1) {
      LevelPlan& prev = levels_[l - 2];
      L.quad = true;
      L.jobs.clear();
      L.segs.clear();
      L.vbase.clear();
      L.fbase.clear();
      L.cbase.clear();
      int64_t vb = 0, fb = 0, cb = 0, eb = 0;
      L.fin = FinalTable{};
      L.pos = FinalTable{};
      L.pos_fused = false;
      for (int j : prev.jobs) {
        MeshJob& J = jobs_[j];
        if (J.levels < l) continue;
        if (J.fused && J.levels == l) {  // formed by k_final, outside the union
          L.fin.jobs.push_back(j);
          continue;
        }
        SegDesc S{};
        S.vb = vb;
        S.fb = fb;
        S.cb = cb;
        S.eb = eb;
        S.nv = (int32_t)J.nv[l - 1];
        S.nf = (int32_t)J.nf[l - 1];
        S.nnz = J.nnz[l - 1];
        L.jobs.push_back(j);
        L.segs.push_back(S);
        L.vbase.push_back(vb);
        L.fbase.push_back(fb);
        L.cbase.push_back(cb);
        vb += S.nv;
        fb += S.nf;
        cb += S.nnz;
        eb += J.ne[l - 1];
      }
      L.vbase.push_back(vb);
      L.fbase.push_back(fb);
      L.cbase.push_back(cb);
      L.nv = vb;
      L.nf = fb;
      L.nnz = cb;
      L.ne = eb;
      check_i32(L.nv, "union vertices of level " + std::to_string(l));
      check_i32(L.nnz, "union corners of level " + std::to_string(l));
      max_nv = std::max(max_nv, L.nv);
      max_nnz = std::max(max_nnz, L.nnz);
    }
    // destinations of this level's segments; a mesh whose next level is its
    // fused last one hands on its positions only, outside the next union
    auto to_final = [l](const MeshJob& J) { return J.fused && J.levels == l + 1; };
    int64_t nv_next = 0, nnz_next = 0;
    for (int j : L.jobs)
      if (jobs_[j].levels > l && !to_final(jobs_[j])) {
        nv_next += jobs_[j].nv[l];
        nnz_next += jobs_[j].nnz[l];
      }
    int64_t p1_total = 0, p1_off = 0;  // one block for the positions handed to k_final
    for (int j : L.jobs)
      if (to_final(jobs_[j])) p1_total += (jobs_[j].nv[l] + 7) & ~(int64_t)7;
    // the first level's positions of a single mesh: the buffer the early
    // numbering has written the face points into
    float4* early = l == 1 && early_faces_ && nv_next + p1_total <= early_p1_cap_ ? early_p1_ : nullptr;
    if (l == 1 && early_faces_ && !early) throw Error(SLAK_ERR_INTERNAL, "first-level positions exceed their buffer");
    float4* Un_P = nv_next ? (early ? early : arena_.get_n<float4>(nv_next)) : nullptr;
    int32_t* Un_ids = nnz_next ? arena_.get_n<int32_t>(nnz_next) : nullptr;
    float4* p1_block = p1_total ? (early ? early : arena_.get_n<float4>(p1_total)) : nullptr;
    int64_t vbn = 0, cbn = 0, ebn = 0;
    for (size_t k = 0; k < L.segs.size(); ++k) {
      MeshJob& J = jobs_[L.jobs[k]];
      SegDesc& S = L.segs[k];
      S.ne = J.ne[l - 1];
      if (l == 1) S.eb = bounds[k];
      L.boundary |= 2 * J.ne[l - 1] != J.nnz[l - 1];  // n_ext = 2 n_e - nnz
      if (J.levels == l) {
        S.dst_verts = J.out_verts;
        S.dst_ids = J.out_ids;
        S.dst_ptr = J.out_ptr;
        S.dst_vals = J.out_vals;
        S.id_offset = 0;
        S.next_vb = -1;
        S.next_cb = -1;
        S.next_eb = -1;
        S.dst_final = 1;
      } else if (to_final(J)) {
        J.p1 = p1_block + p1_off;
        p1_off += (J.nv[l] + 7) & ~(int64_t)7;  // 128-byte lines
        S.dst_verts = J.p1;
        S.dst_ids = nullptr;
        S.dst_ptr = nullptr;
        S.dst_vals = nullptr;
        S.id_offset = 0;
        S.next_vb = -1;
        S.next_cb = -1;
        S.next_eb = -1;
      } else {
        S.dst_verts = Un_P + vbn;
        S.dst_ids = Un_ids ? Un_ids + cbn : nullptr;
        S.dst_ptr = nullptr;
        S.dst_vals = nullptr;
        S.id_offset = vbn;
        S.next_vb = vbn;
        S.next_cb = cbn;
        S.next_eb = ebn;
        vbn += J.nv[l];
        cbn += J.nnz[l];
        ebn += J.ne[l];
      }
    }
    if (l == 1) L.ne = bounds[nseg];
    L.next_nv = nv_next;
    L.next_nnz = nnz_next;
    if (l < Lmax) {
      LevelPlan& nx = levels_[l];
      nx.ids = Un_ids;
      nx.P = Un_P;
    }
  }
  // Levels >= 2 receive their structure from the level before (two
  // alternating buffers of each kind, sized for the largest level); the
  // boundary flags and their scan are one buffer pair shared by every level.
  // Only the last level's union can be empty (every mesh in it fused), and
  // then nothing is written for it.
  int Lstruct = Lmax;  // last level whose input structure is written
  while (Lstruct >= 2 && levels_[Lstruct - 1].segs.empty()) --Lstruct;
  if (Lstruct >= 2) {
    int32_t* mt_ptr[2] = {arena_.get_n<int32_t>(max_nv + 1),
                          Lstruct >= 3 ? arena_.get_n<int32_t>(max_nv + 1) : nullptr};
    int2* mt_row[2] = {arena_.get_n<int2>(max_nnz), Lstruct >= 3 ? arena_.get_n<int2>(max_nnz) : nullptr};
    int2* mt_ci[2] = {arena_.get_n<int2>(max_nnz), Lstruct >= 3 ? arena_.get_n<int2>(max_nnz) : nullptr};
    int32_t* mt_pos[2] = {nullptr, nullptr};
    if (scheme_ == kSLAK) {
      mt_pos[0] = arena_.get_n<int32_t>(max_nnz);
      if (Lstruct >= 3) mt_pos[1] = arena_.get_n<int32_t>(max_nnz);
    }
    for (int l = 2; l <= Lstruct; ++l) {
      LevelPlan& L = levels_[l - 1];
      L.tptr = mt_ptr[l % 2];
      L.trow = mt_row[l % 2];
      L.ci = mt_ci[l % 2];
      L.mpos = mt_pos[l % 2];
      levels_[l - 2].ntptr = L.tptr;
      levels_[l - 2].ntrow = L.trow;
      levels_[l - 2].nci = L.ci;
      levels_[l - 2].nmpos = L.mpos;
    }
  }
  int64_t max_ne = 0;
  for (int l = 1; l <= Lstruct; ++l)
    if (levels_[l - 1].boundary) max_ne = std::max(max_ne, levels_[l - 1].ne);
  if (max_ne > 0) {
    int32_t* bnd = arena_.get_n<int32_t>(max_ne);
    int32_t* blk = arena_.get_n<int32_t>(max_ne + 1);
    for (int l = 1; l <= Lstruct; ++l) {
      LevelPlan& L = levels_[l - 1];
      if (!L.boundary) continue;
      if (!(l == 1 && number_early_)) L.bnd = bnd;  // the early numbering wrote level 1's own
      L.blk = blk;
      if (l >= 2) levels_[l - 2].nbnd = bnd;
    }
    int64_t need2 = scan_partials_len(max_ne);
    if (need2 > partial_len_) {
      d_partial_ = arena_.get_n<int64_t>(need2);
      partial_len_ = need2;
    }
  }
  for (int l = 1; l <= Lmax; ++l) {
    LevelPlan& L = levels_[l - 1];
    if (!have_hist) {
      L.cap = 8;
      continue;
    }
    L.long_rows = l == 1 ? vlong : vlong + flong;
    L.max_row = l == 1 ? vmax : std::max(std::max(vmax, fmax), 4);
    L.cap = autotune_cap(L.quad, L.nv, L.long_rows, L.max_row);  // used by the first level's row kernels
  }
  // the first level's rows per block, now with the row-length spread (the
  // numbering after the read-back and every graph replay use it)
  {
    const RowTiling rt = autotune_rows(levels_[0].nv, levels_[0].nnz, have_hist ? hb.data() : nullptr);
    levels_[0].rows = rt.rows;
    levels_[0].rows_num = rt.rows_num;
    levels_[0].tile = rt.tile;
  }
  // k_final tables: per fused mesh, its segment of M (the level before's
  // union), the positions of M1 and its outputs
  std::vector<int> seg_in_m(jobs_.size(), -1);  // job -> its segment of M
  const char* pe = std::getenv("SLAK_FUSED_POS");
  const bool allow_pos = !pe || std::atoi(pe) != 0;
  for (int l = 2; l <= Lmax; ++l) {
    LevelPlan& L = levels_[l - 1];
    const LevelPlan& M = levels_[l - 2];
    for (size_t k = 0; k < M.jobs.size(); ++k) seg_in_m[M.jobs[k]] = (int)k;
    // meshes whose last level is l
    for (int j : L.fin.jobs) {
      const int k = seg_in_m[j];
      if (k < 0 || !jobs_[j].p1) throw Error(SLAK_ERR_INTERNAL, "fused level without its input");
      add_final_seg(L.fin, M.segs[k], jobs_[j].p1, jobs_[j], true);
    }
    if (final_table(L.fin)) build_final_units(l, L.fin);
    // a level whose every segment hands its positions to k_final at the next
    // level, with a quad mesh two levels back: its positions from that mesh's
    // structure too (positions only)
    bool pos = allow_pos && !L.segs.empty() && (l >= 3 || levels_[0].quad);
    for (int j : L.jobs) pos = pos && jobs_[j].fused && jobs_[j].levels == l + 1 && jobs_[j].p1;
    if (pos) {
      L.pos_fused = true;
      for (size_t k = 0; k < L.segs.size(); ++k) {
        const int j = L.jobs[k];
        const int km = seg_in_m[j];
        if (km < 0) throw Error(SLAK_ERR_INTERNAL, "fused level without its input");
        L.pos.jobs.push_back(j);
        add_final_seg(L.pos, M.segs[km], L.P + L.segs[k].vb, jobs_[j], false);
      }
      if (final_table(L.pos)) build_final_units(l, L.pos);
    }
  }
  // a level whose meshes all come from k_final runs no refine or vertex
  // items over its structure: only k_final reads its M^T rows, and it needs
  // their corners alone
  for (size_t i = 1; i < levels_.size(); ++i)
    levels_[i].rows_x = levels_[i - 1].quad && (levels_[i].segs.empty() || levels_[i].pos_fused);
  upload_tables(0, (int)levels_.size(), staging_all_);
  prepared_ = true;
}

// The first level's edge numbering (E of the mapped SpGEMM: edge id and twin
// of every half-edge, boundary flags), with its face points in the same
// launch when faces is set.
void Plan::launch_number(LevelPlan& L, bool faces) {
  const double poly_b = L.quad ? 0.0 : 4.0 * (L.nf + 1) + 4.0 * L.nnz;
  const double ne = (double)L.ne;
  LevelView V{};
  V.ids = L.ids;
  V.fptr = L.fptr;
  V.cface = L.cface;
  V.P = L.P;
  V.nv = (int32_t)L.nv;
  V.nf = (int32_t)L.nf;
  V.nnz = L.nnz;
  V.tptr = L.tptr;
  V.tids = L.tids;
  V.trow = L.trow;  // sorted rows with in-neighbours (symbolic pass)
  V.eptr = L.eptr;
  V.ci = reinterpret_cast<int32_t*>(L.ci);
  V.bnd = L.bnd;
  V.flags = d_flags_;
  V.seg = L.table();
  V.rows_per_block = L.rows_num;
  V.tile_entries = L.tile;
  ProfScope ps(st_, faces ? (L.quad ? "number_facepoints_quad" : "number_facepoints_poly") : "number",
               12.0 * L.nv + 12.0 * L.nnz + 8.0 * L.nnz + (L.bnd ? 4.0 * ne : 0.0) + poly_b +
                   (faces ? 4.0 * L.nnz + 16.0 * L.nv + 16.0 * L.nf : 0.0));
  launch_number_faces(L.quad, scheme_, L.cap, V, faces, st_);
}

// Vertex classes of the mesh two levels before job J's level l (the mesh M
// of its k_final): the original vertices, then the face points and edge
// points of every level before M; one class if the sizes do not add up.
static std::vector<int64_t> vertex_classes(const MeshJob& J, int l) {
  std::vector<int64_t> cls{J.nv[0]};
  for (int k = 0; k + 2 < l; ++k) {
    cls.push_back(J.nf[k]);
    cls.push_back(J.ne[k]);
  }
  int64_t sum = 0;
  for (int64_t c : cls) sum += c;
  if (sum != J.nv[l - 2]) cls.assign(1, J.nv[l - 2]);
  return cls;
}

// k_final over several meshes: work units of at most one block each, mesh by
// mesh. A mesh's corner units (consecutive corners) and the units of each of
// its vertex classes (consecutive vertices of the class) are interleaved by
// their relative position in their list, so that every class advances over
// the surface together with the corner items (each class is numbered in the
// order of the original faces) and the positions a mesh's vertex items gather
// are still in the L2 its corner items filled.
void Plan::build_final_units(int l, FinalTable& T) {
  T.units.clear();
  struct Key {
    int64_t num, den;  // relative position (2j + 1) / (2 nch) of the chunk
    int c;             // -1: corners, else the vertex class
    int64_t j;
  };
  std::vector<Key> keys;
  for (size_t i = 0; i < T.jobs.size(); ++i) {
    const std::vector<int64_t> cls = vertex_classes(jobs_[T.jobs[i]], l);
    std::vector<int64_t> off(cls.size() + 1, 0);
    for (size_t c = 0; c < cls.size(); ++c) off[c + 1] = off[c] + cls[c];
    keys.clear();
    const int64_t ncc = (T.segs[i].nnz + kBlockItems - 1) / kBlockItems;
    for (int64_t j = 0; j < ncc; ++j) keys.push_back(Key{2 * j + 1, 2 * ncc, -1, j});
    for (size_t c = 0; c < cls.size(); ++c) {
      const int64_t nch = (cls[c] + kBlockItems - 1) / kBlockItems;
      for (int64_t j = 0; j < nch; ++j) keys.push_back(Key{2 * j + 1, 2 * nch, (int)c, j});
    }
    std::stable_sort(keys.begin(), keys.end(),
                     [](const Key& a, const Key& b) { return a.num * b.den < b.num * a.den; });
    for (const Key& k : keys) {
      if (k.c < 0) {
        const int64_t c0 = k.j * kBlockItems;
        T.units.push_back(make_int4((int)i, (int)c0, (int)std::min<int64_t>(kBlockItems, T.segs[i].nnz - c0), 0));
      } else {
        const int64_t first = off[k.c] + k.j * kBlockItems;
        const int64_t count = std::min<int64_t>(kBlockItems, cls[k.c] - k.j * kBlockItems);
        T.units.push_back(make_int4((int)i, (int)first, (int)count, 1));
      }
    }
  }
}

// One mesh of a k_final table: m is its segment of M, P1 the positions of
// M1. The output level writes the job's outputs (quads, LAK form,
// positions); the level before it writes the positions handed to k_final.
void Plan::add_final_seg(FinalTable& T, const SegDesc& m, const float4* P1, const MeshJob& J, bool output_level) {
  FinalSeg F{};
  F.cb = m.cb;
  F.vb = m.vb;
  F.eb = m.eb;
  F.nv = m.nv;
  F.nf = m.nf;
  F.ne = (int32_t)m.ne;
  F.nnz = (int32_t)m.nnz;
  F.id_off = 0;
  F.P1 = P1;
  F.out_verts = output_level ? J.out_verts : J.p1;
  if (output_level) {
    F.out_ids = J.out_ids;
    F.out_ptr = J.out_ptr;
    F.out_vals = J.out_vals;
  }
  if (T.segs.empty()) {
    T.cb.assign(1, 0);
    T.vb.assign(1, 0);
  }
  T.segs.push_back(F);
  T.cb.push_back(T.cb.back() + m.nnz);
  T.vb.push_back(T.vb.back() + m.nv);
}

// k_final over the meshes of table T at level l, from the structure of their
// mesh two levels back (the previous level's union): launched before anything
// of level l overwrites that structure (the boundary scan buffer, the
// alternating M^T buffers). ids: the output level (quads, LAK form,
// positions); otherwise positions only.
void Plan::launch_final_table(int l, FinalTable& T, bool ids, const char* stage) {
  if (T.empty()) return;
  const LevelPlan& M = levels_[l - 2];
  FinalView V{};
  V.ids = M.ids;
  V.ci = M.ci;
  V.tptr = M.tptr;
  V.trow = M.trow;
  V.trow_x = M.rows_x ? reinterpret_cast<const int32_t*>(M.trow) : nullptr;
  V.blk = M.blk;
  V.mpos = M.mpos;
  V.nnz = T.cb.back();
  V.nv = T.vb.back();
  if (final_table(T)) {
    V.segs = T.d_segs;
    V.units = T.d_units;
    V.nunits = (int32_t)T.units.size();
    V.nseg = (int32_t)T.segs.size();
  } else {
    V.one = T.segs[0];
    V.nseg = 1;
    // vertex classes of M: the original vertices, then the face points and
    // edge points of every level before M (leading ones merged past kMaxVSeg)
    std::vector<int64_t> cls = vertex_classes(jobs_[T.jobs[0]], l);
    while (cls.size() > (size_t)kMaxVSeg) {
      cls[1] += cls[0];
      cls.erase(cls.begin());
    }
    V.nvseg = (int32_t)cls.size();
    V.vseg[0] = 0;
    for (size_t i = 0; i < cls.size(); ++i) V.vseg[i + 1] = V.vseg[i] + (int32_t)cls[i];
  }
  // algorithmic bytes: reads the ids, (edge, twin) pairs, M^T rows and
  // offsets of M (boundary scan, SLAK positions) and the positions of M1;
  // writes every output position and, for the output level, the quads (and
  // the LAK form)
  // (several meshes: the LAK form is written by a launch of its own after it)
  const bool multi = final_table(T);
  double reads = 0, writes = 0, lak_form = 0;
  for (int j : T.jobs) {
    const MeshJob& J = jobs_[j];
    const double mnnz = (double)J.nnz[l - 2], mnv = (double)J.nv[l - 2], mne = (double)J.ne[l - 2];
    const double nnz1 = (double)J.nnz[l - 1], nv1 = (double)J.nv[l - 1], nf1 = (double)J.nf[l - 1];
    const double ne1 = (double)J.ne[l - 1];
    reads += (M.rows_x ? 16.0 : 20.0) * mnnz + 4.0 * mnv + (M.blk ? 4.0 * mne : 0.0) + (scheme_ == kSLAK ? 4.0 * mnnz : 0.0) + 16.0 * nv1;
    writes += 16.0 * nv1 + 16.0 * nf1 + 16.0 * ne1;
    if (ids) {
      writes += 16.0 * nnz1;
      if (J.out_ptr) (multi ? lak_form : writes) += 4.0 * (nnz1 + 1) + 16.0 * nnz1;
    }
  }
  {
    ProfScope ps(st_, stage, reads + writes);
    launch_final(scheme_, V, ids, st_);
  }
  if (multi && lak_form > 0) {
    ProfScope ps(st_, "lak_form", lak_form);
    launch_final_lak_form(V, st_);
  }
}

void Plan::record_level(int l) {
  LevelPlan& L = levels_[l - 1];
  if (l > 1) {
    launch_final_table(l, L.fin, true, "refine_vertex_final_quad");
    if (L.pos_fused) launch_final_table(l, L.pos, false, "refine_vertex_pos_quad");
  }
  if (L.segs.empty()) return;  // every mesh of the level was fused
  const double poly_b = L.quad ? 0.0 : 4.0 * (L.nf + 1) + 4.0 * L.nnz;
  const double ne = (double)L.ne;
  if (l == 1 && !number_early_) {
    // edge numbering of the first level (later levels get theirs in closed form)
    launch_number(L, true);
  }
  if (l == 1 && number_early_ && !early_faces_) {
    ProfScope ps(st_, L.quad ? "facepoints_quad" : "facepoints_poly",
                 4.0 * L.nnz + 16.0 * L.nv + 16.0 * L.nf + (L.quad ? 0.0 : 4.0 * (L.nf + 1)));
    launch_faces(L.quad, true, false, L.ids, L.fptr, L.P, (int32_t)L.nf, L.table(), nullptr, nullptr, st_);
  }
  if (L.blk) {
    ProfScope ps(st_, "scan", 8.0 * ne + 4);
    launch_scan(L.bnd, L.blk, L.ne, d_partial_, d_flags_, st_);
  }
  if (L.pos_fused) return;  // its positions came from k_final; the scan above serves the next level's
  if (l > 1) {
    ProfScope ps(st_, L.quad ? "facepoints_quad" : "facepoints_poly",
                 4.0 * L.nnz + 16.0 * L.nv + 16.0 * L.nf + (L.quad ? 0.0 : 4.0 * (L.nf + 1)));
    launch_faces(L.quad, true, false, L.ids, L.fptr, L.P, (int32_t)L.nf, L.table(), nullptr, nullptr, st_);
  }
  RefineView R{};
  R.ids = L.ids;
  R.fptr = L.fptr;
  R.cface = L.cface;
  R.P = L.P;
  R.nv = (int32_t)L.nv;
  R.nf = (int32_t)L.nf;
  R.nnz = L.nnz;
  R.tptr = L.tptr;
  R.trow = L.trow;
  R.ci = L.ci;
  R.blk = L.blk;
  R.mpos = L.mpos;
  R.seg = L.table();
  R.ntptr = L.ntptr;
  R.ntrow = L.ntrow;
  R.nci = L.nci;
  R.nbnd = L.nbnd;
  R.nmpos = L.nmpos;
  R.next_nv = (int32_t)L.next_nv;
  R.next_nnz = (int32_t)L.next_nnz;
  R.next_rows_x = l < (int)levels_.size() && levels_[l].rows_x ? 1 : 0;
  if (L.segs.size() == 1 && l >= 2) {
    // one mesh: its vertex classes (the original vertices, then the face and
    // edge points of every earlier level) as runs in step with the corners
    std::vector<int64_t> cls = vertex_classes(jobs_[L.jobs[0]], l + 1);
    while (cls.size() > (size_t)kMaxVSeg) {
      cls[1] += cls[0];
      cls.erase(cls.begin());
    }
    R.nvseg = (int32_t)cls.size();
    R.vseg[0] = 0;
    for (size_t i = 0; i < cls.size(); ++i) R.vseg[i + 1] = R.vseg[i] + (int32_t)cls[i];
  }
  const bool next = L.ntptr != nullptr;
  {
    // algorithmic bytes of the fused launch, each array once: reads ids,
    // (edge, twin) pairs, positions, face points and the M^T rows; writes
    // edge and vertex points, the refined quads of the meshes that keep them
    // (not those whose next level is formed by k_final) and, for the meshes
    // that feed another level, their (edge, twin) pairs, M^T rows and
    // offsets, boundary flags and (SLAK) M^T positions
    double reads = 4.0 * L.nnz + 8.0 * L.nnz + 16.0 * L.nv + 16.0 * L.nf + 4.0 * L.nv + 8.0 * L.nnz +
                   (L.blk ? 4.0 * ne : 0.0) + poly_b;
    double writes = 16.0 * ne + 16.0 * L.nv;
    for (const SegDesc& S : L.segs) {
      const double snnz = (double)S.nnz, sne = (double)S.ne;
      if (S.dst_ids) writes += 16.0 * snnz;
      if (S.dst_ptr) writes += 4.0 * (snnz + 1) + 16.0 * snnz;  // LAK form of a final level
      if (next && S.next_cb >= 0) {
        writes += (R.next_rows_x ? 48.0 : 64.0) * snnz + 4.0 * ((double)S.nv + (double)S.nf + sne) + (L.nbnd ? 4.0 * (snnz + 2.0 * sne) : 0.0) +
                  (scheme_ == kSLAK ? 16.0 * snnz : 0.0);
        if (scheme_ == kSLAK) reads += 4.0 * snnz;
      }
    }
    ProfScope ps(st_, L.quad ? "refine_vertex_quad" : "refine_vertex_poly", reads + writes);
    launch_level(L.quad, scheme_, R, st_);
  }
}

void Plan::launch(bool include_analysis) {
  if (levels_.empty()) return;
  if (include_analysis) number_early_ = false;  // the analysis runs again: number in order
  LevelPlan& L1 = levels_[0];
  if (include_analysis) {
    launch_fill_i32(d_state_, n_state_, 0, st_);  // a fill kernel: graph-safe
    {
      ProfScope ps(st_, "gather", 4.0 * (L1.nf + 1) + 12.0 * L1.nnz + (L1.quad ? 0.0 : 8.0 * L1.nnz + 4.0 * L1.nf));
      launch_gather(L1.table(), (int32_t)L1.nf, scheme_ == kLAK, const_cast<int32_t*>(L1.ids),
                    const_cast<int32_t*>(L1.fptr), const_cast<int32_t*>(L1.cface), d_flags_, d_cnt_,
                    (int32_t)L1.nv, d_state_ + kStateFaceHist, st_);
    }
    if (P1_) {
      ProfScope ps(st_, "gather_verts", 32.0 * L1.nv);
      launch_gather_verts(L1.table(), (int32_t)L1.nv, P1_, st_);
    }
    record_analysis(L1);
  }
  for (int l = 1; l <= (int)levels_.size(); ++l) record_level(l);
}

void Plan::run_graph() {
  if (levels_.empty()) return;
  // The legacy default stream cannot be captured: graph plans run on a stream
  // of their own, ordered after the caller's default-stream work by an event.
  if (!capture_st_) {
    if (st_) {
      capture_st_ = st_;
    } else {
      SLAK_HIP_CHECK(hipStreamCreateWithFlags(&owned_st_, hipStreamNonBlocking));
      SLAK_HIP_CHECK(hipEventCreateWithFlags(&order_ev_, hipEventDisableTiming));
      capture_st_ = owned_st_;
    }
  }
  if (owned_st_) {
    SLAK_HIP_CHECK(hipEventRecord(order_ev_, nullptr));
    SLAK_HIP_CHECK(hipStreamWaitEvent(owned_st_, order_ev_, 0));
  }
  hipStream_t saved = st_;
  st_ = capture_st_;
  struct Restore {
    hipStream_t& ref;
    hipStream_t val;
    ~Restore() { ref = val; }
  } restore{st_, saved};
  if (!exec_) {
    // event-pair timing stays out of the captured graph
    const bool prof = Profiler::get().enabled();
    Profiler::get().set_enabled(false);
    SLAK_HIP_CHECK(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
    try {
      launch(true);
      Profiler::get().set_enabled(prof);
    } catch (...) {
      Profiler::get().set_enabled(prof);
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(st_, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    SLAK_HIP_CHECK(hipStreamEndCapture(st_, &graph_));
    SLAK_HIP_CHECK(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
  }
  SLAK_HIP_CHECK(hipGraphLaunch(exec_, st_));
  if (owned_st_) {
    // later work on the caller's default stream sees the results
    SLAK_HIP_CHECK(hipEventRecord(order_ev_, owned_st_));
    SLAK_HIP_CHECK(hipStreamWaitEvent(nullptr, order_ev_, 0));
  }
}

void Plan::finish() {
  int fl = readback(nullptr, 0, nullptr);
  if (fl) throw Error(SLAK_ERR_INTERNAL, "refinement produced invalid topology:" + flags_message(fl));
}

void Plan::release_outputs() { outputs_released_ = true; }

size_t Plan::bytes() const { return arena_.bytes(); }

}  // namespace slak
