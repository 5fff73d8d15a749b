// Host-side declarations shared by the kernel launchers and the runtime.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <exception>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <map>
#include <vector>

#include "slak_device.hpp"

namespace slak {

// ---------------------------------------------------------------------------
// errors: every failure becomes a slak::Error carrying a C-ABI status code

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define SLAK_HIP_CHECK(expr)                                                                           \
  do {                                                                                                 \
    hipError_t _e = (expr);                                                                            \
    if (_e != hipSuccess)                                                                              \
      throw ::slak::Error(_e == hipErrorOutOfMemory ? 4 : 3,                                           \
                          std::string("HIP error in ") + #expr + ": " + hipGetErrorString(_e));        \
  } while (0)

// ---------------------------------------------------------------------------
// launchers (slak_kernels.hip)

// reads SLAK_XCD_REMAP once and sets the device-side option (call before capture)
void init_device_options();
void launch_fill_i32(int32_t* p, int64_t n, int32_t value, hipStream_t st);
// also zeroes zero[0..nzero) (the counters of the analysis that follows) and,
// when hist_faces is set, adds the face orders to that zeroed histogram block
// (kHistCopies copies of 17 bins; the host sums the copies)
void launch_gather(const SegTable& seg, int32_t nf_total, bool use_vals, int32_t* ids_pos, int32_t* fptr,
                   int32_t* cface, int* flags, int32_t* zero, int32_t nzero, int32_t* hist_faces, hipStream_t st);
void launch_gather_verts(const SegTable& seg, int32_t nv_total, float4* P, hipStream_t st);
constexpr int kRowHistBins = 17;  // row lengths 0..15, 16 = longer
// LDS tile of the first level's row kernels, in M^T entries (sized per level
// by the plan, at most this): a block's rows are staged through it when their
// total length fits
constexpr int kRowTileMax = 2560;
// Each histogram is kept as kHistCopies copies, one 128-byte line apart; block
// b adds its counts to copy b % kHistCopies. A single copy takes one
// device-scope atomic per block on the same address, and those serialize: on
// a 4096-block launch they cost about 35 us.
constexpr int kHistCopies = 32;
constexpr int kHistStride = 32;  // ints per copy
constexpr int kHistInts = kHistCopies * kHistStride;
// The plan's state block (int32, zeroed before the gather), read back from
// its start after the first level's analysis:
// [flags | pad | row hist copies | face hist copies | edge bounds (nseg + 1) | mesh flags (nseg)]
constexpr int kStateRowHist = 32;
constexpr int kStateFaceHist = kStateRowHist + kHistInts;
constexpr int kStateBounds = kStateFaceHist + kHistInts;
// bounds[k] = eptr[vbase[k]], k = 0..nseg
void launch_seg_bounds(const int64_t* vbase, int32_t nseg, const int32_t* eptr, int32_t* bounds, hipStream_t st);
void launch_copy_probe(const float4* a, float4* b, int64_t n, hipStream_t st);
// *out = order-sensitive 64-bit digest of bytes at p (out: device, 8 bytes)
void launch_digest(const void* p, int64_t bytes, uint64_t* out, hipStream_t st);
// guard (or null): the kernel does nothing when it holds an input-validation flag
void launch_faces(bool quad, bool fp, bool count, const int32_t* ids, const int32_t* fptr, const float4* P, int32_t nf,
                  const SegTable& seg, int32_t* cnt, const int* guard, hipStream_t st);
void launch_transpose(bool quad, const int32_t* ids, const int32_t* fptr, int32_t nf, const int32_t* tptr,
                      int32_t* cnt, int32_t* tids, const int* guard, hipStream_t st);
void launch_symbolic(bool quad, int scheme, int cap, const LevelView& L, hipStream_t st);
// first level: edge numbering and (faces set) face points in one launch; the
// rows come from the symbolic pass (L.trow sorted with in-neighbours, L.tids
// sorted in place for rows above the register capacity)
void launch_number_faces(bool quad, int scheme, int cap, const LevelView& L, bool faces, hipStream_t st);
void launch_spgemm_fde(bool quad, const LevelView& L, hipStream_t st);
// refine (per corner) and vertex (per vertex) items of a level in one launch;
// the refine items also write the LAK form of the segments with dst_ptr set
void launch_level(bool quad, int scheme, const RefineView& R, hipStream_t st);
// last level of one mesh or several (V.segs) from the quad mesh two levels
// back; ids: also the output quads and LAK form (false: positions only)
void launch_final(int scheme, const FinalView& V, bool ids, hipStream_t st);
// the LAK form (ptr, vals) of a several-mesh output level, after its k_final
// launch (which leaves it out)
void launch_final_lak_form(const FinalView& V, hipStream_t st);
// homogeneous-mesh conversions: per-face reorder of (ids, vals) by vals or by
// ids (faces by ptr, or stride when ptr is null; 3 or 4 corners); fills of
// stride ptr / 1-based positions
void launch_sort_faces(bool by_vals, const int32_t* ptr, int32_t stride, int32_t nf, int64_t nnz, int32_t* ids,
                       int32_t* vals, int* flags, hipStream_t st);
void launch_fill_homogeneous(int32_t* ptr, int64_t nptr, int32_t* vals, int64_t nvals, int32_t stride, hipStream_t st);
int64_t scan_partials_len(int64_t n);
// Scratch of the entry points that allocate nothing: a 1 MiB file-scope
// device array per GPU (flags word, then int64 scan partials), shared by
// every such call on that GPU.
constexpr size_t kNoAllocScratchBytes = size_t(1) << 20;
struct NoAllocScratch {
  int* flags;
  int64_t* partial;
  int64_t partial_len;
};
// the current device's scratch (its address is looked up once per device)
NoAllocScratch noalloc_scratch();
// largest scan the scratch holds the partials of
int64_t noalloc_scan_max_items();
// out[i] = sum(in[0..i)), out[n] = total (also to *total_out when given)
void launch_scan(const int32_t* in, int32_t* out, int64_t n, int64_t* partial, int* flags, hipStream_t st,
                 int32_t* total_out = nullptr);

// ---------------------------------------------------------------------------
// caching device allocator (one per device)

class Pool {
 public:
  void* alloc(size_t bytes);
  // Returns the block to the cache. Callers free only after the stream that
  // used the block has been synchronized.
  bool release(void* p);
  bool owns(void* p);
  void trim();
  size_t in_use() const { return in_use_; }
  size_t cached() const { return cached_; }
  size_t peak() const { return peak_; }
  void reset_peak() { peak_ = in_use_; }

 private:
  std::mutex mu_;
  std::multimap<size_t, void*> free_;
  std::unordered_map<void*, size_t> used_;
  size_t in_use_ = 0, cached_ = 0, peak_ = 0;
};

Pool& pool();


// Allocations of one library call: released back to the pool after the
// call's final stream synchronization, except buffers handed to the caller.
class Arena {
 public:
  explicit Arena(Pool& p) : pool_(p) {}
  // On an error path kernels of the call may still be queued: wait for the
  // device before the blocks go back to the pool.
  ~Arena() {
    if (!blocks_.empty() && std::uncaught_exceptions() > 0) (void)hipDeviceSynchronize();
    release_all();
  }
  void* get(size_t bytes);
  template <class T>
  T* get_n(int64_t n) {
    return static_cast<T*>(get(sizeof(T) * (size_t)(n > 0 ? n : 1)));
  }
  // Later get() calls are carved (256-byte aligned) from one pool block of
  // `bytes` while it lasts: one pool round trip instead of dozens while a
  // plan is set up. Blocks carved this way cannot be kept.
  void reserve(size_t bytes);
  void keep(void* p);  // ownership moves to the caller (blocks from get() without a reserve)
  void release_all();
  bool empty() const { return blocks_.empty(); }
  size_t bytes() const { return bytes_; }

 private:
  Pool& pool_;
  std::vector<void*> blocks_;
  size_t bytes_ = 0;
  char* bump_ = nullptr;
  size_t bump_left_ = 0;
};

// ---------------------------------------------------------------------------
// per-stage HIP-event timing (TimeProfileIndividual, LibInclude/profiling.h:182-200):
// when enabled, every launch site records an event pair on its stream and the
// stage's algorithmic bytes (compulsory reads + writes of distinct arrays).

class Profiler {
 public:
  static Profiler& get();
  bool enabled() const { return enabled_; }
  void set_enabled(bool on) { enabled_ = on; }
  void begin(hipStream_t st, const char* name, double bytes);
  void end(hipStream_t st);
  // adds `bytes` to the latest record of stage `name` (names compare by text)
  void amend_last(const char* name, double bytes);
  void reset();
  std::string report();  // JSON

 private:
  struct Rec {
    const char* name;
    double bytes;
    hipEvent_t a, b;
  };
  hipEvent_t event();
  std::mutex mu_;
  bool enabled_ = false;
  std::vector<Rec> recs_;
  std::vector<hipEvent_t> spare_;
};

struct ProfScope {
  hipStream_t st;
  bool on;
  ProfScope(hipStream_t s, const char* name, double bytes) : st(s), on(Profiler::get().enabled()) {
    if (on) Profiler::get().begin(st, name, bytes);
  }
  ~ProfScope() {
    if (on) Profiler::get().end(st);
  }
};

}  // namespace slak
