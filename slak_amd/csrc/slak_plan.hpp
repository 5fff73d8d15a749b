// Multi-mesh level plans: the level chain of subdivideIteratively
// (LAKCatmullClarkImpl.cu:512-571, CatmullClarkImpl.cu:457-515) over a union
// of meshes, with every size after the first level's edge count known in
// closed form, so the whole chain can be launched back to back or captured in
// one hipGraph.
#pragma once

#include <functional>

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "slak_c.h"
#include "slak_internal.hpp"

namespace slak {

// One caller mesh of a plan.
struct MeshJob {
  slak_mesh in{};   // caller's device mesh (never modified)
  int target = 0;   // requested level
  int levels = 0;   // levels to refine (a non-quad input gets at least one)
  // A mesh this library refined itself: valid by construction (no validation
  // round trip) and with its edge count known in closed form (known_ne >= 0).
  bool trusted = false;
  int64_t known_ne = -1;
  bool lak_form = true;  // LAK scheme: also write ptr/vals of the final level
  // closed-form sizes of the input of level l (l = 0..levels) and its edges
  std::vector<int64_t> nv, nf, nnz, ne;
  // final outputs
  float4* out_verts = nullptr;
  int32_t* out_ids = nullptr;
  int32_t* out_ptr = nullptr;
  int32_t* out_vals = nullptr;
  int bad_flags = 0;
  // the last level is formed by k_final from the quad mesh two levels back;
  // p1: the positions of the level before the output (its only output)
  bool fused = false;
  float4* p1 = nullptr;
};

// Meshes whose level is formed by k_final from the structure of their mesh M
// two levels before it: per mesh its segment of M, the positions of M1 and
// the outputs; item bases; work units of the table path (several meshes).
struct FinalTable {
  std::vector<int> jobs;
  std::vector<FinalSeg> segs;
  std::vector<int64_t> cb, vb;  // corner and vertex item bases (n + 1)
  std::vector<int4> units;
  FinalSeg* d_segs = nullptr;
  int4* d_units = nullptr;
  bool empty() const { return segs.empty(); }
};

// rows per block and LDS row tile (entries) of the first level's row kernels
struct RowTiling {
  int rows;      // k_symbolic
  int rows_num;  // the numbering of k_number_faces
  int tile;
};

struct LevelPlan {
  bool quad = true;
  int64_t nv = 0, nf = 0, nnz = 0;  // union sizes
  const int32_t* ids = nullptr;
  const int32_t* fptr = nullptr;
  const int32_t* cface = nullptr;
  const float4* P = nullptr;
  int64_t ne = 0;           // union edges
  bool boundary = false;    // some mesh of the level has boundary edges
  int cap = 8;              // first level: M^T row length held in registers (autotuned)
  int rows = 256;           // first level: M^T rows per block of k_symbolic (autotuned)
  int rows_num = 128;       // first level: M^T rows per block of the numbering (autotuned)
  int tile = 1536;          // first level: LDS row tile of the row kernels, entries (autotuned)
  int64_t long_rows = -1;   // rows longer than 4 (from the first level's histogram)
  int max_row = -1;
  // first-level analysis scratch
  int32_t *cnt = nullptr, *tids = nullptr, *ecnt = nullptr, *eptr = nullptr;
  // the level's structure: M^T rows with in-neighbours, (edge id, twin) per
  // corner, boundary flags per edge and their scan (boundary levels)
  int32_t* tptr = nullptr;
  int2* trow = nullptr;
  int2* ci = nullptr;
  int32_t *bnd = nullptr, *blk = nullptr;
  int32_t* mpos = nullptr;  // M^T storage position per corner (SLAK)
  // the next level's structure, written by this level's refine and vertex kernels
  int32_t* ntptr = nullptr;
  int2* ntrow = nullptr;
  int2* nci = nullptr;
  int32_t* nbnd = nullptr;
  int32_t* nmpos = nullptr;
  int64_t next_nv = 0, next_nnz = 0;  // next level's union sizes
  std::vector<int> jobs;  // segment -> job
  std::vector<SegDesc> segs;
  std::vector<int64_t> vbase, fbase, cbase;
  // meshes whose last level is this one, formed by k_final (output quads,
  // LAK form, positions); they are not segments of this level's union
  FinalTable fin;
  // every segment of this level hands its positions to k_final at the next
  // level: they are formed here by k_final's positions-only variant, and
  // this level launches no face points and no refine or vertex items
  bool pos_fused = false;
  FinalTable pos;
  // this level's M^T rows are read by k_final alone (no refine or vertex
  // items run over them): the level before writes only their corners, as
  // int32 in the row buffer
  bool rows_x = false;
  SegDesc* d_segs = nullptr;
  int64_t* d_vbase = nullptr;
  int64_t* d_fbase = nullptr;
  int64_t* d_cbase = nullptr;
  int* d_mflags = nullptr;
  SegTable table() const { return SegTable{d_segs, d_vbase, d_fbase, (int32_t)segs.size(), d_mflags, d_cbase}; }
};

enum class OutputMode {
  kSeparate,  // one pool block per output array (handed to the caller)
  kPooled,    // one block per array kind for all meshes (owned by the plan)
};

class Plan {
 public:
  Plan(int scheme, std::vector<MeshJob> jobs, OutputMode mode, hipStream_t st);
  ~Plan();
  Plan(const Plan&) = delete;
  Plan& operator=(const Plan&) = delete;

  // Validates the inputs and runs the first level's analysis (one blocking
  // round trip, none for trusted meshes of known edge count), then sizes and
  // allocates everything. Throws on bad input.
  void prepare();
  // Launches the chain on the plan's stream. With include_analysis the first
  // level is recomputed from the caller's meshes (used for graph replays).
  void launch(bool include_analysis);
  // Captures launch(true) into a hipGraph (once) and replays it.
  void run_graph();
  // Waits for the stream and checks the device flag word.
  void finish();
  // Hands the outputs of kSeparate plans to the caller (not freed by the plan).
  void release_outputs();

  std::vector<MeshJob>& jobs() { return jobs_; }
  LevelPlan& level(int i) { return levels_[i]; }
  int* flags() { return d_flags_; }
  Arena& arena() { return arena_; }
  int syncs() const { return syncs_; }
  hipStream_t stream() const { return st_; }
  size_t bytes() const;
  int max_levels() const { return (int)levels_.size(); }
  std::string describe_failure(int fl) const;

 private:
  void plan_first_level();
  void record_analysis(LevelPlan& L);
  void record_level(int l);
  // state_ints > 0: also carries the plan's zeroed state block (first upload)
  // pinned host staging of a table upload (recycled through a process-wide
  // pool; released once the plan's stream has passed the copy)
  struct Staging {
    char* p = nullptr;
    size_t cap = 0;
  };
  void upload_tables(int l0, int l1, Staging& staging, int64_t state_ints = 0);
  // overlap: work enqueued after the copy and before the wait (the host then
  // waits for the copy only)
  int readback(const int32_t* src, int count, int64_t* dst, const std::function<void()>& overlap = {});
  void launch_number(LevelPlan& L, bool faces);

  int scheme_;
  std::vector<MeshJob> jobs_;
  OutputMode mode_;
  hipStream_t st_;
  Arena arena_;
  std::vector<LevelPlan> levels_;
  int* d_flags_ = nullptr;
  int64_t* d_partial_ = nullptr;
  int64_t partial_len_ = 0;
  // per-plan device state, zeroed by the first table upload:
  // [flags | row hist (17) | face hist (17) | edge bounds (nseg + 1) | mesh flags (nseg)]
  int32_t* d_state_ = nullptr;
  int64_t n_state_ = 0;
  int32_t* d_cnt_ = nullptr;  // first level: valence counters (zeroed by the gather)
  Staging staging_first_, staging_all_;  // segment table uploads
  // first-level union gather (multi-mesh)
  float4* P1_ = nullptr;
  bool prepared_ = false;
  bool outputs_released_ = false;
  int syncs_ = 0;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  hipStream_t capture_st_ = nullptr;  // stream the graph is captured on and replayed
  hipStream_t owned_st_ = nullptr;    // created when the caller's stream is the default one
  hipEvent_t order_ev_ = nullptr;
  // some mesh refined two or more levels with a quad mesh two levels before
  // its output: its last level is computed from that mesh's structure
  // (k_final) and the level before it writes positions only
  bool fused_final_ = false;
  // k_final's meshes come from a device table and work units (several meshes)
  bool force_final_table_ = false;
  bool final_table(const FinalTable& T) const { return T.segs.size() > 1 || (force_final_table_ && !T.empty()); }
  void add_final_seg(FinalTable& T, const SegDesc& m, const float4* P1, const MeshJob& J, bool output_level);
  void launch_final_table(int l, FinalTable& T, bool ids, const char* stage);
  void build_final_units(int l, FinalTable& T);
  // single mesh: the first level's edge numbering ran during the analysis
  // read-back (the level then launches its face points only)
  bool number_early_ = false;
  // ... and, when the first level feeds another, its face points, into the
  // first level's position buffer allocated before the read-back
  bool early_faces_ = false;
  float4* early_p1_ = nullptr;
  int64_t early_p1_cap_ = 0;
  std::vector<void*> pooled_outputs_;
};

std::string flags_message(int f);

}  // namespace slak
