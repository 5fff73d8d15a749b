// Device-side building blocks of the Catmull-Clark level pipeline (gfx950).
//
// A level runs over a union of meshes ("segments"): mesh k owns the vertex
// range [vb_k, vb_k + nv_k), faces [fb_k, ...), corners [cb_k, ...) and edges
// [eb_k, eb_k + ne_k) of the level's arrays, which hold global (union)
// indices. Data layout in HBM:
//   ids[nnz]           position-ordered corner -> vertex (quad levels: 4 per face, int4-aligned)
//   fptr[nf+1], cface[nnz]   face offsets and corner -> face (poly levels only)
//   P[nv]              float4 positions (x, y, z, w=1)
//   tptr[nv+1], trow[nnz]    M^T: vertex -> incident corners, ascending, each
//                      with its in-neighbour (the corner's predecessor vertex)
//   ci[nnz]            per corner: mesh-local edge id of the outgoing
//                      half-edge and the twin corner (-1 on the boundary)
//   bnd[ne], blk[ne+1] boundary flag per edge and its exclusive scan (levels
//                      with a boundary only)
// Outputs go to each segment's own destination: refined positions
// [vertex points | face points | edge points] and refined quads in the
// mesh's local numbering (+ the segment's id offset when the result feeds the
// next level's union). A level that feeds another also writes that level's
// tptr/trow/ci/bnd in closed form, so only the first level numbers edges.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace slak {

enum Scheme : int { kLAK = 0, kSLAK = 1 };

// validation / error flags (OR-ed into a device word)
enum : int {
  kFlagBadPtr = 1,
  kFlagBadOrder = 2,
  kFlagBadId = 4,
  kFlagBadVals = 8,
  kFlagBadQuad = 16,
  kFlagDupHalfEdge = 32,
  kFlagDegenerate = 64,
  kFlagOverflow = 128,
};
// flags the first-level gather raises for an invalid input mesh; the analysis
// kernels after it do nothing once one is set (their indices would come from
// the invalid mesh)
constexpr int kFlagsInput = kFlagBadPtr | kFlagBadOrder | kFlagBadId | kFlagBadVals | kFlagBadQuad;

// One mesh of a level's union.
struct SegDesc {
  int64_t vb, fb, cb;      // union bases: vertices, faces, corners
  int32_t nv, nf;          // local sizes
  int64_t nnz;             // local corners
  int64_t ne;              // local edges (edge points the destination holds)
  int64_t eb;              // union base of the mesh's edges (boundary flags)
  float4* dst_verts;       // nv + nf + ne refined positions
  int32_t* dst_ids;        // 4 * nnz refined quad ids
  int32_t* dst_ptr;        // LAK form (final level) or null
  int32_t* dst_vals;       // LAK form (final level) or null
  int64_t id_offset;       // added to local refined ids
  // next level's union bases of this mesh (-1: the mesh finishes here); the
  // level writes the refined mesh's M^T rows there directly
  int64_t next_vb, next_cb, next_eb;
  // first-level sources (the caller's mesh)
  const int32_t* src_ptr;  // null: homogeneous with src_stride
  const int32_t* src_ids;
  const int32_t* src_vals; // null: storage order
  const float4* src_verts;
  int32_t src_stride;
  int32_t require_quad;
  int32_t dst_final;       // 1: the destination is the caller's output (stored non-temporally)
};

struct SegTable {
  const SegDesc* d;
  const int64_t* vbase;  // n + 1 vertex bases
  const int64_t* fbase;  // n + 1 face bases
  int32_t n;
  int* mflags;           // per-segment validation flags
  const int64_t* cbase;  // n + 1 corner bases
};

// A pointer read from device memory (a mesh table entry) is generic to the
// compiler: its accesses become FLAT instructions, whose completion it does
// not count in order, so every wait on one of them waits for all of the
// wave's outstanding memory operations, stores included. Passing the pointer
// through the global address space lets it emit global instructions (every
// pointer in the tables is device memory).
// (Through an integer: a direct generic -> global -> generic cast pair is
// folded away before the address spaces are inferred.)
template <class T>
__device__ __forceinline__ T* to_global(T* p) {
  return (T*)(__attribute__((address_space(1))) T*)(uintptr_t)p;
}

// Segment of a union index: the last k with base[k] <= x (empty segments skipped).
__device__ __forceinline__ int seg_of(const int64_t* base, int n, int64_t x) {
  if (n <= 1) return 0;
  int lo = 0, hi = n;  // base[lo] <= x < base[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (base[mid] <= x)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// seg_of for items that rise with the lane (item = first + lane within a
// wave): one wave-uniform binary search for the first active lane's item
// (scalar loads), then a few steps forward per lane. Falls back to the
// per-lane search when an item lies before the first lane's or more than 8
// segments after it.
__device__ __forceinline__ int seg_of_wave(const int64_t* base, int n, int64_t x) {
  if (n <= 1) return 0;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)x >> 32));
  const int64_t x0 = (int64_t)(((uint64_t)hi << 32) | lo);
  if (x < x0) return seg_of(base, n, x);
  int k = seg_of(base, n, x0);
#pragma unroll 1
  for (int it = 0; it < 8 && k + 1 < n && base[k + 1] <= x; ++it) ++k;
  if (k + 1 < n && base[k + 1] <= x) k = seg_of(base, n, x);
  return k;
}

struct LevelView {
  const int32_t* ids;
  const int32_t* fptr;   // poly levels
  const int32_t* cface;  // poly levels
  const float4* P;
  int32_t nv;
  int32_t nf;
  int64_t nnz;
  const int32_t* tptr;
  int32_t* tids;         // M^T corners (first level: transposed, rows sorted on use)
  int2* trow;            // sorted rows with in-neighbours (written by the symbolic pass), or null
  int32_t* mpos;         // M^T storage position of every corner (SLAK, symbolic pass), or null
  int32_t* ecnt;
  int32_t* hist_rows;    // zeroed histogram copies of row lengths (symbolic pass), or null
  const int32_t* eptr;
  int32_t* ci;           // (edge id, twin) pairs per corner, written by the number pass
  int32_t* bnd;          // union edge boundary flags (number pass), or null: closed level
  int32_t* E_ids;   // LAK edge matrix rows (SpGEMM operator)
  int32_t* E_vals;  // LAK edge matrix values (+-(id+1))
  int* flags;
  SegTable seg;
  // mapped SpGEMM operator outputs (F, D rows share M^T's row offsets)
  int32_t* F_ids;
  int32_t* F_vals;
  int32_t* D_ids;
  int32_t* D_vals;
  int64_t ne_cap;  // E capacity of the SpGEMM operator
  int32_t rows_per_block;  // first-level row kernels: M^T rows per block (64, 128 or 256)
  int32_t tile_entries;    // first-level row kernels: entries of the block's LDS row tile
};

__device__ __forceinline__ float4 f4add(float4 a, float4 b) {
  return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
__device__ __forceinline__ float4 f4mul(float4 a, float s) { return make_float4(a.x * s, a.y * s, a.z * s, a.w * s); }
__device__ __forceinline__ float4 f4div(float4 a, float s) { return make_float4(a.x / s, a.y / s, a.z / s, a.w / s); }

// Corner geometry: face, successor and predecessor corner of corner t.
template <bool QUAD>
__device__ __forceinline__ void corner_geo(const LevelView& L, int32_t t, int32_t& f, int32_t& tn, int32_t& tp) {
  if (QUAD) {
    f = t >> 2;
    int32_t b = t & ~3;
    tn = b | ((t + 1) & 3);
    tp = b | ((t + 3) & 3);
  } else {
    f = L.cface[t];
    int32_t s = L.fptr[f], e = L.fptr[f + 1];
    tn = (t + 1 == e) ? s : t + 1;
    tp = (t == s) ? e - 1 : t - 1;
  }
}

// Static loop helper: CAP > 0 unrolls with a guard (register arrays stay in
// VGPRs), CAP == 0 runs a plain dynamic loop.
template <int CAP, class Fn>
__device__ __forceinline__ void rloop(int n, Fn&& fn) {
  if constexpr (CAP > 0) {
#pragma unroll
    for (int i = 0; i < CAP; ++i)
      if (i < n) fn(i);
  } else {
    for (int i = 0; i < n; ++i) fn(i);
  }
}

// One row of M^T held in registers: corners sorted ascending, with face,
// successor vertex (out-neighbour), predecessor vertex (in-neighbour) and the
// predecessor's corner; jm/im pair each out-half-edge with its twin.
template <int CAP, bool QUAD>
struct RegRow {
  static constexpr int kCap = CAP;
  int n;
  int32_t t_[CAP], f_[CAP], w_[CAP], u_[CAP], tp_[CAP];
  int8_t jm_[CAP], im_[CAP];
  int32_t fj_[CAP], tpj_[CAP];  // face / predecessor corner of the twin of out-half-edge i
  int8_t rko_[CAP], rki_[CAP];  // LAK: rank of w_i / u_j among the owned lower neighbours

  __device__ __forceinline__ void load(const LevelView& L, int32_t s, int cnt) { load(L, L.tids + s, cnt); }
  // row: the row's unsorted corners (global memory, or the block's LDS copy)
  __device__ __forceinline__ void load(const LevelView& L, const int32_t* row, int cnt) {
    n = cnt;
#pragma unroll
    for (int i = 0; i < CAP; ++i) t_[i] = (i < n) ? row[i] : 0x7fffffff;
    // odd-even transposition network: canonical (ascending corner) order
#pragma unroll
    for (int p = 0; p < CAP; ++p) {
#pragma unroll
      for (int i = p & 1; i + 1 < CAP; i += 2) {
        int32_t a = t_[i], b = t_[i + 1];
        t_[i] = a < b ? a : b;
        t_[i + 1] = a < b ? b : a;
      }
    }
#pragma unroll
    for (int i = 0; i < CAP; ++i) {
      if (i < n) {
        int32_t f, tn, tp;
        corner_geo<QUAD>(L, t_[i], f, tn, tp);
        f_[i] = f;
        tp_[i] = tp;
        if (QUAD) {  // the corner's whole quad in one 16-byte load
          const int4 q = reinterpret_cast<const int4*>(L.ids)[f];
          const int k = t_[i] & 3;
          w_[i] = k == 0 ? q.y : k == 1 ? q.z : k == 2 ? q.w : q.x;
          u_[i] = k == 0 ? q.w : k == 1 ? q.x : k == 2 ? q.y : q.z;
        } else {
          w_[i] = L.ids[tn];
          u_[i] = L.ids[tp];
        }
      } else {
        f_[i] = 0;
        tp_[i] = 0;
        w_[i] = -1 - i;
        u_[i] = -100 - i;
      }
    }
    pair();
  }
  // The same row from the symbolic pass's output: corners already sorted,
  // each with its in-neighbour, so only the out-neighbours are gathered.
  __device__ __forceinline__ void load_sorted(const LevelView& L, const int2* __restrict__ row, int cnt) {
    n = cnt;
#pragma unroll
    for (int i = 0; i < CAP; ++i) {
      const int2 e = (i < n) ? row[i] : make_int2(0x7fffffff, -100 - i);
      t_[i] = e.x;
      u_[i] = e.y;
    }
#pragma unroll
    for (int i = 0; i < CAP; ++i) {
      int32_t f = 0, tn = 0, tp = 0;
      if (i < n) corner_geo<QUAD>(L, t_[i], f, tn, tp);
      f_[i] = f;
      tp_[i] = tp;
      w_[i] = (i < n) ? L.ids[tn] : -1 - i;
    }
    pair();
  }
  // jm / im: each out-half-edge's twin among the in-half-edges and back
  __device__ __forceinline__ void pair() {
#pragma unroll
    for (int i = 0; i < CAP; ++i) {
      jm_[i] = -1;
      im_[i] = -1;
      fj_[i] = 0;
      tpj_[i] = 0;
    }
#pragma unroll
    for (int i = 0; i < CAP; ++i)
#pragma unroll
      for (int j = 0; j < CAP; ++j) {
        bool hit = (i < n) && (j < n) && (u_[j] == w_[i]);
        if (hit && jm_[i] < 0) {
          jm_[i] = (int8_t)j;
          fj_[i] = f_[j];
          tpj_[i] = tp_[j];
        }
        if (hit && im_[j] < 0) im_[j] = (int8_t)i;
      }
  }
  // LAK owns {x, v} for x < v: out-neighbours below v and in-only neighbours
  // below v; their ranks in ascending x are the E column v positions.
  __device__ __forceinline__ void rank_lower(int32_t v) {
    int32_t co[CAP], ci[CAP];
#pragma unroll
    for (int i = 0; i < CAP; ++i) {
      co[i] = (i < n && w_[i] < v) ? w_[i] : 0x7fffffff;
      ci[i] = (i < n && u_[i] < v && im_[i] < 0) ? u_[i] : 0x7fffffff;
    }
#pragma unroll
    for (int i = 0; i < CAP; ++i) {
      int a = 0, b = 0;
#pragma unroll
      for (int k = 0; k < CAP; ++k) {
        a += (co[k] < co[i]) + (ci[k] < co[i]);
        b += (co[k] < ci[i]) + (ci[k] < ci[i]);
      }
      rko_[i] = (int8_t)a;
      rki_[i] = (int8_t)b;
    }
  }
  __device__ __forceinline__ int RKO(int i) const { return rko_[i]; }
  __device__ __forceinline__ int RKI(int j) const { return rki_[j]; }
  __device__ __forceinline__ int32_t T(int i) const { return t_[i]; }
  __device__ __forceinline__ int32_t F(int i) const { return f_[i]; }
  __device__ __forceinline__ int32_t W(int i) const { return w_[i]; }
  __device__ __forceinline__ int32_t U(int i) const { return u_[i]; }
  __device__ __forceinline__ int32_t TP(int i) const { return tp_[i]; }
  __device__ __forceinline__ int JM(int i) const { return jm_[i]; }
  __device__ __forceinline__ int32_t FJ(int i) const { return fj_[i]; }
  __device__ __forceinline__ int32_t TPJ(int i) const { return tpj_[i]; }
  __device__ __forceinline__ int IM(int j) const { return im_[j]; }
};

// The same row read from global memory (rows longer than the register
// capacity). The row must already be sorted in tids.
template <bool QUAD>
struct MemRow {
  static constexpr int kCap = 0;
  const int32_t* ids;
  const int32_t* fptr;
  const int32_t* cface;
  const int32_t* tids;
  int32_t s;
  int n;
  __device__ __forceinline__ MemRow(const LevelView& L, int32_t s_, int n_)
      : ids(L.ids), fptr(L.fptr), cface(L.cface), tids(L.tids), s(s_), n(n_) {}
  __device__ __forceinline__ void geo(int i, int32_t& f, int32_t& tn, int32_t& tp) const {
    int32_t t = tids[s + i];
    if (QUAD) {
      f = t >> 2;
      int32_t b = t & ~3;
      tn = b | ((t + 1) & 3);
      tp = b | ((t + 3) & 3);
    } else {
      f = cface[t];
      int32_t a = fptr[f], e = fptr[f + 1];
      tn = (t + 1 == e) ? a : t + 1;
      tp = (t == a) ? e - 1 : t - 1;
    }
  }
  __device__ __forceinline__ int32_t T(int i) const { return tids[s + i]; }
  __device__ __forceinline__ int32_t F(int i) const {
    int32_t f, a, b;
    geo(i, f, a, b);
    return f;
  }
  __device__ __forceinline__ int32_t W(int i) const {
    int32_t f, a, b;
    geo(i, f, a, b);
    return ids[a];
  }
  __device__ __forceinline__ int32_t U(int i) const {
    int32_t f, a, b;
    geo(i, f, a, b);
    return ids[b];
  }
  __device__ __forceinline__ int32_t TP(int i) const {
    int32_t f, a, b;
    geo(i, f, a, b);
    return b;
  }
  __device__ int JM(int i) const {
    int32_t x = W(i);
    for (int j = 0; j < n; ++j)
      if (U(j) == x) return j;
    return -1;
  }
  __device__ int IM(int j) const {
    int32_t x = U(j);
    for (int i = 0; i < n; ++i)
      if (W(i) == x) return i;
    return -1;
  }
  __device__ __forceinline__ void rank_lower(int32_t) {}
  __device__ int32_t FJ(int i) const { return F(JM(i)); }
  __device__ int32_t TPJ(int i) const { return TP(JM(i)); }
  __device__ int rank_of(int32_t x) const {
    int k = 0;
    for (int j = 0; j < n; ++j) {
      int32_t u = U(j);
      k += W(j) < x;
      k += (u < x) && (IM(j) < 0);
    }
    return k;
  }
  __device__ int RKO(int i) const { return rank_of(W(i)); }
  __device__ int RKI(int j) const { return rank_of(U(j)); }
};

// Symbolic pass of the mapped SpGEMM for vertex row v: counts the edges v
// owns and validates the row.
//   LAK  owns the edges {x, v} with x < v (column v of E, spgemm.cuh:71-100);
//   SLAK owns its flagged out-half-edges v->w: twin present and v < w, or no
//        twin (CatmullClarkKernels.cuh:147-236).
template <int SCHEME, class Row>
__device__ __forceinline__ void row_symbolic(const Row& r, int32_t v, int32_t& count, int32_t& bcount, int& bad) {
  constexpr int C = Row::kCap;
  int c = 0, b = 0;
  rloop<C>(r.n, [&](int i) __attribute__((always_inline)) {
    int32_t wi = r.W(i), ui = r.U(i);
    if (wi == v || ui == v) bad |= kFlagDegenerate;
    rloop<C>(r.n, [&](int j) __attribute__((always_inline)) {
      if (j > i) {
        if (r.W(j) == wi) bad |= kFlagDupHalfEdge;
        if (r.U(j) == ui) bad |= kFlagDupHalfEdge;
      }
      if (r.U(j) == wi && r.F(j) == r.F(i)) bad |= kFlagDegenerate;
    });
    if (SCHEME == kLAK) {
      if (wi < v) {
        ++c;
        b += r.JM(i) < 0;
      }
      if (ui < v && r.IM(i) < 0) {
        ++c;
        ++b;
      }
    } else {
      int jm = r.JM(i);
      if (jm < 0 || wi > v) ++c;
      b += jm < 0;
    }
  });
  count = c;
  bcount = b;
}

// Number pass of the first level: the mesh-local id of every edge v owns
// (LAK: column v of E, spgemm.cuh:71-100; SLAK: flagged out-half-edges in
// corner order, CatmullClarkKernels.cuh:147-236) and the twin corner, stored
// as (edge id, twin) at both of its half-edges' corners, and the boundary
// flag of every owned edge by union edge index.
template <int SCHEME, class Row>
__device__ __forceinline__ void row_number(Row& r, const LevelView& L, int32_t v) {
  constexpr int C = Row::kCap;
  const int k = seg_of_wave(L.seg.vbase, L.seg.n, v);
  const SegDesc* __restrict__ S = L.seg.d + k;
  const int64_t ne_cap = S->ne;
  const int64_t eb = S->eb;
  int32_t* __restrict__ ci = L.ci;
  int32_t* __restrict__ bnd = L.bnd;
  int* flags = L.flags;
  const int32_t e0 = L.eptr[v] - L.eptr[S->vb];  // mesh-local edge base of v
  const int n = r.n;
  // Edge ids beyond the segment's edges are clamped (and flagged) so that
  // every later index stays inside the mesh's ranges.
  auto cap = [flags, ne_cap](int32_t e) __attribute__((always_inline)) {
    if (e >= ne_cap) {
      atomicOr(flags, kFlagOverflow);
      return (int32_t)(ne_cap > 0 ? ne_cap - 1 : 0);
    }
    return e;
  };
  // The owner of an edge writes both of its half-edges' (edge id, twin)
  // pairs whole: every corner's pair is written once, by one 8-byte store.
  int2* __restrict__ ci2 = reinterpret_cast<int2*>(ci);
  auto put = [ci2, bnd, eb](int32_t t, int32_t tw, int32_t e, bool boundary) __attribute__((always_inline)) {
    ci2[t] = make_int2(e, tw);
    if (tw >= 0) ci2[tw] = make_int2(e, t);
    if (bnd) bnd[eb + e] = boundary ? 1 : 0;
  };
  if (SCHEME == kLAK) {
    r.rank_lower(v);
    rloop<C>(n, [&](int i) __attribute__((always_inline)) {
      if (r.W(i) < v) {
        const int jm = r.JM(i);
        put(r.T(i), jm >= 0 ? r.TPJ(i) : -1, cap(e0 + r.RKO(i)), jm < 0);
      }
      if (r.U(i) < v && r.IM(i) < 0)  // boundary edge seen only as u -> v
        put(r.TP(i), -1, cap(e0 + r.RKI(i)), true);
    });
  } else {
    int kk = 0;
    rloop<C>(n, [&](int i) __attribute__((always_inline)) {
      const int jm = r.JM(i);
      if (jm < 0 || r.W(i) > v) {
        put(r.T(i), jm >= 0 ? r.TPJ(i) : -1, cap(e0 + kk), jm < 0);
        ++kk;
      }
    });
  }
}

// One level's refinement inputs and the next level's structure.
constexpr int kMaxVSeg = 8;

struct RefineView {
  const int32_t* ids;
  const int32_t* fptr;   // poly levels
  const int32_t* cface;  // poly levels
  const float4* P;
  int32_t nv;
  int32_t nf;
  int64_t nnz;
  const int32_t* tptr;
  const int2* trow;      // (corner, in-neighbour), sorted by corner
  const int2* ci;        // (mesh-local edge id of the outgoing half-edge, twin corner or -1)
  const int32_t* blk;    // exclusive scan of the union edge boundary flags, or null: closed level
  const int32_t* mpos;   // M^T storage position of every corner (SLAK)
  SegTable seg;
  // next level (null when no mesh continues)
  int32_t* ntptr;
  int2* ntrow;
  int2* nci;
  int32_t* nbnd;         // next level's union edge boundary flags, or null: closed
  int32_t* nmpos;        // next level's M^T positions (SLAK)
  int32_t next_nv;       // next level's union vertices and corners (end of its M^T)
  int32_t next_nnz;
  int32_t next_rows_x;   // 1: the next level's M^T rows are corners only (int32 in ntrow)
  // one mesh: its vertex items as up to kMaxVSeg class runs [vseg[s],
  // vseg[s+1]) of vnb[s] blocks, each run spread in proportion over the
  // launch (nvseg 0: one run)
  int32_t nvseg;
  int32_t vseg[kMaxVSeg + 1];
  int32_t vnb[kMaxVSeg];
};

// One mesh of a final level computed from the structure of the quad mesh M
// two levels before the output and the positions of its refinement M1 (whose
// structure is never written): M's bases in the union arrays of its level,
// its sizes, its M1 positions and its outputs.
constexpr int kBlockItems = 256;  // threads (items) per block of the level kernels

struct FinalSeg {
  int64_t cb, vb, eb;    // union bases of M's corners, vertices, edges
  int32_t nv, nf, ne;    // M sizes
  int32_t nnz;
  int32_t id_off;        // added to the output vertex ids (pooled outputs)
  const float4* P1;      // M1 positions = [M vertices | M face points | M edge points]
  float4* out_verts;     // output positions: [M1 vertices | M1 face points | M1 edge points]
  int32_t* out_ids;      // output quads: 4 per M1 corner, 16 per M corner
  int32_t* out_ptr;      // LAK form of the output (ptr[i] = 4i, vals cyclic 1..4), or null
  int32_t* out_vals;
};

struct FinalView {
  // union arrays of M's level
  const int32_t* ids;    // corners -> vertices
  const int2* ci;        // (mesh-local edge id, twin corner or -1) per corner
  const int32_t* tptr;   // M^T row offsets
  const int2* trow;      // M^T rows (corner, in-neighbour), sorted
  const int32_t* trow_x; // or the rows' corners alone (int32), or null
  const int32_t* blk;    // exclusive scan of the boundary edge flags, or null (closed)
  const int32_t* mpos;   // M^T position per corner (SLAK)
  FinalSeg one;          // a single mesh (nseg == 1, no table)
  const FinalSeg* segs;  // several meshes: device table
  // several meshes: work units of one block each, (mesh, first mesh-local
  // item, count, kind): kind 0 consecutive corners of one mesh, kind 1
  // consecutive vertices of one class of one mesh. Mesh by mesh, the corner
  // units and every vertex class's units are interleaved by their relative
  // position, so a mesh's vertex items run beside its corner items.
  const int4* units;
  int32_t nunits;
  int32_t nseg;
  int64_t nnz, nv;       // corner and vertex items of the launch
  // the vertex items as up to kMaxVSeg runs [vseg[s], vseg[s+1]) of vnb[s]
  // blocks, each run spread in proportion over the launch (nvseg 0: one run)
  int32_t nvseg;
  int32_t vseg[kMaxVSeg + 1];
  int32_t vnb[kMaxVSeg];
};

// Face, successor and predecessor corner of corner t (quad: bit arithmetic;
// poly: through cface / fptr).
template <bool QUAD>
__device__ __forceinline__ void corner_nbrs(const int32_t* __restrict__ fptr, const int32_t* __restrict__ cface,
                                            int32_t t, int32_t& f, int32_t& tn, int32_t& tp) {
  if (QUAD) {
    f = t >> 2;
    const int32_t b = t & ~3;
    tn = b | ((t + 1) & 3);
    tp = b | ((t + 3) & 3);
  } else {
    f = cface[t];
    const int32_t s = fptr[f], e = fptr[f + 1];
    tn = (t + 1 == e) ? s : t + 1;
    tp = (t == s) ? e - 1 : t - 1;
  }
}

// Mapped SpGEMM operator for row v (spgemm_mapped + compress): F row =
// out-neighbours w ascending with value f+1 (v->w in f), D row = in-neighbours
// u ascending with value f+1 (u->v in f), E column v = owned lower
// neighbours ascending with -(id+1) interior / +(id+1) boundary.
template <class Row>
__device__ __forceinline__ void row_fde(Row& r, const LevelView& L, int32_t v) {
  constexpr int C = Row::kCap;
  const int32_t s = L.tptr[v];
  const int32_t e0 = L.eptr[v];
  const int64_t ne_cap = L.ne_cap;
  int32_t* __restrict__ F_ids = L.F_ids;
  int32_t* __restrict__ F_vals = L.F_vals;
  int32_t* __restrict__ D_ids = L.D_ids;
  int32_t* __restrict__ D_vals = L.D_vals;
  int32_t* __restrict__ E_ids = L.E_ids;
  int32_t* __restrict__ E_vals = L.E_vals;
  const int n = r.n;
  r.rank_lower(v);
  rloop<C>(n, [&](int i) __attribute__((always_inline)) {
    int32_t w = r.W(i), u = r.U(i);
    int kw = 0, ku = 0;
    rloop<C>(n, [&](int j) __attribute__((always_inline)) {
      kw += r.W(j) < w;
      ku += r.U(j) < u;
    });
    F_ids[s + kw] = w;
    F_vals[s + kw] = r.F(i) + 1;
    D_ids[s + ku] = u;
    D_vals[s + ku] = r.F(i) + 1;
    if (w < v) {
      int32_t e = e0 + r.RKO(i);
      if (e < ne_cap) {
        E_ids[e] = w;
        E_vals[e] = r.JM(i) >= 0 ? -(e + 1) : (e + 1);
      }
    }
    if (u < v && r.IM(i) < 0) {
      int32_t e = e0 + r.RKI(i);
      if (e < ne_cap) {
        E_ids[e] = u;
        E_vals[e] = e + 1;
      }
    }
  });
}

}  // namespace slak
