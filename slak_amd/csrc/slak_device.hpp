// Device-side building blocks of the Catmull-Clark level pipeline (gfx950).
//
// A level runs over a union of meshes ("segments"): mesh k owns the vertex
// range [vb_k, vb_k + nv_k), faces [fb_k, ...) and corners [cb_k, ...) of the
// level's arrays, which hold global (union) indices. Data layout in HBM:
//   ids[nnz]           position-ordered corner -> vertex (quad levels: 4 per face, int4-aligned)
//   fptr[nf+1], cface[nnz]   face offsets and corner -> face (poly levels only)
//   P[nv]              float4 positions (x, y, z, w=1)
//   tptr[nv+1], tids[nnz]    M^T: vertex -> incident corners (corner index t)
//   eptr[nv+1]         per-vertex offset of the edges the vertex owns (global scan)
//   eout[nnz]          mesh-local edge id of every corner's outgoing half-edge
// Outputs go to each segment's own destination: refined positions
// [vertex points | face points | edge points] and refined quads in the
// mesh's local numbering (+ the segment's id offset when the result feeds the
// next level's union).
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

// One mesh of a level's union.
struct SegDesc {
  int64_t vb, fb, cb;      // union bases: vertices, faces, corners
  int32_t nv, nf;          // local sizes
  int64_t nnz;             // local corners
  int64_t ne;              // local edges (edge points the destination holds)
  float4* dst_verts;       // nv + nf + ne refined positions
  int32_t* dst_ids;        // 4 * nnz refined quad ids
  int32_t* dst_ptr;        // LAK form (final level) or null
  int32_t* dst_vals;       // LAK form (final level) or null
  int64_t id_offset;       // added to local refined ids
  // next level's union bases of this mesh (-1: the mesh finishes here); the
  // level writes the refined mesh's M^T rows there directly
  int64_t next_vb, next_cb;
  // first-level sources (the caller's mesh)
  const int32_t* src_ptr;  // null: homogeneous with src_stride
  const int32_t* src_ids;
  const int32_t* src_vals; // null: storage order
  const float4* src_verts;
  int32_t src_stride;
  int32_t require_quad;
};

struct SegTable {
  const SegDesc* d;
  const int64_t* vbase;  // n + 1 vertex bases
  const int64_t* fbase;  // n + 1 face bases
  int32_t n;
  int* mflags;           // per-segment validation flags
};

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

struct LevelView {
  const int32_t* ids;
  const int32_t* fptr;   // poly levels
  const int32_t* cface;  // poly levels
  const float4* P;
  int32_t nv;
  int32_t nf;
  int64_t nnz;
  const int32_t* tptr;
  int32_t* tids;
  int32_t* ecnt;
  const int32_t* eptr;
  int32_t* eout;
  int32_t* E_ids;   // optional LAK edge matrix rows (single segment)
  int32_t* E_vals;  // optional LAK edge matrix values (+-(id+1))
  int* flags;
  SegTable seg;
  // refined incidence for the next level (null when no mesh continues)
  int32_t* ntptr;
  int32_t* ntids;
  int32_t* bcnt;         // owned boundary edges per vertex (symbolic), or null
  const int32_t* bptr;   // its exclusive scan (numeric), or null: no boundary
  // mapped SpGEMM operator outputs (F, D rows share M^T's row offsets)
  int32_t* F_ids;
  int32_t* F_vals;
  int32_t* D_ids;
  int32_t* D_vals;
  int64_t ne_cap;  // E capacity of the SpGEMM operator
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

  __device__ __forceinline__ void load(const LevelView& L, int32_t s, int cnt) {
    n = cnt;
#pragma unroll
    for (int i = 0; i < CAP; ++i) t_[i] = (i < n) ? L.tids[s + i] : 0x7fffffff;
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
        w_[i] = L.ids[tn];
        u_[i] = L.ids[tp];
      } else {
        f_[i] = 0;
        tp_[i] = 0;
        w_[i] = -1 - i;
        u_[i] = -100 - i;
      }
    }
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

// Numeric pass: edge ids, edge points and the vertex point of row v, plus the
// edge id of every half-edge of the edges v owns (eout).
template <int SCHEME, bool QUADT, class Row>
__device__ __forceinline__ void row_numeric(Row& r, const LevelView& L, int32_t v) {
  constexpr int C = Row::kCap;
  // local copies: lambdas below capture locals only, so the kernel argument
  // block never has its address taken (it would be spilled to scratch)
  const int k = seg_of(L.seg.vbase, L.seg.n, v);
  const SegDesc* __restrict__ S = L.seg.d + k;
  const int64_t vb = S->vb, fb = S->fb;
  const int32_t snv = S->nv, snf = S->nf;
  const int64_t ne_cap = S->ne;
  float4* __restrict__ dst = S->dst_verts;
  const float4* __restrict__ P = L.P;
  auto FP = [dst, snv, fb](int32_t f) __attribute__((always_inline)) { return dst[snv + (f - fb)]; };
  float4* __restrict__ ep = dst + snv + snf;
  int32_t* __restrict__ eout = L.eout;
  int32_t* __restrict__ E_ids = L.E_ids;
  int32_t* __restrict__ E_vals = L.E_vals;
  int* flags = L.flags;
  const float4 Pv = P[v];
  const int32_t e0 = L.eptr[v] - L.eptr[vb];  // mesh-local edge base of v
  const int n = r.n;
  // refined incidence of the next level: local corner c of this level maps to
  // the refined quads 4c..4c+3, i.e. next-level corners ncb + 4c + j
  const int64_t ncb = S->next_cb, nvb = S->next_vb, cb = S->cb;
  const int64_t snnz = S->nnz;
  int32_t* __restrict__ ntptr = L.ntptr;
  int32_t* __restrict__ ntids = L.ntids;
  const int32_t* __restrict__ cface = L.cface;
  const int32_t* __restrict__ fptr = L.fptr;
  const int32_t bb = L.bptr ? L.bptr[v] - L.bptr[vb] : 0;  // boundary edges owned by earlier vertices
  auto next_corner = [cface, fptr](int32_t t) __attribute__((always_inline)) {
    if (QUADT) {
      return (t & ~3) | ((t + 1) & 3);
    } else {
      int32_t f = cface[t];
      return (t + 1 == fptr[f + 1]) ? fptr[f] : t + 1;
    }
  };
  // row of edge point e: corners (4 ta + 1, 4 next(ta) + 3[, 4 tb + 1, 4 next(tb) + 3]), sorted
  auto edge_row = [&](int32_t e, int32_t bl, int32_t ta, int32_t tb) __attribute__((always_inline)) {
    if (ncb < 0 || e >= ne_cap) return;
    const int64_t row = ncb + 2 * snnz + 4 * (int64_t)e - 2 * (int64_t)bl;
    ntptr[nvb + snv + snf + e] = (int32_t)row;
    int32_t a0 = 4 * (int32_t)(ta - cb) + 1, a1 = 4 * (int32_t)(next_corner(ta) - cb) + 3;
    if (a1 < a0) {
      int32_t x = a0;
      a0 = a1;
      a1 = x;
    }
    if (tb < 0) {
      ntids[row] = (int32_t)(ncb + a0);
      ntids[row + 1] = (int32_t)(ncb + a1);
      return;
    }
    int32_t b0 = 4 * (int32_t)(tb - cb) + 1, b1 = 4 * (int32_t)(next_corner(tb) - cb) + 3;
    if (b1 < b0) {
      int32_t x = b0;
      b0 = b1;
      b1 = x;
    }
    // merge two sorted pairs
    int32_t m0 = min(a0, b0), m3 = max(a1, b1);
    int32_t lo = max(a0, b0), hi = min(a1, b1);
    int32_t m1 = min(lo, hi), m2 = max(lo, hi);
    ntids[row] = (int32_t)(ncb + m0);
    ntids[row + 1] = (int32_t)(ncb + m1);
    ntids[row + 2] = (int32_t)(ncb + m2);
    ntids[row + 3] = (int32_t)(ncb + m3);
  };
  // LAK: boundary edges among v's owned edges with a smaller neighbour than x
  auto lak_brank = [&](int32_t x) __attribute__((always_inline)) {
    int k = 0;
    rloop<C>(n, [&](int j) __attribute__((always_inline)) {
      k += (r.W(j) < x) && (r.JM(j) < 0);
      k += (r.U(j) < x) && (r.IM(j) < 0);
    });
    return k;
  };
  // Edge ids beyond the segment's edge points are clamped (and flagged) so
  // that refined ids stay inside the refined vertex range.
  auto cap = [flags, ne_cap](int32_t e) __attribute__((always_inline)) {
    if (e >= ne_cap) {
      atomicOr(flags, kFlagOverflow);
      return (int32_t)(ne_cap > 0 ? ne_cap - 1 : 0);
    }
    return e;
  };

  if (SCHEME == kLAK) {
    r.rank_lower(v);
    rloop<C>(n, [&](int i) __attribute__((always_inline)) {
      int32_t x = r.W(i);
      if (x < v) {
        int32_t e = cap(e0 + r.RKO(i));
        int jm = r.JM(i);
        float4 q;
        if (jm >= 0)  // interior: f0 = face of x->v, f1 = face of v->x
          q = f4mul(f4add(f4add(f4add(P[x], Pv), FP(r.FJ(i))), FP(r.F(i))), 0.25f);
        else
          q = f4mul(f4add(P[x], Pv), 0.5f);
        if (e < ne_cap) ep[e] = q;
        eout[r.T(i)] = e;
        if (jm >= 0) eout[r.TPJ(i)] = e;
        if (E_ids && e < ne_cap) {
          E_ids[e] = x;
          E_vals[e] = jm >= 0 ? -(e + 1) : (e + 1);
        }
        if (ncb >= 0) edge_row(e, bb + (L.bptr ? lak_brank(x) : 0), r.T(i), jm >= 0 ? r.TPJ(i) : -1);
      }
      int32_t y = r.U(i);
      if (y < v && r.IM(i) < 0) {  // boundary edge seen only as y -> v
        int32_t e = cap(e0 + r.RKI(i));
        if (e < ne_cap) ep[e] = f4mul(f4add(P[y], Pv), 0.5f);
        eout[r.TP(i)] = e;
        if (E_ids && e < ne_cap) {
          E_ids[e] = y;
          E_vals[e] = e + 1;
        }
        if (ncb >= 0) edge_row(e, bb + (L.bptr ? lak_brank(y) : 0), r.TP(i), -1);
      }
    });
  } else {
    int kk = 0, kb = 0;
    rloop<C>(n, [&](int i) __attribute__((always_inline)) {
      int32_t x = r.W(i);
      int jm = r.JM(i);
      if (jm < 0 || x > v) {
        int32_t e = cap(e0 + kk);
        ++kk;
        if (ncb >= 0) edge_row(e, bb + kb, r.T(i), jm >= 0 ? r.TPJ(i) : -1);
        kb += jm < 0;
        float4 q;
        if (jm >= 0)  // interior, v < x: f0 = face of v->x, f1 = face of x->v
          q = f4mul(f4add(f4add(f4add(Pv, P[x]), FP(r.F(i))), FP(r.FJ(i))), 0.25f);
        else
          q = (v < x) ? f4mul(f4add(Pv, P[x]), 0.5f) : f4mul(f4add(P[x], Pv), 0.5f);
        if (e < ne_cap) ep[e] = q;
        eout[r.T(i)] = e;
        if (jm >= 0) eout[r.TPJ(i)] = e;
      }
    });
  }

  // vertex point (LAKCatmullClarkImpl.cu:200-235, pull form)
  bool bnd = false;
  rloop<C>(n, [&](int i) __attribute__((always_inline)) { bnd |= (r.JM(i) < 0) || (r.IM(i) < 0); });
  float4 q;
  if (!bnd) {
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    rloop<C>(n, [&](int i) __attribute__((always_inline)) { s = f4add(s, FP(r.F(i))); });
    rloop<C>(n, [&](int j) __attribute__((always_inline)) { s = f4add(s, P[r.U(j)]); });
    float nf = (float)n;
    q = f4div(f4add(f4div(s, nf), f4mul(Pv, nf - 2.0f)), nf);
  } else {
    q = f4mul(Pv, 0.75f);
    rloop<C>(n, [&](int i) __attribute__((always_inline)) {
      if (r.JM(i) < 0) q = f4add(q, f4mul(P[r.W(i)], 0.125f));
    });
    rloop<C>(n, [&](int j) __attribute__((always_inline)) {
      if (r.IM(j) < 0) q = f4add(q, f4mul(P[r.U(j)], 0.125f));
    });
  }
  dst[v - vb] = q;
  if (ncb >= 0) {  // row of the vertex itself: corner 0 of each refined quad at v
    const int64_t row = ncb + (L.tptr[v] - cb);
    ntptr[nvb + (v - vb)] = (int32_t)row;
    rloop<C>(n, [&](int i) __attribute__((always_inline)) {
      ntids[row + i] = (int32_t)(ncb + 4 * (r.T(i) - cb));
    });
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
