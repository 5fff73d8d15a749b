// HIP kernels of the Catmull-Clark level pipeline for gfx950 (MI355X).
//
// First level (the caller's mesh):
//   validate      check the input and bring corners into position order
//                 (sortKeyValuePairsSegmentedInPlace, LAKCatmullClarkKernels.cuh:147-180)
//   count         valence per vertex (spmv_right with ones, LAKCatmullClarkImpl.cu:347-356)
//   scan          exclusive scan (cub_wrappers.cuh:30-43)
//   transpose     M -> M^T by counting sort (csr2cscEx2, LAKCatmullClarkImpl.cu:381-423)
//   symbolic      row sizing of the mapped SpGEMM -> edges owned per vertex;
//                 rows sorted and stored with their in-neighbours
//   number        mapped SpGEMM rows -> E numbering: edge id and twin of
//                 every corner's half-edge, boundary flags
// Every level:
//   face points   left mapped SpMV fp = M^T P / order (spmv.cuh:64-81)
//   refine        per corner: edge points (calculateInternal/ExternalEdgepoints),
//                 refined quads (refineTopology[Homogeneous],
//                 LAKCatmullClarkKernels.cuh:320-431) and, for a level that
//                 feeds another, the refined mesh's twins, edge ids and the
//                 face- and edge-point rows of its M^T, all in closed form
//   vertex        per vertex: the two right SpMVs of the vertex update
//                 (LAKCatmullClarkImpl.cu:200-235) and boundary rule
//                 (LAKCatmullClarkKernels.cuh:464-511); the vertex rows of the
//                 refined M^T
// Last two levels of a quad chain:
//   final         from the quad mesh M two levels back: the output quads,
//                 edge, face and vertex points of the level (or its positions
//                 only, for the level before it); several meshes per launch
//                 through work-unit tables
//   LAK form      several meshes: the output level's ptr / vals in a launch
//                 of their own (createQuadColPtr / createQuadVals)
// All kernels are pull-based and atomic-free except the integer valence
// count and transpose slots, whose result does not depend on atomic order,
// so every output is bitwise deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "slak_device.hpp"
#include "slak_internal.hpp"

namespace slak {

namespace {

constexpr int kBlock = kBlockItems;
constexpr int kRowCap = 8;  // M^T rows up to this length stay in registers

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// With the remap, consecutive tiles (which gather from neighbouring parts of
// the mesh) land on the same XCD: XCD x runs the x-th eighth of the grid in
// order. Selected once per process by the bit mask SLAK_XCD_REMAP: bit 0 for
// the level kernels that hand on a structure, bit 1 for k_final (its gathers
// then stay in one L2: 136 -> 113 us on the 4-level chain of the 256x128
// torus), bit 2 for the other level kernels, bit 3 for every other kernel.
// For the level kernels it pays together with vertex items in class runs
// (alone it measured 10-25 % slower). Default 3.
__device__ int g_xcd_remap;
constexpr int kRemapLevel = 1, kRemapFinal = 2, kRemapPositions = 4, kRemapOther = 8;

template <int BIT = kRemapOther>
__device__ __forceinline__ int64_t block_index() {
  const int b = blockIdx.x;
  if (!(g_xcd_remap & BIT)) return b;
  const int nb = gridDim.x;
  const int per = nb >> 3, rem = nb & 7, x = b & 7, i = b >> 3;
  return (int64_t)((x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + i);
}

inline unsigned grid_for(int64_t n, int block = kBlock) {
  int64_t g = (n + block - 1) / block;
  return (unsigned)(g < 1 ? 1 : g);
}

// ---------------------------------------------------------------------------
// fills

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t value) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = value;
}

// ---------------------------------------------------------------------------
// HBM reference: a float4 stream copy (bandwidth probe and counter
// calibration), one float4 per lane over a grid that covers the buffer (on
// MI355X this reaches about 6.3 TB/s read + write; a grid-stride loop over a
// fixed grid stays near 4.5)

__global__ void __launch_bounds__(kBlock) k_copy_probe(const float4* __restrict__ a, float4* __restrict__ b,
                                                        int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// Order-sensitive 64-bit digest of a byte buffer: the sum (mod 2^64) over its
// 32-bit words w_i (the last one zero-padded) of mix64(w_i + (i + 1) * G), a
// splitmix64 finaliser. A sum is independent of the reduction order, so the
// digest is deterministic; each block adds its partial to *out (zeroed first).
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

constexpr int kDigestBlocks = 2048;

__global__ void __launch_bounds__(kBlock) k_digest(const uint8_t* __restrict__ p, int64_t bytes,
                                                   unsigned long long* __restrict__ out) {
  const int64_t nw = (bytes + 3) / 4, full = bytes / 4;
  const uint32_t* __restrict__ w = reinterpret_cast<const uint32_t*>(p);
  uint64_t acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t x;
    if (i < full) {
      x = w[i];
    } else {  // the zero-padded last word
      x = 0;
      for (int64_t b = 4 * i; b < bytes; ++b) x |= (uint32_t)p[b] << (8 * (b - 4 * i));
    }
    acc += mix64((uint64_t)x + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull);
  }
  __shared__ uint64_t part[kBlock / 64];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t s = 0;
    for (int k = 0; k < kBlock / 64; ++k) s += part[k];
    atomicAdd(out, (unsigned long long)s);
  }
}

// ---------------------------------------------------------------------------
// first level: validate each caller mesh and gather it into the union in
// position order (sortKeyValuePairsSegmentedInPlace, LAKCatmullClarkKernels.cuh:147-180)

constexpr int kHistBins = kRowHistBins;  // row lengths 0..15, 16 = longer
static_assert(kHistBins <= kHistStride, "a histogram copy holds every bin");

// Adds 1 to bin `bin` of an LDS histogram for every calling lane, with one
// LDS atomic per distinct bin of the wave: same-address LDS atomics from the
// lanes of a wave serialize (on a quad mesh every lane has the same bin).
__device__ __forceinline__ void hist_add_wave(int32_t* lds, int bin) {
  unsigned long long todo = __ballot(1);
  while (todo) {
    const int leader = __ffsll(todo) - 1;
    const int b = __shfl(bin, leader);
    const unsigned long long same = __ballot(bin == b) & todo;
    if ((int)(threadIdx.x & 63) == leader) atomicAdd(&lds[b], __popcll(same));
    todo &= ~same;
  }
}

// Block-local histogram in LDS, added to copy blockIdx % kHistCopies of a
// zeroed global histogram block (one atomic per nonzero bin).
struct BlockHist {
  int32_t* lds;
  __device__ __forceinline__ void init() {
    for (int i = threadIdx.x; i < kHistBins; i += blockDim.x) lds[i] = 0;
    __syncthreads();
  }
  __device__ __forceinline__ void flush(int32_t* global) {
    __syncthreads();
    int32_t* copy = global + (blockIdx.x % kHistCopies) * kHistStride;
    for (int i = threadIdx.x; i < kHistBins; i += blockDim.x)
      if (lds[i]) atomicAdd(&copy[i], lds[i]);
  }
};

__device__ void gather_face(const SegTable& seg, int32_t f, int32_t nf_total, int use_vals, int32_t* __restrict__ ids_pos,
                            int32_t* __restrict__ fptr, int32_t* __restrict__ cface, int* __restrict__ flags,
                            int32_t* hist);

__global__ void __launch_bounds__(kBlock) k_gather(SegTable seg, int32_t nf_total, int use_vals, int32_t* __restrict__ ids_pos,
                         int32_t* __restrict__ fptr, int32_t* __restrict__ cface, int* __restrict__ flags,
                         int32_t* __restrict__ zero, int32_t nzero, int32_t* __restrict__ hist_faces) {
  __shared__ int32_t h[kHistBins];
  BlockHist bh{h};
  if (hist_faces) bh.init();
  const int32_t f = (int32_t)(block_index() * blockDim.x + threadIdx.x);
  if (f < nzero) zero[f] = 0;  // counters of the analysis that follows
  if (f < nf_total) gather_face(seg, f, nf_total, use_vals, ids_pos, fptr, cface, flags, hist_faces ? h : nullptr);
  if (hist_faces) bh.flush(hist_faces);
}

// first level: validate caller face f and gather it into the union in
// position order; face order into the block's LDS histogram (hist, or null)
__device__ void gather_face(const SegTable& seg, int32_t f, int32_t nf_total, int use_vals, int32_t* __restrict__ ids_pos,
                            int32_t* __restrict__ fptr, int32_t* __restrict__ cface, int* __restrict__ flags,
                            int32_t* hist) {
  const int k = seg_of_wave(seg.fbase, seg.n, f);
  const SegDesc* S = seg.d + k;
  const int32_t lf = (int32_t)(f - S->fb);
  const int32_t* sp = to_global(S->src_ptr);
  const int32_t* sid = to_global(S->src_ids);
  const int32_t* sval = use_vals ? to_global(S->src_vals) : nullptr;
  const int64_t snnz = S->nnz;
  int64_t s = sp ? sp[lf] : (int64_t)S->src_stride * lf;
  int64_t e = sp ? sp[lf + 1] : (int64_t)S->src_stride * (lf + 1);
  int bad = 0;
  if (lf == 0 && s != 0) bad |= kFlagBadPtr;
  if (lf == S->nf - 1 && e != snnz) bad |= kFlagBadPtr;
  if (s < 0 || e < s || e > snnz) {
    bad |= kFlagBadPtr;
    atomicOr(seg.mflags + k, bad);
    atomicOr(flags, bad);
    return;
  }
  const int32_t order = (int32_t)(e - s);
  if (hist) hist_add_wave(hist, order < kHistBins - 1 ? order : kHistBins - 1);
  if (order < 3) bad |= kFlagBadOrder;
  if (S->require_quad && (order != 4 || s != 4ll * lf)) bad |= kFlagBadQuad;
  const int64_t cb = S->cb;
  const int32_t vb = (int32_t)S->vb;
  if (order == 4 && ((reinterpret_cast<uintptr_t>(sid + s) | (sval ? reinterpret_cast<uintptr_t>(sval + s) : 0)) & 15) == 0) {
    // a quad with 16-byte aligned corners: one load of its ids (and vals),
    // one store when the union position is aligned too; anything unusual
    // takes the per-corner path below
    const int4 q = *reinterpret_cast<const int4*>(sid + s);
    const int4 p = sval ? *reinterpret_cast<const int4*>(sval + s) : make_int4(1, 2, 3, 4);
    const uint32_t in_range = (uint32_t)q.x < (uint32_t)S->nv && (uint32_t)q.y < (uint32_t)S->nv &&
                              (uint32_t)q.z < (uint32_t)S->nv && (uint32_t)q.w < (uint32_t)S->nv;
    const bool pos_ok = (uint32_t)(p.x - 1) < 4u && (uint32_t)(p.y - 1) < 4u && (uint32_t)(p.z - 1) < 4u &&
                        (uint32_t)(p.w - 1) < 4u;
    if (in_range && pos_ok && ((1u << (p.x - 1)) | (1u << (p.y - 1)) | (1u << (p.z - 1)) | (1u << (p.w - 1))) == 15u) {
      // the corner at position k (1-based) is the one whose vals entry is k
      auto at = [&](int k) __attribute__((always_inline)) {
        return vb + (p.x == k ? q.x : p.y == k ? q.y : p.z == k ? q.z : q.w);
      };
      const int4 o = make_int4(at(1), at(2), at(3), at(4));
      int32_t* dst = ids_pos + cb + s;
      if ((reinterpret_cast<uintptr_t>(dst) & 15) == 0) {
        *reinterpret_cast<int4*>(dst) = o;
      } else {
        dst[0] = o.x;
        dst[1] = o.y;
        dst[2] = o.z;
        dst[3] = o.w;
      }
      if (cface) {
        int32_t* cf = cface + cb + s;
        if ((reinterpret_cast<uintptr_t>(cf) & 15) == 0) {
          *reinterpret_cast<int4*>(cf) = make_int4(f, f, f, f);
        } else {
          cf[0] = f;
          cf[1] = f;
          cf[2] = f;
          cf[3] = f;
        }
      }
      if (fptr) {
        fptr[f] = (int32_t)(cb + s);
        if (f == nf_total - 1) fptr[nf_total] = (int32_t)(cb + e);
      }
      if (bad) {
        atomicOr(seg.mflags + k, bad);
        atomicOr(flags, bad);
      }
      return;
    }
  }
  // any other face: its first kGatherRegs corners are loaded before the
  // first store (a load issued after a store would wait for it), the rest
  // one by one
  constexpr int kGatherRegs = 8;
  const int32_t snv = S->nv;
  uint32_t mask = 0;
  auto corner = [&](int64_t c, int32_t id, int32_t pos) __attribute__((always_inline)) {
    if (id < 0 || id >= snv) {
      bad |= kFlagBadId;
      id = 0;
    }
    if (pos < 1 || pos > order) {
      bad |= kFlagBadVals;
      return;
    }
    if (order <= 32) {
      uint32_t bit = 1u << (pos - 1);
      if (mask & bit) bad |= kFlagBadVals;
      mask |= bit;
    } else {
      for (int64_t c2 = s; c2 < c; ++c2)
        if ((sval ? sval[c2] : (int32_t)(c2 - s + 1)) == pos) bad |= kFlagBadVals;
    }
    ids_pos[cb + s + pos - 1] = vb + id;
    if (cface) cface[cb + c] = f;
  };
  int32_t rid[kGatherRegs], rpos[kGatherRegs];
#pragma unroll
  for (int i = 0; i < kGatherRegs; ++i)
    if (i < order) {
      rid[i] = sid[s + i];
      rpos[i] = sval ? sval[s + i] : i + 1;
    }
#pragma unroll
  for (int i = 0; i < kGatherRegs; ++i)
    if (i < order) corner(s + i, rid[i], rpos[i]);
  for (int64_t c = s + kGatherRegs; c < e; ++c) corner(c, sid[c], sval ? sval[c] : (int32_t)(c - s + 1));
  if (fptr) {
    fptr[f] = (int32_t)(cb + s);
    if (f == nf_total - 1) fptr[nf_total] = (int32_t)(cb + e);
  }
  if (bad) {
    atomicOr(seg.mflags + k, bad);
    atomicOr(flags, bad);
  }
}

__global__ void k_gather_verts(SegTable seg, int32_t nv_total, float4* __restrict__ P) {
  int32_t v = (int32_t)(block_index() * blockDim.x + threadIdx.x);
  if (v >= nv_total) return;
  const int k = seg_of_wave(seg.vbase, seg.n, v);
  const SegDesc* S = seg.d + k;
  P[v] = to_global(S->src_verts)[v - S->vb];
}

// eptr at every segment's first vertex: bounds[k] = eptr[vbase[k]], k = 0..n
// (multi-mesh plans; a single mesh takes its total from the scan)
__global__ void k_seg_bounds(const int64_t* __restrict__ vbase, int32_t n, const int32_t* __restrict__ eptr,
                             int32_t* __restrict__ bounds) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k <= n) bounds[k] = eptr[vbase[k]];
}

// ---------------------------------------------------------------------------
// valence count (+ face points into each segment's destination)

template <bool QUAD, bool FP, bool COUNT>
__device__ __forceinline__ void faces_item(const int32_t* __restrict__ ids, const int32_t* __restrict__ fptr,
                                           const float4* __restrict__ P, int32_t nf, const SegTable& seg,
                                           int32_t* __restrict__ cnt, const int* __restrict__ guard, int32_t f) {
  if (f >= nf) return;
  if (guard && (*guard & kFlagsInput)) return;
  float4* dst = nullptr;
  if (FP) {
    const SegDesc* S = seg.d + seg_of_wave(seg.fbase, seg.n, f);
    dst = to_global(S->dst_verts) + S->nv + (f - S->fb);
  }
  if (QUAD) {
    int4 c = reinterpret_cast<const int4*>(ids)[f];
    if (FP) *dst = f4div(f4add(f4add(f4add(P[c.x], P[c.y]), P[c.z]), P[c.w]), 4.0f);
    if (COUNT) {
      atomicAdd(&cnt[c.x], 1);
      atomicAdd(&cnt[c.y], 1);
      atomicAdd(&cnt[c.z], 1);
      atomicAdd(&cnt[c.w], 1);
    }
  } else {
    int32_t s = fptr[f], e = fptr[f + 1];
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int32_t c = s; c < e; ++c) {
      int32_t v = ids[c];
      if (FP) acc = f4add(acc, P[v]);
      if (COUNT) atomicAdd(&cnt[v], 1);
    }
    if (FP) *dst = f4div(acc, (float)(e - s));
  }
}

template <bool QUAD, bool FP, bool COUNT>
__global__ void k_faces(const int32_t* __restrict__ ids, const int32_t* __restrict__ fptr, const float4* __restrict__ P,
                        int32_t nf, SegTable seg, int32_t* __restrict__ cnt, const int* __restrict__ guard) {
  faces_item<QUAD, FP, COUNT>(ids, fptr, P, nf, seg, cnt, guard, (int32_t)(block_index() * blockDim.x + threadIdx.x));
}

// ---------------------------------------------------------------------------
// Valence count and transpose through a block-local vertex window. The corners
// of a block's faces usually reference a narrow range of vertex ids (grids,
// refined meshes, each mesh of a batch): the block then counts into LDS over
// [vmin, vmax] and adds each touched vertex's count with one device atomic,
// and the transpose reserves each touched vertex's slots with one atomic and
// hands them out by LDS atomics. Device-scope atomics are what both passes
// cost (one per corner otherwise); a block whose window is wider than
// kCountWindow vertices takes one per corner.
constexpr int kCountWindow = 2048;

// min / max over the block (mm: kBlock / 64 LDS entries)
__device__ __forceinline__ int2 block_minmax(int lo, int hi, int2* mm) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    lo = min(lo, __shfl_xor(lo, d));
    hi = max(hi, __shfl_xor(hi, d));
  }
  if ((threadIdx.x & 63) == 0) mm[threadIdx.x >> 6] = make_int2(lo, hi);
  __syncthreads();
  int2 r = mm[0];
#pragma unroll
  for (int w = 1; w < kBlock / 64; ++w) r = make_int2(min(r.x, mm[w].x), max(r.y, mm[w].y));
  return r;
}

// the vertex range of face f's corners (quad: c holds them)
template <bool QUAD>
__device__ __forceinline__ int2 face_range(const int32_t* __restrict__ ids, const int32_t* __restrict__ fptr,
                                           int32_t f, int4& c, int32_t& s, int32_t& e) {
  if (QUAD) {
    c = reinterpret_cast<const int4*>(ids)[f];
    return make_int2(min(min(c.x, c.y), min(c.z, c.w)), max(max(c.x, c.y), max(c.z, c.w)));
  }
  s = fptr[f];
  e = fptr[f + 1];
  int lo = INT_MAX, hi = INT_MIN;
  for (int32_t k = s; k < e; ++k) {
    const int32_t v = ids[k];
    lo = min(lo, v);
    hi = max(hi, v);
  }
  return make_int2(lo, hi);
}

// cnt[v] += valence of v over the faces [0, nf) (cnt zeroed before)
template <bool QUAD>
__global__ void __launch_bounds__(kBlock) k_count(const int32_t* __restrict__ ids, const int32_t* __restrict__ fptr,
                                                  int32_t nf, int32_t* __restrict__ cnt, const int* __restrict__ guard) {
  __shared__ int32_t win[kCountWindow];
  __shared__ int2 mm[kBlock / 64];
  if (guard && (*guard & kFlagsInput)) return;  // block-uniform
  const int32_t f = (int32_t)(block_index() * blockDim.x + threadIdx.x);
  const bool act = f < nf;
  int4 c = make_int4(0, 0, 0, 0);
  int32_t s = 0, e = 0;
  int2 r = make_int2(INT_MAX, INT_MIN);
  if (act) r = face_range<QUAD>(ids, fptr, f, c, s, e);
  const int2 w = block_minmax(r.x, r.y, mm);
  const int64_t n = (int64_t)w.y - w.x + 1;  // <= 0 when no lane is active
  if (n <= kCountWindow) {                    // block-uniform
    for (int i = threadIdx.x; i < n; i += blockDim.x) win[i] = 0;
    __syncthreads();
    if (act) {
      if (QUAD) {
        atomicAdd(&win[c.x - w.x], 1);
        atomicAdd(&win[c.y - w.x], 1);
        atomicAdd(&win[c.z - w.x], 1);
        atomicAdd(&win[c.w - w.x], 1);
      } else {
        for (int32_t k = s; k < e; ++k) atomicAdd(&win[ids[k] - w.x], 1);
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const int32_t k = win[i];
      if (k) atomicAdd(&cnt[w.x + i], k);
    }
  } else if (act) {
    if (QUAD) {
      atomicAdd(&cnt[c.x], 1);
      atomicAdd(&cnt[c.y], 1);
      atomicAdd(&cnt[c.z], 1);
      atomicAdd(&cnt[c.w], 1);
    } else {
      for (int32_t k = s; k < e; ++k) atomicAdd(&cnt[ids[k]], 1);
    }
  }
}

// ---------------------------------------------------------------------------
// transpose scatter: tids[tptr[v] + slot] = corner; cnt[] counts back to zero
// (slot order within a row is arbitrary: the row-sizing pass sorts the rows)

template <bool QUAD>
__global__ void __launch_bounds__(kBlock) k_transpose(const int32_t* __restrict__ ids, const int32_t* __restrict__ fptr,
                                                      int32_t nf, const int32_t* __restrict__ tptr,
                                                      int32_t* __restrict__ cnt, int32_t* __restrict__ tids,
                                                      const int* __restrict__ guard, int window) {
  __shared__ int32_t lc[kCountWindow];    // the block's corners per window vertex, then counted back
  __shared__ int32_t base[kCountWindow];  // the first tids slot the block reserved for each
  __shared__ int2 mm[kBlock / 64];
  if (guard && (*guard & kFlagsInput)) return;  // block-uniform
  const int32_t f = (int32_t)(block_index() * blockDim.x + threadIdx.x);
  const bool act = f < nf;
  int4 c = make_int4(0, 0, 0, 0);
  int32_t s = 0, e = 0;
  int2 r = make_int2(INT_MAX, INT_MIN);
  if (act) r = face_range<QUAD>(ids, fptr, f, c, s, e);
  const int2 w = block_minmax(r.x, r.y, mm);
  const int64_t n = (int64_t)w.y - w.x + 1;
  if (window && n <= kCountWindow) {  // block-uniform
    for (int i = threadIdx.x; i < n; i += blockDim.x) lc[i] = 0;
    __syncthreads();
    if (act) {
      if (QUAD) {
        atomicAdd(&lc[c.x - w.x], 1);
        atomicAdd(&lc[c.y - w.x], 1);
        atomicAdd(&lc[c.z - w.x], 1);
        atomicAdd(&lc[c.w - w.x], 1);
      } else {
        for (int32_t k = s; k < e; ++k) atomicAdd(&lc[ids[k] - w.x], 1);
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const int32_t k = lc[i];
      if (k) base[i] = tptr[w.x + i] + atomicSub(&cnt[w.x + i], k) - k;
    }
    __syncthreads();
    if (act) {
      if (QUAD) {
        const int32_t t = 4 * f;
        tids[base[c.x - w.x] + atomicSub(&lc[c.x - w.x], 1) - 1] = t;
        tids[base[c.y - w.x] + atomicSub(&lc[c.y - w.x], 1) - 1] = t + 1;
        tids[base[c.z - w.x] + atomicSub(&lc[c.z - w.x], 1) - 1] = t + 2;
        tids[base[c.w - w.x] + atomicSub(&lc[c.w - w.x], 1) - 1] = t + 3;
      } else {
        for (int32_t k = s; k < e; ++k) {
          const int32_t i = ids[k] - w.x;
          tids[base[i] + atomicSub(&lc[i], 1) - 1] = k;
        }
      }
    }
    return;
  }
  if (!act) return;
  if (QUAD) {
    int4 c = reinterpret_cast<const int4*>(ids)[f];
    int32_t t = 4 * f;
    tids[tptr[c.x] + atomicSub(&cnt[c.x], 1) - 1] = t;
    tids[tptr[c.y] + atomicSub(&cnt[c.y], 1) - 1] = t + 1;
    tids[tptr[c.z] + atomicSub(&cnt[c.z], 1) - 1] = t + 2;
    tids[tptr[c.w] + atomicSub(&cnt[c.w], 1) - 1] = t + 3;
  } else {
    for (int32_t k = fptr[f]; k < fptr[f + 1]; ++k) {
      int32_t v = ids[k];
      tids[tptr[v] + atomicSub(&cnt[v], 1) - 1] = k;
    }
  }
}

// ---------------------------------------------------------------------------
// mapped SpGEMM, symbolic and numeric

__device__ void sort_row_in_place(int32_t* a, int n) {
  for (int i = 1; i < n; ++i) {
    int32_t x = a[i];
    int j = i;
    while (j > 0 && a[j - 1] > x) {
      a[j] = a[j - 1];
      --j;
    }
    a[j] = x;
  }
}

// The M^T rows of a block's R consecutive vertices (R = L.rows_per_block:
// 256, 128 or 64) are one contiguous segment [tptr[v0], tptr[v0 + R]) of tids
// (and of trow): when it fits the block's LDS tile (L.tile_entries entries,
// dynamic LDS; the plan sizes R and the tile from the row lengths) it is read
// in with all 256 lanes on consecutive entries, the rows are taken from LDS
// (one lane per row), and the sorted rows with their in-neighbours go back out
// the same way. Longer segments use global memory.
extern __shared__ __align__(16) unsigned char row_tile_lds[];

template <bool QUAD, int SCHEME, int CAP, bool STAGED>
__device__ __forceinline__ void symbolic_row(const LevelView& L, int32_t v, int32_t* hist, const int32_t* rows_in,
                                             int2* rows_out, int32_t s0);

template <bool QUAD, int SCHEME, int CAP>
__global__ void __launch_bounds__(kBlock) k_symbolic(LevelView L) {
  __shared__ int32_t h[kHistBins];
  const int T = L.tile_entries;  // even: tile_out starts 8-byte aligned
  int32_t* const tile_in = reinterpret_cast<int32_t*>(row_tile_lds);       // T corners
  int2* const tile_out = reinterpret_cast<int2*>(row_tile_lds + 4 * T);  // T (corner, in-neighbour)
  BlockHist bh{h};
  if (L.hist_rows) bh.init();
  const int R = L.rows_per_block;
  const int64_t v0 = block_index() * R;
  const int32_t v = (int32_t)(v0 + threadIdx.x);
  const bool valid = !(*L.flags & kFlagsInput);
  int32_t s0 = 0, len = 0;
  if (valid && v0 < L.nv) {
    s0 = L.tptr[v0];
    len = L.tptr[min(v0 + (int64_t)R, (int64_t)L.nv)] - s0;
  }
  const bool staged = len > 0 && len <= T;  // block-uniform
  if (staged)
    for (int i = threadIdx.x; i < len; i += blockDim.x) tile_in[i] = L.tids[s0 + i];
  __syncthreads();
  // (staged or not as a template argument: a pointer that may address LDS
  // or global memory makes every access through it a FLAT instruction)
  if ((int)threadIdx.x < R && v < L.nv && valid) {
    if (staged)
      symbolic_row<QUAD, SCHEME, CAP, true>(L, v, L.hist_rows ? h : nullptr, tile_in, L.trow ? tile_out : nullptr, s0);
    else
      symbolic_row<QUAD, SCHEME, CAP, false>(L, v, L.hist_rows ? h : nullptr, nullptr, nullptr, s0);
  }
  __syncthreads();
  if (staged && L.trow)
    for (int i = threadIdx.x; i < len; i += blockDim.x) L.trow[s0 + i] = tile_out[i];
  if (L.hist_rows) bh.flush(L.hist_rows);
}

// mapped-SpGEMM row sizing of vertex v (edges it owns), validation, the
// sorted row with in-neighbours; row length into the block histogram.
// STAGED: rows_in / rows_out are the block's LDS tiles from position s0
// (rows_out null: no sorted rows wanted); else both null
template <bool QUAD, int SCHEME, int CAP, bool STAGED>
__device__ __forceinline__ void symbolic_row(const LevelView& L, int32_t v, int32_t* hist, const int32_t* rows_in,
                                             int2* rows_out, int32_t s0) {
  int32_t s = L.tptr[v];
  int n = L.tptr[v + 1] - s;
  if (hist) hist_add_wave(hist, n < kHistBins - 1 ? n : kHistBins - 1);
  int32_t count = 0, bcount = 0;
  int bad = 0;
  // the sorted row goes to the block's LDS tile (or straight to trow)
  auto put = [&](int i, int2 x) __attribute__((always_inline)) {
    if (STAGED) {
      if (rows_out) rows_out[s - s0 + i] = x;
    } else if (L.trow) {
      L.trow[s + i] = x;
    }
  };
  if (n <= CAP) {
    RegRow<CAP, QUAD> r;
    if (STAGED)
      r.load(L, rows_in + (s - s0), n);
    else
      r.load(L, L.tids + s, n);
    row_symbolic<SCHEME>(r, v, count, bcount, bad);
#pragma unroll
    for (int i = 0; i < CAP; ++i)
      if (i < n) put(i, make_int2(r.T(i), r.U(i)));
    if (L.mpos)
#pragma unroll
      for (int i = 0; i < CAP; ++i)
        if (i < n) L.mpos[r.T(i)] = s + i;
  } else {
    sort_row_in_place(L.tids + s, n);
    MemRow<QUAD> r(L, s, n);
    row_symbolic<SCHEME>(r, v, count, bcount, bad);
    for (int i = 0; i < n; ++i) put(i, make_int2(r.T(i), r.U(i)));
    if (L.mpos)
      for (int i = 0; i < n; ++i) L.mpos[r.T(i)] = s + i;
  }
  L.ecnt[v] = count;
  (void)bcount;
  if (bad) {
    atomicOr(L.flags, bad);
    if (L.seg.mflags) atomicOr(L.seg.mflags + seg_of(L.seg.vbase, L.seg.n, v), bad);
  }
}

template <bool QUAD, int SCHEME, int CAP, bool STAGED>
__device__ __forceinline__ void number_row(const LevelView& L, int32_t v, const int2* rows, int32_t s0);
template <bool QUAD, bool FP, bool COUNT>
__device__ __forceinline__ void faces_item(const int32_t* __restrict__ ids, const int32_t* __restrict__ fptr,
                                           const float4* __restrict__ P, int32_t nf, const SegTable& seg,
                                           int32_t* __restrict__ cnt, const int* __restrict__ guard, int32_t f);

// First level: the edge numbering (blocks [0, nbn)) and the face points (the
// other blocks) are independent; one launch runs both.
template <bool QUAD, int SCHEME, int CAP>
__global__ void __launch_bounds__(kBlock) k_number_faces(LevelView L, int32_t nbn) {
  const int64_t b = block_index();
  if (b < nbn) {
    // the block's sorted rows through LDS, as in k_symbolic
    int2* const tile = reinterpret_cast<int2*>(row_tile_lds);  // L.tile_entries entries
    const int R = L.rows_per_block;
    const int64_t v0 = b * R;
    int32_t s0 = 0, len = 0;
    if (v0 < L.nv) {
      s0 = L.tptr[v0];
      len = L.tptr[min(v0 + (int64_t)R, (int64_t)L.nv)] - s0;
    }
    const bool staged = len > 0 && len <= L.tile_entries;  // block-uniform
    if (staged)
      for (int i = threadIdx.x; i < len; i += blockDim.x) tile[i] = L.trow[s0 + i];
    __syncthreads();
    if ((int)threadIdx.x < R) {
      if (staged)
        number_row<QUAD, SCHEME, CAP, true>(L, (int32_t)(v0 + threadIdx.x), tile, s0);
      else
        number_row<QUAD, SCHEME, CAP, false>(L, (int32_t)(v0 + threadIdx.x), nullptr, s0);
    }
  } else
    faces_item<QUAD, true, false>(L.ids, L.fptr, L.P, L.nf, L.seg, nullptr, nullptr,
                                  (int32_t)((b - nbn) * blockDim.x + threadIdx.x));
}

// STAGED: rows is the block's LDS copy of trow from position s0 (else null)
template <bool QUAD, int SCHEME, int CAP, bool STAGED>
__device__ __forceinline__ void number_row(const LevelView& L, int32_t v, const int2* rows, int32_t s0) {
  if (v >= L.nv) return;
  int32_t s = L.tptr[v];
  int n = L.tptr[v + 1] - s;
  if (n <= CAP) {
    RegRow<CAP, QUAD> r;
    // the symbolic pass's sorted rows
    if (STAGED)
      r.load_sorted(L, rows + (s - s0), n);
    else
      r.load_sorted(L, L.trow + s, n);
    row_number<SCHEME>(r, L, v);
  } else {
    // the symbolic pass sorted every row longer than its own capacity; the
    // in-memory path needs this one sorted whatever that capacity was (idempotent)
    sort_row_in_place(L.tids + s, n);
    MemRow<QUAD> r(L, s, n);
    row_number<SCHEME>(r, L, v);
  }
}

template <bool QUAD>
__global__ void __launch_bounds__(kBlock) k_spgemm_fde(LevelView L) {
  int32_t v = (int32_t)(block_index() * blockDim.x + threadIdx.x);
  if (v >= L.nv) return;
  int32_t s = L.tptr[v];
  int n = L.tptr[v + 1] - s;
  if (n <= kRowCap) {
    RegRow<kRowCap, QUAD> r;
    r.load(L, s, n);
    row_fde(r, L, v);
  } else {
    MemRow<QUAD> r(L, s, n);
    row_fde(r, L, v);
  }
}

__device__ __forceinline__ int32_t clampi(int32_t x, int32_t hi) { return min(max(x, 0), hi); }

// Streamed outputs of the final level are written once and not read back by
// this pipeline: non-temporal stores keep them from displacing the gathered
// positions in the caches.
typedef int v4i_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void store_nt(int32_t* p, int a, int b, int c, int d) {
  v4i_t x = {a, b, c, d};
  __builtin_nontemporal_store(x, reinterpret_cast<v4i_t*>(p));
}
__device__ __forceinline__ void store_nt(int32_t* p, int a) { __builtin_nontemporal_store(a, p); }
typedef float v4f_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void store_nt(float4* p, float4 q) {
  __builtin_nontemporal_store(__builtin_bit_cast(v4f_t, q), reinterpret_cast<v4f_t*>(p));
}

// DPP quad permutations (the four lanes of a quad hold the corners of one
// quad face): every lane of the quad reads lane K / lane i+1 / lane i-1.
template <int K>
__device__ __forceinline__ float4 quad_lane(float4 x) {
  constexpr int ctrl = K * 0x55;
  return make_float4(__int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x.x), ctrl, 0xF, 0xF, false)),
                     __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x.y), ctrl, 0xF, 0xF, false)),
                     __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x.z), ctrl, 0xF, 0xF, false)),
                     __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x.w), ctrl, 0xF, 0xF, false)));
}
__device__ __forceinline__ float4 quad_next(float4 x) {
  constexpr int ctrl = 0x39;  // [1, 2, 3, 0]
  return make_float4(__int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x.x), ctrl, 0xF, 0xF, false)),
                     __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x.y), ctrl, 0xF, 0xF, false)),
                     __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x.z), ctrl, 0xF, 0xF, false)),
                     __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x.w), ctrl, 0xF, 0xF, false)));
}
__device__ __forceinline__ float4 quad_prev(float4 x) {
  constexpr int ctrl = 0x93;  // [3, 0, 1, 2]
  return make_float4(__int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x.x), ctrl, 0xF, 0xF, false)),
                     __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x.y), ctrl, 0xF, 0xF, false)),
                     __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x.z), ctrl, 0xF, 0xF, false)),
                     __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x.w), ctrl, 0xF, 0xF, false)));
}
__device__ __forceinline__ int32_t quad_next_i(int32_t x) { return __builtin_amdgcn_mov_dpp(x, 0x39, 0xF, 0xF, false); }
__device__ __forceinline__ int32_t quad_prev_i(int32_t x) { return __builtin_amdgcn_mov_dpp(x, 0x93, 0xF, 0xF, false); }

// entry k of the next level's M^T rows: (corner, in-neighbour), or the corner
// alone when those rows are read by the last level's kernel only
template <bool QUAD>
__device__ __forceinline__ void put_next_row(const RefineView& R, int64_t k, int2 x) {
  if (QUAD && R.next_rows_x)  // (only a quad level writes such rows)
    reinterpret_cast<int32_t*>(R.ntrow)[k] = x.x;
  else
    R.ntrow[k] = x;
}

// Per-segment constants of a refine thread, all read from the segment table
// up front: a table read issued after a store would wait for that store.
struct SegK {
  int32_t vb, fb, cb, snv, snf, snnz, emax, off;
  int64_t eb;
  float4* dst_verts;
  int32_t* dst_ids;
  int32_t* dst_ptr;
  int32_t* dst_vals;
  int32_t ncb, nvb, neb;  // next level's union bases (ncb < 0: the mesh finishes here)
  bool next, dst_final;
};

__device__ __forceinline__ SegK seg_consts(const SegTable& seg, int32_t f) {
  SegK K;
  const SegDesc* __restrict__ S = seg.d + seg_of_wave(seg.fbase, seg.n, f);
  K.vb = (int32_t)S->vb;
  K.fb = (int32_t)S->fb;
  K.cb = (int32_t)S->cb;
  K.snv = S->nv;
  K.snf = S->nf;
  K.snnz = (int32_t)S->nnz;
  K.emax = S->ne > 0 ? (int32_t)S->ne - 1 : 0;
  K.off = (int32_t)S->id_offset;
  K.eb = S->eb;
  K.dst_verts = to_global(S->dst_verts);
  K.dst_ids = to_global(S->dst_ids);
  K.dst_ptr = to_global(S->dst_ptr);
  K.dst_vals = to_global(S->dst_vals);
  K.ncb = (int32_t)S->next_cb;
  K.nvb = (int32_t)S->next_vb;
  K.neb = (int32_t)S->next_eb;
  K.next = S->next_cb >= 0;
  K.dst_final = S->dst_final != 0;
  return K;
}

// Everything one corner's stores need, loaded beforehand.
struct CornerIn {
  int32_t t, f, tn, tp;   // corner, face, successor / predecessor corner
  int32_t v, w, u;        // its vertex, out- and in-neighbour
  int32_t e1, e0, s1, s0; // out / in edge ids and twin corners
  int32_t g1, g0, n1, ns0;  // twin faces, successors of the twins
  int32_t bl1, bl0;       // boundary edges before e1 / e0
  int32_t mp_t, mp_s0;    // M^T positions of t and s0 (SLAK, next level)
  bool own;               // this half-edge writes edge e1
  float4 plo, phi, fa, fz;  // edge point inputs (own only)
};

// First round of a corner's loads: ids and (edge, twin) pairs already in
// hand; derives twins' faces and decides ownership.
template <bool QUAD>
__device__ __forceinline__ void corner_topology(const RefineView& R, const SegK& K, CornerIn& c, int2 c1, int2 c0) {
  c.e1 = clampi(c1.x, K.emax);
  c.e0 = clampi(c0.x, K.emax);
  c.s1 = c1.y;
  c.s0 = c0.y;
  c.g1 = c.f;
  c.g0 = c.f;
  c.n1 = 0;
  c.ns0 = 0;
  if (c.s1 >= 0) {
    int32_t x;
    corner_nbrs<QUAD>(R.fptr, R.cface, c.s1, c.g1, c.n1, x);
  }
  if (c.s0 >= 0) {
    int32_t x;
    corner_nbrs<QUAD>(R.fptr, R.cface, c.s0, c.g0, c.ns0, x);
  }
  c.own = c.s1 < 0 || c.t < c.s1;
}

// Second round: boundary ranks and SLAK positions.
template <int SCHEME>
__device__ __forceinline__ void corner_ranks(const RefineView& R, const SegK& K, CornerIn& c) {
  c.bl1 = 0;
  c.bl0 = 0;
  if (R.blk) {
    const int32_t bb = R.blk[K.eb];
    c.bl1 = R.blk[K.eb + c.e1] - bb;
    c.bl0 = R.blk[K.eb + c.e0] - bb;
  }
  c.mp_t = 0;
  c.mp_s0 = 0;
  if (SCHEME == kSLAK && K.next) {
    c.mp_t = R.mpos[c.t];
    if (c.s0 >= 0) c.mp_s0 = R.mpos[c.s0];
  }
}

// Every store of one corner t: refined quad, LAK form, edge point (owner)
// and, for a level that feeds another, the refined mesh's structure.
template <bool QUAD, int SCHEME>
__device__ __forceinline__ void corner_store(const RefineView& R, const SegK& K, const CornerIn& c) {
  const int32_t lt = c.t - K.cb, lf = c.f - K.fb;
  const int32_t epb = K.snv + K.snf + K.off, fpid = K.snv + lf + K.off;
  const bool final_level = K.dst_final;  // the caller's output: not read again here
  if (final_level) {
    store_nt(K.dst_ids + 4 * lt, c.v - K.vb + K.off, epb + c.e1, fpid, epb + c.e0);
    if (K.dst_ptr) {
      // LAK form of refined face lt (createQuadColPtr / createQuadVals,
      // LAKCatmullClarkKernels.cuh:396-431): ptr = 4 lt, vals 1..4
      store_nt(K.dst_ptr + lt, 4 * lt);
      if (lt == K.snnz - 1) store_nt(K.dst_ptr + K.snnz, 4 * K.snnz);
      if (K.dst_vals) store_nt(K.dst_vals + 4 * (int64_t)lt, 1, 2, 3, 4);
    }
  } else if (K.dst_ids) {  // null: the next level is computed without them (k_final)
    reinterpret_cast<int4*>(K.dst_ids)[lt] = make_int4(c.v - K.vb + K.off, epb + c.e1, fpid, epb + c.e0);
  }
  if (c.own) {
    float4 q;
    if (c.s1 >= 0)
      q = f4mul(f4add(f4add(f4add(c.plo, c.phi), c.fa), c.fz), 0.25f);
    else
      q = f4mul(f4add(c.plo, c.phi), 0.5f);
    float4* const dv = K.dst_verts + (K.snv + K.snf + c.e1);
    if (final_level) store_nt(dv, q);
    else *dv = q;
  }
  if (!K.next) return;

  // ---- the refined mesh's structure
  const int32_t cb = K.cb, snnz = K.snnz, snv = K.snv, snf = K.snf;
  const int32_t ncb = K.ncb, nvb = K.nvb;
  const int32_t c4 = ncb + 4 * lt;
  const int32_t tw0 = c.s1 >= 0 ? ncb + 4 * (c.n1 - cb) + 3 : -1;
  const int32_t tw1 = ncb + 4 * (c.tn - cb) + 2;
  const int32_t tw2 = ncb + 4 * (c.tp - cb) + 1;
  const int32_t tw3 = c.s0 >= 0 ? ncb + 4 * (c.s0 - cb) : -1;
  // edge point rows (the clamp only matters for inconsistent edge ids: it
  // keeps a row inside the mesh's refined M^T)
  const int32_t row1 = min(ncb + 2 * snnz + 4 * c.e1 - 2 * c.bl1, ncb + 4 * snnz - (c.s1 < 0 ? 2 : 4));
  const int32_t row0 = min(ncb + 2 * snnz + 4 * c.e0 - 2 * c.bl0, ncb + 4 * snnz - (c.s0 < 0 ? 2 : 4));
  int32_t ex, ey, ez, ew;
  if (SCHEME == kLAK) {
    const int32_t base1 = 4 * c.e1 - c.bl1, base0 = 4 * c.e0 - c.bl0;
    ex = base1 + (c.v > c.w ? 1 : 0);
    ey = base1 + 2 + ((c.s1 >= 0 && c.f > c.g1) ? 1 : 0);
    ez = base0 + 2 + ((c.s0 >= 0 && c.f > c.g0) ? 1 : 0);
    ew = base0 + (c.v > c.u ? 1 : 0);
  } else {
    ex = c.mp_t - cb;
    ey = snnz + (c.tn - cb);
    ez = snnz + lt;
    ew = c.s0 < 0 ? 2 * snnz + c.bl0 : c.mp_s0 - cb;
    // M^T positions of the refined corners
    const int32_t a1 = c4 + 1, a3 = c4 + 3;
    int32_t r1 = (ncb + 4 * (c.tn - cb) + 3 < a1) ? 1 : 0;
    if (c.s1 >= 0) r1 += (ncb + 4 * (c.s1 - cb) + 1 < a1 ? 1 : 0) + (ncb + 4 * (c.n1 - cb) + 3 < a1 ? 1 : 0);
    int32_t r3 = (ncb + 4 * (c.tp - cb) + 1 < a3) ? 1 : 0;
    if (c.s0 >= 0) r3 += (ncb + 4 * (c.s0 - cb) + 1 < a3 ? 1 : 0) + (ncb + 4 * (c.ns0 - cb) + 3 < a3 ? 1 : 0);
    reinterpret_cast<int4*>(R.nmpos)[c4 >> 2] =
        make_int4(ncb + (c.mp_t - cb), row1 + r1, ncb + snnz + lt, row0 + r3);
  }
  int4* __restrict__ o = reinterpret_cast<int4*>(R.nci + c4);
  o[0] = make_int4(ex, tw0, ey, tw1);
  o[1] = make_int4(ez, tw2, ew, tw3);
  // face point row: corner 2 of each refined quad of f, in corner order;
  // in-neighbour ep(e1)
  const int32_t nep = nvb + snv + snf;
  put_next_row<QUAD>(R, ncb + snnz + lt, make_int2(c4 + 2, nep + c.e1));
  bool first;
  if (QUAD) first = (c.t & 3) == 0;
  else first = c.t == R.fptr[c.f];
  if (first) R.ntptr[nvb + snv + lf] = ncb + snnz + lt;
  int32_t* __restrict__ nbnd = R.nbnd;
  const int32_t neb = K.neb;
  if (SCHEME == kSLAK && nbnd) nbnd[neb + snnz + lt] = 0;  // face point edges are interior
  if (c.own) {
    R.ntptr[nep + c.e1] = row1;
    int2 a0 = make_int2(c4 + 1, nvb + (c.v - K.vb));
    int2 a1 = make_int2(ncb + 4 * (c.tn - cb) + 3, nvb + snv + lf);
    if (a1.x < a0.x) {
      int2 x = a0;
      a0 = a1;
      a1 = x;
    }
    auto q = [&](int k, int2 x) __attribute__((always_inline)) { put_next_row<QUAD>(R, row1 + k, x); };
    if (c.s1 < 0) {
      q(0, a0);
      q(1, a1);
    } else {
      int2 b0 = make_int2(ncb + 4 * (c.s1 - cb) + 1, nvb + (c.w - K.vb));
      int2 b1 = make_int2(ncb + 4 * (c.n1 - cb) + 3, nvb + snv + (c.g1 - K.fb));
      if (b1.x < b0.x) {
        int2 x = b0;
        b0 = b1;
        b1 = x;
      }
      // merge two sorted pairs
      const bool alo = a0.x < b0.x, ahi = a1.x > b1.x;
      const int2 m0 = alo ? a0 : b0, m3 = ahi ? a1 : b1;
      const int2 l = alo ? b0 : a0, h = ahi ? b1 : a1;
      const bool lh = l.x < h.x;
      q(0, m0);
      q(1, lh ? l : h);
      q(2, lh ? h : l);
      q(3, m3);
    }
    if (nbnd) {
      // the refined mesh has snnz + 2 ne edges
      const int32_t nlast = neb + snnz + 2 * (K.emax + 1) - 1;
      if (SCHEME == kLAK) {
        const int32_t base1 = min(neb + 4 * c.e1 - c.bl1, nlast - (c.s1 < 0 ? 2 : 3));
        const int32_t b = c.s1 < 0 ? 1 : 0;
        nbnd[base1] = b;
        nbnd[base1 + 1] = b;
        nbnd[base1 + 2] = 0;
        if (c.s1 >= 0) nbnd[base1 + 3] = 0;
      } else if (c.s1 < 0) {
        nbnd[min(neb + 2 * snnz + c.bl1, nlast)] = 1;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// refine: for every corner t of face f, v -> w the corner's out-half-edge
// (edge e1, twin s1), u -> v the in-half-edge (edge e0, twin s0).
//   refined quad t = (v, ep(e1), fp(f), ep(e0))   in the mesh's local numbering
//   edge point e1  (written by the half-edge min(t, s1), or t on the boundary)
//                  = (P_lo + P_hi + fp(face of lo->hi) + fp(face of hi->lo)) / 4,
//                    (P_lo + P_hi) / 2 on the boundary
// and for a level that feeds another, in closed form:
//   twins          4t -> 4 next(s1) + 3, 4t+1 -> 4 next(t) + 2,
//                  4t+2 -> 4 prev(t) + 1, 4t+3 -> 4 s0
//   LAK edge ids   edge point e owns its edges to the two end points, then to
//                  the face points, ranked by the smaller id: base(e) = 4e - bl(e)
//   SLAK edge ids  vertices own one edge per corner, numbered by the corner's
//                  M^T storage position; face points one per corner
//                  (nnz + corner); boundary edge points one (2 nnz + bl(e))
//   M^T rows       face point f: its corners 4t+2; edge point e: 4t+1,
//                  4 next(t) + 3 of both half-edges, sorted (row 2 nnz + 4e - 2 bl(e))
//   M^T positions  (SLAK) 4t keeps t's, 4t+2 -> nnz + t, 4t+1 / 4t+3 -> rank
//                  in their edge point's row
// bl(e) = boundary edges with a smaller id (0 on closed levels). One thread
// per corner, so that each store instruction of a wave covers contiguous
// memory; every load is issued before the first store.

template <bool QUAD, int SCHEME>
__device__ __forceinline__ void refine_item(const RefineView& R, int64_t t64) {
  if (t64 >= R.nnz) return;
  CornerIn c;
  c.t = (int32_t)t64;
  corner_nbrs<QUAD>(R.fptr, R.cface, c.t, c.f, c.tn, c.tp);
  const SegK K = seg_consts(R.seg, c.f);
  if (R.ntptr && t64 == R.nnz - 1) R.ntptr[R.next_nv] = R.next_nnz;  // end of the next level's M^T
  c.v = R.ids[c.t];
  const int2 c1 = R.ci[c.t];
  if (QUAD) {
    // the face's other corners are the quad's other lanes (nnz = 4 nf, and
    // blocks start on a face): w, u and the in-half-edge by DPP, not loads
    c.w = quad_next_i(c.v);
    c.u = quad_prev_i(c.v);
    corner_topology<QUAD>(R, K, c, c1, make_int2(quad_prev_i(c1.x), quad_prev_i(c1.y)));
  } else {
    c.w = R.ids[c.tn];
    c.u = R.ids[c.tp];
    corner_topology<QUAD>(R, K, c, c1, R.ci[c.tp]);
  }
  const bool vlo = c.v < c.w;
  const float4* __restrict__ fp = K.dst_verts + K.snv;  // face points, by mesh-local face
  if (c.own) {
    const float4 Pv = R.P[c.v], Pw = R.P[c.w];
    c.plo = vlo ? Pv : Pw;
    c.phi = vlo ? Pw : Pv;
    if (c.s1 >= 0) {
      c.fa = fp[(vlo ? c.f : c.g1) - K.fb];
      c.fz = fp[(vlo ? c.g1 : c.f) - K.fb];
    }
  }
  corner_ranks<SCHEME>(R, K, c);
  corner_store<QUAD, SCHEME>(R, K, c);
}

// vertex: one thread per vertex v with M^T row (t_i, u_i), i < n, sorted.
//   interior  q = ((sum_i fp(face t_i) + sum_i P[u_i]) / n + P[v] (n - 2)) / n
//   boundary  q = 3/4 P[v] + 1/8 (neighbours across boundary edges)
// and for a level that feeds another: the vertex's refined M^T row (corners
// 4 t_i, in-neighbours ep(e0(t_i))) and, for SLAK, the boundary flags of the
// edges it owns. The first kVertexRegs row entries and everything they gather
// are loaded before any store; longer rows loop.
constexpr int kVertexRegs = 4;

template <bool QUAD, int SCHEME, bool NEXT, bool BND>
__device__ __forceinline__ void vertex_item(const RefineView& R, int32_t v) {
  if (v >= R.nv) return;
  // the segment's constants, all read before the first store (a table read
  // issued after a store would wait for that store)
  const SegDesc* __restrict__ S = R.seg.d + seg_of_wave(R.seg.vbase, R.seg.n, v);
  const int32_t vb = (int32_t)S->vb, fb = (int32_t)S->fb, cb = (int32_t)S->cb;
  const int32_t snv = S->nv, snf = S->nf;
  const int32_t semax = S->ne > 0 ? (int32_t)S->ne - 1 : 0;
  float4* __restrict__ dst = to_global(S->dst_verts);
  const bool dst_final = S->dst_final != 0;
  const int32_t ncb = NEXT ? (int32_t)S->next_cb : -1, nvb = NEXT ? (int32_t)S->next_vb : 0;
  const int32_t neb = NEXT ? (int32_t)S->next_eb : 0;
  const float4* __restrict__ fpts = dst + snv;  // face points, by mesh-local face
  const float4* __restrict__ P = R.P;
  const int2* __restrict__ ci = R.ci;
  // NEXT: some mesh of the level continues (R.ntptr set); BND: the level has a
  // boundary (R.blk set)
  const bool next = NEXT && ncb >= 0;
  constexpr bool bnd_level = BND;
  constexpr bool need_prev = NEXT || BND;  // (edge, twin) of the in-half-edge
  constexpr bool need_own = BND;           // (edge, twin) of the out-half-edge
  const int32_t p0 = R.tptr[v];
  const int n = R.tptr[v + 1] - p0;
  const int2* __restrict__ row = R.trow + p0;
  const float4 Pv = P[v];
  auto face = [&](int32_t t) __attribute__((always_inline)) { return (QUAD ? (t >> 2) : R.cface[t]) - fb; };
  auto prevc = [&](int32_t t) __attribute__((always_inline)) {
    int32_t f, tn, tp;
    corner_nbrs<QUAD>(R.fptr, R.cface, t, f, tn, tp);
    return tp;
  };
  // the in-half-edges give the refined row's in-neighbours (and the boundary
  // rule): not needed when the next level's rows are corners alone
  const bool load_cp = BND || (NEXT && !(QUAD && R.next_rows_x));
  int2 e[kVertexRegs], cp[kVertexRegs], co[kVertexRegs];
  float4 a[kVertexRegs], b[kVertexRegs];
#pragma unroll
  for (int i = 0; i < kVertexRegs; ++i)
    if (i < n) e[i] = row[i];
#pragma unroll
  for (int i = 0; i < kVertexRegs; ++i)
    if (i < n) {
      a[i] = fpts[face(e[i].x)];
      b[i] = P[e[i].y];
      if (need_prev && load_cp) cp[i] = ci[prevc(e[i].x)];
      if (need_own) co[i] = ci[e[i].x];
    }
  bool boundary = false;
  if (bnd_level) {
#pragma unroll
    for (int i = 0; i < kVertexRegs; ++i)
      if (i < n) boundary |= co[i].y < 0 || cp[i].y < 0;
    for (int i = kVertexRegs; i < n; ++i) boundary |= ci[row[i].x].y < 0 || ci[prevc(row[i].x)].y < 0;
  }
  float4 q;
  if (!boundary) {
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int i = 0; i < kVertexRegs; ++i)
      if (i < n) s = f4add(s, a[i]);
    for (int i = kVertexRegs; i < n; ++i) s = f4add(s, fpts[face(row[i].x)]);
#pragma unroll
    for (int i = 0; i < kVertexRegs; ++i)
      if (i < n) s = f4add(s, b[i]);
    for (int i = kVertexRegs; i < n; ++i) s = f4add(s, P[row[i].y]);
    const float nf = (float)n;
    q = f4div(f4add(f4div(s, nf), f4mul(Pv, nf - 2.0f)), nf);
  } else {
    q = f4mul(Pv, 0.75f);
    for (int i = 0; i < n; ++i) {
      const int32_t t = row[i].x;
      if (ci[t].y < 0) {
        int32_t f, tn, tp;
        corner_nbrs<QUAD>(R.fptr, R.cface, t, f, tn, tp);
        q = f4add(q, f4mul(P[R.ids[tn]], 0.125f));
      }
    }
    for (int i = 0; i < n; ++i)
      if (ci[prevc(row[i].x)].y < 0) q = f4add(q, f4mul(P[row[i].y], 0.125f));
  }
  if (!next && dst_final) {
    store_nt(dst + (v - vb), q);
    return;
  }
  dst[v - vb] = q;
  if (!next) return;

  const int32_t nep = nvb + snv + snf;
  const int32_t emax = semax;
  const int32_t pos = p0 - cb;  // mesh-local M^T storage position of the row
  R.ntptr[nvb + (v - vb)] = ncb + pos;
  const int64_t nrow = (int64_t)ncb + pos;  // first entry of the refined row
  int32_t* __restrict__ nbnd = (SCHEME == kSLAK) ? R.nbnd : nullptr;
#pragma unroll
  for (int i = 0; i < kVertexRegs; ++i)
    if (i < n) {
      put_next_row<QUAD>(R, nrow + i, make_int2(ncb + 4 * (e[i].x - cb), load_cp ? nep + clampi(cp[i].x, emax) : 0));
      if (nbnd) nbnd[neb + pos + i] = co[i].y < 0 ? 1 : 0;
    }
  for (int i = kVertexRegs; i < n; ++i) {
    const int32_t t = row[i].x;
    put_next_row<QUAD>(R, nrow + i, make_int2(ncb + 4 * (t - cb), load_cp ? nep + clampi(ci[prevc(t)].x, emax) : 0));
    if (nbnd) nbnd[neb + pos + i] = ci[t].y < 0 ? 1 : 0;
  }
}

// One launch per level: the refine items (one per corner) and the vertex
// items (one per vertex) are independent once the face points exist, and they
// write disjoint outputs. Their blocks are interleaved in proportion over the
// grid, so write-bound refine waves and gather-bound vertex waves share the
// machine. Block b is a vertex block when floor((b+1) nbv / T) > floor(b nbv / T)
// (T = nbr + nbv): vertex block floor(b nbv / T), else refine block
// b - floor(b nbv / T); a bijection onto both block ranges.
template <bool QUAD, int SCHEME, bool NEXT, bool BND>
__global__ void __launch_bounds__(kBlock) k_level(RefineView R, int32_t nbr, int32_t nbv) {
  const int64_t b = NEXT ? block_index<kRemapLevel>() : block_index<kRemapPositions>();
  const int64_t T = (int64_t)nbr + nbv;
  const int64_t before = b * nbv / T, after = (b + 1) * nbv / T;
  if (after > before) {
    // vertex block `before`: its class run and its block within the run
    int64_t u = before, rest = nbv;
    int s = 0;
    for (; s + 1 < R.nvseg; ++s) {
      const int64_t nb = R.vnb[s];
      const int64_t lo = u * nb / rest, hi = (u + 1) * nb / rest;
      if (hi > lo) {
        u = lo;
        break;
      }
      u -= lo;
      rest -= nb;
    }
    const int64_t v = (int64_t)R.vseg[s] + u * blockDim.x + threadIdx.x;
    vertex_item<QUAD, SCHEME, NEXT, BND>(R, v < R.vseg[s + 1] ? (int32_t)v : R.nv);
  } else {
    refine_item<QUAD, SCHEME>(R, (b - before) * blockDim.x + threadIdx.x);
  }
}

// ---------------------------------------------------------------------------
// A level in closed form from the quad mesh M two levels before it (one mesh,
// or several through a device table and work units): the output level (IDS:
// quads, LAK form, positions) or the level before it (positions only). M1
// (the refinement of M, whose positions P1 are in hand; its ids and
// structure are never written) has corners 4t + k for the corners t of M:
//   quad t of M1   (v, ep(e1), fp(F), ep(e0))  in M1 vertex ids
//   M1 edge ids    corner 4t + k leaves along E1[k] = (ex, ey, ez, ew) and
//                  enters along E0[k] = (ew, ex, ey, ez), the ids the refine
//                  items give M1 (LAK: 4e - bl(e) + rank, SLAK: M^T positions)
//   output quad    of M1 corner 4t + k: (vertex, ep'(E1[k]), fp'(t), ep'(E0[k]))
// One lane per corner t of M computes the face point of M1 face t from its
// four corners (the fourth comes from the previous lane) and its four output
// quads (64 contiguous bytes, written out by the block through LDS). The
// owner of edge e1 of M (t < twin, or no twin) writes the edge points of the
// three or four M1 edges around ep(e1) (both halves of e1 and the edges to
// fp(F) and fp(G); for SLAK each lane writes its own edge to fp(F) instead),
// contiguous for LAK, and the vertex point of ep(e1), from the same nine
// positions. Lane 0 of each quad writes the vertex
// point of fp(F) from its four lanes. Vertex items remain only for the
// vertices of M. Every sum adds the same terms in the same order as the
// refine and vertex items of M1 would: the results are the same bits.

__device__ __forceinline__ float4 f4sel(bool c, float4 a, float4 b) { return c ? a : b; }
// face point of a quad from its corners' positions in corner order (the sum
// and division k_faces uses)
__device__ __forceinline__ float4 face_point(float4 a, float4 b, float4 c, float4 d) {
  return f4div(f4add(f4add(f4add(a, b), c), d), 4.0f);
}

constexpr int kStageStride = 5;  // int4 per lane in the staging area (4 used; 5 spreads the banks)

struct Quads4 {
  int4 a, b, c, d;
};

// Corner t (mesh-local) of M in mesh S: returns its 16 output ids (IDS),
// which the block stages in LDS.
template <int SCHEME, bool BND, bool IDS, bool EARLY>
__device__ __forceinline__ Quads4 final_corner(const FinalView& V, const FinalSeg& S, int32_t t) {
  // every lane of a quad runs (nnz = 4 nf, meshes and blocks hold whole quads)
  const int64_t cb = S.cb;
  const int32_t vb = (int32_t)S.vb;
  const int32_t F = t >> 2, q0 = t & ~3;
  const int32_t tn = q0 | ((t + 1) & 3);
  const int32_t* __restrict__ ids = V.ids + cb;
  const int2* __restrict__ ci = V.ci + cb;
  // the quad's four lanes load their own corner once; w (next corner's
  // vertex), u and the in-half-edge's (edge, twin) come from the
  // neighbouring lanes
  const int32_t v = ids[t] - vb;
  const int32_t w = quad_next_i(v), u = quad_prev_i(v);
  const int2 c1 = ci[t];
  const int2 c0 = make_int2(quad_prev_i(c1.x), quad_prev_i(c1.y));
  const int32_t emax = S.ne > 0 ? S.ne - 1 : 0;
  const int32_t e1 = clampi(c1.x, emax), e0 = clampi(c0.x, emax);
  // twins are union corners
  const int32_t s1 = c1.y >= 0 ? (int32_t)(c1.y - cb) : -1, s0 = c0.y >= 0 ? (int32_t)(c0.y - cb) : -1;
  const int32_t g1 = s1 >= 0 ? (s1 >> 2) : F, g0 = s0 >= 0 ? (s0 >> 2) : F;
  const int32_t n1 = s1 >= 0 ? ((s1 & ~3) | ((s1 + 1) & 3)) : 0;  // next(s1)
  const bool own = s1 < 0 || t < s1;
  const int32_t nv = S.nv, nnz = S.nnz;
  const int32_t ep0 = nv + S.nf;     // M1 vertex of the edge point of edge 0 of M
  const int32_t nv1 = ep0 + S.ne;    // M1 vertices
  const int32_t epb2 = nv1 + nnz;    // output vertex of the edge point of M1 edge 0
  const float4* __restrict__ P1 = S.P1;
  float4* __restrict__ out = S.out_verts;
  int32_t bl1 = 0, bl0 = 0;
  if (BND) {
    const int32_t* __restrict__ blk = V.blk + S.eb;
    const int32_t bb = blk[0];
    bl1 = blk[e1] - bb;
    bl0 = blk[e0] - bb;
  }
  int32_t mp_t = 0, mp_s0 = 0, mp_s1 = 0;  // mesh-local M^T positions
  if (SCHEME == kSLAK) {
    const int32_t* __restrict__ mpos = V.mpos + cb;
    mp_t = mpos[t] - (int32_t)cb;
    if (s0 >= 0) mp_s0 = mpos[s0] - (int32_t)cb;
    if (own && s1 >= 0) mp_s1 = mpos[s1] - (int32_t)cb;
  }
  int2 cs0 = make_int2(0, 0), cn1 = make_int2(0, 0);  // (edge, twin) of prev(s1) and next(s1)
  if (own && s1 >= 0) {
    cs0 = ci[(s1 & ~3) | ((s1 + 3) & 3)];
    cn1 = ci[n1];
  }
  const float4 Pe = P1[ep0 + e1];   // ep(e1)
  const float4 Pv = P1[v];
  const float4 PF = P1[nv + F];     // fp(F)
  // across e1: fp(G) and the edge points of prev(s1) and next(s1). EARLY:
  // loaded before the first store, since a wait for a load also waits for
  // the stores issued before it (more registers: one wave fewer per SIMD)
  float4 PG = make_float4(0.f, 0.f, 0.f, 0.f), Pa = PG, Pb = PG;
  if (EARLY && own && s1 >= 0) {
    PG = P1[nv + g1];
    Pa = P1[ep0 + clampi(cs0.x, emax)];
    Pb = P1[ep0 + clampi(cn1.x, emax)];
  }
  // from the other lanes of the quad (all four run this): w is the vertex of
  // the next corner, ep(e0) the edge point of the previous one
  const float4 Pw = quad_next(Pv), Pu = quad_prev(Pe);
  // face point of M1 face t = (v, ep(e1), fp(F), ep(e0)), as k_faces forms it;
  // it is an output position too
  const float4 fpt = face_point(Pv, Pe, PF, Pu);
  out[nv1 + t] = fpt;
  const float4 fptn = quad_next(fpt);
  // vertex point of fp(F): row 4t'+2 for the corners t' of F in order,
  // in-neighbours ep(e1(t')); never on the boundary. Every lane of the quad
  // forms the sum from the quad's lanes in corner order (each term consumed
  // as it arrives), lane 0 stores it.
  {
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    s = f4add(s, quad_lane<0>(fpt));
    s = f4add(s, quad_lane<1>(fpt));
    s = f4add(s, quad_lane<2>(fpt));
    s = f4add(s, quad_lane<3>(fpt));
    s = f4add(s, quad_lane<0>(Pe));
    s = f4add(s, quad_lane<1>(Pe));
    s = f4add(s, quad_lane<2>(Pe));
    s = f4add(s, quad_lane<3>(Pe));
    if ((t & 3) == 0) out[nv + F] = f4div(f4add(f4div(s, 4.0f), f4mul(PF, 2.0f)), 4.0f);
  }
  // across e1: M1 faces s1 = (w, ep(e1), fp(G), ep(e0(s1))) and
  // next(s1) = (v, ep(e1(next(s1))), fp(G), ep(e1))
  float4 fps = make_float4(0.f, 0.f, 0.f, 0.f), fpsn = fps;
  if (own && s1 >= 0) {
    if (!EARLY) {
      PG = P1[nv + g1];
      Pa = P1[ep0 + clampi(cs0.x, emax)];
      Pb = P1[ep0 + clampi(cn1.x, emax)];
    }
    fps = face_point(Pw, Pe, PG, Pa);
    fpsn = face_point(Pv, Pb, PG, Pe);
  }

  // M1 edge ids of the corners 4t .. 4t+3 (as the refine items number M1)
  int32_t ex, ey, ez, ew;
  if (!IDS) {
    ex = ey = ez = ew = 0;
  } else if (SCHEME == kLAK) {
    const int32_t base1 = 4 * e1 - bl1, base0 = 4 * e0 - bl0;
    ex = base1 + (v > w ? 1 : 0);
    ey = base1 + 2 + ((s1 >= 0 && F > g1) ? 1 : 0);
    ez = base0 + 2 + ((s0 >= 0 && F > g0) ? 1 : 0);
    ew = base0 + (v > u ? 1 : 0);
  } else {
    ex = mp_t;
    ey = nnz + tn;
    ez = nnz + t;
    ew = s0 < 0 ? 2 * nnz + bl0 : mp_s0;
  }
  // the four output quads of M1 face t
  Quads4 q4{};
  if (IDS) {
    const int32_t off = S.id_off, eb2 = epb2 + off;
    const int32_t x0 = v + off, x1 = ep0 + e1 + off, x2 = nv + F + off, x3 = ep0 + e0 + off, fid = nv1 + t + off;
    // the block stages them in LDS and writes them out as 1 KB runs
    // (64-byte pieces per lane, stored directly, cost several times more)
    q4.a = make_int4(x0, eb2 + ex, fid, eb2 + ew);
    q4.b = make_int4(x1, eb2 + ey, fid, eb2 + ex);
    q4.c = make_int4(x2, eb2 + ez, fid, eb2 + ey);
    q4.d = make_int4(x3, eb2 + ew, fid, eb2 + ez);
  }
  if (SCHEME == kSLAK) {
    // the M1 edge from ep(e1) to fp(F), SLAK slot nnz + next(t): every lane
    // writes its own, so a wave stores 1 KB of consecutive slots (the owner
    // of e1 would form the same sum, in the same order, for both faces)
    const float4 q = f4mul(f4add(f4add(f4add(PF, Pe), fptn), fpt), 0.25f);
    float4* const d = out + epb2 + nnz + tn;
    if (IDS)
      store_nt(d, q);
    else
      *d = q;
  }
  if (!own) return q4;
  // edge points of the M1 edges around ep(e1), by lower end point then the
  // faces of the half-edges lower -> upper and upper -> lower
  const bool inner = !BND || s1 >= 0;
  const float4 q_vh = inner ? f4mul(f4add(f4add(f4add(Pv, Pe), fpt), fpsn), 0.25f) : f4mul(f4add(Pv, Pe), 0.5f);
  const float4 q_wh = inner ? f4mul(f4add(f4add(f4add(Pw, Pe), fps), fptn), 0.25f) : f4mul(f4add(Pw, Pe), 0.5f);
  const float4 q_fF = f4mul(f4add(f4add(f4add(PF, Pe), fptn), fpt), 0.25f);
  int32_t i_vh, i_wh, i_fF, i_fG;
  if (SCHEME == kLAK) {
    const int32_t base1 = 4 * e1 - bl1;
    i_vh = base1 + (v > w ? 1 : 0);
    i_wh = base1 + (w > v ? 1 : 0);
    i_fF = base1 + 2 + ((s1 >= 0 && F > g1) ? 1 : 0);
    i_fG = base1 + 2 + (g1 > F ? 1 : 0);
  } else {
    i_vh = mp_t;
    i_wh = s1 >= 0 ? mp_s1 : 2 * nnz + bl1;
    i_fF = nnz + tn;
    i_fG = nnz + n1;
  }
  // LAK: one 64-byte group, cached; SLAK: the two halves of e1 (the edges
  // to the face points are written above), non-temporal at the output level
  // (cached when the next kernel reads them)
  float4* __restrict__ eo = out + epb2;
  const float4 q_fG = f4mul(f4add(f4add(f4add(PG, Pe), fpsn), fps), 0.25f);
  if (SCHEME == kLAK) {
    eo[i_vh] = q_vh;
    eo[i_wh] = q_wh;
    eo[i_fF] = q_fF;
    if (inner) eo[i_fG] = q_fG;
  } else if (!IDS) {
    eo[i_vh] = q_vh;
    eo[i_wh] = q_wh;
  } else {
    store_nt(eo + i_vh, q_vh);
    store_nt(eo + i_wh, q_wh);
  }
  // vertex point of ep(e1)
  float4 q;
  if (inner) {
    // row: corners 4t+1 (face t, in-neighbour v), 4tn+3 (tn, fp(F)),
    // 4s1+1 (s1, w), 4n1+3 (n1, fp(G)), sorted
    const int32_t ka = 4 * t + 1, kb = 4 * tn + 3, kc = 4 * s1 + 1, kd = 4 * n1 + 3;
    const int ra = (kb < ka) + (kc < ka) + (kd < ka);
    const int rb = (ka < kb) + (kc < kb) + (kd < kb);
    const int rc = (ka < kc) + (kb < kc) + (kd < kc);
    // the face points, then the in-neighbours, each in row order
    float4 sm = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int i = 0; i < 4; ++i) sm = f4add(sm, f4sel(ra == i, fpt, f4sel(rb == i, fptn, f4sel(rc == i, fps, fpsn))));
#pragma unroll
    for (int i = 0; i < 4; ++i) sm = f4add(sm, f4sel(ra == i, Pv, f4sel(rb == i, PF, f4sel(rc == i, Pw, PG))));
    q = f4div(f4add(f4div(sm, 4.0f), f4mul(Pe, 2.0f)), 4.0f);
  } else {
    q = f4add(f4add(f4mul(Pe, 0.75f), f4mul(Pw, 0.125f)), f4mul(Pv, 0.125f));
  }
  out[ep0 + e1] = q;
  return q4;
}

// vertex point of a vertex v of M in the output: row 4t for its corners t in
// order, faces t, in-neighbours ep(e0(t)); on the boundary 3/4 P1[v] plus 1/8
// of ep(e1(t)) across every boundary out-edge and of ep(e0(t)) across every
// boundary in-edge
template <int SCHEME, bool BND>
__device__ __forceinline__ void final_vertex(const FinalView& V, const FinalSeg& S, int32_t v) {
  const int32_t cb = (int32_t)S.cb;
  const int32_t ep0 = S.nv + S.nf;
  const int32_t emax = S.ne > 0 ? S.ne - 1 : 0;
  const float4* __restrict__ P1 = S.P1;
  const int2* __restrict__ ci = V.ci + cb;  // mesh-local corners from here on
  const int32_t p0 = V.tptr[S.vb + v];
  const int n = V.tptr[S.vb + v + 1] - p0;
  // the row's corners (union corners): (corner, in-neighbour) pairs, or
  // corners alone
  const int32_t* __restrict__ rx = V.trow_x ? V.trow_x + p0 : reinterpret_cast<const int32_t*>(V.trow + p0);
  const int rs = V.trow_x ? 1 : 2;
  const float4 Pv = P1[v];
  auto prevc = [](int32_t t) __attribute__((always_inline)) { return (t & ~3) | ((t + 3) & 3); };
  // face point of M1 face t = (v, ep(e1(t)), fp(face of t), ep(e0(t)))
  auto fp_of = [&](int32_t t, int2 c1, float4 pe0) __attribute__((always_inline)) {
    return face_point(Pv, P1[ep0 + clampi(c1.x, emax)], P1[S.nv + (t >> 2)], pe0);
  };
  int32_t tt[kVertexRegs];
  int2 cp[kVertexRegs], co[kVertexRegs];
  float4 a[kVertexRegs], b[kVertexRegs];
#pragma unroll
  for (int i = 0; i < kVertexRegs; ++i)
    if (i < n) tt[i] = rx[rs * i] - cb;
#pragma unroll
  for (int i = 0; i < kVertexRegs; ++i)
    if (i < n) {
      cp[i] = ci[prevc(tt[i])];
      co[i] = ci[tt[i]];
    }
#pragma unroll
  for (int i = 0; i < kVertexRegs; ++i)
    if (i < n) b[i] = P1[ep0 + clampi(cp[i].x, emax)];
  // ep(e1(t_i)) is ep(e0) of the face across e1(t_i), another face of the
  // row (but on the boundary): taken from b when an in-edge of the row has
  // the same edge id (the same position, so the same bits), else loaded
#pragma unroll
  for (int i = 0; i < kVertexRegs; ++i)
    if (i < n) {
      int m = -1;
#pragma unroll
      for (int j = 0; j < kVertexRegs; ++j)
        if (j < n && cp[j].x == co[i].x) m = j;
      float4 e1 = b[0];
#pragma unroll
      for (int j = 1; j < kVertexRegs; ++j) e1 = f4sel(m == j, b[j], e1);
      if (m < 0) e1 = P1[ep0 + clampi(co[i].x, emax)];
      a[i] = face_point(Pv, e1, P1[S.nv + (tt[i] >> 2)], b[i]);
    }
  bool boundary = false;
  if (BND) {
#pragma unroll
    for (int i = 0; i < kVertexRegs; ++i)
      if (i < n) boundary |= co[i].y < 0 || cp[i].y < 0;
    for (int i = kVertexRegs; i < n; ++i) {
      const int32_t t = rx[rs * i] - cb;
      boundary |= ci[t].y < 0 || ci[prevc(t)].y < 0;
    }
  }
  float4 q;
  if (!boundary) {
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int i = 0; i < kVertexRegs; ++i)
      if (i < n) s = f4add(s, a[i]);
    for (int i = kVertexRegs; i < n; ++i) {
      const int32_t t = rx[rs * i] - cb;
      s = f4add(s, fp_of(t, ci[t], P1[ep0 + clampi(ci[prevc(t)].x, emax)]));
    }
#pragma unroll
    for (int i = 0; i < kVertexRegs; ++i)
      if (i < n) s = f4add(s, b[i]);
    for (int i = kVertexRegs; i < n; ++i) s = f4add(s, P1[ep0 + clampi(ci[prevc(rx[rs * i] - cb)].x, emax)]);
    const float nf = (float)n;
    q = f4div(f4add(f4div(s, nf), f4mul(Pv, nf - 2.0f)), nf);
  } else {
    q = f4mul(Pv, 0.75f);
    for (int i = 0; i < n; ++i) {
      const int2 c = ci[rx[rs * i] - cb];
      if (c.y < 0) q = f4add(q, f4mul(P1[ep0 + clampi(c.x, emax)], 0.125f));
    }
    for (int i = 0; i < n; ++i) {
      const int2 c = ci[prevc(rx[rs * i] - cb)];
      if (c.y < 0) q = f4add(q, f4mul(P1[ep0 + clampi(c.x, emax)], 0.125f));
    }
  }
  S.out_verts[v] = q;
}

// Corner items and vertex items interleaved in proportion, as in k_level.
// The vertex blocks are split the same way between the runs of V.vseg (run s
// against the runs after it): the vertex classes of M (the original vertices,
// then the face points and edge points of every earlier level), each numbered
// in the order of the original faces it descends from, so that every run
// advances over the same part of the surface as the corner blocks and the
// positions they gather are fetched once for both. Several meshes take their
// blocks from V.units, which interleaves each mesh's corner and vertex units
// the same way. A corner block stages its 16 KB of output quads in LDS and
// stores them with consecutive lanes on consecutive 16-byte pieces.
template <int SCHEME, bool BND, bool MULTI, bool IDS>
__global__ void __launch_bounds__(kBlock) k_final(FinalView V, int32_t nbc, int32_t nbv) {
  const int64_t b = block_index<kRemapFinal>();
  FinalSeg S;
  int32_t t0, cnt;
  if (MULTI) {  // work unit b (block-uniform): corners or vertices of one mesh
    const int4 U = V.units[b];
    S = V.segs[U.x];
    S.P1 = to_global(S.P1);
    S.out_verts = to_global(S.out_verts);
    S.out_ids = to_global(S.out_ids);
    S.out_ptr = to_global(S.out_ptr);
    S.out_vals = to_global(S.out_vals);
    if (U.w) {
      if ((int)threadIdx.x < U.z) final_vertex<SCHEME, BND>(V, S, U.y + (int32_t)threadIdx.x);
      return;
    }
    t0 = U.y;
    cnt = U.z;
  } else {
    const int64_t T = (int64_t)nbc + nbv;
    const int64_t before = b * nbv / T, after = (b + 1) * nbv / T;
    if (after > before) {
      int64_t u = before, rest = nbv;
      int s = 0;
      for (; s + 1 < V.nvseg; ++s) {
        const int64_t nb = V.vnb[s];
        const int64_t lo = u * nb / rest, hi = (u + 1) * nb / rest;
        if (hi > lo) {
          u = lo;
          break;
        }
        u -= lo;
        rest -= nb;
      }
      const int64_t v = (int64_t)V.vseg[s] + u * blockDim.x + threadIdx.x;
      if (v < V.vseg[s + 1]) final_vertex<SCHEME, BND>(V, V.one, (int32_t)v);
      return;
    }
    S = V.one;
    t0 = (int32_t)((b - before) * kBlock);
    cnt = (int32_t)min((int64_t)kBlock, V.nnz - t0);
  }
  const int32_t t = t0 + (int32_t)threadIdx.x;
  if (!IDS) {  // positions only
    if ((int)threadIdx.x < cnt) final_corner<SCHEME, BND, false, MULTI>(V, S, t);
    return;
  }
  __shared__ int4 stage[kStageStride * kBlock];
  Quads4 q4{};
  if ((int)threadIdx.x < cnt) q4 = final_corner<SCHEME, BND, true, MULTI>(V, S, t);
  if ((int)threadIdx.x < cnt) {
    int4* __restrict__ o = stage + kStageStride * threadIdx.x;
    o[0] = q4.a;
    o[1] = q4.b;
    o[2] = q4.c;
    o[3] = q4.d;
  }
  // the barrier keeps the block's waves together: their stores leave as one
  // 16 KB run (wave-local staging measured 1 % slower on the batch)
  __syncthreads();
  const int nq = 4 * cnt;
  int32_t* __restrict__ dst = S.out_ids + 16 * (int64_t)t0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int i = j * kBlock + threadIdx.x;
    if (i < nq) {
      const int4 x = stage[kStageStride * (i >> 2) + (i & 3)];
      store_nt(dst + 4 * i, x.x, x.y, x.z, x.w);
    }
  }
  // LAK form of the block's 4 x 256 output quads (createQuadColPtr /
  // createQuadVals): every int4 of vals is (1, 2, 3, 4) and int4 q of ptr
  // is 16q + (0, 4, 8, 12), so consecutive lanes take consecutive pieces.
  // Several meshes: written by k_lak_form after this launch instead.
  if (MULTI) return;
  if (S.out_vals) {
    int32_t* __restrict__ vals = S.out_vals + 16 * (int64_t)t0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i = j * kBlock + threadIdx.x;
      if (i < nq) store_nt(vals + 4 * i, 1, 2, 3, 4);
    }
  }
  if (S.out_ptr) {
    if ((int)threadIdx.x < cnt) store_nt(S.out_ptr + 4 * (int64_t)t, 16 * t, 16 * t + 4, 16 * t + 8, 16 * t + 12);
    if (t == S.nnz - 1) store_nt(S.out_ptr + 4 * (int64_t)S.nnz, 16 * S.nnz);
  }
}


// The LAK form of the output quads of a multi-mesh k_final launch
// (createQuadColPtr / createQuadVals, LAKCatmullClarkKernels.cuh:396-431),
// from the same corner units: a block writes the 1024 quads of its unit's
// 256 corners, vals as int4 (1, 2, 3, 4), ptr as int4 16t + (0, 4, 8, 12).
// A pure streaming write, apart from the gathers and scattered stores of
// k_final; each block takes every gridDim.x-th unit (a block per unit spends
// more on dispatch than on its 20 KB of stores).
__global__ void __launch_bounds__(kBlock) k_lak_form(FinalView V) {
  for (int64_t u = block_index(); u < V.nunits; u += gridDim.x) {
    const int4 U = V.units[u];
    if (U.w) continue;  // vertex unit
    const FinalSeg& Sg = V.segs[U.x];
    int32_t* const out_vals = to_global(Sg.out_vals);
    int32_t* const out_ptr = to_global(Sg.out_ptr);
    const int32_t t0 = U.y, cnt = U.z, nnz = Sg.nnz;
    const int32_t t = t0 + (int32_t)threadIdx.x;
    const int nq = 4 * cnt;
    if (out_vals) {
      int32_t* __restrict__ vals = out_vals + 16 * (int64_t)t0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int i = j * kBlock + threadIdx.x;
        if (i < nq) store_nt(vals + 4 * i, 1, 2, 3, 4);
      }
    }
    if (out_ptr) {
      if ((int)threadIdx.x < cnt) store_nt(out_ptr + 4 * (int64_t)t, 16 * t, 16 * t + 4, 16 * t + 8, 16 * t + 12);
      if (t == nnz - 1) store_nt(out_ptr + 4 * (int64_t)nnz, 16 * nnz);
    }
  }
}

// ---------------------------------------------------------------------------
// reduced / unreduced conversions of homogeneous meshes (GPU::CSCMesh::reduce /
// unreduce, include/Mesh.h:305-377): each face's (ids, vals) reordered by vals
// (reduce: ids in corner-position order) or by ids (unreduce), in place.

template <bool KEY_VALS>
__global__ void __launch_bounds__(kBlock) k_sort_faces(const int32_t* __restrict__ ptr, int32_t stride, int32_t nf,
                                                       int64_t nnz, int32_t* __restrict__ ids,
                                                       int32_t* __restrict__ vals, int* __restrict__ flags) {
  const int32_t f = (int32_t)(block_index() * blockDim.x + threadIdx.x);
  if (f >= nf) return;
  const int64_t s = ptr ? ptr[f] : (int64_t)stride * f;
  const int64_t e = ptr ? ptr[f + 1] : s + stride;
  const int n = (int)(e - s);
  if (s < 0 || e > nnz || n < 0 || n > 4) {  // homogeneous faces have 3 or 4 corners
    atomicOr(flags, kFlagBadPtr);
    return;
  }
  int32_t k[4], a[4], b[4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (i < n) {
      a[i] = ids[s + i];
      b[i] = vals[s + i];
      k[i] = KEY_VALS ? b[i] : a[i];
    }
  // insertion sort by key, carrying (id, val)
#pragma unroll
  for (int i = 1; i < 4; ++i)
#pragma unroll
    for (int j = i; j > 0; --j)
      if (i < n && k[j - 1] > k[j]) {
        int32_t x = k[j]; k[j] = k[j - 1]; k[j - 1] = x;
        x = a[j]; a[j] = a[j - 1]; a[j - 1] = x;
        x = b[j]; b[j] = b[j - 1]; b[j - 1] = x;
      }
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (i < n) {
      ids[s + i] = a[i];
      vals[s + i] = b[i];
    }
}

// vals[i] = 1-based corner position (i mod stride) + 1; ptr[f] = stride f
__global__ void k_fill_homogeneous(int32_t* __restrict__ ptr, int64_t nptr, int32_t* __restrict__ vals, int64_t nvals,
                                   int32_t stride) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ptr && i < nptr) ptr[i] = (int32_t)(stride * i);
  if (vals && i < nvals) vals[i] = (int32_t)(i % stride) + 1;
}

// ---------------------------------------------------------------------------
// exclusive scan with total: out[i] = sum(in[0..i)), out[n] = sum(in)

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

__device__ __forceinline__ int64_t block_exclusive_scan(int64_t x, int64_t* lds, int64_t& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t v = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int64_t y = __shfl_up(v, d, 64);
    if (lane >= d) v += y;
  }
  if (lane == 63) lds[wave] = v;
  __syncthreads();
  int64_t off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    int64_t t = lds[w];
    off += (w < wave) ? t : 0;
    tot += t;
  }
  __syncthreads();
  total = tot;
  return off + v - x;
}

// Inputs up to kScanSmall items: one block of 1024 threads in one launch
// (instead of three).
constexpr int kScanSmallThreads = 1024;
constexpr int kScanSmallRounds = 8;  // int4 loads per thread
constexpr int kScanSmallRound = 4 * kScanSmallThreads;
constexpr int64_t kScanSmall = (int64_t)kScanSmallRounds * kScanSmallRound;

// 64-bit lane exchange by DPP (row_shr / row_bcast); lanes without a source,
// and rows outside ROWMASK, read 0.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int64_t dpp64(int64_t x) {
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)x, CTRL, ROWMASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(x >> 32), CTRL, ROWMASK, 0xF, false);
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
// Inclusive scan over the 64 lanes of a wave: shifts by 1, 2, 4, 8 inside
// each row of 16 lanes, then lane 15 of rows 0 / 2 into rows 1 / 3 and lane
// 31 into rows 2 and 3.
__device__ __forceinline__ int64_t wave_scan64(int64_t x) {
  x += dpp64<0x111, 0xF>(x);
  x += dpp64<0x112, 0xF>(x);
  x += dpp64<0x114, 0xF>(x);
  x += dpp64<0x118, 0xF>(x);
  x += dpp64<0x142, 0xA>(x);
  x += dpp64<0x143, 0xC>(x);
  return x;
}

// Item (r, tid, j) = r * 4096 + 4 tid + j: every load and store instruction
// of a wave covers 1 KB of consecutive ints, all loads are in flight before
// the first add, the eight rounds are scanned side by side in registers, and
// wave 0 scans the 128 wave totals (rounds major) between two barriers.
__global__ void __launch_bounds__(kScanSmallThreads) k_scan_small(const int32_t* __restrict__ in, int64_t n,
                                                                  int32_t* __restrict__ out, int* __restrict__ flags,
                                                                  int32_t* __restrict__ total_out) {
  constexpr int kWaves = kScanSmallThreads / 64;
  static_assert(kScanSmallRounds * kWaves == 128, "wave 0 scans two wave totals per lane");
  __shared__ int64_t wsum[kScanSmallRounds * kWaves];
  __shared__ int64_t total;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool vin = (reinterpret_cast<uintptr_t>(in) & 15) == 0;
  const bool vout = (reinterpret_cast<uintptr_t>(out) & 15) == 0;
  int4 x[kScanSmallRounds];
#pragma unroll
  for (int r = 0; r < kScanSmallRounds; ++r) {
    const int64_t i = (int64_t)r * kScanSmallRound + 4 * tid;
    if (vin && i + 3 < n) {
      x[r] = *reinterpret_cast<const int4*>(in + i);
    } else {
      x[r].x = i < n ? in[i] : 0;
      x[r].y = i + 1 < n ? in[i + 1] : 0;
      x[r].z = i + 2 < n ? in[i + 2] : 0;
      x[r].w = i + 3 < n ? in[i + 3] : 0;
    }
  }
  int64_t g[kScanSmallRounds], v[kScanSmallRounds];
#pragma unroll
  for (int r = 0; r < kScanSmallRounds; ++r) {
    g[r] = (int64_t)x[r].x + x[r].y + x[r].z + x[r].w;
    v[r] = wave_scan64(g[r]);
  }
  if (lane == 63) {
#pragma unroll
    for (int r = 0; r < kScanSmallRounds; ++r) wsum[r * kWaves + wave] = v[r];
  }
  __syncthreads();
  if (wave == 0) {  // 128 wave totals, two per lane: exclusive prefixes in place
    const int64_t a = wsum[2 * lane], b = wsum[2 * lane + 1];
    const int64_t inc = wave_scan64(a + b);
    wsum[2 * lane] = inc - a - b;
    wsum[2 * lane + 1] = inc - b;
    if (lane == 63) total = inc;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kScanSmallRounds; ++r) {
    const int64_t e0 = wsum[r * kWaves + wave] + v[r] - g[r];
    const int4 o = make_int4((int32_t)e0, (int32_t)(e0 + x[r].x), (int32_t)(e0 + x[r].x + x[r].y),
                             (int32_t)(e0 + x[r].x + x[r].y + x[r].z));
    const int64_t i = (int64_t)r * kScanSmallRound + 4 * tid;
    if (vout && i + 3 < n) {
      *reinterpret_cast<int4*>(out + i) = o;
    } else {
      if (i < n) out[i] = o.x;
      if (i + 1 < n) out[i + 1] = o.y;
      if (i + 2 < n) out[i + 2] = o.z;
      if (i + 3 < n) out[i + 3] = o.w;
    }
  }
  if (tid == 0) {
    const int64_t carry = total;
    if (carry > 0x7fffffffll) atomicOr(flags, kFlagOverflow);
    out[n] = (int32_t)carry;
    if (total_out) *total_out = (int32_t)carry;
  }
}

// Tiles of kScanTile items: item (r, tid, j) of a tile is r * 4 kBlock + 4 tid
// + j (r < kScanRounds, j < 4), so that every load and store instruction of a
// wave covers 1 KB of consecutive ints.
constexpr int kScanRounds = kScanItems / 4;
// up to this many tiles (8 M items) the down pass sums the preceding tile
// sums itself (at most 16 per lane) instead of a partials launch
constexpr int64_t kScanFoldTiles = 4096;
static_assert(kScanRounds * 4 * kBlock == kScanTile, "tile layout");

__device__ __forceinline__ int4 scan_load4(const int32_t* __restrict__ in, int64_t i, int64_t n, bool vec) {
  if (vec && i + 3 < n) return *reinterpret_cast<const int4*>(in + i);
  return make_int4(i < n ? in[i] : 0, i + 1 < n ? in[i + 1] : 0, i + 2 < n ? in[i + 2] : 0, i + 3 < n ? in[i + 3] : 0);
}

__global__ void __launch_bounds__(kBlock) k_scan_reduce(const int32_t* __restrict__ in, int64_t n,
                                                        int64_t* __restrict__ partial) {
  __shared__ int64_t lds[kBlock / 64];
  const bool vec = (reinterpret_cast<uintptr_t>(in) & 15) == 0;
  const int64_t base = (int64_t)blockIdx.x * kScanTile + 4 * threadIdx.x;
  int4 x[kScanRounds];
#pragma unroll
  for (int r = 0; r < kScanRounds; ++r) x[r] = scan_load4(in, base + (int64_t)r * 4 * kBlock, n, vec);
  int64_t s = 0;
#pragma unroll
  for (int r = 0; r < kScanRounds; ++r) s += (int64_t)x[r].x + x[r].y + x[r].z + x[r].w;
  int64_t tot;
  block_exclusive_scan(s, lds, tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kBlock) k_scan_partials(int64_t* __restrict__ partial, int64_t ntiles,
                                                          int32_t* __restrict__ out, int64_t n, int* __restrict__ flags,
                                                          int32_t* __restrict__ total_out) {
  __shared__ int64_t lds[kBlock / 64];
  int64_t carry = 0;
  for (int64_t b = 0; b < ntiles; b += kBlock) {
    int64_t i = b + threadIdx.x;
    int64_t x = i < ntiles ? partial[i] : 0;
    int64_t tot;
    int64_t ex = block_exclusive_scan(x, lds, tot);
    if (i < ntiles) partial[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) {
    if (carry > 0x7fffffffll) atomicOr(flags, kFlagOverflow);
    out[n] = (int32_t)carry;
    if (total_out) *total_out = (int32_t)carry;
  }
}

// FOLD: partial holds the tile sums (no partials pass); each block adds up
// the sums of the tiles before it, and block 0 also forms the total
// (out[n], total_out, the overflow flag). Without FOLD partial holds the
// exclusive tile prefixes.
template <bool FOLD>
__global__ void __launch_bounds__(kBlock) k_scan_down(const int32_t* __restrict__ in, int64_t n,
                                                      const int64_t* __restrict__ partial, int32_t* __restrict__ out,
                                                      int64_t ntiles, int* __restrict__ flags,
                                                      int32_t* __restrict__ total_out) {
  constexpr int kWaves = kBlock / 64;
  __shared__ int64_t wsum[kScanRounds * kWaves];
  __shared__ int64_t red[kBlock / 64];
  int64_t prefix = 0;
  if (FOLD) {
    int64_t acc = 0;
    for (int64_t i = threadIdx.x; i < (int64_t)blockIdx.x; i += kBlock) acc += partial[i];
    block_exclusive_scan(acc, red, prefix);
    if (blockIdx.x == 0) {
      int64_t all = 0, carry;
      for (int64_t i = threadIdx.x; i < ntiles; i += kBlock) all += partial[i];
      block_exclusive_scan(all, red, carry);
      if (threadIdx.x == 0) {
        if (carry > 0x7fffffffll) atomicOr(flags, kFlagOverflow);
        out[n] = (int32_t)carry;
        if (total_out) *total_out = (int32_t)carry;
      }
    }
  } else {
    (void)ntiles;
    (void)flags;
    (void)total_out;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool vin = (reinterpret_cast<uintptr_t>(in) & 15) == 0;
  const bool vout = (reinterpret_cast<uintptr_t>(out) & 15) == 0;
  const int64_t base = (int64_t)blockIdx.x * kScanTile + 4 * threadIdx.x;
  int4 x[kScanRounds];
#pragma unroll
  for (int r = 0; r < kScanRounds; ++r) x[r] = scan_load4(in, base + (int64_t)r * 4 * kBlock, n, vin);
  int64_t g[kScanRounds], v[kScanRounds];
#pragma unroll
  for (int r = 0; r < kScanRounds; ++r) {
    g[r] = (int64_t)x[r].x + x[r].y + x[r].z + x[r].w;
    v[r] = wave_scan64(g[r]);
  }
  if (lane == 63)
#pragma unroll
    for (int r = 0; r < kScanRounds; ++r) wsum[r * kWaves + wave] = v[r];
  __syncthreads();
  if (threadIdx.x == 0) {  // the wave totals, rounds major: exclusive prefixes in place
    int64_t run = FOLD ? prefix : partial[blockIdx.x];
    for (int k = 0; k < kScanRounds * kWaves; ++k) {
      const int64_t t = wsum[k];
      wsum[k] = run;
      run += t;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kScanRounds; ++r) {
    const int64_t e0 = wsum[r * kWaves + wave] + v[r] - g[r];
    const int4 o = make_int4((int32_t)e0, (int32_t)(e0 + x[r].x), (int32_t)(e0 + x[r].x + x[r].y),
                             (int32_t)(e0 + x[r].x + x[r].y + x[r].z));
    const int64_t i = base + (int64_t)r * 4 * kBlock;
    if (vout && i + 3 < n) {
      *reinterpret_cast<int4*>(out + i) = o;
    } else {
      if (i < n) out[i] = o.x;
      if (i + 1 < n) out[i + 1] = o.y;
      if (i + 2 < n) out[i + 2] = o.z;
      if (i + 3 < n) out[i + 3] = o.w;
    }
  }
}

}  // namespace

void init_device_options() {
  static std::once_flag once;
  std::call_once(once, [] {
    const char* e = std::getenv("SLAK_XCD_REMAP");
    int on = e ? std::atoi(e) : (kRemapLevel | kRemapFinal);
    (void)hipMemcpyToSymbol(HIP_SYMBOL(g_xcd_remap), &on, sizeof(int));
  });
}

// ---------------------------------------------------------------------------
// launchers

void launch_fill_i32(int32_t* p, int64_t n, int32_t value, hipStream_t st) {
  if (n <= 0) return;
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  k_fill_i32<<<(unsigned)g, kBlock, 0, st>>>(p, n, value);
}

void launch_gather(const SegTable& seg, int32_t nf_total, bool use_vals, int32_t* ids_pos, int32_t* fptr,
                   int32_t* cface, int* flags, int32_t* zero, int32_t nzero, int32_t* hist_faces, hipStream_t st) {
  const int32_t n = std::max(nf_total, nzero);
  if (n <= 0) return;
  k_gather<<<grid_for(n), kBlock, 0, st>>>(seg, nf_total, use_vals ? 1 : 0, ids_pos, fptr, cface, flags, zero,
                                          nzero, hist_faces);
}

void launch_gather_verts(const SegTable& seg, int32_t nv_total, float4* P, hipStream_t st) {
  if (nv_total <= 0) return;
  k_gather_verts<<<grid_for(nv_total), kBlock, 0, st>>>(seg, nv_total, P);
}

void launch_seg_bounds(const int64_t* vbase, int32_t n, const int32_t* eptr, int32_t* bounds, hipStream_t st) {
  k_seg_bounds<<<grid_for(n + 1), kBlock, 0, st>>>(vbase, n, eptr, bounds);
}

void launch_copy_probe(const float4* a, float4* b, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  if (n > 0x7fffffffll) throw Error(1, "copy probe too large");  // SLAK_ERR_INVALID_ARGUMENT
  k_copy_probe<<<grid_for(n), kBlock, 0, st>>>(a, b, n);
}

void launch_digest(const void* p, int64_t bytes, uint64_t* out, hipStream_t st) {
  k_fill_i32<<<1, 64, 0, st>>>(reinterpret_cast<int32_t*>(out), 2, 0);
  if (bytes <= 0) return;
  const int64_t nw = (bytes + 3) / 4;
  const unsigned g = (unsigned)std::min<int64_t>(kDigestBlocks, (nw + kBlock - 1) / kBlock);
  k_digest<<<g, kBlock, 0, st>>>(static_cast<const uint8_t*>(p), bytes, reinterpret_cast<unsigned long long*>(out));
}

// the valence count and transpose through the block's vertex window (default);
// SLAK_COUNT_WINDOW=0 takes one device atomic per corner (read once)
static bool count_window() {
  static const bool on = [] {
    const char* e = std::getenv("SLAK_COUNT_WINDOW");
    return !e || std::atoi(e) != 0;
  }();
  return on;
}

void launch_faces(bool quad, bool fp, bool count, const int32_t* ids, const int32_t* fptr, const float4* P, int32_t nf,
                  const SegTable& seg, int32_t* cnt, const int* guard, hipStream_t st) {
  if (nf <= 0) return;
  unsigned g = grid_for(nf);
  if (count && !fp && count_window()) {  // the valence count alone: through the block's vertex window
    if (quad)
      k_count<true><<<g, kBlock, 0, st>>>(ids, fptr, nf, cnt, guard);
    else
      k_count<false><<<g, kBlock, 0, st>>>(ids, fptr, nf, cnt, guard);
    return;
  }
#define SLAK_FACES(Q, A, B) k_faces<Q, A, B><<<g, kBlock, 0, st>>>(ids, fptr, P, nf, seg, cnt, guard)
  if (quad) {
    if (fp && count) SLAK_FACES(true, true, true);
    else if (fp) SLAK_FACES(true, true, false);
    else if (count) SLAK_FACES(true, false, true);
  } else {
    if (fp && count) SLAK_FACES(false, true, true);
    else if (fp) SLAK_FACES(false, true, false);
    else if (count) SLAK_FACES(false, false, true);
  }
#undef SLAK_FACES
}

void launch_transpose(bool quad, const int32_t* ids, const int32_t* fptr, int32_t nf, const int32_t* tptr,
                      int32_t* cnt, int32_t* tids, const int* guard, hipStream_t st) {
  if (nf <= 0) return;
  const int window = count_window() ? 1 : 0;
  if (quad)
    k_transpose<true><<<grid_for(nf), kBlock, 0, st>>>(ids, fptr, nf, tptr, cnt, tids, guard, window);
  else
    k_transpose<false><<<grid_for(nf), kBlock, 0, st>>>(ids, fptr, nf, tptr, cnt, tids, guard, window);
}

// cap: M^T row length held in registers (4 or 8); longer rows take the
// in-memory path. Chosen per level from the row-length histogram.
#define SLAK_ROW_DISPATCH(KERNEL, quad, scheme, cap, g, smem, L, st)                  \
  do {                                                                                \
    if (quad) {                                                                       \
      if (scheme == kLAK) {                                                           \
        if (cap <= 4) KERNEL<true, kLAK, 4><<<g, kBlock, smem, st>>>(L);              \
        else KERNEL<true, kLAK, 8><<<g, kBlock, smem, st>>>(L);                       \
      } else {                                                                        \
        if (cap <= 4) KERNEL<true, kSLAK, 4><<<g, kBlock, smem, st>>>(L);             \
        else KERNEL<true, kSLAK, 8><<<g, kBlock, smem, st>>>(L);                      \
      }                                                                               \
    } else {                                                                          \
      if (scheme == kLAK) {                                                           \
        if (cap <= 4) KERNEL<false, kLAK, 4><<<g, kBlock, smem, st>>>(L);             \
        else KERNEL<false, kLAK, 8><<<g, kBlock, smem, st>>>(L);                      \
      } else {                                                                        \
        if (cap <= 4) KERNEL<false, kSLAK, 4><<<g, kBlock, smem, st>>>(L);            \
        else KERNEL<false, kSLAK, 8><<<g, kBlock, smem, st>>>(L);                     \
      }                                                                               \
    }                                                                                 \
  } while (0)

// rows per block (64, 128 or 256) and LDS tile (even, 256..kRowTileMax
// entries) of the first level's row kernels
static void check_row_tiling(const LevelView& L) {
  if (L.rows_per_block != 64 && L.rows_per_block != 128 && L.rows_per_block != kBlock)
    throw Error(5, "rows per block must be 64, 128 or 256");  // SLAK_ERR_INTERNAL
  if (L.tile_entries < 256 || L.tile_entries > kRowTileMax || (L.tile_entries & 1))
    throw Error(5, "row tile out of range");  // SLAK_ERR_INTERNAL
}

void launch_symbolic(bool quad, int scheme, int cap, const LevelView& L, hipStream_t st) {
  if (L.nv <= 0) return;
  check_row_tiling(L);
  unsigned g = grid_for(L.nv, L.rows_per_block);
  const size_t smem = (size_t)L.tile_entries * (sizeof(int32_t) + sizeof(int2));
  SLAK_ROW_DISPATCH(k_symbolic, quad, scheme, cap, g, smem, L, st);
}

void launch_number_faces(bool quad, int scheme, int cap, const LevelView& L, bool faces, hipStream_t st) {
  if (!L.trow) throw Error(5, "edge numbering without the symbolic pass's rows");  // SLAK_ERR_INTERNAL
  check_row_tiling(L);
  const int32_t nbn = L.nv > 0 ? (int32_t)grid_for(L.nv, L.rows_per_block) : 0;
  const size_t smem = (size_t)L.tile_entries * sizeof(int2);
  const int32_t nbf = faces && L.nf > 0 ? (int32_t)grid_for(L.nf) : 0;
  if (nbn + nbf == 0) return;
  const unsigned g = (unsigned)(nbn + nbf);
  if (quad) {
    if (scheme == kLAK) {
      if (cap <= 4) k_number_faces<true, kLAK, 4><<<g, kBlock, smem, st>>>(L, nbn);
      else k_number_faces<true, kLAK, 8><<<g, kBlock, smem, st>>>(L, nbn);
    } else {
      if (cap <= 4) k_number_faces<true, kSLAK, 4><<<g, kBlock, smem, st>>>(L, nbn);
      else k_number_faces<true, kSLAK, 8><<<g, kBlock, smem, st>>>(L, nbn);
    }
  } else {
    if (scheme == kLAK) {
      if (cap <= 4) k_number_faces<false, kLAK, 4><<<g, kBlock, smem, st>>>(L, nbn);
      else k_number_faces<false, kLAK, 8><<<g, kBlock, smem, st>>>(L, nbn);
    } else {
      if (cap <= 4) k_number_faces<false, kSLAK, 4><<<g, kBlock, smem, st>>>(L, nbn);
      else k_number_faces<false, kSLAK, 8><<<g, kBlock, smem, st>>>(L, nbn);
    }
  }
}

void launch_spgemm_fde(bool quad, const LevelView& L, hipStream_t st) {
  if (L.nv <= 0) return;
  if (quad)
    k_spgemm_fde<true><<<grid_for(L.nv), kBlock, 0, st>>>(L);
  else
    k_spgemm_fde<false><<<grid_for(L.nv), kBlock, 0, st>>>(L);
}

template <bool QUAD, int SCHEME>
void launch_level_t(const RefineView& R, unsigned g, int32_t nbr, int32_t nbv, hipStream_t st) {
  const bool next = R.ntptr != nullptr, bnd = R.blk != nullptr;
  if (next) {
    if (bnd) k_level<QUAD, SCHEME, true, true><<<g, kBlock, 0, st>>>(R, nbr, nbv);
    else k_level<QUAD, SCHEME, true, false><<<g, kBlock, 0, st>>>(R, nbr, nbv);
  } else {
    if (bnd) k_level<QUAD, SCHEME, false, true><<<g, kBlock, 0, st>>>(R, nbr, nbv);
    else k_level<QUAD, SCHEME, false, false><<<g, kBlock, 0, st>>>(R, nbr, nbv);
  }
}

void launch_level(bool quad, int scheme, const RefineView& R0, hipStream_t st) {
  RefineView R = R0;
  static const bool runs = [] {  // SLAK_LEVEL_RUNS=0: vertex items in plain order
    const char* e = std::getenv("SLAK_LEVEL_RUNS");
    return !e || std::atoi(e) != 0;
  }();
  if (!runs || R.nvseg <= 0 || R.nvseg > kMaxVSeg || R.vseg[0] != 0 || R.vseg[R.nvseg] != R.nv) {
    R.nvseg = 1;
    R.vseg[0] = 0;
    R.vseg[1] = R.nv;
  }
  const int32_t nbr = R.nnz > 0 ? (int32_t)grid_for(R.nnz) : 0;
  int32_t nbv = 0;
  for (int s = 0; s < R.nvseg; ++s) {
    const int64_t n = (int64_t)R.vseg[s + 1] - R.vseg[s];
    if (n < 0) throw Error(5, "level: vertex runs out of order");  // SLAK_ERR_INTERNAL
    R.vnb[s] = (int32_t)((n + kBlock - 1) / kBlock);
    nbv += R.vnb[s];
  }
  if (R.nv <= 0) nbv = 0;
  if (nbr + nbv == 0) return;
  const unsigned g = (unsigned)(nbr + nbv);
  if (quad) {
    if (scheme == kLAK) launch_level_t<true, kLAK>(R, g, nbr, nbv, st);
    else launch_level_t<true, kSLAK>(R, g, nbr, nbv, st);
  } else {
    if (scheme == kLAK) launch_level_t<false, kLAK>(R, g, nbr, nbv, st);
    else launch_level_t<false, kSLAK>(R, g, nbr, nbv, st);
  }
}

void launch_final(int scheme, const FinalView& V0, bool ids, hipStream_t st) {
  FinalView V = V0;
  const bool multi = V.segs != nullptr;
  if (!multi) {
    if (V.one.nnz <= 0 || (V.one.nnz & 3) || (int64_t)V.one.nnz != 4 * (int64_t)V.one.nf)
      throw Error(5, "final level: the mesh two levels back must be quads");  // SLAK_ERR_INTERNAL
    V.nnz = V.one.nnz;
    V.nv = V.one.nv;
  }
  if (V.nnz <= 0) return;
  if (V.nv > 0x7fffffffll) throw Error(1, "level too large for one launch");
  static const bool runs = [] {  // SLAK_FINAL_RUNS=0: vertex items in plain order
    const char* e = std::getenv("SLAK_FINAL_RUNS");
    return !e || std::atoi(e) != 0;
  }();
  if (multi || !runs || V.nvseg <= 0 || V.nvseg > kMaxVSeg || V.vseg[0] != 0 || V.vseg[V.nvseg] != V.nv) {
    V.nvseg = 1;
    V.vseg[0] = 0;
    V.vseg[1] = (int32_t)V.nv;
  }
  int64_t nbv = 0;
  for (int s = 0; s < V.nvseg; ++s) {
    const int64_t n = (int64_t)V.vseg[s + 1] - V.vseg[s];
    if (n < 0) throw Error(5, "final level: vertex runs out of order");  // SLAK_ERR_INTERNAL
    V.vnb[s] = (int32_t)((n + kBlock - 1) / kBlock);
    nbv += V.vnb[s];
  }
  int64_t nbc = grid_for(V.nnz);
  if (multi) {  // one block per work unit of V.units
    if (!V.units || V.nunits <= 0) throw Error(5, "final level: no work units");  // SLAK_ERR_INTERNAL
    nbc = V.nunits;
    nbv = 0;
  }
  if ((nbc + nbv) * kBlock > 0xffffffffll) throw Error(1, "level too large for one launch");
  const unsigned g = (unsigned)(nbc + nbv);
  const bool bnd = V.blk != nullptr;
#define SLAK_FINAL(S, B, M)                                                    \
  do {                                                                         \
    if (ids)                                                                   \
      k_final<S, B, M, true><<<g, kBlock, 0, st>>>(V, (int32_t)nbc, (int32_t)nbv);  \
    else                                                                       \
      k_final<S, B, M, false><<<g, kBlock, 0, st>>>(V, (int32_t)nbc, (int32_t)nbv); \
  } while (0)
  if (scheme == kLAK) {
    if (multi) {
      if (bnd) SLAK_FINAL(kLAK, true, true);
      else SLAK_FINAL(kLAK, false, true);
    } else {
      if (bnd) SLAK_FINAL(kLAK, true, false);
      else SLAK_FINAL(kLAK, false, false);
    }
  } else {
    if (multi) {
      if (bnd) SLAK_FINAL(kSLAK, true, true);
      else SLAK_FINAL(kSLAK, false, true);
    } else {
      if (bnd) SLAK_FINAL(kSLAK, true, false);
      else SLAK_FINAL(kSLAK, false, false);
    }
  }
#undef SLAK_FINAL
}


constexpr int64_t kLakFormBlocks = 4096;  // 16 per CU

void launch_final_lak_form(const FinalView& V, hipStream_t st) {
  if (!V.segs || !V.units || V.nunits <= 0) throw Error(5, "final level: no work units");  // SLAK_ERR_INTERNAL
  k_lak_form<<<(unsigned)std::min<int64_t>(V.nunits, kLakFormBlocks), kBlock, 0, st>>>(V);
}

void launch_sort_faces(bool by_vals, const int32_t* ptr, int32_t stride, int32_t nf, int64_t nnz, int32_t* ids,
                       int32_t* vals, int* flags, hipStream_t st) {
  if (nf <= 0) return;
  if (by_vals)
    k_sort_faces<true><<<grid_for(nf), kBlock, 0, st>>>(ptr, stride, nf, nnz, ids, vals, flags);
  else
    k_sort_faces<false><<<grid_for(nf), kBlock, 0, st>>>(ptr, stride, nf, nnz, ids, vals, flags);
}

void launch_fill_homogeneous(int32_t* ptr, int64_t nptr, int32_t* vals, int64_t nvals, int32_t stride, hipStream_t st) {
  const int64_t n = std::max(ptr ? nptr : 0, vals ? nvals : 0);
  if (n <= 0) return;
  k_fill_homogeneous<<<grid_for(n), kBlock, 0, st>>>(ptr, nptr, vals, nvals, stride);
}

int64_t scan_partials_len(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

// ---------------------------------------------------------------------------
// Scratch of the entry points that allocate nothing: one file-scope device
// array (each GPU holds its own copy from module load to process exit).
// Layout: the flags word in the first 256 bytes, then the int64 scan
// partials. Host code uses the per-device address hipGetSymbolAddress
// returns, never the array's name.

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

NoAllocScratch noalloc_scratch() {
  static std::mutex mu;
  static std::vector<void*> per_dev;
  int dev = 0;
  SLAK_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(mu);
  if ((int)per_dev.size() <= dev) per_dev.resize(dev + 1, nullptr);
  if (!per_dev[dev]) {
    void* p = nullptr;
    SLAK_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    per_dev[dev] = p;
  }
  char* base = static_cast<char*>(per_dev[dev]);
  NoAllocScratch s;
  s.flags = reinterpret_cast<int*>(base);
  s.partial = reinterpret_cast<int64_t*>(base + 256);
  s.partial_len = (int64_t)((kNoAllocScratchBytes - 256) / sizeof(int64_t));
  return s;
}

int64_t noalloc_scan_max_items() {
  const int64_t partials = (int64_t)((kNoAllocScratchBytes - 256) / sizeof(int64_t));
  return (partials - 1) * kScanTile;  // scan_partials_len(n) <= partials
}

void launch_scan(const int32_t* in, int32_t* out, int64_t n, int64_t* partial, int* flags, hipStream_t st,
                 int32_t* total_out) {
  if (n <= kScanSmall) {
    k_scan_small<<<1, kScanSmallThreads, 0, st>>>(in, n, out, flags, total_out);
    return;
  }
  int64_t ntiles = (n + kScanTile - 1) / kScanTile;
  k_scan_reduce<<<(unsigned)ntiles, kBlock, 0, st>>>(in, n, partial);
  if (ntiles <= kScanFoldTiles) {  // two launches: each tile adds up the sums before it
    k_scan_down<true><<<(unsigned)ntiles, kBlock, 0, st>>>(in, n, partial, out, ntiles, flags, total_out);
    return;
  }
  k_scan_partials<<<1, kBlock, 0, st>>>(partial, ntiles, out, n, flags, total_out);
  k_scan_down<false><<<(unsigned)ntiles, kBlock, 0, st>>>(in, n, partial, out, ntiles, flags, total_out);
}

}  // namespace slak
