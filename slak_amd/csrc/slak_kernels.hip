// HIP kernels of the Catmull-Clark level pipeline for gfx950 (MI355X).
//
// One level is:
//   validate      (first level only) check the input and bring corners into
//                 position order (sortKeyValuePairsSegmentedInPlace,
//                 LAKCatmullClarkKernels.cuh:147-180)
//   count         valence per vertex (spmv_right with ones, LAKCatmullClarkImpl.cu:347-356)
//   scan          exclusive scan (cub_wrappers.cuh:30-43)
//   transpose     M -> M^T by counting sort (csr2cscEx2, LAKCatmullClarkImpl.cu:381-423)
//   face points   left mapped SpMV fp = M^T P / order (spmv.cuh:64-81)
//   symbolic      row sizing of the mapped SpGEMM -> edges owned per vertex
//   numeric       mapped SpGEMM rows with fused epilogue: E numbering, edge
//                 points and the two right SpMVs of the vertex update
//   topology      refined quads (refineTopology[Homogeneous], LAKCatmullClarkKernels.cuh:320-431)
// All kernels are pull-based and atomic-free except the integer valence
// count and transpose slots, whose result does not depend on atomic order,
// so every output is bitwise deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <mutex>

#include "slak_device.hpp"
#include "slak_internal.hpp"

namespace slak {

namespace {

constexpr int kBlock = 256;
constexpr int kRowCap = 8;  // M^T rows up to this length stay in registers

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// With the remap, consecutive tiles (which gather from neighbouring parts of
// the mesh) land on the same XCD. Selected once per process (SLAK_XCD_REMAP).
__device__ int g_xcd_remap;

__device__ __forceinline__ int64_t block_index() {
  const int b = blockIdx.x;
  if (!g_xcd_remap) return b;
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
// HBM reference: a float4 stream copy (bandwidth probe and counter calibration)

__global__ void __launch_bounds__(kBlock) k_copy_probe(const float4* __restrict__ a, float4* __restrict__ b,
                                                        int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = a[i];
}

// ---------------------------------------------------------------------------
// first level: validate each caller mesh and gather it into the union in
// position order (sortKeyValuePairsSegmentedInPlace, LAKCatmullClarkKernels.cuh:147-180)

__global__ void k_gather(SegTable seg, int32_t nf_total, int use_vals, int32_t* __restrict__ ids_pos,
                         int32_t* __restrict__ fptr, int32_t* __restrict__ cface, int* __restrict__ flags) {
  int32_t f = (int32_t)(block_index() * blockDim.x + threadIdx.x);
  if (f >= nf_total) return;
  const int k = seg_of(seg.fbase, seg.n, f);
  const SegDesc* S = seg.d + k;
  const int32_t lf = (int32_t)(f - S->fb);
  const int32_t* sp = S->src_ptr;
  const int32_t* sid = S->src_ids;
  const int32_t* sval = use_vals ? S->src_vals : nullptr;
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
  if (order < 3) bad |= kFlagBadOrder;
  if (S->require_quad && (order != 4 || s != 4ll * lf)) bad |= kFlagBadQuad;
  const int64_t cb = S->cb;
  const int32_t vb = (int32_t)S->vb;
  uint32_t mask = 0;
  for (int64_t c = s; c < e; ++c) {
    int32_t id = sid[c];
    if (id < 0 || id >= S->nv) {
      bad |= kFlagBadId;
      id = 0;
    }
    int32_t pos = sval ? sval[c] : (int32_t)(c - s + 1);
    if (pos < 1 || pos > order) {
      bad |= kFlagBadVals;
      continue;
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
  }
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
  const int k = seg_of(seg.vbase, seg.n, v);
  const SegDesc* S = seg.d + k;
  P[v] = S->src_verts[v - S->vb];
}

// Row-length histograms of a level: M^T rows (valence, bins 0..15, 16 = more)
// and face orders (same bins), block-local in LDS then one atomic per bin.
constexpr int kHistBins = 17;

__global__ void __launch_bounds__(kBlock) k_row_hist(const int32_t* __restrict__ tptr, int32_t nv,
                                                      const int32_t* __restrict__ fptr, int32_t nf, int32_t face_order,
                                                      int32_t* __restrict__ hist) {
  __shared__ int32_t h[2 * kHistBins];
  for (int i = threadIdx.x; i < 2 * kHistBins; i += blockDim.x) h[i] = 0;
  __syncthreads();
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = i; v < nv; v += stride) {
    int n = tptr[v + 1] - tptr[v];
    atomicAdd(&h[n < kHistBins - 1 ? n : kHistBins - 1], 1);
  }
  for (int64_t f = i; f < nf; f += stride) {
    int n = fptr ? fptr[f + 1] - fptr[f] : face_order;
    atomicAdd(&h[kHistBins + (n < kHistBins - 1 ? n : kHistBins - 1)], 1);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < 2 * kHistBins; b += blockDim.x)
    if (h[b]) atomicAdd(&hist[b], h[b]);
}

// eptr at every segment's first vertex: out[k] = eptr[vbase[k]], k = 0..n
__global__ void k_seg_bounds(const int64_t* __restrict__ vbase, int32_t n, const int32_t* __restrict__ eptr,
                             int32_t* __restrict__ out) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > n) return;
  out[k] = eptr[vbase[k]];
}

// ---------------------------------------------------------------------------
// valence count (+ face points into each segment's destination)

template <bool QUAD, bool FP, bool COUNT>
__global__ void k_faces(const int32_t* __restrict__ ids, const int32_t* __restrict__ fptr, const float4* __restrict__ P,
                        int32_t nf, SegTable seg, int32_t* __restrict__ cnt) {
  int32_t f = (int32_t)(block_index() * blockDim.x + threadIdx.x);
  if (f >= nf) return;
  float4* dst = nullptr;
  if (FP) {
    const SegDesc* S = seg.d + seg_of(seg.fbase, seg.n, f);
    dst = S->dst_verts + S->nv + (f - S->fb);
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

// ---------------------------------------------------------------------------
// transpose scatter: tids[tptr[v] + slot] = corner; cnt[] counts back to zero

template <bool QUAD>
__global__ void k_transpose(const int32_t* __restrict__ ids, const int32_t* __restrict__ fptr, int32_t nf,
                            const int32_t* __restrict__ tptr, int32_t* __restrict__ cnt, int32_t* __restrict__ tids) {
  int32_t f = (int32_t)(block_index() * blockDim.x + threadIdx.x);
  if (f >= nf) return;
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

template <bool QUAD, int SCHEME, int CAP>
__global__ void __launch_bounds__(kBlock) k_symbolic(LevelView L) {
  int32_t v = (int32_t)(block_index() * blockDim.x + threadIdx.x);
  if (v >= L.nv) return;
  int32_t s = L.tptr[v];
  int n = L.tptr[v + 1] - s;
  int32_t count = 0, bcount = 0;
  int bad = 0;
  if (n <= CAP) {
    RegRow<CAP, QUAD> r;
    r.load(L, s, n);
    row_symbolic<SCHEME>(r, v, count, bcount, bad);
  } else {
    sort_row_in_place(L.tids + s, n);
    MemRow<QUAD> r(L, s, n);
    row_symbolic<SCHEME>(r, v, count, bcount, bad);
  }
  L.ecnt[v] = count;
  if (L.bcnt) L.bcnt[v] = bcount;
  if (bad) {
    atomicOr(L.flags, bad);
    if (L.seg.mflags) atomicOr(L.seg.mflags + seg_of(L.seg.vbase, L.seg.n, v), bad);
  }
}

template <bool QUAD, int SCHEME, int CAP>
__global__ void __launch_bounds__(kBlock) k_numeric(LevelView L) {
  int32_t v = (int32_t)(block_index() * blockDim.x + threadIdx.x);
  if (v >= L.nv) return;
  int32_t s = L.tptr[v];
  int n = L.tptr[v + 1] - s;
  if (n <= CAP) {
    RegRow<CAP, QUAD> r;
    r.load(L, s, n);
    row_numeric<SCHEME, QUAD>(r, L, v);
  } else {
    // the symbolic pass may have held this row in registers (another
    // capacity); the in-memory path needs it sorted (idempotent)
    sort_row_in_place(L.tids + s, n);
    MemRow<QUAD> r(L, s, n);
    row_numeric<SCHEME, QUAD>(r, L, v);
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

// ---------------------------------------------------------------------------
// topology: refined face t = (v, ep(v->next), nv + f, ep(prev->v)) in the
// mesh's local numbering, plus the segment's id offset

__global__ void k_topology_quad(const int32_t* __restrict__ ids, const int32_t* __restrict__ eout, int32_t nf,
                                SegTable seg, int32_t* __restrict__ L_ntptr, int32_t* __restrict__ L_ntids) {
  int32_t f = (int32_t)(block_index() * blockDim.x + threadIdx.x);
  if (f >= nf) return;
  const SegDesc* S = seg.d + seg_of(seg.fbase, seg.n, f);
  const int32_t lf = (int32_t)(f - S->fb);
  const int32_t vb = (int32_t)S->vb;
  const int32_t off = (int32_t)S->id_offset;
  const int32_t eb = S->nv + S->nf + off;
  const int32_t fpid = S->nv + lf + off;
  int4 c = reinterpret_cast<const int4*>(ids)[f];
  int4 e = reinterpret_cast<const int4*>(eout)[f];
  // every corner's edge id was written by the numeric pass; the clamp keeps
  // refined ids inside the refined vertex range whatever eout holds
  const int32_t emax = S->ne > 0 ? (int32_t)S->ne - 1 : 0;
  e.x = min(max(e.x, 0), emax);
  e.y = min(max(e.y, 0), emax);
  e.z = min(max(e.z, 0), emax);
  e.w = min(max(e.w, 0), emax);
  c.x += off - vb;
  c.y += off - vb;
  c.z += off - vb;
  c.w += off - vb;
  int4* o = reinterpret_cast<int4*>(S->dst_ids) + 4 * (int64_t)lf;
  o[0] = make_int4(c.x, eb + e.x, fpid, eb + e.w);
  o[1] = make_int4(c.y, eb + e.y, fpid, eb + e.x);
  o[2] = make_int4(c.z, eb + e.z, fpid, eb + e.y);
  o[3] = make_int4(c.w, eb + e.w, fpid, eb + e.z);
  if (S->next_cb >= 0) {  // face point's row of the next level: corner 2 of its 4 refined quads
    const int64_t row = S->next_cb + S->nnz + 4 * (int64_t)lf;
    L_ntptr[S->next_vb + S->nv + lf] = (int32_t)row;
    const int32_t c0 = (int32_t)(S->next_cb + 16 * (int64_t)lf + 2);
    *reinterpret_cast<int4*>(L_ntids + row) = make_int4(c0, c0 + 4, c0 + 8, c0 + 12);
  }
  if (S->dst_ptr) {  // createQuadColPtr: ptr[t] = 4t
    int32_t b = 16 * lf;
    reinterpret_cast<int4*>(S->dst_ptr)[lf] = make_int4(b, b + 4, b + 8, b + 12);
    if (lf == S->nf - 1) S->dst_ptr[4 * (int64_t)S->nf] = 16 * S->nf;
  }
  if (S->dst_vals) {  // createQuadVals: vals[i] = i % 4 + 1
    int4* ov = reinterpret_cast<int4*>(S->dst_vals) + 4 * (int64_t)lf;
    int4 one = make_int4(1, 2, 3, 4);
    ov[0] = one;
    ov[1] = one;
    ov[2] = one;
    ov[3] = one;
  }
}

__global__ void k_topology_poly(const int32_t* __restrict__ ids, const int32_t* __restrict__ fptr,
                                const int32_t* __restrict__ eout, int32_t nf, SegTable seg,
                                int32_t* __restrict__ L_ntptr, int32_t* __restrict__ L_ntids) {
  int32_t f = (int32_t)(block_index() * blockDim.x + threadIdx.x);
  if (f >= nf) return;
  const SegDesc* S = seg.d + seg_of(seg.fbase, seg.n, f);
  const int32_t lf = (int32_t)(f - S->fb);
  const int32_t vb = (int32_t)S->vb;
  const int32_t off = (int32_t)S->id_offset;
  const int32_t eb = S->nv + S->nf + off;
  const int32_t fpid = S->nv + lf + off;
  const int64_t cb = S->cb;
  int32_t s = fptr[f], e = fptr[f + 1];
  int4* o = reinterpret_cast<int4*>(S->dst_ids);
  const int32_t emax = S->ne > 0 ? (int32_t)S->ne - 1 : 0;
  for (int32_t t = s; t < e; ++t) {
    int32_t tp = (t == s) ? e - 1 : t - 1;
    int64_t lt = t - cb;
    int32_t e0 = min(max(eout[t], 0), emax), e1 = min(max(eout[tp], 0), emax);
    o[lt] = make_int4(ids[t] - vb + off, eb + e0, fpid, eb + e1);
    if (S->dst_ptr) S->dst_ptr[lt] = (int32_t)(4 * lt);
    if (S->dst_vals) reinterpret_cast<int4*>(S->dst_vals)[lt] = make_int4(1, 2, 3, 4);
    if (S->next_cb >= 0) L_ntids[S->next_cb + S->nnz + lt] = (int32_t)(S->next_cb + 4 * lt + 2);
  }
  if (S->next_cb >= 0) L_ntptr[S->next_vb + S->nv + lf] = (int32_t)(S->next_cb + S->nnz + (s - cb));
  if (S->dst_ptr && lf == S->nf - 1) S->dst_ptr[S->nnz] = (int32_t)(4 * S->nnz);
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

__global__ void __launch_bounds__(kBlock) k_scan_reduce(const int32_t* __restrict__ in, int64_t n,
                                                        int64_t* __restrict__ partial) {
  __shared__ int64_t lds[kBlock / 64];
  int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) s += in[base + k];
  int64_t tot;
  block_exclusive_scan(s, lds, tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kBlock) k_scan_partials(int64_t* __restrict__ partial, int64_t ntiles,
                                                          int32_t* __restrict__ out, int64_t n, int* __restrict__ flags) {
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
  }
}

__global__ void __launch_bounds__(kBlock) k_scan_down(const int32_t* __restrict__ in, int64_t n,
                                                      const int64_t* __restrict__ partial, int32_t* __restrict__ out) {
  __shared__ int64_t lds[kBlock / 64];
  int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int32_t x[kScanItems];
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    x[k] = (base + k < n) ? in[base + k] : 0;
    s += x[k];
  }
  int64_t tot;
  int64_t run = block_exclusive_scan(s, lds, tot) + partial[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = (int32_t)run;
    run += x[k];
  }
}

}  // namespace

void init_device_options() {
  static std::once_flag once;
  std::call_once(once, [] {
    const char* e = std::getenv("SLAK_XCD_REMAP");
    int on = e ? std::atoi(e) : 1;
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
                   int32_t* cface, int* flags, hipStream_t st) {
  if (nf_total <= 0) return;
  k_gather<<<grid_for(nf_total), kBlock, 0, st>>>(seg, nf_total, use_vals ? 1 : 0, ids_pos, fptr, cface, flags);
}

void launch_gather_verts(const SegTable& seg, int32_t nv_total, float4* P, hipStream_t st) {
  if (nv_total <= 0) return;
  k_gather_verts<<<grid_for(nv_total), kBlock, 0, st>>>(seg, nv_total, P);
}

void launch_row_hist(const int32_t* tptr, int32_t nv, const int32_t* fptr, int32_t nf, int32_t face_order,
                     int32_t* hist, hipStream_t st) {
  int64_t g = (std::max<int64_t>(nv, nf) + kBlock - 1) / kBlock;
  if (g > 1024) g = 1024;
  if (g < 1) g = 1;
  k_fill_i32<<<1, 64, 0, st>>>(hist, 2 * kHistBins, 0);
  k_row_hist<<<(unsigned)g, kBlock, 0, st>>>(tptr, nv, fptr, nf, face_order, hist);
}

void launch_copy_probe(const float4* a, float4* b, int64_t n, hipStream_t st) {
  k_copy_probe<<<8192, kBlock, 0, st>>>(a, b, n);
}

void launch_seg_bounds(const int64_t* vbase, int32_t n, const int32_t* eptr, int32_t* out, hipStream_t st) {
  k_seg_bounds<<<grid_for(n + 1), kBlock, 0, st>>>(vbase, n, eptr, out);
}

void launch_faces(bool quad, bool fp, bool count, const int32_t* ids, const int32_t* fptr, const float4* P, int32_t nf,
                  const SegTable& seg, int32_t* cnt, hipStream_t st) {
  if (nf <= 0) return;
  unsigned g = grid_for(nf);
#define SLAK_FACES(Q, A, B) k_faces<Q, A, B><<<g, kBlock, 0, st>>>(ids, fptr, P, nf, seg, cnt)
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
                      int32_t* cnt, int32_t* tids, hipStream_t st) {
  if (nf <= 0) return;
  if (quad)
    k_transpose<true><<<grid_for(nf), kBlock, 0, st>>>(ids, fptr, nf, tptr, cnt, tids);
  else
    k_transpose<false><<<grid_for(nf), kBlock, 0, st>>>(ids, fptr, nf, tptr, cnt, tids);
}

// cap: M^T row length held in registers (4 or 8); longer rows take the
// in-memory path. Chosen per level from the row-length histogram.
#define SLAK_ROW_DISPATCH(KERNEL, quad, scheme, cap, g, L, st)                       \
  do {                                                                               \
    if (quad) {                                                                      \
      if (scheme == kLAK) {                                                          \
        if (cap <= 4) KERNEL<true, kLAK, 4><<<g, kBlock, 0, st>>>(L);                \
        else KERNEL<true, kLAK, 8><<<g, kBlock, 0, st>>>(L);                         \
      } else {                                                                       \
        if (cap <= 4) KERNEL<true, kSLAK, 4><<<g, kBlock, 0, st>>>(L);               \
        else KERNEL<true, kSLAK, 8><<<g, kBlock, 0, st>>>(L);                        \
      }                                                                              \
    } else {                                                                         \
      if (scheme == kLAK) {                                                          \
        if (cap <= 4) KERNEL<false, kLAK, 4><<<g, kBlock, 0, st>>>(L);               \
        else KERNEL<false, kLAK, 8><<<g, kBlock, 0, st>>>(L);                        \
      } else {                                                                       \
        if (cap <= 4) KERNEL<false, kSLAK, 4><<<g, kBlock, 0, st>>>(L);              \
        else KERNEL<false, kSLAK, 8><<<g, kBlock, 0, st>>>(L);                       \
      }                                                                              \
    }                                                                                \
  } while (0)

void launch_symbolic(bool quad, int scheme, int cap, const LevelView& L, hipStream_t st) {
  if (L.nv <= 0) return;
  unsigned g = grid_for(L.nv);
  SLAK_ROW_DISPATCH(k_symbolic, quad, scheme, cap, g, L, st);
}

void launch_numeric(bool quad, int scheme, int cap, const LevelView& L, hipStream_t st) {
  if (L.nv <= 0) return;
  unsigned g = grid_for(L.nv);
  SLAK_ROW_DISPATCH(k_numeric, quad, scheme, cap, g, L, st);
}

void launch_spgemm_fde(bool quad, const LevelView& L, hipStream_t st) {
  if (L.nv <= 0) return;
  if (quad)
    k_spgemm_fde<true><<<grid_for(L.nv), kBlock, 0, st>>>(L);
  else
    k_spgemm_fde<false><<<grid_for(L.nv), kBlock, 0, st>>>(L);
}

void launch_topology(bool quad, const int32_t* ids, const int32_t* fptr, const int32_t* eout, int32_t nf,
                     const SegTable& seg, int32_t* ntptr, int32_t* ntids, hipStream_t st) {
  if (nf <= 0) return;
  if (quad)
    k_topology_quad<<<grid_for(nf), kBlock, 0, st>>>(ids, eout, nf, seg, ntptr, ntids);
  else
    k_topology_poly<<<grid_for(nf), kBlock, 0, st>>>(ids, fptr, eout, nf, seg, ntptr, ntids);
}

int64_t scan_partials_len(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

void launch_scan(const int32_t* in, int32_t* out, int64_t n, int64_t* partial, int* flags, hipStream_t st) {
  int64_t ntiles = (n + kScanTile - 1) / kScanTile;
  if (ntiles == 0) {
    k_fill_i32<<<1, 64, 0, st>>>(out, 1, 0);
    return;
  }
  k_scan_reduce<<<(unsigned)ntiles, kBlock, 0, st>>>(in, n, partial);
  k_scan_partials<<<1, kBlock, 0, st>>>(partial, ntiles, out, n, flags);
  k_scan_down<<<(unsigned)ntiles, kBlock, 0, st>>>(in, n, partial, out);
}

}  // namespace slak
