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

#include "slak_device.hpp"
#include "slak_internal.hpp"

namespace slak {

namespace {

constexpr int kBlock = 256;
constexpr int kRowCap = 8;  // M^T rows up to this length stay in registers

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
// first-level validation + position ordering

__global__ void k_validate(const int32_t* __restrict__ ptr, const int32_t* __restrict__ ids,
                           const int32_t* __restrict__ vals, int32_t stride, int32_t nf, int32_t nv, int64_t nnz,
                           int require_quad, int32_t* __restrict__ ids_pos, int32_t* __restrict__ cface,
                           int* __restrict__ flags) {
  int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int64_t s = ptr ? ptr[f] : (int64_t)stride * f;
  int64_t e = ptr ? ptr[f + 1] : (int64_t)stride * (f + 1);
  int bad = 0;
  if (f == 0 && s != 0) bad |= kFlagBadPtr;
  if (f == nf - 1 && e != nnz) bad |= kFlagBadPtr;
  if (s < 0 || e < s || e > nnz) {
    atomicOr(flags, bad | kFlagBadPtr);
    return;
  }
  int32_t order = (int32_t)(e - s);
  if (order < 3) bad |= kFlagBadOrder;
  if (require_quad && (order != 4 || s != 4ll * f)) bad |= kFlagBadQuad;
  uint32_t mask = 0;
  for (int64_t k = s; k < e; ++k) {
    int32_t id = ids[k];
    if (id < 0 || id >= nv) {
      bad |= kFlagBadId;
      id = 0;
    }
    int32_t pos = vals ? vals[k] : (int32_t)(k - s + 1);
    if (pos < 1 || pos > order) {
      bad |= kFlagBadVals;
      continue;
    }
    if (order <= 32) {
      uint32_t bit = 1u << (pos - 1);
      if (mask & bit) bad |= kFlagBadVals;
      mask |= bit;
    } else {
      for (int64_t k2 = s; k2 < k; ++k2)
        if ((vals ? vals[k2] : (int32_t)(k2 - s + 1)) == pos) bad |= kFlagBadVals;
    }
    ids_pos[s + pos - 1] = id;
    if (cface) cface[k] = f;
  }
  if (bad) atomicOr(flags, bad);
}

// ---------------------------------------------------------------------------
// valence count (+ face points when the output buffer exists)

template <bool QUAD, bool FP, bool COUNT>
__global__ void k_faces(const int32_t* __restrict__ ids, const int32_t* __restrict__ fptr, const float4* __restrict__ P,
                        int32_t nf, int32_t nv, float4* __restrict__ out, int32_t* __restrict__ cnt) {
  int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  if (QUAD) {
    int4 c = reinterpret_cast<const int4*>(ids)[f];
    if (FP) {
      float4 s = f4add(f4add(f4add(P[c.x], P[c.y]), P[c.z]), P[c.w]);
      out[(int64_t)nv + f] = f4div(s, 4.0f);
    }
    if (COUNT) {
      atomicAdd(&cnt[c.x], 1);
      atomicAdd(&cnt[c.y], 1);
      atomicAdd(&cnt[c.z], 1);
      atomicAdd(&cnt[c.w], 1);
    }
  } else {
    int32_t s = fptr[f], e = fptr[f + 1];
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int32_t k = s; k < e; ++k) {
      int32_t v = ids[k];
      if (FP) acc = f4add(acc, P[v]);
      if (COUNT) atomicAdd(&cnt[v], 1);
    }
    if (FP) out[(int64_t)nv + f] = f4div(acc, (float)(e - s));
  }
}

// ---------------------------------------------------------------------------
// transpose scatter: tids[tptr[v] + slot] = corner; cnt[] counts back to zero

template <bool QUAD>
__global__ void k_transpose(const int32_t* __restrict__ ids, const int32_t* __restrict__ fptr, int32_t nf,
                            const int32_t* __restrict__ tptr, int32_t* __restrict__ cnt, int32_t* __restrict__ tids) {
  int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
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

template <bool QUAD, int SCHEME>
__global__ void __launch_bounds__(kBlock) k_symbolic(LevelView L) {
  int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= L.nv) return;
  int32_t s = L.tptr[v];
  int n = L.tptr[v + 1] - s;
  int32_t count = 0;
  int bad = 0;
  if (n <= kRowCap) {
    RegRow<kRowCap, QUAD> r;
    r.load(L, s, n);
    row_symbolic<SCHEME>(r, v, count, bad);
  } else {
    sort_row_in_place(L.tids + s, n);
    MemRow<QUAD> r(L, s, n);
    row_symbolic<SCHEME>(r, v, count, bad);
  }
  L.ecnt[v] = count;
  if (bad) atomicOr(L.flags, bad);
}

template <bool QUAD, int SCHEME>
__global__ void __launch_bounds__(kBlock) k_numeric(LevelView L) {
  int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= L.nv) return;
  int32_t s = L.tptr[v];
  int n = L.tptr[v + 1] - s;
  if (n <= kRowCap) {
    RegRow<kRowCap, QUAD> r;
    r.load(L, s, n);
    row_numeric<SCHEME>(r, L, v);
  } else {
    MemRow<QUAD> r(L, s, n);
    row_numeric<SCHEME>(r, L, v);
  }
}

template <bool QUAD>
__global__ void __launch_bounds__(kBlock) k_spgemm_fde(LevelView L) {
  int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
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
// topology: refined face t = (v, ep(v->next), nv + f, ep(prev->v))

__global__ void k_topology_quad(const int32_t* __restrict__ ids, const int32_t* __restrict__ eout, int32_t nv,
                                int32_t nf, int32_t* __restrict__ rids, int32_t* __restrict__ rptr,
                                int32_t* __restrict__ rvals) {
  int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int4 c = reinterpret_cast<const int4*>(ids)[f];
  int4 e = reinterpret_cast<const int4*>(eout)[f];
  const int32_t eb = nv + nf;
  const int32_t fpid = nv + f;
  int4* o = reinterpret_cast<int4*>(rids) + 4 * (int64_t)f;
  o[0] = make_int4(c.x, eb + e.x, fpid, eb + e.w);
  o[1] = make_int4(c.y, eb + e.y, fpid, eb + e.x);
  o[2] = make_int4(c.z, eb + e.z, fpid, eb + e.y);
  o[3] = make_int4(c.w, eb + e.w, fpid, eb + e.z);
  if (rptr) {  // createQuadColPtr: ptr[t] = 4t
    int32_t b = 16 * f;
    reinterpret_cast<int4*>(rptr)[f] = make_int4(b, b + 4, b + 8, b + 12);
    if (f == nf - 1) rptr[4 * (int64_t)nf] = 16 * nf;
  }
  if (rvals) {  // createQuadVals: vals[i] = i % 4 + 1
    int4* ov = reinterpret_cast<int4*>(rvals) + 4 * (int64_t)f;
    int4 one = make_int4(1, 2, 3, 4);
    ov[0] = one;
    ov[1] = one;
    ov[2] = one;
    ov[3] = one;
  }
}

__global__ void k_topology_poly(const int32_t* __restrict__ ids, const int32_t* __restrict__ fptr,
                                const int32_t* __restrict__ eout, int32_t nv, int32_t nf, int64_t nnz,
                                int32_t* __restrict__ rids, int32_t* __restrict__ rptr, int32_t* __restrict__ rvals) {
  int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int32_t eb = nv + nf;
  int32_t s = fptr[f], e = fptr[f + 1];
  int4* o = reinterpret_cast<int4*>(rids);
  for (int32_t t = s; t < e; ++t) {
    int32_t tp = (t == s) ? e - 1 : t - 1;
    o[t] = make_int4(ids[t], eb + eout[t], nv + f, eb + eout[tp]);
    if (rptr) rptr[t] = 4 * t;
    if (rvals) reinterpret_cast<int4*>(rvals)[t] = make_int4(1, 2, 3, 4);
  }
  if (rptr && f == nf - 1) rptr[nnz] = (int32_t)(4 * nnz);
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

// ---------------------------------------------------------------------------
// launchers

void launch_fill_i32(int32_t* p, int64_t n, int32_t value, hipStream_t st) {
  if (n <= 0) return;
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  k_fill_i32<<<(unsigned)g, kBlock, 0, st>>>(p, n, value);
}

void launch_validate(const int32_t* ptr, const int32_t* ids, const int32_t* vals, int32_t stride, int32_t nf,
                     int32_t nv, int64_t nnz, bool require_quad, int32_t* ids_pos, int32_t* cface, int* flags,
                     hipStream_t st) {
  if (nf <= 0) return;
  k_validate<<<grid_for(nf), kBlock, 0, st>>>(ptr, ids, vals, stride, nf, nv, nnz, require_quad ? 1 : 0, ids_pos,
                                              cface, flags);
}

void launch_faces(bool quad, bool fp, bool count, const int32_t* ids, const int32_t* fptr, const float4* P, int32_t nf,
                  int32_t nv, float4* out, int32_t* cnt, hipStream_t st) {
  if (nf <= 0) return;
  unsigned g = grid_for(nf);
#define SLAK_FACES(Q, A, B) k_faces<Q, A, B><<<g, kBlock, 0, st>>>(ids, fptr, P, nf, nv, out, cnt)
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

void launch_symbolic(bool quad, int scheme, const LevelView& L, hipStream_t st) {
  if (L.nv <= 0) return;
  unsigned g = grid_for(L.nv);
  if (quad) {
    if (scheme == kLAK) k_symbolic<true, kLAK><<<g, kBlock, 0, st>>>(L);
    else k_symbolic<true, kSLAK><<<g, kBlock, 0, st>>>(L);
  } else {
    if (scheme == kLAK) k_symbolic<false, kLAK><<<g, kBlock, 0, st>>>(L);
    else k_symbolic<false, kSLAK><<<g, kBlock, 0, st>>>(L);
  }
}

void launch_numeric(bool quad, int scheme, const LevelView& L, hipStream_t st) {
  if (L.nv <= 0) return;
  unsigned g = grid_for(L.nv);
  if (quad) {
    if (scheme == kLAK) k_numeric<true, kLAK><<<g, kBlock, 0, st>>>(L);
    else k_numeric<true, kSLAK><<<g, kBlock, 0, st>>>(L);
  } else {
    if (scheme == kLAK) k_numeric<false, kLAK><<<g, kBlock, 0, st>>>(L);
    else k_numeric<false, kSLAK><<<g, kBlock, 0, st>>>(L);
  }
}

void launch_spgemm_fde(bool quad, const LevelView& L, hipStream_t st) {
  if (L.nv <= 0) return;
  if (quad)
    k_spgemm_fde<true><<<grid_for(L.nv), kBlock, 0, st>>>(L);
  else
    k_spgemm_fde<false><<<grid_for(L.nv), kBlock, 0, st>>>(L);
}

void launch_topology(bool quad, const int32_t* ids, const int32_t* fptr, const int32_t* eout, int32_t nv, int32_t nf,
                     int64_t nnz, int32_t* rids, int32_t* rptr, int32_t* rvals, hipStream_t st) {
  if (nf <= 0) return;
  if (quad)
    k_topology_quad<<<grid_for(nf), kBlock, 0, st>>>(ids, eout, nv, nf, rids, rptr, rvals);
  else
    k_topology_poly<<<grid_for(nf), kBlock, 0, st>>>(ids, fptr, eout, nv, nf, nnz, rids, rptr, rvals);
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
