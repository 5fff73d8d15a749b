// CPU oracle for SLAK's Catmull-Clark subdivision (LAK and SLAK schemes).
//
// TEST INFRASTRUCTURE ONLY. This file is a plain C++ restatement of the
// reference algorithm, used by the tests and the CPU-baseline leg of the
// benchmark as the checker. The shipped library never includes, links or
// calls it.
//
// It follows the reference kernels with every atomic serialized in thread-id
// order, which makes the reference's run-dependent edge numbering canonical:
//   * LAK: edges ranked by (max vertex, min vertex), the order in which
//     createEVals (LibInclude/device/math/spgemm.cuh:71-100) hands out ids when
//     its atomicAdd executes in thread order;
//   * SLAK: edges ranked by their flagged half-edge in (vertex, corner) order,
//     the order of evaluateValences' atomic slots (CatmullClarkKernels.cuh:137-145)
//     and the exclusive scan of the flags (CatmullClarkImpl.cu:174-218).
//
// Everything is templated on the real type: double serially is the FP64
// oracle; float with 1 or N OpenMP threads is the CPU baseline of the same
// LAK algorithm (pull-form SpMVs, per-row sort-based mapped SpGEMM), which is
// deterministic, so its topology equals the FP64 run bit for bit.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace slak_oracle {

enum MeshType { TRI = 0, QUAD = 1, POLY = 2 };
enum Scheme { LAK = 0, SLAK = 1 };

template <class Real>
struct Mesh {
  int type = POLY;
  int64_t nv = 0, nf = 0;
  std::vector<int32_t> ptr;   // nf + 1
  std::vector<int32_t> ids;   // nnz
  std::vector<int32_t> vals;  // nnz, 1-based corner position (may be empty)
  std::vector<Real> P;        // 4 * nv (x, y, z, w)
  int64_t nnz() const { return (int64_t)ids.size(); }
};

// CSR matrix stored column-per-vertex ("CSC" in the reference's wording).
struct Csc {
  std::vector<int32_t> ptr, ids, vals;
};

// Intermediates of one level, kept for operator-level parity tests.
struct LevelDebug {
  std::vector<int32_t> order;                 // valence (nv) then face order (nf)
  Csc Mt;                                     // M^T: row v -> faces, vals = position
  Csc F, D, E;                                // compressed F, D (=F^T), E
  int64_t nedges = 0, nextern = 0;
  std::vector<int32_t> int0, int1, intid, f0, f1, ext0, ext1, extid;
  // SLAK adjacency S (CatmullClarkImpl.cu:151-226)
  std::vector<int32_t> S_ptr, S_ids, S_epids;
};

struct Options {
  int nthreads = 1;
  bool keep_debug = false;
};

inline int nthreads_of(const Options& o) { return o.nthreads < 1 ? 1 : o.nthreads; }

// Run f(lo, hi) over contiguous chunks of [0, n), one per thread.
template <class Fn>
void pchunks(int64_t n, int nt, Fn&& f) {
#ifdef _OPENMP
  if (nt > 1 && n > 1) {
#pragma omp parallel num_threads(nt)
    {
      int64_t t = omp_get_thread_num(), T = omp_get_num_threads();
      f(n * t / T, n * (t + 1) / T);
    }
    return;
  }
#endif
  f(0, n);
}

template <class Fn>
void pfor(int64_t n, int nt, Fn&& f) {
  pchunks(n, nt, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) f(i);
  });
}

// Exclusive scan with the total appended: out[n] = sum(in).
inline void exclusive_scan(const std::vector<int32_t>& in, std::vector<int32_t>& out) {
  out.assign(in.size() + 1, 0);
  int64_t acc = 0;
  for (size_t i = 0; i < in.size(); ++i) {
    out[i] = (int32_t)acc;
    acc += in[i];
  }
  if (acc > std::numeric_limits<int32_t>::max()) throw std::runtime_error("int32 overflow in scan");
  out[in.size()] = (int32_t)acc;
}

// ---------------------------------------------------------------------------
// Input checks (the reference has none; these keep the oracle well defined).
template <class Real>
void validate(const Mesh<Real>& m) {
  if ((int64_t)m.ptr.size() != m.nf + 1) throw std::runtime_error("ptr size mismatch");
  if (m.nf > 0 && m.ptr[0] != 0) throw std::runtime_error("ptr[0] != 0");
  if (m.ptr[m.nf] != m.nnz()) throw std::runtime_error("ptr[nf] != nnz");
  for (int64_t f = 0; f < m.nf; ++f)
    if (m.ptr[f + 1] - m.ptr[f] < 3) throw std::runtime_error("face with fewer than 3 vertices");
  for (int32_t v : m.ids)
    if (v < 0 || v >= m.nv) throw std::runtime_error("vertex id out of range");
  if ((int64_t)m.P.size() != 4 * m.nv) throw std::runtime_error("positions size mismatch");
}

// Corner order: LAK sorts each face by its vals (sortKeyValuePairsSegmentedInPlace,
// LAKCatmullClarkImpl.cu:326 then :101/:277 -- sort by id, then back by value,
// net effect: position order). SLAK ignores vals (CatmullClarkImpl.cu never
// reads cmesh.vals) and uses storage order.
template <class Real>
std::vector<int32_t> position_ordered_ids(const Mesh<Real>& m, bool use_vals) {
  std::vector<int32_t> out(m.ids);
  if (!use_vals || m.vals.empty()) return out;
  for (int64_t f = 0; f < m.nf; ++f) {
    int32_t s = m.ptr[f], e = m.ptr[f + 1], n = e - s;
    std::vector<char> seen(n, 0);
    for (int32_t k = s; k < e; ++k) {
      int32_t p = m.vals[k];
      if (p < 1 || p > n || seen[p - 1]) throw std::runtime_error("vals is not a permutation of 1..order");
      seen[p - 1] = 1;
      out[s + p - 1] = m.ids[k];
    }
  }
  return out;
}

// ---------------------------------------------------------------------------
// LAK (LibSource/device/LAKCatmullClarkImpl.cu)

// Valences and face orders: spmv_right / spmv_left with a ones vector
// (LAKCatmullClarkImpl.cu:347-369).
inline void lak_orders(int64_t nv, int64_t nf, const std::vector<int32_t>& ptr, const std::vector<int32_t>& ids,
                       std::vector<int32_t>& order) {
  order.assign(nv + nf, 0);
  for (int32_t v : ids) order[v] += 1;
  for (int64_t f = 0; f < nf; ++f) order[nv + f] = ptr[f + 1] - ptr[f];
}

// Stable transpose M (face rows) -> M^T (vertex rows), as cusparseCsr2cscEx2
// (LAKCatmullClarkImpl.cu:381-423): rows of M^T list faces ascending, vals
// carry the 1-based corner position.
inline void lak_transpose(int64_t nv, int64_t nf, const std::vector<int32_t>& ptr, const std::vector<int32_t>& ids,
                          Csc& Mt) {
  std::vector<int32_t> cnt(nv, 0);
  for (int32_t v : ids) cnt[v]++;
  exclusive_scan(cnt, Mt.ptr);
  Mt.ids.assign(ids.size(), 0);
  Mt.vals.assign(ids.size(), 0);
  std::vector<int32_t> fill(Mt.ptr.begin(), Mt.ptr.end() - 1);
  for (int64_t f = 0; f < nf; ++f)
    for (int32_t k = ptr[f]; k < ptr[f + 1]; ++k) {
      int32_t v = ids[k];
      int32_t slot = fill[v]++;
      Mt.ids[slot] = (int32_t)f;
      Mt.vals[slot] = k - ptr[f] + 1;
    }
}

// Mapped SpGEMM of spgemm_mapped (spgemm.cuh:102-309) with bhSPARSE's
// expand-sort-compress semantics (bhsparse_cuda.cuh:1039-1112, 951-1037):
// row v expands every product (face f of v, vertex w of f) into key w with
// value (f+1) if MappedMulFunc holds (spgemm.cuh:62-69) else 0, sorts by key
// and sums duplicate keys. mapPwr = 0 gives F (w follows v in f), mapPwr = 1
// gives D (w precedes v in f). Both share the pattern of M*M^T.
// Output (pre-compress): per row, keys ascending with F and D values.
struct RowProducts {
  std::vector<int32_t> ptr, keys, fvals, dvals;
};

inline void lak_spgemm_rows(int64_t nv, const std::vector<int32_t>& ptr, const std::vector<int32_t>& ids,
                            const Csc& Mt, RowProducts& R, int nt) {
  // upper bound on products per row (compute_nnzCt, bhsparse_cuda.cuh:364-392)
  std::vector<int32_t> ub(nv, 0);
  pfor(nv, nt, [&](int64_t v) {
    int32_t s = 0;
    for (int32_t i = Mt.ptr[v]; i < Mt.ptr[v + 1]; ++i) {
      int32_t f = Mt.ids[i];
      s += ptr[f + 1] - ptr[f];
    }
    ub[v] = s;
  });
  std::vector<int32_t> ubp;
  exclusive_scan(ub, ubp);
  std::vector<int32_t> tk(ubp[nv]), tf(ubp[nv]), td(ubp[nv]);
  std::vector<int32_t> cnt(nv, 0);
  pchunks(nv, nt, [&](int64_t lo, int64_t hi) {
    std::vector<int32_t> k, a, b;  // expanded products of one row
    for (int64_t v = lo; v < hi; ++v) {
      k.clear();
      a.clear();
      b.clear();
      for (int32_t i = Mt.ptr[v]; i < Mt.ptr[v + 1]; ++i) {
        int32_t f = Mt.ids[i];
        int32_t pv = Mt.vals[i];
        int32_t n = ptr[f + 1] - ptr[f];
        for (int32_t j = ptr[f]; j < ptr[f + 1]; ++j) {
          int32_t pw = j - ptr[f] + 1;
          k.push_back(ids[j]);
          a.push_back((pw == (pv % n) + 1) ? f + 1 : 0);                 // mapPwr = 0
          b.push_back((pw == ((pv + 1 * (n - 2)) % n) + 1) ? f + 1 : 0);  // mapPwr = 1
        }
      }
      // stable insertion sort by key (rows hold a few dozen products)
      for (size_t x = 1; x < k.size(); ++x) {
        int32_t kk = k[x], aa = a[x], bb = b[x];
        size_t y = x;
        while (y > 0 && k[y - 1] > kk) {
          k[y] = k[y - 1];
          a[y] = a[y - 1];
          b[y] = b[y - 1];
          --y;
        }
        k[y] = kk;
        a[y] = aa;
        b[y] = bb;
      }
      int32_t o = ubp[v], c = 0;
      for (size_t j = 0; j < k.size(); ++j) {
        if (c > 0 && tk[o + c - 1] == k[j]) {  // duplicate key: sum values
          tf[o + c - 1] += a[j];
          td[o + c - 1] += b[j];
        } else {
          tk[o + c] = k[j];
          tf[o + c] = a[j];
          td[o + c] = b[j];
          ++c;
        }
      }
      cnt[v] = c;
    }
  });
  exclusive_scan(cnt, R.ptr);
  R.keys.assign(R.ptr[nv], 0);
  R.fvals.assign(R.ptr[nv], 0);
  R.dvals.assign(R.ptr[nv], 0);
  pfor(nv, nt, [&](int64_t v) {
    for (int32_t c = 0; c < cnt[v]; ++c) {
      R.keys[R.ptr[v] + c] = tk[ubp[v] + c];
      R.fvals[R.ptr[v] + c] = tf[ubp[v] + c];
      R.dvals[R.ptr[v] + c] = td[ubp[v] + c];
    }
  });
}

// compress (LAKCatmullClarkKernels.cuh:217-252): drop zero values.
inline void compress(const std::vector<int32_t>& ptr, const std::vector<int32_t>& keys, const std::vector<int32_t>& vals,
                     Csc& out) {
  int64_t n = (int64_t)ptr.size() - 1;
  std::vector<int32_t> cnt(n, 0);
  for (int64_t c = 0; c < n; ++c)
    for (int32_t i = ptr[c]; i < ptr[c + 1]; ++i) cnt[c] += vals[i] != 0;
  exclusive_scan(cnt, out.ptr);
  out.ids.resize(out.ptr[n]);
  out.vals.resize(out.ptr[n]);
  for (int64_t c = 0; c < n; ++c) {
    int32_t o = out.ptr[c];
    for (int32_t i = ptr[c]; i < ptr[c + 1]; ++i)
      if (vals[i] != 0) {
        out.ids[o] = keys[i];
        out.vals[o] = vals[i];
        ++o;
      }
  }
}

inline int32_t find_in_col(const Csc& A, int32_t col, int32_t row) {
  for (int32_t j = A.ptr[col]; j < A.ptr[col + 1]; ++j)
    if (A.ids[j] == row) return j;
  return -1;
}

template <class Real>
void lak_level(const Mesh<Real>& in, Mesh<Real>& out, const Options& opt, LevelDebug* dbg) {
  const int nt = nthreads_of(opt);
  const int64_t nv = in.nv, nf = in.nf, nnz = in.nnz();
  const std::vector<int32_t>& ptr = in.ptr;
  const std::vector<int32_t> ids = position_ordered_ids(in, true);

  std::vector<int32_t> order;
  lak_orders(nv, nf, ptr, ids, order);
  Csc Mt;
  lak_transpose(nv, nf, ptr, ids, Mt);

  RowProducts R;
  lak_spgemm_rows(nv, ptr, ids, Mt, R, nt);

  // createEVals serialized in thread order (spgemm.cuh:71-100): column c,
  // rows r < c ascending; id = ++counter when F or D is nonzero; interior
  // (both nonzero) stored as -id, boundary as +id.
  std::vector<int32_t> evals(R.keys.size(), 0);
  std::vector<int32_t> ecnt(nv, 0);
  pfor(nv, nt, [&](int64_t c) {
    int32_t k = 0;
    for (int32_t i = R.ptr[c]; i < R.ptr[c + 1]; ++i) {
      if (c <= R.keys[i]) break;
      if (R.fvals[i] != 0 || R.dvals[i] != 0) ++k;
    }
    ecnt[c] = k;
  });
  std::vector<int32_t> ebase;
  exclusive_scan(ecnt, ebase);
  int64_t nedges = ebase[nv];
  pfor(nv, nt, [&](int64_t c) {
    int32_t id = ebase[c];
    for (int32_t i = R.ptr[c]; i < R.ptr[c + 1]; ++i) {
      if (c <= R.keys[i]) break;
      int32_t v0 = R.fvals[i], v1 = R.dvals[i];
      if (v0 != 0 || v1 != 0) {
        ++id;
        evals[i] = (v0 != 0 && v1 != 0) ? -id : id;
      }
    }
  });
  int64_t nextern = 0;
  for (size_t i = 0; i < evals.size(); ++i) nextern += evals[i] > 0;

  Csc F, D, E;
  compress(R.ptr, R.keys, R.fvals, F);
  compress(R.ptr, R.keys, R.dvals, D);
  compress(R.ptr, R.keys, evals, E);

  // extractEdgeInfoFromE + extractFaceInfoFromFFt (LAKCatmullClarkKernels.cuh:257-317),
  // serialized: lists in E storage order.
  std::vector<int32_t> int0, int1, intid, f0, f1, ext0, ext1, extid;
  for (int64_t c = 0; c < nv; ++c)
    for (int32_t j = E.ptr[c]; j < E.ptr[c + 1]; ++j) {
      int32_t r = E.ids[j], val = E.vals[j];
      if (val > 0) {
        ext0.push_back(r);
        ext1.push_back((int32_t)c);
        extid.push_back(val - 1);
      } else {
        int0.push_back(r);
        int1.push_back((int32_t)c);
        intid.push_back(-val - 1);
      }
    }
  const int64_t nint = (int64_t)int0.size();
  f0.assign(nint, 0);
  f1.assign(nint, 0);
  pfor(nint, nt, [&](int64_t t) {
    int32_t col = int0[t], row = int1[t];
    int32_t j = find_in_col(F, col, row);
    f0[t] = F.vals[j < 0 ? 0 : j] - 1;
    j = find_in_col(D, col, row);
    f1[t] = D.vals[j < 0 ? 0 : j] - 1;
  });

  // refineTopology (LAKCatmullClarkKernels.cuh:320-374).
  out.type = QUAD;
  out.nf = nnz;
  out.nv = nv + nf + nedges;
  out.ptr.resize(out.nf + 1);
  out.ids.resize(4 * nnz);
  out.vals.resize(4 * nnz);
  const int32_t base = (int32_t)(nv + nf);
  auto edge_of = [&](int32_t a, int32_t b) -> int32_t {
    int32_t j = find_in_col(E, std::max(a, b), std::min(a, b));
    int32_t val = E.vals[j < 0 ? 0 : j];
    return val > 0 ? val - 1 + base : -val - 1 + base;
  };
  pfor(nf, nt, [&](int64_t f) {
    int32_t s = ptr[f], e = ptr[f + 1];
    for (int32_t i = s; i < e; ++i) {
      int32_t v0 = ids[i];
      int32_t vout = (i == e - 1) ? ids[s] : ids[i + 1];
      int32_t vin = (i == s) ? ids[e - 1] : ids[i - 1];
      int32_t* q = &out.ids[4 * (int64_t)i];
      q[0] = v0;
      q[1] = edge_of(v0, vout);
      q[2] = (int32_t)(nv + f);
      q[3] = edge_of(v0, vin);
    }
  });
  for (int64_t i = 0; i <= out.nf; ++i) out.ptr[i] = (int32_t)(4 * i);  // createQuadColPtr
  for (int64_t i = 0; i < 4 * nnz; ++i) out.vals[i] = (int32_t)(i % 4 + 1);  // createQuadVals

  // subdivideVertexDataPolyMesh (LAKCatmullClarkImpl.cu:150-253), pull form.
  const std::vector<Real>& P = in.P;
  std::vector<Real>& Q = out.P;
  Q.assign(4 * out.nv, Real(0));
  Real* fp = Q.data() + 4 * nv;
  Real* ep = Q.data() + 4 * (nv + nf);
  pfor(nf, nt, [&](int64_t f) {  // spmv_left_mapped_f4 + divElementWise
    Real acc[4] = {0, 0, 0, 0};
    for (int32_t i = ptr[f]; i < ptr[f + 1]; ++i)
      for (int c = 0; c < 4; ++c) acc[c] += P[4 * (int64_t)ids[i] + c];
    Real ord = (Real)order[nv + f];
    for (int c = 0; c < 4; ++c) fp[4 * f + c] = acc[c] / ord;
  });
  pfor(nint, nt, [&](int64_t t) {  // calculateInternalEdgepoints
    for (int c = 0; c < 4; ++c)
      ep[4 * (int64_t)intid[t] + c] = (P[4 * (int64_t)int0[t] + c] + P[4 * (int64_t)int1[t] + c] +
                                       fp[4 * (int64_t)f0[t] + c] + fp[4 * (int64_t)f1[t] + c]) /
                                      Real(4);
  });
  const int64_t next = (int64_t)ext0.size();
  for (int64_t t = 0; t < next; ++t)  // calculateExternalEdgepoints
    for (int c = 0; c < 4; ++c)
      ep[4 * (int64_t)extid[t] + c] = (P[4 * (int64_t)ext0[t] + c] + P[4 * (int64_t)ext1[t] + c]) / Real(2);
  pfor(nv, nt, [&](int64_t v) {
    // multElementWise(P, order, -2); spmv_right(M, fp); spmv_right(F, P);
    // divElementWise(order); addElementWise; divElementWise(order).
    Real n = (Real)order[v];
    Real s[4] = {0, 0, 0, 0};
    for (int32_t i = Mt.ptr[v]; i < Mt.ptr[v + 1]; ++i)
      for (int c = 0; c < 4; ++c) s[c] += fp[4 * (int64_t)Mt.ids[i] + c];
    for (int32_t i = D.ptr[v]; i < D.ptr[v + 1]; ++i)  // F column w holds row v  <=>  w -> v
      for (int c = 0; c < 4; ++c) s[c] += P[4 * (int64_t)D.ids[i] + c];
    for (int c = 0; c < 4; ++c) {
      Real pn = P[4 * v + c] * (n + Real(-2));
      Q[4 * v + c] = (s[c] / n + pn) / n;
    }
  });
  // prepareExternalVertexUpdate / calculateExternalVertexUpdate
  // (LAKCatmullClarkKernels.cuh:464-511): 3/4 P, then +1/8 of the boundary neighbours.
  for (int64_t t = 0; t < next; ++t)
    for (int c = 0; c < 4; ++c) {
      Q[4 * (int64_t)ext0[t] + c] = P[4 * (int64_t)ext0[t] + c] * Real(3.0 / 4.0);
      Q[4 * (int64_t)ext1[t] + c] = P[4 * (int64_t)ext1[t] + c] * Real(3.0 / 4.0);
    }
  for (int64_t t = 0; t < next; ++t)
    for (int c = 0; c < 4; ++c) {
      Q[4 * (int64_t)ext1[t] + c] += P[4 * (int64_t)ext0[t] + c] * Real(1.0 / 8.0);
      Q[4 * (int64_t)ext0[t] + c] += P[4 * (int64_t)ext1[t] + c] * Real(1.0 / 8.0);
    }

  if (dbg) {
    dbg->order = order;
    dbg->Mt = Mt;
    dbg->F = F;
    dbg->D = D;
    dbg->E = E;
    dbg->nedges = nedges;
    dbg->nextern = nextern;
    dbg->int0 = int0;
    dbg->int1 = int1;
    dbg->intid = intid;
    dbg->f0 = f0;
    dbg->f1 = f1;
    dbg->ext0 = ext0;
    dbg->ext1 = ext1;
    dbg->extid = extid;
  }
}

// ---------------------------------------------------------------------------
// SLAK (LibSource/device/CatmullClarkImpl.cu, LibInclude/device/CatmullClarkKernels.cuh)

template <class Real>
void slak_level(const Mesh<Real>& in, Mesh<Real>& out, bool& has_boundary, const Options& opt, LevelDebug* dbg) {
  const int nt = nthreads_of(opt);
  const int64_t nv = in.nv, nf = in.nf, nnz = in.nnz();
  const std::vector<int32_t>& ptr = in.ptr;
  const std::vector<int32_t>& ids = in.ids;  // storage order (vals unused)

  // corner -> face (createThreadFaceMapping, :189-198)
  std::vector<int32_t> cface(nnz);
  for (int64_t f = 0; f < nf; ++f)
    for (int32_t i = ptr[f]; i < ptr[f + 1]; ++i) cface[i] = (int32_t)f;
  auto next_of = [&](int64_t t) -> int32_t {
    int32_t f = cface[t], s = ptr[f], n = ptr[f + 1] - s;
    return ids[s + (t + 1 - s) % n];
  };
  auto prev_of = [&](int64_t t) -> int32_t {
    int32_t f = cface[t], s = ptr[f], n = ptr[f + 1] - s;
    return ids[s + (t + n - 1 - s) % n];
  };

  // evaluateValences (:137-145), atomic slots serialized by corner index.
  std::vector<int32_t> val(nv, 0), slot(nnz);
  for (int64_t t = 0; t < nnz; ++t) slot[t] = val[ids[t]]++;
  std::vector<int32_t> Sptr;
  exclusive_scan(val, Sptr);  // CUB scan (:180, :396)
  // evaluateAdjacencies (:147-187): S row v lists next(t) in slot order, flag v < next.
  std::vector<int32_t> Sids(nnz), flags(nnz);
  for (int64_t t = 0; t < nnz; ++t) {
    int32_t v0 = ids[t], v1 = next_of(t);
    Sids[Sptr[v0] + slot[t]] = v1;
    flags[Sptr[v0] + slot[t]] = v0 < v1;
  }
  auto find_S = [&](int32_t col, int32_t row) -> int32_t {
    for (int32_t i = Sptr[col]; i < Sptr[col + 1]; ++i)
      if (Sids[i] == row) return i;
    return -1;
  };
  // checkForBoundary (:200-321): half-edge v0 -> vout without a twin.
  int64_t nextern = 0;
  if (has_boundary) {
    for (int64_t t = 0; t < nnz; ++t) {
      int32_t v0 = ids[t], vout = next_of(t);
      int32_t eid = find_S(v0, vout);
      if (find_S(vout, v0) >= 0) continue;
      ++nextern;
      flags[eid < 0 ? 0 : eid] = 1;
    }
    has_boundary = nextern != 0;
  }
  std::vector<int32_t> epid;
  exclusive_scan(flags, epid);  // scan of flags (:210, :433)
  epid.resize(nnz);
  std::vector<int32_t> everts;
  if (has_boundary) {  // markBoundary (:324-403), serialized: loc = nextern-1, nextern-2, ...
    everts.assign(nextern, 0);
    int64_t ctr = nextern;
    for (int64_t t = 0; t < nnz; ++t) {
      int32_t v0 = ids[t], vout = next_of(t);
      int32_t eid = find_S(v0, vout);
      if (find_S(vout, v0) >= 0) continue;
      everts[--ctr] = v0;
      epid[eid] = -(epid[eid] + 1);
    }
  }
  const int64_t nedges = (nnz - nextern) / 2 + nextern;

  // refineTopology[Quad][Extern] (:405-575): quads (v, ep(v,next), nv+f, ep(prev,v)).
  out.type = QUAD;
  out.nf = nnz;
  out.nv = nv + nf + nedges;
  out.ptr.resize(nnz + 1);  // kept for the next level; SLAK form drops it at the end
  for (int64_t i = 0; i <= nnz; ++i) out.ptr[i] = (int32_t)(4 * i);
  out.vals.clear();
  out.ids.resize(4 * nnz);
  const int32_t base = (int32_t)(nv + nf);
  auto edge_of = [&](int32_t a, int32_t b) -> int32_t {
    int32_t j = find_S(std::min(a, b), std::max(a, b));
    if (j < 0) j = find_S(std::max(a, b), std::min(a, b));
    int32_t s = epid[j < 0 ? 0 : j];
    return s >= 0 ? s + base : -s + base - 1;
  };
  pfor(nnz, nt, [&](int64_t t) {
    int32_t v0 = ids[t];
    int32_t* q = &out.ids[4 * t];
    q[0] = v0;
    q[1] = edge_of(v0, next_of(t));
    q[2] = (int32_t)(nv + cface[t]);
    q[3] = edge_of(v0, prev_of(t));
  });

  // Vertex data (CatmullClarkImpl.cu:37-105, 231-309).
  const std::vector<Real>& P = in.P;
  std::vector<Real>& Q = out.P;
  Q.assign(4 * out.nv, Real(0));
  Real* fp = Q.data() + 4 * nv;
  Real* ep = Q.data() + 4 * (nv + nf);
  for (int64_t f = 0; f < nf; ++f) {  // calculateFacepoints (:602-618)
    Real rcp = Real(1) / Real(ptr[f + 1] - ptr[f]);
    Real acc[4] = {0, 0, 0, 0};
    for (int32_t i = ptr[f]; i < ptr[f + 1]; ++i)
      for (int c = 0; c < 4; ++c) acc[c] += P[4 * (int64_t)ids[i] + c];
    for (int c = 0; c < 4; ++c) fp[4 * f + c] = acc[c] * rcp;
  }
  for (int64_t t = 0; t < nnz; ++t) {  // calculateEdgepoints (:620-632)
    int64_t e = (int64_t)out.ids[4 * t + 1] - base;
    int32_t v = ids[t];
    int32_t f = cface[t];
    for (int c = 0; c < 4; ++c) ep[4 * e + c] += (P[4 * (int64_t)v + c] + fp[4 * (int64_t)f + c]) * Real(1.0 / 4.0);
  }
  for (int64_t v = 0; v < nv; ++v) {  // vertexSum (:634-652)
    Real n = (Real)val[v];
    Real s[4] = {0, 0, 0, 0};
    for (int32_t i = Sptr[v]; i < Sptr[v + 1]; ++i)
      for (int c = 0; c < 4; ++c) s[c] += P[4 * (int64_t)Sids[i] + c];
    for (int c = 0; c < 4; ++c) Q[4 * v + c] = s[c] / (n * n) + P[4 * v + c] * (n - Real(2)) / n;
  }
  for (int64_t f = 0; f < nf; ++f)  // facepointSum (:671-686)
    for (int32_t i = ptr[f]; i < ptr[f + 1]; ++i) {
      int32_t r = ids[i];
      Real n = (Real)val[r];
      for (int c = 0; c < 4; ++c) Q[4 * (int64_t)r + c] += fp[4 * f + c] / (n * n);
    }
  if (has_boundary) {
    // correctBoundaryS1 / S2 (:688-740)
    std::vector<int32_t> nxt(nextern);
    for (int64_t k = 0; k < nextern; ++k) {
      int32_t v = everts[k];
      int32_t loc = 0;
      for (int32_t i = Sptr[v]; i < Sptr[v + 1]; ++i)
        if (epid[i] < 0) {
          loc = i;
          break;
        }
      int32_t e = -epid[loc] - 1;
      int32_t w = Sids[loc];
      nxt[k] = w;
      for (int c = 0; c < 4; ++c) {
        ep[4 * (int64_t)e + c] = (P[4 * (int64_t)v + c] + P[4 * (int64_t)w + c]) / Real(2);
        Q[4 * (int64_t)v + c] = P[4 * (int64_t)v + c] * Real(3.0 / 4.0) + P[4 * (int64_t)w + c] * Real(1.0 / 8.0);
      }
    }
    for (int64_t k = 0; k < nextern; ++k) {
      int32_t v = everts[k], w = nxt[k];
      for (int c = 0; c < 4; ++c) Q[4 * (int64_t)w + c] += P[4 * (int64_t)v + c] * Real(1.0 / 8.0);
    }
  }

  if (dbg) {
    dbg->order.assign(val.begin(), val.end());
    dbg->S_ptr = Sptr;
    dbg->S_ids = Sids;
    dbg->S_epids = epid;
    dbg->nedges = nedges;
    dbg->nextern = nextern;
  }
}

// ---------------------------------------------------------------------------
// subdivideIteratively (LAKCatmullClarkImpl.cu:512-571, CatmullClarkImpl.cu:457-515):
// a non-quad input always gets one poly level, then quad levels up to target.
template <class Real>
int subdivide(int scheme, const Mesh<Real>& cmesh, Mesh<Real>& rmesh, int target_level, const Options& opt,
              std::vector<LevelDebug>* dbg) {
  validate(cmesh);
  Mesh<Real> cur = cmesh;
  int level = 0;
  bool has_boundary = true;
  auto step = [&]() {
    Mesh<Real> nxt;
    LevelDebug* d = nullptr;
    if (dbg) {
      dbg->emplace_back();
      d = &dbg->back();
    }
    if (scheme == LAK)
      lak_level(cur, nxt, opt, d);
    else
      slak_level(cur, nxt, has_boundary, opt, d);
    cur = std::move(nxt);
  };
  if (cur.type != QUAD) {
    step();
    ++level;
  }
  for (; level < target_level; ++level) step();
  if (level == 0) {
    rmesh = Mesh<Real>();  // nothing refined: output untouched in the reference
    return 0;
  }
  rmesh = std::move(cur);
  if (scheme == SLAK) {  // SLAK form: ids only, ptr = vals = nullptr (CatmullClarkImpl.cu:124-125)
    rmesh.ptr.clear();
    rmesh.vals.clear();
  }
  return level;
}

}  // namespace slak_oracle
