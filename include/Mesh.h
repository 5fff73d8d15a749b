// Host container of a device mesh with the interface of the reference's
// include/Mesh.h: CPU::Helper permutation helpers (Mesh.h:36-86),
// GPU::CSCMeshCCMeshInfoConversion (Mesh.h:96-126) and GPU::CSCMesh
// (Mesh.h:128-413): construction from an OBJ file, copies, CPU access,
// reduce / unreduce and toObj. Device memory goes through the library's C-ABI
// (slak_malloc / slak_memcpy_* / slak_free), reduce / unreduce run on the GPU
// (slak_mesh_reduce / slak_mesh_unreduce). Errors throw like the reference:
// std::runtime_error for mesh-type errors, const char* for device failures.
// Differences: a missing vals array is filled with 1-based corner positions
// by unreduce (the reference writes (c - 1) mod order), and the copy
// constructor allocates all four arrays.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <numeric>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "math/vector.h"
#include "obj.h"
#include "slak_c.h"

namespace CPU {
namespace Helper {

// permutation p that sorts vec[begin, end) by compare (p holds offsets from begin)
template <typename T, typename Compare>
std::vector<std::size_t> sort_permutation(const std::vector<T>& vec, Compare& compare, int begin, int end) {
  std::vector<std::size_t> p(static_cast<std::size_t>(end - begin));
  std::iota(p.begin(), p.end(), std::size_t{0});
  std::sort(p.begin(), p.end(), [&](std::size_t i, std::size_t j) { return compare(vec[i + begin], vec[j + begin]); });
  return p;
}

// a copy of vec with vec[begin, end) reordered by p
template <typename T>
std::vector<T> apply_permutation(const std::vector<T>& vec, const std::vector<std::size_t>& p, int begin, int end) {
  std::vector<T> out = vec;
  for (int k = 0; k < end - begin; ++k) out[begin + k] = vec[begin + p[k]];
  return out;
}

// vec[begin, end) reordered by p in place, cycle by cycle
template <typename T>
void apply_permutation_in_place(std::vector<T>& vec, const std::vector<std::size_t>& p, int begin, int end) {
  std::vector<bool> done(static_cast<std::size_t>(end - begin), false);
  for (std::size_t i = 0; i < done.size(); ++i) {
    if (done[i]) continue;
    done[i] = true;
    std::size_t prev = i, j = p[i];
    while (j != i) {
      std::swap(vec[begin + prev], vec[begin + j]);
      done[j] = true;
      prev = j;
      j = p[j];
    }
  }
}

}  // namespace Helper
}  // namespace CPU

namespace GPU {

namespace detail {
inline void succeed(int status) {
  if (status != SLAK_OK) throw slak_last_error();
}
}  // namespace detail

// Shallow conversion between the mesh structs (pointers are shared, the
// receiving struct does not own them).
template <template <typename, typename, typename, typename> class FROM_TYPE,
          template <typename, typename, typename, typename> class TO_TYPE, typename OFFSET_TYPE, typename INDEX_TYPE,
          typename VALUE_TYPE, typename VERTEX_TYPE>
static void CSCMeshCCMeshInfoConversion(const FROM_TYPE<OFFSET_TYPE, INDEX_TYPE, VALUE_TYPE, VERTEX_TYPE>& in,
                                        TO_TYPE<OFFSET_TYPE, INDEX_TYPE, VALUE_TYPE, VERTEX_TYPE>& out) {
  using InType = FROM_TYPE<OFFSET_TYPE, INDEX_TYPE, VALUE_TYPE, VERTEX_TYPE>;
  using OutType = TO_TYPE<OFFSET_TYPE, INDEX_TYPE, VALUE_TYPE, VERTEX_TYPE>;
  out.nverts = in.nverts;
  out.nfaces = in.nfaces;
  out.nnz = in.nnz;
  out.max_face_order = in.max_face_order;
  out.ptr = in.ptr;
  out.ids = in.ids;
  out.vals = in.vals;
  out.verts = in.verts;
  if (in.type == InType::MeshType::QUAD)
    out.type = OutType::MeshType::QUAD;
  else if (in.type == InType::MeshType::TRI)
    out.type = OutType::MeshType::TRI;
  else
    out.type = OutType::MeshType::POLY;
  out.is_reduced = in.is_reduced;
}

template <typename OFFSET_TYPE = int, typename INDEX_TYPE = int, typename VALUE_TYPE = int, typename VERTEX_TYPE = float>
struct CSCMesh {
  using offset_t = OFFSET_TYPE;
  using index_t = INDEX_TYPE;
  using value_t = VALUE_TYPE;
  using vertex_t = VERTEX_TYPE;
  static_assert(sizeof(offset_t) == 4 && sizeof(index_t) == 4 && sizeof(value_t) == 4 && sizeof(vertex_t) == 4,
                "the library's meshes hold 32-bit offsets, indices, values and float positions");

  using MeshType = OBJ::Data::FaceType;
  MeshType type{MeshType::POLY};

  size_t nfaces{0};
  size_t nverts{0};
  size_t nnz{0};
  size_t max_face_order{0};

  bool is_reduced{false};
  offset_t* ptr{nullptr};
  index_t* ids{nullptr};
  value_t* vals{nullptr};
  vertex_t* verts{nullptr};

  CSCMesh() = default;

  // Reads an OBJ file (include/Mesh.h:156-181): vals are the 1-based corner
  // positions, positions get w = 1, the mesh is marked reduced.
  explicit CSCMesh(const std::string& path) {
    OBJ::Data obj;
    OBJ::read(obj, path.c_str());
    type = obj.type;
    nnz = obj.f_vi.size();
    nfaces = obj.f_offs.size() - 1;
    nverts = obj.v.size();
    max_face_order = static_cast<size_t>(obj.max_face_order);
    is_reduced = true;
    std::vector<value_t> h_vals(nnz);
    for (size_t f = 0; f < nfaces; ++f)
      for (int c = obj.f_offs[f]; c < obj.f_offs[f + 1]; ++c) h_vals[c] = static_cast<value_t>(c - obj.f_offs[f] + 1);
    std::vector<math::float4> h_verts(nverts);
    for (size_t i = 0; i < nverts; ++i) h_verts[i] = math::float4(obj.v[i], 1.0f);
    std::vector<offset_t> h_ptr(obj.f_offs.begin(), obj.f_offs.end());
    std::vector<index_t> h_ids(obj.f_vi.begin(), obj.f_vi.end());
    setFromCPU(h_ptr, h_ids, h_vals, h_verts);
  }

  CSCMesh(const CSCMesh& rhs)
      : type(rhs.type), max_face_order(rhs.max_face_order), is_reduced(rhs.is_reduced) {
    alloc(rhs.nfaces, rhs.nverts, rhs.nnz, rhs.ptr != nullptr, rhs.vals != nullptr);
    copy(*this, rhs);
  }

  ~CSCMesh() {
    try {
      free();
    } catch (...) {
    }
  }

  CSCMesh& operator=(const CSCMesh& rhs) {
    if (&rhs == this) return *this;
    free();
    type = rhs.type;
    is_reduced = rhs.is_reduced;
    alloc(rhs.nfaces, rhs.nverts, rhs.nnz, rhs.ptr != nullptr, rhs.vals != nullptr);
    copy(*this, rhs);
    return *this;
  }

  void alloc(size_t nf, size_t nv, size_t nz, bool with_ptr = true, bool with_vals = true) {
    if (with_ptr) detail::succeed(slak_malloc(reinterpret_cast<void**>(&ptr), (nf + 1) * sizeof(offset_t)));
    detail::succeed(slak_malloc(reinterpret_cast<void**>(&ids), nz * sizeof(index_t)));
    if (with_vals) detail::succeed(slak_malloc(reinterpret_cast<void**>(&vals), nz * sizeof(value_t)));
    detail::succeed(slak_malloc(reinterpret_cast<void**>(&verts), nv * 4 * sizeof(vertex_t)));
  }

  void free() {
    for (void* p : {(void*)ptr, (void*)ids, (void*)vals, (void*)verts})
      if (p) detail::succeed(slak_free(p));
    ptr = nullptr, ids = nullptr, vals = nullptr, verts = nullptr;
    nfaces = 0, nverts = 0, nnz = 0;
  }

  static void copy(CSCMesh& dst, const CSCMesh& src) {
    if (&dst == &src) return;
    dst.nfaces = src.nfaces, dst.nverts = src.nverts, dst.nnz = src.nnz, dst.max_face_order = src.max_face_order;
    if (dst.ptr && src.ptr) detail::succeed(slak_memcpy_d2d(dst.ptr, src.ptr, (src.nfaces + 1) * sizeof(offset_t)));
    if (dst.ids && src.ids) detail::succeed(slak_memcpy_d2d(dst.ids, src.ids, src.nnz * sizeof(index_t)));
    if (dst.vals && src.vals) detail::succeed(slak_memcpy_d2d(dst.vals, src.vals, src.nnz * sizeof(value_t)));
    if (dst.verts && src.verts)
      detail::succeed(slak_memcpy_d2d(dst.verts, src.verts, src.nverts * 4 * sizeof(vertex_t)));
  }

  void accessFromCpu(std::vector<offset_t>& cpu_ptr, std::vector<index_t>& cpu_ids, std::vector<value_t>& cpu_vals,
                     std::vector<math::float4>& cpu_verts) const {
    if (ptr) {
      cpu_ptr.resize(nfaces + 1);
      detail::succeed(slak_memcpy_d2h(cpu_ptr.data(), ptr, (nfaces + 1) * sizeof(offset_t)));
    }
    if (ids) {
      cpu_ids.resize(nnz);
      detail::succeed(slak_memcpy_d2h(cpu_ids.data(), ids, nnz * sizeof(index_t)));
    }
    if (vals) {
      cpu_vals.resize(nnz);
      detail::succeed(slak_memcpy_d2h(cpu_vals.data(), vals, nnz * sizeof(value_t)));
    }
    if (verts) {
      cpu_verts.resize(nverts);
      detail::succeed(slak_memcpy_d2h(cpu_verts.data(), verts, nverts * sizeof(math::float4)));
    }
  }

  void accessFromCpu(std::vector<offset_t>& cpu_ptr, std::vector<index_t>& cpu_ids, std::vector<value_t>& cpu_vals,
                     std::vector<math::float3>& cpu_verts) const {
    std::vector<math::float4> v4;
    accessFromCpu(cpu_ptr, cpu_ids, cpu_vals, v4);
    cpu_verts.reserve(v4.size());
    for (const auto& p : v4) cpu_verts.emplace_back(p.xyz());
  }

  // Replaces the mesh with host arrays (include/Mesh.h:275-290): nverts is
  // one past the largest index, as in the reference.
  void setFromCPU(std::vector<offset_t>& cpu_ptr, std::vector<index_t>& cpu_ids, std::vector<value_t>& cpu_vals,
                  std::vector<math::float4>& cpu_verts) {
    free();
    nfaces = cpu_ptr.empty() ? 0 : cpu_ptr.size() - 1;
    nverts = cpu_ids.empty() ? 0 : static_cast<size_t>(*std::max_element(cpu_ids.begin(), cpu_ids.end())) + 1;
    nnz = cpu_ids.size();
    if (cpu_verts.size() < nverts) throw std::runtime_error("setFromCPU: fewer positions than referenced vertices");
    alloc(nfaces, nverts, nnz, !cpu_ptr.empty(), !cpu_vals.empty());
    if (ptr) detail::succeed(slak_memcpy_h2d(ptr, cpu_ptr.data(), (nfaces + 1) * sizeof(offset_t)));
    detail::succeed(slak_memcpy_h2d(ids, cpu_ids.data(), nnz * sizeof(index_t)));
    if (vals) detail::succeed(slak_memcpy_h2d(vals, cpu_vals.data(), nnz * sizeof(value_t)));
    detail::succeed(slak_memcpy_h2d(verts, cpu_verts.data(), nverts * sizeof(math::float4)));
  }

  // GPU::CSCMesh::reduce (include/Mesh.h:305-334): every face's (ids, vals)
  // in vals order
  void reduce() {
    if (type == MeshType::POLY) throw std::runtime_error("mesh type not homogeneous");
    if (is_reduced) return;
    slak_mesh m = view();
    detail::succeed(slak_mesh_reduce(&m, nullptr));
    adopt(m);
  }

  // GPU::CSCMesh::unreduce (include/Mesh.h:336-377): missing ptr / vals are
  // created, every face's (ids, vals) in ids order
  void unreduce() {
    if (type == MeshType::POLY) throw std::runtime_error("mesh type not homogeneous");
    if (!is_reduced) return;
    slak_mesh m = view();
    detail::succeed(slak_mesh_unreduce(&m, nullptr));
    adopt(m);
  }

  // GPU::CSCMesh::toObj (include/Mesh.h:380-411): written in reduced form;
  // an unreduced mesh is reduced for the write and unreduced afterwards.
  void toObj(const std::string& fname) {
    OBJ::Data data;
    data.type = type;
    const bool was_reduced = is_reduced;
    if (!is_reduced) reduce();
    if (data.type == MeshType::POLY && ptr) {
      data.f_offs.resize(nfaces + 1);
      detail::succeed(slak_memcpy_d2h(data.f_offs.data(), ptr, (nfaces + 1) * sizeof(offset_t)));
    }
    if (ids) {
      data.f_vi.resize(nnz);
      detail::succeed(slak_memcpy_d2h(data.f_vi.data(), ids, nnz * sizeof(index_t)));
    }
    if (verts) {
      std::vector<math::float4> v4(nverts);
      detail::succeed(slak_memcpy_d2h(v4.data(), verts, nverts * sizeof(math::float4)));
      data.v.reserve(v4.size());
      for (const auto& p : v4) data.v.emplace_back(p.xyz());
    }
    OBJ::write(data, fname.c_str());
    if (!was_reduced) unreduce();
  }

 private:
  slak_mesh view() const {
    slak_mesh m{};
    m.type = type == MeshType::QUAD ? SLAK_QUAD : type == MeshType::TRI ? SLAK_TRI : SLAK_POLY;
    m.is_reduced = is_reduced ? 1 : 0;
    m.nfaces = nfaces;
    m.nverts = nverts;
    m.nnz = nnz;
    m.max_face_order = max_face_order;
    m.ptr = reinterpret_cast<int32_t*>(ptr);
    m.ids = reinterpret_cast<int32_t*>(ids);
    m.vals = reinterpret_cast<int32_t*>(vals);
    m.verts = reinterpret_cast<float*>(verts);
    return m;
  }
  void adopt(const slak_mesh& m) {
    is_reduced = m.is_reduced != 0;
    ptr = reinterpret_cast<offset_t*>(m.ptr);
    ids = reinterpret_cast<index_t*>(m.ids);
    vals = reinterpret_cast<value_t*>(m.vals);
    verts = reinterpret_cast<vertex_t*>(m.verts);
  }
};

}  // namespace GPU
