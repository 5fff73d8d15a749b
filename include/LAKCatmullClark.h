// Drop-in replacement of the reference's LibInclude/LAKCatmullClark.h.
//
// Same struct name, MeshInfo template and subdivideIteratively signature
// (LibInclude/LAKCatmullClark.h:33-192); the implementation calls the MI355X
// library through its C-ABI (slak_c.h) instead of CUDA. Mesh pointers are
// device pointers, the caller owns both meshes, outputs come from the
// library's allocator and are released by freeAndReset (slak_free). Errors are
// thrown as `const char*`, as the reference does (cuda_host_helpers.h:34-41).
// Differences: the input mesh is never modified (the reference re-sorts
// ids/vals in place), the stdout banner is printed only when the environment
// variable SLAK_BANNER is set, and the library is re-entrant.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <type_traits>

#include "slak_c.h"

namespace slak_dropin {

inline void succeed(int status) {
  if (status != SLAK_OK) throw slak_last_error();
}

template <class MI>
slak_mesh to_c(const MI& m) {
  slak_mesh c{};
  c.type = static_cast<int32_t>(m.type);
  c.is_reduced = m.is_reduced ? 1 : 0;
  c.nfaces = m.nfaces;
  c.nverts = m.nverts;
  c.nnz = m.nnz;
  c.max_face_order = m.max_face_order;
  c.ptr = reinterpret_cast<int32_t*>(m.ptr);
  c.ids = reinterpret_cast<int32_t*>(m.ids);
  c.vals = reinterpret_cast<int32_t*>(m.vals);
  c.verts = reinterpret_cast<float*>(m.verts);
  return c;
}

template <class MI>
void from_c(const slak_mesh& c, MI& m) {
  using MT = typename MI::MeshType;
  m.type = static_cast<MT>(c.type);
  m.is_reduced = c.is_reduced != 0;
  m.nfaces = c.nfaces;
  m.nverts = c.nverts;
  m.nnz = c.nnz;
  m.max_face_order = c.max_face_order;
  m.ptr = reinterpret_cast<typename MI::offset_t*>(c.ptr);
  m.ids = reinterpret_cast<typename MI::index_t*>(c.ids);
  m.vals = reinterpret_cast<typename MI::value_t*>(c.vals);
  m.verts = reinterpret_cast<typename MI::vertex_t*>(c.verts);
}

template <class MI>
void check_types() {
  static_assert(sizeof(typename MI::offset_t) == 4 && std::is_integral<typename MI::offset_t>::value,
                "offsets must be 32-bit integers (MeshInfo<int, int, int, float>)");
  static_assert(sizeof(typename MI::index_t) == 4 && std::is_integral<typename MI::index_t>::value,
                "indices must be 32-bit integers (MeshInfo<int, int, int, float>)");
  static_assert(sizeof(typename MI::value_t) == 4 && std::is_integral<typename MI::value_t>::value,
                "values must be 32-bit integers (MeshInfo<int, int, int, float>)");
  static_assert(std::is_same<typename MI::vertex_t, float>::value, "vertices must be float (float4 per vertex)");
}

template <class MI>
void subdivide(int scheme, const char* banner, MI const& cmesh, MI& rmesh, int target_level) {
  check_types<MI>();
  slak_mesh cin = to_c(cmesh);
  slak_mesh cout = to_c(rmesh);
  slak_stats st{};
  succeed(slak_subdivide(scheme, &cin, &cout, target_level, &st, nullptr));
  if (st.levels > 0) from_c(cout, rmesh);
  if (std::getenv("SLAK_BANNER")) {
    std::printf("%s\nSubdivision to level %d took %.2f ms. peak mem %llu MB \nCtrl. Mesh:  nf: %zu  nv: %zu "
                "\nSubd. Mesh:  nf: %zu  nv: %zu\n\n",
                banner, target_level, st.wall_ms, (unsigned long long)(st.peak_bytes / (1000 * 1000)), cmesh.nfaces,
                cmesh.nverts, rmesh.nfaces, rmesh.nverts);
  }
}

// The fields and methods of MeshInfo (LibInclude/LAKCatmullClark.h:95-156,
// LibInclude/CatmullClark.h:69-130). LAKCatmullClark::MeshInfo and
// CatmullClark::MeshInfo derive from it with themselves as Self, so they are
// distinct types with the same layout, as in the reference.
template <class Self, typename OFFSET_TYPE, typename INDEX_TYPE, typename VALUE_TYPE, typename VERTEX_TYPE>
struct MeshInfoFields {
  enum class MeshType { TRI, QUAD, POLY };
  MeshType type;

  size_t nfaces{0};
  size_t nverts{0};
  size_t nnz{0};
  size_t max_face_order{0};

  bool is_reduced{false};

  using offset_t = OFFSET_TYPE;
  OFFSET_TYPE* ptr{nullptr};

  using index_t = INDEX_TYPE;
  INDEX_TYPE* ids{nullptr};

  using value_t = VALUE_TYPE;
  VALUE_TYPE* vals{nullptr};

  using vertex_t = VERTEX_TYPE;
  VERTEX_TYPE* verts{nullptr};

  void reset() { static_cast<Self&>(*this) = Self(); }

  void freeAndReset() {
    if (ptr != nullptr) succeed(slak_free(ptr));
    if (ids != nullptr) succeed(slak_free(ids));
    if (vals != nullptr) succeed(slak_free(vals));
    if (verts != nullptr) succeed(slak_free(verts));
    static_cast<Self&>(*this) = Self();
  }

  // Buffers the caller's memory manager tracks go back to it, the rest to the library.
  template <typename MEM_MANAGER>
  void freeAndReset(MEM_MANAGER& mem) {
    if (ptr != nullptr && !mem.freeMemory(ptr)) succeed(slak_free(ptr));
    if (ids != nullptr && !mem.freeMemory(ids)) succeed(slak_free(ids));
    if (vals != nullptr && !mem.freeMemory(vals)) succeed(slak_free(vals));
    if (verts != nullptr && !mem.freeMemory(verts)) succeed(slak_free(verts));
    static_cast<Self&>(*this) = Self();
  }
};

}  // namespace slak_dropin

struct LAKCatmullClark {
  template <typename OFFSET_TYPE, typename INDEX_TYPE, typename VALUE_TYPE, typename VERTEX_TYPE>
  struct MeshInfo : slak_dropin::MeshInfoFields<MeshInfo<OFFSET_TYPE, INDEX_TYPE, VALUE_TYPE, VERTEX_TYPE>,
                                                OFFSET_TYPE, INDEX_TYPE, VALUE_TYPE, VERTEX_TYPE> {};

  template <typename MESH_INFO>
  static void subdivideIteratively(MESH_INFO const& cmesh, MESH_INFO& rmesh, int target_level) {
    slak_dropin::subdivide(SLAK_SCHEME_LAK, "==========LAK===========", cmesh, rmesh, target_level);
  }
};
