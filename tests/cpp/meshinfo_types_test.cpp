// Compile-time checks of the drop-in headers: LAKCatmullClark::MeshInfo and
// CatmullClark::MeshInfo are distinct types (as in the reference,
// LibInclude/LAKCatmullClark.h:95-156 and LibInclude/CatmullClark.h:69-130)
// with the same fields, so code that overloads or specialises on both
// compiles. Built with -fsyntax-only by tests/test_headers.py.
#include <cstddef>
#include <type_traits>

#include "CatmullClark.h"
#include "LAKCatmullClark.h"

using LMesh = LAKCatmullClark::MeshInfo<int, int, int, float>;
using SMesh = CatmullClark::MeshInfo<int, int, int, float>;

static_assert(!std::is_same<LMesh, SMesh>::value, "the two MeshInfo types must be distinct");
static_assert(!std::is_same<LMesh::MeshType, SMesh::MeshType>::value, "each MeshInfo has its own MeshType");
static_assert(sizeof(LMesh) == sizeof(SMesh), "same layout");
static_assert(std::is_default_constructible<LMesh>::value && std::is_copy_assignable<SMesh>::value, "value types");
static_assert(std::is_same<LMesh::offset_t, int>::value && std::is_same<SMesh::vertex_t, float>::value, "typedefs");

// overloads on both types
constexpr int which(const LMesh&) { return 0; }
constexpr int which(const SMesh&) { return 1; }

// a specialisation on each
template <class T>
struct Scheme;
template <>
struct Scheme<LMesh> {
  static constexpr int value = 0;
};
template <>
struct Scheme<SMesh> {
  static constexpr int value = 1;
};

int check() {
  LMesh a;
  SMesh b;
  a.nfaces = 3;
  b.nnz = 4;
  a.type = LMesh::MeshType::QUAD;
  b.type = SMesh::MeshType::POLY;
  a.reset();
  static_assert(Scheme<LMesh>::value == 0 && Scheme<SMesh>::value == 1, "specialisations");
  return which(a) + which(b) + (int)a.nfaces + (int)b.nnz;
}
