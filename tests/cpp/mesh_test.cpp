// GPU::CSCMesh (include/Mesh.h of this repository) used the way the
// reference's demo and tools use it: construction from an OBJ, copy
// construction and assignment, CPU access, unreduce / reduce round trip,
// conversion to MeshInfo, refinement and toObj.
//   mesh_test <in.obj> <out.obj>   prints "mesh_test ok" on success
#include <cstdio>
#include <string>
#include <vector>

#include "LAKCatmullClark.h"
#include "Mesh.h"

using Mesh = GPU::CSCMesh<int, int, int, float>;

static int fail(const char* what) {
  std::printf("mesh_test FAILED: %s\n", what);
  return 1;
}

int main(int argc, char** argv) {
  if (argc != 3) return fail("usage: mesh_test <in.obj> <out.obj>");
  try {
    const Mesh a(argv[1]);
    std::vector<int> p0, i0, v0;
    std::vector<math::float4> x0;
    a.accessFromCpu(p0, i0, v0, x0);
    if (!a.is_reduced || a.nfaces + 1 != p0.size() || a.nnz != i0.size() || a.nverts != x0.size())
      return fail("sizes after reading");
    for (size_t f = 0; f < a.nfaces; ++f)
      for (int c = p0[f]; c < p0[f + 1]; ++c)
        if (v0[c] != c - p0[f] + 1) return fail("vals are not the corner positions");
    Mesh b(a);  // copy construction
    Mesh c;
    c = a;  // assignment
    for (const Mesh* m : {&b, &c}) {
      std::vector<int> p, i, v;
      std::vector<math::float4> x;
      m->accessFromCpu(p, i, v, x);
      if (p != p0 || i != i0 || v != v0 || x.size() != x0.size()) return fail("copy differs");
      for (size_t k = 0; k < x.size(); ++k)
        if (x[k].x != x0[k].x || x[k].y != x0[k].y || x[k].z != x0[k].z || x[k].w != 1.0f)
          return fail("copied positions differ");
    }
    if (a.type != Mesh::MeshType::POLY) {
      b.unreduce();  // each face by ids, vals keep the positions
      if (b.is_reduced) return fail("unreduce left the mesh reduced");
      b.reduce();
      std::vector<int> p, i, v;
      std::vector<math::float4> x;
      b.accessFromCpu(p, i, v, x);
      if (!b.is_reduced || i != i0 || v != v0) return fail("unreduce + reduce is not the identity");
    }
    // refinement through the MeshInfo conversion, written by toObj
    using MeshInfo = LAKCatmullClark::MeshInfo<int, int, int, float>;
    MeshInfo ci, ri;
    GPU::CSCMeshCCMeshInfoConversion(a, ci);
    LAKCatmullClark::subdivideIteratively(ci, ri, 2);
    Mesh r;
    GPU::CSCMeshCCMeshInfoConversion(ri, r);
    if (r.is_reduced || r.type != Mesh::MeshType::QUAD) return fail("LAK result form");
    r.toObj(argv[2]);
    if (r.is_reduced) return fail("toObj did not restore the unreduced form");
  } catch (const char* m) {
    return fail(m);
  } catch (const std::exception& e) {
    return fail(e.what());
  }
  std::printf("mesh_test ok\n");
  return 0;
}
