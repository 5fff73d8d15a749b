// Uses the drop-in headers exactly like the reference's demo does
// (source/main.cpp:84-112): build a mesh on the device, refine it with both
// schemes, read the result back, free with freeAndReset. Prints one line per
// check; exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <vector>

#include "CatmullClark.h"
#include "LAKCatmullClark.h"

using LMesh = LAKCatmullClark::MeshInfo<int, int, int, float>;
using SMesh = CatmullClark::MeshInfo<int, int, int, float>;

#define CHECK(cond)                                                     \
  do {                                                                  \
    if (!(cond)) {                                                      \
      std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond);        \
      return 1;                                                         \
    }                                                                   \
  } while (0)

template <class T>
T* upload(const std::vector<T>& h) {
  void* p = nullptr;
  slak_dropin::succeed(slak_malloc(&p, h.size() * sizeof(T)));
  slak_dropin::succeed(slak_memcpy_h2d(p, h.data(), h.size() * sizeof(T)));
  return static_cast<T*>(p);
}

template <class T>
std::vector<T> download(const T* d, size_t n) {
  std::vector<T> h(n);
  slak_dropin::succeed(slak_memcpy_d2h(h.data(), d, n * sizeof(T)));
  return h;
}

template <class MI>
MI make_torus(int a, int b) {
  std::vector<int> ptr, ids, vals;
  std::vector<float> verts;
  for (int j = 0; j < b; ++j)
    for (int i = 0; i < a; ++i) {
      int i1 = (i + 1) % a, j1 = (j + 1) % b;
      int q[4] = {j * a + i, j * a + i1, j1 * a + i1, j1 * a + i};
      for (int k = 0; k < 4; ++k) {
        ids.push_back(q[k]);
        vals.push_back(k + 1);
      }
    }
  for (int f = 0; f <= a * b; ++f) ptr.push_back(4 * f);
  for (int j = 0; j < b; ++j)
    for (int i = 0; i < a; ++i) {
      double u = 2 * M_PI * i / a, v = 2 * M_PI * j / b;
      verts.push_back((float)((1 + 0.3 * std::cos(v)) * std::cos(u)));
      verts.push_back((float)((1 + 0.3 * std::cos(v)) * std::sin(u)));
      verts.push_back((float)(0.3 * std::sin(v)));
      verts.push_back(1.0f);
    }
  MI m;
  m.type = MI::MeshType::QUAD;
  m.nfaces = a * b;
  m.nverts = a * b;
  m.nnz = 4 * a * b;
  m.max_face_order = 4;
  m.is_reduced = false;
  m.ptr = upload(ptr);
  m.ids = upload(ids);
  m.vals = upload(vals);
  m.verts = upload(verts);
  return m;
}

int main() {
  const int a = 32, b = 16, levels = 3;
  LMesh cmesh = make_torus<LMesh>(a, b);
  LMesh rmesh;
  LAKCatmullClark::subdivideIteratively(cmesh, rmesh, levels);
  const size_t nf = (size_t)a * b * 64;
  CHECK(rmesh.nfaces == nf && rmesh.nverts == nf && rmesh.nnz == 4 * nf);
  CHECK(rmesh.type == LMesh::MeshType::QUAD && !rmesh.is_reduced && rmesh.max_face_order == 4);
  CHECK(rmesh.ptr && rmesh.vals && rmesh.ids && rmesh.verts);
  auto ptr = download(rmesh.ptr, rmesh.nfaces + 1);
  auto vals = download(rmesh.vals, rmesh.nnz);
  auto verts = download(rmesh.verts, 4 * rmesh.nverts);
  for (size_t i = 0; i <= rmesh.nfaces; ++i) CHECK(ptr[i] == (int)(4 * i));
  for (size_t i = 0; i < rmesh.nnz; ++i) CHECK(vals[i] == (int)(i % 4 + 1));
  for (size_t v = 0; v < rmesh.nverts; ++v) CHECK(verts[4 * v + 3] == 1.0f);
  std::printf("ok LAK: nf=%zu nv=%zu\n", rmesh.nfaces, rmesh.nverts);

  SMesh scmesh = make_torus<SMesh>(a, b);  // CatmullClark has a MeshInfo type of its own
  SMesh smesh;
  CatmullClark::subdivideIteratively(scmesh, smesh, levels);
  CHECK(smesh.nfaces == nf && smesh.is_reduced && smesh.ptr == nullptr && smesh.vals == nullptr);
  auto sverts = download(smesh.verts, 4 * smesh.nverts);
  // same refined positions up to the edge-point permutation: the control
  // vertices' limit points and the last level's face points (indexed by
  // coarse corner in both schemes) are bitwise identical
  const size_t nin = (size_t)a * b * 16;  // vertices = faces of the last level's input
  for (size_t i = 0; i < 4 * (size_t)a * b; ++i) CHECK(sverts[i] == verts[i]);
  for (size_t i = 4 * nin; i < 8 * nin; ++i) CHECK(sverts[i] == verts[i]);
  std::printf("ok SLAK: nf=%zu nv=%zu\n", smesh.nfaces, smesh.nverts);

  // errors surface as const char* throws
  LMesh bad = cmesh;
  bad.nnz = 3;  // inconsistent with ptr
  bool thrown = false;
  try {
    LMesh out;
    LAKCatmullClark::subdivideIteratively(bad, out, 1);
  } catch (const char* msg) {
    thrown = true;
    std::printf("ok error: %s\n", msg);
  }
  CHECK(thrown);

  rmesh.freeAndReset();
  smesh.freeAndReset();
  scmesh.freeAndReset();
  cmesh.freeAndReset();
  CHECK(rmesh.ids == nullptr && cmesh.verts == nullptr);
  std::printf("ok dropin\n");
  return 0;
}
