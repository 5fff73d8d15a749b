// Drop-in replacement of the reference's LibInclude/CatmullClark.h (the
// specialised SLAK scheme, LibInclude/CatmullClark.h:33-165) over the C-ABI.
// Output in SLAK form: refined quads in ids only (ptr = vals = nullptr,
// is_reduced = true). See LAKCatmullClark.h for the shared conventions.
#pragma once

#include "LAKCatmullClark.h"

struct CatmullClark {
  template <typename OFFSET_TYPE, typename INDEX_TYPE, typename VALUE_TYPE, typename VERTEX_TYPE>
  using MeshInfo = LAKCatmullClark::MeshInfo<OFFSET_TYPE, INDEX_TYPE, VALUE_TYPE, VERTEX_TYPE>;

  template <typename MESH_INFO>
  static void subdivideIteratively(MESH_INFO const& cmesh, MESH_INFO& rmesh, int target_level) {
    slak_dropin::subdivide(SLAK_SCHEME_SLAK, "==========SLAK==========", cmesh, rmesh, target_level);
  }
};
