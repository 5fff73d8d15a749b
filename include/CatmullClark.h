// Drop-in replacement of the reference's LibInclude/CatmullClark.h (the
// specialised SLAK scheme, LibInclude/CatmullClark.h:33-165) over the C-ABI.
// Output in SLAK form: refined quads in ids only (ptr = vals = nullptr,
// is_reduced = true). See LAKCatmullClark.h for the shared conventions.
#pragma once

#include "LAKCatmullClark.h"

struct CatmullClark {
  // A type of its own with LAKCatmullClark::MeshInfo's layout, as in the
  // reference (LibInclude/CatmullClark.h:69-130).
  template <typename OFFSET_TYPE, typename INDEX_TYPE, typename VALUE_TYPE, typename VERTEX_TYPE>
  struct MeshInfo : slak_dropin::MeshInfoFields<MeshInfo<OFFSET_TYPE, INDEX_TYPE, VALUE_TYPE, VERTEX_TYPE>,
                                                OFFSET_TYPE, INDEX_TYPE, VALUE_TYPE, VERTEX_TYPE> {};

  template <typename MESH_INFO>
  static void subdivideIteratively(MESH_INFO const& cmesh, MESH_INFO& rmesh, int target_level) {
    slak_dropin::subdivide(SLAK_SCHEME_SLAK, "==========SLAK==========", cmesh, rmesh, target_level);
  }
};
