// OBJ reader and writer with the behaviour of the reference's include/obj.h
// (OBJ::Data, OBJ::read, OBJ::write; include/obj.h:70-297), host-only C++.
//
// Reader (include/obj.h:107-253):
//   - "v x y z" positions; "vn" normals with z negated (obj.h:148); "vt";
//   - "f a b c ..." with a, a/t, a//n, a/t/n tokens; indices are 1-based and
//     negative ones count back from the elements read so far (obj.h:167);
//   - "t type nint/nfloat/nstring ints floats" tags; "#" and "g" lines and
//     unknown entries are skipped;
//   - face type: TRI or QUAD when every face has 3 or 4 corners, POLY when the
//     orders differ or exceed 4, an error for faces below 3 corners
//     (obj.h:237-253).
// Writer (include/obj.h:260-297): "v x y z" in the default stream format (six
// significant digits, lossy), "f" lines with 1-based indices, faces from
// f_offs or, without it, at stride 3 / 4 for TRI / QUAD, and the same progress
// lines on stdout.
#pragma once

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "math/vector.h"

namespace OBJ {

struct Data {
  enum class FaceType { TRI, QUAD, POLY };
  FaceType type{FaceType::POLY};
  int max_face_order{-1};

  std::vector<int> f_offs;

  std::vector<int> f_vi;
  std::vector<math::float3> v;

  std::vector<int> f_vni;
  std::vector<math::float3> vn;

  std::vector<int> f_vti;
  std::vector<math::float2> vt;

  struct Tag {
    std::string type;
    std::vector<int> intargs;
    std::vector<float> floatargs;
  };
  std::vector<Tag> t;
};

namespace detail {

// 1-based index, or negative: relative to the count read so far
inline int resolve(int i, size_t count) { return i < 0 ? i + static_cast<int>(count) : i - 1; }

// an optional index after '/'; failure leaves the stream usable
inline bool slash_index(std::istringstream& in, int& out) {
  in >> out;
  if (in.fail()) {
    in.clear();
    return false;
  }
  return true;
}

inline void face_line(Data& d, std::istringstream& in) {
  int nverts = 0;
  int vi = -1;
  while (in >> vi) {
    d.f_vi.push_back(resolve(vi, d.v.size()));
    if (in.peek() == '/') {
      in.get();
      int ti = -1;
      if (slash_index(in, ti)) d.f_vti.push_back(resolve(ti, d.vt.size()));
      if (in.peek() == '/') {
        in.get();
        int ni = -1;
        if (slash_index(in, ni)) d.f_vni.push_back(resolve(ni, d.vn.size()));
      }
    }
    ++nverts;
  }
  d.f_offs.push_back(d.f_offs.back() + nverts);
}

inline void tag_line(Data& d, std::istringstream& in) {
  Data::Tag tag;
  in >> tag.type;
  int nint = 0, nfloat = 0, nstring = 0;
  in >> nint;
  in.get();  // '/'
  in >> nfloat;
  in.get();  // '/'
  in >> nstring;
  tag.intargs.assign(std::max(nint, 0), 0);
  for (int& x : tag.intargs) in >> x;
  tag.floatargs.assign(std::max(nfloat, 0), 0.0f);
  for (float& x : tag.floatargs) in >> x;
  d.t.push_back(std::move(tag));
}

inline void classify(Data& d) {
  int lo = std::numeric_limits<int>::max();
  d.max_face_order = 0;
  for (size_t f = 1; f < d.f_offs.size(); ++f) {
    const int order = d.f_offs[f] - d.f_offs[f - 1];
    lo = std::min(lo, order);
    d.max_face_order = std::max(d.max_face_order, order);
  }
  if (lo != d.max_face_order || d.max_face_order > 4)
    d.type = Data::FaceType::POLY;
  else if (d.max_face_order == 4)
    d.type = Data::FaceType::QUAD;
  else if (d.max_face_order == 3)
    d.type = Data::FaceType::TRI;
  else
    throw std::runtime_error("Error loading obj: Face with less than three vertices detected.\n");
}

}  // namespace detail

inline void read(Data& data, const char* fname) {
  std::ifstream file(fname);
  if (!file.is_open()) throw std::runtime_error("Could not open obj file\n");
  data.f_offs.push_back(0);
  std::string line;
  while (std::getline(file, line)) {
    if (line.empty()) continue;
    std::istringstream in(line);
    std::string id;
    in >> id;
    if (id == "#" || id == "g") continue;
    if (id == "v") {
      math::float3 p;
      in >> p.x >> p.y >> p.z;
      data.v.push_back(p);
    } else if (id == "vn") {
      math::float3 n;
      in >> n.x >> n.y >> n.z;
      n.z = -n.z;
      data.vn.push_back(n);
    } else if (id == "vt") {
      math::float2 t;
      in >> t.x >> t.y;
      data.vt.push_back(t);
    } else if (id == "f") {
      detail::face_line(data, in);
    } else if (id == "t") {
      detail::tag_line(data, in);
    }
  }
  detail::classify(data);
}

inline void write(const Data& data, const char* fname) {
  std::FILE* out = std::fopen(fname, "w");
  if (!out) {
    std::cerr << "Error opening mesh file for writing!" << std::endl;
    return;
  }
  std::cout << "Writing file " << fname << std::endl;
  std::cout << "Writing " << data.v.size() << " vertices" << std::endl;
  // "%g" with six significant digits: the default formatting of a float
  // inserted into an ostream
  for (const math::float3& p : data.v)
    std::fprintf(out, "v %g %g %g\n", (double)p.x, (double)p.y, (double)p.z);
  auto face = [out](const int* first, int n) {
    std::fputc('f', out);
    for (int k = 0; k < n; ++k) std::fprintf(out, " %d", first[k] + 1);
    std::fputc('\n', out);
  };
  if (!data.f_offs.empty()) {
    std::cout << "Writing " << data.f_offs.size() - 1 << " faces" << std::endl;
    for (size_t f = 0; f + 1 < data.f_offs.size(); ++f)
      face(data.f_vi.data() + data.f_offs[f], data.f_offs[f + 1] - data.f_offs[f]);
  } else {
    if (data.type == Data::FaceType::POLY) {
      std::cerr << "Error: Face-offsets required to write a poly mesh!" << std::endl;
      std::fclose(out);
      return;
    }
    const int k = data.type == Data::FaceType::QUAD ? 4 : 3;
    std::cout << "Writing " << data.f_vi.size() / k << " faces" << std::endl;
    for (size_t first = 0; first + k <= data.f_vi.size(); first += k) face(data.f_vi.data() + first, k);
  }
  std::cout << "Done writing" << std::endl;
  std::fclose(out);
}

}  // namespace OBJ
