// Command-line driver with the reference demo's interface (source/main.cpp:33-119):
//   slak_cli <mesh.obj> <scheme 0=LAK,1=SLAK> <level> <save 0|1>
//   slak_cli <config file>        (lines "<mesh.obj> <scheme> <level> <save>", '#' comments)
// Reads the OBJ like include/obj.h:107-254 (1-based and negative indices,
// v/vt/vn face forms, TRI/QUAD/POLY detection, uniform face order < 3
// rejected), builds the device mesh like GPU::CSCMesh (include/Mesh.h:153-178,
// 289-300: vals = 1-based corner position, w = 1, nverts = max index + 1),
// refines through the drop-in headers and writes
// <prefix>_<level>_<LAK|SLAK>.obj like GPU::CSCMesh::toObj / OBJ::write
// (include/Mesh.h:379-412, include/obj.h:260-297: default stream precision).
// With --parse-only it only reads the mesh and prints its sizes (no GPU).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "CatmullClark.h"
#include "LAKCatmullClark.h"

namespace {

struct Obj {
  int type = SLAK_POLY;
  int max_face_order = 0;
  std::vector<int> f_offs{0};
  std::vector<int> f_vi;
  std::vector<float> v;  // 3 per vertex
};

Obj read_obj(const std::string& path) {
  std::ifstream in(path);
  if (!in.is_open()) throw std::runtime_error("Could not open obj file\n");
  Obj o;
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty()) continue;
    std::stringstream ss(line);
    std::string id;
    ss >> id;
    if (id == "v") {
      float x = 0, y = 0, z = 0;
      ss >> x >> y >> z;
      o.v.insert(o.v.end(), {x, y, z});
    } else if (id == "f") {
      int n = 0, vi = -1;
      const int nv = (int)(o.v.size() / 3);
      while (ss >> vi && !ss.fail()) {
        o.f_vi.push_back(vi < 0 ? vi + nv : vi - 1);
        if (ss.peek() == '/') {  // skip texture / normal indices
          ss.get();
          int t = 0;
          if (!(ss >> t)) ss.clear();
          if (ss.peek() == '/') {
            ss.get();
            int nn = 0;
            if (!(ss >> nn)) ss.clear();
          }
        }
        vi = -1;
        ++n;
      }
      o.f_offs.push_back(o.f_offs.back() + n);
    }
  }
  int lo = 1 << 30;
  for (size_t f = 1; f < o.f_offs.size(); ++f) {
    int order = o.f_offs[f] - o.f_offs[f - 1];
    lo = std::min(lo, order);
    o.max_face_order = std::max(o.max_face_order, order);
  }
  if (o.f_offs.size() <= 1 || lo != o.max_face_order || o.max_face_order > 4)
    o.type = SLAK_POLY;
  else if (o.max_face_order == 4)
    o.type = SLAK_QUAD;
  else if (o.max_face_order == 3)
    o.type = SLAK_TRI;
  else
    throw std::runtime_error("Error loading obj: Face with less than three vertices detected.\n");
  return o;
}

template <class T>
T* upload(const std::vector<T>& h) {
  void* p = nullptr;
  slak_dropin::succeed(slak_malloc(&p, std::max<size_t>(1, h.size()) * sizeof(T)));
  if (!h.empty()) slak_dropin::succeed(slak_memcpy_h2d(p, h.data(), h.size() * sizeof(T)));
  return static_cast<T*>(p);
}

template <class T>
std::vector<T> download(const T* d, size_t n) {
  std::vector<T> h(n);
  if (n) slak_dropin::succeed(slak_memcpy_d2h(h.data(), d, n * sizeof(T)));
  return h;
}

using LMesh = LAKCatmullClark::MeshInfo<int, int, int, float>;

LMesh to_device(const Obj& o) {
  LMesh m;
  m.type = static_cast<LMesh::MeshType>(o.type);
  m.nfaces = o.f_offs.size() - 1;
  m.nnz = o.f_vi.size();
  m.max_face_order = (size_t)o.max_face_order;
  m.is_reduced = true;
  // GPU::CSCMesh::setFromCPU sizes the vertex array by the largest referenced index
  int vmax = -1;
  for (int id : o.f_vi) vmax = std::max(vmax, id);
  m.nverts = (size_t)(vmax + 1);
  std::vector<int> vals(o.f_vi.size());
  for (size_t f = 0; f + 1 < o.f_offs.size(); ++f)
    for (int k = o.f_offs[f]; k < o.f_offs[f + 1]; ++k) vals[k] = k - o.f_offs[f] + 1;
  std::vector<float> v4(4 * m.nverts, 1.0f);
  for (size_t i = 0; i < m.nverts && 3 * i + 2 < o.v.size(); ++i)
    for (int c = 0; c < 3; ++c) v4[4 * i + c] = o.v[3 * i + c];
  m.ptr = upload(o.f_offs);
  m.ids = upload(o.f_vi);
  m.vals = upload(vals);
  m.verts = upload(v4);
  return m;
}

void write_obj(const LMesh& r, const std::string& fname) {
  std::ofstream out(fname);
  if (!out.is_open()) {
    std::cerr << "Error opening mesh file for writing!" << std::endl;
    return;
  }
  auto verts = download(r.verts, 4 * r.nverts);
  auto ids = download(r.ids, r.nnz);
  std::cout << "Writing file " << fname << std::endl;
  std::cout << "Writing " << r.nverts << " vertices" << std::endl;
  for (size_t i = 0; i < r.nverts; ++i)
    out << "v " << verts[4 * i] << " " << verts[4 * i + 1] << " " << verts[4 * i + 2] << "\n";
  // refined meshes are quads; LAK form is sorted by vals first (already in order)
  std::cout << "Writing " << r.nnz / 4 << " faces" << std::endl;
  for (size_t f = 0; f < r.nnz; f += 4)
    out << "f " << ids[f] + 1 << " " << ids[f + 1] + 1 << " " << ids[f + 2] + 1 << " " << ids[f + 3] + 1 << "\n";
  std::cout << "Done writing" << std::endl;
}

}  // namespace

int main(int argc, char* argv[]) {
  std::vector<std::string> paths;
  std::vector<int> schemes, levels;
  std::vector<bool> saves;
  bool parse_only = argc >= 2 && std::string(argv[1]) == "--parse-only";
  if (parse_only) {
    for (int i = 2; i < argc; ++i) {
      Obj o = read_obj(argv[i]);
      int vmax = -1;
      for (int id : o.f_vi) vmax = std::max(vmax, id);
      std::printf("%s type=%d nf=%zu nnz=%zu nv=%d max_face_order=%d\n", argv[i], o.type, o.f_offs.size() - 1,
                  o.f_vi.size(), vmax + 1, o.max_face_order);
    }
    return 0;
  }
  if (argc == 2) {
    std::ifstream file(argv[1]);
    if (!file.is_open()) std::cerr << "Error: could not open file\n" << std::endl;
    std::string line;
    while (std::getline(file, line)) {
      if (line.empty() || line[0] == '#') continue;
      std::istringstream ss(line);
      std::string path;
      int scheme = 0, level = 0, save = 0;
      if (!(ss >> path >> scheme >> level >> save)) continue;
      paths.push_back(path);
      schemes.push_back(scheme);
      levels.push_back(level);
      saves.push_back(save != 0);
    }
  } else if (argc == 5) {
    paths.emplace_back(argv[1]);
    schemes.push_back(std::atoi(argv[2]));
    levels.push_back(std::atoi(argv[3]));
    saves.push_back(std::atoi(argv[4]) != 0);
  } else {
    std::cerr << "Usage: slak_cli <path_to_obj> <scheme> <target_level> <save_result>" << std::endl;
    std::cerr << "Usage: slak_cli config_file" << std::endl;
    return -1;
  }
  if (!std::getenv("SLAK_BANNER")) setenv("SLAK_BANNER", "1", 0);  // the reference always prints it
  int status = 0;
  for (size_t i = 0; i < paths.size(); ++i) {
    try {
      std::cout << "Mesh: " << paths[i] << std::endl;
      Obj o = read_obj(paths[i]);
      LMesh cmesh = to_device(o), rmesh;
      if (schemes[i] == 0)
        LAKCatmullClark::subdivideIteratively(cmesh, rmesh, levels[i]);
      else
        CatmullClark::subdivideIteratively(cmesh, rmesh, levels[i]);
      if (saves[i] && rmesh.ids) {
        std::string prefix = paths[i].substr(0, paths[i].find_last_of('.'));
        write_obj(rmesh, prefix + "_" + std::to_string(levels[i]) + "_" + (schemes[i] == 0 ? "LAK" : "SLAK") + ".obj");
      }
      rmesh.freeAndReset();
      cmesh.freeAndReset();
    } catch (const char* msg) {  // the reference swallows errors; report them
      std::cerr << "error: " << msg << std::endl;
      status = 1;
    } catch (const std::exception& e) {
      std::cerr << "error: " << e.what() << std::endl;
      status = 1;
    }
  }
  return status;
}
