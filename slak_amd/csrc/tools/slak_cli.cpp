// Command-line driver with the reference demo's interface (source/main.cpp:33-119):
//   slak_cli <mesh.obj> <scheme 0=LAK,1=SLAK> <level> <save 0|1>
//   slak_cli <config file>        (lines "<mesh.obj> <scheme> <level> <save>", '#' comments)
// the same steps as the demo: GPU::CSCMesh from the OBJ (include/obj.h,
// include/Mesh.h of this repository), the conversion to MeshInfo, the drop-in
// subdivideIteratively and CSCMesh::toObj to <prefix>_<level>_<LAK|SLAK>.obj.
// Differences: an error is printed and the exit status is 1 (the reference
// swallows it and stops), blank or malformed config lines are skipped.
// With --parse-only it only runs OBJ::read and prints what it read (no GPU).
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
#include "Mesh.h"

int main(int argc, char* argv[]) {
  std::vector<std::string> paths;
  std::vector<int> schemes, levels;
  std::vector<bool> saves;
  bool parse_only = argc >= 2 && std::string(argv[1]) == "--parse-only";
  if (parse_only) {  // OBJ::read only (no GPU): sizes, type and the other elements read
    for (int i = 2; i < argc; ++i) {
      OBJ::Data o;
      OBJ::read(o, argv[i]);
      int vmax = -1;
      for (int id : o.f_vi) vmax = std::max(vmax, id);
      std::printf("%s type=%d nf=%zu nnz=%zu nv=%d max_face_order=%d v=%zu vt=%zu vn=%zu f_vti=%zu f_vni=%zu tags=%zu\n",
                  argv[i], (int)o.type, o.f_offs.size() - 1, o.f_vi.size(), vmax + 1, o.max_face_order, o.v.size(),
                  o.vt.size(), o.vn.size(), o.f_vti.size(), o.f_vni.size(), o.t.size());
      for (const auto& n : o.vn) std::printf("vn %.9g %.9g %.9g\n", n.x, n.y, n.z);
      for (int x : o.f_vti) std::printf("vti %d\n", x);
      for (int x : o.f_vni) std::printf("vni %d\n", x);
      for (const auto& t : o.t) {
        std::printf("tag %s", t.type.c_str());
        for (int x : t.intargs) std::printf(" i%d", x);
        for (float x : t.floatargs) std::printf(" f%.9g", x);
        std::printf("\n");
      }
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
      const GPU::CSCMesh<int, int, int, float> cmesh(paths[i]);
      GPU::CSCMesh<int, int, int, float> rmesh;
      if (schemes[i] == 0) {
        using MeshInfo = LAKCatmullClark::MeshInfo<int, int, int, float>;
        MeshInfo cmesh_info, rmesh_info;
        GPU::CSCMeshCCMeshInfoConversion(cmesh, cmesh_info);
        LAKCatmullClark::subdivideIteratively(cmesh_info, rmesh_info, levels[i]);
        GPU::CSCMeshCCMeshInfoConversion(rmesh_info, rmesh);  // rmesh owns the result
      } else if (schemes[i] == 1) {
        using MeshInfo = CatmullClark::MeshInfo<int, int, int, float>;
        MeshInfo cmesh_info, rmesh_info;
        GPU::CSCMeshCCMeshInfoConversion(cmesh, cmesh_info);
        CatmullClark::subdivideIteratively(cmesh_info, rmesh_info, levels[i]);
        GPU::CSCMeshCCMeshInfoConversion(rmesh_info, rmesh);
      } else {
        throw std::runtime_error("scheme must be 0 (LAK) or 1 (SLAK)");
      }
      if (saves[i]) {
        const std::string prefix = paths[i].substr(0, paths[i].find_last_of('.'));
        rmesh.toObj(prefix + "_" + std::to_string(levels[i]) + "_" + (schemes[i] == 0 ? "LAK" : "SLAK") + ".obj");
      }
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
