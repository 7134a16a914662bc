// Drives sensemap_amd::mvs::Model (the C++ mirror of the reference's sparse
// model) on inputs written by tests/test_gpu_cpp_mirror.py:
//   <dir>/model.txt : I; per image R[9] T[3] rig; P; per point x y z L track[L];
//                     max_num_images min_triangulation_angle; subset size and indices (-1 = no subset)
// and writes <dir>/lists.txt (one line per image: count then indices) and
// <dir>/ranges.txt (depth_min depth_max per image, as float bit patterns).
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "../../sensemap_amd/host/patch_match_cuda.h"

using namespace sensemap_amd::mvs;

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string dir = argv[1];
  std::ifstream in(dir + "/model.txt");
  Model model;
  size_t I = 0, P = 0;
  in >> I;
  const float K[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (size_t i = 0; i < I; ++i) {
    float R[9], T[3];
    int rig = 0;
    for (float& x : R) in >> x;
    for (float& x : T) in >> x;
    in >> rig;
    model.images.emplace_back(1, 1, K, R, T, std::vector<uint8_t>(1, 0), rig != 0);
  }
  in >> P;
  model.points.resize(P);
  for (auto& p : model.points) {
    size_t L = 0;
    in >> p.x >> p.y >> p.z >> L;
    p.track.resize(L);
    for (int& t : p.track) in >> t;
  }
  size_t max_n = 0;
  double angle = 0;
  long nsub = 0;
  in >> max_n >> angle >> nsub;
  try {
    std::vector<std::vector<int>> lists;
    if (nsub < 0) {
      lists = model.GetMaxOverlappingImages(max_n, angle);
    } else {
      std::unordered_set<int> sub;
      for (long k = 0; k < nsub; ++k) {
        int i;
        in >> i;
        sub.insert(i);
      }
      lists = model.GetMaxOverlappingImages(max_n, angle, sub);
    }
    std::ofstream lo(dir + "/lists.txt");
    for (const auto& l : lists) {
      lo << l.size();
      for (int i : l) lo << ' ' << i;
      lo << '\n';
    }
    std::ofstream ro(dir + "/ranges.txt");
    for (const auto& r : model.ComputeDepthRanges()) {
      uint32_t a, b;
      std::memcpy(&a, &r.first, 4);
      std::memcpy(&b, &r.second, 4);
      ro << a << ' ' << b << '\n';
    }
  } catch (const std::exception& e) {
    std::cerr << "shim_model: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
