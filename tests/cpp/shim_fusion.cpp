// Drives sensemap_amd::mvs::StereoFusion (the C++ mirror of the reference's
// depth-map fusion) on inputs written by tests/test_gpu_cpp_mirror.py:
//   <dir>/meta.txt : n_images; per image W H fx cx fy cy R[9] T[3] is_rig used
//                    num_overlapping overlapping...; then step_size
//                    min_num_pixels max_num_pixels min_num_visible_images
//                    max_traversal_depth max_reproj_error max_depth_error
//                    with_normal max_normal_error
//   <dir>/depth<k>.bin, <dir>/normal<k>.bin : DepthMap / NormalMap files
//   <dir>/grey<k>.raw : the W*H grey bitmap
// and writes <dir>/points.bin: the point count (int64), then per point
// x y z nx ny nz (float32), r g b (uint8), the scores (float32), the
// visibility offsets (int64, count + 1) and the visible images (uint32).
#include <cstdint>
#include <fstream>
#include <iostream>
#include <iterator>
#include <vector>

#include "../../sensemap_amd/host/patch_match_cuda.h"

using namespace sensemap_amd::mvs;

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string dir = argv[1];
  std::ifstream meta(dir + "/meta.txt");
  int n = 0;
  meta >> n;
  std::vector<Image> images;
  std::vector<DepthMap> depths(n);
  std::vector<NormalMap> normals(n);
  std::vector<std::vector<int>> overlapping(n);
  std::vector<bool> used(n);
  for (int k = 0; k < n; ++k) {
    int W, H, rig, use, nov;
    float K4[4], R[9], T[3];
    meta >> W >> H;
    for (float& x : K4) meta >> x;
    for (float& x : R) meta >> x;
    for (float& x : T) meta >> x;
    meta >> rig >> use >> nov;
    overlapping[k].resize(nov);
    for (int& j : overlapping[k]) meta >> j;
    used[k] = use != 0;
    std::ifstream g(dir + "/grey" + std::to_string(k) + ".raw", std::ios::binary);
    std::vector<uint8_t> grey((std::istreambuf_iterator<char>(g)), std::istreambuf_iterator<char>());
    if (grey.size() != (size_t)W * H) return 3;
    const float K[9] = {K4[0], 0.f, K4[1], 0.f, K4[2], K4[3], 0.f, 0.f, 1.f};
    images.emplace_back(W, H, K, R, T, std::move(grey), rig != 0);
    depths[k].Read(dir + "/depth" + std::to_string(k) + ".bin");
    normals[k].Read(dir + "/normal" + std::to_string(k) + ".bin");
  }
  StereoFusion::Options o;
  int with_normal = 1;
  meta >> o.step_size >> o.min_num_pixels >> o.max_num_pixels >> o.min_num_visible_images >> o.max_traversal_depth >>
      o.max_reproj_error >> o.max_depth_error >> with_normal >> o.max_normal_error;
  o.with_normal = with_normal != 0;
  try {
    StereoFusion fusion(o, &images, &depths, &normals, &overlapping, used);
    fusion.Run();
    const std::vector<PlyPoint>& pts = fusion.GetFusedPoints();
    const std::vector<std::vector<uint32_t>>& vis = fusion.GetFusedPointsVisibility();
    const std::vector<float>& score = fusion.GetFusedPointsScore();
    std::ofstream out(dir + "/points.bin", std::ios::binary);
    const int64_t np = (int64_t)pts.size();
    out.write((const char*)&np, 8);
    for (const PlyPoint& p : pts) {
      const float v[6] = {p.x, p.y, p.z, p.nx, p.ny, p.nz};
      out.write((const char*)v, sizeof(v));
    }
    for (const PlyPoint& p : pts) {
      const uint8_t c[3] = {p.r, p.g, p.b};
      out.write((const char*)c, 3);
    }
    out.write((const char*)score.data(), (std::streamsize)(score.size() * 4));
    int64_t off = 0;
    out.write((const char*)&off, 8);
    for (const auto& v : vis) {
      off += (int64_t)v.size();
      out.write((const char*)&off, 8);
    }
    for (const auto& v : vis) out.write((const char*)v.data(), (std::streamsize)(v.size() * 4));
  } catch (const std::exception& e) {
    std::cerr << e.what() << "\n";
    return 4;
  }
  return 0;
}
