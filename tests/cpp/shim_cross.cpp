// Drives sensemap_amd::mvs::CrossFilterCuda (the C++ mirror of the
// reference's cross-view filter) on inputs written by tests/test_gpu_cpp_mirror.py:
//   <dir>/meta.txt : n_images; per image W H fx cx fy cy R[9] T[3];
//                    n_views; per view ref num_src src...;
//                    filter_geom_consistency_max_cost max_normal_error depth_diff_threshold min_consistent
//   <dir>/depth<k>.bin, <dir>/normal<k>.bin : the image's maps as DepthMap / NormalMap files
// and writes <dir>/refined<ref>.bin (DepthMap::Write) for every view; then
// runs DedupRun (the views' references flagged) and writes, per image k,
// <dir>/dedup<k>.bin, <dir>/vis<k>.bin (Mat<uint32_t>) and <dir>/delt<k>.bin.
#include <fstream>
#include <iostream>
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
  for (int k = 0; k < n; ++k) {
    int W, H;
    float K4[4], R[9], T[3];
    meta >> W >> H;
    for (float& x : K4) meta >> x;
    for (float& x : R) meta >> x;
    for (float& x : T) meta >> x;
    const float K[9] = {K4[0], 0.f, K4[1], 0.f, K4[2], K4[3], 0.f, 0.f, 1.f};
    images.emplace_back(W, H, K, R, T, std::vector<uint8_t>((size_t)W * H, 0));
    depths[k].Read(dir + "/depth" + std::to_string(k) + ".bin");
    normals[k].Read(dir + "/normal" + std::to_string(k) + ".bin");
  }
  int nv = 0;
  meta >> nv;
  std::vector<Problem> problems(nv);
  for (int v = 0; v < nv; ++v) {
    int m = 0;
    meta >> problems[v].ref_image_idx >> m;
    problems[v].src_image_extend_idxs.resize(m);
    for (int& s : problems[v].src_image_extend_idxs) meta >> s;
    problems[v].images = &images;
    problems[v].depth_maps = &depths;
    problems[v].normal_maps = &normals;
  }
  PatchMatchOptions options;
  meta >> options.filter_geom_consistency_max_cost >> options.max_normal_error >> options.depth_diff_threshold >>
      options.filter_min_num_consistent;
  options.gpu_index = "0";
  std::vector<int> ids(n);
  for (int k = 0; k < n; ++k) ids[k] = k;
  try {
    std::vector<bool> ref_flags(n, false);
    for (const Problem& p : problems) ref_flags[p.ref_image_idx] = true;
    CrossFilterCuda filter(options, problems, ids, ref_flags);
    filter.Run();
    for (const Problem& p : problems) {
      DepthMapInfo info;
      info.width = (int)images[p.ref_image_idx].GetWidth();
      info.height = (int)images[p.ref_image_idx].GetHeight();
      filter.GetRefineDepthMap(p.ref_image_idx, info).Write(dir + "/refined" + std::to_string(p.ref_image_idx) + ".bin");
    }
    filter.DedupRun();
    for (int k = 0; k < n; ++k) {
      DepthMapInfo info;
      info.width = (int)images[k].GetWidth();
      info.height = (int)images[k].GetHeight();
      filter.GetDeduplicDepthMap(k, info).Write(dir + "/dedup" + std::to_string(k) + ".bin");
      filter.GetVisibilyMap(k).Write(dir + "/vis" + std::to_string(k) + ".bin");
      filter.GetDeltDepthMap(k).Write(dir + "/delt" + std::to_string(k) + ".bin");
    }
  } catch (const std::exception& e) {
    std::cerr << e.what() << "\n";
    return 4;
  }
  return 0;
}
