// Drives sensemap_amd::mvs::PatchMatchACMMCuda (the C++ mirror of the
// reference interface) on a view written by tests/test_gpu_cpp_mirror.py:
//   <dir>/meta.txt   : W H N depth_min depth_max radius step sigma_s sigma_c iters thk seed
//                      then per image (ref first): fx cx fy cy R[9] T[3]
//   <dir>/img<k>.u8  : W*H bytes per image
//   <dir>/mask.u8    : optional W*H mask of the reference image (Problem::mask_maps)
//   <dir>/geom.txt   : optional; present: a geometric pass with the curvature map and the
//                      median filter, reading <dir>/depth<k>.f32 (W*H per image) and
//                      <dir>/normal0.f32 (3*W*H, the reference's initial normals)
// and writes <dir>/depth.f32 and <dir>/normal.f32 (and <dir>/curvature.u16 in the
// geometric pass).
#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

#include "../../sensemap_amd/host/patch_match_cuda.h"

using namespace sensemap_amd::mvs;

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string dir = argv[1];
  std::ifstream meta(dir + "/meta.txt");
  int W, H, N, radius, step, iters, thk;
  double dmin, dmax, ss, sc;
  unsigned long long seed;
  meta >> W >> H >> N >> dmin >> dmax >> radius >> step >> ss >> sc >> iters >> thk >> seed;
  std::vector<Image> images;
  for (int k = 0; k <= N; ++k) {
    float K4[4], R[9], T[3];
    for (float& x : K4) meta >> x;
    for (float& x : R) meta >> x;
    for (float& x : T) meta >> x;
    const float K[9] = {K4[0], 0.f, K4[1], 0.f, K4[2], K4[3], 0.f, 0.f, 1.f};
    std::vector<uint8_t> px((size_t)W * H);
    std::ifstream f(dir + "/img" + std::to_string(k) + ".u8", std::ios::binary);
    f.read((char*)px.data(), (std::streamsize)px.size());
    images.emplace_back(W, H, K, R, T, std::move(px));
  }
  PatchMatchOptions options;
  options.window_radius = radius;
  options.window_step = step;
  options.sigma_spatial = ss;
  options.sigma_color = sc;
  options.num_iterations = iters;
  options.thk = thk;
  options.seed = seed;
  const bool geom = (bool)std::ifstream(dir + "/geom.txt");
  options.geom_consistency = geom;
  options.init_depth_random = !geom;
  options.est_curvature = geom;
  options.median_filter = geom;
  options.depth_min = dmin;
  options.depth_max = dmax;
  options.gpu_index = "0";
  if (!options.Check()) {
    std::cerr << "invalid options\n";
    return 3;
  }
  Problem problem;
  problem.ref_image_idx = 0;
  for (int k = 1; k <= N; ++k) problem.src_image_idxs.push_back(k);
  problem.images = &images;
  std::vector<DepthMap> depth_maps;
  std::vector<NormalMap> normal_maps;
  if (geom) {
    for (int k = 0; k <= N; ++k) {
      DepthMap d(W, H, (float)dmin, (float)dmax);
      std::ifstream(dir + "/depth" + std::to_string(k) + ".f32", std::ios::binary)
          .read((char*)d.GetPtr(), (std::streamsize)(sizeof(float) * W * H));
      depth_maps.push_back(std::move(d));
      NormalMap n(W, H);
      if (k == 0)
        std::ifstream(dir + "/normal0.f32", std::ios::binary)
            .read((char*)n.GetPtr(), (std::streamsize)(sizeof(float) * 3 * W * H));
      normal_maps.push_back(std::move(n));
    }
    problem.depth_maps = &depth_maps;
    problem.normal_maps = &normal_maps;
  }
  std::vector<Bitmap> masks(images.size());
  std::ifstream mf(dir + "/mask.u8", std::ios::binary);
  if (mf) {
    std::vector<uint8_t> m((size_t)W * H);
    mf.read((char*)m.data(), (std::streamsize)m.size());
    masks[0] = Bitmap(W, H, std::move(m));
    problem.mask_maps = &masks;
  }
  try {
    PatchMatchACMMCuda engine(options, problem);
    engine.Run();
    const DepthMap depth = engine.GetDepthMap();
    const NormalMap normal = engine.GetNormalMap();
    std::ofstream(dir + "/depth.f32", std::ios::binary)
        .write((const char*)depth.GetPtr(), (std::streamsize)(sizeof(float) * W * H));
    std::ofstream(dir + "/normal.f32", std::ios::binary)
        .write((const char*)normal.GetPtr(), (std::streamsize)(sizeof(float) * 3 * W * H));
    if (geom) {
      const Mat<unsigned short> curv = engine.GetCurvatureMap();
      std::ofstream(dir + "/curvature.u16", std::ios::binary)
          .write((const char*)curv.GetPtr(), (std::streamsize)(sizeof(unsigned short) * 2 * W * H));
    }
  } catch (const std::exception& e) {
    std::cerr << e.what() << "\n";
    return 4;
  }
  return 0;
}
