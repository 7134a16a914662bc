// pm_cross.hip -- the cross-view depth filter: CrossFilterCuda::Run
// (src/mvs/patch_match_cuda.cu:1493-1551) and CrossFilterKernel
// (src/mvs/kernel_functions.cu:3561-3778), all views in one launch.
//
// Layout.  Every image used by a view has its depth map and normal codes in
// HBM once (the reference copies all of them into two layered textures of
// the largest size).  A table of views (ref maps, output, poses, sources)
// drives one launch of 16x16-pixel workgroups; each workgroup finds its view
// by binary search over the views' first tiles.  The pass reads 8 bytes
// (depth, code) per source per pixel and does about 150 FP32 operations per
// source: it is bound by memory latency, not by the ALUs.
#include "pm_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

using namespace pmh;

namespace {

constexpr int kCrossMaxSrc = 31;  // MAX_NUM_CROSS_SRC (utils.h:39)
constexpr int kCrossTile = 16;    // a workgroup covers 16x16 pixels

struct CrossSrc {
  const float* depth;
  const uint32_t* codes;
  int32_t width, height;
};

struct CrossViewDev {
  int32_t W, H, N, tile_start;
  int32_t min_consistent;  // min(filter_min_num_consistent, N)
  int32_t pad0;
  float K[4], iK[4];
  const float* depth;
  const uint32_t* codes;
  float* out;
  const float* poses;  // N * PM_POSE_STRIDE
  CrossSrc src[kCrossMaxSrc];
};

struct CrossParams {
  float max_reproj;      // filter_geom_consistency_max_cost
  float sqr_depth_thr;   // depth_diff_threshold^2
  float max_cos_normal;  // cosf(DEG2RAD(max_normal_error))
  float max_view_cos;    // cosf(DEG2RAD(100))
  int32_t n_views;
};

using CV = const __attribute__((address_space(4))) CrossViewDev;

// The decode inside CrossFilterKernel (kernel_functions.cu:3593-3603).
__device__ __forceinline__ void decode_code(uint32_t code, float n[3]) {
  const float q1 = (float)(code >> 16) / 32767.0f - 1.0f;
  const float q2 = (float)(code & 0xffffu) / 32767.0f - 1.0f;
  const float s = q1 * q1 + q2 * q2;
  n[0] = (float)(1.0 - 2.0 * (double)s);
  const float r = sqrtf(1.0f - s);
  n[1] = (2.0f * q1) * r;
  n[2] = (2.0f * q2) * r;
}

__device__ float cross_pixel(CV& v, const CrossParams& prm, int row, int col) {
  const __attribute__((address_space(1))) float* dmap = (const __attribute__((address_space(1))) float*)v.depth;
  const int p = row * v.W + col;
  const float depth = dmap[p];
  if ((double)depth < 1e-6 || col < 3 || col > v.W - 3 || row < 3 || row > v.H - 3) return 0.0f;
  float nr[3];
  decode_code(((const __attribute__((address_space(1))) uint32_t*)v.codes)[p], nr);
  const float fcol = (float)col, frow = (float)row;
  float vr0 = v.iK[0] * fcol + v.iK[1];
  float vr1 = v.iK[2] * frow + v.iK[3];
  float vr2 = 1.0f;
  const float nv = sqrtf(vr0 * vr0 + vr1 * vr1 + vr2 * vr2);
  vr0 = vr0 / nv;
  vr1 = vr1 / nv;
  vr2 = vr2 / nv;
  if (nr[0] * vr0 + nr[1] * vr1 + nr[2] * vr2 >= prm.max_view_cos) return 0.0f;
  const float fx = depth * (v.iK[0] * fcol + v.iK[1]);
  const float fy = depth * (v.iK[2] * frow + v.iK[3]);
  const float fz = depth;
  float depths[kCrossMaxSrc + 1];
  depths[0] = depth;
  int ng = 0;
  for (int i = 0; i < v.N; ++i) {
    const __attribute__((address_space(1))) float* pose =
        (const __attribute__((address_space(1))) float*)v.poses + i * PM_POSE_STRIDE;
    const __attribute__((address_space(1))) float* P = pose + 16;
    const __attribute__((address_space(1))) float* iP = pose + 28;
    const __attribute__((address_space(1))) float* R = pose + 4;
    const float inv_fz = 1.0f / (P[8] * fx + P[9] * fy + P[10] * fz + P[11]);
    float sc = inv_fz * (P[0] * fx + P[1] * fy + P[2] * fz + P[3]);
    float sr = inv_fz * (P[4] * fx + P[5] * fy + P[6] * fz + P[7]);
    // point sampling at (sc + 0.5, sr + 0.5) with a zero border
    const float xf = sc + 0.5f, yf = sr + 0.5f;
    const int sw = v.src[i].width, sh = v.src[i].height;
    float sd = 0.0f;
    uint32_t scode = 0u;
    if (xf >= 0.0f && xf < (float)sw && yf >= 0.0f && yf < (float)sh) {
      const int q = (int)floorf(yf) * sw + (int)floorf(xf);
      sd = ((const __attribute__((address_space(1))) float*)v.src[i].depth)[q];
      scode = ((const __attribute__((address_space(1))) uint32_t*)v.src[i].codes)[q];
    }
    if ((double)sd < 1e-4) continue;
    sc = sc * sd;
    sr = sr * sd;
    const float bx = iP[0] * sc + iP[1] * sr + iP[2] * sd + iP[3];
    const float by = iP[4] * sc + iP[5] * sr + iP[6] * sd + iP[7];
    const float bz = iP[8] * sc + iP[9] * sr + iP[10] * sd + iP[11];
    const float inv_bz = 1.0f / bz;
    const float bcol = inv_bz * (v.K[0] * bx + v.K[1] * bz);
    const float brow = inv_bz * (v.K[2] * by + v.K[3] * bz);
    const float dc = fcol - bcol, dr = frow - brow;
    const float reproj = sqrtf(dc * dc + dr * dr);
    if (reproj > prm.max_reproj || (depth - bz) * (depth - bz) > depth * depth * prm.sqr_depth_thr) continue;
    float sn[3];
    decode_code(scode, sn);
    const float r0 = R[0] * sn[0] + R[3] * sn[1] + R[6] * sn[2];
    const float r1 = R[1] * sn[0] + R[4] * sn[1] + R[7] * sn[2];
    const float r2 = R[2] * sn[0] + R[5] * sn[1] + R[8] * sn[2];
    if (nr[0] * r0 + nr[1] * r1 + nr[2] * r2 < prm.max_cos_normal) continue;
    ng++;
    depths[ng] = bz;
  }
  if (ng < v.min_consistent) return 0.0f;
  // BubbleSort(depths, num_good_src_depth) (utils.h:206-220): the first ng
  // of the ng + 1 collected depths, as the reference calls it
  for (int a = 0; a < ng - 1; ++a)
    for (int b = 0; b < ng - 1 - a; ++b)
      if (depths[b] > depths[b + 1]) {
        const float t = depths[b];
        depths[b] = depths[b + 1];
        depths[b + 1] = t;
      }
  return depths[ng / 2];
}

__global__ __launch_bounds__(256) void k_cross_filter(const CrossViewDev* __restrict__ views, CrossParams prm) {
  const int t = blockIdx.x;
  const __attribute__((address_space(4))) CrossViewDev* tab = (const __attribute__((address_space(4))) CrossViewDev*)views;
  int lo = 0, hi = prm.n_views - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tab[mid].tile_start <= t) lo = mid;
    else hi = mid - 1;
  }
  CV& v = tab[lo];
  const int lt = t - v.tile_start;
  const int tiles_x = (v.W + kCrossTile - 1) / kCrossTile;
  const int ty = lt / tiles_x, tx = lt - ty * tiles_x;
  const int col = tx * kCrossTile + (int)(threadIdx.x & (kCrossTile - 1));
  const int row = ty * kCrossTile + (int)(threadIdx.x / kCrossTile);
  if (col >= v.W || row >= v.H) return;
  ((__attribute__((address_space(1))) float*)v.out)[row * v.W + col] = cross_pixel(v, prm, row, col);
}

// cosf(DEG2RAD(deg)) as the reference forms it: the double product rounded
// to the float argument of cosf (kernel_functions.h:39).
float cos_deg(double deg) { return (float)std::cos((double)(float)(deg * 0.0174532925199432)); }

// A device buffer of one call, borrowed from the engine's block cache; the
// stream is synchronised before the block goes back (a call that fails may
// leave work on it).
struct DevBuf {
  void* p = nullptr;
  pmh::DevCache* cache = nullptr;
  hipStream_t s = nullptr;
  size_t bytes = 0;
  hipError_t get(pm_engine* e, hipStream_t st, size_t b) {
    cache = &e->cache;
    s = st;
    bytes = b;
    return cache->get(b, &p);
  }
  ~DevBuf() {
    if (!p) return;
    (void)hipStreamSynchronize(s);
    cache->put(p, bytes);
  }
};

}  // namespace

PM_API pm_status pm_cross_options_init(pm_cross_options* o) {
  if (!o) return fail(PM_ERR_INVALID_ARG, "null options");
  o->filter_geom_consistency_max_cost = 1.0f;
  o->max_normal_error = 15.0f;
  o->depth_diff_threshold = 0.02f;
  o->filter_min_num_consistent = 1;
  return PM_OK;
}

namespace {

// One image as the launch reads it: device depth map and normal codes.
struct CrossImgDev {
  const float* depth;
  const uint32_t* codes;
  int32_t width, height;
  pm_camera cam;
};

pm_status check_cross_args(const pm_cross_options* o, int32_t num_images, int32_t num_views,
                           const pm_cross_view* views) {
  if (!o || !views || num_images <= 0 || num_views <= 0) return fail(PM_ERR_INVALID_ARG, "invalid cross-filter arguments");
  if (!(o->filter_geom_consistency_max_cost >= 0.0f) || !(o->depth_diff_threshold >= 0.0f) ||
      !(o->max_normal_error >= 0.0f) || o->filter_min_num_consistent < 0)
    return fail(PM_ERR_INVALID_ARG, "invalid cross-filter options");
  for (int32_t k = 0; k < num_views; ++k) {
    const pm_cross_view& cv = views[k];
    if (cv.ref < 0 || cv.ref >= num_images || !cv.src)
      return fail(PM_ERR_INVALID_ARG, "invalid cross-filter view " + std::to_string(k));
    if (cv.num_src < 1 || cv.num_src > kCrossMaxSrc)
      return fail(PM_ERR_TOO_MANY_SOURCES, "cross-filter views take 1..31 sources");
    for (int32_t s = 0; s < cv.num_src; ++s)
      if (cv.src[s] < 0 || cv.src[s] >= num_images)
        return fail(PM_ERR_INVALID_ARG, "cross-filter source index out of range in view " + std::to_string(k));
  }
  return PM_OK;
}

// The launch over device-resident images: builds the view table and runs
// k_cross_filter, writing view k's refined map at d_out + out_off[k].
// Leaves the work enqueued on the engine's stream.
pm_status run_cross(pm_engine* e, const pm_cross_options* o, const std::vector<CrossImgDev>& imgs, int32_t num_views,
                    const pm_cross_view* views, float* d_out, const std::vector<int64_t>& out_off, DevBuf& d_poses,
                    DevBuf& d_tab) {
  hipStream_t st = e->stream;
  std::vector<float> poses;
  std::vector<CrossViewDev> tab(num_views);
  std::vector<size_t> pose_off(num_views);
  int64_t tiles = 0;
  for (int32_t k = 0; k < num_views; ++k) {
    const pm_cross_view& cv = views[k];
    const CrossImgDev& ref = imgs[cv.ref];
    pm_view_desc vd{};
    std::vector<pm_image> srcs(cv.num_src);
    std::vector<pm_camera> cams(cv.num_src);
    for (int32_t s = 0; s < cv.num_src; ++s) {
      const CrossImgDev& si = imgs[cv.src[s]];
      srcs[s].width = si.width;
      srcs[s].height = si.height;
      cams[s] = si.cam;
    }
    vd.num_src = cv.num_src;
    vd.ref_cam = ref.cam;
    vd.src = srcs.data();
    vd.src_cams = cams.data();
    float k8[8];
    pose_off[k] = poses.size();
    poses.resize(poses.size() + (size_t)cv.num_src * PM_POSE_STRIDE);
    pack_poses(&vd, poses.data() + pose_off[k], k8);
    CrossViewDev& t = tab[k];
    std::memset(&t, 0, sizeof(t));
    t.W = ref.width;
    t.H = ref.height;
    t.N = cv.num_src;
    t.min_consistent = std::min(o->filter_min_num_consistent, cv.num_src);
    std::memcpy(t.K, k8, sizeof(t.K));
    std::memcpy(t.iK, k8 + 4, sizeof(t.iK));
    t.depth = ref.depth;
    t.codes = ref.codes;
    t.out = d_out + out_off[k];
    for (int32_t s = 0; s < cv.num_src; ++s) {
      const CrossImgDev& si = imgs[cv.src[s]];
      t.src[s].depth = si.depth;
      t.src[s].codes = si.codes;
      t.src[s].width = si.width;
      t.src[s].height = si.height;
    }
    t.tile_start = (int32_t)tiles;
    tiles += (int64_t)((ref.width + kCrossTile - 1) / kCrossTile) * ((ref.height + kCrossTile - 1) / kCrossTile);
    if (tiles > INT32_MAX) return fail(PM_ERR_INVALID_ARG, "cross filter: too many pixels in one call");
  }
  PM_HIP(d_poses.get(e, st, std::max<size_t>(poses.size(), 1) * sizeof(float)));
  PM_HIP(hipMemcpyAsync(d_poses.p, poses.data(), poses.size() * sizeof(float), hipMemcpyHostToDevice, st));
  for (int32_t k = 0; k < num_views; ++k) tab[k].poses = (const float*)d_poses.p + pose_off[k];
  PM_HIP(d_tab.get(e, st, tab.size() * sizeof(CrossViewDev)));
  PM_HIP(hipMemcpyAsync(d_tab.p, tab.data(), tab.size() * sizeof(CrossViewDev), hipMemcpyHostToDevice, st));

  CrossParams prm;
  prm.max_reproj = o->filter_geom_consistency_max_cost;
  prm.sqr_depth_thr = o->depth_diff_threshold * o->depth_diff_threshold;
  prm.max_cos_normal = cos_deg((double)o->max_normal_error);
  prm.max_view_cos = cos_deg(100.0);
  prm.n_views = num_views;
  hipLaunchKernelGGL(k_cross_filter, dim3((unsigned)tiles), dim3(kCrossTile * kCrossTile), 0, st,
                     (const CrossViewDev*)d_tab.p, prm);
  PM_HIP(hipGetLastError());
  return PM_OK;
}

// Output offsets of the views' refined maps in one device buffer.
std::vector<int64_t> out_offsets(const std::vector<CrossImgDev>& imgs, int32_t num_views, const pm_cross_view* views,
                                 int64_t* total) {
  std::vector<int64_t> off(num_views);
  int64_t t = 0;
  for (int32_t k = 0; k < num_views; ++k) {
    off[k] = t;
    t += (int64_t)imgs[views[k].ref].width * imgs[views[k].ref].height;
  }
  *total = t;
  return off;
}

}  // namespace

PM_API pm_status pm_cross_filter(pm_engine* e, const pm_cross_options* o, int32_t num_images,
                                 const pm_cross_image* images, int32_t num_views, const pm_cross_view* views) {
  if (!e || !images) return fail(PM_ERR_INVALID_ARG, "invalid cross-filter arguments");
  if (pm_status s = check_cross_args(o, num_images, num_views, views)) return s;
  for (int32_t k = 0; k < num_views; ++k)
    if (!views[k].refined) return fail(PM_ERR_INVALID_ARG, "invalid cross-filter view " + std::to_string(k));
  for (int32_t i = 0; i < num_images; ++i) {
    const pm_cross_image& im = images[i];
    if (im.width <= 0 || im.height <= 0 || (int64_t)im.width * im.height > ((int64_t)1 << 28) || !im.depth ||
        !im.normal_codes || !finite_camera(im.cam))
      return fail(PM_ERR_INVALID_ARG, "invalid cross-filter image " + std::to_string(i));
  }
  PM_HIP(hipSetDevice(e->device));
  hipStream_t st = e->stream;

  // images used by some view: depth and codes uploaded once each
  std::vector<int64_t> img_off(num_images, -1);
  int64_t img_total = 0;
  auto use = [&](int32_t i) {
    if (img_off[i] < 0) {
      img_off[i] = img_total;
      img_total += (int64_t)images[i].width * images[i].height;
    }
  };
  for (int32_t k = 0; k < num_views; ++k) {
    use(views[k].ref);
    for (int32_t s = 0; s < views[k].num_src; ++s) use(views[k].src[s]);
  }
  DevBuf d_depth, d_codes, d_out, d_poses, d_tab;
  PM_HIP(d_depth.get(e, st, (size_t)img_total * sizeof(float)));
  PM_HIP(d_codes.get(e, st, (size_t)img_total * sizeof(uint32_t)));
  std::vector<CrossImgDev> imgs(num_images);
  for (int32_t i = 0; i < num_images; ++i) {
    imgs[i].width = images[i].width;
    imgs[i].height = images[i].height;
    imgs[i].cam = images[i].cam;
    if (img_off[i] < 0) continue;
    const size_t n = (size_t)images[i].width * images[i].height;
    imgs[i].depth = (const float*)d_depth.p + img_off[i];
    imgs[i].codes = (const uint32_t*)d_codes.p + img_off[i];
    if (pm_status s = copy_h2d(e, (float*)d_depth.p + img_off[i], images[i].depth, n * sizeof(float), st)) return s;
    if (pm_status s = copy_h2d(e, (uint32_t*)d_codes.p + img_off[i], images[i].normal_codes, n * sizeof(uint32_t), st))
      return s;
  }
  int64_t out_total = 0;
  const std::vector<int64_t> off = out_offsets(imgs, num_views, views, &out_total);
  PM_HIP(d_out.get(e, st, (size_t)out_total * sizeof(float)));
  if (pm_status s = run_cross(e, o, imgs, num_views, views, (float*)d_out.p, off, d_poses, d_tab)) return s;
  for (int32_t k = 0; k < num_views; ++k) {
    const size_t n = (size_t)images[views[k].ref].width * images[views[k].ref].height;
    if (pm_status s = copy_d2h(e, views[k].refined, (const float*)d_out.p + off[k], n * sizeof(float), st)) return s;
  }
  PM_HIP(hipStreamSynchronize(st));
  return PM_OK;
}

PM_API pm_status pm_cross_filter_problems(pm_engine* e, const pm_cross_options* o, int32_t num_problems,
                                          pm_problem* const* problems, int32_t num_views, const pm_cross_view* views) {
  if (!e || !problems) return fail(PM_ERR_INVALID_ARG, "invalid cross-filter arguments");
  if (pm_status s = check_cross_args(o, num_problems, num_views, views)) return s;
  for (int32_t i = 0; i < num_problems; ++i) {
    const pm_problem* p = problems[i];
    if (!p || p->engine != e || !p->host.depth || !p->host.normal || !finite_camera(p->ref_cam))
      return fail(PM_ERR_INVALID_ARG, "invalid cross-filter problem " + std::to_string(i) +
                                          " (null, on another engine, or without maps)");
  }
  for (int32_t a = 0; a < num_views; ++a)
    for (int32_t b = a + 1; b < num_views; ++b)
      if (!views[a].refined && views[a].ref == views[b].ref && !views[b].refined)
        return fail(PM_ERR_INVALID_ARG, "two views write back into the same problem");
  PM_HIP(hipSetDevice(e->device));
  hipStream_t st = e->stream;

  // normal codes of the problems that some view reads, encoded on the device
  std::vector<int64_t> code_off(num_problems, -1);
  int64_t code_total = 0;
  auto use = [&](int32_t i) {
    if (code_off[i] < 0) {
      code_off[i] = code_total;
      code_total += (int64_t)problems[i]->W * problems[i]->H;
    }
  };
  for (int32_t k = 0; k < num_views; ++k) {
    use(views[k].ref);
    for (int32_t s = 0; s < views[k].num_src; ++s) use(views[k].src[s]);
  }
  DevBuf d_codes, d_out, d_poses, d_tab;
  PM_HIP(d_codes.get(e, st, (size_t)std::max<int64_t>(code_total, 1) * sizeof(uint32_t)));
  std::vector<CrossImgDev> imgs(num_problems);
  for (int32_t i = 0; i < num_problems; ++i) {
    const pm_problem* p = problems[i];
    imgs[i].width = p->W;
    imgs[i].height = p->H;
    imgs[i].cam = p->ref_cam;
    imgs[i].depth = p->host.depth;
    if (code_off[i] < 0) continue;
    uint32_t* c = (uint32_t*)d_codes.p + code_off[i];
    imgs[i].codes = c;
    PM_HIP(encode_normals_async(p->host.normal, (int64_t)p->W * p->H, c, st));
  }
  int64_t out_total = 0;
  const std::vector<int64_t> off = out_offsets(imgs, num_views, views, &out_total);
  PM_HIP(d_out.get(e, st, (size_t)std::max<int64_t>(out_total, 1) * sizeof(float)));
  if (pm_status s = run_cross(e, o, imgs, num_views, views, (float*)d_out.p, off, d_poses, d_tab)) return s;
  // after the launch (every view has read its sources): results out
  for (int32_t k = 0; k < num_views; ++k) {
    const pm_problem* p = problems[views[k].ref];
    const size_t bytes = (size_t)p->W * p->H * sizeof(float);
    if (views[k].refined)
      PM_HIP(hipMemcpyAsync(views[k].refined, (const float*)d_out.p + off[k], bytes, hipMemcpyDeviceToHost, st));
    else
      PM_HIP(hipMemcpyAsync(p->host.depth, (const float*)d_out.p + off[k], bytes, hipMemcpyDeviceToDevice, st));
  }
  PM_HIP(hipStreamSynchronize(st));
  return PM_OK;
}

// ---------------------------------------------------------------------------
// CrossFilterCuda::DedupRun (patch_match_cuda.cu:1561-1612): visibility words
// per image, DeduplicFilterKernel (kernel_functions.cu:3780-3916) per view in
// the given order.  Consecutive views that touch disjoint images (no view's
// reference is another's reference or source) share one launch; that is the
// same as running them one after another, since a view writes only its
// reference's words and reads only its sources'.

namespace {

struct DedupSrc {
  const float* depth;
  const uint32_t* vis;
  int32_t width, height;
  int32_t flag, pad;
};

struct DedupViewDev {
  int32_t W, H, N, tile_start;
  int32_t min_consistent, pad0;
  float K[4], iK[4];
  const float* depth;
  uint32_t* vis;
  float* delt;         // may be null
  const float* poses;  // N * PM_POSE_STRIDE
  DedupSrc src[kCrossMaxSrc];
};

struct DedupParams {
  float max_reproj;     // filter_geom_consistency_max_cost
  float depth_thr;      // depth_diff_threshold
  float sqr_depth_thr;  // depth_diff_threshold^2
  int32_t first, count; // the group's views in the table
};

using DV = const __attribute__((address_space(4))) DedupViewDev;
using gfloat = const __attribute__((address_space(1))) float;
using guint = __attribute__((address_space(1))) uint32_t;

__device__ void dedup_pixel(DV& v, const DedupParams& prm, int row, int col) {
  const int p = row * v.W + col;
  guint* rv = (guint*)v.vis;
  __attribute__((address_space(1))) float* delt = (__attribute__((address_space(1))) float*)v.delt;
  if (delt) delt[p] = 0.0f;
  const float depth = ((gfloat*)v.depth)[p];
  if ((double)depth < 1e-6 || col < 3 || col > v.W - 3 || row < 3 || row > v.H - 3) {
    rv[p] = 0u;
    return;
  }
  const uint32_t word = rv[p];
  if (!(word >> 3)) {  // visbility_value >> (sizeof(uint32_t) - 1) (:3803)
    rv[p] = 0u;
    return;
  }
  const float fcol = (float)col, frow = (float)row;
  const float fx = depth * (v.iK[0] * fcol + v.iK[1]);
  const float fy = depth * (v.iK[2] * frow + v.iK[3]);
  const float fz = depth;
  float deltas[kCrossMaxSrc];
  uint32_t bits = word;
  int ng = 0, nvis = 0;
  for (int i = 0; i < v.N; ++i) {
    gfloat* pose = (gfloat*)v.poses + i * PM_POSE_STRIDE;
    gfloat* P = pose + 16;
    gfloat* iP = pose + 28;
    const float inv_fz = 1.0f / (P[8] * fx + P[9] * fy + P[10] * fz + P[11]);
    float sc = inv_fz * (P[0] * fx + P[1] * fy + P[2] * fz + P[3]);
    float sr = inv_fz * (P[4] * fx + P[5] * fy + P[6] * fz + P[7]);
    // point sampling at (sc + 0.5, sr + 0.5) with a zero border
    const float xf = sc + 0.5f, yf = sr + 0.5f;
    const int sw = v.src[i].width, sh = v.src[i].height;
    if (!(xf >= 0.0f && xf < (float)sw && yf >= 0.0f && yf < (float)sh)) continue;
    const int q = (int)floorf(yf) * sw + (int)floorf(xf);
    const float sd = ((gfloat*)v.src[i].depth)[q];
    if (sd == 0.0f) continue;
    sc = sc * sd;
    sr = sr * sd;
    const float bx = iP[0] * sc + iP[1] * sr + iP[2] * sd + iP[3];
    const float by = iP[4] * sc + iP[5] * sr + iP[6] * sd + iP[7];
    const float bz = iP[8] * sc + iP[9] * sr + iP[10] * sd + iP[11];
    const float inv_bz = 1.0f / bz;
    const float bcol = inv_bz * (v.K[0] * bx + v.K[1] * bz);
    const float brow = inv_bz * (v.K[2] * by + v.K[3] * bz);
    const float dc = fcol - bcol, dr = frow - brow;
    const float reproj = sqrtf(dc * dc + dr * dr);
    if (reproj > prm.max_reproj || (depth - bz) * (depth - bz) > depth * depth * prm.sqr_depth_thr) continue;
    deltas[ng++] = fabsf(bz - depth);
    bits |= 1u << (kCrossMaxSrc - 1 - i);  // MAX_NUM_CROSS_SRC - 1 - idx (:3907)
    // the source's word at the sampled texel ((size_t)(x + 0.5), x + 0.5 >= 0)
    if ((((const __attribute__((address_space(1))) uint32_t*)v.src[i].vis)[q] & 0x80000000u) && v.src[i].flag) nvis++;
  }
  if (ng < v.min_consistent || nvis > 0) {
    rv[p] = 0u;
    return;
  }
  rv[p] = bits;
  if (ng == 0 || !delt) return;
  // BubbleSort(delt_depths, num_good_src_depth - 1): all but the last entry;
  // median index int((num_good_src_depth - 1) * 0.5) (:3912-3914)
  const int len = ng - 1;
  for (int a = 0; a < len - 1; ++a)
    for (int b = 0; b < len - 1 - a; ++b)
      if (deltas[b] > deltas[b + 1]) {
        const float t = deltas[b];
        deltas[b] = deltas[b + 1];
        deltas[b + 1] = t;
      }
  delt[p] = deltas[(ng - 1) >> 1] / (prm.depth_thr * depth);
}

__global__ __launch_bounds__(256) void k_dedup(const DedupViewDev* __restrict__ views, DedupParams prm) {
  const int t = blockIdx.x;
  const __attribute__((address_space(4))) DedupViewDev* tab =
      (const __attribute__((address_space(4))) DedupViewDev*)views;
  int lo = prm.first, hi = prm.first + prm.count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tab[mid].tile_start <= t) lo = mid;
    else hi = mid - 1;
  }
  DV& v = tab[lo];
  const int lt = t - v.tile_start;
  const int tiles_x = (v.W + kCrossTile - 1) / kCrossTile;
  const int ty = lt / tiles_x, tx = lt - ty * tiles_x;
  const int col = tx * kCrossTile + (int)(threadIdx.x & (kCrossTile - 1));
  const int row = ty * kCrossTile + (int)(threadIdx.x / kCrossTile);
  if (col >= v.W || row >= v.H) return;
  dedup_pixel(v, prm, row, col);
}

// FillGpuMat (kernel_functions.cu:3930-3941): every word 0x80000000.
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t value) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = value;
}

// GetDeduplicDepthMapsLayer (kernel_functions.cu:3943-3959) over all images.
__global__ __launch_bounds__(256) void k_dedup_depth(const float* __restrict__ depth, const uint32_t* __restrict__ vis,
                                                     float* __restrict__ out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (vis[i] & 0x80000000u) ? depth[i] : 0.0f;
}

unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096)); }

}  // namespace

PM_API pm_status pm_cross_dedup(pm_engine* e, const pm_cross_options* o, int32_t num_images,
                                const pm_dedup_image* images, int32_t num_views, const pm_dedup_view* views) {
  if (!e || !o || !images || num_images <= 0 || num_views < 0 || (num_views > 0 && !views))
    return fail(PM_ERR_INVALID_ARG, "invalid dedup arguments");
  if (!(o->filter_geom_consistency_max_cost >= 0.0f) || !(o->depth_diff_threshold >= 0.0f) ||
      o->filter_min_num_consistent < 0)
    return fail(PM_ERR_INVALID_ARG, "invalid dedup options");
  for (int32_t i = 0; i < num_images; ++i) {
    const pm_dedup_image& im = images[i];
    if (im.width <= 0 || im.height <= 0 || (int64_t)im.width * im.height > ((int64_t)1 << 28) || !im.depth ||
        !finite_camera(im.cam))
      return fail(PM_ERR_INVALID_ARG, "invalid dedup image " + std::to_string(i));
  }
  for (int32_t k = 0; k < num_views; ++k) {
    const pm_dedup_view& v = views[k];
    if (v.ref < 0 || v.ref >= num_images || v.num_src < 0 || (v.num_src > 0 && !v.src))
      return fail(PM_ERR_INVALID_ARG, "invalid dedup view " + std::to_string(k));
    if (v.num_src > kCrossMaxSrc) return fail(PM_ERR_TOO_MANY_SOURCES, "dedup views take 0..31 sources");
    for (int32_t s = 0; s < v.num_src; ++s)
      if (v.src[s] < 0 || v.src[s] >= num_images || v.src[s] == v.ref)
        return fail(PM_ERR_INVALID_ARG, "dedup source index out of range or equal to the reference in view " +
                                            std::to_string(k));
  }
  PM_HIP(hipSetDevice(e->device));
  hipStream_t st = e->stream;

  // every image's depth and visibility words on the device
  std::vector<int64_t> off(num_images);
  int64_t total = 0;
  for (int32_t i = 0; i < num_images; ++i) {
    off[i] = total;
    total += (int64_t)images[i].width * images[i].height;
  }
  int64_t delt_total = 0;
  std::vector<int64_t> delt_off(num_views, -1);
  for (int32_t k = 0; k < num_views; ++k) {
    if (!views[k].delt || views[k].num_src < 1) continue;
    delt_off[k] = delt_total;
    delt_total += (int64_t)images[views[k].ref].width * images[views[k].ref].height;
  }
  DevBuf d_depth, d_vis, d_out, d_delt, d_poses, d_tab;
  PM_HIP(d_depth.get(e, st, (size_t)total * sizeof(float)));
  PM_HIP(d_vis.get(e, st, (size_t)total * sizeof(uint32_t)));
  PM_HIP(d_out.get(e, st, (size_t)total * sizeof(float)));
  PM_HIP(d_delt.get(e, st, (size_t)std::max<int64_t>(delt_total, 1) * sizeof(float)));
  float* depth = (float*)d_depth.p;
  uint32_t* vis = (uint32_t*)d_vis.p;
  for (int32_t i = 0; i < num_images; ++i)
    if (pm_status s = copy_h2d(e, depth + off[i], images[i].depth,
                               (size_t)images[i].width * images[i].height * sizeof(float), st))
      return s;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(total)), dim3(256), 0, st, vis, total, 0x80000000u);
  PM_HIP(hipGetLastError());

  // view table and poses; groups of consecutive independent views
  std::vector<DedupViewDev> tab;
  std::vector<float> poses;
  std::vector<size_t> pose_off;
  std::vector<std::pair<int, int>> groups;  // (first, count) in tab
  std::vector<int32_t> touched(num_images, -1);  // group that reads or writes the image
  std::vector<uint8_t> written(num_images, 0);   // the image is a reference in the current group
  int64_t tiles = 0;
  for (int32_t k = 0; k < num_views; ++k) {
    const pm_dedup_view& cv = views[k];
    if (cv.num_src < 1) continue;  // FindSrcImageLists gave < 2 entries (:1586-1588)
    // conflicts with the current group: its references vs this view's
    // images, this view's reference vs every image the group touches
    const int g = (int)groups.size() - 1;
    bool conflict = g < 0 || touched[cv.ref] == g;
    for (int32_t s = 0; s < cv.num_src && !conflict; ++s) conflict = touched[cv.src[s]] == g && written[cv.src[s]];
    if (conflict) {
      groups.push_back({(int)tab.size(), 0});
      std::fill(written.begin(), written.end(), 0);
      tiles = 0;
    }
    const int cur = (int)groups.size() - 1;
    groups.back().second++;
    touched[cv.ref] = cur;
    written[cv.ref] = 1;
    for (int32_t s = 0; s < cv.num_src; ++s) touched[cv.src[s]] = cur;

    const pm_dedup_image& ref = images[cv.ref];
    pm_view_desc vd{};
    std::vector<pm_image> srcs(cv.num_src);
    std::vector<pm_camera> cams(cv.num_src);
    for (int32_t s = 0; s < cv.num_src; ++s) {
      srcs[s].width = images[cv.src[s]].width;
      srcs[s].height = images[cv.src[s]].height;
      cams[s] = images[cv.src[s]].cam;
    }
    vd.num_src = cv.num_src;
    vd.ref_cam = ref.cam;
    vd.src = srcs.data();
    vd.src_cams = cams.data();
    float k8[8];
    pose_off.push_back(poses.size());
    poses.resize(poses.size() + (size_t)cv.num_src * PM_POSE_STRIDE);
    pack_poses(&vd, poses.data() + pose_off.back(), k8);
    DedupViewDev t;
    std::memset(&t, 0, sizeof(t));
    t.W = ref.width;
    t.H = ref.height;
    t.N = cv.num_src;
    t.min_consistent = std::min(o->filter_min_num_consistent, cv.num_src);
    std::memcpy(t.K, k8, sizeof(t.K));
    std::memcpy(t.iK, k8 + 4, sizeof(t.iK));
    t.depth = depth + off[cv.ref];
    t.vis = vis + off[cv.ref];
    t.delt = delt_off[k] >= 0 ? (float*)d_delt.p + delt_off[k] : nullptr;
    for (int32_t s = 0; s < cv.num_src; ++s) {
      const int32_t si = cv.src[s];
      t.src[s].depth = depth + off[si];
      t.src[s].vis = vis + off[si];
      t.src[s].width = images[si].width;
      t.src[s].height = images[si].height;
      t.src[s].flag = images[si].ref_flag ? 1 : 0;
    }
    t.tile_start = (int32_t)tiles;
    tiles += (int64_t)((ref.width + kCrossTile - 1) / kCrossTile) * ((ref.height + kCrossTile - 1) / kCrossTile);
    if (tiles > INT32_MAX) return fail(PM_ERR_INVALID_ARG, "dedup: too many pixels in one group");
    tab.push_back(t);
  }
  if (!tab.empty()) {
    PM_HIP(d_poses.get(e, st, std::max<size_t>(poses.size(), 1) * sizeof(float)));
    PM_HIP(hipMemcpyAsync(d_poses.p, poses.data(), poses.size() * sizeof(float), hipMemcpyHostToDevice, st));
    for (size_t k = 0; k < tab.size(); ++k) tab[k].poses = (const float*)d_poses.p + pose_off[k];
    PM_HIP(d_tab.get(e, st, tab.size() * sizeof(DedupViewDev)));
    PM_HIP(hipMemcpyAsync(d_tab.p, tab.data(), tab.size() * sizeof(DedupViewDev), hipMemcpyHostToDevice, st));
    DedupParams prm;
    prm.max_reproj = o->filter_geom_consistency_max_cost;
    prm.depth_thr = o->depth_diff_threshold;
    prm.sqr_depth_thr = o->depth_diff_threshold * o->depth_diff_threshold;
    for (const auto& g : groups) {
      const DedupViewDev& last = tab[g.first + g.second - 1];
      const int64_t n_tiles = (int64_t)last.tile_start +
                              (int64_t)((last.W + kCrossTile - 1) / kCrossTile) * ((last.H + kCrossTile - 1) / kCrossTile);
      prm.first = g.first;
      prm.count = g.second;
      hipLaunchKernelGGL(k_dedup, dim3((unsigned)n_tiles), dim3(kCrossTile * kCrossTile), 0, st,
                         (const DedupViewDev*)d_tab.p, prm);
      PM_HIP(hipGetLastError());
    }
  }
  hipLaunchKernelGGL(k_dedup_depth, dim3(grid_for(total)), dim3(256), 0, st, (const float*)depth, (const uint32_t*)vis,
                     (float*)d_out.p, total);
  PM_HIP(hipGetLastError());
  for (int32_t i = 0; i < num_images; ++i) {
    const size_t n = (size_t)images[i].width * images[i].height;
    if (images[i].visibility)
      if (pm_status s = copy_d2h(e, images[i].visibility, vis + off[i], n * sizeof(uint32_t), st)) return s;
    if (images[i].dedup_depth)
      if (pm_status s = copy_d2h(e, images[i].dedup_depth, (const float*)d_out.p + off[i], n * sizeof(float), st))
        return s;
  }
  for (int32_t k = 0; k < num_views; ++k) {
    if (!views[k].delt) continue;
    const size_t n = (size_t)images[views[k].ref].width * images[views[k].ref].height;
    if (delt_off[k] >= 0) {
      if (pm_status s = copy_d2h(e, views[k].delt, (const float*)d_delt.p + delt_off[k], n * sizeof(float), st))
        return s;
    } else
      std::memset(views[k].delt, 0, n * sizeof(float));
  }
  PM_HIP(hipStreamSynchronize(st));
  return PM_OK;
}
