// pm_engine.hip -- host side of the gfx950 PatchMatch engine and its C-ABI
// (include/sensemap_pm.h).
//
// One pm_engine owns a device and a HIP stream.  One pm_problem is one
// reference view with its sources resident in HBM: the constructor of
// SenseMap's PatchMatchACMMCuda (src/mvs/patch_match_cuda.cu:422-432) becomes
// pm_problem_create, Run() (:748-1057) becomes pm_problem_run and the
// Get*Map() getters become pm_problem_get_*.  Errors are returned as
// pm_status with a per-thread message, never by exit() as the reference's
// CUDA_SAFE_CALL does (src/util/cudacc.cc:30-38).
#include "pm_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <new>

#include "pm_math.h"

using namespace pmh;

namespace pmh {

namespace {
thread_local std::string g_last_error;
}  // namespace

pm_status fail(pm_status st, const std::string& msg) {
  g_last_error = msg;
  return st;
}

const std::string& last_error() { return g_last_error; }

// ---------------------------------------------------------------------------
// arena

Arena::~Arena() {
  if (base_) (void)hipFree(base_);
}

pm_status Arena::init(size_t bytes) {
  bytes = round(bytes);
  PM_HIP(hipMalloc(&base_, bytes));
  cap_ = bytes;
  used_ = 0;
  holes_.clear();
  holes_[0] = bytes;
  return PM_OK;
}

bool Arena::alloc(size_t bytes, size_t* off) {
  bytes = round(bytes > 0 ? bytes : 1);
  std::lock_guard<std::mutex> lk(mu_);
  for (auto it = holes_.begin(); it != holes_.end(); ++it) {
    if (it->second < bytes) continue;
    *off = it->first;
    const size_t rest = it->second - bytes;
    const size_t roff = it->first + bytes;
    holes_.erase(it);
    if (rest) holes_[roff] = rest;
    used_ += bytes;
    return true;
  }
  return false;
}

void Arena::release(size_t off, size_t bytes) {
  bytes = round(bytes > 0 ? bytes : 1);
  std::lock_guard<std::mutex> lk(mu_);
  used_ -= bytes;
  auto it = holes_.emplace(off, bytes).first;
  auto nx = std::next(it);
  if (nx != holes_.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    holes_.erase(nx);
  }
  if (it != holes_.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      holes_.erase(it);
    }
  }
}

hipError_t DevCache::get(size_t bytes, void** out) {
  const size_t b = round(bytes);
  {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = blocks_.find(b);
    if (it != blocks_.end()) {
      *out = it->second;
      blocks_.erase(it);
      bytes_ -= b;
      return hipSuccess;
    }
  }
  return hipMalloc(out, b);
}

void DevCache::put(void* p, size_t bytes) {
  if (!p) return;
  const size_t b = round(bytes);
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (bytes_ + b <= kMaxBytes) {
      blocks_.emplace(b, p);
      bytes_ += b;
      return;
    }
  }
  (void)hipFree(p);
}

DevCache::~DevCache() {
  for (auto& kv : blocks_) (void)hipFree(kv.second);
}

size_t Arena::used() const {
  std::lock_guard<std::mutex> lk(mu_);
  return used_;
}

size_t Arena::largest_hole() const {
  std::lock_guard<std::mutex> lk(mu_);
  size_t m = 0;
  for (const auto& h : holes_) m = std::max(m, h.second);
  return m;
}

}  // namespace pmh

// ---------------------------------------------------------------------------
// options

PM_API pm_status pm_options_init(pm_options* o) {
  if (!o) return fail(PM_ERR_INVALID_ARG, "pm_options_init: null options");
  std::memset(o, 0, sizeof(*o));
  o->window_radius = 5;
  o->window_step = 2;
  o->sigma_spatial = 5.0f;  // sigma_spatial = window_radius (patch_match_options.h:89)
  o->sigma_color = 0.2f;
  o->num_samples = 15;
  o->num_iterations = 5;
  o->geom_consistency = 1;
  o->geom_consistency_regularizer = 0.3f;
  o->geom_consistency_max_cost = 3.0f;
  o->random_optimization = 1;
  o->local_optimization = 1;
  o->random_depth_ratio = 0.004f;
  o->random_angle1_range = (float)(20.0 * M_PI / 180.0);
  o->random_angle2_range = (float)(12.0 * M_PI / 180.0);
  o->init_ncc_matching_cost = 0.8f;
  o->max_ncc_matching_cost = 1.2f;
  o->alpha = 90.0f;
  o->beta = 0.3f;
  o->thk = 4;
  o->num_good_hypothesis = 2;
  o->num_bad_hypothesis = 3;
  o->init_depth_random = 0;  // = !geom_consistency
  o->filter = 1;
  o->conf_filter = 0;
  o->est_curvature = 0;
  o->median_filter = 0;
  o->conf_threshold = 0.5f;
  o->depth_diff_threshold = 0.02f;
  o->filter_geom_consistency_max_cost = 1.0f;
  o->filter_min_num_consistent = 1;
  o->init_samples = 1;
  o->compat_lo_geo_bug = 1;
  o->image_storage = PM_STORAGE_F16;  // the 8-bit values exactly, fewer instructions per sample
  o->accum_mode = PM_ACCUM_PLAIN;
  o->rcp_mode = PM_RCP_EXACT;
  o->seed = 0;
  return PM_OK;
}

PM_API pm_status pm_options_check(const pm_options* o) {
  if (!o) return fail(PM_ERR_INVALID_ARG, "null options");
  // PatchMatchOptions::Check (patch_match_options.cc:75-101)
  if (o->window_radius <= 0 || o->window_radius > PM_MAX_WINDOW_RADIUS)
    return fail(PM_ERR_INVALID_ARG, "window_radius must be in [1, 32]");
  if (o->window_step <= 0 || o->window_step > 2) return fail(PM_ERR_INVALID_ARG, "window_step must be 1 or 2");
  if (!(o->sigma_color > 0.0f)) return fail(PM_ERR_INVALID_ARG, "sigma_color must be > 0");
  if (!(o->sigma_spatial > 0.0f)) return fail(PM_ERR_INVALID_ARG, "sigma_spatial must be > 0");
  if (o->num_samples <= 0) return fail(PM_ERR_INVALID_ARG, "num_samples must be > 0");
  if (o->num_iterations < 0) return fail(PM_ERR_INVALID_ARG, "num_iterations must be >= 0");
  if (!(o->geom_consistency_regularizer >= 0.0f))
    return fail(PM_ERR_INVALID_ARG, "geom_consistency_regularizer must be >= 0");
  if (!(o->geom_consistency_max_cost >= 0.0f)) return fail(PM_ERR_INVALID_ARG, "geom_consistency_max_cost must be >= 0");
  // the reference CHECKs that they are not both set (patch_match_cuda.cu:950); this engine builds neither
  if (o->has_prior_depth)
    return fail(PM_ERR_INVALID_ARG, "has_prior_depth (RGB-D depth priors) is not supported by this engine");
  if (o->propagate_depth)
    return fail(PM_ERR_INVALID_ARG,
                "propagate_depth is not supported: the reference's sweep reads prior normal and weight maps it never "
                "allocates for it (patch_match_cuda.cu:589-591, kernel_functions.cu:2191-2199)");
  // limits of this engine
  if (o->thk < 1 || o->thk > 4) return fail(PM_ERR_INVALID_ARG, "thk must be in [1, 4] (mweights has 5 entries)");
  if (!(o->beta > 0.0f)) return fail(PM_ERR_INVALID_ARG, "beta must be > 0");
  if (!(o->alpha > 0.0f)) return fail(PM_ERR_INVALID_ARG, "alpha must be > 0");
  if (o->init_samples < 1 || o->init_samples > 64) return fail(PM_ERR_INVALID_ARG, "init_samples must be in [1, 64]");
  if (o->image_storage != PM_STORAGE_U8 && o->image_storage != PM_STORAGE_F16)
    return fail(PM_ERR_INVALID_ARG, "image_storage must be PM_STORAGE_U8 or PM_STORAGE_F16");
  if (o->accum_mode != PM_ACCUM_PLAIN && o->accum_mode != PM_ACCUM_COMPENSATED)
    return fail(PM_ERR_INVALID_ARG, "accum_mode must be PM_ACCUM_PLAIN or PM_ACCUM_COMPENSATED");
  if (o->rcp_mode != PM_RCP_EXACT && o->rcp_mode != PM_RCP_FAST)
    return fail(PM_ERR_INVALID_ARG, "rcp_mode must be PM_RCP_EXACT or PM_RCP_FAST");
  return PM_OK;
}

PM_API const char* pm_last_error(void) { return pmh::last_error().c_str(); }
PM_API int32_t pm_abi_version(void) { return PM_ABI_VERSION; }

// ---------------------------------------------------------------------------
// engine

PM_API pm_status pm_engine_create(int32_t device, pm_engine** out) {
  if (!out) return fail(PM_ERR_INVALID_ARG, "null out");
  *out = nullptr;
  int count = 0;
  PM_HIP(hipGetDeviceCount(&count));
  if (device < 0) device = 0;
  if (device >= count) return fail(PM_ERR_INVALID_ARG, "device index out of range");
  PM_HIP(hipSetDevice(device));
  pm_engine* e = new (std::nothrow) pm_engine();
  if (!e) return fail(PM_ERR_OUT_OF_MEMORY, "host allocation");
  e->device = device;
  hipError_t he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  if (he != hipSuccess) {
    delete e;
    return fail(PM_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(he));
  }
  *out = e;
  return PM_OK;
}

namespace pmh {

void engine_retain(pm_engine* e) { e->refs.fetch_add(1, std::memory_order_relaxed); }

namespace {

// Engines holding pinned staging buffers.  At exit, before the HIP runtime's
// own teardown (this handler is registered after the runtime started, and
// exit handlers run last-in first-out), the buffers and their events of
// engines never destroyed are freed.
std::mutex g_staged_mu;
std::vector<pm_engine*>* g_staged = nullptr;  // never freed: outlives every engine

void free_staging(pm_engine* e) {
  for (int b = 0; b < 2; ++b) {
    if (e->stage_ev[b]) (void)hipEventDestroy(e->stage_ev[b]);
    if (e->stage[b]) (void)hipHostFree(e->stage[b]);
    e->stage_ev[b] = nullptr;
    e->stage[b] = nullptr;
  }
}

void free_staging_at_exit() {
  std::lock_guard<std::mutex> lk(g_staged_mu);
  for (pm_engine* e : *g_staged) {
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    free_staging(e);
  }
  g_staged->clear();
}

}  // namespace

void engine_release(pm_engine* e) {
  if (e->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  {
    std::lock_guard<std::mutex> lk(g_staged_mu);
    if (g_staged) g_staged->erase(std::remove(g_staged->begin(), g_staged->end(), e), g_staged->end());
    free_staging(e);
  }
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

namespace {

constexpr size_t kStageBytes = size_t(16) << 20;  // per pinned buffer
constexpr size_t kDirectBytes = size_t(1) << 20;  // below: straight through hipMemcpyAsync

// memcpy split over up to eight threads of at least 2 MB each (the staging
// copies are CPU-bound)
void par_memcpy(void* d, const void* s, size_t n) {
  constexpr size_t kPart = size_t(2) << 20;
  constexpr size_t kThreads = 8;
  const size_t parts = std::min<size_t>(kThreads, (n + kPart - 1) / kPart);
  if (parts <= 1) {
    std::memcpy(d, s, n);
    return;
  }
  const size_t per = (n + parts - 1) / parts;
  std::thread th[kThreads - 1];
  for (size_t p = 1; p < parts; ++p) {
    const size_t o = p * per;
    th[p - 1] = std::thread([=] { std::memcpy((char*)d + o, (const char*)s + o, std::min(per, n - o)); });
  }
  std::memcpy(d, s, std::min(per, n));
  for (size_t p = 1; p < parts; ++p) th[p - 1].join();
}

pm_status stage_init(pm_engine* e) {
  if (e->stage[0] && e->stage[1] && e->stage_ev[0] && e->stage_ev[1]) return PM_OK;
  std::lock_guard<std::mutex> lk(g_staged_mu);
  for (int b = 0; b < 2; ++b) {
    if (!e->stage[b]) PM_HIP(hipHostMalloc((void**)&e->stage[b], kStageBytes, hipHostMallocDefault));
    if (!e->stage_ev[b]) PM_HIP(hipEventCreateWithFlags(&e->stage_ev[b], hipEventDisableTiming));
  }
  if (!g_staged) {
    g_staged = new std::vector<pm_engine*>();
    std::atexit(free_staging_at_exit);
  }
  if (std::find(g_staged->begin(), g_staged->end(), e) == g_staged->end()) g_staged->push_back(e);
  return PM_OK;
}

}  // namespace

pm_status copy_h2d(pm_engine* e, void* dst, const void* src, size_t bytes, hipStream_t st) {
  if (bytes < kDirectBytes) {
    PM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    return PM_OK;
  }
  std::lock_guard<std::mutex> lk(e->stage_mu);
  if (pm_status s = stage_init(e)) return s;
  int b = 0;
  for (size_t off = 0; off < bytes; off += kStageBytes, b ^= 1) {
    const size_t n = std::min(kStageBytes, bytes - off);
    PM_HIP(hipEventSynchronize(e->stage_ev[b]));  // its last DMA has read it
    par_memcpy(e->stage[b], (const char*)src + off, n);
    PM_HIP(hipMemcpyAsync((char*)dst + off, e->stage[b], n, hipMemcpyHostToDevice, st));
    PM_HIP(hipEventRecord(e->stage_ev[b], st));
  }
  return PM_OK;
}

pm_status copy_d2h(pm_engine* e, void* dst, const void* src, size_t bytes, hipStream_t st) {
  if (bytes < kDirectBytes) {
    PM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    PM_HIP(hipStreamSynchronize(st));
    return PM_OK;
  }
  std::lock_guard<std::mutex> lk(e->stage_mu);
  if (pm_status s = stage_init(e)) return s;
  const size_t chunks = (bytes + kStageBytes - 1) / kStageBytes;
  auto issue = [&](size_t k) -> pm_status {
    const int b = (int)(k & 1);
    const size_t off = k * kStageBytes, n = std::min(kStageBytes, bytes - off);
    PM_HIP(hipStreamWaitEvent(st, e->stage_ev[b], 0));  // after its last use on any stream
    PM_HIP(hipMemcpyAsync(e->stage[b], (const char*)src + off, n, hipMemcpyDeviceToHost, st));
    PM_HIP(hipEventRecord(e->stage_ev[b], st));
    return PM_OK;
  };
  for (size_t k = 0; k < std::min<size_t>(2, chunks); ++k)
    if (pm_status s = issue(k)) return s;
  for (size_t k = 0; k < chunks; ++k) {
    const int b = (int)(k & 1);
    const size_t off = k * kStageBytes, n = std::min(kStageBytes, bytes - off);
    PM_HIP(hipEventSynchronize(e->stage_ev[b]));
    par_memcpy((char*)dst + off, e->stage[b], n);
    if (k + 2 < chunks)
      if (pm_status s = issue(k + 2)) return s;
  }
  return PM_OK;
}

}  // namespace pmh

PM_API void pm_engine_destroy(pm_engine* e) {
  if (e) pmh::engine_release(e);
}

PM_API pm_status pm_engine_set_counting(pm_engine* e, int32_t enable) {
  if (!e) return fail(PM_ERR_INVALID_ARG, "null engine");
  e->counting = enable ? 1 : 0;
  return PM_OK;
}

// ---------------------------------------------------------------------------
// problem

namespace pmh {

// InitTransforms (patch_match_cuda.cu:493-580) with ComputeRelativePose,
// ComposeProjectionMatrix and ComposeInverseProjectionMatrix
// (utils.cc:14-52), in FP64, rounded once to FP32.  inv_P is the closed
// form [R^T K^-1 | -R^T T] of the inverse of [K R | K T; 0 0 0 1].
void pack_poses(const pm_view_desc* view, float* poses, float k8[8]) {
  const pm_camera* rc = &view->ref_cam;
  const float fx = rc->K[0], cx = rc->K[1], fy = rc->K[2], cy = rc->K[3];
  k8[0] = fx;
  k8[1] = cx;
  k8[2] = fy;
  k8[3] = cy;
  k8[4] = 1.0f / fx;
  k8[5] = -cx / fx;
  k8[6] = 1.0f / fy;
  k8[7] = -cy / fy;
  for (int s = 0; s < view->num_src; ++s) {
    const pm_camera* sc = &view->src_cams[s];
    double R[9], T[3];
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) {
        double acc = 0.0;
        for (int k = 0; k < 3; ++k) acc += (double)sc->R[3 * i + k] * (double)rc->R[3 * j + k];
        R[3 * i + j] = acc;
      }
    }
    for (int i = 0; i < 3; ++i) {
      double acc = 0.0;
      for (int k = 0; k < 3; ++k) acc += R[3 * i + k] * (double)rc->T[k];
      T[i] = (double)sc->T[i] - acc;
    }
    const double sfx = sc->K[0], scx = sc->K[1], sfy = sc->K[2], scy = sc->K[3];
    double P[12];
    for (int j = 0; j < 4; ++j) {
      const double r0 = j < 3 ? R[j] : T[0];
      const double r1 = j < 3 ? R[3 + j] : T[1];
      const double r2 = j < 3 ? R[6 + j] : T[2];
      P[j] = sfx * r0 + scx * r2;
      P[4 + j] = sfy * r1 + scy * r2;
      P[8 + j] = r2;
    }
    const double Kinv[9] = {1.0 / sfx, 0.0, -scx / sfx, 0.0, 1.0 / sfy, -scy / sfy, 0.0, 0.0, 1.0};
    double iP[12];
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) {
        double acc = 0.0;
        for (int k = 0; k < 3; ++k) acc += R[3 * k + i] * Kinv[3 * k + j];
        iP[4 * i + j] = acc;
      }
      double acc = 0.0;
      for (int k = 0; k < 3; ++k) acc += R[3 * k + i] * T[k];
      iP[4 * i + 3] = -acc;
    }
    float* o = poses + (size_t)s * PM_POSE_STRIDE;
    std::memset(o, 0, sizeof(float) * PM_POSE_STRIDE);
    o[0] = sc->K[0];
    o[1] = sc->K[1];
    o[2] = sc->K[2];
    o[3] = sc->K[3];
    for (int i = 0; i < 9; ++i) o[4 + i] = (float)R[i];
    for (int i = 0; i < 3; ++i) o[13 + i] = (float)T[i];
    for (int i = 0; i < 12; ++i) o[16 + i] = (float)P[i];
    for (int i = 0; i < 12; ++i) o[28 + i] = (float)iP[i];
    o[40] = (float)view->src[s].width;
    o[41] = (float)view->src[s].height;
  }
}

// BilateralWeightComputer (utils.h:72-92) as spatial x colour tables, FP64.
void weight_tables(const pm_options* o, std::vector<float>& sp, std::vector<float>& ct, int* nwin) {
  const int R = o->window_radius, S = o->window_step;
  int n = 0;
  for (int d = -R; d <= R; d += S) ++n;
  *nwin = n;
  sp.resize((size_t)n * n);
  const double ss = (double)o->sigma_spatial;
  const double sn = 1.0 / (2.0 * ss * ss);
  const double sc = (double)o->sigma_color;
  const double cn = 1.0 / (2.0 * sc * sc);
  int k = 0;
  for (int dr = -R; dr <= R; dr += S)
    for (int dc = -R; dc <= R; dc += S) sp[k++] = (float)std::exp(-(double)(dr * dr + dc * dc) * sn);
  ct.resize(256);
  for (int d = 0; d < 256; ++d) {
    const double x = (double)d / 255.0;
    ct[d] = (float)std::exp(-x * x * cn);
  }
}

bool finite_camera(const pm_camera& c) {
  for (float k : c.K)
    if (!std::isfinite(k)) return false;
  for (float r : c.R)
    if (!std::isfinite(r)) return false;
  for (float t : c.T)
    if (!std::isfinite(t)) return false;
  return c.K[0] != 0.0f && c.K[2] != 0.0f;
}

pm_status check_view(const pm_options* o, const pm_view_desc* v) {
  if (!v) return fail(PM_ERR_INVALID_ARG, "null view");
  if (v->num_src < 1) return fail(PM_ERR_INVALID_ARG, "view needs at least one source image");
  if (v->num_src > PM_MAX_NUM_SRC) return fail(PM_ERR_TOO_MANY_SOURCES, "more than 32 source images");
  if (!v->ref.data || v->ref.width <= 0 || v->ref.height <= 0 || v->ref.stride < v->ref.width)
    return fail(PM_ERR_INVALID_ARG, "invalid reference image");
  if ((int64_t)v->ref.width * v->ref.height > (int64_t)1 << 28)
    return fail(PM_ERR_INVALID_ARG, "reference image too large");
  if (!v->src || !v->src_cams) return fail(PM_ERR_INVALID_ARG, "null source arrays");
  if (!finite_camera(v->ref_cam)) return fail(PM_ERR_INVALID_ARG, "non-finite or degenerate reference camera");
  for (int i = 0; i < v->num_src; ++i) {
    const pm_image& s = v->src[i];
    if (!s.data || s.width <= 0 || s.height <= 0 || s.stride < s.width)
      return fail(PM_ERR_INVALID_ARG, "invalid source image " + std::to_string(i));
    if ((int64_t)s.width * s.height > (int64_t)1 << 28) return fail(PM_ERR_INVALID_ARG, "source image too large");
    if (!finite_camera(v->src_cams[i])) return fail(PM_ERR_INVALID_ARG, "non-finite source camera " + std::to_string(i));
  }
  if (!(v->depth_min > 0.0f) || !(v->depth_max >= v->depth_min) || !std::isfinite(v->depth_max))
    return fail(PM_ERR_INVALID_ARG, "depth range must satisfy 0 < depth_min <= depth_max");
  if (o->geom_consistency) {
    if (!v->src_depth_maps) return fail(PM_ERR_INVALID_ARG, "geom_consistency needs src_depth_maps");
    for (int i = 0; i < v->num_src; ++i)
      if (!v->src_depth_maps[i]) return fail(PM_ERR_INVALID_ARG, "missing source depth map " + std::to_string(i));
  }
  if (!o->init_depth_random && (!v->init_depth || !v->init_normal))
    return fail(PM_ERR_INVALID_ARG, "init_depth_random = 0 needs init_depth and init_normal");
  if (v->ref_mask.data && (v->ref_mask.width != v->ref.width || v->ref_mask.height != v->ref.height ||
                           v->ref_mask.stride < v->ref_mask.width))
    return fail(PM_ERR_INVALID_ARG, "mask map must have the reference image's size");
  return PM_OK;
}

}  // namespace pmh

namespace pmh {

size_t src_quad_bytes(const pm_image& im, int storage) {
  const size_t n = (size_t)(im.width + 2 * pm::kQuadPad) * (im.height + 2 * pm::kQuadPad);
  return n * (storage == PM_STORAGE_U8 ? sizeof(uint32_t) : sizeof(uint2));
}

size_t job_block_bytes(const pm_options* o, const pm_view_desc* v, const bool* src_cached) {
  const size_t plane = (size_t)v->ref.width * v->ref.height;
  const int N = v->num_src;
  int nwin = 0;
  for (int d = -o->window_radius; d <= o->window_radius; d += o->window_step) ++nwin;
  size_t b = 0;
  auto add = [&](size_t bytes) { b += Arena::round(bytes > 0 ? bytes : 1); };
  add(plane);                                         // ref
  add((size_t)nwin * nwin * sizeof(float));           // spatial weights
  add(256 * sizeof(float));                           // colour weights
  add((size_t)N * PM_POSE_STRIDE * sizeof(float));    // poses
  for (int i = 0; i < N; ++i)
    if (!src_cached || !src_cached[i]) add(src_quad_bytes(v->src[i], o->image_storage));
  if (o->geom_consistency)
    for (int i = 0; i < N; ++i)
      add((size_t)(v->src[i].width + 2 * pm::kQuadPad) * (v->src[i].height + 2 * pm::kQuadPad) * sizeof(float4));
  add(plane * sizeof(float4));                        // moments
  add(plane);                                         // gradient gate
  add(plane * sizeof(float));                         // depth
  add(3 * plane * sizeof(float));                     // normal
  add(plane * sizeof(float));                         // conf
  add((size_t)N * plane * sizeof(float));             // cost
  add(plane * sizeof(uint32_t));                      // view_sel
  add(6 * sizeof(unsigned long long));                // counters
  if (o->geom_consistency && o->est_curvature) add(2 * plane * sizeof(uint16_t));  // curvature map
  if (v->ref_mask.data) add(plane);                   // mask
  if (o->median_filter) add(plane * sizeof(float));   // relative-depth gradient
  if (!o->init_depth_random) {
    add(plane * sizeof(float));
    add(3 * plane * sizeof(float));
  }
  add(sizeof(pm::ViewDev));
  return b;
}

// An 8-bit host image into a dense W x H device buffer.  A contiguous image
// goes as one copy: from pageable memory a 2-D copy moves row by row, which
// for 1080 rows costs about 10 ms per image.
pm_status upload_image(uint8_t* dst, const uint8_t* src, int W, int H, int stride, hipStream_t s) {
  if (stride == W) PM_HIP(hipMemcpyAsync(dst, src, (size_t)W * H, hipMemcpyHostToDevice, s));
  else PM_HIP(hipMemcpy2DAsync(dst, W, src, stride, W, H, hipMemcpyHostToDevice, s));
  return PM_OK;
}

pm_status upload_quads(const pm_image& im, int storage, void* dst, uint8_t* staging, hipStream_t s) {
  pm_status rs = upload_image(staging, im.data, im.width, im.height, im.stride, s);
  if (rs != PM_OK) return rs;
  if (storage == PM_STORAGE_U8)
    PM_HIP(pm::launch_build_quads_u8(staging, im.width, im.height, im.width, (uint32_t*)dst, s));
  else
    PM_HIP(pm::launch_build_quads_f16(staging, im.width, im.height, im.width, (uint2*)dst, s));
  return PM_OK;
}

pm_status create_problem(pm_engine* e, const pm_options* o, const pm_view_desc* v, const CreateEnv& env,
                         pm_problem** out) {
  if (!e || !o || !out) return fail(PM_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  pm_status st = pm_options_check(o);
  if (st != PM_OK) return st;
  st = check_view(o, v);
  if (st != PM_OK) return st;
  PM_HIP(hipSetDevice(e->device));
  pm_problem* p = new (std::nothrow) pm_problem();
  if (!p) return fail(PM_ERR_OUT_OF_MEMORY, "host allocation");
  hipStream_t s = env.stream ? env.stream : e->stream;
  // on an error path uploads may still read the staging buffer and write the
  // problem's buffers: wait for the stream before handing them back
  struct Guard {
    pm_problem*& p;
    pm_engine* e;
    hipStream_t s;
    uint8_t* staging = nullptr;  // temporary staging (engine problems)
    size_t staging_bytes = 0;
    bool keep = false;
    ~Guard() {
      if (staging || !keep) (void)hipStreamSynchronize(s);
      if (staging) e->cache.put(staging, staging_bytes);
      if (!keep) delete p;
    }
  } guard{p, e, s};
  p->engine = e;
  pmh::engine_retain(e);
  p->opt = *o;
  p->W = v->ref.width;
  p->H = v->ref.height;
  p->N = v->num_src;
  p->view_id = v->view_id;
  p->ref_cam = v->ref_cam;
  p->depth_min = v->depth_min;
  p->depth_max = v->depth_max;
  if (env.arena) {
    p->arena = env.arena;
    p->block_off = env.block_off;
    p->block_bytes = env.block_bytes;
  }
  const int W = p->W, H = p->H, N = p->N;
  const size_t plane = (size_t)W * H;
  pm::ViewDev& h = p->host;
  std::memset(&h, 0, sizeof(h));

  // reference image and tables
  uint8_t* ref = nullptr;
  if ((st = p->alloc(&ref, plane)) != PM_OK) return st;
  if ((st = upload_image(ref, v->ref.data, W, H, v->ref.stride, s)) != PM_OK) return st;
  std::vector<float> sp, ct;
  int nwin = 0;
  weight_tables(o, sp, ct, &nwin);
  float *d_sp = nullptr, *d_ct = nullptr;
  if ((st = p->alloc(&d_sp, sp.size())) != PM_OK) return st;
  if ((st = p->alloc(&d_ct, ct.size())) != PM_OK) return st;
  PM_HIP(hipMemcpyAsync(d_sp, sp.data(), sp.size() * sizeof(float), hipMemcpyHostToDevice, s));
  PM_HIP(hipMemcpyAsync(d_ct, ct.data(), ct.size() * sizeof(float), hipMemcpyHostToDevice, s));
  p->poses.assign((size_t)N * PM_POSE_STRIDE, 0.0f);
  pack_poses(v, p->poses.data(), p->refk8);
  float* d_poses = nullptr;
  if ((st = p->alloc(&d_poses, p->poses.size())) != PM_OK) return st;
  PM_HIP(hipMemcpyAsync(d_poses, p->poses.data(), p->poses.size() * sizeof(float), hipMemcpyHostToDevice, s));

  // staging for the uploads below: the caller's, or a temporary one
  size_t max_src = 0;
  for (int i = 0; i < N; ++i) {
    const size_t n = (size_t)v->src[i].width * v->src[i].height;
    max_src = std::max(max_src, o->geom_consistency ? n * sizeof(float) : n);
  }
  uint8_t* staging = env.staging;
  if (!staging) {
    void* q = nullptr;
    PM_HIP(e->cache.get(max_src, &q));
    guard.staging = (uint8_t*)q;
    guard.staging_bytes = max_src;
    staging = guard.staging;
  } else if (env.staging_bytes < max_src) {
    return fail(PM_ERR_INTERNAL, "staging buffer too small");
  }

  // source images in the padded quad layout (stream order keeps the staging
  // buffer's reuse safe: each copy follows the previous quad kernel)
  for (int i = 0; i < N; ++i) {
    const pm_image& si = v->src[i];
    pm::SrcDev& sd = h.src[i];
    sd.width = si.width;
    sd.height = si.height;
    sd.pitch = si.width + 2 * pm::kQuadPad;
    if (env.src_tex && env.src_tex[i]) {
      sd.tex = env.src_tex[i];
      continue;
    }
    void* q = nullptr;
    if (o->image_storage == PM_STORAGE_U8) {
      uint32_t* t = nullptr;
      if ((st = p->alloc(&t, src_quad_bytes(si, o->image_storage) / sizeof(uint32_t))) != PM_OK) return st;
      q = t;
    } else {
      uint2* t = nullptr;
      if ((st = p->alloc(&t, src_quad_bytes(si, o->image_storage) / sizeof(uint2))) != PM_OK) return st;
      q = t;
    }
    if ((st = upload_quads(si, o->image_storage, q, staging, s)) != PM_OK) return st;
    sd.tex = q;
  }

  // source depth maps (geometric pass)
  if (o->geom_consistency) {
    for (int i = 0; i < N; ++i) {
      const pm_image& si = v->src[i];
      const size_t n = (size_t)si.width * si.height;
      float4* dq = nullptr;
      const size_t pn = (size_t)(si.width + 2 * pm::kQuadPad) * (si.height + 2 * pm::kQuadPad);
      if ((st = p->alloc(&dq, pn)) != PM_OK) return st;
      PM_HIP(hipMemcpyAsync(staging, v->src_depth_maps[i], n * sizeof(float), hipMemcpyHostToDevice, s));
      PM_HIP(pm::launch_build_depth_quads((const float*)staging, si.width, si.height, dq, s));
      h.src[i].depth_q = dq;
    }
  }

  // state, moments
  float4* refm = nullptr;
  uint8_t* grad = nullptr;
  if ((st = p->alloc(&refm, plane)) != PM_OK) return st;
  if ((st = p->alloc(&grad, plane)) != PM_OK) return st;
  if ((st = p->alloc(&h.depth, plane)) != PM_OK) return st;
  if ((st = p->alloc(&h.normal, 3 * plane)) != PM_OK) return st;
  if ((st = p->alloc(&h.conf, plane)) != PM_OK) return st;
  if ((st = p->alloc(&h.cost, (size_t)N * plane)) != PM_OK) return st;
  if ((st = p->alloc(&h.view_sel, plane)) != PM_OK) return st;
  if ((st = p->alloc(&p->counters, 6)) != PM_OK) return st;
  if (o->geom_consistency && o->est_curvature) {
    if ((st = p->alloc(&h.curv, 2 * plane)) != PM_OK) return st;
    PM_HIP(hipMemsetAsync(h.curv, 0, 2 * plane * sizeof(uint16_t), s));
  }
  if (o->median_filter && (st = p->alloc(&h.geom_grad, plane)) != PM_OK) return st;
  if (v->ref_mask.data) {
    uint8_t* m = nullptr;
    if ((st = p->alloc(&m, plane)) != PM_OK) return st;
    if ((st = upload_image(m, v->ref_mask.data, W, H, v->ref_mask.stride, s)) != PM_OK) return st;
    h.mask = m;
  }
  if (!o->init_depth_random) {
    if ((st = p->alloc(&p->init_depth, plane)) != PM_OK) return st;
    if ((st = p->alloc(&p->init_normal, 3 * plane)) != PM_OK) return st;
    PM_HIP(hipMemcpyAsync(p->init_depth, v->init_depth, plane * sizeof(float), hipMemcpyHostToDevice, s));
    PM_HIP(hipMemcpyAsync(p->init_normal, v->init_normal, 3 * plane * sizeof(float), hipMemcpyHostToDevice, s));
  }

  h.W = W;
  h.H = H;
  h.N = N;
  h.nwin = nwin;
  h.radius = o->window_radius;
  h.step = o->window_step;
  h.view_id = (int32_t)v->view_id;
  h.geom = o->geom_consistency ? 1 : 0;
  std::memcpy(h.refK, p->refk8, 4 * sizeof(float));
  std::memcpy(h.refInvK, p->refk8 + 4, 4 * sizeof(float));
  h.depth_min = v->depth_min;
  h.depth_max = v->depth_max;
  h.depth_max_eff = (float)(1.2 * (double)v->depth_max);
  h.geom_regularizer = o->geom_consistency_regularizer;
  h.geom_max_cost = o->geom_consistency_max_cost;
  h.random_depth_ratio = o->random_depth_ratio;
  h.angle1 = std::fmin(o->random_angle1_range, PM_PI_F);
  h.angle2 = std::fmin(o->random_angle2_range, PM_HALF_PI_F);
  h.max_ncc_matching_cost = o->max_ncc_matching_cost;
  h.inv_beta2 = 0.5f / (o->beta * o->beta);
  h.num_good = o->num_good_hypothesis;
  h.num_bad = o->num_bad_hypothesis;
  h.num_samples = o->num_samples;
  h.random_optimization = o->random_optimization ? 1 : 0;
  h.thk = o->thk;
  h.init_samples = o->init_samples;
  h.init_random = o->init_depth_random ? 1 : 0;
  h.compat_lo_geo_bug = o->compat_lo_geo_bug ? 1 : 0;
  h.accum_mode = o->accum_mode;
  h.conf_threshold = o->conf_threshold;
  h.depth_diff_threshold = o->depth_diff_threshold;
  h.seed = o->seed;
  h.ref = ref;
  h.refm = refm;
  h.grad_ok = grad;
  h.poses = d_poses;
  h.sp = d_sp;
  h.ct = d_ct;
  h.counters = nullptr;
  if ((st = p->alloc(&p->dev, 1)) != PM_OK) return st;
  if (p->arena && p->block_used != p->block_bytes) return fail(PM_ERR_INTERNAL, "job memory block size mismatch");
  PM_HIP(hipMemcpyAsync(p->dev, &h, sizeof(h), hipMemcpyHostToDevice, s));
  PM_HIP(pm::launch_ref_moments(p->dev, W, H, o->accum_mode, refm, grad, s));
  // host sources of the copies above are pageable locals and caller memory
  PM_HIP(hipStreamSynchronize(s));
  guard.keep = true;
  *out = p;
  return PM_OK;
}

}  // namespace pmh

PM_API pm_status pm_problem_create(pm_engine* e, const pm_options* o, const pm_view_desc* v, pm_problem** out) {
  return create_problem(e, o, v, CreateEnv{}, out);
}

namespace pmh {

pm_status EventTimer::rec(hipStream_t s) {
  hipEvent_t e;
  PM_HIP(hipEventCreate(&e));
  ev.push_back(e);
  PM_HIP(hipEventRecord(e, s));
  return PM_OK;
}

float EventTimer::between(size_t a, size_t b) const {
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, ev[a], ev[b]) != hipSuccess) return 0.0f;
  return ms;
}

EventTimer::~EventTimer() {
  for (hipEvent_t e : ev) (void)hipEventDestroy(e);
}

// Good-matching threshold of iteration `iter` (patch_match_cuda.cu:853).
float th_mc_of(const pm_options& o, int iter) {
  return (float)((double)o.init_ncc_matching_cost * std::exp((double)(-(float)(iter * iter) / o.alpha)));
}

std::vector<std::vector<PhaseRec>> view_ticks(const pm_problem& p) {
  const pm_options& o = p.opt;
  std::vector<std::vector<PhaseRec>> t;
  const uint32_t goff = o.geom_consistency ? PM_PASS_GEOM_OFFSET : 0u;
  // both colours of the init in one tick (init touches only its own pixel)
  t.push_back({{pm::kPhaseInit, 0, PM_PASS_INIT + goff, 0.0f}, {pm::kPhaseInit, 1, PM_PASS_INIT + goff, 0.0f}});
  for (int iter = 0; iter < o.num_iterations; ++iter) {
    const float th_mc = th_mc_of(o, iter);
    for (int c = 0; c < 2; ++c) t.push_back({{pm::kPhaseSweep, c, PM_PASS_SWEEP(iter, c) + goff, th_mc}});
    if (o.local_optimization && iter > 2)
      for (int c = 0; c < 2; ++c) t.push_back({{pm::kPhaseLocal, c, PM_PASS_LOCAL(iter, c) + goff, 0.0f}});
  }
  // after PropagationAndOptim: curvature (:985-989), median filter
  // (:991-1008), FilterDepthMap (:1015-1037), mask (:1040-1052); the
  // curvature, gradient and mask passes write only their own pixel's outputs
  // and read nothing another entry of the tick writes, so both colours share
  // a tick
  if (p.host.curv) t.push_back({{pm::kPhaseCurvature, 0, 0u, 0.0f}, {pm::kPhaseCurvature, 1, 0u, 0.0f}});
  if (p.host.geom_grad) {  // median filter (:991-1008): gradient of all pixels, then black, then red
    t.push_back({{pm::kPhaseGeomGrad, 0, 0u, 0.0f}, {pm::kPhaseGeomGrad, 1, 0u, 0.0f}});
    for (int c = 0; c < 2; ++c) t.push_back({{pm::kPhaseMedian, c, 0u, 0.0f}});
  }
  if (o.filter && o.conf_filter)
    for (int c = 0; c < 2; ++c) t.push_back({{pm::kPhaseFilter, c, 0u, 0.0f}});
  if (p.host.mask) t.push_back({{pm::kPhaseMask, 0, 0u, 0.0f}, {pm::kPhaseMask, 1, 0u, 0.0f}});
  return t;
}

}  // namespace pmh

pm_problem::~pm_problem() {
  delete schedule;
  if (!engine) return;
  for (const auto& a : allocs) engine->cache.put(a.first, a.second);
  pmh::engine_release(engine);
}

PM_API void pm_problem_destroy(pm_problem* p) {
  if (!p) return;
  (void)hipSetDevice(p->engine->device);
  (void)hipStreamSynchronize(p->engine->stream);
  delete p;
}

namespace pmh {

pm::LaunchCfg cfg_of(const pm_options& o) { return pm::LaunchCfg{o.image_storage, o.rcp_mode, o.window_radius, o.window_step}; }

int nbucket(int N) { return N <= 8 ? 0 : (N <= 16 ? 1 : 2); }

TuneKey tune_key(int kind, const pm_options& o, int N) {
  return TuneKey{kind, o.window_radius, o.window_step, o.image_storage, o.rcp_mode, nbucket(N)};
}

pm_tuning shape_of(pm_engine* e, int kind, const pm_options& o, int N) {
  std::lock_guard<std::mutex> lk(e->tune_mu);
  auto it = e->tuning.find(tune_key(kind, o, N));
  if (it != e->tuning.end()) return it->second;
  return pm_tuning{pm::kTileLog2WDefault, 0};
}

pm_status plan_launches(const std::vector<std::pair<pm_problem*, const std::vector<PhaseRec>*>>& phases,
                        TickPlan& plan, const pm_tuning* force) {
  struct Group {
    pm::PhaseKind kind;
    pm::LaunchCfg cfg;
    pm_tuning shape;
    std::vector<pm::LaunchEntry> e;
    int tiles = 0;
    int nmax = 1;  // most sources of a view in the launch (workspace size)
  };
  std::vector<Group> groups;
  for (const auto& ph : phases) {
    pm_problem* p = ph.first;
    for (const PhaseRec& r : *ph.second) {
      const pm::LaunchCfg cfg = pm::plain_phase(r.kind) ? pm::LaunchCfg{0, 0, 0, 0} : cfg_of(p->opt);
      const pm_tuning shape = force ? *force : shape_of(p->engine, r.kind, p->opt, p->N);
      Group* g = nullptr;
      for (Group& gg : groups)
        if (gg.kind == r.kind && gg.cfg == cfg && gg.shape.tile_lw == shape.tile_lw &&
            gg.shape.blocks_per_cu == shape.blocks_per_cu)
          g = &gg;
      if (!g) {
        groups.push_back(Group{r.kind, cfg, shape, {}, 0, 1});
        g = &groups.back();
      }
      pm::LaunchEntry en{};
      en.v = p->dev;
      en.tile_start = g->tiles;
      en.color = r.color;
      en.pass = r.pass;
      en.th_mc = r.th_mc;
      g->e.push_back(en);
      const long long tiles = (long long)g->tiles + pm::view_tiles(p->W, p->H, g->shape.tile_lw);
      if (tiles > INT32_MAX) return fail(PM_ERR_INVALID_ARG, "too many tiles in one launch");
      g->tiles = (int)tiles;
      g->nmax = std::max(g->nmax, p->N);
    }
  }
  for (Group& g : groups) {
    const int grid = pm::phase_grid(g.kind, g.cfg, g.tiles, g.shape.blocks_per_cu);
    if (grid <= 0) return fail(PM_ERR_DEVICE, "no launch configuration for a phase kernel");
    const size_t off = (plan.tables.size() + 15) & ~(size_t)15;
    plan.tables.resize(off + pm::launch_table_bytes((int)g.e.size()));
    pm::LaunchTable* lt = (pm::LaunchTable*)(plan.tables.data() + off);
    lt->n = (int32_t)g.e.size();
    lt->total_tiles = g.tiles;
    lt->ws_slots = grid * pm::kSweepThreads;
    lt->tile_lw = g.shape.tile_lw;
    lt->ws = nullptr;
    lt->queue = nullptr;
    std::memcpy(lt->e, g.e.data(), g.e.size() * sizeof(pm::LaunchEntry));
    plan.launches.push_back(pm_schedule::Launch{g.kind, g.cfg, off, g.tiles, grid});
    plan.ws_floats = std::max(plan.ws_floats, (size_t)lt->ws_slots * pm::ws_floats_per_slot(g.nmax));
  }
  return PM_OK;
}

// Build the launches of `ticks` over `views`; `force` overrides the shapes
// (autotuning).
pm_status build_schedule(const std::vector<pm_problem*>& views, const Ticks& ticks, pm_schedule& sch, hipStream_t s,
                         const pm_tuning* force) {
  size_t T = 0;
  for (const auto& tv : ticks) T = std::max(T, tv.size());
  TickPlan plan;
  for (size_t t = 0; t < T; ++t) {
    std::vector<std::pair<pm_problem*, const std::vector<PhaseRec>*>> phases;
    for (size_t vi = 0; vi < views.size(); ++vi)
      if (t < ticks[vi].size()) phases.emplace_back(views[vi], &ticks[vi][t]);
    pm_status rs = plan_launches(phases, plan, force);
    if (rs != PM_OK) return rs;
  }
  sch.tables = std::move(plan.tables);
  sch.launches = std::move(plan.launches);
  sch.ws_floats = plan.ws_floats;
  sch.release();
  auto get = [&](void** p, size_t bytes) -> hipError_t {
    if (sch.cache) return sch.cache->get(bytes, p);
    return hipMalloc(p, bytes);
  };
  sch.ws_bytes = std::max<size_t>(sch.ws_floats, 1) * sizeof(float);
  sch.queue_bytes = std::max<size_t>(8 * sch.launches.size(), 1) * sizeof(uint32_t);
  sch.dev_bytes = std::max<size_t>(sch.tables.size(), 16);
  PM_HIP(get((void**)&sch.ws, sch.ws_bytes));
  PM_HIP(get((void**)&sch.queues, sch.queue_bytes));
  for (size_t k = 0; k < sch.launches.size(); ++k) {
    pm::LaunchTable* lt = (pm::LaunchTable*)(sch.tables.data() + sch.launches[k].offset);
    lt->ws = sch.ws;
    lt->queue = sch.queues + 8 * k;
  }
  PM_HIP(get(&sch.dev, sch.dev_bytes));
  PM_HIP(hipMemcpyAsync(sch.dev, sch.tables.data(), sch.tables.size(), hipMemcpyHostToDevice, s));
  return PM_OK;
}

Ticks all_ticks(const std::vector<pm_problem*>& views) {
  Ticks t;
  for (pm_problem* p : views) t.push_back(view_ticks(*p));
  return t;
}

// Per-view work before the launches: counting on/off and the initial state
// of views created with init_depth_random = 0.
pm_status prepare_views(const std::vector<pm_problem*>& views, bool counting, hipStream_t s, bool allow_memset) {
  for (pm_problem* p : views) {
    unsigned long long* cnt = counting ? p->counters : nullptr;
    if (cnt != p->host.counters) {
      p->host.counters = cnt;
      PM_HIP(hipMemcpyAsync(p->dev, &p->host, sizeof(p->host), hipMemcpyHostToDevice, s));
    }
    if (cnt && allow_memset) PM_HIP(hipMemsetAsync(cnt, 0, 6 * sizeof(unsigned long long), s));
    if (!p->opt.init_depth_random) {
      const size_t plane = (size_t)p->W * p->H;
      PM_HIP(hipMemcpyAsync(p->host.depth, p->init_depth, plane * sizeof(float), hipMemcpyDeviceToDevice, s));
      PM_HIP(hipMemcpyAsync(p->host.normal, p->init_normal, 3 * plane * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
  }
  return PM_OK;
}

// Issue every launch of the schedule on stream s, bracketing each with
// events when t is given.
pm_status launch_all(const pm_schedule& sch, hipStream_t s, EventTimer* t, pm_run_stats* st) {
  pm_status rs;
  PM_HIP(pm::launch_zero_u32(sch.queues, (int)(8 * sch.launches.size()), s));
  for (const pm_schedule::Launch& l : sch.launches) {
    if (t && (rs = t->rec(s)) != PM_OK) return rs;
    PM_HIP(pm::launch_phase(l.kind, l.cfg, (const pm::LaunchTable*)((const uint8_t*)sch.dev + l.offset), l.grid, s));
    if (t && (rs = t->rec(s)) != PM_OK) return rs;
    if (st) {
      if (l.kind == pm::kPhaseSweep) st->sweep_launches++;
      if (l.kind == pm::kPhaseLocal) st->local_opt_launches++;
    }
  }
  return PM_OK;
}

void collect_times(const pm_schedule& sch, const EventTimer& t, size_t first, pm_run_stats* st) {
  size_t k = first;
  for (const pm_schedule::Launch& l : sch.launches) {
    const double ms = t.between(k, k + 1);
    if (l.kind == pm::kPhaseSweep) st->sweep_ms += ms;
    if (l.kind == pm::kPhaseLocal) st->local_opt_ms += ms;
    if (l.kind == pm::kPhaseInit) st->init_ms += ms;
    if (pm::plain_phase(l.kind)) st->post_ms += ms;
    k += 2;
  }
}

pm_status read_counters(const std::vector<pm_problem*>& views, pm_run_stats* st) {
  for (pm_problem* p : views) {
    if (!p->host.counters) continue;
    unsigned long long c[6];
    PM_HIP(hipMemcpy(c, p->counters, sizeof(c), hipMemcpyDeviceToHost));
    st->ncc_evaluations += c[0];
    st->ncc_samples += c[1];
    st->geo_evaluations += c[2];
    st->sweep_ncc_evaluations += c[3];
    st->sweep_ncc_samples += c[4];
    st->sweep_geo_evaluations += c[5];
  }
  return PM_OK;
}

// Eager run of a set of views with per-launch hipEvent timings.
pm_status run_views(pm_engine* e, const std::vector<pm_problem*>& views, pm_schedule& sch, pm_run_stats* st) {
  PM_HIP(hipSetDevice(e->device));
  hipStream_t s = e->stream;
  pm_status rs;
  if (!sch.dev && (rs = build_schedule(views, all_ticks(views), sch, s)) != PM_OK) return rs;
  if ((rs = prepare_views(views, e->counting != 0, s, true)) != PM_OK) return rs;
  EventTimer t;
  std::memset(st, 0, sizeof(*st));
  if ((rs = t.rec(s)) != PM_OK) return rs;
  if ((rs = launch_all(sch, s, &t, st)) != PM_OK) return rs;
  if ((rs = t.rec(s)) != PM_OK) return rs;
  PM_HIP(hipStreamSynchronize(s));
  st->total_ms = t.between(0, t.ev.size() - 1);
  collect_times(sch, t, 1, st);
  return read_counters(views, st);
}

}  // namespace pmh

// ---------------------------------------------------------------------------
// launch shapes and the autotuner

namespace pmh {

pm_status check_phase(int32_t phase) {
  if (phase < PM_PHASE_INIT || phase > PM_PHASE_MEDIAN) return fail(PM_ERR_INVALID_ARG, "unknown phase");
  return PM_OK;
}

pm_status check_tuning(const pm_tuning* t) {
  if (!t) return fail(PM_ERR_INVALID_ARG, "null tuning");
  if (t->tile_lw < pm::kTileLog2WMin || t->tile_lw > pm::kTileLog2WMax)
    return fail(PM_ERR_INVALID_ARG, "tile_lw must be 4, 5 or 6");
  if (t->blocks_per_cu < 0 || t->blocks_per_cu > 64) return fail(PM_ERR_INVALID_ARG, "blocks_per_cu must be 0..64");
  return PM_OK;
}

// Device time of one run of a schedule (queue reset + its launches).
pm_status time_schedule(const pm_schedule& sch, hipStream_t s, double* ms) {
  EventTimer t;
  pm_status rs;
  if ((rs = t.rec(s)) != PM_OK) return rs;
  if ((rs = launch_all(sch, s, nullptr, nullptr)) != PM_OK) return rs;
  if ((rs = t.rec(s)) != PM_OK) return rs;
  PM_HIP(hipStreamSynchronize(s));
  *ms = t.between(0, 1);
  return PM_OK;
}

std::vector<pm_tuning> tune_candidates(pm::PhaseKind kind, const pm_options& o) {
  const int occ = pm::phase_occupancy(kind, cfg_of(o));
  std::vector<pm_tuning> c;
  for (int lw = pm::kTileLog2WMin; lw <= pm::kTileLog2WMax; ++lw) {
    c.push_back(pm_tuning{lw, 0});
    for (int b = occ - 1; b >= 1 && b >= occ - 2; --b) c.push_back(pm_tuning{lw, b});
  }
  return c;
}

// Time every candidate shape on one phase of view p (two rounds, the
// minimum counts) and return the fastest.
pm_status tune_phase(pm_problem* p, const PhaseRec& r, pm_tuning* best, double* ms_default, double* ms_best,
                     int32_t* ncand) {
  hipStream_t s = p->engine->stream;
  const std::vector<pm_tuning> cands = tune_candidates(r.kind, p->opt);
  std::vector<double> t(cands.size(), 1e30);
  const std::vector<pm_problem*> one{p};
  Ticks tk(1);
  tk[0].push_back({r});
  pm_status rs;
  for (int round = 0; round < 2; ++round) {
    for (size_t k = 0; k < cands.size(); ++k) {
      pm_schedule sch;
      if ((rs = build_schedule(one, tk, sch, s, &cands[k])) != PM_OK) return rs;
      double ms = 0.0;
      if ((rs = time_schedule(sch, s, &ms)) != PM_OK) return rs;
      t[k] = std::min(t[k], ms);
    }
  }
  size_t kb = 0;
  *ms_default = 0.0;
  for (size_t k = 0; k < cands.size(); ++k) {
    if (t[k] < t[kb]) kb = k;
    if (cands[k].tile_lw == pm::kTileLog2WDefault && cands[k].blocks_per_cu == 0) *ms_default = t[k];
  }
  *best = cands[kb];
  *ms_best = t[kb];
  *ncand = (int32_t)cands.size();
  return PM_OK;
}

pm_status autotune(pm_problem* p, pm_tune_report* rep) {
  pm_engine* e = p->engine;
  PM_HIP(hipSetDevice(e->device));
  hipStream_t s = e->stream;
  const pm_options& o = p->opt;
  const uint32_t goff = o.geom_consistency ? PM_PASS_GEOM_OFFSET : 0u;
  pm_status rs;
  // a realistic state: init and two iterations of sweeps, default shapes
  {
    Ticks warm(1);
    warm[0] = {{{pm::kPhaseInit, 0, PM_PASS_INIT + goff, 0.0f}, {pm::kPhaseInit, 1, PM_PASS_INIT + goff, 0.0f}}};
    for (int it = 0; it < 2; ++it)
      for (int c = 0; c < 2; ++c) warm[0].push_back({{pm::kPhaseSweep, c, PM_PASS_SWEEP(it, c) + goff, th_mc_of(o, it)}});
    const std::vector<pm_problem*> one{p};
    if ((rs = prepare_views(one, false, s, true)) != PM_OK) return rs;
    pm_schedule sch;
    if ((rs = build_schedule(one, warm, sch, s, nullptr)) != PM_OK) return rs;
    if ((rs = launch_all(sch, s, nullptr, nullptr)) != PM_OK) return rs;
    PM_HIP(hipStreamSynchronize(s));
  }
  pm_tune_report r{};
  const PhaseRec sw{pm::kPhaseSweep, 0, PM_PASS_SWEEP(2, 0) + goff, th_mc_of(o, 2)};
  if ((rs = tune_phase(p, sw, &r.sweep, &r.sweep_ms_default, &r.sweep_ms_best, &r.candidates)) != PM_OK) return rs;
  r.local_opt = pm_tuning{pm::kTileLog2WDefault, 0};
  if (o.local_optimization) {
    const PhaseRec lo{pm::kPhaseLocal, 0, PM_PASS_LOCAL(3, 0) + goff, 0.0f};
    int32_t nc = 0;
    if ((rs = tune_phase(p, lo, &r.local_opt, &r.local_opt_ms_default, &r.local_opt_ms_best, &nc)) != PM_OK) return rs;
  }
  {
    std::lock_guard<std::mutex> lk(e->tune_mu);
    e->tuning[tune_key(pm::kPhaseSweep, o, p->N)] = r.sweep;
    if (o.local_optimization) e->tuning[tune_key(pm::kPhaseLocal, o, p->N)] = r.local_opt;
  }
  if (rep) *rep = r;
  return PM_OK;
}

// Autotune, on its largest view, every sweep class of `views` the engine
// holds no shape for.
pm_status tune_batch(const std::vector<pm_problem*>& views) {
  std::map<TuneKey, pm_problem*> largest;
  for (pm_problem* p : views) {
    const TuneKey k = tune_key(pm::kPhaseSweep, p->opt, p->N);
    {
      std::lock_guard<std::mutex> lk(p->engine->tune_mu);
      if (p->engine->tuning.count(k)) continue;
    }
    auto it = largest.find(k);
    if (it == largest.end() || (long long)p->W * p->H > (long long)it->second->W * it->second->H) largest[k] = p;
  }
  for (auto& kv : largest) {
    pm_status rs = autotune(kv.second, nullptr);
    if (rs != PM_OK) return rs;
  }
  return PM_OK;
}

std::string device_arch(int device) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return "unknown";
  std::string a = prop.gcnArchName;
  const size_t c = a.find(':');
  return c == std::string::npos ? a : a.substr(0, c);
}

}  // namespace pmh

PM_API pm_status pm_engine_set_tuning(pm_engine* e, int32_t phase, const pm_options* o, int32_t num_src,
                                      const pm_tuning* t) {
  if (!e || !o) return fail(PM_ERR_INVALID_ARG, "null argument");
  pm_status rs;
  if ((rs = check_phase(phase)) != PM_OK || (rs = check_tuning(t)) != PM_OK) return rs;
  std::lock_guard<std::mutex> lk(e->tune_mu);
  e->tuning[tune_key(phase, *o, num_src)] = *t;
  return PM_OK;
}

PM_API pm_status pm_engine_get_tuning(pm_engine* e, int32_t phase, const pm_options* o, int32_t num_src,
                                      pm_tuning* out) {
  if (!e || !o || !out) return fail(PM_ERR_INVALID_ARG, "null argument");
  pm_status rs;
  if ((rs = check_phase(phase)) != PM_OK) return rs;
  *out = shape_of(e, phase, *o, num_src);
  return PM_OK;
}

PM_API pm_status pm_problem_autotune(pm_problem* p, pm_tune_report* report) {
  if (!p) return fail(PM_ERR_INVALID_ARG, "null problem");
  return autotune(p, report);
}

PM_API pm_status pm_engine_save_tuning(pm_engine* e, const char* path) {
  if (!e || !path) return fail(PM_ERR_INVALID_ARG, "null argument");
  const std::string arch = device_arch(e->device);
  FILE* f = std::fopen(path, "w");
  if (!f) return fail(PM_ERR_INVALID_ARG, std::string("cannot write ") + path);
  std::lock_guard<std::mutex> lk(e->tune_mu);
  std::fprintf(f, "# arch phase radius step storage rcp nbucket tile_lw blocks_per_cu\n");
  for (const auto& kv : e->tuning) {
    const TuneKey& k = kv.first;
    std::fprintf(f, "%s %d %d %d %d %d %d %d %d\n", arch.c_str(), k.kind, k.radius, k.step, k.storage, k.rcp,
                 k.nbucket, kv.second.tile_lw, kv.second.blocks_per_cu);
  }
  const bool ok = std::fclose(f) == 0;
  return ok ? PM_OK : fail(PM_ERR_INVALID_ARG, std::string("cannot write ") + path);
}

PM_API pm_status pm_engine_load_tuning(pm_engine* e, const char* path) {
  if (!e || !path) return fail(PM_ERR_INVALID_ARG, "null argument");
  const std::string arch = device_arch(e->device);
  FILE* f = std::fopen(path, "r");
  if (!f) return fail(PM_ERR_INVALID_ARG, std::string("cannot read ") + path);
  char line[256];
  int bad = 0;
  std::map<TuneKey, pm_tuning> got;
  while (std::fgets(line, sizeof(line), f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    char a[64];
    TuneKey k{};
    pm_tuning t{};
    if (std::sscanf(line, "%63s %d %d %d %d %d %d %d %d", a, &k.kind, &k.radius, &k.step, &k.storage, &k.rcp,
                    &k.nbucket, &t.tile_lw, &t.blocks_per_cu) != 9 ||
        check_phase(k.kind) != PM_OK || check_tuning(&t) != PM_OK) {
      bad++;
      continue;
    }
    if (arch == a) got[k] = t;
  }
  std::fclose(f);
  if (bad) return fail(PM_ERR_INVALID_ARG, std::string("malformed lines in ") + path);
  std::lock_guard<std::mutex> lk(e->tune_mu);
  for (auto& kv : got) e->tuning[kv.first] = kv.second;
  return PM_OK;
}

// ---------------------------------------------------------------------------
// batch

struct pm_batch {
  pmh::EngineRef engine_ref;  // first member: released after the others
  pm_engine* engine = nullptr;
  std::vector<pm_problem*> views;
  pm_schedule schedule;
  pm_run_stats stats{};
  bool use_graph = false;
  bool autotune = false;
  bool tuned = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~pm_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (pm_problem* p : views) delete p;
  }
};

PM_API pm_status pm_batch_create(pm_engine* e, int32_t num_views, const pm_options* options, const pm_view_desc* views,
                                 pm_batch** out) {
  if (!e || !options || !views || !out || num_views <= 0) return fail(PM_ERR_INVALID_ARG, "invalid batch arguments");
  *out = nullptr;
  pm_batch* b = new (std::nothrow) pm_batch();
  if (!b) return fail(PM_ERR_OUT_OF_MEMORY, "host allocation");
  b->engine_ref.hold(e);
  b->engine = e;
  for (int32_t i = 0; i < num_views; ++i) {
    pm_problem* p = nullptr;
    pm_status st = pm_problem_create(e, &options[i], &views[i], &p);
    if (st != PM_OK) {
      const std::string msg = "view " + std::to_string(i) + ": " + g_last_error;
      delete b;
      return fail(st, msg);
    }
    b->views.push_back(p);
  }
  *out = b;
  return PM_OK;
}

PM_API pm_status pm_batch_set_graph(pm_batch* b, int32_t enable) {
  if (!b) return fail(PM_ERR_INVALID_ARG, "null batch");
  b->use_graph = enable != 0;
  return PM_OK;
}

PM_API pm_status pm_batch_set_autotune(pm_batch* b, int32_t enable) {
  if (!b) return fail(PM_ERR_INVALID_ARG, "null batch");
  b->autotune = enable != 0;
  return PM_OK;
}

PM_API pm_status pm_batch_run(pm_batch* b) {
  if (!b) return fail(PM_ERR_INVALID_ARG, "null batch");
  pm_engine* e = b->engine;
  if (b->autotune && !b->tuned) {
    pm_status rs = tune_batch(b->views);
    if (rs != PM_OK) return rs;
    b->tuned = true;
    // shapes may have changed: rebuild the launches and any captured graph
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    PM_HIP(hipStreamSynchronize(e->stream));
    b->schedule.release();
  }
  if (!b->use_graph || e->counting) return run_views(e, b->views, b->schedule, &b->stats);
  PM_HIP(hipSetDevice(e->device));
  hipStream_t s = e->stream;
  pm_status rs;
  if (!b->exec) {
    if (!b->schedule.dev && (rs = build_schedule(b->views, all_ticks(b->views), b->schedule, s)) != PM_OK) return rs;
    // counters off before capture (their pointer lives in the view block)
    if ((rs = prepare_views(b->views, false, s, false)) != PM_OK) return rs;
    PM_HIP(hipStreamSynchronize(s));
    PM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rs = prepare_views(b->views, false, s, false);
    if (rs == PM_OK) rs = launch_all(b->schedule, s, nullptr, nullptr);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(s, &g);
    if (rs != PM_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rs;
    }
    if (ce != hipSuccess) return fail(PM_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
    b->graph = g;
    PM_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  EventTimer t;
  if ((rs = t.rec(s)) != PM_OK) return rs;
  PM_HIP(hipGraphLaunch(b->exec, s));
  if ((rs = t.rec(s)) != PM_OK) return rs;
  PM_HIP(hipStreamSynchronize(s));
  std::memset(&b->stats, 0, sizeof(b->stats));
  b->stats.total_ms = t.between(0, 1);
  for (const pm_schedule::Launch& l : b->schedule.launches) {
    if (l.kind == pm::kPhaseSweep) b->stats.sweep_launches++;
    if (l.kind == pm::kPhaseLocal) b->stats.local_opt_launches++;
  }
  return PM_OK;
}

PM_API pm_status pm_batch_get_view(const pm_batch* b, int32_t index, pm_problem** view) {
  if (!b || !view) return fail(PM_ERR_INVALID_ARG, "null argument");
  if (index < 0 || index >= (int32_t)b->views.size()) return fail(PM_ERR_INVALID_ARG, "view index out of range");
  *view = b->views[index];
  return PM_OK;
}

PM_API pm_status pm_batch_get_stats(const pm_batch* b, pm_run_stats* out) {
  if (!b || !out) return fail(PM_ERR_INVALID_ARG, "null argument");
  *out = b->stats;
  return PM_OK;
}

PM_API void pm_batch_destroy(pm_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->engine->device);
  (void)hipStreamSynchronize(b->engine->stream);
  delete b;
}

PM_API pm_status pm_problem_run(pm_problem* p) {
  if (!p) return fail(PM_ERR_INVALID_ARG, "null problem");
  if (!p->schedule) {
    p->schedule = new (std::nothrow) pm_schedule();
    if (!p->schedule) return fail(PM_ERR_OUT_OF_MEMORY, "host allocation");
    p->schedule->cache = &p->engine->cache;
  }
  return run_views(p->engine, {p}, *p->schedule, &p->stats);
}

namespace {

pm_status copy_out(const pm_problem* p, void* dst, const void* src, size_t bytes) {
  if (!p || !dst) return fail(PM_ERR_INVALID_ARG, "null argument");
  PM_HIP(hipSetDevice(p->engine->device));
  PM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, p->engine->stream));
  PM_HIP(hipStreamSynchronize(p->engine->stream));
  return PM_OK;
}

}  // namespace

PM_API pm_status pm_problem_get_depth_map(const pm_problem* p, float* out) {
  if (!p) return fail(PM_ERR_INVALID_ARG, "null problem");
  return copy_out(p, out, p->host.depth, (size_t)p->W * p->H * sizeof(float));
}
PM_API pm_status pm_problem_get_normal_map(const pm_problem* p, float* out) {
  if (!p) return fail(PM_ERR_INVALID_ARG, "null problem");
  return copy_out(p, out, p->host.normal, 3 * (size_t)p->W * p->H * sizeof(float));
}
PM_API pm_status pm_problem_get_conf_map(const pm_problem* p, float* out) {
  if (!p) return fail(PM_ERR_INVALID_ARG, "null problem");
  return copy_out(p, out, p->host.conf, (size_t)p->W * p->H * sizeof(float));
}
PM_API pm_status pm_problem_get_cost_map(const pm_problem* p, float* out) {
  if (!p) return fail(PM_ERR_INVALID_ARG, "null problem");
  return copy_out(p, out, p->host.cost, (size_t)p->N * p->W * p->H * sizeof(float));
}
PM_API pm_status pm_problem_get_view_sel(const pm_problem* p, uint32_t* out) {
  if (!p) return fail(PM_ERR_INVALID_ARG, "null problem");
  return copy_out(p, out, p->host.view_sel, (size_t)p->W * p->H * sizeof(uint32_t));
}
PM_API pm_status pm_problem_get_curvature_map(const pm_problem* p, uint16_t* out) {
  if (!p) return fail(PM_ERR_INVALID_ARG, "null problem");
  if (!p->host.curv) return fail(PM_ERR_INVALID_ARG, "no curvature map: needs geom_consistency and est_curvature");
  return copy_out(p, out, p->host.curv, 2 * (size_t)p->W * p->H * sizeof(uint16_t));
}
PM_API pm_status pm_problem_get_stats(const pm_problem* p, pm_run_stats* out) {
  if (!p || !out) return fail(PM_ERR_INVALID_ARG, "null argument");
  *out = p->stats;
  return PM_OK;
}
PM_API pm_status pm_problem_get_poses(const pm_problem* p, float* poses, float ref_k8[8]) {
  if (!p || !poses || !ref_k8) return fail(PM_ERR_INVALID_ARG, "null argument");
  std::memcpy(poses, p->poses.data(), p->poses.size() * sizeof(float));
  std::memcpy(ref_k8, p->refk8, sizeof(p->refk8));
  return PM_OK;
}

PM_API pm_status pm_problem_run_filter(pm_problem* p, int32_t deduplication) {
  if (!p) return fail(PM_ERR_INVALID_ARG, "null problem");
  if (!p->opt.geom_consistency || !p->init_depth)
    return fail(PM_ERR_INVALID_ARG, "RunFilter needs geom_consistency = 1 and a given depth map (init_depth_random = 0)");
  PM_HIP(hipSetDevice(p->engine->device));
  hipStream_t s = p->engine->stream;
  const size_t plane = (size_t)p->W * p->H;
  PM_HIP(hipMemcpyAsync(p->host.depth, p->init_depth, plane * sizeof(float), hipMemcpyDeviceToDevice, s));
  const int min_c = std::min(p->opt.filter_min_num_consistent, p->N);
  PM_HIP(pm::launch_run_filter(p->dev, p->W, p->H, deduplication ? 1 : 0, p->opt.filter_geom_consistency_max_cost,
                               min_c, s));
  PM_HIP(hipStreamSynchronize(s));
  return PM_OK;
}

PM_API pm_status pm_problem_init_from(pm_problem* dst, const pm_problem* from) {
  if (!dst || !from) return fail(PM_ERR_INVALID_ARG, "null problem");
  if (dst->engine != from->engine) return fail(PM_ERR_INVALID_ARG, "problems live on different engines");
  if (dst->W != from->W || dst->H != from->H) return fail(PM_ERR_INVALID_ARG, "map sizes differ");
  if (!dst->init_depth) return fail(PM_ERR_INVALID_ARG, "destination was created with init_depth_random");
  const size_t plane = (size_t)dst->W * dst->H;
  hipStream_t s = dst->engine->stream;
  PM_HIP(hipSetDevice(dst->engine->device));
  PM_HIP(hipMemcpyAsync(dst->init_depth, from->host.depth, plane * sizeof(float), hipMemcpyDeviceToDevice, s));
  PM_HIP(hipMemcpyAsync(dst->init_normal, from->host.normal, 3 * plane * sizeof(float), hipMemcpyDeviceToDevice, s));
  return PM_OK;
}

PM_API pm_status pm_problem_set_src_depth(pm_problem* dst, int32_t src_index, const pm_problem* from) {
  if (!dst || !from) return fail(PM_ERR_INVALID_ARG, "null problem");
  if (dst->engine != from->engine) return fail(PM_ERR_INVALID_ARG, "problems live on different engines");
  if (!dst->opt.geom_consistency) return fail(PM_ERR_INVALID_ARG, "destination has no geometric-consistency term");
  if (src_index < 0 || src_index >= dst->N) return fail(PM_ERR_INVALID_ARG, "source index out of range");
  const pm::SrcDev& sd = dst->host.src[src_index];
  if (sd.width != from->W || sd.height != from->H) return fail(PM_ERR_INVALID_ARG, "depth map size differs from source");
  PM_HIP(hipSetDevice(dst->engine->device));
  PM_HIP(pm::launch_build_depth_quads(from->host.depth, from->W, from->H, const_cast<float4*>(sd.depth_q),
                                      dst->engine->stream));
  return PM_OK;
}

PM_API pm_status pm_run_view(pm_engine* e, const pm_options* o, const pm_view_desc* v, const pm_outputs* out) {
  if (!out) return fail(PM_ERR_INVALID_ARG, "null outputs");
  pm_problem* p = nullptr;
  pm_status st = pm_problem_create(e, o, v, &p);
  if (st != PM_OK) return st;
  st = pm_problem_run(p);
  if (st == PM_OK && out->depth) st = pm_problem_get_depth_map(p, out->depth);
  if (st == PM_OK && out->normal) st = pm_problem_get_normal_map(p, out->normal);
  if (st == PM_OK && out->conf) st = pm_problem_get_conf_map(p, out->conf);
  if (st == PM_OK && out->cost) st = pm_problem_get_cost_map(p, out->cost);
  if (st == PM_OK && out->view_sel) st = pm_problem_get_view_sel(p, out->view_sel);
  if (st == PM_OK && out->curvature && p->host.curv) st = pm_problem_get_curvature_map(p, out->curvature);
  pm_problem_destroy(p);
  return st;
}
