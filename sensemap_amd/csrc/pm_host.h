// pm_host.h -- host-side internals shared by the engine (pm_engine.hip) and
// the serving loop (pm_server.hip).  Not part of the public C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/sensemap_pm.h"
#include "pm_device.h"

#define PM_API extern "C" __attribute__((visibility("default")))

namespace pmh {

// Record `msg` as this thread's pm_last_error() text and return `st`.
pm_status fail(pm_status st, const std::string& msg);
const std::string& last_error();

}  // namespace pmh

#define PM_HIP(call)                                                                           \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) {                                                                    \
      return ::pmh::fail(e_ == hipErrorOutOfMemory ? PM_ERR_OUT_OF_MEMORY : PM_ERR_DEVICE,     \
                         std::string(#call) + ": " + hipGetErrorString(e_));                   \
    }                                                                                          \
  } while (0)

namespace pmh {

// Launch-shape class: phase kernel, window, storage, reciprocal mode and
// source-count bucket.
struct TuneKey {
  int kind, radius, step, storage, rcp, nbucket;
  bool operator<(const TuneKey& o) const {
    const int a[6] = {kind, radius, step, storage, rcp, nbucket};
    const int b[6] = {o.kind, o.radius, o.step, o.storage, o.rcp, o.nbucket};
    for (int k = 0; k < 6; ++k)
      if (a[k] != b[k]) return a[k] < b[k];
    return false;
  }
};

// First-fit allocator over one device allocation: the HBM a serving loop
// may hold.  Offsets are 256-byte aligned.  Thread-safe.
class Arena {
 public:
  static constexpr size_t kAlign = 256;
  Arena() = default;
  Arena(const Arena&) = delete;
  Arena& operator=(const Arena&) = delete;
  ~Arena();
  pm_status init(size_t bytes);
  // false when no hole is large enough
  bool alloc(size_t bytes, size_t* off);
  void release(size_t off, size_t bytes);
  uint8_t* base() const { return base_; }
  size_t capacity() const { return cap_; }
  size_t used() const;
  size_t largest_hole() const;
  static size_t round(size_t bytes) { return (bytes + kAlign - 1) & ~(kAlign - 1); }

 private:
  mutable std::mutex mu_;
  std::map<size_t, size_t> holes_;  // offset -> length
  uint8_t* base_ = nullptr;
  size_t cap_ = 0, used_ = 0;
};

// Where a problem's device memory and uploads come from.
struct CreateEnv {
  hipStream_t stream = nullptr;          // uploads and set-up kernels (null: the engine's stream)
  Arena* arena = nullptr;                // non-null: all job memory in one arena range, reserved by the caller
  size_t block_off = 0, block_bytes = 0;  //   that range
  const void* const* src_tex = nullptr;  // non-null: per-source quads already on the device (null entries: upload)
  uint8_t* staging = nullptr;            // non-null: upload staging buffer of staging_bytes
  size_t staging_bytes = 0;
};

}  // namespace pmh

namespace pmh {

// Device blocks an engine keeps for reuse.  Problems and their schedules
// hand their buffers back here instead of freeing them, so creating a
// problem of a size seen before allocates nothing (the reference's caller
// builds a new engine per problem, paying every allocation).  Blocks are
// matched by size rounded to 64 KiB; at most kMaxBytes are kept, the rest is
// freed.  Everything is freed with the engine.
struct DevCache {
  static constexpr size_t kMaxBytes = (size_t)16 << 30;
  static size_t round(size_t b) { return (std::max<size_t>(b, 1) + 0xffff) & ~(size_t)0xffff; }
  hipError_t get(size_t bytes, void** out);
  void put(void* p, size_t bytes);
  ~DevCache();

 private:
  std::mutex mu_;
  std::multimap<size_t, void*> blocks_;
  size_t bytes_ = 0;
};

// An engine is reference counted: the caller's handle holds one reference
// and every problem, batch and server created on it holds one more, so
// pm_engine_destroy may come before the objects made from it (a garbage
// collector finalises in any order) and the stream and cached blocks they
// still use stay valid until the last of them is destroyed.
void engine_retain(pm_engine* e);
void engine_release(pm_engine* e);

// An owner's reference on its engine.  Declared as the owner's first member,
// it is released after every other member (members are destroyed in reverse
// order of declaration), so their destructors may still use the engine.
struct EngineRef {
  pm_engine* e = nullptr;
  void hold(pm_engine* engine) {
    engine_retain(engine);
    e = engine;
  }
  ~EngineRef() {
    if (e) engine_release(e);
  }
};

}  // namespace pmh

struct pm_engine {
  int device = 0;
  std::atomic<int> refs{1};
  pmh::DevCache cache;
  hipStream_t stream = nullptr;
  int counting = 0;
  std::mutex tune_mu;
  std::map<pmh::TuneKey, pm_tuning> tuning;  // absent = default shape
  // two pinned host buffers for large copies from and to pageable caller
  // memory (pmh::copy_h2d / copy_d2h), made on first use
  std::mutex stage_mu;
  uint8_t* stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
};

struct pm_problem {
  pm_engine* engine = nullptr;
  pm_options opt{};
  int W = 0, H = 0, N = 0;
  uint32_t view_id = 0;
  float depth_min = 0.f, depth_max = 0.f;
  std::vector<std::pair<void*, size_t>> allocs;  // from the engine's block cache (engine problems)
  pmh::Arena* arena = nullptr;         // arena-backed (server jobs): one range, carved in order
  size_t block_off = 0, block_bytes = 0, block_used = 0;
  pm::ViewDev host{};
  pm::ViewDev* dev = nullptr;
  float* init_depth = nullptr;   // device copies of the given initial state
  float* init_normal = nullptr;
  unsigned long long* counters = nullptr;
  std::vector<float> poses;      // N * PM_POSE_STRIDE
  float refk8[8] = {};
  pm_camera ref_cam{};           // the reference view's camera as given
  pm_run_stats stats{};
  size_t device_bytes = 0;
  struct pm_schedule* schedule = nullptr;

  template <typename T>
  pm_status alloc(T** p, size_t count) {
    const size_t bytes = count * sizeof(T);
    if (arena) {
      const size_t off = block_used;
      const size_t need = pmh::Arena::round(bytes > 0 ? bytes : 1);
      if (off + need > block_bytes) return pmh::fail(PM_ERR_INTERNAL, "job memory block too small");
      block_used += need;
      device_bytes += need;
      *p = (T*)(arena->base() + block_off + off);
      return PM_OK;
    }
    void* q = nullptr;
    hipError_t e = engine->cache.get(bytes, &q);
    if (e != hipSuccess) {
      return pmh::fail(e == hipErrorOutOfMemory ? PM_ERR_OUT_OF_MEMORY : PM_ERR_DEVICE,
                       "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    allocs.emplace_back(q, bytes);
    device_bytes += bytes;
    *p = (T*)q;
    return PM_OK;
  }
  ~pm_problem();
};

namespace pmh {

struct EventTimer {
  std::vector<hipEvent_t> ev;
  pm_status rec(hipStream_t s);
  float between(size_t a, size_t b) const;
  ~EventTimer();
};

// One phase of one view: the iteration schedule of Run()
// (patch_match_cuda.cu:774-909): init, then per iteration the black and red
// sweeps and, from iteration 3, the black and red local optimisation; the
// optional FilterDepthMap pair at the end (:1007-1033).
struct PhaseRec {
  pm::PhaseKind kind;
  int color;
  uint32_t pass;
  float th_mc;
};

using Ticks = std::vector<std::vector<std::vector<PhaseRec>>>;  // [view][tick][phase]

float th_mc_of(const pm_options& o, int iter);
std::vector<std::vector<PhaseRec>> view_ticks(const pm_problem& p);
Ticks all_ticks(const std::vector<pm_problem*>& views);

pm::LaunchCfg cfg_of(const pm_options& o);
TuneKey tune_key(int kind, const pm_options& o, int N);
pm_tuning shape_of(pm_engine* e, int kind, const pm_options& o, int N);

// Validation shared by every entry point that takes a view.
pm_status check_view(const pm_options* o, const pm_view_desc* v);
bool finite_camera(const pm_camera& c);

// Copies between caller (pageable) host memory and the device through the
// engine's two pinned staging buffers, in chunks and double buffered: the
// CPU copy of one chunk (split over a few threads) overlaps the DMA of the
// other.  The host buffer may be reused on return; copy_d2h also waits for
// the data.  Small copies go straight through hipMemcpyAsync.
pm_status copy_h2d(pm_engine* e, void* dst, const void* src, size_t bytes, hipStream_t st);
pm_status copy_d2h(pm_engine* e, void* dst, const void* src, size_t bytes, hipStream_t st);
// NormalMap codes of a planar normal map on the device (k_encode_normals),
// enqueued on s.
hipError_t encode_normals_async(const float* normal, int64_t plane, uint32_t* out, hipStream_t s);
// InitTransforms (patch_match_cuda.cu:493-580): PM_POSE_STRIDE floats per
// source, FP64 algebra rounded once, and the ref {fx,cx,fy,cy, 1/fx,...}.
void pack_poses(const pm_view_desc* view, float* poses, float k8[8]);

// Device bytes of one job's memory block (state, tables, reference and the
// sources not given in `src_cached`), in Arena::round units.
size_t job_block_bytes(const pm_options* o, const pm_view_desc* v, const bool* src_cached);
// Device bytes of one source's quad texels at the given storage.
size_t src_quad_bytes(const pm_image& im, int storage);
// Upload one 8-bit image into quads at `dst` (staging: a device buffer of at
// least width*height bytes; stream-ordered, no wait).
pm_status upload_quads(const pm_image& im, int storage, void* dst, uint8_t* staging, hipStream_t s);

pm_status create_problem(pm_engine* e, const pm_options* o, const pm_view_desc* v, const CreateEnv& env,
                         pm_problem** out);

}  // namespace pmh

// All launches of one run over a set of views: at tick t every view runs its
// phase t, and views whose phase shares the kernel (kind, window shape and
// launch shape) share one launch, described by a LaunchTable.  The tables of
// all launches are built on the host once and uploaded in one copy.
struct pm_schedule {
  struct Launch {
    pm::PhaseKind kind;
    pm::LaunchCfg cfg;
    size_t offset;  // of the launch's table in `tables`
    int tiles;
    int grid;       // persistent workgroups
  };
  std::vector<uint8_t> tables;
  std::vector<Launch> launches;
  void* dev = nullptr;         // device copy of `tables`
  uint32_t* queues = nullptr;  // 8 tile counters per launch
  float* ws = nullptr;         // workspace sized for the largest launch
  size_t ws_floats = 0;
  pmh::DevCache* cache = nullptr;  // non-null: buffers come from and go back to this cache
  size_t dev_bytes = 0, queue_bytes = 0, ws_bytes = 0;
  void release() {
    void* p[3] = {dev, queues, ws};
    const size_t b[3] = {dev_bytes, queue_bytes, ws_bytes};
    for (int k = 0; k < 3; ++k) {
      if (!p[k]) continue;
      if (cache) cache->put(p[k], b[k]);
      else (void)hipFree(p[k]);
    }
    dev = nullptr;
    queues = nullptr;
    ws = nullptr;
  }
  ~pm_schedule() { release(); }
};

namespace pmh {

// Host image of the launches of one tick (or of a whole run): groups the
// phases by kernel class and writes LaunchTables.  `ws` and `queue` pointers
// are patched by the caller once the device buffers are known.
struct TickPlan {
  std::vector<uint8_t> tables;
  std::vector<pm_schedule::Launch> launches;
  size_t ws_floats = 0;  // workspace the launches need
};
pm_status plan_launches(const std::vector<std::pair<pm_problem*, const std::vector<PhaseRec>*>>& phases,
                        TickPlan& plan, const pm_tuning* force);

pm_status build_schedule(const std::vector<pm_problem*>& views, const Ticks& ticks, pm_schedule& sch, hipStream_t s,
                         const pm_tuning* force = nullptr);
pm_status prepare_views(const std::vector<pm_problem*>& views, bool counting, hipStream_t s, bool allow_memset);
pm_status launch_all(const pm_schedule& sch, hipStream_t s, EventTimer* t, pm_run_stats* st);
pm_status run_views(pm_engine* e, const std::vector<pm_problem*>& views, pm_schedule& sch, pm_run_stats* st);
pm_status tune_batch(const std::vector<pm_problem*>& views);

}  // namespace pmh
