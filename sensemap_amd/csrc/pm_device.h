// pm_device.h -- device-side view descriptor and launch interface of the
// gfx950 PatchMatch kernels.  Shared by the kernels (pm_kernels.hip) and the
// host engine (pm_engine.hip); not part of the public C-ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sensemap_pm.h"

namespace pm {

constexpr int kMaxHyp = 11;          // 1 current + 2 neighbours + 8 directional
constexpr int kNumRefine = 5;        // current + 4 refinement hypotheses
constexpr int kQuadPad = 4;          // zero texels around every source image
// A tile is 2^lw columns x 512 / 2^lw rows of one view; its 256 threads
// cover the pixels of one checkerboard colour (2^(lw-1) threads per row).
// Shapes: lw = 4 (16x32), 5 (32x16, the default), 6 (64x8).
constexpr int kTileLog2WMin = 4, kTileLog2WMax = 6, kTileLog2WDefault = 5;
constexpr int kTilePixels = 512;
constexpr int kMaxBlocksPerCU = 8;   // resident workgroups of a persistent launch per CU, at most
constexpr int kSweepThreads = 256;
constexpr float kMaxCost = 2.0f;
constexpr float kVarThreshold = 0.65025f;  // 1e-5 * 255^2
constexpr float kLowTexVar = 25.0f;        // compensated-mode cut (0..255 units)

// Per-source device data.
struct SrcDev {
  const void* tex;          // quad texels: uint32 (u8 x4) or uint2 (f16 x4)
  const float4* depth_q;    // quad texels of the source depth map (geometric pass)
  int32_t width, height;    // source image size
  int32_t pitch;            // width + 2*kQuadPad (texels per padded row)
  int32_t pad0;
};

// One view as the kernels see it: lives in device memory, read through
// uniform (scalar) loads.
struct ViewDev {
  int32_t W, H, N;
  int32_t nwin;             // samples per window row
  int32_t radius, step;
  int32_t view_id;
  int32_t geom;             // geometric-consistency term on
  float refK[4];
  float refInvK[4];
  float depth_min, depth_max, depth_max_eff;  // eff = 1.2 * depth_max (sweep random depth)
  float geom_regularizer, geom_max_cost;
  float random_depth_ratio, angle1, angle2;
  float max_ncc_matching_cost, inv_beta2;
  int32_t num_good, num_bad, num_samples, random_optimization;
  int32_t thk, init_samples, init_random, compat_lo_geo_bug;
  int32_t accum_mode, pad_ws;
  float conf_threshold, depth_diff_threshold;
  uint64_t seed;
  const uint8_t* ref;       // W*H
  const float4* refm;       // W*H {mu_r, var_r, 1/sum_w, err_r}
  const uint8_t* grad_ok;   // W*H
  const float* poses;       // N * PM_POSE_STRIDE
  const float* sp;          // nwin*nwin spatial weights
  const float* ct;          // 256 colour weights
  float* depth;             // W*H
  float* normal;            // 3*W*H planar
  float* conf;              // W*H
  float* cost;              // N*W*H planar
  uint32_t* view_sel;       // W*H
  uint16_t* curv;           // 2*W*H planar {gradient, curvature} or null (no curvature pass)
  const uint8_t* mask;      // W*H depth mask or null
  float* geom_grad;         // W*H relative-depth gradient (median filter) or null
  unsigned long long* counters;  // [ncc_eval, ncc_samples, geo_eval] or null
  SrcDev src[PM_MAX_NUM_SRC];
};

// Workspace fields per thread of a launch (SoA over the launch's threads:
// field f of slot s at ws[f * ws_slots + s], ws_slots = blocks * 256).
//   hyp   : 4 * (kMaxHyp + kNumRefine) floats  (nx, ny, nz, d per hypothesis)
//   M     : kMaxHyp * N floats      (candidate x source costs)
//   RM    : kNumRefine * N floats   (refinement x source costs)
//   W     : N floats                (view weights, later CDF)
inline int ws_floats_per_slot(int N) {
  return 4 * (kMaxHyp + kNumRefine) + kMaxHyp * N + kNumRefine * N + N;
}

// One entry of a launch table: one view advanced by one phase.  A launch
// covers the entries' tiles back to back; tile_start is the entry's first
// tile within the launch.
struct LaunchEntry {
  const ViewDev* v;
  int32_t tile_start;
  int32_t color;        // checkerboard colour of the pixels this entry updates
  uint32_t pass;        // random-stream pass id
  float th_mc;          // good-matching threshold of the iteration (sweep)
};

// A launch runs `grid` persistent workgroups over `total_tiles` tiles.  Each
// workgroup owns 256 workspace slots (one per thread) for the whole launch,
// so no two resident workgroups ever share workspace lines, and the launch
// boundary makes the workspace coherent across the XCDs' L2 caches.
struct LaunchTable {
  int32_t n;             // entries
  int32_t total_tiles;   // tiles of the launch
  int32_t ws_slots;      // grid * 256
  int32_t tile_lw;       // log2 of the tile width (kTileLog2WMin..kTileLog2WMax)
  float* ws;             // workspace of the launch
  uint32_t* queue;       // 8 tile counters (one per XCD range), zero at launch start
  LaunchEntry e[1];      // n entries follow (allocated with the table)
};

inline size_t launch_table_bytes(int n) {
  return sizeof(LaunchTable) + (size_t)(n > 1 ? n - 1 : 0) * sizeof(LaunchEntry);
}

// Tiles of a view at tile width 2^lw (one workgroup iteration each per
// colour pass).
inline int view_tiles(int W, int H, int lw) {
  const int tw = 1 << lw, th = kTilePixels >> lw;
  return ((W + tw - 1) / tw) * ((H + th - 1) / th);
}

struct LaunchCfg {
  int storage;   // PM_STORAGE_*
  int rcp;       // PM_RCP_*
  int radius, step;
  bool operator==(const LaunchCfg& o) const {
    return storage == o.storage && rcp == o.rcp && radius == o.radius && step == o.step;
  }
};

// Kernel launch wrappers (pm_kernels.hip).  Every launch checks the grid it
// builds against the view sizes first.
hipError_t launch_build_quads_u8(const uint8_t* img, int W, int H, int stride, uint32_t* out,
                                 hipStream_t s);
hipError_t launch_build_quads_f16(const uint8_t* img, int W, int H, int stride, uint2* out,
                                  hipStream_t s);
hipError_t launch_build_depth_quads(const float* depth, int W, int H, float4* out, hipStream_t s);
hipError_t launch_ref_moments(const ViewDev* v, int W, int H, int accum_mode, float4* refm,
                              uint8_t* grad_ok, hipStream_t s);
// Table-driven colour-pass launches: `table` is a device copy of a
// LaunchTable whose queue counters are zero; `grid` comes from phase_grid
// for the table's total_tiles (and sized its ws_slots).
enum PhaseKind {
  kPhaseInit = 0,
  kPhaseSweep = 1,
  kPhaseLocal = 2,
  kPhaseFilter = 3,     // FilterDepthMap, one colour per entry
  kPhaseCurvature = 4,  // ComputeCurvatureMap, one colour per entry
  kPhaseMask = 5,       // MaskDepthMapKernel, one colour per entry
  kPhaseGeomGrad = 6,   // ComputeGeomGradientMap, one colour per entry
  kPhaseMedian = 7      // Black/RedPixelFilter, one colour per entry
};
// Phases with no window, storage or reciprocal variants.
inline bool plain_phase(PhaseKind k) { return k >= kPhaseFilter; }
// Persistent grid for a launch of `tiles` tiles on the current device: the
// number of CUs times the workgroups one CU holds (capped by blocks_per_cu
// when > 0), and never more than the tiles.
int phase_grid(PhaseKind kind, const LaunchCfg& cfg, int tiles, int blocks_per_cu);
// Workgroups of the phase kernel one CU holds at once.
int phase_occupancy(PhaseKind kind, const LaunchCfg& cfg);
hipError_t launch_phase(PhaseKind kind, const LaunchCfg& cfg, const LaunchTable* table, int grid, hipStream_t s);
// RunFilter over the view's current depth map (a 2-D grid of 16x16 pixels).
hipError_t launch_run_filter(const ViewDev* v, int W, int H, int dedup, float filter_max_cost, int min_consistent,
                             hipStream_t s);
// Zero n uint32 words with a kernel (safe inside stream capture).
hipError_t launch_zero_u32(uint32_t* p, int n, hipStream_t s);

}  // namespace pm
