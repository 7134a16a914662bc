// pm_fusion.hip -- depth-map fusion: StereoFusion's walk (src/mvs/fusion.cc:
// 1730-2183, FuseImage and Fuse) over the images in the order of
// StereoFusion::Run (fusion.cc:697-727, 1150-1215), as one fusing thread per
// image runs it.
//
// The walks are sequential by definition: a pixel belongs to the first walk
// that accepts it, and each seed's walk sees the pixels earlier walks took.
// The device keeps that order exactly while it runs walks side by side:
//
//   * Graphs.  Whether a pixel passes a walk's tests depends only on the
//     seed (its point and normal) and on the pixel, and the pixels a passing
//     pixel pushes depend only on the pixel.  k_graph records, one wave per
//     seed of a window, every pixel the walk could reach against the masks
//     of the window's start (breadth first; a masked pixel stays masked) and
//     keeps the ones that pass, each with its pushes of passing pixels in
//     push order (a pushed pixel that fails ends in `continue` when popped).
//     Replaying the depth-first walk over the graph against later masks
//     gives the same pixels, order and cuts as walking the images again.
//   * Dependencies.  A walk's result depends on the masks only at the pixels
//     of its graph, and a walk sets masks only at pixels it accepts, which
//     are in its own graph.  So walk w needs, for each pixel of its graph,
//     the last earlier walk whose graph holds the pixel to have finished
//     (that walk needed the ones before it in turn).  The walks holding each
//     pixel are listed (count, prefix sum, scatter); k_preds turns the lists
//     into each walk's distinct predecessors and k_succ into successor lists.
//   * Dataflow.  k_flow's persistent waves (one per CU) run a walk once its
//     last predecessor has finished, so no wave ever waits on a walk: the
//     wave that finishes that predecessor runs the first walk it made ready
//     itself and puts the others in a ready queue, from which the other
//     waves take them.  A wave replays its walk over the graph (the whole
//     wave: the set of available nodes in scalar registers, one lane per
//     push), sets the accepted pixels' mask bits, publishes them with a
//     release fence and counts itself off at each successor.  Everything a
//     wave reads that another wrote (queue slots, mask words) is an
//     agent-scope atomic load, since the waves run on every XCD.  A wave
//     waiting for a queue slot gives up only when no walk anywhere finished
//     for 4 s, which fails the call.
//   * A graph that outgrows the wave's LDS marks its walk to run alone
//     (k_walk, pixel by pixel, with buffers for max_num_pixels) between two
//     dataflow segments.
//
// The points (medians by rank, the mean gradient in visiting order, the
// score) are formed afterwards, one wave per walk that accepted pixels, from
// values recomputed as k_graph computed them, and gathered in seed order on
// the device.
#include "pm_host.h"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

using namespace pmh;

namespace {

constexpr int kVals = 11;                         // x y z nx ny nz r g b grad err
constexpr int kNodeCap = 256;                     // passing pixels per stored walk graph
constexpr int kChildCap = 2048;                   // pushes per stored walk graph
constexpr int kProvCap = 512;                     // pixels (passing or not) met building a graph
constexpr int kProvChild = 4096;                  // push slots met building a graph
constexpr int kHash = 1024;                       // open-addressing slots (gid -> pixel)
constexpr int64_t kMaxWindow = int64_t(1) << 23;  // seeds per window
// Build knobs of the dataflow, defaults measured on the 1920x1080 11-view
// scene (8: one wave per CU; 16 s_sleep units between polls).
#ifndef PM_FLOW_WAVES_PER_8CU
#define PM_FLOW_WAVES_PER_8CU 8
#endif
constexpr int kFlowWavesPer8CU = PM_FLOW_WAVES_PER_8CU;  // dataflow waves per 8 CUs
#ifndef PM_POLL_SLEEP
#define PM_POLL_SLEEP 16
#endif
constexpr int kPollSleep = PM_POLL_SLEEP;  // s_sleep units (64 clocks) between polls of a queue slot
constexpr int kPollCheck = 64;             // polls between looks at the shared counters
constexpr double kWaitSeconds = 4.0;  // one wait's bound

struct FusImgDev {
  const float* depth;
  const float* normal;  // null without normals
  const uint8_t* rgb;
  const int32_t* ov;
  int64_t base;  // first global pixel id of the image
  int32_t W, H, nov, used;
  float P[12], iP[12], Rt[9];
  int32_t ch;  // colour channels: 3 (r, g, b interleaved) or 1 (grey)
};

struct FusParams {
  int32_t cur, seed_cols, step, n_images;
  int32_t min_num_pixels, max_num_pixels, min_num_visible, max_traversal;
  int32_t with_normal, acc_cap, stack_cap, pad0;
  float max_sq_reproj, min_cos_normal;
  double max_depth_error;
};

// An accepted pixel of a walk run alone: its image and values.
struct AccEntry {
  int32_t img, pad;
  float v[kVals];
};

struct WalkRec {
  int32_t n_acc, pad;
};

struct StackEntry {
  int32_t img, row, col, depth;
};

// A fused point as formed on the device.
struct PointRec {
  float xyz[3], nrm[3], rgb[3], score;
  int32_t nvis, pad;
};

// The mask: one bit per pixel of every image (global pixel ids).
__device__ inline bool masked(const uint32_t* __restrict__ mask, int64_t gid) {
  return (mask[gid >> 5] >> (gid & 31)) & 1u;
}

__device__ inline float row4(const float* m, float x, float y, float z) { return ((m[0] * x + m[1] * y) + m[2] * z) + m[3]; }

// Fuse (fusion.cc:1794-2182) from one seed, against the committed masks,
// pixel by pixel: the accepted pixels (ids in G, values in A) in visiting
// order.  Run for one walk at a time, with buffers for max_num_pixels
// accepted pixels and for every push they can make.
__global__ void __launch_bounds__(64) k_walk(const FusImgDev* __restrict__ imgs, const uint32_t* __restrict__ mask,
                                             const uint8_t* __restrict__ fused, FusParams prm, int32_t seed,
                                             WalkRec* __restrict__ rec_out, int64_t* __restrict__ G,
                                             AccEntry* __restrict__ A, StackEntry* __restrict__ S) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int seed_row = 3 + (seed / prm.seed_cols) * prm.step;
  const int seed_col = 3 + (seed % prm.seed_cols) * prm.step;
  int sp = 0, na = 0;
  S[sp++] = {prm.cur, seed_row, seed_col, 0};
  float ref_pt[3] = {0, 0, 0}, ref_n[3] = {0, 0, 0};
  while (sp > 0) {
    const StackEntry e = S[--sp];
    const FusImgDev& I = imgs[e.img];
    const int W = I.W, H = I.H;
    const int64_t p = (int64_t)e.row * W + e.col;
    const int64_t gid = I.base + p;
    if (masked(mask, gid)) continue;
    bool mine = false;  // taken earlier by this walk
    for (int j = 0; j < na && !mine; ++j) mine = G[j] == gid;
    if (mine) continue;
    const float depth = I.depth[p];
    if (depth <= 0.0f) continue;
    float depth_error = 0.0f;
    if (e.depth > 0) {
      const float p0 = row4(I.P, ref_pt[0], ref_pt[1], ref_pt[2]);
      const float p1 = row4(I.P + 4, ref_pt[0], ref_pt[1], ref_pt[2]);
      const float p2 = row4(I.P + 8, ref_pt[0], ref_pt[1], ref_pt[2]);
      const float cd = p0 / p2 - (float)e.col;
      const float rd = p1 / p2 - (float)e.row;
      if (cd * cd + rd * rd > prm.max_sq_reproj) continue;
      depth_error = fabsf(p2 - depth) / depth;
      if ((double)depth_error > prm.max_depth_error) continue;
    }
    float nrm[3] = {0, 0, 0};
    if (prm.with_normal) {
      const int64_t plane = (int64_t)W * H;
      const float n0 = I.normal[p], n1 = I.normal[plane + p], n2 = I.normal[2 * plane + p];
      for (int i = 0; i < 3; ++i) nrm[i] = (I.Rt[3 * i] * n0 + I.Rt[3 * i + 1] * n1) + I.Rt[3 * i + 2] * n2;
      if (e.depth > 0) {
        const float c = (ref_n[0] * nrm[0] + ref_n[1] * nrm[1]) + ref_n[2] * nrm[2];
        if (c < prm.min_cos_normal) continue;
      }
    }
    const float u = (float)e.col * depth, v = (float)e.row * depth;
    const float xyz[3] = {row4(I.iP, u, v, depth), row4(I.iP + 4, u, v, depth), row4(I.iP + 8, u, v, depth)};
    float grad = 0.0f;
    if (e.col > 1 && e.col < W - 1 && e.row > 1 && e.row < H - 1) {
      const int l = I.rgb[I.ch * (p - 1)], r = I.rgb[I.ch * (p + 1)];
      const int t = I.rgb[I.ch * (p - W)], b = I.rgb[I.ch * (p + W)];
      const float gx = (float)(l - r) / 255.0f, gy = (float)(t - b) / 255.0f;
      grad = sqrtf(gx * gx + gy * gy);
    }
    G[na] = gid;
    AccEntry& a = A[na++];
    a.img = e.img;
    a.v[0] = xyz[0];
    a.v[1] = xyz[1];
    a.v[2] = xyz[2];
    a.v[3] = nrm[0];
    a.v[4] = nrm[1];
    a.v[5] = nrm[2];
    const uint8_t* cp = I.rgb + I.ch * p;  // grey: channel 0 three times
    a.v[6] = (float)cp[0];
    a.v[7] = (float)cp[I.ch == 3 ? 1 : 0];
    a.v[8] = (float)cp[I.ch == 3 ? 2 : 0];
    a.v[9] = grad;
    a.v[10] = depth_error;
    if (e.depth == 0) {
      ref_pt[0] = xyz[0];
      ref_pt[1] = xyz[1];
      ref_pt[2] = xyz[2];
      ref_n[0] = nrm[0];
      ref_n[1] = nrm[1];
      ref_n[2] = nrm[2];
    }
    if (na >= prm.max_num_pixels) break;
    const int nd = e.depth + 1;
    if (nd >= prm.max_traversal) continue;
    // at most one push per (accepted pixel, overlapping image) plus the
    // seed: the stack (stack_cap entries) never fills
    for (int k = 0; k < I.nov && sp < prm.stack_cap; ++k) {
      const int j = I.ov[k];
      if (j < 0 || j >= prm.n_images) continue;
      const FusImgDev& J = imgs[j];
      if (!J.used || fused[j]) continue;
      const float q0 = row4(J.P, xyz[0], xyz[1], xyz[2]);
      const float q1 = row4(J.P + 4, xyz[0], xyz[1], xyz[2]);
      const float q2 = row4(J.P + 8, xyz[0], xyz[1], xyz[2]);
      const float fc = roundf(q0 / q2), fr = roundf(q1 / q2);
      if (!(fabsf(fc) < 1073741824.0f) || !(fabsf(fr) < 1073741824.0f)) continue;
      const int nc = (int)fc, nr = (int)fr;
      if (nc < 3 || nr < 3 || nc >= J.W - 3 || nr >= J.H - 3) continue;
      S[sp++] = {j, nr, nc, nd};
    }
  }
  rec_out->n_acc = na;
}

// the tests and values of one pixel against the seed's point and normal
// (fusion.cc:1884-2040); returns whether it passes.  Without check_mask the
// mask is not read (the values of a pixel already accepted).
__device__ bool test_pixel(const FusImgDev& I, const uint32_t* __restrict__ mask, int row, int col, bool seed,
                           const float ref_pt[3], const float ref_n[3], const FusParams& prm, float* v,
                           bool check_mask = true) {
  const int W = I.W, H = I.H;
  const int64_t p = (int64_t)row * W + col;
  if (check_mask && masked(mask, I.base + p)) return false;
  const float depth = I.depth[p];
  if (depth <= 0.0f) return false;
  float depth_error = 0.0f;
  if (!seed) {
    const float p0 = row4(I.P, ref_pt[0], ref_pt[1], ref_pt[2]);
    const float p1 = row4(I.P + 4, ref_pt[0], ref_pt[1], ref_pt[2]);
    const float p2 = row4(I.P + 8, ref_pt[0], ref_pt[1], ref_pt[2]);
    const float cd = p0 / p2 - (float)col;
    const float rd = p1 / p2 - (float)row;
    if (cd * cd + rd * rd > prm.max_sq_reproj) return false;
    depth_error = fabsf(p2 - depth) / depth;
    if ((double)depth_error > prm.max_depth_error) return false;
  }
  float nrm[3] = {0, 0, 0};
  if (prm.with_normal) {
    const int64_t plane = (int64_t)W * H;
    const float n0 = I.normal[p], n1 = I.normal[plane + p], n2 = I.normal[2 * plane + p];
    for (int i = 0; i < 3; ++i) nrm[i] = (I.Rt[3 * i] * n0 + I.Rt[3 * i + 1] * n1) + I.Rt[3 * i + 2] * n2;
    if (!seed) {
      const float c = (ref_n[0] * nrm[0] + ref_n[1] * nrm[1]) + ref_n[2] * nrm[2];
      if (c < prm.min_cos_normal) return false;
    }
  }
  const float u = (float)col * depth, w = (float)row * depth;
  v[0] = row4(I.iP, u, w, depth);
  v[1] = row4(I.iP + 4, u, w, depth);
  v[2] = row4(I.iP + 8, u, w, depth);
  v[3] = nrm[0];
  v[4] = nrm[1];
  v[5] = nrm[2];
  const uint8_t* cp = I.rgb + I.ch * p;  // grey: channel 0 three times
  v[6] = (float)cp[0];
  v[7] = (float)cp[I.ch == 3 ? 1 : 0];
  v[8] = (float)cp[I.ch == 3 ? 2 : 0];
  float grad = 0.0f;
  if (col > 1 && col < W - 1 && row > 1 && row < H - 1) {
    const int l = I.rgb[I.ch * (p - 1)], r = I.rgb[I.ch * (p + 1)];
    const int t = I.rgb[I.ch * (p - W)], b = I.rgb[I.ch * (p + W)];
    const float gx = (float)(l - r) / 255.0f, gy = (float)(t - b) / 255.0f;
    grad = sqrtf(gx * gx + gy * gy);
  }
  v[9] = grad;
  v[10] = depth_error;
  return true;
}

// The window's walk graphs, packed one after another.  A stored graph holds
// only the pixels that pass the walk's tests (node 0 is the seed), and per
// pixel only its pushes of passing pixels, in push order: a pushed pixel
// that fails ends in `continue` when popped and changes nothing.
struct Pool {
  int32_t* n;          // [seed] nodes; 0: the seed fails; -1: outgrew LDS; -2: no room in the pool
  int32_t* nc;         // [seed] pushes
  int32_t* node_off;   // [seed] first node
  int32_t* child_off;  // [seed] first push
  int64_t* gid;        // [node]
  int32_t* img;        // [node]
  int16_t* cbase;      // [node] first push within the graph's; -1 when not expanded
  int16_t* ccnt;       // [node] pushes
  int16_t* child;      // [push] node index
  unsigned long long* used;  // [2] nodes and pushes handed out
  int32_t* no_room;          // first seed of the window that found no room
  int64_t node_cap, child_cap;
};

__device__ inline int hash_slot(int64_t gid) { return (int)((uint64_t)gid * 0x9E3779B97F4A7C15ull >> 54); }

// One wave (one block of 64) per seed of the window.
__global__ void __launch_bounds__(64) k_graph(const FusImgDev* __restrict__ imgs, const uint32_t* __restrict__ mask,
                                              const uint8_t* __restrict__ fused, FusParams prm, int32_t window_first,
                                              int32_t count, Pool g) {
  __shared__ int64_t hkey[kHash];
  __shared__ int16_t hval[kHash];
  __shared__ uint16_t nimg[kProvCap];
  __shared__ int32_t nrc[kProvCap];  // row << 16 | col
  __shared__ uint8_t npass[kProvCap];
  __shared__ int16_t cbase[kProvCap], ccnt[kProvCap];
  __shared__ float nxyz[kProvCap][3];
  __shared__ int16_t child[kProvChild];
  __shared__ int16_t front[kProvCap];  // the frontier; after the walk, each pixel's node (-1: fails)
  __shared__ int16_t vfirst[kProvCap];  // first stored push of each pixel
  __shared__ uint8_t sown[128], sk[128];  // a push slot's frontier pixel (in the batch) and candidate
  __shared__ long long offs[2];
  const int ws = blockIdx.x, lane = threadIdx.x;
  if (ws >= count) return;
  // past a seed that found the pool full: no room either (that seed ends the window)
  if (ws > __hip_atomic_load(g.no_room, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
    if (lane == 0) g.n[ws] = -2;
    return;
  }
  const int s = window_first + ws;
  const int seed_row = 3 + (s / prm.seed_cols) * prm.step;
  const int seed_col = 3 + (s % prm.seed_cols) * prm.step;
  // the seed (pixel 0); a seed that fails has an empty walk and no graph
  float ref_pt[3], ref_n[3];
  {
    const FusImgDev& I = imgs[prm.cur];
    float v[kVals];
    if (!test_pixel(I, mask, seed_row, seed_col, true, nullptr, nullptr, prm, v)) {
      if (lane == 0) g.n[ws] = 0;
      return;
    }
    for (int k = 0; k < 3; ++k) {
      ref_pt[k] = v[k];
      ref_n[k] = v[3 + k];
    }
    for (int i = lane; i < kHash; i += 64) hkey[i] = -1;
    __syncthreads();
    if (lane == 0) {
      const int64_t gid = I.base + (int64_t)seed_row * I.W + seed_col;
      nimg[0] = (uint16_t)prm.cur;
      nrc[0] = (seed_row << 16) | seed_col;
      npass[0] = 1;
      const int h = hash_slot(gid);
      hkey[h] = gid;
      hval[h] = 0;
      for (int k = 0; k < 3; ++k) nxyz[0][k] = v[k];
      cbase[0] = -1;
      ccnt[0] = 0;
      front[0] = 0;
    }
  }
  int nn = 1, nch = 0, nf = 1 < prm.max_traversal ? 1 : 0, level = 0;
  bool overflow = false;
  __syncthreads();
  // Pushes, two slots per lane (slots lane and 64 + lane of a batch): a push
  // of frontier pixel f is its projection into its k-th overlapping image.
  // The projections' loads are in flight together; the candidates' pixels
  // are claimed all lanes at once, a compare-and-swap on the hash key
  // claiming a new pixel; the new ones get numbers in slot order (the
  // numbering does not affect the walk, only the push order does, and each
  // push keeps its slot).  idx is the pushed pixel's number (-1: none);
  // overflow is set when the pixels outgrow the graph.
  auto project = [&](int f, int k, bool active, int& cj, int& crow, int& ccol) __attribute__((always_inline)) {
    cj = -1;
    crow = 0;
    ccol = 0;
    if (!active) return;
    const FusImgDev& F = imgs[nimg[f]];
    const int j = F.ov[k];
    if (j >= 0 && j < prm.n_images && imgs[j].used && !fused[j]) {
      const FusImgDev& J = imgs[j];
      const float x = nxyz[f][0], y = nxyz[f][1], z = nxyz[f][2];
      const float q0 = row4(J.P, x, y, z), q1 = row4(J.P + 4, x, y, z), q2 = row4(J.P + 8, x, y, z);
      const float fc = roundf(q0 / q2), fr = roundf(q1 / q2);
      if (fabsf(fc) < 1073741824.0f && fabsf(fr) < 1073741824.0f) {
        const int nc = (int)fc, nr = (int)fr;
        if (!(nc < 3 || nr < 3 || nc >= J.W - 3 || nr >= J.H - 3)) {
          cj = j;
          crow = nr;
          ccol = nc;
        }
      }
    }
  };
  auto claim = [&](int cj, int crow, int ccol, int& h, bool& mine) __attribute__((always_inline)) {
    h = 0;
    mine = false;
    if (cj < 0) return;
    const int64_t gid = imgs[cj].base + (int64_t)crow * imgs[cj].W + ccol;
    h = hash_slot(gid);
    for (;;) {
      const unsigned long long old = atomicCAS((unsigned long long*)&hkey[h], ~0ull, (unsigned long long)gid);
      if (old == ~0ull) {
        mine = true;
        break;
      }
      if ((int64_t)old == gid) break;
      h = (h + 1) & (kHash - 1);
    }
  };
  auto push2 = [&](int f0, int k0, bool a0, int f1, int k1, bool a1, int& idx0, int& idx1)
                   __attribute__((always_inline)) {
    int cj0, r0, c0, cj1, r1, c1, h0, h1;
    bool m0, m1;
    project(f0, k0, a0, cj0, r0, c0);
    project(f1, k1, a1, cj1, r1, c1);
    claim(cj0, r0, c0, h0, m0);
    claim(cj1, r1, c1, h1, m1);
    const uint64_t mm0 = __ballot(m0), mm1 = __ballot(m1);
    const int fresh0 = __popcll(mm0), fresh = fresh0 + __popcll(mm1);
    idx0 = -1;
    idx1 = -1;
    if (nn + fresh > kProvCap) {
      overflow = true;
      return;
    }
    const uint64_t lt = (1ull << lane) - 1;
    if (m0) {
      idx0 = nn + __popcll(mm0 & lt);
      hval[h0] = (int16_t)idx0;
      nimg[idx0] = (uint16_t)cj0;
      nrc[idx0] = (r0 << 16) | c0;
    }
    if (m1) {
      idx1 = nn + fresh0 + __popcll(mm1 & lt);
      hval[h1] = (int16_t)idx1;
      nimg[idx1] = (uint16_t)cj1;
      nrc[idx1] = (r1 << 16) | c1;
    }
    nn += fresh;
    __builtin_amdgcn_wave_barrier();
    if (cj0 >= 0 && !m0) idx0 = hval[h0];
    if (cj1 >= 0 && !m1) idx1 = hval[h1];
  };
  constexpr int kBatch = 128;  // push slots of one batch
  while (nf > 0 && !overflow) {
    const int level_start = nn;
    for (int fi = 0; fi < nf && !overflow;) {
      // the next frontier pixels whose pushes fit the batch's 128 slots
      // together, in frontier order: their pushes are the same slots as
      // one pixel after another
      const int bf = fi + lane < nf ? front[fi + lane] : -1;
      const int bnov = bf >= 0 ? imgs[nimg[bf]].nov : 0;
      int cum = bnov;  // inclusive prefix over the lanes
      for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(cum, d);
        if (lane >= d) cum += t;
      }
      const int nb = __popcll(__ballot(bf >= 0 && cum <= kBatch));  // a prefix of the lanes
      if (nb == 0) {
        // a pixel with more pushes than a batch: 128 at a time
        const int f = __builtin_amdgcn_readfirstlane(bf);
        const int nov = __builtin_amdgcn_readfirstlane(bnov);
        if (nch + nov > kProvChild) {
          overflow = true;
          break;
        }
        if (lane == 0) {
          cbase[f] = (int16_t)nch;
          ccnt[f] = (int16_t)nov;
        }
        for (int k0 = 0; k0 < nov && !overflow; k0 += kBatch) {
          const bool a0 = k0 + lane < nov, a1 = k0 + 64 + lane < nov;
          int idx0, idx1;
          push2(f, k0 + lane, a0, f, k0 + 64 + lane, a1, idx0, idx1);
          if (overflow) break;
          if (a0) child[nch + k0 + lane] = (int16_t)idx0;
          if (a1) child[nch + k0 + 64 + lane] = (int16_t)idx1;
          __builtin_amdgcn_wave_barrier();
        }
        nch += nov;
        fi += 1;
      } else {
        const int total = __shfl(cum, nb - 1);
        if (nch + total > kProvChild) {
          overflow = true;
          break;
        }
        const int start = cum - bnov;
        if (lane < nb) {
          cbase[bf] = (int16_t)(nch + start);
          ccnt[bf] = (int16_t)bnov;
          for (int j = 0; j < bnov; ++j) {
            sown[start + j] = (uint8_t)lane;
            sk[start + j] = (uint8_t)j;
          }
        }
        __builtin_amdgcn_wave_barrier();
        const bool a0 = lane < total, a1 = 64 + lane < total;
        const int f0 = a0 ? front[fi + sown[lane]] : 0, k0 = a0 ? sk[lane] : 0;
        const int f1 = a1 ? front[fi + sown[64 + lane]] : 0, k1 = a1 ? sk[64 + lane] : 0;
        int idx0, idx1;
        push2(f0, k0, a0, f1, k1, a1, idx0, idx1);
        if (overflow) break;
        if (a0) child[nch + lane] = (int16_t)idx0;
        if (a1) child[nch + 64 + lane] = (int16_t)idx1;
        __builtin_amdgcn_wave_barrier();
        nch += total;
        fi += nb;
      }
      __syncthreads();
    }
    if (overflow) break;
    // test the new pixels (in parallel); the passing ones below the depth
    // cap form the next frontier, in pixel order
    int nnext = 0;
    for (int b = level_start; b < nn; b += 64) {
      const int i = b + lane;
      bool ok = false;
      if (i < nn) {
        float v[kVals];
        const int rc = nrc[i];
        ok = test_pixel(imgs[nimg[i]], mask, rc >> 16, rc & 0xffff, false, ref_pt, ref_n, prm, v);
        npass[i] = ok ? 1 : 0;
        if (ok)
          for (int k = 0; k < 3; ++k) nxyz[i][k] = v[k];
        cbase[i] = -1;
        ccnt[i] = 0;
      }
      const bool grow = ok && level + 2 < prm.max_traversal;
      const uint64_t m = __ballot(grow);
      if (grow) front[nnext + __popcll(m & ((1ull << lane) - 1))] = (int16_t)i;
      nnext += __popcll(m);
      __syncthreads();
    }
    nf = nnext;
    level++;
    __syncthreads();
  }
  // the stored graph: the passing pixels in order, and their pushes of
  // passing pixels
  int np = 0, nch2 = 0;
  if (!overflow) {
    for (int b = 0; b < nn; b += 64) {
      const int i = b + lane;
      const bool p = i < nn && npass[i];
      const uint64_t m = __ballot(p);
      if (i < nn) front[i] = p ? (int16_t)(np + __popcll(m & ((1ull << lane) - 1))) : (int16_t)-1;
      np += __popcll(m);
    }
    __syncthreads();
    for (int b = 0; b < nn; b += 64) {
      const int i = b + lane;
      int x = 0;
      if (i < nn && cbase[i] >= 0)
        for (int k = 0; k < ccnt[i]; ++k) {
          const int ch = child[cbase[i] + k];
          x += ch >= 0 && npass[ch];
        }
      int v = x;  // inclusive prefix over the wave
      for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(v, d);
        if (lane >= d) v += t;
      }
      if (i < nn) vfirst[i] = (int16_t)(nch2 + v - x);
      nch2 += __shfl(v, 63);
    }
    overflow = np > kNodeCap || nch2 > kChildCap;
  }
  if (overflow) {
    if (lane == 0) g.n[ws] = -1;
    return;
  }
  if (lane == 0) {
    // nodes and pushes start at multiples of 8 elements (k_flow loads them
    // 16 bytes at a time)
    const unsigned long long np8 = (unsigned long long)((np + 7) & ~7), nc8 = (unsigned long long)((nch2 + 7) & ~7);
    const unsigned long long no = atomicAdd(g.used, np8);
    const unsigned long long co = atomicAdd(g.used + 1, nc8);
    if (no + np8 > (unsigned long long)g.node_cap || co + nc8 > (unsigned long long)g.child_cap) {
      offs[0] = -1;
      atomicMin(g.no_room, ws);
      g.n[ws] = -2;
    } else {
      offs[0] = (long long)no;
      offs[1] = (long long)co;
    }
  }
  __syncthreads();
  if (offs[0] < 0) return;
  const int64_t no = offs[0], co = offs[1];
  for (int i = lane; i < nn; i += 64) {
    const int j = front[i];
    if (j < 0) continue;
    const int rc = nrc[i], im = nimg[i];
    const FusImgDev& I = imgs[im];
    g.gid[no + j] = I.base + (int64_t)(rc >> 16) * I.W + (rc & 0xffff);
    g.img[no + j] = im;
    int w = 0;
    if (cbase[i] >= 0)
      for (int k = 0; k < ccnt[i]; ++k) {
        const int ch = child[cbase[i] + k];
        if (ch >= 0 && npass[ch]) g.child[co + vfirst[i] + w++] = front[ch];
      }
    g.cbase[no + j] = cbase[i] >= 0 ? vfirst[i] : (int16_t)-1;
    g.ccnt[no + j] = (int16_t)w;
  }
  if (lane == 0) {
    g.n[ws] = np;
    g.nc[ws] = nch2;
    g.node_off[ws] = (int32_t)no;
    g.child_off[ws] = (int32_t)co;
  }
}

// The walks whose graphs pass each pixel (hold it as a node): counted, then
// (after an exclusive prefix sum turns the counts into starts) scattered
// into per-pixel lists.  One wave per seed.
__global__ void __launch_bounds__(256) k_count(Pool g, int32_t count, int32_t* __restrict__ cnt) {
  const int ws = blockIdx.x * 4 + threadIdx.x / 64, lane = threadIdx.x % 64;
  if (ws >= count) return;
  const int n = g.n[ws];
  if (n <= 0) return;
  const int64_t no = g.node_off[ws];
  for (int i = lane; i < n; i += 64) atomicAdd(cnt + g.gid[no + i], 1);
}

__global__ void __launch_bounds__(256) k_scatter(Pool g, int32_t count, int32_t* __restrict__ end,
                                                 int32_t* __restrict__ lists) {
  const int ws = blockIdx.x * 4 + threadIdx.x / 64, lane = threadIdx.x % 64;
  if (ws >= count) return;
  const int n = g.n[ws];
  if (n <= 0) return;
  const int64_t no = g.node_off[ws];
  for (int i = lane; i < n; i += 64) lists[atomicAdd(end + g.gid[no + i], 1)] = ws;
}

// Exclusive prefix sums of int32 counts (totals stay below 2^31): per-tile
// sums, a scan of the tile sums in one block, then each tile's scan.
constexpr int kScanThreads = 256, kScanItems = 8, kScanTile = kScanThreads * kScanItems;

__device__ int32_t block_exclusive(int32_t x, int32_t* sh, int32_t* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int32_t v = x;
  for (int d = 1; d < 64; d <<= 1) {
    const int32_t t = __shfl_up(v, d);
    if (lane >= d) v += t;
  }
  if (lane == 63) sh[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t s = 0;
    for (int w = 0; w < kScanThreads / 64; ++w) {
      const int32_t t = sh[w];
      sh[w] = s;
      s += t;
    }
    sh[kScanThreads / 64] = s;
  }
  __syncthreads();
  const int32_t r = sh[wid] + v - x;
  *total = sh[kScanThreads / 64];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(const int32_t* __restrict__ in, int64_t n,
                                                             int32_t* __restrict__ sums) {
  __shared__ int32_t sh[kScanThreads / 64 + 1];
  const int64_t b = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int32_t s = 0;
  for (int k = 0; k < kScanItems; ++k)
    if (b + k < n) s += in[b + k];
  int32_t tot;
  (void)block_exclusive(s, sh, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_sums(int32_t* __restrict__ sums, int64_t ntiles,
                                                            int32_t* __restrict__ total) {
  __shared__ int32_t sh[kScanThreads / 64 + 1];
  int32_t carry = 0;
  for (int64_t b0 = 0; b0 < ntiles; b0 += kScanTile) {
    const int64_t b = b0 + (int64_t)threadIdx.x * kScanItems;
    int32_t x[kScanItems], s = 0;
    for (int k = 0; k < kScanItems; ++k) {
      x[k] = b + k < ntiles ? sums[b + k] : 0;
      s += x[k];
    }
    int32_t tot;
    int32_t run = carry + block_exclusive(s, sh, &tot);
    for (int k = 0; k < kScanItems; ++k)
      if (b + k < ntiles) {
        sums[b + k] = run;
        run += x[k];
      }
    carry += tot;
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_apply(const int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                             int64_t n, const int32_t* __restrict__ sums) {
  __shared__ int32_t sh[kScanThreads / 64 + 1];
  const int64_t b = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int32_t x[kScanItems], s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    x[k] = b + k < n ? in[b + k] : 0;
    s += x[k];
  }
  int32_t tot;
  int32_t run = sums[blockIdx.x] + block_exclusive(s, sh, &tot);
  for (int k = 0; k < kScanItems; ++k)
    if (b + k < n) {
      out[b + k] = run;
      run += x[k];
    }
}

// The dataflow's state, for one segment of a window's walks (the window's
// walks between two that run alone).
struct Flow {
  const int32_t* tickets;  // the segment's seeds with a graph, in seed order
  int32_t n_tickets;
  int32_t seg_lo;          // the segment's first seed: earlier walks have finished
  int32_t* queue;          // [n_tickets] ready walks (seeds), -1 until written
  int32_t* counters;       // [0] slots claimed [1] slots written [2] walks finished [3] stop [4] longest chain
  int32_t* pending;        // [seed] earlier walks it depends on that have not finished
  int32_t* npred;          // [seed] earlier walks it depends on
  int32_t* preds;          // [node] those walks, at the graph's node offset
  int32_t* nsucc;          // [seed] later walks that depend on it
  int32_t* succ_end;       // [seed] end of its successor list (the prefix sum, advanced by the scatter)
  int32_t* succ;           // successor lists
  int32_t* depth;          // [seed] 1 + the longest chain of walks it depends on
  int32_t* res_cnt;        // [seed] pixels accepted
  int16_t* res_nodes;      // [node] accepted nodes in visiting order, at the graph's node offset
  const int32_t* cnt;      // [pixel] walks whose graphs hold the pixel
  const int32_t* end;      // [pixel] end of its list
  const int32_t* lists;
  long long wait_ticks;    // how long no walk may finish before the waves give up
  int64_t seed_base;       // the image's first global pixel id
  int32_t seed_w;          // its width
  int32_t window_first;    // the window's first seed
};

// A set of up to kNodeCap = 256 nodes in four 64-bit words (named, not an
// array: a dynamically indexed array would live in scratch memory).
struct NodeSet {
  uint64_t w0, w1, w2, w3;
  __device__ bool has(int i) const {  // any lane's i: the word by two selects on its bits, one shift
    const bool hi = ((uint32_t)i & 128u) != 0, odd = ((uint32_t)i & 64u) != 0;
    const uint64_t a = odd ? w1 : w0, b = odd ? w3 : w2;
    const uint64_t w = hi ? b : a;
    return ((w >> ((uint32_t)i & 63u)) & 1ull) != 0;
  }
  __device__ void remove(int i) {  // a wave-uniform i
    const uint64_t b = ~(1ull << (i & 63));
    const int q = i >> 6;
    w0 &= q == 0 ? b : ~0ull;
    w1 &= q == 1 ? b : ~0ull;
    w2 &= q == 2 ? b : ~0ull;
    w3 &= q == 3 ? b : ~0ull;
  }
};

// The depth-first walk of Fuse over a graph in LDS, run by the whole wave:
// the accepted nodes in visiting order (accl), their count returned.
// `ok` holds the nodes not masked; the available set (not masked and not yet
// taken) lives in scalar registers.
// A push of a node that is masked or already taken would end in `continue`
// when popped (a node is never untaken), so it is not stacked; the others
// are stacked in push order, one lane per push.  Every node is taken at
// most once, so at most (pushes + 1) entries are stacked.
//   * The last entry a node pushes is the next one popped, and it is neither
//     masked nor taken (nothing is taken between its push and its pop): it is
//     taken straight from the pushing lane's registers.
//   * Otherwise the stack's top 64 entries are read at once and the topmost
//     one neither masked nor taken is popped; the ones above it would have
//     been popped and skipped one by one.
//   * A node's first push and push count come from the lanes' registers:
//     info[q] of lane l holds node 64 q + l as (first << 16) | count.
__device__ __attribute__((always_inline)) inline int replay(const FusParams& prm, const NodeSet ok,
                                                             const int32_t info[4], const int16_t* child,
                                                             int32_t* stack, int16_t* accl) {
  const int lane = threadIdx.x % 64;
  NodeSet avail = ok;
  int sp = 0, na = 0;
  int e = 0;                 // the entry to visit next: node 0 at depth 0
  bool have = avail.has(0);
  for (;;) {
    if (!have) {
      while (sp > 0) {
        const int k = sp - 1 - lane;
        int x = 0;
        bool v = false;
        if (k >= 0) {
          x = stack[k];
          v = avail.has(x & 0xffff);
        }
        const uint64_t m = __ballot(v);
        if (m) {
          const int l = __ffsll((long long)m) - 1;
          e = __builtin_amdgcn_readlane(x, l);
          sp -= l + 1;
          have = true;
          break;
        }
        sp = sp > 64 ? sp - 64 : 0;
      }
      if (!have) break;
    }
    have = false;
    const int node = e & 0xffff, d = e >> 16;
    avail.remove(node);
    if (lane == 0) accl[na] = (int16_t)node;
    ++na;
    if (na >= prm.max_num_pixels) break;
    if (d + 1 >= prm.max_traversal) continue;
    const int q = node >> 6;
    const int iv = q == 0 ? info[0] : (q == 1 ? info[1] : (q == 2 ? info[2] : info[3]));
    const int nfo = __builtin_amdgcn_readlane(iv, node & 63);
    const int b = nfo >> 16, cnt = nfo & 0xffff;
    const int val0 = (d + 1) << 16;
    for (int k0 = 0; k0 < cnt; k0 += 64) {
      const int k = k0 + lane;
      int c = 0;
      bool push = false;
      if (k < cnt) {
        c = child[b + k];
        push = avail.has(c);
      }
      const uint64_t m = __ballot(push);
      if (push) stack[sp + __popcll(m & ((1ull << lane) - 1))] = val0 | c;
      sp += __popcll(m);
      if (m) {
        e = __builtin_amdgcn_readlane(val0 | c, 63 - __clzll((long long)m));
        have = true;
      }
    }
    if (have) --sp;  // the last push, popped from the registers
    __builtin_amdgcn_wave_barrier();  // the wave's LDS writes precede its next reads
  }
  return na;
}

__device__ inline bool masked_now(const uint32_t* mask, int64_t gid) {
  return (__hip_atomic_load(mask + (gid >> 5), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> (gid & 31)) & 1u;
}

// The walks a walk depends on: for each pixel of its graph, the last earlier
// walk of the segment whose graph holds it (that one depended in turn on the
// ones before it).  Stored once each; each is counted as having one more
// successor.  One wave per walk of the segment.
constexpr int kPredWaves = 4;
__global__ void __launch_bounds__(64 * kPredWaves) k_preds(Pool g, Flow f) {
  __shared__ int32_t pl[kPredWaves][kNodeCap];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int t = blockIdx.x * kPredWaves + wave;
  if (t >= f.n_tickets) return;
  const int ws = f.tickets[t];
  const int nn = g.n[ws];
  const int64_t no = g.node_off[ws];
  int32_t* P = pl[wave];
  for (int i = lane; i < nn; i += 64) {
    const int64_t gid = g.gid[no + i];
    const int e = f.end[gid], c = f.cnt[gid];
    int pred = -1;
    for (int j = e - c; j < e; ++j) {
      const int v = f.lists[j];
      if (v < ws && v > pred) pred = v;
    }
    P[i] = pred >= f.seg_lo ? pred : -1;
  }
  __builtin_amdgcn_wave_barrier();
  int np = 0;
  for (int b = 0; b < nn; b += 64) {
    const int i = b + lane;
    const int p = i < nn ? P[i] : -1;
    bool first = p >= 0;
    for (int j = 0; j < i && first; ++j) first = P[j] != p;
    const uint64_t m = __ballot(first);
    if (first) {
      f.preds[no + np + __popcll(m & ((1ull << lane) - 1))] = p;
      atomicAdd(f.nsucc + p, 1);
    }
    np += __popcll(m);
  }
  if (lane == 0) {
    f.npred[ws] = np;
    f.pending[ws] = np;
  }
}

// Successor lists (after a prefix sum of nsucc into succ_end), and the walks
// ready at the start in the queue.
__global__ void __launch_bounds__(256) k_succ(Pool g, Flow f) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= f.n_tickets) return;
  const int ws = f.tickets[t];
  const int np = f.npred[ws];
  const int64_t no = g.node_off[ws];
  for (int k = 0; k < np; ++k) {
    const int p = f.preds[no + k];
    f.succ[atomicAdd(f.succ_end + p, 1)] = ws;
  }
  if (np == 0) f.queue[atomicAdd(f.counters + 1, 1)] = ws;
}

// The dataflow: persistent waves take queue slots in order; a slot's walk is
// written once all the walks it depends on have finished.  A wave replays
// its walk over the graph against the masks, sets the accepted pixels' mask
// bits, publishes them (release fence), and then counts itself off at each
// successor: the first one it was the last predecessor of, it runs next
// itself; the others it queues.  All reads of what other waves wrote (queue
// slots, mask words, depths) are agent-scope atomic loads.  Every walk
// becomes ready (a walk depends only on earlier walks) and is either run by
// the wave that made it ready or written to a slot, so a wave waiting on a
// slot gets a walk or sees every walk finished; it gives up only when no walk
// anywhere finished for wait_ticks.
__global__ void __launch_bounds__(64) k_flow(FusParams prm, Pool g, uint32_t* mask, Flow f) {
  __shared__ __attribute__((aligned(16))) int64_t ngid[kNodeCap];
  __shared__ __attribute__((aligned(16))) int16_t nbase[kNodeCap];
  __shared__ __attribute__((aligned(16))) int16_t ncnt[kNodeCap];
  __shared__ __attribute__((aligned(16))) int16_t child[kChildCap];
  __shared__ int32_t stack[kChildCap + 1];
  __shared__ int16_t accl[kNodeCap];
  const int lane = threadIdx.x;
  int next = -1;  // a walk this wave made ready and runs next, past the queue (wave-uniform)
  for (;;) {
    int ws = next;
    next = -1;
    if (ws < 0) {
    // the slot: every lane takes part in the atomic (only lane 0 adds), the
    // result is wave-uniform.  A lane-0-only branch here lets the compiler
    // run lane 0 and the other lanes round different loops.
    const int h = __builtin_amdgcn_readfirstlane(atomicAdd(f.counters, lane == 0 ? 1 : 0));
    if (h >= f.n_tickets) break;
    ws = __hip_atomic_load(f.queue + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (ws < 0) {
      // Most waves wait here at any time: each polls only its own slot, and
      // looks at the shared counters (stop flag, walks finished) once every
      // kPollCheck polls, so that the waiting adds little to the memory
      // traffic the running walks wait on.
      long long t0 = wall_clock64();
      int seen = __hip_atomic_load(f.counters + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (int it = 1; ws < 0; ++it) {
        __builtin_amdgcn_s_sleep(kPollSleep);
        ws = __hip_atomic_load(f.queue + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (ws >= 0 || it % kPollCheck != 0) continue;
        if (__hip_atomic_load(f.counters + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        const int fin = __hip_atomic_load(f.counters + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (fin >= f.n_tickets) break;  // every walk has finished: this slot stays empty
        if (fin != seen) {
          seen = fin;
          t0 = wall_clock64();
        } else if (wall_clock64() - t0 > f.wait_ticks) {
          __hip_atomic_store(f.counters + 3, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
      }
    }
    ws = __builtin_amdgcn_readfirstlane(ws);
    if (ws < 0) break;  // stopped, or every walk finished
    }
    const int nn = g.n[ws], nch = g.nc[ws];
    const int64_t no = g.node_off[ws], co = g.child_off[ws];
    const int np = f.npred[ws], ns = f.nsucc[ws], se = f.succ_end[ws];
    // what the end of the walk needs, loaded early: the longest chain behind
    // it (a statistic: a predecessor's depth is stored without ordering
    // against its countdown and may still read as the segment's 0, so the
    // chain counted is a lower bound) and its successors
    int dep = 0;
    for (int k = lane; k < np; k += 64)
      dep = max(dep, __hip_atomic_load(f.depth + f.preds[no + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    const int sw0 = lane < ns ? f.succ[se - ns + lane] : -1;
    // the graph into LDS, every load in flight at once and together with
    // the seed's mask word: a graph's nodes and pushes start at multiples of
    // 8 elements, so 16-byte loads cover them (the tails past nn and nch are
    // never read)
    const int ol = lane & 31;
    const bool lg0 = 2 * lane < nn, lg1 = 2 * (64 + lane) < nn, lc = 8 * ol < nn;
    const bool lh0 = 8 * lane < nch, lh1 = 8 * (64 + lane) < nch, lh2 = 8 * (128 + lane) < nch,
               lh3 = 8 * (192 + lane) < nch;
    const uint4 z4 = {0u, 0u, 0u, 0u};
    const uint4 g0 = lg0 ? *(const uint4*)(g.gid + no + 2 * lane) : z4;
    const uint4 g1 = lg1 ? *(const uint4*)(g.gid + no + 2 * (64 + lane)) : z4;
    const uint4 c0 = lc ? *(const uint4*)((lane < 32 ? g.cbase : g.ccnt) + no + 8 * ol) : z4;
    const uint4 h0 = lh0 ? *(const uint4*)(g.child + co + 8 * lane) : z4;
    const uint4 h1 = lh1 ? *(const uint4*)(g.child + co + 8 * (64 + lane)) : z4;
    const uint4 h2 = lh2 ? *(const uint4*)(g.child + co + 8 * (128 + lane)) : z4;
    const uint4 h3 = lh3 ? *(const uint4*)(g.child + co + 8 * (192 + lane)) : z4;
    // the seed (node 0, its pixel from the seed's number): taken by an
    // earlier walk, this walk accepts nothing.  The masks are final for the
    // walk's pixels: every earlier walk that could take one has finished.
    const int sd = f.window_first + ws;
    const int64_t seed_gid = f.seed_base + (int64_t)(3 + (sd / prm.seed_cols) * prm.step) * f.seed_w +
                             (3 + (sd % prm.seed_cols) * prm.step);
    const bool seed_free = !masked_now(mask, seed_gid);
    if (lg0) *(uint4*)(ngid + 2 * lane) = g0;
    if (lg1) *(uint4*)(ngid + 2 * (64 + lane)) = g1;
    if (lc) *(uint4*)((lane < 32 ? nbase : ncnt) + 8 * ol) = c0;
    if (lh0) *(uint4*)(child + 8 * lane) = h0;
    if (lh1) *(uint4*)(child + 8 * (64 + lane)) = h1;
    if (lh2) *(uint4*)(child + 8 * (128 + lane)) = h2;
    if (lh3) *(uint4*)(child + 8 * (192 + lane)) = h3;
    int na = 0;
    if (seed_free) {
      __builtin_amdgcn_wave_barrier();
      int32_t info[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 64 * r + lane;
        info[r] = i < nn ? (int32_t)(((uint32_t)(uint16_t)nbase[i] << 16) | (uint16_t)ncnt[i]) : 0;
      }
      NodeSet ok;
      const bool m0 = lane < nn && masked_now(mask, ngid[lane]);
      const bool m1 = 64 + lane < nn && masked_now(mask, ngid[64 + lane]);
      const bool m2 = 128 + lane < nn && masked_now(mask, ngid[128 + lane]);
      const bool m3 = 192 + lane < nn && masked_now(mask, ngid[192 + lane]);
      ok.w0 = __ballot(lane < nn && !m0);
      ok.w1 = __ballot(64 + lane < nn && !m1);
      ok.w2 = __ballot(128 + lane < nn && !m2);
      ok.w3 = __ballot(192 + lane < nn && !m3);
      na = replay(prm, ok, info, child, stack, accl);
      __builtin_amdgcn_wave_barrier();
      for (int j = lane; j < na; j += 64) {
        const int node = accl[j];
        const int64_t gid = ngid[node];
        __hip_atomic_fetch_or(mask + (gid >> 5), 1u << (gid & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        f.res_nodes[no + j] = (int16_t)node;
      }
    }
    // publish the mask bits (a walk that took no pixel wrote none)
    if (na > 0) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    // the successors this walk was the last one for: the first one this
    // wave runs next itself (no queue slot and no poll on its way), the
    // others go to the queue.  The queue then fills fewer slots than there
    // are walks, and the waves waiting on the rest leave once every walk has
    // finished.
    for (int k0 = 0; k0 < ns; k0 += 64) {
      const int sw = k0 == 0 ? sw0 : (k0 + lane < ns ? f.succ[se - ns + k0 + lane] : -1);
      const bool ready = sw >= 0 && atomicSub(f.pending + sw, 1) == 1;
      const uint64_t rm = __ballot(ready);
      const int kl = next < 0 && rm ? __ffsll((long long)rm) - 1 : -1;
      if (ready && lane != kl)
        __hip_atomic_store(f.queue + atomicAdd(f.counters + 1, 1), sw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (kl >= 0) next = __builtin_amdgcn_readlane(sw, kl);
    }
    // the statistic after the countdown: its loads are then waited for
    // together with the countdown's, not before it
    for (int d = 32; d > 0; d >>= 1) dep = max(dep, __shfl_xor(dep, d));
    if (lane == 0) {
      f.res_cnt[ws] = na;
      __hip_atomic_store(f.depth + ws, dep + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      atomicMax(f.counters + 4, dep + 1);
    }
    if (lane == 0) atomicAdd(f.counters + 2, 1);
  }
}

// Mask bits of the pixels a walk run alone accepted.
__global__ void __launch_bounds__(256) k_commit_alone(const WalkRec* __restrict__ rec, const int64_t* __restrict__ gids,
                                                      uint32_t* __restrict__ mask) {
  const int na = rec->n_acc;
  for (int j = blockIdx.x * 256 + threadIdx.x; j < na; j += gridDim.x * 256) {
    const int64_t g = gids[j];
    atomicOr(mask + (g >> 5), 1u << (g & 31));
  }
}

// The point of a walk (fusion.cc:2098-2180), one wave per walk.  The medians
// (internal::Median, fusion.cc:47-60: the element of rank n/2 and, for an
// even count, the mean with the element of rank n/2 - 1) come from each
// entry's rank, counted in LDS; the mean gradient is summed in visiting
// order.  A graph walk's values are recomputed from its pixels and its seed
// (the same function, so the same bits as while building the graph); a walk
// run alone has them in its record, and one with more entries than LDS
// holds is read from there directly.
constexpr int kPtWaves = 2;
constexpr int kLdsAcc = kNodeCap;           // accepted pixels staged in LDS per wave
constexpr int kMaxImages = 65536;           // fusion takes at most this many images
constexpr int kImgWords = kMaxImages / 32;  // visibility bitmap per wave

__device__ inline float pick(bool c0, float x0, bool c1, float x1) {
  const uint64_t m0 = __ballot(c0), m1 = __ballot(c1);
  if (m0) return __shfl(x0, __ffsll((long long)m0) - 1);
  return __shfl(x1, __ffsll((long long)m1) - 1);
}

struct GraphSrc {  // accepted pixels of a graph walk, values recomputed
  const FusImgDev* imgs;
  const int64_t* gid;
  const int32_t* img;
  const int16_t* nodes;
  int64_t no;
  const FusParams* prm;
  float ref_pt[3], ref_n[3];
  __device__ void fill(int i, float* v, int32_t* im) const {
    const int node = nodes[no + i];
    const int m = img[no + node];
    const FusImgDev& I = imgs[m];
    const int64_t p = gid[no + node] - I.base;
    test_pixel(I, nullptr, (int)(p / I.W), (int)(p % I.W), node == 0, ref_pt, ref_n, *prm, v, false);
    *im = m;
  }
  static constexpr bool kLdsOnly = true;  // at most kNodeCap = kLdsAcc pixels
};

struct AccSrc {  // accepted pixels of a walk run alone
  const AccEntry* a;
  __device__ void fill(int i, float* v, int32_t* im) const {
    for (int k = 0; k < kVals; ++k) v[k] = a[i].v[k];
    *im = a[i].img;
  }
  __device__ float v(int i, int k) const { return a[i].v[k]; }
  static constexpr bool kLdsOnly = false;
};

// Called by every thread of the block; the `active` waves form their walk's
// point.  Returns whether a point was formed (uniform over the wave).
template <class Src>
__device__ bool form_point(bool active, int na, const Src& src, const FusParams& prm, float (*lv)[kLdsAcc],
                           uint32_t* bits, PointRec* out, int32_t* vis_out) {
  const int lane = threadIdx.x % 64;
  const bool in_lds = na <= kLdsAcc;
  const int words = (prm.n_images + 31) / 32;
  for (int k = lane; k < words; k += 64) bits[k] = 0u;
  __syncthreads();
  if (active) {
    for (int i = lane; i < na; i += 64) {
      float x[kVals];
      int32_t m;
      src.fill(i, x, &m);
      if (in_lds)
        for (int k = 0; k < kVals; ++k) lv[k][i] = x[k];
      atomicOr(&bits[m >> 5], 1u << (m & 31));
    }
  }
  __syncthreads();
  if (!active) return false;
  auto val = [&](int i, int k) -> float {
    if constexpr (Src::kLdsOnly)
      return lv[k][i];
    else
      return in_lds ? lv[k][i] : src.v(i, k);
  };
  // rank-based median of value k over entries [b, na)
  auto median = [&](int k, int b) -> float {
    const int cnt = na - b, mid = cnt / 2;
    float hi = 0.0f, lo = 0.0f;
    for (int base = b; base < na; base += 128) {
      const int i0 = base + lane, i1 = base + 64 + lane;
      float x0 = 0.0f, x1 = 0.0f;
      int l0 = 0, e0 = 0, l1 = 0, e1 = 0;
      if (i0 < na) x0 = val(i0, k);
      if (i1 < na) x1 = val(i1, k);
      for (int j = b; j < na; ++j) {
        const float y = val(j, k);
        l0 += y < x0;
        e0 += y <= x0;
        l1 += y < x1;
        e1 += y <= x1;
      }
      const bool h0 = i0 < na && l0 <= mid && mid < e0, h1 = i1 < na && l1 <= mid && mid < e1;
      const bool o0 = i0 < na && l0 <= mid - 1 && mid - 1 < e0, o1 = i1 < na && l1 <= mid - 1 && mid - 1 < e1;
      if (__ballot(h0) | __ballot(h1)) hi = pick(h0, x0, h1, x1);
      if (__ballot(o0) | __ballot(o1)) lo = pick(o0, x0, o1, x1);
    }
    return (cnt % 2 == 0) ? (hi + lo) / 2.0f : hi;
  };
  // visibility: the distinct images in ascending order, from the bitmap
  // (each lane a contiguous run of words, then a prefix over the lanes)
  const int per = (words + 63) / 64;
  const int k0 = min(lane * per, words), k1 = min(k0 + per, words);
  int mine = 0;
  for (int k = k0; k < k1; ++k) mine += __popc(bits[k]);
  int before = 0, nvis = 0;
  for (int l = 0; l < 64; ++l) {
    const int c = __shfl(mine, l);
    before += l < lane ? c : 0;
    nvis += c;
  }
  if (nvis < prm.min_num_visible) return false;
  PointRec r{};
  r.nvis = nvis;
  bool ok = true;
  for (int c = 0; c < 3; ++c) {
    r.xyz[c] = median(c, 0);
    if (isnan(r.xyz[c]) || isinf(r.xyz[c])) ok = false;
  }
  if (ok && prm.with_normal) {
    const float nx = median(3, 0), ny = median(4, 0), nz = median(5, 0);
    const float norm = sqrtf(nx * nx + (ny * ny + nz * nz));
    if (norm < FLT_EPSILON) {
      ok = false;
    } else {
      r.nrm[0] = nx / norm;
      r.nrm[1] = ny / norm;
      r.nrm[2] = nz / norm;
    }
  }
  if (ok) {
    for (int c = 0; c < 3; ++c) {
      const float m = roundf(median(6 + c, 0));
      r.rgb[c] = m < 0.0f ? 0.0f : (m > 255.0f ? 255.0f : m);
    }
    float m_conf = 0.0f;
    for (int j = 0; j < na; ++j) m_conf += val(j, 9);
    m_conf /= (float)na;
    if (nvis >= 3) m_conf = 1.0f;
    m_conf = 1.0f - m_conf;
    // the depth errors are those of every accepted pixel but the seed
    if (na < 2) {
      ok = false;
    } else {
      const float med = median(10, 1);
      float score = (float)((double)med / prm.max_depth_error);
      score = fminf(0.6f * score + 0.4f * m_conf, 1.0f);
      if (isnan(score) || isinf(score)) ok = false;
      r.score = score;
    }
  }
  if (!ok) return false;
  int pos = before;
  for (int k = k0; k < k1; ++k)
    for (uint32_t m = bits[k]; m; m &= m - 1) vis_out[pos++] = 32 * k + __ffs(m) - 1;
  if (lane == 0) *out = r;
  return true;
}

// The points of the window's graph walks: a PointRec and a flag per seed,
// the visibility at the graph's node offset (a point sees at most as many
// images as its walk accepted pixels).
__global__ void __launch_bounds__(64 * kPtWaves) k_points(const FusImgDev* __restrict__ imgs, FusParams prm,
                                                          int32_t window_first, Pool g, Flow f,
                                                          PointRec* __restrict__ pts, int32_t* __restrict__ has,
                                                          int32_t* __restrict__ vis_pool,
                                                          unsigned long long* __restrict__ kept) {
  __shared__ float lv[kPtWaves][kVals][kLdsAcc];
  __shared__ uint32_t bits[kPtWaves][kImgWords];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int t = blockIdx.x * kPtWaves + wave;
  const int ws = __builtin_amdgcn_readfirstlane(t < f.n_tickets ? f.tickets[t] : 0);
  const int na = __builtin_amdgcn_readfirstlane(t < f.n_tickets ? f.res_cnt[ws] : 0);
  if (lane == 0 && na > 0) atomicAdd(kept, 1ull);
  const bool active = na >= prm.min_num_pixels;
  const int64_t no = active ? g.node_off[ws] : 0;
  GraphSrc src{imgs, g.gid, g.img, f.res_nodes, no, &prm, {0, 0, 0}, {0, 0, 0}};
  if (active) {  // the seed's point and normal
    const int s = window_first + ws;
    float v[kVals];
    test_pixel(imgs[prm.cur], nullptr, 3 + (s / prm.seed_cols) * prm.step, 3 + (s % prm.seed_cols) * prm.step, true,
               nullptr, nullptr, prm, v, false);
    for (int k = 0; k < 3; ++k) {
      src.ref_pt[k] = v[k];
      src.ref_n[k] = v[3 + k];
    }
  }
  if (form_point(active, na, src, prm, lv[wave], bits[wave], pts + ws, vis_pool + no) && lane == 0) has[ws] = 1;
}

// The point of a walk run alone (the block's first wave).
__global__ void __launch_bounds__(64 * kPtWaves) k_points_alone(FusParams prm, const WalkRec* __restrict__ rec,
                                                                const AccEntry* __restrict__ acc,
                                                                PointRec* __restrict__ out, int32_t* __restrict__ has,
                                                                int32_t* __restrict__ vis_out) {
  __shared__ float lv[kPtWaves][kVals][kLdsAcc];
  __shared__ uint32_t bits[kPtWaves][kImgWords];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int na = __builtin_amdgcn_readfirstlane(wave == 0 ? rec->n_acc : 0);
  const bool active = wave == 0 && na >= prm.min_num_pixels;
  const AccSrc src{acc};
  const bool made = form_point(active, na, src, prm, lv[wave], bits[wave], out, vis_out);
  if (wave == 0 && lane == 0) *has = made ? 1 : 0;
}

// The window's points in seed order: PointRec, seed and visibility count at
// the point's rank among the seeds with a point, then the visibility lists.
__global__ void __launch_bounds__(256) k_gather_points(const PointRec* __restrict__ pts, const int32_t* __restrict__ has,
                                                       const int32_t* __restrict__ rank, int32_t count,
                                                       PointRec* __restrict__ out, int32_t* __restrict__ out_seed,
                                                       int32_t* __restrict__ out_nvis) {
  const int ws = blockIdx.x * 256 + threadIdx.x;
  if (ws >= count || !has[ws]) return;
  const int k = rank[ws];
  out[k] = pts[ws];
  out_seed[k] = ws;
  out_nvis[k] = pts[ws].nvis;
}

__global__ void __launch_bounds__(256) k_gather_vis(const int32_t* __restrict__ out_seed,
                                                    const int32_t* __restrict__ out_nvis,
                                                    const int32_t* __restrict__ vis_at, int32_t n_points, Pool g,
                                                    const int32_t* __restrict__ vis_pool,
                                                    int32_t* __restrict__ out_vis) {
  const int k = blockIdx.x * 4 + threadIdx.x / 64, lane = threadIdx.x % 64;
  if (k >= n_points) return;
  const int64_t no = g.node_off[out_seed[k]];
  const int nv = out_nvis[k], at = vis_at[k];
  for (int i = lane; i < nv; i += 64) out_vis[at + i] = vis_pool[no + i];
}

// The window's seeds with a graph (tickets, in seed order) and the seeds
// whose walks run alone (barriers, each with the number of tickets before
// it), listed on the device: flags, exclusive prefix sums, a scatter.  Seeds
// from the first one that found the pool full (no_room) on are left out.
__global__ void __launch_bounds__(256) k_ticket_flags(const int32_t* __restrict__ n, int32_t count,
                                                      const int32_t* __restrict__ no_room, int32_t* __restrict__ ft,
                                                      int32_t* __restrict__ fb) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= count) return;
  const int v = k < *no_room ? n[k] : 0;
  ft[k] = v > 0 ? 1 : 0;
  fb[k] = v == -1 ? 1 : 0;
}
__global__ void __launch_bounds__(256) k_ticket_scatter(const int32_t* __restrict__ n, int32_t count,
                                                        const int32_t* __restrict__ no_room,
                                                        const int32_t* __restrict__ pt, const int32_t* __restrict__ pb,
                                                        int32_t* __restrict__ tickets, int32_t* __restrict__ bars) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= count || k >= *no_room) return;
  const int v = n[k];
  if (v > 0) tickets[pt[k]] = k;
  if (v == -1) {
    bars[2 * pb[k]] = k;
    bars[2 * pb[k] + 1] = pt[k];
  }
}

__global__ void __launch_bounds__(256) k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = v;
}
__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = 0u;
}
__global__ void k_window_reset(unsigned long long* used, int32_t* no_room, int32_t count, int32_t* status) {
  used[0] = used[1] = 0ull;
  *no_room = count;
  status[0] = status[1] = status[2] = 0;
}
// a segment's stop flag and longest chain into the window's status
__global__ void k_segment_end(const int32_t* counters, int32_t* status) {
  status[0] |= counters[3];
  status[2] += counters[4];
}
__global__ void k_zero_u64(unsigned long long* p) { *p = 0ull; }

unsigned fill_grid(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }

// A device buffer freed with its owner.
struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  pm_status get(size_t b) {
    if (b <= bytes && p) return PM_OK;
    if (p) PM_HIP(hipFree(p));
    p = nullptr;
    bytes = 0;
    PM_HIP(hipMalloc(&p, std::max<size_t>(b, 1)));
    bytes = std::max<size_t>(b, 1);
    return PM_OK;
  }
  template <class T>
  T* as() const {
    return (T*)p;
  }
};

// Exclusive prefix sum of n int32 values into out; the total lands in *total.
void scan(const int32_t* in, int32_t* out, int64_t n, int32_t* sums, int32_t* total, hipStream_t st) {
  const int64_t tiles = std::max<int64_t>(1, (n + kScanTile - 1) / kScanTile);
  hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(kScanThreads), 0, st, in, n, sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, st, sums, tiles, total);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)tiles), dim3(kScanThreads), 0, st, in, out, n, (const int32_t*)sums);
}

// FindNextImage (fusion.cc:697-727).
int find_next(const pm_fusion_image* im, int n, const std::vector<uint8_t>& fused, int prev) {
  for (int pass = 0; pass < 2; ++pass)
    for (int k = 0; k < im[prev].num_overlapping; ++k) {
      const int j = im[prev].overlapping[k];
      if (j < 0 || j >= n) continue;
      if (im[j].used && !fused[j] && (pass == 1 || !im[j].is_rig)) return j;
    }
  for (int pass = 0; pass < 2; ++pass)
    for (int j = 0; j < n; ++j)
      if (im[j].used && !fused[j] && (pass == 1 || !im[j].is_rig)) return j;
  return -1;
}

}  // namespace

struct pm_fusion {
  std::vector<float> xyz, normal, score;
  std::vector<uint8_t> rgb;
  std::vector<int64_t> vis_offsets{0};
  std::vector<int32_t> vis_images;
  pm_fusion_stats stats{};
  void add(const PointRec& r, const int32_t* vis) {
    for (int c = 0; c < 3; ++c) {
      xyz.push_back(r.xyz[c]);
      normal.push_back(r.nrm[c]);
      rgb.push_back((uint8_t)r.rgb[c]);
    }
    score.push_back(r.score);
    vis_images.insert(vis_images.end(), vis, vis + r.nvis);
    vis_offsets.push_back((int64_t)vis_images.size());
  }
};

PM_API pm_status pm_fusion_options_init(pm_fusion_options* o) {
  if (!o) return fail(PM_ERR_INVALID_ARG, "null options");
  o->step_size = 1;
  o->min_num_pixels = 5;
  o->max_num_pixels = 10000;
  o->min_num_visible_images = 3;
  o->max_traversal_depth = 100;
  o->with_normal = 1;
  o->max_reproj_error = 2.0;
  o->max_depth_error = (double)0.01f;
  o->max_normal_error = (double)10.0f;
  return PM_OK;
}

PM_API pm_status pm_fusion_run(pm_engine* e, const pm_fusion_options* o, int32_t n, const pm_fusion_image* im,
                               pm_fusion** out) {
  if (!e || !o || !im || !out || n < 1) return fail(PM_ERR_INVALID_ARG, "invalid fusion arguments");
  *out = nullptr;
  if (o->step_size < 1 || o->min_num_pixels < 2 || o->max_num_pixels < 1 || o->max_traversal_depth < 1 ||
      !(o->max_reproj_error >= 0.0) || !(o->max_depth_error > 0.0) || !(o->max_normal_error >= 0.0))
    return fail(PM_ERR_INVALID_ARG, "invalid fusion options (min_num_pixels >= 2, step_size >= 1, ...)");
  if (n > kMaxImages) return fail(PM_ERR_INVALID_ARG, "fusion takes at most 65536 images");
  int max_nov = 0;
  int64_t max_seeds = 1;
  for (int32_t i = 0; i < n; ++i) {
    const pm_fusion_image& x = im[i];
    if (!x.used) continue;
    if (x.width <= 6 || x.height <= 6 || x.width > 65535 || x.height > 32767 || !x.depth || !x.rgb ||
        (o->with_normal && !x.normal) || x.num_overlapping < 0 || (x.num_overlapping > 0 && !x.overlapping) ||
        !(x.channels == 0 || x.channels == 1 || x.channels == 3) || !finite_camera(x.cam))
      return fail(PM_ERR_INVALID_ARG, "invalid fusion image " + std::to_string(i));
    max_nov = std::max(max_nov, x.num_overlapping);
    const int64_t sc = (x.width - 6 + o->step_size - 1) / o->step_size;
    const int64_t sr = (x.height - 6 + o->step_size - 1) / o->step_size;
    max_seeds = std::max(max_seeds, sc * sr);
  }
  PM_HIP(hipSetDevice(e->device));
  hipStream_t st = e->stream;
  int cus = 0, clock_khz = 0;
  PM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device));
  PM_HIP(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, e->device));
  struct Events {  // destroyed on every return
    hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
    ~Events() {
      for (hipEvent_t x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } ev;
  for (hipEvent_t& x : ev.e) PM_HIP(hipEventCreate(&x));
  hipEvent_t t0 = ev.e[0], t1 = ev.e[1], f0 = ev.e[2], f1 = ev.e[3];
  PM_HIP(hipEventRecord(t0, st));

  // every used image's maps on the device, mask bits zero
  std::vector<FusImgDev> tab(n);
  std::vector<int64_t> base(n, 0), cbase(n, 0);  // first pixel, first colour byte
  int64_t total = 0, ov_total = 0, col_total = 0;
  for (int32_t i = 0; i < n; ++i) {
    if (!im[i].used) continue;
    base[i] = total;
    cbase[i] = col_total;
    total += (int64_t)im[i].width * im[i].height;
    col_total += (int64_t)im[i].width * im[i].height * (im[i].channels == 1 ? 1 : 3);
    ov_total += im[i].num_overlapping;
  }
  const int64_t mask_words = (total + 31) / 32;
  Buf d_depth, d_normal, d_rgb, d_mask, d_ov, d_tab, d_fused;
  if (pm_status s = d_depth.get((size_t)total * 4)) return s;
  if (o->with_normal)
    if (pm_status s = d_normal.get((size_t)total * 12)) return s;
  if (pm_status s = d_rgb.get((size_t)col_total)) return s;
  if (pm_status s = d_mask.get((size_t)mask_words * 4)) return s;
  if (pm_status s = d_ov.get((size_t)std::max<int64_t>(ov_total, 1) * 4)) return s;
  if (pm_status s = d_fused.get((size_t)n)) return s;
  int64_t ov_at = 0;
  for (int32_t i = 0; i < n; ++i) {
    FusImgDev& t = tab[i];
    std::memset(&t, 0, sizeof(t));
    t.used = im[i].used ? 1 : 0;
    t.W = im[i].width;
    t.H = im[i].height;
    if (!t.used) continue;
    const int64_t px = (int64_t)t.W * t.H;
    t.base = base[i];
    t.depth = d_depth.as<const float>() + base[i];
    t.ch = im[i].channels == 1 ? 1 : 3;
    t.rgb = d_rgb.as<const uint8_t>() + cbase[i];
    t.ov = d_ov.as<const int32_t>() + ov_at;
    t.nov = im[i].num_overlapping;
    if (pm_status s = copy_h2d(e, d_depth.as<float>() + base[i], im[i].depth, (size_t)px * 4, st)) return s;
    if (pm_status s = copy_h2d(e, d_rgb.as<uint8_t>() + cbase[i], im[i].rgb, (size_t)(px * t.ch), st)) return s;
    if (o->with_normal) {
      t.normal = d_normal.as<const float>() + 3 * base[i];
      if (pm_status s = copy_h2d(e, d_normal.as<float>() + 3 * base[i], im[i].normal, (size_t)px * 12, st)) return s;
    }
    if (t.nov > 0)
      PM_HIP(hipMemcpyAsync(d_ov.as<int32_t>() + ov_at, im[i].overlapping, (size_t)t.nov * 4, hipMemcpyHostToDevice,
                            st));
    ov_at += t.nov;
    // P and its inverse: the pose packing of a source seen from an identity reference
    pm_image si{};
    si.width = t.W;
    si.height = t.H;
    pm_view_desc vd{};
    vd.num_src = 1;
    vd.ref_cam.K[0] = vd.ref_cam.K[2] = 1.0f;
    vd.ref_cam.R[0] = vd.ref_cam.R[4] = vd.ref_cam.R[8] = 1.0f;
    vd.src = &si;
    vd.src_cams = &im[i].cam;
    float pose[PM_POSE_STRIDE], k8[8];
    pack_poses(&vd, pose, k8);
    std::memcpy(t.P, pose + 16, sizeof(t.P));
    std::memcpy(t.iP, pose + 28, sizeof(t.iP));
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) t.Rt[3 * a + b] = im[i].cam.R[3 * b + a];
  }
  if (pm_status s = d_tab.get(tab.size() * sizeof(FusImgDev))) return s;
  PM_HIP(hipMemcpyAsync(d_tab.p, tab.data(), tab.size() * sizeof(FusImgDev), hipMemcpyHostToDevice, st));
  uint32_t* mask = d_mask.as<uint32_t>();
  hipLaunchKernelGGL(k_fill_u32, dim3(fill_grid(mask_words)), dim3(256), 0, st, mask, mask_words);
  std::vector<uint8_t> fused(n, 0);
  PM_HIP(hipMemcpyAsync(d_fused.p, fused.data(), (size_t)n, hipMemcpyHostToDevice, st));
  PM_HIP(hipGetLastError());

  FusParams prm{};
  prm.n_images = n;
  prm.step = o->step_size;
  prm.min_num_pixels = o->min_num_pixels;
  prm.max_num_pixels = o->max_num_pixels;
  prm.min_num_visible = o->min_num_visible_images;
  prm.max_traversal = o->max_traversal_depth;
  prm.with_normal = o->with_normal ? 1 : 0;
  // max_squared_reproj_error_ and min_cos_normal_error_ (fusion.cc:657-659)
  prm.max_sq_reproj = (float)(o->max_reproj_error * o->max_reproj_error);
  prm.min_cos_normal =
      (float)std::cos(o->max_normal_error * 0.0174532925199432954743716805978692718781530857086181640625);
  prm.max_depth_error = o->max_depth_error;

  // window buffers: per seed, the graph pool, per pixel the walk lists.  A
  // window is as many seeds as the pool holds graphs for (at most one
  // image).  child_cap = 8 * node_cap holds node_cap / kNodeCap graphs of
  // any size.
  const int64_t wcap = std::min(max_seeds, kMaxWindow);
  const int64_t node_cap = std::min<int64_t>(std::max<int64_t>(wcap * 48, int64_t(1) << 16), int64_t(1) << 27);
  const int64_t child_cap = node_cap * (kChildCap / kNodeCap);
  Buf p_n, p_nc, p_noff, p_coff, p_gid, p_img, p_cbase, p_ccnt, p_child, p_used, p_room;
  Buf w_cnt, w_end, w_lists, w_sums, w_total, w_tickets, w_queue, w_counters, w_pending, w_npred, w_preds, w_nsucc,
      w_succend, w_succ, w_depth, w_rescnt, w_resnodes, w_status, w_pts, w_has, w_rank, w_visp, w_out, w_oseed, w_onvis,
      w_visat, w_ovis, w_kept, w_bars;
  const size_t ws_b = (size_t)wcap;
  const size_t nd_b = (size_t)node_cap;
  if (pm_status s = p_n.get(ws_b * 4)) return s;
  if (pm_status s = p_nc.get(ws_b * 4)) return s;
  if (pm_status s = p_noff.get(ws_b * 4)) return s;
  if (pm_status s = p_coff.get(ws_b * 4)) return s;
  if (pm_status s = p_gid.get(nd_b * 8)) return s;
  if (pm_status s = p_img.get(nd_b * 4)) return s;
  if (pm_status s = p_cbase.get(nd_b * 2)) return s;
  if (pm_status s = p_ccnt.get(nd_b * 2)) return s;
  if (pm_status s = p_child.get((size_t)child_cap * 2)) return s;
  if (pm_status s = p_used.get(16)) return s;
  if (pm_status s = p_room.get(4)) return s;
  if (pm_status s = w_cnt.get((size_t)total * 4)) return s;
  if (pm_status s = w_end.get((size_t)total * 4)) return s;
  if (pm_status s = w_lists.get(nd_b * 4)) return s;
  const int64_t scan_n = std::max(total, wcap);
  if (pm_status s = w_sums.get((size_t)((scan_n + kScanTile - 1) / kScanTile + 1) * 4)) return s;
  if (pm_status s = w_total.get(8)) return s;
  if (pm_status s = w_tickets.get(ws_b * 4)) return s;
  if (pm_status s = w_bars.get(ws_b * 8)) return s;
  if (pm_status s = w_queue.get(ws_b * 4)) return s;
  if (pm_status s = w_counters.get(32)) return s;
  if (pm_status s = w_pending.get(ws_b * 4)) return s;
  if (pm_status s = w_npred.get(ws_b * 4)) return s;
  if (pm_status s = w_preds.get(nd_b * 4)) return s;
  if (pm_status s = w_nsucc.get(ws_b * 4)) return s;
  if (pm_status s = w_succend.get(ws_b * 4)) return s;
  if (pm_status s = w_succ.get(nd_b * 4)) return s;
  if (pm_status s = w_rescnt.get(ws_b * 4)) return s;
  if (pm_status s = w_resnodes.get(nd_b * 2)) return s;
  if (pm_status s = w_status.get(16)) return s;
  if (pm_status s = w_depth.get(ws_b * 4)) return s;
  if (pm_status s = w_pts.get(ws_b * sizeof(PointRec))) return s;
  if (pm_status s = w_has.get(ws_b * 4)) return s;
  if (pm_status s = w_rank.get(ws_b * 4)) return s;
  if (pm_status s = w_visp.get(nd_b * 4)) return s;
  if (pm_status s = w_out.get(ws_b * sizeof(PointRec))) return s;
  if (pm_status s = w_oseed.get(ws_b * 4)) return s;
  if (pm_status s = w_onvis.get(ws_b * 4)) return s;
  if (pm_status s = w_visat.get(ws_b * 4)) return s;
  if (pm_status s = w_ovis.get(nd_b * 4)) return s;
  if (pm_status s = w_kept.get(8)) return s;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, w_kept.as<unsigned long long>());
  Pool g{};
  g.n = p_n.as<int32_t>();
  g.nc = p_nc.as<int32_t>();
  g.node_off = p_noff.as<int32_t>();
  g.child_off = p_coff.as<int32_t>();
  g.gid = p_gid.as<int64_t>();
  g.img = p_img.as<int32_t>();
  g.cbase = p_cbase.as<int16_t>();
  g.ccnt = p_ccnt.as<int16_t>();
  g.child = p_child.as<int16_t>();
  g.used = p_used.as<unsigned long long>();
  g.no_room = p_room.as<int32_t>();
  g.node_cap = node_cap;
  g.child_cap = child_cap;
  Flow f{};
  f.tickets = w_tickets.as<const int32_t>();
  f.queue = w_queue.as<int32_t>();
  f.counters = w_counters.as<int32_t>();
  f.pending = w_pending.as<int32_t>();
  f.npred = w_npred.as<int32_t>();
  f.preds = w_preds.as<int32_t>();
  f.nsucc = w_nsucc.as<int32_t>();
  f.succ_end = w_succend.as<int32_t>();
  f.succ = w_succ.as<int32_t>();
  f.depth = w_depth.as<int32_t>();
  f.res_cnt = w_rescnt.as<int32_t>();
  f.res_nodes = w_resnodes.as<int16_t>();
  f.cnt = w_cnt.as<const int32_t>();
  f.end = w_end.as<const int32_t>();
  f.lists = w_lists.as<const int32_t>();
  f.wait_ticks = (long long)(kWaitSeconds * 1000.0 * std::max(clock_khz, 1000));
  const unsigned flow_grid = (unsigned)std::max(1, cus * kFlowWavesPer8CU / 8);

  // a walk run alone: buffers for max_num_pixels
  const int big_acc = o->max_num_pixels;
  const int64_t big_stack = (int64_t)o->max_num_pixels * std::max(max_nov, 1) + 1;
  if (big_stack > INT32_MAX) return fail(PM_ERR_INVALID_ARG, "fusion: max_num_pixels x overlapping images too large");
  Buf b_rec, b_gids, b_acc, b_stack, b_pt, b_has, b_vis;

  std::unique_ptr<pm_fusion> res(new (std::nothrow) pm_fusion());
  if (!res) return fail(PM_ERR_OUT_OF_MEMORY, "host allocation");
  pm_fusion_stats& stt = res->stats;
  std::vector<int32_t> bars_h;
  std::vector<PointRec> pts_h;
  std::vector<int32_t> seed_h, vis_h;
  struct Alone {
    int32_t ws;
    PointRec r;
    std::vector<int32_t> vis;
  };
  std::vector<Alone> alone;

  int start = 0;
  for (int32_t i = 0; i < n; ++i)
    if (!im[i].is_rig) {
      start = i;
      break;
    }
  for (int cur = start; cur >= 0; cur = find_next(im, n, fused, cur)) {
    if (!im[cur].used) continue;
    const int W = im[cur].width, H = im[cur].height, step = o->step_size;
    const int seed_cols = (W - 6 + step - 1) / step, seed_rows = (H - 6 + step - 1) / step;
    const int64_t seeds = (int64_t)seed_cols * seed_rows;
    prm.cur = cur;
    prm.seed_cols = seed_cols;
    int64_t w0 = 0;
    int64_t limit = wcap;  // seeds of the next window: about as many as the pool held last
    while (w0 < seeds) {
      const int nwin = (int)std::min<int64_t>(limit, seeds - w0);
      // every seed's walk graph, against the masks of the window's start
      hipLaunchKernelGGL(k_window_reset, dim3(1), dim3(1), 0, st, g.used, g.no_room, nwin, w_status.as<int32_t>());
      hipLaunchKernelGGL(k_graph, dim3(nwin), dim3(64), 0, st, (const FusImgDev*)d_tab.p, (const uint32_t*)mask,
                         (const uint8_t*)d_fused.p, prm, (int32_t)w0, nwin, g);
      // the window's tickets and barriers (uses w_has, w_rank, w_onvis and
      // w_visat before the points do)
      const unsigned tgrid = (unsigned)((nwin + 255) / 256);
      hipLaunchKernelGGL(k_ticket_flags, dim3(tgrid), dim3(256), 0, st, (const int32_t*)g.n, nwin,
                         (const int32_t*)g.no_room, w_has.as<int32_t>(), w_onvis.as<int32_t>());
      scan(w_has.as<const int32_t>(), w_rank.as<int32_t>(), nwin, w_sums.as<int32_t>(), w_total.as<int32_t>(), st);
      scan(w_onvis.as<const int32_t>(), w_visat.as<int32_t>(), nwin, w_sums.as<int32_t>(), w_total.as<int32_t>() + 1,
           st);
      hipLaunchKernelGGL(k_ticket_scatter, dim3(tgrid), dim3(256), 0, st, (const int32_t*)g.n, nwin,
                         (const int32_t*)g.no_room, (const int32_t*)w_rank.p, (const int32_t*)w_visat.p,
                         w_tickets.as<int32_t>(), w_bars.as<int32_t>());
      PM_HIP(hipGetLastError());
      int32_t cnts[3] = {0, 0, 0};  // seeds with room, tickets, barriers
      PM_HIP(hipMemcpyAsync(&cnts[0], g.no_room, 4, hipMemcpyDeviceToHost, st));
      PM_HIP(hipMemcpyAsync(&cnts[1], w_total.p, 8, hipMemcpyDeviceToHost, st));
      PM_HIP(hipStreamSynchronize(st));
      const int32_t m = cnts[0], nt = cnts[1], nb = cnts[2];
      if (m == 0) {
        // the pool went to later seeds first: a window whose graphs all fit
        limit = std::max<int64_t>(1, node_cap / kNodeCap);
        continue;
      }
      const bool full = m < nwin;
      bars_h.resize((size_t)2 * nb);
      if (nb > 0) PM_HIP(hipMemcpy(bars_h.data(), w_bars.p, (size_t)nb * 8, hipMemcpyDeviceToHost));
      // the walks passing each pixel
      if (nt > 0) {
        hipLaunchKernelGGL(k_fill_i32, dim3(fill_grid(total)), dim3(256), 0, st, w_cnt.as<int32_t>(), total, 0);
        hipLaunchKernelGGL(k_count, dim3((m + 3) / 4), dim3(256), 0, st, g, m, w_cnt.as<int32_t>());
        scan(w_cnt.as<const int32_t>(), w_end.as<int32_t>(), total, w_sums.as<int32_t>(), w_total.as<int32_t>(), st);
        hipLaunchKernelGGL(k_scatter, dim3((m + 3) / 4), dim3(256), 0, st, g, m, w_end.as<int32_t>(),
                           w_lists.as<int32_t>());
        hipLaunchKernelGGL(k_fill_i32, dim3(fill_grid(m)), dim3(256), 0, st, f.res_cnt, (int64_t)m, 0);
      }
      hipLaunchKernelGGL(k_fill_i32, dim3(fill_grid(m)), dim3(256), 0, st, w_has.as<int32_t>(), (int64_t)m, 0);
      PM_HIP(hipGetLastError());
      // the dataflow, split at the walks that run alone
      PM_HIP(hipEventRecord(f0, st));
      int32_t t_begin = 0;
      for (int32_t bi = 0; bi <= nb; ++bi) {
        const int64_t stop = bi < nb ? bars_h[2 * bi] : m;
        const int32_t t_end = bi < nb ? bars_h[2 * bi + 1] : nt;
        if (t_end > t_begin) {
          // the segment's dependencies: predecessors, successor lists, the
          // walks ready at the start; then the dataflow
          Flow fs = f;
          fs.seed_base = tab[cur].base;
          fs.seed_w = W;
          fs.window_first = (int32_t)w0;
          fs.tickets = f.tickets + t_begin;
          fs.n_tickets = t_end - t_begin;
          // the seed after the last walk run alone: no earlier seed's walk
          // is a predecessor here, and none from it to the segment's first
          // ticket has a graph
          fs.seg_lo = bi == 0 ? 0 : bars_h[2 * (bi - 1)] + 1;
          const int nts = fs.n_tickets;
          const int64_t span = stop - fs.seg_lo;  // seeds [seg_lo, stop) hold every predecessor and successor
          hipLaunchKernelGGL(k_fill_i32, dim3(fill_grid(span)), dim3(256), 0, st, f.nsucc + fs.seg_lo, span, 0);
          hipLaunchKernelGGL(k_fill_i32, dim3(fill_grid(span)), dim3(256), 0, st, f.depth + fs.seg_lo, span, 0);
          hipLaunchKernelGGL(k_fill_i32, dim3(fill_grid(nts)), dim3(256), 0, st, f.queue, (int64_t)nts, -1);
          hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, st, f.counters, (int64_t)8, 0);
          hipLaunchKernelGGL(k_preds, dim3((nts + kPredWaves - 1) / kPredWaves), dim3(64 * kPredWaves), 0, st, g, fs);
          scan(f.nsucc + fs.seg_lo, f.succ_end + fs.seg_lo, span, w_sums.as<int32_t>(), w_total.as<int32_t>(), st);
          hipLaunchKernelGGL(k_succ, dim3((nts + 255) / 256), dim3(256), 0, st, g, fs);
          hipLaunchKernelGGL(k_flow, dim3(std::min<unsigned>(flow_grid, (unsigned)nts)), dim3(64), 0, st, prm, g,
                             mask, fs);
          hipLaunchKernelGGL(k_segment_end, dim3(1), dim3(1), 0, st, (const int32_t*)f.counters,
                             w_status.as<int32_t>());
          PM_HIP(hipGetLastError());
          stt.rounds++;
          stt.walks += nts;
        }
        t_begin = t_end;
        if (bi == nb) break;
        // the walk itself, pixel by pixel, with buffers for max_num_pixels
        const int32_t bws = bars_h[2 * bi];
        if (pm_status s = b_rec.get(sizeof(WalkRec))) return s;
        if (pm_status s = b_gids.get((size_t)big_acc * sizeof(int64_t))) return s;
        if (pm_status s = b_acc.get((size_t)big_acc * sizeof(AccEntry))) return s;
        if (pm_status s = b_stack.get((size_t)big_stack * sizeof(StackEntry))) return s;
        if (pm_status s = b_pt.get(sizeof(PointRec))) return s;
        if (pm_status s = b_has.get(4)) return s;
        if (pm_status s = b_vis.get((size_t)std::min(big_acc, n) * 4)) return s;
        FusParams pb = prm;
        pb.acc_cap = big_acc;
        pb.stack_cap = (int32_t)big_stack;
        hipLaunchKernelGGL(k_walk, dim3(1), dim3(64), 0, st, (const FusImgDev*)d_tab.p, (const uint32_t*)mask,
                           (const uint8_t*)d_fused.p, pb, (int32_t)(w0 + bws), b_rec.as<WalkRec>(),
                           b_gids.as<int64_t>(), b_acc.as<AccEntry>(), b_stack.as<StackEntry>());
        hipLaunchKernelGGL(k_commit_alone, dim3(fill_grid(big_acc)), dim3(256), 0, st, (const WalkRec*)b_rec.p,
                           (const int64_t*)b_gids.p, mask);
        hipLaunchKernelGGL(k_points_alone, dim3(1), dim3(64 * kPtWaves), 0, st, pb, (const WalkRec*)b_rec.p,
                           (const AccEntry*)b_acc.p, b_pt.as<PointRec>(), b_has.as<int32_t>(), b_vis.as<int32_t>());
        PM_HIP(hipGetLastError());
        Alone a;
        a.ws = bws;
        int32_t made = 0;
        WalkRec wr{};
        PM_HIP(hipMemcpyAsync(&made, b_has.p, 4, hipMemcpyDeviceToHost, st));
        PM_HIP(hipMemcpyAsync(&a.r, b_pt.p, sizeof(PointRec), hipMemcpyDeviceToHost, st));
        PM_HIP(hipMemcpyAsync(&wr, b_rec.p, sizeof(WalkRec), hipMemcpyDeviceToHost, st));
        PM_HIP(hipStreamSynchronize(st));
        stt.big_walks++;
        stt.walks++;
        if (wr.n_acc > 0) stt.kept++;
        if (made) {
          a.vis.resize(a.r.nvis);
          if (a.r.nvis > 0) PM_HIP(hipMemcpy(a.vis.data(), b_vis.p, (size_t)a.r.nvis * 4, hipMemcpyDeviceToHost));
          alone.push_back(std::move(a));
        }
      }
      PM_HIP(hipEventRecord(f1, st));
      // the graph walks' points, gathered in seed order
      int32_t np = 0, nv = 0;
      if (nt > 0) {
        Flow fa = f;
        fa.n_tickets = nt;
        hipLaunchKernelGGL(k_points, dim3((nt + kPtWaves - 1) / kPtWaves), dim3(64 * kPtWaves), 0, st,
                           (const FusImgDev*)d_tab.p, prm, (int32_t)w0, g, fa, w_pts.as<PointRec>(),
                           w_has.as<int32_t>(), w_visp.as<int32_t>(), w_kept.as<unsigned long long>());
        scan(w_has.as<const int32_t>(), w_rank.as<int32_t>(), m, w_sums.as<int32_t>(), w_total.as<int32_t>(), st);
        hipLaunchKernelGGL(k_gather_points, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st,
                           (const PointRec*)w_pts.p, (const int32_t*)w_has.p, (const int32_t*)w_rank.p, m,
                           w_out.as<PointRec>(), w_oseed.as<int32_t>(), w_onvis.as<int32_t>());
        PM_HIP(hipGetLastError());
        PM_HIP(hipMemcpyAsync(&np, w_total.p, 4, hipMemcpyDeviceToHost, st));
        PM_HIP(hipStreamSynchronize(st));
        if (np > 0) {
          scan(w_onvis.as<const int32_t>(), w_visat.as<int32_t>(), np, w_sums.as<int32_t>(), w_total.as<int32_t>(),
               st);
          hipLaunchKernelGGL(k_gather_vis, dim3((unsigned)((np + 3) / 4)), dim3(256), 0, st,
                             (const int32_t*)w_oseed.p, (const int32_t*)w_onvis.p, (const int32_t*)w_visat.p, np, g,
                             (const int32_t*)w_visp.p, w_ovis.as<int32_t>());
          PM_HIP(hipGetLastError());
          PM_HIP(hipMemcpyAsync(&nv, w_total.p, 4, hipMemcpyDeviceToHost, st));
          PM_HIP(hipStreamSynchronize(st));
          pts_h.resize(np);
          seed_h.resize(np);
          vis_h.resize(std::max(nv, 1));
          PM_HIP(hipMemcpyAsync(pts_h.data(), w_out.p, (size_t)np * sizeof(PointRec), hipMemcpyDeviceToHost, st));
          PM_HIP(hipMemcpyAsync(seed_h.data(), w_oseed.p, (size_t)np * 4, hipMemcpyDeviceToHost, st));
          if (nv > 0) PM_HIP(hipMemcpyAsync(vis_h.data(), w_ovis.p, (size_t)nv * 4, hipMemcpyDeviceToHost, st));
        }
      }
      int32_t status[3] = {0, 0, 0};
      PM_HIP(hipMemcpyAsync(status, w_status.p, 12, hipMemcpyDeviceToHost, st));
      PM_HIP(hipStreamSynchronize(st));
      if (status[0]) return fail(PM_ERR_INTERNAL, "fusion: no walk finished for 4 s; the dataflow gave up");
      float flow_ms = 0.0f;
      PM_HIP(hipEventElapsedTime(&flow_ms, f0, f1));
      stt.flow_ms += flow_ms;
      stt.chain += status[2];
      // merged with the walks run alone, by seed
      size_t ai = 0;
      int64_t vat = 0;
      for (int k = 0; k < np; ++k) {
        for (; ai < alone.size() && alone[ai].ws < seed_h[k]; ++ai) res->add(alone[ai].r, alone[ai].vis.data());
        res->add(pts_h[k], vis_h.data() + vat);
        vat += pts_h[k].nvis;
      }
      for (; ai < alone.size(); ++ai) res->add(alone[ai].r, alone[ai].vis.data());
      alone.clear();
      w0 += m;
      limit = full ? std::min<int64_t>(wcap, (int64_t)m + m / 4 + 1) : wcap;
    }
    fused[cur] = 255;
    const uint8_t one = 255;
    PM_HIP(hipMemcpyAsync((uint8_t*)d_fused.p + cur, &one, 1, hipMemcpyHostToDevice, st));
  }
  unsigned long long kept = 0;
  PM_HIP(hipMemcpyAsync(&kept, w_kept.p, 8, hipMemcpyDeviceToHost, st));
  PM_HIP(hipEventRecord(t1, st));
  PM_HIP(hipEventSynchronize(t1));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, t0, t1);
  stt.kept += (int64_t)kept;
  stt.total_ms = ms;
  stt.points = (int64_t)res->score.size();
  *out = res.release();
  return PM_OK;
}

PM_API pm_status pm_fusion_get_counts(const pm_fusion* f, int64_t* num_points, int64_t* num_visibility) {
  if (!f || !num_points || !num_visibility) return fail(PM_ERR_INVALID_ARG, "null argument");
  *num_points = (int64_t)f->score.size();
  *num_visibility = (int64_t)f->vis_images.size();
  return PM_OK;
}

PM_API pm_status pm_fusion_get_points(const pm_fusion* f, float* xyz, float* normal, uint8_t* rgb, float* score,
                                      int64_t* vis_offsets, int32_t* vis_images) {
  if (!f) return fail(PM_ERR_INVALID_ARG, "null fusion");
  const size_t n = f->score.size();
  if (xyz) std::copy(f->xyz.begin(), f->xyz.end(), xyz);
  if (normal) std::copy(f->normal.begin(), f->normal.end(), normal);
  if (rgb) std::copy(f->rgb.begin(), f->rgb.end(), rgb);
  if (score) std::copy(f->score.begin(), f->score.end(), score);
  if (vis_offsets) std::copy(f->vis_offsets.begin(), f->vis_offsets.begin() + n + 1, vis_offsets);
  if (vis_images) std::copy(f->vis_images.begin(), f->vis_images.end(), vis_images);
  return PM_OK;
}

PM_API pm_status pm_fusion_get_stats(const pm_fusion* f, pm_fusion_stats* out) {
  if (!f || !out) return fail(PM_ERR_INVALID_ARG, "null argument");
  *out = f->stats;
  return PM_OK;
}

PM_API void pm_fusion_destroy(pm_fusion* f) { delete f; }
