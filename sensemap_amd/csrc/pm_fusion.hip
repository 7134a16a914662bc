// pm_fusion.hip -- depth-map fusion: StereoFusion's walk (src/mvs/fusion.cc:
// 1730-2183, FuseImage and Fuse) over the images in the order of
// StereoFusion::Run (fusion.cc:697-727, 1150-1215), as one fusing thread per
// image runs it.
//
// The walks are sequential by definition: a pixel belongs to the first walk
// that accepts it, and each seed's walk sees the pixels earlier walks took.
// The device runs them speculatively instead, a window of seeds at a time,
// one thread per walk, each walk against the masks committed so far (plus
// its own accepted pixels), and then keeps exactly the walks whose result
// the sequential order would give:
//
//   * A walk's result depends on the masks only through the pixels it would
//     accept: a popped pixel that is masked and one that fails the
//     consistency tests both end in `continue` with no other effect.
//   * So a walk is exact when no earlier walk of the round accepted any of
//     the pixels it accepted (one atomicMin per accepted pixel finds the
//     earliest walk that accepted it).
//   * A walk that is not exact reruns in the next round, when every earlier
//     walk has committed.  Its rerun accepts a subset of what it accepted
//     speculatively (the masks only grew, and the accepted set is the
//     closure of the seed through the unmasked consistent pixels), so a later
//     walk that is exact now stays exact -- unless the speculative walk was
//     cut by max_traversal_depth or max_num_pixels (then a rerun may go
//     elsewhere): a round keeps no walk after such a cut walk that is not
//     exact.
//   * The first walk of a round is always exact, so every round commits at
//     least one walk; a walk that outgrew its record buffers reruns alone
//     with buffers sized for max_num_pixels.
//
// Per walk the kernel keeps the accepted pixels' values in its record, in
// visiting order, and computes the point there (medians by rank, the mean
// gradient in visiting order, the score).  Committing sets the masks and
// copies the point into its seed's slot; the host collects the slots of a
// window in seed order.
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

constexpr int kBlock = 128;
constexpr int kWindow = 65536;   // walks in flight
constexpr int kAccCap = 128;     // accepted pixels per walk record
constexpr int kStackCap = 2048;  // walk stack entries
constexpr int kVals = 11;        // x y z nx ny nz r g b grad err

enum : int32_t { kOverflow = 1, kCut = 2 };

struct FusImgDev {
  const float* depth;
  const float* normal;  // null without normals
  const uint8_t* rgb;
  uint8_t* mask;
  const int32_t* ov;
  int64_t base;  // first global pixel id of the image
  int32_t W, H, nov, used;
  float P[12], iP[12], Rt[9];
  int32_t pad[1];
};

struct FusParams {
  int32_t cur, seed_cols, step, n_images;
  int32_t min_num_pixels, max_num_pixels, min_num_visible, max_traversal;
  int32_t with_normal, acc_cap, stack_cap, pad0;
  float max_sq_reproj, min_cos_normal;
  double max_depth_error;
};

struct AccEntry {
  int32_t img, pad;
  float v[kVals];
};

struct WalkRec {
  int32_t flags, n_acc, has_point, nvis;
  float xyz[3], nrm[3], rgb[3], score;
};

struct StackEntry {
  int32_t img, row, col, depth;
};

__device__ inline float row4(const float* m, float x, float y, float z) { return ((m[0] * x + m[1] * y) + m[2] * z) + m[3]; }

// Fuse (fusion.cc:1794-2182) from one seed, against the committed masks:
// the accepted pixels (ids in gids, values in acc) in visiting order.  The
// point is formed later, for the walks that are kept.
__global__ void __launch_bounds__(kBlock) k_walk(const FusImgDev* __restrict__ imgs,
                                                 const uint8_t* __restrict__ fused, FusParams prm,
                                                 const int32_t* __restrict__ seeds, int32_t n, WalkRec* __restrict__ recs,
                                                 int64_t* __restrict__ gids, AccEntry* __restrict__ acc,
                                                 StackEntry* __restrict__ stacks) {
  const int w = blockIdx.x * kBlock + threadIdx.x;
  if (w >= n) return;
  int64_t* G = gids + (size_t)w * prm.acc_cap;
  AccEntry* A = acc + (size_t)w * prm.acc_cap;
  StackEntry* S = stacks + (size_t)w * prm.stack_cap;
  WalkRec rec;
  rec.flags = 0;
  rec.has_point = 0;
  rec.nvis = 0;
  const int s = seeds[w];
  const int seed_row = 3 + (s / prm.seed_cols) * prm.step;
  const int seed_col = 3 + (s % prm.seed_cols) * prm.step;
  int sp = 0, na = 0;
  S[sp++] = {prm.cur, seed_row, seed_col, 0};
  float ref_pt[3] = {0, 0, 0}, ref_n[3] = {0, 0, 0};
  while (sp > 0) {
    const StackEntry e = S[--sp];
    const FusImgDev& I = imgs[e.img];
    const int W = I.W, H = I.H;
    const int64_t p = (int64_t)e.row * W + e.col;
    if (I.mask[p]) continue;
    const int64_t gid = I.base + p;
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
      const int l = I.rgb[3 * (p - 1)], r = I.rgb[3 * (p + 1)];
      const int t = I.rgb[3 * (p - W)], b = I.rgb[3 * (p + W)];
      const float gx = (float)(l - r) / 255.0f, gy = (float)(t - b) / 255.0f;
      grad = sqrtf(gx * gx + gy * gy);
    }
    if (na == prm.acc_cap) {
      rec.flags |= kOverflow;
      break;
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
    a.v[6] = (float)I.rgb[3 * p];
    a.v[7] = (float)I.rgb[3 * p + 1];
    a.v[8] = (float)I.rgb[3 * p + 2];
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
    if (na >= prm.max_num_pixels) {
      rec.flags |= kCut;
      break;
    }
    const int nd = e.depth + 1;
    if (nd >= prm.max_traversal) {
      rec.flags |= kCut;
      continue;
    }
    bool full = false;
    for (int k = 0; k < I.nov && !full; ++k) {
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
      if (sp == prm.stack_cap) {
        full = true;
        break;
      }
      S[sp++] = {j, nr, nc, nd};
    }
    if (full) {
      rec.flags |= kOverflow;
      break;
    }
  }
  rec.n_acc = na;
  recs[w] = rec;
}

// ---------------------------------------------------------------------------
// The walk as a graph.  Whether a pixel passes the walk's tests depends only
// on the seed (its point and normal) and on the pixel, and the pixels a
// passing pixel pushes depend only on the pixel; the masks decide which of
// them the depth-first walk takes.  k_graph therefore records, once per
// window and one wave per seed, every pixel the walk could reach (breadth
// first, with the masks of the window's start: masked pixels stay masked),
// whether it passes, its values and its pushes in order.  Each round,
// k_replay replays the depth-first walk over that graph in LDS against the
// masks of the round -- the same pixels, order and cuts as walking the
// images again, without touching them.  A graph that does not fit is marked
// and its walk runs alone through k_walk.

constexpr int kNodeCap = 256;    // pixels per walk graph
constexpr int kChildCap = 2048;  // pushes per walk graph
constexpr int kHash = 512;       // open-addressing slots (gid -> node)

struct Graph {  // structure of arrays over the window's seeds
  int32_t* n;        // nodes; -1 when the graph did not fit
  int32_t* nc;       // child slots used
  int64_t* gid;      // [seed][kNodeCap]
  int32_t* img;      // [seed][kNodeCap]
  uint8_t* pass;     // [seed][kNodeCap]
  int16_t* cbase;    // [seed][kNodeCap]: first child slot, -1 when not expanded
  uint8_t* ccnt;     // [seed][kNodeCap]: child slots (the overlapping list's length)
  int16_t* child;    // [seed][kChildCap]: node index or -1 (no push)
  float* val;        // [seed][kNodeCap][kVals]
};

// the tests and values of one pixel against the seed's point and normal
// (fusion.cc:1884-2040); returns whether it passes
__device__ bool test_pixel(const FusImgDev& I, int row, int col, bool seed, const float ref_pt[3], const float ref_n[3],
                           const FusParams& prm, float* v) {
  const int W = I.W, H = I.H;
  const int64_t p = (int64_t)row * W + col;
  if (I.mask[p]) return false;
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
  v[6] = (float)I.rgb[3 * p];
  v[7] = (float)I.rgb[3 * p + 1];
  v[8] = (float)I.rgb[3 * p + 2];
  float grad = 0.0f;
  if (col > 1 && col < W - 1 && row > 1 && row < H - 1) {
    const int l = I.rgb[3 * (p - 1)], r = I.rgb[3 * (p + 1)];
    const int t = I.rgb[3 * (p - W)], b = I.rgb[3 * (p + W)];
    const float gx = (float)(l - r) / 255.0f, gy = (float)(t - b) / 255.0f;
    grad = sqrtf(gx * gx + gy * gy);
  }
  v[9] = grad;
  v[10] = depth_error;
  return true;
}

// One wave (one block of 64) per seed of the window.
__global__ void __launch_bounds__(64) k_graph(const FusImgDev* __restrict__ imgs, const uint8_t* __restrict__ fused,
                                              FusParams prm, int32_t window_first, int32_t count, Graph g) {
  __shared__ int64_t hkey[kHash];
  __shared__ int16_t hval[kHash];
  __shared__ int32_t nimg[kNodeCap];
  __shared__ int32_t nrc[kNodeCap];  // row << 16 | col
  __shared__ uint8_t npass[kNodeCap], ndep[kNodeCap];
  __shared__ int16_t cbase[kNodeCap];
  __shared__ uint8_t ccnt[kNodeCap];
  __shared__ float nxyz[kNodeCap][3];
  __shared__ int16_t child[kChildCap];
  __shared__ int16_t front[kNodeCap];
  const int ws = blockIdx.x, lane = threadIdx.x;
  if (ws >= count) return;
  const int s = window_first + ws;
  const int seed_row = 3 + (s / prm.seed_cols) * prm.step;
  const int seed_col = 3 + (s % prm.seed_cols) * prm.step;
  const size_t nb = (size_t)ws * kNodeCap;
  for (int i = lane; i < kHash; i += 64) hkey[i] = -1;
  __syncthreads();
  // the seed (node 0)
  float ref_pt[3], ref_n[3];
  {
    const FusImgDev& I = imgs[prm.cur];
    float v[kVals];
    const bool ok = test_pixel(I, seed_row, seed_col, true, nullptr, nullptr, prm, v);
    const int64_t gid = I.base + (int64_t)seed_row * I.W + seed_col;
    if (lane == 0) {
      nimg[0] = prm.cur;
      nrc[0] = (seed_row << 16) | seed_col;
      npass[0] = ok ? 1 : 0;
      ndep[0] = 0;
      const int h = (int)((uint64_t)gid * 0x9E3779B97F4A7C15ull >> 55);
      hkey[h] = gid;
      hval[h] = 0;
      if (ok)
        for (int k = 0; k < kVals; ++k) g.val[nb * kVals + k] = v[k];
    }
    for (int k = 0; k < 3; ++k) {
      ref_pt[k] = v[k];
      ref_n[k] = v[3 + k];
      if (lane == 0 && ok) nxyz[0][k] = v[k];
    }
    if (!ok) {
      if (lane == 0) {
        g.n[ws] = 1;
        g.nc[ws] = 0;
        g.gid[nb] = gid;
        g.img[nb] = prm.cur;
        g.pass[nb] = 0;
        g.cbase[nb] = -1;
        g.ccnt[nb] = 0;
      }
      return;
    }
  }
  int nn = 1, nch = 0, nf = 0, level = 0;
  bool overflow = false;
  if (1 < prm.max_traversal) {
    if (lane == 0) front[0] = 0;
    nf = 1;
  }
  if (lane == 0) cbase[0] = -1, ccnt[0] = 0;
  __syncthreads();
  while (nf > 0 && !overflow) {
    const int level_start = nn;
    for (int fi = 0; fi < nf && !overflow; ++fi) {
      const int f = front[fi];
      const FusImgDev& F = imgs[nimg[f]];
      const int nov = F.nov;
      if (nch + nov > kChildCap) {
        overflow = true;
        break;
      }
      if (lane == 0) {
        cbase[f] = (int16_t)nch;
        ccnt[f] = (uint8_t)nov;
      }
      const float x = nxyz[f][0], y = nxyz[f][1], z = nxyz[f][2];
      for (int k0 = 0; k0 < nov && !overflow; k0 += 64) {
        // candidate of push k (the projection into overlapping image k)
        const int k = k0 + lane;
        int cj = -1, crow = 0, ccol = 0;
        if (k < nov) {
          const int j = F.ov[k];
          if (j >= 0 && j < prm.n_images && imgs[j].used && !fused[j]) {
            const FusImgDev& J = imgs[j];
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
        }
        // insertion in push order, one candidate at a time (uniform)
        const int kn = min(64, nov - k0);
        for (int l = 0; l < kn; ++l) {
          const int j = __shfl(cj, l), r = __shfl(crow, l), c = __shfl(ccol, l);
          int idx = -1;
          if (j >= 0) {
            const int64_t gid = imgs[j].base + (int64_t)r * imgs[j].W + c;
            int h = (int)((uint64_t)gid * 0x9E3779B97F4A7C15ull >> 55);
            while (hkey[h] != -1 && hkey[h] != gid) h = (h + 1) & (kHash - 1);
            if (hkey[h] == gid) {
              idx = hval[h];
            } else if (nn < kNodeCap) {
              idx = nn++;
              __syncthreads();
              if (lane == 0) {
                hkey[h] = gid;
                hval[h] = (int16_t)idx;
                nimg[idx] = j;
                nrc[idx] = (r << 16) | c;
                ndep[idx] = (uint8_t)(level + 1);
              }
              __syncthreads();
            } else {
              overflow = true;
              break;
            }
          }
          if (lane == 0) child[nch + k0 + l] = (int16_t)idx;
        }
      }
      nch += nov;
      __syncthreads();
    }
    if (overflow) break;
    // test the new pixels (in parallel); the passing ones below the depth
    // cap form the next frontier, in node order
    int nnext = 0;
    for (int b = level_start; b < nn; b += 64) {
      const int i = b + lane;
      bool ok = false;
      if (i < nn) {
        float v[kVals];
        const int rc = nrc[i];
        ok = test_pixel(imgs[nimg[i]], rc >> 16, rc & 0xffff, false, ref_pt, ref_n, prm, v);
        npass[i] = ok ? 1 : 0;
        if (ok) {
          for (int k = 0; k < kVals; ++k) g.val[(nb + i) * kVals + k] = v[k];
          nxyz[i][0] = v[0];
          nxyz[i][1] = v[1];
          nxyz[i][2] = v[2];
        }
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
  if (overflow) {
    if (lane == 0) g.n[ws] = -1;
    return;
  }
  for (int i = lane; i < nn; i += 64) {
    const int rc = nrc[i];
    const FusImgDev& I = imgs[nimg[i]];
    g.gid[nb + i] = I.base + (int64_t)(rc >> 16) * I.W + (rc & 0xffff);
    g.img[nb + i] = nimg[i];
    g.pass[nb + i] = npass[i];
    g.cbase[nb + i] = cbase[i];
    g.ccnt[nb + i] = ccnt[i];
  }
  for (int c = lane; c < nch; c += 64) g.child[(size_t)ws * kChildCap + c] = child[c];
  if (lane == 0) {
    g.n[ws] = nn;
    g.nc[ws] = nch;
  }
}

// The depth-first walk of Fuse over a recorded graph against the masks of
// the round, then the walk's claims; one wave per pending walk.  Writes the
// accepted pixels in visiting order, as k_walk does.
__global__ void __launch_bounds__(64) k_replay(const FusImgDev* __restrict__ imgs, FusParams prm, Graph g,
                                               const int32_t* __restrict__ seeds, int32_t window_first, int32_t n,
                                               WalkRec* __restrict__ recs, int64_t* __restrict__ gids,
                                               AccEntry* __restrict__ acc, int32_t* __restrict__ first) {
  __shared__ int64_t ngid[kNodeCap];
  __shared__ uint8_t nok[kNodeCap], taken[kNodeCap], ncnt[kNodeCap];
  __shared__ int16_t nbase[kNodeCap];
  __shared__ int16_t child[kChildCap];
  __shared__ int32_t stack[kChildCap + 1];
  __shared__ int16_t accl[kAccCap];
  __shared__ int32_t sh[2];
  const int w = blockIdx.x, lane = threadIdx.x;
  if (w >= n) return;
  const int ws = seeds[w] - window_first;
  const int nn = g.n[ws];
  if (nn < 0) {
    if (lane == 0) {
      WalkRec r{};
      r.flags = kOverflow;
      recs[w] = r;
    }
    return;
  }
  const size_t nb = (size_t)ws * kNodeCap;
  const int nch = g.nc[ws];
  for (int i = lane; i < nn; i += 64) {
    const int64_t gid = g.gid[nb + i];
    const FusImgDev& I = imgs[g.img[nb + i]];
    ngid[i] = gid;
    nok[i] = g.pass[nb + i] && !I.mask[gid - I.base];
    taken[i] = 0;
    nbase[i] = g.cbase[nb + i];
    ncnt[i] = g.ccnt[nb + i];
  }
  for (int c = lane; c < nch; c += 64) child[c] = g.child[(size_t)ws * kChildCap + c];
  __syncthreads();
  if (lane == 0) {
    int sp = 0, na = 0, flags = 0;
    stack[sp++] = 0;  // node 0 at depth 0
    while (sp > 0) {
      const int e = stack[--sp];
      const int node = e & 0xffff, d = e >> 16;
      if (!nok[node] || taken[node]) continue;
      if (na == kAccCap) {
        flags |= kOverflow;
        break;
      }
      taken[node] = 1;
      accl[na++] = (int16_t)node;
      if (na >= prm.max_num_pixels) {
        flags |= kCut;
        break;
      }
      if (d + 1 >= prm.max_traversal) {
        flags |= kCut;
        continue;
      }
      const int b = nbase[node], cnt = ncnt[node];
      for (int k = 0; k < cnt; ++k) {
        const int c = child[b + k];
        if (c >= 0) stack[sp++] = ((d + 1) << 16) | c;
      }
    }
    sh[0] = na;
    sh[1] = flags;
  }
  __syncthreads();
  const int na = sh[0];
  for (int j = lane; j < na; j += 64) {
    const int node = accl[j];
    const int64_t gid = ngid[node];
    gids[(size_t)w * kAccCap + j] = gid;
    AccEntry& a = acc[(size_t)w * kAccCap + j];
    a.img = g.img[nb + node];
    for (int k = 0; k < kVals; ++k) a.v[k] = g.val[(nb + node) * kVals + k];
    atomicMin(first + gid, w);
  }
  if (lane == 0) {
    WalkRec r{};
    r.flags = sh[1];
    r.n_acc = na;
    recs[w] = r;
  }
}

// The earliest walk of the round that accepted each pixel.
__global__ void __launch_bounds__(kBlock) k_claim(const WalkRec* __restrict__ recs, const int64_t* __restrict__ gids,
                                                  int32_t n, int32_t acc_cap, int32_t* __restrict__ first) {
  const int w = blockIdx.x * kBlock + threadIdx.x;
  if (w >= n) return;
  const int na = min(recs[w].n_acc, acc_cap);
  for (int j = 0; j < na; ++j) atomicMin(first + gids[(size_t)w * acc_cap + j], w);
}

// Exact walks; the first inexact walk that was cut or overflowed bounds the
// round (stop).
__global__ void __launch_bounds__(kBlock) k_validate(const WalkRec* __restrict__ recs, const int64_t* __restrict__ gids,
                                                     int32_t n, int32_t acc_cap, const int32_t* __restrict__ first,
                                                     uint8_t* __restrict__ exact, int32_t* __restrict__ stop) {
  const int w = blockIdx.x * kBlock + threadIdx.x;
  if (w >= n) return;
  const WalkRec r = recs[w];
  bool ok = !(r.flags & kOverflow);
  for (int j = 0; j < r.n_acc && ok; ++j) ok = first[gids[(size_t)w * acc_cap + j]] == w;
  exact[w] = ok ? 1 : 0;
  if (!ok && (r.flags & (kOverflow | kCut))) atomicMin(stop, w);
}

// Commit the kept walks (masks; their seed slots start without a point) and
// reset the claims.
__global__ void __launch_bounds__(kBlock) k_commit(const FusImgDev* __restrict__ imgs, const WalkRec* __restrict__ recs,
                                                   const int64_t* __restrict__ gids, const AccEntry* __restrict__ acc,
                                                   const int32_t* __restrict__ seeds, int32_t n, int32_t acc_cap,
                                                   int32_t window_first, const int32_t* __restrict__ stop,
                                                   uint8_t* __restrict__ exact, int32_t* __restrict__ first,
                                                   WalkRec* __restrict__ slots) {
  const int w = blockIdx.x * kBlock + threadIdx.x;
  if (w >= n) return;
  const WalkRec r = recs[w];
  const int na = min(r.n_acc, acc_cap);
  const bool keep = exact[w] && w < *stop;
  for (int j = 0; j < na; ++j) {
    const int64_t g = gids[(size_t)w * acc_cap + j];
    first[g] = INT_MAX;
    if (keep) {
      const FusImgDev& I = imgs[acc[(size_t)w * acc_cap + j].img];
      I.mask[g - I.base] = 1;
    }
  }
  exact[w] = keep ? 1 : 0;
  if (!keep) return;
  WalkRec out = r;
  out.has_point = 0;
  slots[seeds[w] - window_first] = out;
}

// The point of a kept walk (fusion.cc:2098-2180), one wave per walk.  The
// medians (internal::Median, fusion.cc:47-60: the element of rank n/2 and,
// for an even count, the mean with the element of rank n/2 - 1) come from
// each entry's rank, counted in LDS; the mean gradient is summed in
// visiting order.  Walks with more entries than LDS holds (run alone) are
// read from the record directly.
constexpr int kPtWaves = 4;
constexpr int kMaxImages = 65536;           // fusion takes at most this many images
constexpr int kImgWords = kMaxImages / 32;  // visibility bitmap per wave

__device__ inline float pick(bool c0, float x0, bool c1, float x1) {
  const uint64_t m0 = __ballot(c0), m1 = __ballot(c1);
  if (m0) return __shfl(x0, __ffsll((long long)m0) - 1);
  return __shfl(x1, __ffsll((long long)m1) - 1);
}

__global__ void __launch_bounds__(64 * kPtWaves) k_points(const WalkRec* __restrict__ recs,
                                                          const AccEntry* __restrict__ acc,
                                                          const int32_t* __restrict__ seeds, int32_t n, FusParams prm,
                                                          int32_t window_first, const uint8_t* __restrict__ kept,
                                                          WalkRec* __restrict__ slots, int32_t* __restrict__ vis_out,
                                                          int32_t vis_stride) {
  __shared__ float lv[kPtWaves][kVals][kAccCap];
  __shared__ uint32_t bits[kPtWaves][kImgWords];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int w = blockIdx.x * kPtWaves + wave;
  WalkRec r{};
  bool active = w < n && kept[w];
  if (active) {
    r = recs[w];
    active = !(r.flags & kOverflow) && r.n_acc >= prm.min_num_pixels;
  }
  const int na = active ? r.n_acc : 0;
  const AccEntry* A = acc + (size_t)(active ? w : 0) * prm.acc_cap;
  const bool in_lds = na <= kAccCap;
  const int words = (prm.n_images + 31) / 32;
  for (int k = lane; k < words; k += 64) bits[wave][k] = 0u;
  __syncthreads();
  if (active) {
    for (int i = lane; i < na; i += 64) {
      if (in_lds)
        for (int v = 0; v < kVals; ++v) lv[wave][v][i] = A[i].v[v];
      const int32_t m = A[i].img;
      atomicOr(&bits[wave][m >> 5], 1u << (m & 31));
    }
  }
  __syncthreads();
  if (!active) return;
  auto val = [&](int i, int v) -> float { return in_lds ? lv[wave][v][i] : A[i].v[v]; };
  // rank-based median of value v over entries [b, na)
  auto median = [&](int v, int b) -> float {
    const int cnt = na - b, mid = cnt / 2;
    float hi = 0.0f, lo = 0.0f;
    for (int base = b; base < na; base += 128) {
      const int i0 = base + lane, i1 = base + 64 + lane;
      float x0 = 0.0f, x1 = 0.0f;
      int l0 = 0, e0 = 0, l1 = 0, e1 = 0;
      if (i0 < na) x0 = val(i0, v);
      if (i1 < na) x1 = val(i1, v);
      for (int j = b; j < na; ++j) {
        const float y = val(j, v);
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
  for (int k = k0; k < k1; ++k) mine += __popc(bits[wave][k]);
  int before = 0, nvis = 0;
  for (int l = 0; l < 64; ++l) {
    const int c = __shfl(mine, l);
    before += l < lane ? c : 0;
    nvis += c;
  }
  if (nvis < prm.min_num_visible) return;
  if (vis_out) {
    int32_t* V = vis_out + (size_t)(seeds[w] - window_first) * vis_stride;
    int pos = before;
    for (int k = k0; k < k1; ++k)
      for (uint32_t m = bits[wave][k]; m; m &= m - 1) V[pos++] = 32 * k + __ffs(m) - 1;
  }
  WalkRec out = r;
  out.nvis = nvis;
  bool ok = true;
  for (int c = 0; c < 3; ++c) {
    out.xyz[c] = median(c, 0);
    if (isnan(out.xyz[c]) || isinf(out.xyz[c])) ok = false;
  }
  out.nrm[0] = out.nrm[1] = out.nrm[2] = 0.0f;
  if (ok && prm.with_normal) {
    const float nx = median(3, 0), ny = median(4, 0), nz = median(5, 0);
    const float norm = sqrtf(nx * nx + (ny * ny + nz * nz));
    if (norm < FLT_EPSILON) {
      ok = false;
    } else {
      out.nrm[0] = nx / norm;
      out.nrm[1] = ny / norm;
      out.nrm[2] = nz / norm;
    }
  }
  if (ok) {
    for (int c = 0; c < 3; ++c) {
      const float m = roundf(median(6 + c, 0));
      out.rgb[c] = m < 0.0f ? 0.0f : (m > 255.0f ? 255.0f : m);
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
      out.score = score;
    }
  }
  out.has_point = ok ? 1 : 0;
  if (lane == 0) slots[seeds[w] - window_first] = out;
}

__global__ void __launch_bounds__(256) k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = v;
}
__global__ void __launch_bounds__(256) k_fill_u8(uint8_t* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = 0;
}

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
};

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
  int max_nov = 0;
  for (int32_t i = 0; i < n; ++i) {
    const pm_fusion_image& x = im[i];
    if (!x.used) continue;
    if (x.width <= 6 || x.height <= 6 || x.width > 65535 || x.height > 32767 || !x.depth || !x.rgb ||
        (o->with_normal && !x.normal) || x.num_overlapping < 0 || (x.num_overlapping > 0 && !x.overlapping) ||
        !finite_camera(x.cam))
      return fail(PM_ERR_INVALID_ARG, "invalid fusion image " + std::to_string(i));
    max_nov = std::max(max_nov, x.num_overlapping);
  }
  if (n > kMaxImages) return fail(PM_ERR_INVALID_ARG, "fusion takes at most 65536 images");
  PM_HIP(hipSetDevice(e->device));
  hipStream_t st = e->stream;
  struct Events {  // destroyed on every return
    hipEvent_t a = nullptr, b = nullptr;
    ~Events() {
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
    }
  } ev;
  PM_HIP(hipEventCreate(&ev.a));
  PM_HIP(hipEventCreate(&ev.b));
  hipEvent_t t0 = ev.a, t1 = ev.b;
  PM_HIP(hipEventRecord(t0, st));

  // every used image's maps on the device, masks zero, claims INT_MAX
  std::vector<FusImgDev> tab(n);
  std::vector<int64_t> base(n, 0);
  int64_t total = 0, rgb_total = 0, ov_total = 0;
  for (int32_t i = 0; i < n; ++i) {
    if (!im[i].used) continue;
    base[i] = total;
    total += (int64_t)im[i].width * im[i].height;
    ov_total += im[i].num_overlapping;
  }
  rgb_total = 3 * total;
  Buf d_depth, d_normal, d_rgb, d_mask, d_first, d_ov, d_tab, d_fused;
  if (pm_status s = d_depth.get((size_t)total * 4)) return s;
  if (o->with_normal)
    if (pm_status s = d_normal.get((size_t)total * 12)) return s;
  if (pm_status s = d_rgb.get((size_t)rgb_total)) return s;
  if (pm_status s = d_mask.get((size_t)total)) return s;
  if (pm_status s = d_first.get((size_t)total * 4)) return s;
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
    t.depth = (const float*)d_depth.p + base[i];
    t.rgb = (const uint8_t*)d_rgb.p + 3 * base[i];
    t.mask = (uint8_t*)d_mask.p + base[i];
    t.ov = (const int32_t*)d_ov.p + ov_at;
    t.nov = im[i].num_overlapping;
    PM_HIP(hipMemcpyAsync((float*)d_depth.p + base[i], im[i].depth, (size_t)px * 4, hipMemcpyHostToDevice, st));
    PM_HIP(hipMemcpyAsync((uint8_t*)d_rgb.p + 3 * base[i], im[i].rgb, (size_t)px * 3, hipMemcpyHostToDevice, st));
    if (o->with_normal) {
      t.normal = (const float*)d_normal.p + 3 * base[i];
      PM_HIP(hipMemcpyAsync((float*)d_normal.p + 3 * base[i], im[i].normal, (size_t)px * 12, hipMemcpyHostToDevice,
                            st));
    }
    if (t.nov > 0)
      PM_HIP(hipMemcpyAsync((int32_t*)d_ov.p + ov_at, im[i].overlapping, (size_t)t.nov * 4, hipMemcpyHostToDevice, st));
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
  hipLaunchKernelGGL(k_fill_u8, dim3(fill_grid(total)), dim3(256), 0, st, (uint8_t*)d_mask.p, total);
  hipLaunchKernelGGL(k_fill_i32, dim3(fill_grid(total)), dim3(256), 0, st, (int32_t*)d_first.p, total, INT_MAX);
  hipLaunchKernelGGL(k_fill_u8, dim3(1), dim3(256), 0, st, (uint8_t*)d_fused.p, (int64_t)n);
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

  // walk records: a window of kWindow walks, and one large record for a walk
  // that outgrew its buffers
  const int big_acc = o->max_num_pixels;
  const int64_t big_stack = (int64_t)o->max_num_pixels * std::max(max_nov, 1) + 1;
  if (big_stack > INT32_MAX) return fail(PM_ERR_INVALID_ARG, "fusion: max_num_pixels x overlapping images too large");
  Buf d_recs, d_gids, d_acc, d_stack, d_seeds, d_exact, d_stop, d_slots, d_slot_vis, b_gids, b_acc, b_stack, b_vis;
  if (pm_status s = d_recs.get((size_t)kWindow * sizeof(WalkRec))) return s;
  if (pm_status s = d_gids.get((size_t)kWindow * kAccCap * sizeof(int64_t))) return s;
  if (pm_status s = d_acc.get((size_t)kWindow * kAccCap * sizeof(AccEntry))) return s;
  if (pm_status s = d_stack.get((size_t)kWindow * kStackCap * sizeof(StackEntry))) return s;
  if (pm_status s = d_seeds.get((size_t)kWindow * 4)) return s;
  if (pm_status s = d_exact.get((size_t)kWindow)) return s;
  if (pm_status s = d_stop.get(4)) return s;
  if (pm_status s = d_slots.get((size_t)kWindow * sizeof(WalkRec))) return s;
  if (pm_status s = d_slot_vis.get((size_t)kWindow * kAccCap * 4)) return s;
  Buf g_n, g_nc, g_gid, g_img, g_pass, g_cbase, g_ccnt, g_child, g_val;
  const size_t nodes = (size_t)kWindow * kNodeCap;
  if (pm_status s = g_n.get((size_t)kWindow * 4)) return s;
  if (pm_status s = g_nc.get((size_t)kWindow * 4)) return s;
  if (pm_status s = g_gid.get(nodes * 8)) return s;
  if (pm_status s = g_img.get(nodes * 4)) return s;
  if (pm_status s = g_pass.get(nodes)) return s;
  if (pm_status s = g_cbase.get(nodes * 2)) return s;
  if (pm_status s = g_ccnt.get(nodes)) return s;
  if (pm_status s = g_child.get((size_t)kWindow * kChildCap * 2)) return s;
  if (pm_status s = g_val.get(nodes * kVals * 4)) return s;
  Graph graph{(int32_t*)g_n.p,   (int32_t*)g_nc.p,    (int64_t*)g_gid.p,  (int32_t*)g_img.p, (uint8_t*)g_pass.p,
              (int16_t*)g_cbase.p, (uint8_t*)g_ccnt.p, (int16_t*)g_child.p, (float*)g_val.p};

  std::unique_ptr<pm_fusion> res(new (std::nothrow) pm_fusion());
  if (!res) return fail(PM_ERR_OUT_OF_MEMORY, "host allocation");
  pm_fusion_stats& stt = res->stats;
  std::vector<uint8_t> fused(n, 0);
  std::vector<int32_t> pending, next;
  std::vector<uint8_t> exact(kWindow);
  std::vector<WalkRec> slots(kWindow);
  std::vector<int32_t> slot_vis((size_t)kWindow * kAccCap);
  std::vector<std::vector<int32_t>> big_vis(kWindow);  // visibility of walks run alone (may exceed kAccCap)
  std::vector<uint8_t> big_slot(kWindow);

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
    if (seeds > INT32_MAX) return fail(PM_ERR_INVALID_ARG, "fusion: too many seeds in one image");
    prm.cur = cur;
    prm.seed_cols = seed_cols;
    for (int64_t w0 = 0; w0 < seeds; w0 += kWindow) {
      const int nwin = (int)std::min<int64_t>(kWindow, seeds - w0);
      pending.resize(nwin);
      for (int k = 0; k < nwin; ++k) pending[k] = (int32_t)(w0 + k);
      std::fill(big_slot.begin(), big_slot.begin() + nwin, 0);
      bool big_next = false;
      // every seed's walk graph, against the masks of the window's start
      hipLaunchKernelGGL(k_graph, dim3(nwin), dim3(64), 0, st, (const FusImgDev*)d_tab.p, (const uint8_t*)d_fused.p,
                         prm, (int32_t)w0, nwin, graph);
      PM_HIP(hipGetLastError());
      while (!pending.empty()) {
        const int m = big_next ? 1 : (int)pending.size();
        int64_t* gid = (int64_t*)d_gids.p;
        AccEntry* acc = (AccEntry*)d_acc.p;
        StackEntry* stk = (StackEntry*)d_stack.p;
        int32_t* vis_out = (int32_t*)d_slot_vis.p;
        int32_t vis_stride = kAccCap;
        prm.acc_cap = kAccCap;
        prm.stack_cap = kStackCap;
        if (big_next) {
          if (pm_status s = b_gids.get((size_t)big_acc * sizeof(int64_t))) return s;
          if (pm_status s = b_acc.get((size_t)big_acc * sizeof(AccEntry))) return s;
          if (pm_status s = b_stack.get((size_t)big_stack * sizeof(StackEntry))) return s;
          if (pm_status s = b_vis.get((size_t)std::min(big_acc, n) * 4)) return s;
          gid = (int64_t*)b_gids.p;
          acc = (AccEntry*)b_acc.p;
          stk = (StackEntry*)b_stack.p;
          vis_out = (int32_t*)b_vis.p;
          vis_stride = 0;
          prm.acc_cap = big_acc;
          prm.stack_cap = (int32_t)big_stack;
        }
        PM_HIP(hipMemcpyAsync(d_seeds.p, pending.data(), (size_t)m * 4, hipMemcpyHostToDevice, st));
        const int32_t no_stop = INT_MAX;
        PM_HIP(hipMemcpyAsync(d_stop.p, &no_stop, 4, hipMemcpyHostToDevice, st));
        const dim3 grid((m + kBlock - 1) / kBlock);
        if (big_next) {
          // the walk itself, pixel by pixel, with buffers for max_num_pixels
          hipLaunchKernelGGL(k_walk, grid, dim3(kBlock), 0, st, (const FusImgDev*)d_tab.p,
                             (const uint8_t*)d_fused.p, prm, (const int32_t*)d_seeds.p, m, (WalkRec*)d_recs.p, gid,
                             acc, stk);
          hipLaunchKernelGGL(k_claim, grid, dim3(kBlock), 0, st, (const WalkRec*)d_recs.p, (const int64_t*)gid, m,
                             prm.acc_cap, (int32_t*)d_first.p);
        } else {
          hipLaunchKernelGGL(k_replay, dim3(m), dim3(64), 0, st, (const FusImgDev*)d_tab.p, prm, graph,
                             (const int32_t*)d_seeds.p, (int32_t)w0, m, (WalkRec*)d_recs.p, gid, acc,
                             (int32_t*)d_first.p);
        }
        hipLaunchKernelGGL(k_validate, grid, dim3(kBlock), 0, st, (const WalkRec*)d_recs.p, (const int64_t*)gid, m,
                           prm.acc_cap, (const int32_t*)d_first.p, (uint8_t*)d_exact.p, (int32_t*)d_stop.p);
        hipLaunchKernelGGL(k_commit, grid, dim3(kBlock), 0, st, (const FusImgDev*)d_tab.p, (const WalkRec*)d_recs.p,
                           (const int64_t*)gid, (const AccEntry*)acc, (const int32_t*)d_seeds.p, m, prm.acc_cap,
                           (int32_t)w0, (const int32_t*)d_stop.p, (uint8_t*)d_exact.p, (int32_t*)d_first.p,
                           (WalkRec*)d_slots.p);
        hipLaunchKernelGGL(k_points, dim3((m + kPtWaves - 1) / kPtWaves), dim3(64 * kPtWaves), 0, st,
                           (const WalkRec*)d_recs.p, (const AccEntry*)acc, (const int32_t*)d_seeds.p, m, prm,
                           (int32_t)w0, (const uint8_t*)d_exact.p, (WalkRec*)d_slots.p, vis_out, vis_stride);
        PM_HIP(hipGetLastError());
        PM_HIP(hipMemcpyAsync(exact.data(), d_exact.p, (size_t)m, hipMemcpyDeviceToHost, st));
        WalkRec first_rec{};
        PM_HIP(hipMemcpyAsync(&first_rec, d_recs.p, sizeof(WalkRec), hipMemcpyDeviceToHost, st));
        WalkRec first_slot{};
        if (big_next)
          PM_HIP(hipMemcpyAsync(&first_slot, (const WalkRec*)d_slots.p + (pending[0] - (int32_t)w0), sizeof(WalkRec),
                                hipMemcpyDeviceToHost, st));
        PM_HIP(hipStreamSynchronize(st));
        stt.rounds++;
        stt.walks += m;
        if (big_next) {
          // run alone with buffers for max_num_pixels: always kept
          if (!exact[0]) return fail(PM_ERR_INTERNAL, "fusion: a walk run alone was not kept");
          stt.big_walks++;
          const int slot = pending[0] - (int32_t)w0;
          big_slot[slot] = 1;
          big_vis[slot].assign(first_slot.has_point ? first_slot.nvis : 0, 0);
          if (!big_vis[slot].empty())
            PM_HIP(hipMemcpy(big_vis[slot].data(), b_vis.p, big_vis[slot].size() * 4, hipMemcpyDeviceToHost));
          pending.erase(pending.begin());
          stt.kept++;
          big_next = false;
          continue;
        }
        next.clear();
        for (int k = 0; k < m; ++k)
          if (!exact[k]) next.push_back(pending[k]);
        stt.kept += m - (int64_t)next.size();
        // the first walk of a round is kept unless it outgrew its buffers
        big_next = !exact[0] && (first_rec.flags & kOverflow);
        if (!exact[0] && !big_next) return fail(PM_ERR_INTERNAL, "fusion: the first walk of a round was not kept");
        pending.swap(next);
      }
      // the window's points, in seed order
      PM_HIP(hipMemcpyAsync(slots.data(), d_slots.p, (size_t)nwin * sizeof(WalkRec), hipMemcpyDeviceToHost, st));
      PM_HIP(hipMemcpyAsync(slot_vis.data(), d_slot_vis.p, (size_t)nwin * kAccCap * 4, hipMemcpyDeviceToHost, st));
      PM_HIP(hipStreamSynchronize(st));
      for (int k = 0; k < nwin; ++k) {
        const WalkRec& r = slots[k];
        if (!r.has_point) continue;
        for (int c = 0; c < 3; ++c) {
          res->xyz.push_back(r.xyz[c]);
          res->normal.push_back(r.nrm[c]);
          res->rgb.push_back((uint8_t)r.rgb[c]);
        }
        res->score.push_back(r.score);
        if (big_slot[k]) {
          res->vis_images.insert(res->vis_images.end(), big_vis[k].begin(), big_vis[k].end());
          big_vis[k].clear();
        } else {
          res->vis_images.insert(res->vis_images.end(), slot_vis.begin() + (size_t)k * kAccCap,
                                 slot_vis.begin() + (size_t)k * kAccCap + r.nvis);
        }
        res->vis_offsets.push_back((int64_t)res->vis_images.size());
      }
    }
    fused[cur] = 255;
    const uint8_t one = 255;
    PM_HIP(hipMemcpyAsync((uint8_t*)d_fused.p + cur, &one, 1, hipMemcpyHostToDevice, st));
  }
  PM_HIP(hipEventRecord(t1, st));
  PM_HIP(hipEventSynchronize(t1));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, t0, t1);
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
