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
  int64_t gid;
  int32_t img, pad;
  float v[kVals];
  float pad2;
};

struct WalkRec {
  int32_t flags, n_acc, has_point, nvis;
  float xyz[3], nrm[3], rgb[3], score;
};

struct StackEntry {
  int32_t img, row, col, depth;
};

__device__ inline float row4(const float* m, float x, float y, float z) { return ((m[0] * x + m[1] * y) + m[2] * z) + m[3]; }

// internal::Median (fusion.cc:47-60): the element of rank n/2 and, for an
// even count, the mean with the element of rank n/2 - 1.
__device__ float median_rank(const AccEntry* a, int n, int v) {
  const int mid = n / 2;
  float hi = 0.0f, lo = 0.0f;
  for (int i = 0; i < n; ++i) {
    const float x = a[i].v[v];
    int less = 0, le = 0;
    for (int j = 0; j < n; ++j) {
      const float y = a[j].v[v];
      less += y < x;
      le += y <= x;
    }
    if (less <= mid && mid < le) hi = x;
    if (less <= mid - 1 && mid - 1 < le) lo = x;
  }
  return (n % 2 == 0) ? (hi + lo) / 2.0f : hi;
}

// Fuse (fusion.cc:1794-2182) from one seed, against the committed masks.
__global__ void __launch_bounds__(kBlock) k_walk(const FusImgDev* __restrict__ imgs,
                                                 const uint8_t* __restrict__ fused, FusParams prm,
                                                 const int32_t* __restrict__ seeds, int32_t n, WalkRec* __restrict__ recs,
                                                 AccEntry* __restrict__ acc, StackEntry* __restrict__ stacks,
                                                 int32_t* __restrict__ vis) {
  const int w = blockIdx.x * kBlock + threadIdx.x;
  if (w >= n) return;
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
    for (int j = 0; j < na && !mine; ++j) mine = A[j].gid == gid;
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
    AccEntry& a = A[na++];
    a.gid = gid;
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
  if (!(rec.flags & kOverflow) && na >= prm.min_num_pixels) {
    // distinct images, ascending (the visibility set)
    int32_t* V = vis + (size_t)w * prm.acc_cap;
    int nv = 0;
    for (int j = 0; j < na; ++j) {
      const int32_t img = A[j].img;
      int pos = 0;
      bool dup = false;
      for (; pos < nv; ++pos) {
        if (V[pos] == img) {
          dup = true;
          break;
        }
        if (V[pos] > img) break;
      }
      if (dup) continue;
      for (int q = nv; q > pos; --q) V[q] = V[q - 1];
      V[pos] = img;
      nv++;
    }
    rec.nvis = nv;
    if (nv >= prm.min_num_visible) {
      bool ok = true;
      for (int c = 0; c < 3; ++c) {
        rec.xyz[c] = median_rank(A, na, c);
        if (isnan(rec.xyz[c]) || isinf(rec.xyz[c])) ok = false;
      }
      rec.nrm[0] = rec.nrm[1] = rec.nrm[2] = 0.0f;
      if (ok && prm.with_normal) {
        const float nx = median_rank(A, na, 3), ny = median_rank(A, na, 4), nz = median_rank(A, na, 5);
        const float norm = sqrtf(nx * nx + (ny * ny + nz * nz));
        if (norm < FLT_EPSILON) {
          ok = false;
        } else {
          rec.nrm[0] = nx / norm;
          rec.nrm[1] = ny / norm;
          rec.nrm[2] = nz / norm;
        }
      }
      if (ok) {
        for (int c = 0; c < 3; ++c) {
          const float m = roundf(median_rank(A, na, 6 + c));
          rec.rgb[c] = m < 0.0f ? 0.0f : (m > 255.0f ? 255.0f : m);
        }
        float m_conf = 0.0f;
        for (int j = 0; j < na; ++j) m_conf += A[j].v[9];
        m_conf /= (float)na;
        if (nv >= 3) m_conf = 1.0f;
        m_conf = 1.0f - m_conf;
        // the depth errors are those of every accepted pixel but the seed
        const int ne = na - 1;
        if (ne == 0) {
          ok = false;
        } else {
          float med = 0.0f;
          {
            const int mid = ne / 2;
            float hi = 0.0f, lo = 0.0f;
            for (int i = 1; i < na; ++i) {
              const float x = A[i].v[10];
              int less = 0, le = 0;
              for (int j = 1; j < na; ++j) {
                const float y = A[j].v[10];
                less += y < x;
                le += y <= x;
              }
              if (less <= mid && mid < le) hi = x;
              if (less <= mid - 1 && mid - 1 < le) lo = x;
            }
            med = (ne % 2 == 0) ? (hi + lo) / 2.0f : hi;
          }
          float score = (float)((double)med / prm.max_depth_error);
          score = fminf(0.6f * score + 0.4f * m_conf, 1.0f);
          if (isnan(score) || isinf(score)) ok = false;
          rec.score = score;
        }
      }
      rec.has_point = ok ? 1 : 0;
    }
  }
  recs[w] = rec;
}

// The earliest walk of the round that accepted each pixel.
__global__ void __launch_bounds__(kBlock) k_claim(const WalkRec* __restrict__ recs, const AccEntry* __restrict__ acc,
                                                  int32_t n, int32_t acc_cap, int32_t* __restrict__ first) {
  const int w = blockIdx.x * kBlock + threadIdx.x;
  if (w >= n) return;
  const int na = min(recs[w].n_acc, acc_cap);
  for (int j = 0; j < na; ++j) atomicMin(first + acc[(size_t)w * acc_cap + j].gid, w);
}

// Exact walks; the first inexact walk that was cut or overflowed bounds the
// round (stop).
__global__ void __launch_bounds__(kBlock) k_validate(const WalkRec* __restrict__ recs, const AccEntry* __restrict__ acc,
                                                     int32_t n, int32_t acc_cap, const int32_t* __restrict__ first,
                                                     uint8_t* __restrict__ exact, int32_t* __restrict__ stop) {
  const int w = blockIdx.x * kBlock + threadIdx.x;
  if (w >= n) return;
  const WalkRec r = recs[w];
  bool ok = !(r.flags & kOverflow);
  for (int j = 0; j < r.n_acc && ok; ++j) ok = first[acc[(size_t)w * acc_cap + j].gid] == w;
  exact[w] = ok ? 1 : 0;
  if (!ok && (r.flags & (kOverflow | kCut))) atomicMin(stop, w);
}

// Commit the kept walks (masks, point slots) and reset the claims.
__global__ void __launch_bounds__(kBlock) k_commit(const FusImgDev* __restrict__ imgs, const WalkRec* __restrict__ recs,
                                                   const AccEntry* __restrict__ acc, const int32_t* __restrict__ vis,
                                                   const int32_t* __restrict__ seeds, int32_t n, int32_t acc_cap,
                                                   int32_t window_first, const int32_t* __restrict__ stop,
                                                   uint8_t* __restrict__ exact, int32_t* __restrict__ first,
                                                   WalkRec* __restrict__ slots, int32_t* __restrict__ slot_vis) {
  const int w = blockIdx.x * kBlock + threadIdx.x;
  if (w >= n) return;
  const WalkRec r = recs[w];
  const int na = min(r.n_acc, acc_cap);
  const bool keep = exact[w] && w < *stop;
  for (int j = 0; j < na; ++j) {
    const AccEntry& a = acc[(size_t)w * acc_cap + j];
    first[a.gid] = INT_MAX;
    if (keep) imgs[a.img].mask[a.gid - imgs[a.img].base] = 1;
  }
  exact[w] = keep ? 1 : 0;
  if (!keep) return;
  const int slot = seeds[w] - window_first;
  slots[slot] = r;
  if (r.has_point && slot_vis)
    for (int j = 0; j < r.nvis; ++j) slot_vis[(size_t)slot * kAccCap + j] = vis[(size_t)w * acc_cap + j];
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
    if (x.width <= 6 || x.height <= 6 || (int64_t)x.width * x.height > ((int64_t)1 << 30) || !x.depth || !x.rgb ||
        (o->with_normal && !x.normal) || x.num_overlapping < 0 || (x.num_overlapping > 0 && !x.overlapping) ||
        !finite_camera(x.cam))
      return fail(PM_ERR_INVALID_ARG, "invalid fusion image " + std::to_string(i));
    max_nov = std::max(max_nov, x.num_overlapping);
  }
  PM_HIP(hipSetDevice(e->device));
  hipStream_t st = e->stream;
  hipEvent_t t0, t1;
  PM_HIP(hipEventCreate(&t0));
  PM_HIP(hipEventCreate(&t1));
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
  Buf d_recs, d_acc, d_stack, d_vis, d_seeds, d_exact, d_stop, d_slots, d_slot_vis, b_acc, b_stack, b_vis;
  if (pm_status s = d_recs.get((size_t)kWindow * sizeof(WalkRec))) return s;
  if (pm_status s = d_acc.get((size_t)kWindow * kAccCap * sizeof(AccEntry))) return s;
  if (pm_status s = d_stack.get((size_t)kWindow * kStackCap * sizeof(StackEntry))) return s;
  if (pm_status s = d_vis.get((size_t)kWindow * kAccCap * 4)) return s;
  if (pm_status s = d_seeds.get((size_t)kWindow * 4)) return s;
  if (pm_status s = d_exact.get((size_t)kWindow)) return s;
  if (pm_status s = d_stop.get(4)) return s;
  if (pm_status s = d_slots.get((size_t)kWindow * sizeof(WalkRec))) return s;
  if (pm_status s = d_slot_vis.get((size_t)kWindow * kAccCap * 4)) return s;

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
      while (!pending.empty()) {
        const int m = big_next ? 1 : (int)pending.size();
        AccEntry* acc = (AccEntry*)d_acc.p;
        StackEntry* stk = (StackEntry*)d_stack.p;
        int32_t* vis = (int32_t*)d_vis.p;
        prm.acc_cap = kAccCap;
        prm.stack_cap = kStackCap;
        if (big_next) {
          if (pm_status s = b_acc.get((size_t)big_acc * sizeof(AccEntry))) return s;
          if (pm_status s = b_stack.get((size_t)big_stack * sizeof(StackEntry))) return s;
          if (pm_status s = b_vis.get((size_t)big_acc * 4)) return s;
          acc = (AccEntry*)b_acc.p;
          stk = (StackEntry*)b_stack.p;
          vis = (int32_t*)b_vis.p;
          prm.acc_cap = big_acc;
          prm.stack_cap = (int32_t)big_stack;
        }
        PM_HIP(hipMemcpyAsync(d_seeds.p, pending.data(), (size_t)m * 4, hipMemcpyHostToDevice, st));
        const int32_t no_stop = INT_MAX;
        PM_HIP(hipMemcpyAsync(d_stop.p, &no_stop, 4, hipMemcpyHostToDevice, st));
        const dim3 grid((m + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_walk, grid, dim3(kBlock), 0, st, (const FusImgDev*)d_tab.p, (const uint8_t*)d_fused.p,
                           prm, (const int32_t*)d_seeds.p, m, (WalkRec*)d_recs.p, acc, stk, vis);
        hipLaunchKernelGGL(k_claim, grid, dim3(kBlock), 0, st, (const WalkRec*)d_recs.p, (const AccEntry*)acc, m,
                           prm.acc_cap, (int32_t*)d_first.p);
        hipLaunchKernelGGL(k_validate, grid, dim3(kBlock), 0, st, (const WalkRec*)d_recs.p, (const AccEntry*)acc, m,
                           prm.acc_cap, (const int32_t*)d_first.p, (uint8_t*)d_exact.p, (int32_t*)d_stop.p);
        hipLaunchKernelGGL(k_commit, grid, dim3(kBlock), 0, st, (const FusImgDev*)d_tab.p, (const WalkRec*)d_recs.p,
                           (const AccEntry*)acc, (const int32_t*)vis, (const int32_t*)d_seeds.p, m, prm.acc_cap,
                           (int32_t)w0, (const int32_t*)d_stop.p, (uint8_t*)d_exact.p, (int32_t*)d_first.p,
                           (WalkRec*)d_slots.p, big_next ? (int32_t*)nullptr : (int32_t*)d_slot_vis.p);
        PM_HIP(hipGetLastError());
        PM_HIP(hipMemcpyAsync(exact.data(), d_exact.p, (size_t)m, hipMemcpyDeviceToHost, st));
        WalkRec first_rec{};
        PM_HIP(hipMemcpyAsync(&first_rec, d_recs.p, sizeof(WalkRec), hipMemcpyDeviceToHost, st));
        PM_HIP(hipStreamSynchronize(st));
        stt.rounds++;
        stt.walks += m;
        if (big_next) {
          // run alone with buffers for max_num_pixels: always kept
          if (!exact[0]) return fail(PM_ERR_INTERNAL, "fusion: a walk run alone was not kept");
          stt.big_walks++;
          const int slot = pending[0] - (int32_t)w0;
          big_slot[slot] = 1;
          big_vis[slot].assign(first_rec.has_point ? first_rec.nvis : 0, 0);
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
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
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
