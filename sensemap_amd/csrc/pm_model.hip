// pm_model.hip -- source-view selection and depth ranges over the sparse
// model: the setup step ahead of the patch-match path.
//
// Model::GetMaxOverlappingImages (src/mvs/model.cc:474-699, :701-934) ranks,
// for every reference image, the images sharing sparse points with it.  The
// reference walks a hash map per image; here the work is three passes over
// one flat array of pair records:
//   1. k_pair_emit: one thread per observation (track entry) writes one
//      64-bit record per co-observation of the same point:
//        (image1 * I + image2) << 32 | flag << 31 | angle bits
//      where angle is the triangulation angle of the point seen from the two
//      camera centres (CalculateTriangulationAngle, base/triangulation.cc:
//      203-226, FP64 rounded to FP32 as the reference stores it) and flag
//      marks pairs outside the requested image subset.  Same-image pairs get
//      the all-ones record, which sorts last.
//   2. a stable radix sort (rocPRIM) groups records by pair and orders each
//      pair's angles; run-length encoding of the pair ids gives each pair's
//      shared-point count and first record, so the 75th percentile angle
//      (Percentile, util/math.h:259-274: element round(0.75 (n-1)) of the
//      sorted angles) is one load (k_pair_reduce).  The same pass folds the
//      nearest co-observing camera distance per image (ComputeSharedPoints,
//      model.cc:962-1000, and the mean-distance loop, model.cc:500-520).
//   3. pairs passing the angle threshold get a rank key
//      image1 << 32 | ~score (score = shared points, x1.5 truncated when both
//      rig flags agree, model.cc:564-570); a second stable sort orders each
//      image's candidates by descending score, then ascending image index,
//      and k_rank_bounds / k_gather cut each image's first max_num_images.
// Images that observe no point take SelectSpatialNeighborViews
// (model.cc:592-680) on the host: O(num_images) each and sequential (a
// greedy walk over a sorted list).
//
// ComputeDepthRanges (model.cc:941-960): the bounding box of the points and
// camera centres (Model::Read, model.cc:221-267) by a block reduction with
// order-preserving integer atomics, then each image's 8 corner depths.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "pm_host.h"

using pmh::fail;

namespace {

constexpr uint64_t kInvalidRecord = ~0ull;
constexpr uint32_t kOutsideSubset = 0x80000000u;
constexpr int kBlock = 256;

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

template <class T>
pm_status dev_alloc(DevBuf& b, size_t n, const char* what) {
  const hipError_t e = hipMalloc(&b.p, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? PM_ERR_OUT_OF_MEMORY : PM_ERR_DEVICE,
                std::string("model: cannot allocate ") + what + ": " + hipGetErrorString(e));
  return PM_OK;
}

// Sums of 3-vectors follow Eigen's fixed-size reductions on x86-64 (SSE2
// packets): FP32 (4-wide packets, none fits) a0 + (a1 + a2); FP64 (2-wide)
// (a0 + a1) + a2.
__host__ __device__ inline float dot3f(float a0, float a1, float a2, float b0, float b1, float b2) {
  return a0 * b0 + (a1 * b1 + a2 * b2);
}
__host__ __device__ inline double sq3(double x, double y, double z) { return (x * x + y * y) + z * z; }

// ComputeProjectionCenter (mvs/utils.cc:28-33): C = -R^T T in FP32.
__host__ __device__ void projection_center(const float* R, const float* T, float C[3]) {
  for (int i = 0; i < 3; ++i) C[i] = -dot3f(R[i], R[3 + i], R[6 + i], T[0], T[1], T[2]);
}

// CalculateTriangulationAngle (base/triangulation.cc:203-226) in FP64.
__device__ inline double tri_angle(const double* c1, const double* c2, double x, double y, double z) {
  const double base2 = sq3(c1[0] - c2[0], c1[1] - c2[1], c1[2] - c2[2]);
  const double ray1 = sq3(x - c1[0], y - c1[1], z - c1[2]);
  const double ray2 = sq3(x - c2[0], y - c2[1], z - c2[2]);
  const double den = 2.0 * sqrt(ray1 * ray2);
  if (den == 0.0) return 0.0;
  const double nom = (ray1 + ray2) - base2;
  const double a = fabs(acos(nom / den));
  const double b = M_PI - a;
  return (b < a) ? b : a;
}

// Exclusive-scan input: records of point p = L (L - 1); 0 past the end.
struct RecordsPerPoint {
  const int64_t* off;
  int64_t P;
  __device__ int64_t operator()(int64_t p) const {
    if (p >= P) return 0;
    const int64_t L = off[p + 1] - off[p];
    return L * (L - 1);
  }
};

struct PairId {
  __device__ uint32_t operator()(uint64_t k) const { return (uint32_t)(k >> 32); }
};

// Largest p in [0, P] with off[p] <= e (off non-decreasing, off[0] = 0).
__device__ inline int64_t point_of(const int64_t* __restrict__ off, int64_t P, int64_t e) {
  int64_t lo = 0, hi = P;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (off[mid] <= e)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(kBlock) k_pair_emit(const int64_t* __restrict__ off, const int32_t* __restrict__ trk,
                                                      int64_t P, int64_t T, const float* __restrict__ xyz,
                                                      const double* __restrict__ C, const uint8_t* __restrict__ mask,
                                                      uint32_t I, const int64_t* __restrict__ rec_base,
                                                      uint64_t* __restrict__ rec, uint8_t* __restrict__ observed) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= T) return;
  const int64_t p = point_of(off, P, e);
  const int64_t o = off[p];
  const int64_t L = off[p + 1] - o;
  const int64_t a = e - o;
  const uint32_t img1 = (uint32_t)trk[e];
  observed[img1] = 1;
  const double x = xyz[3 * p], y = xyz[3 * p + 1], z = xyz[3 * p + 2];
  const double* c1 = C + 3 * (size_t)img1;
  const bool in1 = mask == nullptr || mask[img1] != 0;
  uint64_t* out = rec + rec_base[p] + a * (L - 1);
  for (int64_t b = 0; b < L; ++b) {
    if (b == a) continue;
    const uint32_t img2 = (uint32_t)trk[o + b];
    uint64_t r = kInvalidRecord;
    if (img2 != img1) {
      const uint64_t pair = (uint64_t)img1 * I + img2;
      if (in1 && (mask == nullptr || mask[img2] != 0)) {
        const float ang = (float)tri_angle(c1, C + 3 * (size_t)img2, x, y, z);
        r = (pair << 32) | __float_as_uint(ang);
      } else {
        r = (pair << 32) | kOutsideSubset;
      }
    }
    *out++ = r;
  }
}

// Index of the first all-ones record in a sorted array (binary search).
__global__ void k_first_invalid(const uint64_t* __restrict__ k, int64_t n, int64_t* __restrict__ out) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (k[mid] == kInvalidRecord)
      hi = mid;
    else
      lo = mid + 1;
  }
  *out = lo;
}

// One thread per image pair (a run of sorted records): nearest co-observing
// centre per image, the pair's 75th percentile angle and its rank key.
__global__ void __launch_bounds__(kBlock) k_pair_reduce(
    const uint64_t* __restrict__ rec, const uint32_t* __restrict__ uniq, const uint32_t* __restrict__ cnt,
    const uint32_t* __restrict__ start, const uint32_t* __restrict__ nruns, uint32_t I, const double* __restrict__ C,
    const uint8_t* __restrict__ rig, float min_angle, uint32_t* __restrict__ min_dist_bits,
    uint64_t* __restrict__ rank_key, uint32_t* __restrict__ rank_img) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= *nruns) return;
  const uint32_t pair = uniq[s];
  const uint32_t img1 = pair / I, img2 = pair - img1 * I;
  const double* c1 = C + 3 * (size_t)img1;
  const double* c2 = C + 3 * (size_t)img2;
  const float dist = (float)sqrt(sq3(c1[0] - c2[0], c1[1] - c2[1], c1[2] - c2[2]));
  if (!((double)dist <= 1e-5)) atomicMin(min_dist_bits + img1, __float_as_uint(dist));
  const uint32_t n = cnt[s];
  const uint32_t first = start[s];
  uint64_t key = kInvalidRecord;
  if (!(rec[first] & kOutsideSubset)) {
    const uint32_t idx = (uint32_t)((3ull * (n - 1) + 2) / 4);  // round(0.75 (n - 1)), halves up
    const float ang = __uint_as_float((uint32_t)rec[first + idx]);
    if (ang >= min_angle) {
      const bool same_rig = (rig ? rig[img1] : 0) == (rig ? rig[img2] : 0);
      const uint64_t score = same_rig ? (3ull * n) / 2 : (uint64_t)n;
      const uint32_t sc = score > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)score;
      key = ((uint64_t)img1 << 32) | (0xFFFFFFFFu - sc);
    }
  }
  rank_key[s] = key;
  rank_img[s] = img2;
}

// First and one-past-last candidate of each image in the ranked list.
__global__ void __launch_bounds__(kBlock) k_rank_bounds(const uint64_t* __restrict__ key, const int64_t* __restrict__ nv,
                                                        int32_t* __restrict__ first, int32_t* __restrict__ last) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t n = *nv;
  if (k >= n) return;
  const uint32_t img = (uint32_t)(key[k] >> 32);
  if (k == 0 || (uint32_t)(key[k - 1] >> 32) != img) first[img] = (int32_t)k;
  if (k + 1 == n || (uint32_t)(key[k + 1] >> 32) != img) last[img] = (int32_t)(k + 1);
}

// out[out_off[i] + j] = ranked[first[i] + j] for j < take[i] (take[i] <=
// last[i] - first[i], the image's ranked candidates; 0 for the images the
// host fills).
__global__ void __launch_bounds__(kBlock) k_gather(const uint32_t* __restrict__ ranked, const int32_t* __restrict__ first,
                                                   const int32_t* __restrict__ out_off, const int32_t* __restrict__ take,
                                                   int32_t I, int32_t max_n, int32_t* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (int64_t)I * max_n) return;
  const int32_t i = (int32_t)(t / max_n), j = (int32_t)(t - (int64_t)i * max_n);
  if (j < take[i]) out[out_off[i] + j] = (int32_t)ranked[first[i] + j];
}

__device__ inline uint32_t ordered_bits(float f) {
  const uint32_t u = __float_as_uint(f + 0.0f);  // -0 -> +0
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float from_ordered(uint32_t o) {
  const uint32_t u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// Bounding box of the points and camera centres: per block LDS min/max, then
// one integer atomic per coordinate.  NaN coordinates are skipped, as
// std::min / std::max skip a NaN second argument.
__global__ void __launch_bounds__(kBlock) k_bbox(const float* __restrict__ xyz, int64_t P,
                                                 const float* __restrict__ centres, int32_t I,
                                                 uint32_t* __restrict__ box) {
  __shared__ uint32_t lo[3][kBlock], hi[3][kBlock];
  uint32_t l[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, h[3] = {0u, 0u, 0u};
  const int64_t n = P + I;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBlock) {
    const float* v = k < P ? xyz + 3 * k : centres + 3 * (k - P);
    for (int c = 0; c < 3; ++c) {
      const float f = v[c];
      if (f != f) continue;
      const uint32_t o = ordered_bits(f);
      l[c] = min(l[c], o);
      h[c] = max(h[c], o);
    }
  }
  for (int c = 0; c < 3; ++c) {
    lo[c][threadIdx.x] = l[c];
    hi[c][threadIdx.x] = h[c];
  }
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < 3; ++c) {
        lo[c][threadIdx.x] = min(lo[c][threadIdx.x], lo[c][threadIdx.x + w]);
        hi[c][threadIdx.x] = max(hi[c][threadIdx.x], hi[c][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; ++c) {
      atomicMin(box + c, lo[c][0]);
      atomicMax(box + 3 + c, hi[c][0]);
    }
}

// ComputeDepthRangeFromBBox (model.cc:22-35) per image.
__global__ void __launch_bounds__(kBlock) k_depth_ranges(const uint32_t* __restrict__ box, const float* __restrict__ R,
                                                         const float* __restrict__ T, int32_t I,
                                                         float* __restrict__ out) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= I) return;
  float lt[3], rb[3];
  for (int c = 0; c < 3; ++c) {
    lt[c] = box[c] == 0xFFFFFFFFu ? FLT_MAX : from_ordered(box[c]);
    rb[c] = box[3 + c] == 0u ? -FLT_MAX : from_ordered(box[3 + c]);
  }
  const float* r = R + 9 * (size_t)i;
  const float t = T[3 * (size_t)i + 2];
  // corner order of BoundingBox::GetCorners (util/bb.h:86-96)
  const int sel[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  float dmin = 0.0f, dmax = 0.0f;
  for (int k = 0; k < 8; ++k) {
    const float x = sel[k][0] ? rb[0] : lt[0], y = sel[k][1] ? rb[1] : lt[1], zc = sel[k][2] ? rb[2] : lt[2];
    const float d = dot3f(r[6], r[7], r[8], x, y, zc) + t;
    if (k == 0) {
      dmin = dmax = d;
    } else {
      dmin = (d < dmin) ? d : dmin;
      dmax = (dmax < d) ? d : dmax;
    }
  }
  out[2 * i] = (0.000001f < dmin) ? dmin : 0.000001f;
  out[2 * i + 1] = dmax;
}

pm_status check_model(const pm_model* m) {
  if (!m || m->num_images <= 0 || m->num_images > 65535 || !m->R || !m->T || m->num_points < 0 ||
      (m->num_points > 0 && (!m->xyz || !m->track_offsets)))
    return fail(PM_ERR_INVALID_ARG, "invalid sparse model (1..65535 images, poses, points and track offsets)");
  return PM_OK;
}

// Order-preserving unsigned key of a float; every NaN maps above +inf.
uint32_t total_order(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if (f != f) return 0xFFFFFFFFu;
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// SelectSpatialNeighborViews (model.cc:592-680) for reference image `ref`.
std::vector<int32_t> spatial_neighbours(const pm_model* m, const std::vector<double>& Cd, int32_t ref, float m_dist,
                                        size_t max_num_images) {
  const int32_t I = m->num_images;
  const float min_dist = (float)(0.1 * m_dist), max_dist = (float)(100 * m_dist);
  const float min_angle = (float)(0.0 * 0.0174532925199432954743716805978692718781530857086181640625);
  const float max_angle = (float)(60.0 * 0.0174532925199432954743716805978692718781530857086181640625);
  const float* rr = m->R + 9 * (size_t)ref;
  auto norm = [&](int32_t a, int32_t b) {
    const double* ca = &Cd[3 * (size_t)a];
    const double* cb = &Cd[3 * (size_t)b];
    return std::sqrt(sq3(ca[0] - cb[0], ca[1] - cb[1], ca[2] - cb[2]));
  };
  struct Scored {
    uint32_t key;
    int32_t idx;
  };
  std::vector<Scored> scored;
  for (int32_t i = 0; i < I; ++i) {
    if (i == ref) continue;
    const float* sr = m->R + 9 * (size_t)i;
    const float dist = (float)norm(i, ref);
    if (dist < min_dist || dist > max_dist) continue;
    const float cosv = dot3f(rr[6], rr[7], rr[8], sr[6], sr[7], sr[8]);
    const float angle = std::acos(cosv);
    if (angle < min_angle || angle > max_angle) continue;
    const float wangle = 1.1f - cosv;
    const float q = dist / m_dist;
    const float wdist = (q < 1.0f) ? 1.0f : q;
    scored.push_back({total_order(wangle * wdist), i});
  }
  std::sort(scored.begin(), scored.end(),
            [](const Scored& a, const Scored& b) { return a.key != b.key ? a.key < b.key : a.idx < b.idx; });
  const size_t eff = std::min(scored.size(), max_num_images);
  std::vector<int32_t> chosen;
  if (eff < 1) return chosen;
  chosen.push_back(ref);
  while (chosen.size() <= eff && !scored.empty()) {
    size_t k = 0;
    for (; k < scored.size(); ++k) {
      bool spurious = false;
      for (int32_t c : chosen)
        if (norm(scored[k].idx, c) < (double)min_dist) {
          spurious = true;
          break;
        }
      if (!spurious) {
        chosen.push_back(scored[k].idx);
        scored.erase(scored.begin() + (ptrdiff_t)k);
        break;
      }
    }
    if (k >= scored.size()) break;
  }
  chosen.erase(chosen.begin());
  return chosen;
}

}  // namespace

PM_API pm_status pm_model_max_overlapping_images(pm_engine* e, const pm_model* m, int32_t max_num_images,
                                                 double min_triangulation_angle, const uint8_t* image_mask,
                                                 int32_t* out_offsets, int32_t* out_images, pm_overlap_stats* stats) {
  if (!e || !out_offsets || !out_images || max_num_images < 0)
    return fail(PM_ERR_INVALID_ARG, "invalid overlap-selection arguments");
  if (pm_status s = check_model(m)) return s;
  const int32_t I = m->num_images;
  const int64_t P = m->num_points;
  const int64_t T = P > 0 ? m->track_offsets[P] : 0;
  max_num_images = std::min(max_num_images, I);  // a list holds at most I - 1 images
  if (P > 0 && m->track_offsets[0] != 0) return fail(PM_ERR_INVALID_ARG, "track offsets must start at 0");
  int64_t R = 0;
  for (int64_t p = 0; p < P; ++p) {
    const int64_t L = m->track_offsets[p + 1] - m->track_offsets[p];
    if (L < 0) return fail(PM_ERR_INVALID_ARG, "track offsets must be non-decreasing");
    R += L * (L - 1);
    if (R >= ((int64_t)1 << 32)) return fail(PM_ERR_INVALID_ARG, "more than 2^32 pair records in one call");
  }
  if (T > 0 && !m->track_images) return fail(PM_ERR_INVALID_ARG, "missing track image indices");
  for (int64_t t = 0; t < T; ++t)
    if (m->track_images[t] < 0 || m->track_images[t] >= I)
      return fail(PM_ERR_INVALID_ARG, "track image index out of range at entry " + std::to_string(t));
  PM_HIP(hipSetDevice(e->device));
  hipStream_t st = e->stream;
  const float min_angle = (float)(min_triangulation_angle * 0.0174532925199432954743716805978692718781530857086181640625);

  // camera centres (FP32, then FP64 as the reference holds them)
  std::vector<double> Cd(3 * (size_t)I);
  for (int32_t i = 0; i < I; ++i) {
    float c[3];
    projection_center(m->R + 9 * (size_t)i, m->T + 3 * (size_t)i, c);
    for (int k = 0; k < 3; ++k) Cd[3 * (size_t)i + k] = c[k];
  }

  DevBuf d_off, d_trk, d_xyz, d_C, d_mask, d_rig, d_base, d_rec, d_rec2, d_obs, d_mind, d_tmp;
  DevBuf d_uniq, d_cnt, d_start, d_nruns, d_key, d_img, d_key2, d_img2, d_nv, d_first, d_last, d_ooff, d_take, d_out;
  pm_status s;
  if ((s = dev_alloc<int64_t>(d_off, (size_t)P + 1, "track offsets")) || (s = dev_alloc<int32_t>(d_trk, T, "tracks")) ||
      (s = dev_alloc<float>(d_xyz, 3 * (size_t)P, "points")) || (s = dev_alloc<double>(d_C, 3 * (size_t)I, "centres")) ||
      (s = dev_alloc<uint8_t>(d_mask, I, "subset")) || (s = dev_alloc<uint8_t>(d_rig, I, "rig flags")) ||
      (s = dev_alloc<int64_t>(d_base, (size_t)P + 1, "record offsets")) ||
      (s = dev_alloc<uint64_t>(d_rec, R, "records")) || (s = dev_alloc<uint64_t>(d_rec2, R, "sorted records")) ||
      (s = dev_alloc<uint8_t>(d_obs, I, "observed flags")) || (s = dev_alloc<uint32_t>(d_mind, I, "distances")) ||
      (s = dev_alloc<int64_t>(d_nv, 2, "counts")) || (s = dev_alloc<uint32_t>(d_nruns, 1, "run count")))
    return s;
  if (P > 0) {
    PM_HIP(hipMemcpyAsync(d_off.p, m->track_offsets, ((size_t)P + 1) * 8, hipMemcpyHostToDevice, st));
    PM_HIP(hipMemcpyAsync(d_xyz.p, m->xyz, 3 * (size_t)P * 4, hipMemcpyHostToDevice, st));
  }
  if (T > 0) PM_HIP(hipMemcpyAsync(d_trk.p, m->track_images, (size_t)T * 4, hipMemcpyHostToDevice, st));
  PM_HIP(hipMemcpyAsync(d_C.p, Cd.data(), Cd.size() * 8, hipMemcpyHostToDevice, st));
  if (image_mask) PM_HIP(hipMemcpyAsync(d_mask.p, image_mask, (size_t)I, hipMemcpyHostToDevice, st));
  if (m->is_rig) PM_HIP(hipMemcpyAsync(d_rig.p, m->is_rig, (size_t)I, hipMemcpyHostToDevice, st));
  PM_HIP(hipMemsetAsync(d_obs.p, 0, (size_t)I, st));
  {
    std::vector<uint32_t> init(I);
    const float fmax = FLT_MAX;
    uint32_t fb;
    std::memcpy(&fb, &fmax, 4);
    std::fill(init.begin(), init.end(), fb);
    PM_HIP(hipMemcpyAsync(d_mind.p, init.data(), (size_t)I * 4, hipMemcpyHostToDevice, st));
  }

  const unsigned pair_bits = 32 + (unsigned)std::ceil(std::log2((double)I * I + 1.0));
  const unsigned key_bits = std::min(64u, pair_bits);
  const unsigned rank_bits = 32 + (unsigned)std::ceil(std::log2((double)I + 1.0));
  pmh::EventTimer tm;
  if ((s = tm.rec(st))) return s;
  // 1. records
  auto scan_in = rocprim::make_transform_iterator(rocprim::make_counting_iterator<int64_t>(0),
                                                  RecordsPerPoint{(const int64_t*)d_off.p, P});
  size_t tmp_bytes = 0, need = 0;
  PM_HIP(rocprim::exclusive_scan(nullptr, need, scan_in, (int64_t*)d_base.p, (int64_t)0, (size_t)P + 1,
                                 rocprim::plus<int64_t>(), st));
  tmp_bytes = std::max(tmp_bytes, need);
  PM_HIP(rocprim::radix_sort_keys(nullptr, need, (const uint64_t*)d_rec.p, (uint64_t*)d_rec2.p, (size_t)R, 0, key_bits,
                                  st));
  tmp_bytes = std::max(tmp_bytes, need);
  PM_HIP(rocprim::run_length_encode(nullptr, need, rocprim::make_transform_iterator((const uint64_t*)d_rec2.p, PairId{}),
                                    (unsigned)R, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr, st));
  tmp_bytes = std::max(tmp_bytes, need);
  PM_HIP(rocprim::exclusive_scan(nullptr, need, (const uint32_t*)nullptr, (uint32_t*)nullptr, 0u, (size_t)R,
                                 rocprim::plus<uint32_t>(), st));
  tmp_bytes = std::max(tmp_bytes, need);
  PM_HIP(rocprim::radix_sort_pairs(nullptr, need, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                   (const uint32_t*)nullptr, (uint32_t*)nullptr, (size_t)R, 0, rank_bits, st));
  tmp_bytes = std::max(tmp_bytes, need);
  if ((s = dev_alloc<uint8_t>(d_tmp, tmp_bytes, "sort workspace"))) return s;

  PM_HIP(rocprim::exclusive_scan(d_tmp.p, tmp_bytes, scan_in, (int64_t*)d_base.p, (int64_t)0, (size_t)P + 1,
                                 rocprim::plus<int64_t>(), st));
  if (T > 0) {
    hipLaunchKernelGGL(k_pair_emit, dim3((unsigned)((T + kBlock - 1) / kBlock)), dim3(kBlock), 0, st,
                       (const int64_t*)d_off.p, (const int32_t*)d_trk.p, P, T, (const float*)d_xyz.p,
                       (const double*)d_C.p, image_mask ? (const uint8_t*)d_mask.p : nullptr, (uint32_t)I,
                       (const int64_t*)d_base.p, (uint64_t*)d_rec.p, (uint8_t*)d_obs.p);
    PM_HIP(hipGetLastError());
  }
  if ((s = tm.rec(st))) return s;
  // 2. group by pair
  size_t b = tmp_bytes;
  if (R > 0)
    PM_HIP(rocprim::radix_sort_keys(d_tmp.p, b, (const uint64_t*)d_rec.p, (uint64_t*)d_rec2.p, (size_t)R, 0, key_bits,
                                    st));
  hipLaunchKernelGGL(k_first_invalid, dim3(1), dim3(1), 0, st, (const uint64_t*)d_rec2.p, R, (int64_t*)d_nv.p);
  PM_HIP(hipGetLastError());
  int64_t nvalid = 0;
  PM_HIP(hipMemcpyAsync(&nvalid, d_nv.p, 8, hipMemcpyDeviceToHost, st));
  PM_HIP(hipStreamSynchronize(st));
  if ((s = tm.rec(st))) return s;
  // 3. per pair
  if ((s = dev_alloc<uint32_t>(d_uniq, nvalid, "pairs")) || (s = dev_alloc<uint32_t>(d_cnt, nvalid, "pair counts")) ||
      (s = dev_alloc<uint32_t>(d_start, nvalid, "pair starts")) ||
      (s = dev_alloc<uint64_t>(d_key, nvalid, "rank keys")) || (s = dev_alloc<uint32_t>(d_img, nvalid, "rank images")) ||
      (s = dev_alloc<uint64_t>(d_key2, nvalid, "ranked keys")) ||
      (s = dev_alloc<uint32_t>(d_img2, nvalid, "ranked images")))
    return s;
  uint32_t nruns = 0;
  if (nvalid > 0) {
    b = tmp_bytes;
    PM_HIP(rocprim::run_length_encode(d_tmp.p, b, rocprim::make_transform_iterator((const uint64_t*)d_rec2.p, PairId{}),
                                      (unsigned)nvalid, (uint32_t*)d_uniq.p, (uint32_t*)d_cnt.p, (uint32_t*)d_nruns.p,
                                      st));
    PM_HIP(hipMemcpyAsync(&nruns, d_nruns.p, 4, hipMemcpyDeviceToHost, st));
    PM_HIP(hipStreamSynchronize(st));
    b = tmp_bytes;
    PM_HIP(rocprim::exclusive_scan(d_tmp.p, b, (const uint32_t*)d_cnt.p, (uint32_t*)d_start.p, 0u, (size_t)nruns,
                                   rocprim::plus<uint32_t>(), st));
    hipLaunchKernelGGL(k_pair_reduce, dim3((nruns + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                       (const uint64_t*)d_rec2.p, (const uint32_t*)d_uniq.p, (const uint32_t*)d_cnt.p,
                       (const uint32_t*)d_start.p, (const uint32_t*)d_nruns.p, (uint32_t)I, (const double*)d_C.p,
                       m->is_rig ? (const uint8_t*)d_rig.p : nullptr, min_angle, (uint32_t*)d_mind.p,
                       (uint64_t*)d_key.p, (uint32_t*)d_img.p);
    PM_HIP(hipGetLastError());
  }
  if ((s = tm.rec(st))) return s;
  // 4. rank
  if ((s = dev_alloc<int32_t>(d_first, I, "ranked bounds")) || (s = dev_alloc<int32_t>(d_last, I, "ranked bounds")) ||
      (s = dev_alloc<int32_t>(d_ooff, (size_t)I + 1, "output offsets")) ||
      (s = dev_alloc<int32_t>(d_take, (size_t)I, "output counts")) ||
      (s = dev_alloc<int32_t>(d_out, (size_t)I * std::max(max_num_images, 1), "output")))
    return s;
  PM_HIP(hipMemsetAsync(d_first.p, 0, (size_t)I * 4, st));
  PM_HIP(hipMemsetAsync(d_last.p, 0, (size_t)I * 4, st));
  if (nruns > 0) {
    b = tmp_bytes;
    PM_HIP(rocprim::radix_sort_pairs(d_tmp.p, b, (const uint64_t*)d_key.p, (uint64_t*)d_key2.p,
                                     (const uint32_t*)d_img.p, (uint32_t*)d_img2.p, (size_t)nruns, 0, rank_bits,
                                     st));
    hipLaunchKernelGGL(k_first_invalid, dim3(1), dim3(1), 0, st, (const uint64_t*)d_key2.p, (int64_t)nruns,
                       (int64_t*)d_nv.p + 1);
    PM_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_rank_bounds, dim3((nruns + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                       (const uint64_t*)d_key2.p, (const int64_t*)d_nv.p + 1, (int32_t*)d_first.p,
                       (int32_t*)d_last.p);
    PM_HIP(hipGetLastError());
  }
  std::vector<int32_t> first(I), last(I);
  std::vector<uint8_t> observed(I);
  std::vector<uint32_t> mind(I);
  PM_HIP(hipMemcpyAsync(first.data(), d_first.p, (size_t)I * 4, hipMemcpyDeviceToHost, st));
  PM_HIP(hipMemcpyAsync(last.data(), d_last.p, (size_t)I * 4, hipMemcpyDeviceToHost, st));
  PM_HIP(hipMemcpyAsync(observed.data(), d_obs.p, (size_t)I, hipMemcpyDeviceToHost, st));
  PM_HIP(hipMemcpyAsync(mind.data(), d_mind.p, (size_t)I * 4, hipMemcpyDeviceToHost, st));
  PM_HIP(hipStreamSynchronize(st));

  // spatial fallback for images that observe no point
  std::vector<std::vector<int32_t>> spatial(I);
  int32_t nfallback = 0;
  bool have_mdist = false;
  float m_dist = 0.0f;
  std::vector<int32_t> ooff((size_t)I + 1, 0), take(I, 0);
  for (int32_t i = 0; i < I; ++i) {
    int32_t n = 0;
    if (!image_mask || image_mask[i]) {
      if (observed[i]) {
        n = std::min(last[i] - first[i], max_num_images);
        take[i] = n;
      } else {
        if (!have_mdist) {
          int m_cnt = 0;
          for (int32_t k = 0; k < I; ++k) {
            float d;
            std::memcpy(&d, &mind[k], 4);
            if (d < FLT_MAX) {
              m_dist += d;
              m_cnt++;
            }
          }
          m_dist /= m_cnt;
          have_mdist = true;
        }
        spatial[i] = spatial_neighbours(m, Cd, i, m_dist, (size_t)max_num_images);
        n = (int32_t)spatial[i].size();
        nfallback++;
      }
    }
    ooff[i + 1] = ooff[i] + n;
  }
  PM_HIP(hipMemcpyAsync(d_ooff.p, ooff.data(), ooff.size() * 4, hipMemcpyHostToDevice, st));
  PM_HIP(hipMemcpyAsync(d_take.p, take.data(), take.size() * 4, hipMemcpyHostToDevice, st));
  if (max_num_images > 0 && nruns > 0) {
    hipLaunchKernelGGL(k_gather, dim3((unsigned)(((int64_t)I * max_num_images + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       st, (const uint32_t*)d_img2.p, (const int32_t*)d_first.p, (const int32_t*)d_ooff.p,
                       (const int32_t*)d_take.p, I, max_num_images, (int32_t*)d_out.p);
    PM_HIP(hipGetLastError());
  }
  if ((s = tm.rec(st))) return s;
  if (ooff[I] > 0) PM_HIP(hipMemcpyAsync(out_images, d_out.p, (size_t)ooff[I] * 4, hipMemcpyDeviceToHost, st));
  PM_HIP(hipStreamSynchronize(st));
  // the gather skipped the fallback images (no ranked candidates): fill them
  for (int32_t i = 0; i < I; ++i)
    if (!spatial[i].empty()) std::copy(spatial[i].begin(), spatial[i].end(), out_images + ooff[i]);
  std::copy(ooff.begin(), ooff.end(), out_offsets);
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->pair_records = R;
    stats->valid_records = nvalid;
    stats->image_pairs = nruns;
    stats->fallback_images = nfallback;
    stats->emit_ms = tm.between(0, 1);
    stats->sort_ms = tm.between(1, 2);
    stats->reduce_ms = tm.between(2, 3);
    stats->rank_ms = tm.between(3, 4);
    stats->total_ms = tm.between(0, 4);
  }
  return PM_OK;
}

PM_API pm_status pm_model_depth_ranges(pm_engine* e, const pm_model* m, float* out) {
  if (!e || !out) return fail(PM_ERR_INVALID_ARG, "invalid depth-range arguments");
  if (pm_status s = check_model(m)) return s;
  const int32_t I = m->num_images;
  const int64_t P = m->num_points;
  PM_HIP(hipSetDevice(e->device));
  hipStream_t st = e->stream;
  std::vector<float> centres(3 * (size_t)I);
  for (int32_t i = 0; i < I; ++i) projection_center(m->R + 9 * (size_t)i, m->T + 3 * (size_t)i, &centres[3 * (size_t)i]);
  DevBuf d_xyz, d_c, d_R, d_T, d_box, d_out;
  pm_status s;
  if ((s = dev_alloc<float>(d_xyz, 3 * (size_t)P, "points")) || (s = dev_alloc<float>(d_c, 3 * (size_t)I, "centres")) ||
      (s = dev_alloc<float>(d_R, 9 * (size_t)I, "rotations")) ||
      (s = dev_alloc<float>(d_T, 3 * (size_t)I, "translations")) || (s = dev_alloc<uint32_t>(d_box, 6, "box")) ||
      (s = dev_alloc<float>(d_out, 2 * (size_t)I, "ranges")))
    return s;
  if (P > 0) PM_HIP(hipMemcpyAsync(d_xyz.p, m->xyz, 3 * (size_t)P * 4, hipMemcpyHostToDevice, st));
  PM_HIP(hipMemcpyAsync(d_c.p, centres.data(), centres.size() * 4, hipMemcpyHostToDevice, st));
  PM_HIP(hipMemcpyAsync(d_R.p, m->R, 9 * (size_t)I * 4, hipMemcpyHostToDevice, st));
  PM_HIP(hipMemcpyAsync(d_T.p, m->T, 3 * (size_t)I * 4, hipMemcpyHostToDevice, st));
  const uint32_t init[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
  PM_HIP(hipMemcpyAsync(d_box.p, init, sizeof(init), hipMemcpyHostToDevice, st));
  const int64_t n = P + I;
  const unsigned blocks = (unsigned)std::min<int64_t>((n + kBlock - 1) / kBlock, 2048);
  hipLaunchKernelGGL(k_bbox, dim3(blocks), dim3(kBlock), 0, st, (const float*)d_xyz.p, P, (const float*)d_c.p, I,
                     (uint32_t*)d_box.p);
  PM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_depth_ranges, dim3((I + kBlock - 1) / kBlock), dim3(kBlock), 0, st, (const uint32_t*)d_box.p,
                     (const float*)d_R.p, (const float*)d_T.p, I, (float*)d_out.p);
  PM_HIP(hipGetLastError());
  PM_HIP(hipMemcpyAsync(out, d_out.p, 2 * (size_t)I * 4, hipMemcpyDeviceToHost, st));
  PM_HIP(hipStreamSynchronize(st));
  return PM_OK;
}

// ---------------------------------------------------------------------------
// pm_model_depth_ranges_noalloc: the same ranges from device-resident inputs,
// with the camera centres and the bounding box in the library's fixed device
// array `scratch` instead of allocated buffers.  Layout of `scratch` (32-bit
// words): box at [0, 6) (order-preserving min x, y, z, then max x, y, z),
// camera centres at [kScratchCentres, kScratchCentres + 3 * num_images).

namespace {

constexpr int32_t kNoAllocMaxImages = PM_DEPTH_RANGES_NOALLOC_MAX_IMAGES;
constexpr size_t kScratchCentres = 64;  // 256 bytes: the centres start on their own line
constexpr size_t kScratchWords = kScratchCentres + 3 * (size_t)kNoAllocMaxImages;

// Each GPU holds its own copy until the process ends.
static __device__ __align__(256) uint32_t scratch[kScratchWords];

// The per-device address of `scratch` (the host-side name is a stand-in and
// must never reach a kernel), looked up once per device.
std::mutex g_scratch_mu;
void* g_scratch_ptr[256] = {};

pm_status scratch_address(int device, uint32_t** out) {
  if (device < 0 || device >= 256) return fail(PM_ERR_INVALID_ARG, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(PM_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = p;
  }
  *out = (uint32_t*)g_scratch_ptr[device];
  return PM_OK;
}

// The box starts empty (a kernel, not a memset).
__global__ void k_box_init(uint32_t* __restrict__ box) {
  if (threadIdx.x < 6) box[threadIdx.x] = threadIdx.x < 3 ? 0xFFFFFFFFu : 0u;
}

__global__ void __launch_bounds__(kBlock) k_centres(const float* __restrict__ R, const float* __restrict__ T, int32_t I,
                                                    float* __restrict__ centres) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= I) return;
  float c[3];
  projection_center(R + 9 * (size_t)i, T + 3 * (size_t)i, c);
  for (int k = 0; k < 3; ++k) centres[3 * (size_t)i + k] = c[k];
}

}  // namespace

PM_API pm_status pm_model_depth_ranges_noalloc(pm_engine* e, int32_t num_images, const float* d_R, const float* d_T,
                                               int64_t num_points, const float* d_xyz, float* d_out) {
  if (!e || !d_R || !d_T || !d_out || num_points < 0 || (num_points > 0 && !d_xyz))
    return fail(PM_ERR_INVALID_ARG, "invalid depth-range arguments");
  if (num_images < 1 || num_images > kNoAllocMaxImages)
    return fail(PM_ERR_INVALID_ARG, "num_images must be in [1, " + std::to_string(kNoAllocMaxImages) +
                                        "] (the size of the fixed device array)");
  PM_HIP(hipSetDevice(e->device));
  uint32_t* words = nullptr;
  if (pm_status s = scratch_address(e->device, &words)) return s;
  hipStream_t st = e->stream;
  uint32_t* box = words;
  float* centres = (float*)(words + kScratchCentres);
  hipLaunchKernelGGL(k_box_init, dim3(1), dim3(64), 0, st, box);
  PM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_centres, dim3((num_images + kBlock - 1) / kBlock), dim3(kBlock), 0, st, d_R, d_T, num_images,
                     centres);
  PM_HIP(hipGetLastError());
  const int64_t n = num_points + num_images;
  const unsigned blocks = (unsigned)std::min<int64_t>((n + kBlock - 1) / kBlock, 2048);
  hipLaunchKernelGGL(k_bbox, dim3(blocks), dim3(kBlock), 0, st, d_xyz, num_points, (const float*)centres, num_images,
                     box);
  PM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_depth_ranges, dim3((num_images + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                     (const uint32_t*)box, d_R, d_T, num_images, d_out);
  PM_HIP(hipGetLastError());
  PM_HIP(hipStreamSynchronize(st));
  return PM_OK;
}
