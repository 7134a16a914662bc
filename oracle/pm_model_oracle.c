/* pm_model_oracle.c -- CPU transcription of the sparse model's source-view
 * selection and depth ranges.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the device selection
 * (sensemap_amd/csrc/pm_model.hip) is compared against, and a CPU timing
 * beside it.  The product never links or calls it.
 *
 * It keeps the reference's structure (a per-image walk over the image's
 * points with per-neighbour counters and angle lists) where the device sorts
 * one flat record array, so the two share no code path:
 *   src/mvs/model.cc:701-934    Model::GetMaxOverlappingImages (image subset;
 *                               :474-699 is the same with every image)
 *   src/mvs/model.cc:962-1000   ComputeSharedPoints (which images co-observe)
 *   src/mvs/model.cc:941-960    ComputeDepthRanges
 *   src/mvs/model.cc:22-35      ComputeDepthRangeFromBBox
 *   src/mvs/model.cc:221-267    the bounding box Model::Read builds
 *   src/mvs/utils.cc:28-33      ComputeProjectionCenter
 *   src/base/triangulation.cc:203-226  CalculateTriangulationAngle
 *   src/util/math.h:259-274     Percentile
 * Where the reference leaves an order open (equal scores: hash-map order
 * and an unstable sort) the lower image index ranks first.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pm_oracle.h"

#define DEG2RAD 0.0174532925199432954743716805978692718781530857086181640625

/* Eigen fixed-size reduction orders (SSE2): FP32 a0 + (a1 + a2), FP64
 * (a0 + a1) + a2. */
static float dot3f(float a0, float a1, float a2, float b0, float b1, float b2) { return a0 * b0 + (a1 * b1 + a2 * b2); }
static double sq3(double x, double y, double z) { return (x * x + y * y) + z * z; }

static void proj_center(const float* R, const float* T, float C[3]) {
  for (int i = 0; i < 3; ++i) C[i] = -dot3f(R[i], R[3 + i], R[6 + i], T[0], T[1], T[2]);
}

static double tri_angle(const double* c1, const double* c2, const double* X) {
  const double base2 = sq3(c1[0] - c2[0], c1[1] - c2[1], c1[2] - c2[2]);
  const double ray1 = sq3(X[0] - c1[0], X[1] - c1[1], X[2] - c1[2]);
  const double ray2 = sq3(X[0] - c2[0], X[1] - c2[1], X[2] - c2[2]);
  const double den = 2.0 * sqrt(ray1 * ray2);
  if (den == 0.0) return 0.0;
  const double nom = (ray1 + ray2) - base2;
  const double angle = fabs(acos(nom / den));
  const double other = M_PI - angle;
  return (other < angle) ? other : angle; /* std::min(angle, pi - angle) */
}

static double centre_dist(const double* C, int a, int b) {
  const double* ca = C + 3 * (size_t)a;
  const double* cb = C + 3 * (size_t)b;
  return sqrt(sq3(ca[0] - cb[0], ca[1] - cb[1], ca[2] - cb[2]));
}

/* float order with every NaN above +inf (nth_element's order for non-NaN). */
static uint32_t fkey(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if (f != f) return 0xFFFFFFFFu;
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
static int cmp_fkey(const void* a, const void* b) {
  const uint32_t x = fkey(*(const float*)a), y = fkey(*(const float*)b);
  return x < y ? -1 : x > y;
}

typedef struct {
  int idx;
  long long score;
} ranked_t;
static int cmp_desc(const void* a, const void* b) {
  const ranked_t *x = a, *y = b;
  if (x->score != y->score) return x->score > y->score ? -1 : 1;
  return x->idx - y->idx;
}
typedef struct {
  int idx;
  uint32_t key;
} scored_t;
static int cmp_asc(const void* a, const void* b) {
  const scored_t *x = a, *y = b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->idx - y->idx;
}

/* SelectSpatialNeighborViews (model.cc:592-680).  Returns the count written. */
static int spatial(const pm_model* m, const double* C, int ref, float m_dist, int max_num, int32_t* out) {
  const int I = m->num_images;
  const float min_dist = (float)(0.1 * m_dist), max_dist = (float)(100 * m_dist);
  const float min_angle = (float)(0.0 * DEG2RAD), max_angle = (float)(60.0 * DEG2RAD);
  const float* rr = m->R + 9 * (size_t)ref;
  scored_t* sc = malloc(sizeof(scored_t) * (size_t)I);
  int ns = 0;
  for (int i = 0; i < I; ++i) {
    if (i == ref) continue;
    const float* sr = m->R + 9 * (size_t)i;
    const float dist = (float)centre_dist(C, i, ref);
    if (dist < min_dist || dist > max_dist) continue;
    const float cosv = dot3f(rr[6], rr[7], rr[8], sr[6], sr[7], sr[8]);
    const float angle = acosf(cosv);
    if (angle < min_angle || angle > max_angle) continue;
    const float wangle = 1.1f - cosv;
    const float q = dist / m_dist;
    const float wdist = (q < 1.0f) ? 1.0f : q; /* std::max(q, 1.0f) */
    sc[ns].idx = i;
    sc[ns].key = fkey(wangle * wdist);
    ns++;
  }
  qsort(sc, (size_t)ns, sizeof(scored_t), cmp_asc);
  const int eff = ns < max_num ? ns : max_num;
  int n = 0;
  if (eff >= 1) {
    /* chosen[0] is the reference itself, dropped at the end */
    int* chosen = malloc(sizeof(int) * ((size_t)eff + 2));
    int nc = 1;
    chosen[0] = ref;
    while (nc <= eff && ns > 0) {
      int k;
      for (k = 0; k < ns; ++k) {
        int spurious = 0;
        for (int c = 0; c < nc; ++c)
          if (centre_dist(C, sc[k].idx, chosen[c]) < (double)min_dist) {
            spurious = 1;
            break;
          }
        if (!spurious) {
          chosen[nc++] = sc[k].idx;
          memmove(sc + k, sc + k + 1, sizeof(scored_t) * (size_t)(ns - k - 1));
          ns--;
          break;
        }
      }
      if (k >= ns) break;
    }
    for (int c = 1; c < nc; ++c) out[n++] = chosen[c];
    free(chosen);
  }
  free(sc);
  return n;
}

int pmo_model_max_overlapping_images(const pm_model* m, int max_num_images, double min_triangulation_angle,
                                     const uint8_t* mask, int32_t* out_offsets, int32_t* out_images) {
  const int I = m->num_images;
  const int64_t P = m->num_points;
  const float min_rad = (float)(min_triangulation_angle * DEG2RAD);

  /* images_points (model.cc:710-717): the points each image observes,
   * once per observation, in point order */
  int64_t* ip_off = calloc((size_t)I + 1, sizeof(int64_t));
  for (int64_t t = 0; t < (P > 0 ? m->track_offsets[P] : 0); ++t) ip_off[m->track_images[t] + 1]++;
  for (int i = 0; i < I; ++i) ip_off[i + 1] += ip_off[i];
  int64_t* ip = malloc(sizeof(int64_t) * (size_t)(ip_off[I] + 1));
  int64_t* fill = malloc(sizeof(int64_t) * (size_t)I);
  memcpy(fill, ip_off, sizeof(int64_t) * (size_t)I);
  for (int64_t p = 0; p < P; ++p)
    for (int64_t t = m->track_offsets[p]; t < m->track_offsets[p + 1]; ++t) ip[fill[m->track_images[t]]++] = p;

  /* projection centres (model.cc:719-726) */
  double* C = malloc(sizeof(double) * 3 * (size_t)I);
  for (int i = 0; i < I; ++i) {
    float c[3];
    proj_center(m->R + 9 * (size_t)i, m->T + 3 * (size_t)i, c);
    for (int k = 0; k < 3; ++k) C[3 * (size_t)i + k] = c[k];
  }

  /* per-image scratch: neighbour counters, angle buckets */
  int64_t* cnt = calloc((size_t)I, sizeof(int64_t));
  int64_t* boff = calloc((size_t)I + 1, sizeof(int64_t));
  int* seen = malloc(sizeof(int) * (size_t)I);
  for (int i = 0; i < I; ++i) seen[i] = -1;

  /* mean nearest co-observing distance (model.cc:728-746 over
   * ComputeSharedPoints): every image, the subset ignored */
  float m_dist = 0.0f;
  int m_cnt = 0;
  for (int i = 0; i < I; ++i) {
    float min_dist = FLT_MAX;
    for (int64_t q = ip_off[i]; q < ip_off[i + 1]; ++q) {
      const int64_t p = ip[q];
      for (int64_t t = m->track_offsets[p]; t < m->track_offsets[p + 1]; ++t) {
        const int j = m->track_images[t];
        if (j == i || seen[j] == i) continue;
        seen[j] = i;
        const float dist = (float)centre_dist(C, i, j);
        if (dist <= 1e-5) continue;
        min_dist = dist < min_dist ? dist : min_dist;
      }
    }
    if (min_dist < FLT_MAX) {
      m_dist += min_dist;
      m_cnt++;
    }
  }
  m_dist /= m_cnt;

  int32_t total = 0;
  out_offsets[0] = 0;
  for (int i = 0; i < I; ++i) {
    int n = 0;
    int32_t* out = out_images + total;
    if (!mask || mask[i]) {
      if (ip_off[i + 1] == ip_off[i]) {
        n = spatial(m, C, i, m_dist, max_num_images, out);
      } else {
        /* SelectNeighborViews (model.cc:757-830) */
        memset(cnt, 0, sizeof(int64_t) * (size_t)I);
        for (int64_t q = ip_off[i]; q < ip_off[i + 1]; ++q) {
          const int64_t p = ip[q];
          for (int64_t t = m->track_offsets[p]; t < m->track_offsets[p + 1]; ++t) {
            const int j = m->track_images[t];
            if (mask && !mask[j]) continue;
            if (j != i) cnt[j]++;
          }
        }
        boff[0] = 0;
        for (int j = 0; j < I; ++j) boff[j + 1] = boff[j] + cnt[j];
        float* ang = malloc(sizeof(float) * (size_t)(boff[I] + 1));
        int64_t* pos = malloc(sizeof(int64_t) * (size_t)I);
        memcpy(pos, boff, sizeof(int64_t) * (size_t)I);
        for (int64_t q = ip_off[i]; q < ip_off[i + 1]; ++q) {
          const int64_t p = ip[q];
          const double X[3] = {m->xyz[3 * p], m->xyz[3 * p + 1], m->xyz[3 * p + 2]};
          for (int64_t t = m->track_offsets[p]; t < m->track_offsets[p + 1]; ++t) {
            const int j = m->track_images[t];
            if (mask && !mask[j]) continue;
            if (j != i) ang[pos[j]++] = (float)tri_angle(C + 3 * (size_t)i, C + 3 * (size_t)j, X);
          }
        }
        ranked_t* ord = malloc(sizeof(ranked_t) * (size_t)I);
        int no = 0;
        const int rig_i = m->is_rig ? m->is_rig[i] : 0;
        for (int j = 0; j < I; ++j) {
          const int64_t c = cnt[j];
          if (c == 0) continue;
          /* Percentile(angles, 75): element round(0.75 (c - 1)) in order */
          qsort(ang + boff[j], (size_t)c, sizeof(float), cmp_fkey);
          const int64_t idx = (int64_t)llround(0.75 * (double)(c - 1));
          const float pct = ang[boff[j] + idx];
          if (pct >= min_rad) {
            const double factor = ((m->is_rig ? m->is_rig[j] : 0) == rig_i) ? 1.5 : 1.0;
            ord[no].idx = j;
            ord[no].score = (long long)((double)c * factor);
            no++;
          }
        }
        qsort(ord, (size_t)no, sizeof(ranked_t), cmp_desc);
        n = no < max_num_images ? no : max_num_images;
        for (int k = 0; k < n; ++k) out[k] = ord[k].idx;
        free(ord);
        free(pos);
        free(ang);
      }
    }
    total += n;
    out_offsets[i + 1] = total;
  }
  free(seen);
  free(boff);
  free(cnt);
  free(C);
  free(fill);
  free(ip);
  free(ip_off);
  return 0;
}

int pmo_model_depth_ranges(const pm_model* m, float* out) {
  const int I = m->num_images;
  float lt[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, rb[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  /* std::min(lt, x) / std::max(rb, x): the first argument unless x beats it */
  for (int64_t p = 0; p < m->num_points; ++p)
    for (int c = 0; c < 3; ++c) {
      const float x = m->xyz[3 * p + c];
      lt[c] = (x < lt[c]) ? x : lt[c];
      rb[c] = (rb[c] < x) ? x : rb[c];
    }
  for (int i = 0; i < I; ++i) {
    float C[3];
    proj_center(m->R + 9 * (size_t)i, m->T + 3 * (size_t)i, C);
    for (int c = 0; c < 3; ++c) {
      lt[c] = (C[c] < lt[c]) ? C[c] : lt[c];
      rb[c] = (rb[c] < C[c]) ? C[c] : rb[c];
    }
  }
  for (int i = 0; i < I; ++i) {
    const float* r = m->R + 9 * (size_t)i;
    const float t = m->T[3 * (size_t)i + 2];
    float lo = 0.0f, hi = 0.0f;
    for (int k = 0; k < 8; ++k) {
      /* BoundingBox::GetCorners (util/bb.h:86-96) */
      const int bx = (k == 1 || k == 2 || k == 5 || k == 6);
      const int by = (k == 2 || k == 3 || k == 6 || k == 7);
      const int bz = (k >= 4);
      const float d = dot3f(r[6], r[7], r[8], bx ? rb[0] : lt[0], by ? rb[1] : lt[1], bz ? rb[2] : lt[2]) + t;
      if (k == 0) {
        lo = hi = d;
      } else {
        lo = (d < lo) ? d : lo;
        hi = (hi < d) ? d : hi;
      }
    }
    out[2 * (size_t)i] = (0.000001f < lo) ? lo : 0.000001f;
    out[2 * (size_t)i + 1] = hi;
  }
  return 0;
}
