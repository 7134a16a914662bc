/* pm_fusion_oracle.c -- CPU transcription of StereoFusion's depth-map fusion
 * (src/mvs/fusion.cc).
 *
 * TEST INFRASTRUCTURE ONLY: the checker the device fusion is compared with.
 * The product never links or calls it.
 *
 * It restates, single-threaded, in plain C with the canonical FP32
 * evaluation order of this build (left-to-right sums, no contraction):
 *   fusion.cc:697-727     FindNextImage
 *   fusion.cc:1150-1215   Run()'s image loop (start at the first non-rig image)
 *   fusion.cc:1730-1792   FuseImage with num_eff_threads = 1
 *   fusion.cc:1794-2182   Fuse: depth-first walk, consistency with the first
 *                         pixel, medians, colour, score
 *   fusion.cc:47-60       internal::Median (nth_element; the mean of the two
 *                         middle elements for even sizes)
 * Projection matrices follow ComposeProjectionMatrix /
 * ComposeInverseProjectionMatrix (utils.cc:35-52) computed in FP64 and
 * rounded to FP32, as the rest of this build packs poses.
 *
 * Pinned in tests/test_oracle_fusion.py by closed forms on designed scenes
 * (integer pixel shifts between cameras of a plane), see there. */
#include "pm_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int32_t img, row, col, depth;
} walk_entry;

typedef struct {
  void* p;
  size_t n, cap, elem;
} vec;

static int vec_push(vec* v, const void* x) {
  if (v->n == v->cap) {
    const size_t nc = v->cap ? 2 * v->cap : 64;
    void* q = realloc(v->p, nc * v->elem);
    if (!q) return -1;
    v->p = q;
    v->cap = nc;
  }
  memcpy((char*)v->p + v->n * v->elem, x, v->elem);
  v->n++;
  return 0;
}

struct pmo_fusion {
  vec xyz, normal, rgb, score, vis_off, vis_img;
};

static int cmp_f(const void* a, const void* b) {
  const float x = *(const float*)a, y = *(const float*)b;
  return (x > y) - (x < y);
}
static int cmp_d(const void* a, const void* b) {
  const double x = *(const double*)a, y = *(const double*)b;
  return (x > y) - (x < y);
}

/* internal::Median (fusion.cc:47-60) of n floats (sorted in place). */
static float median_f(float* v, size_t n) {
  qsort(v, n, sizeof(float), cmp_f);
  const size_t mid = n / 2;
  if (n % 2 == 0) return (v[mid] + v[mid - 1]) / 2.0f;
  return v[mid];
}
static float median_d(double* v, size_t n) {
  qsort(v, n, sizeof(double), cmp_d);
  const size_t mid = n / 2;
  if (n % 2 == 0) return ((float)v[mid] + (float)v[mid - 1]) / 2.0f;
  return (float)v[mid];
}

/* row of a 3x4 matrix times (x, y, z, 1) */
static inline float row4(const float* m, float x, float y, float z) { return ((m[0] * x + m[1] * y) + m[2] * z) + m[3]; }

/* FindNextImage (fusion.cc:697-727). */
static int find_next(const pm_fusion_image* im, int n, const uint8_t* fused, int prev) {
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

typedef struct {
  const pm_fusion_options* o;
  const pm_fusion_image* im;
  int n;
  float (*P)[12];
  float (*iP)[12];
  float (*Rt)[9];
  uint8_t** mask;
  const uint8_t* fused;
  float max_sq_reproj, min_cos_normal;
  vec stack;
  vec fx, fy, fz, fnx, fny, fnz, fr, fg, fb, fc, ferr;
  uint8_t* visible; /* per image, this walk */
  int* vis_list;
  int nvis;
} walk_ctx;

/* Fuse (fusion.cc:1794-2182) from one seed. */
static int fuse(walk_ctx* w, struct pmo_fusion* out, int img0, int row0, int col0) {
  const pm_fusion_options* o = w->o;
  const pm_fusion_image* im = w->im;
  w->stack.n = 0;
  w->fx.n = w->fy.n = w->fz.n = w->fnx.n = w->fny.n = w->fnz.n = 0;
  w->fr.n = w->fg.n = w->fb.n = w->fc.n = w->ferr.n = 0;
  for (int k = 0; k < w->nvis; ++k) w->visible[w->vis_list[k]] = 0;
  w->nvis = 0;
  float ref_pt[3] = {0, 0, 0}, ref_n[3] = {0, 0, 0};
  const walk_entry seed = {img0, row0, col0, 0};
  if (vec_push(&w->stack, &seed)) return -1;
  while (w->stack.n > 0) {
    const walk_entry e = ((walk_entry*)w->stack.p)[--w->stack.n];
    const pm_fusion_image* I = &im[e.img];
    const int W = I->width, H = I->height;
    const size_t p = (size_t)e.row * W + e.col;
    if (w->mask[e.img][p]) continue;
    const float depth = I->depth[p];
    if (depth <= 0.0f) continue;
    float depth_error = 0.0f;
    if (e.depth > 0) {
      const float* P = w->P[e.img];
      const float p0 = row4(P, ref_pt[0], ref_pt[1], ref_pt[2]);
      const float p1 = row4(P + 4, ref_pt[0], ref_pt[1], ref_pt[2]);
      const float p2 = row4(P + 8, ref_pt[0], ref_pt[1], ref_pt[2]);
      const float cd = p0 / p2 - (float)e.col;
      const float rd = p1 / p2 - (float)e.row;
      if (cd * cd + rd * rd > w->max_sq_reproj) continue;
      depth_error = fabsf(p2 - depth) / depth;
      if ((double)depth_error > o->max_depth_error) continue;
    }
    float nrm[3] = {0, 0, 0};
    if (o->with_normal) {
      const float n0 = I->normal[p], n1 = I->normal[(size_t)W * H + p], n2 = I->normal[2 * (size_t)W * H + p];
      const float* Rt = w->Rt[e.img];
      for (int i = 0; i < 3; ++i) nrm[i] = (Rt[3 * i] * n0 + Rt[3 * i + 1] * n1) + Rt[3 * i + 2] * n2;
      if (e.depth > 0) {
        const float c = (ref_n[0] * nrm[0] + ref_n[1] * nrm[1]) + ref_n[2] * nrm[2];
        if (c < w->min_cos_normal) continue;
      }
    }
    const float* iP = w->iP[e.img];
    const float u = (float)e.col * depth, v = (float)e.row * depth;
    const float xyz[3] = {row4(iP, u, v, depth), row4(iP + 4, u, v, depth), row4(iP + 8, u, v, depth)};
    /* channels 1: a grey image, its value as r, g and b */
    const int ch = I->channels == 1 ? 1 : 3;
    const uint8_t* px = I->rgb + (int64_t)ch * p;
    float grad = 0.0f;
    if (e.col > 1 && e.col < W - 1 && e.row > 1 && e.row < H - 1) {
      const int l = I->rgb[ch * (p - 1)], r = I->rgb[ch * (p + 1)];
      const int t = I->rgb[ch * (p - W)], b = I->rgb[ch * (p + W)];
      const float gx = (float)(l - r) / 255.0f, gy = (float)(t - b) / 255.0f;
      grad = sqrtf(gx * gx + gy * gy);
    }
    w->mask[e.img][p] = 1;
    const float col_r = px[0], col_g = px[ch == 3 ? 1 : 0], col_b = px[ch == 3 ? 2 : 0];
    if (vec_push(&w->fc, &grad) || vec_push(&w->fx, &xyz[0]) || vec_push(&w->fy, &xyz[1]) ||
        vec_push(&w->fz, &xyz[2]) || vec_push(&w->fr, &col_r) || vec_push(&w->fg, &col_g) ||
        vec_push(&w->fb, &col_b))
      return -1;
    if (o->with_normal && (vec_push(&w->fnx, &nrm[0]) || vec_push(&w->fny, &nrm[1]) || vec_push(&w->fnz, &nrm[2])))
      return -1;
    if (!w->visible[e.img]) {
      w->visible[e.img] = 1;
      w->vis_list[w->nvis++] = e.img;
    }
    if (e.depth > 0) {
      const double de = depth_error;
      if (vec_push(&w->ferr, &de)) return -1;
    } else {
      memcpy(ref_pt, xyz, sizeof(ref_pt));
      memcpy(ref_n, nrm, sizeof(ref_n));
    }
    if (w->fx.n >= (size_t)o->max_num_pixels) break;
    const int nd = e.depth + 1;
    if (nd >= o->max_traversal_depth) continue;
    for (int k = 0; k < I->num_overlapping; ++k) {
      const int j = I->overlapping[k];
      if (j < 0 || j >= w->n || !im[j].used || w->fused[j]) continue;
      const float* P = w->P[j];
      const float q0 = row4(P, xyz[0], xyz[1], xyz[2]);
      const float q1 = row4(P + 4, xyz[0], xyz[1], xyz[2]);
      const float q2 = row4(P + 8, xyz[0], xyz[1], xyz[2]);
      const float fc = roundf(q0 / q2), fr = roundf(q1 / q2);
      /* static_cast<int> of a value out of int range is taken as outside */
      if (!(fabsf(fc) < 1073741824.0f) || !(fabsf(fr) < 1073741824.0f)) continue;
      const int nc = (int)fc, nr = (int)fr;
      if (nc < 3 || nr < 3 || nc >= im[j].width - 3 || nr >= im[j].height - 3) continue;
      const walk_entry ne = {j, nr, nc, nd};
      if (vec_push(&w->stack, &ne)) return -1;
    }
  }
  const size_t npx = w->fx.n;
  if (npx < (size_t)o->min_num_pixels || w->nvis < o->min_num_visible_images) return 0;
  float pt[6] = {median_f((float*)w->fx.p, npx), median_f((float*)w->fy.p, npx), median_f((float*)w->fz.p, npx),
                 0.0f, 0.0f, 0.0f};
  for (int i = 0; i < 3; ++i)
    if (isnan(pt[i]) || isinf(pt[i])) return 0;
  if (o->with_normal) {
    const float nx = median_f((float*)w->fnx.p, npx), ny = median_f((float*)w->fny.p, npx),
                nz = median_f((float*)w->fnz.p, npx);
    const float norm = sqrtf(nx * nx + (ny * ny + nz * nz));
    if (norm < FLT_EPSILON) return 0;
    pt[3] = nx / norm;
    pt[4] = ny / norm;
    pt[5] = nz / norm;
  }
  uint8_t rgb[3];
  vec* chan[3] = {&w->fr, &w->fg, &w->fb};
  for (int c = 0; c < 3; ++c) {
    const float m = roundf(median_f((float*)chan[c]->p, npx));
    rgb[c] = (uint8_t)(m < 0.0f ? 0.0f : (m > 255.0f ? 255.0f : m));
  }
  float m_conf = 0.0f;
  for (size_t i = 0; i < w->fc.n; ++i) m_conf += ((float*)w->fc.p)[i];
  m_conf /= (float)w->fc.n;
  if (w->nvis >= 3) m_conf = 1.0f;
  m_conf = 1.0f - m_conf;
  if (w->ferr.n == 0) return 0; /* the reference's Median CHECKs a non-empty list */
  float score = (float)((double)median_d((double*)w->ferr.p, w->ferr.n) / o->max_depth_error);
  score = fminf(0.6f * score + 0.4f * m_conf, 1.0f);
  if (isnan(score) || isinf(score)) return 0;
  for (int i = 0; i < 3; ++i)
    if (vec_push(&out->xyz, &pt[i]) || vec_push(&out->normal, &pt[3 + i]) || vec_push(&out->rgb, &rgb[i])) return -1;
  if (vec_push(&out->score, &score)) return -1;
  /* the visibility set, ascending */
  for (int j = 0, m = 0; j < w->n && m < w->nvis; ++j)
    if (w->visible[j]) {
      const int32_t id = j;
      if (vec_push(&out->vis_img, &id)) return -1;
      ++m;
    }
  const int64_t off = (int64_t)out->vis_img.n;
  return vec_push(&out->vis_off, &off);
}

static void free_out(struct pmo_fusion* f) {
  if (!f) return;
  free(f->xyz.p);
  free(f->normal.p);
  free(f->rgb.p);
  free(f->score.p);
  free(f->vis_off.p);
  free(f->vis_img.p);
  free(f);
}

struct pmo_fusion* pmo_fusion_run(const pm_fusion_options* o, int n, const pm_fusion_image* im) {
  if (!o || !im || n < 1 || o->min_num_pixels < 2 || o->step_size < 1) return NULL;
  struct pmo_fusion* out = (struct pmo_fusion*)calloc(1, sizeof(*out));
  if (!out) return NULL;
  out->xyz.elem = out->normal.elem = out->score.elem = sizeof(float);
  out->rgb.elem = 1;
  out->vis_off.elem = sizeof(int64_t);
  out->vis_img.elem = sizeof(int32_t);
  const int64_t zero = 0;
  vec_push(&out->vis_off, &zero);
  walk_ctx w;
  memset(&w, 0, sizeof(w));
  w.o = o;
  w.im = im;
  w.n = n;
  w.P = calloc((size_t)n, sizeof(*w.P));
  w.iP = calloc((size_t)n, sizeof(*w.iP));
  w.Rt = calloc((size_t)n, sizeof(*w.Rt));
  w.mask = (uint8_t**)calloc((size_t)n, sizeof(uint8_t*));
  uint8_t* fused = (uint8_t*)calloc((size_t)n, 1);
  w.fused = fused;
  w.visible = (uint8_t*)calloc((size_t)n, 1);
  w.vis_list = (int*)calloc((size_t)n, sizeof(int));
  w.stack.elem = sizeof(walk_entry);
  w.fx.elem = w.fy.elem = w.fz.elem = w.fnx.elem = w.fny.elem = w.fnz.elem = sizeof(float);
  w.fr.elem = w.fg.elem = w.fb.elem = w.fc.elem = sizeof(float);
  w.ferr.elem = sizeof(double);
  /* max_squared_reproj_error_ and min_cos_normal_error_ (fusion.cc:657-659) */
  w.max_sq_reproj = (float)(o->max_reproj_error * o->max_reproj_error);
  w.min_cos_normal = (float)cos(o->max_normal_error * 0.0174532925199432954743716805978692718781530857086181640625);
  int status = (w.P && w.iP && w.Rt && w.mask && fused && w.visible && w.vis_list) ? 0 : -1;
  for (int i = 0; i < n && status == 0; ++i) {
    /* P and its inverse: the pose packing of a source seen from an identity reference */
    pm_image si;
    memset(&si, 0, sizeof(si));
    si.width = im[i].width;
    si.height = im[i].height;
    pm_view_desc vd;
    memset(&vd, 0, sizeof(vd));
    vd.num_src = 1;
    vd.ref_cam.K[0] = vd.ref_cam.K[2] = 1.0f;
    vd.ref_cam.R[0] = vd.ref_cam.R[4] = vd.ref_cam.R[8] = 1.0f;
    vd.src = &si;
    vd.src_cams = &im[i].cam;
    float pose[PM_POSE_STRIDE], k8[8];
    pmo_pack_poses(&vd, pose, k8);
    memcpy(w.P[i], pose + 16, sizeof(w.P[i]));
    memcpy(w.iP[i], pose + 28, sizeof(w.iP[i]));
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) w.Rt[i][3 * a + b] = im[i].cam.R[3 * b + a];
    if (im[i].used) {
      w.mask[i] = (uint8_t*)calloc((size_t)im[i].width * im[i].height, 1);
      if (!w.mask[i]) status = -1;
    }
  }
  if (status == 0) {
    int start = 0;
    for (int i = 0; i < n; ++i)
      if (!im[i].is_rig) {
        start = i;
        break;
      }
    for (int i = start; i >= 0 && status == 0; i = find_next(im, n, fused, i)) {
      if (!im[i].used) continue;
      const int W = im[i].width, H = im[i].height, s = o->step_size;
      for (int r = 3; r < H - 3 && status == 0; r += s)
        for (int c = 3; c < W - 3 && status == 0; c += s) {
          if (w.mask[i][(size_t)r * W + c]) continue;
          if (fuse(&w, out, i, r, c) < 0) status = -1;
        }
      fused[i] = 255;
    }
  }
  for (int i = 0; i < n; ++i) free(w.mask ? w.mask[i] : NULL);
  free(w.mask);
  free(w.P);
  free(w.iP);
  free(w.Rt);
  free(fused);
  free(w.visible);
  free(w.vis_list);
  vec* vs[] = {&w.stack, &w.fx, &w.fy, &w.fz, &w.fnx, &w.fny, &w.fnz, &w.fr, &w.fg, &w.fb, &w.fc, &w.ferr};
  for (size_t k = 0; k < sizeof(vs) / sizeof(vs[0]); ++k) free(vs[k]->p);
  if (status) {
    free_out(out);
    return NULL;
  }
  return out;
}

int pmo_fusion_counts(const struct pmo_fusion* f, int64_t* num_points, int64_t* num_vis) {
  if (!f) return -1;
  *num_points = (int64_t)f->score.n;
  *num_vis = (int64_t)f->vis_img.n;
  return 0;
}

int pmo_fusion_points(const struct pmo_fusion* f, float* xyz, float* normal, uint8_t* rgb, float* score,
                      int64_t* vis_offsets, int32_t* vis_images) {
  if (!f) return -1;
  const size_t n = f->score.n;
  if (xyz) memcpy(xyz, f->xyz.p, 3 * n * sizeof(float));
  if (normal) memcpy(normal, f->normal.p, 3 * n * sizeof(float));
  if (rgb) memcpy(rgb, f->rgb.p, 3 * n);
  if (score) memcpy(score, f->score.p, n * sizeof(float));
  if (vis_offsets) memcpy(vis_offsets, f->vis_off.p, (n + 1) * sizeof(int64_t));
  if (vis_images) memcpy(vis_images, f->vis_img.p, f->vis_img.n * sizeof(int32_t));
  return 0;
}

void pmo_fusion_free(struct pmo_fusion* f) { free_out(f); }
