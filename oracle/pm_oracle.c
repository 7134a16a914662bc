/* pm_oracle.c -- CPU transcription of SenseMap's ACMM PatchMatch stage.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the gfx950 engine and the CPU
 * baseline timed by bench.py.  Nothing in the product links or calls this.
 *
 * Canonical arithmetic: FP32, compiled with -ffp-contract=off, every fused
 * multiply-add written as an explicit fmaf, every division and sqrtf
 * correctly rounded, transcendental functions and random draws from
 * sensemap_amd/csrc/pm_math.h.  The HIP kernels follow the same order, so for
 * the same inputs the maps agree bit for bit.
 *
 * Where the reference's arithmetic cannot be reproduced (CUDA texture
 * filtering, fast-math intrinsics, cuRAND) this file documents the
 * restatement next to the code.
 */
#include "pm_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../sensemap_amd/csrc/pm_math.h"

#define K_NUM_SAMPS 8          /* kernel_functions.h:31 */
#define K_NUM_EACH_SAMP 10     /* kernel_functions.h:32 */
#define MAX_HYP (K_NUM_SAMPS + 3)
#define KMAXCOST 2.0f          /* kernel_functions.cu:1261 */
/* var < 1e-5 in [0,1] intensity units (kernel_functions.cu:1331), expressed
 * in the 0..255 units used here: 1e-5 * 255^2. */
#define VAR_THRESHOLD 0.65025f
/* Compensated mode: windows whose ref variance is below this (0..255 units,
 * sigma < 5 grey levels) accumulate shifted, Neumaier-compensated moments. */
#define LOWTEX_VAR 25.0f

/* neighbor_offs (kernel_functions.cu:23-25), {drow, dcol}. */
static const int kNeighborOffs[4][2] = {{0, -1}, {-1, 0}, {0, 1}, {1, 0}};
/* dirSamples (kernel_functions.cu:27-44), {drow, dcol} pairs. */
static const int kDirSamples[K_NUM_SAMPS][K_NUM_EACH_SAMP * 2] = {
    {-1, -2, 1, -2, -2, -3, 2, -3, -3, -4, 3, -4, -4, -5, 4, -5, -5, -6, 5, -6},
    {-1, 2, 1, 2, -2, 3, 2, 3, -3, 4, 3, 4, -4, 5, 4, 5, -5, 6, 5, 6},
    {-2, -1, -2, 1, -3, -2, -3, 2, -4, -3, -4, 3, -5, -4, -5, 4, -6, -5, -6, 5},
    {2, -1, 2, 1, 3, -2, 3, 2, 4, -3, 4, 3, 5, -4, 5, 4, 6, -5, 6, 5},
    {0, -3, 0, -5, 0, -7, 0, -9, 0, -11, 0, -13, 0, -15, 0, -17, 0, -19, 0, -21},
    {0, 3, 0, 5, 0, 7, 0, 9, 0, 11, 0, 13, 0, 15, 0, 17, 0, 19, 0, 21},
    {-3, 0, -5, 0, -7, 0, -9, 0, -11, 0, -13, 0, -15, 0, -17, 0, -19, 0, -21, 0},
    {3, 0, 5, 0, 7, 0, 9, 0, 11, 0, 13, 0, 15, 0, 17, 0, 19, 0, 21, 0}};
/* mweights (kernel_functions.cu:22). */
static const float kMWeights[5] = {0.f, 1.f, 2.01f, 3.03f, 4.06f};
/* Scharr-like kernels (kernel_functions.cu:46-50). */
static const float kFilterX[3][3] = {{-3, 0, 3}, {-10, 0, 10}, {-3, 0, 3}};
static const float kFilterY[3][3] = {{-3, -10, -3}, {0, 0, 0}, {3, 10, 3}};

typedef struct hyp4 { float n[3]; float d; } hyp4;

/* ------------------------------------------------------------------------ */
/* exported math wrappers */
float pmo_expf(float x) { return pm_expf(x); }
float pmo_sinf(float x) { return pm_sinf(x); }
float pmo_cosf(float x) { return pm_cosf(x); }
float pmo_atan2f(float y, float x) { return pm_atan2f(y, x); }
float pmo_acosf(float x) { return pm_acosf(x); }
void pmo_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  pm_philox4x32_10(ctr, key, out);
}
float pmo_uniform(uint64_t seed, uint32_t view, uint32_t pixel, uint32_t pass, uint32_t draw) {
  pm_rng g;
  pm_rng_init(&g, seed, view, pixel, pass);
  float u = 0.0f;
  for (uint32_t i = 0; i <= draw; ++i) u = pm_rng_uniform(&g);
  return u;
}

/* ------------------------------------------------------------------------ */
/* Pose packing: InitTransforms (patch_match_cuda.cu:493-580) with
 * ComputeRelativePose / ComposeProjectionMatrix /
 * ComposeInverseProjectionMatrix (utils.cc:14-52), computed in FP64 and
 * rounded once to FP32.  inv_P uses the closed form of the inverse of
 * [K R | K T; 0 0 0 1], i.e. [R^T K^-1 | -R^T T], instead of a general 4x4
 * inverse.  Layout per source: K(4) R_rel(9) T_rel(3) P(12) invP(12) W H pad pad. */
int pmo_pack_poses(const pm_view_desc* view, float* poses, float ref_k8[8]) {
  const pm_camera* rc = &view->ref_cam;
  const float fx = rc->K[0], cx = rc->K[1], fy = rc->K[2], cy = rc->K[3];
  ref_k8[0] = fx;
  ref_k8[1] = cx;
  ref_k8[2] = fy;
  ref_k8[3] = cy;
  ref_k8[4] = 1.0f / fx;
  ref_k8[5] = -cx / fx;
  ref_k8[6] = 1.0f / fy;
  ref_k8[7] = -cy / fy;
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
    memset(o, 0, sizeof(float) * PM_POSE_STRIDE);
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
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Bilateral weights.  BilateralWeightComputer (utils.h:72-92) evaluates
 * exp(-(dr^2+dc^2)/(2 ss^2) - (c1-c2)^2/(2 sc^2)) per sample.  Restated as the
 * product of a spatial table and a colour table (exp(a+b) = exp(a) exp(b)),
 * both evaluated in FP64 and rounded to FP32; the colour difference of two u8
 * intensities takes 256 values. */
static void build_weight_tables(const pm_options* opt, int* nwin, float** sp, float ct[256]) {
  const int R = opt->window_radius, S = opt->window_step;
  int n = 0;
  for (int d = -R; d <= R; d += S) ++n;
  *nwin = n;
  float* t = (float*)malloc(sizeof(float) * n * n);
  const double ss = (double)opt->sigma_spatial;
  const double sn = 1.0 / (2.0 * ss * ss);
  const double sc = (double)opt->sigma_color;
  const double cn = 1.0 / (2.0 * sc * sc);
  int k = 0;
  for (int dr = -R; dr <= R; dr += S) {
    for (int dc = -R; dc <= R; dc += S) {
      t[k++] = (float)exp(-(double)(dr * dr + dc * dc) * sn);
    }
  }
  for (int d = 0; d < 256; ++d) {
    const double x = (double)d / 255.0;
    ct[d] = (float)exp(-x * x * cn);
  }
  *sp = t;
}

static inline int ref_px(const pmo_view* v, int r, int c) {
  return (r < 0 || r >= v->H || c < 0 || c >= v->W) ? 0 : (int)v->ref[(size_t)r * v->W + c];
}

/* FilterKernel (gpu_mat_ref_image.cu:16-56) and ComputeGradientMap
 * (kernel_functions.cu:1655-1675) for one pixel.  Moments are in 0..255
 * units; the gradient is in [0,1] units as the reference's normalized-float
 * texture read, gated at 1.0 (kernel_functions.cu:2237). */
static void ref_moments_px(const pm_options* opt, const pmo_view* v, int row, int col,
                           float* mu_o, float* var_o, float* inv_sw_o, float* err_o, uint8_t* grad_o) {
  const int R = opt->window_radius, S = opt->window_step;
  const int c = ref_px(v, row, col);
  float sw = 0.0f, sr = 0.0f, srr = 0.0f;
  int k = 0;
  for (int dr = -R; dr <= R; dr += S) {
    for (int dc = -R; dc <= R; dc += S) {
      const int ri = ref_px(v, row + dr, col + dc);
      const float r = (float)ri;
      const float w = v->sp[k++] * v->ct[abs(c - ri)];
      const float wr = w * r;
      sw = sw + w;
      sr = sr + wr;
      srr = fmaf(wr, r, srr);
    }
  }
  const float mu = sr / sw;
  const float m2 = srr / sw;
  float var = m2 - mu * mu;
  float err = 0.0f;
  if (opt->accum_mode == PM_ACCUM_COMPENSATED) {
    /* second pass around the mean: var = E[(r-mu)^2] - E[r-mu]^2 */
    float s1 = 0.0f, s2 = 0.0f;
    k = 0;
    for (int dr = -R; dr <= R; dr += S) {
      for (int dc = -R; dc <= R; dc += S) {
        const int ri = ref_px(v, row + dr, col + dc);
        const float d = (float)ri - mu;
        const float w = v->sp[k++] * v->ct[abs(c - ri)];
        const float wd = w * d;
        s1 = s1 + wd;
        s2 = fmaf(wd, d, s2);
      }
    }
    err = s1 / sw;
    var = s2 / sw - err * err;
  }
  *mu_o = mu;
  *var_o = var;
  *inv_sw_o = 1.0f / sw;
  *err_o = err;
  float gx = 0.0f, gy = 0.0f;
  for (int dr = -1; dr <= 1; ++dr) {
    for (int dc = -1; dc <= 1; ++dc) {
      const float color = (float)ref_px(v, row + dr, col + dc) / 255.0f;
      gx = gx + color * kFilterX[dr + 1][dc + 1];
      gy = gy + color * kFilterY[dr + 1][dc + 1];
    }
  }
  const float grad = sqrtf(gx * gx + gy * gy);
  *grad_o = grad >= 1.0f ? 1 : 0;
}

/* ------------------------------------------------------------------------ */
/* Geometry */

/* ComposeHomography (kernel_functions.cu:431-503): H = K_s (R + T n^T / dist) K_r^-1,
 * dist = depth * (n . ray). */
static void compose_homography(const float* pose, const float iK[4], int row, int col,
                               const hyp4* h, float H[9]) {
  const float* K = pose;
  const float* R = pose + 4;
  const float* T = pose + 13;
  const float ray0 = iK[0] * (float)col + iK[1];
  const float ray1 = iK[2] * (float)row + iK[3];
  const float dist = h->d * (h->n[0] * ray0 + h->n[1] * ray1 + h->n[2]);
  const float inv_dist = 1.0f / dist;
  const float nd0 = inv_dist * h->n[0], nd1 = inv_dist * h->n[1], nd2 = inv_dist * h->n[2];
  float A[9];
  for (int i = 0; i < 3; ++i) {
    A[3 * i + 0] = R[3 * i + 0] + nd0 * T[i];
    A[3 * i + 1] = R[3 * i + 1] + nd1 * T[i];
    A[3 * i + 2] = R[3 * i + 2] + nd2 * T[i];
  }
  float B[9];
  for (int j = 0; j < 3; ++j) {
    B[j] = K[0] * A[j] + K[1] * A[6 + j];
    B[3 + j] = K[2] * A[3 + j] + K[3] * A[6 + j];
    B[6 + j] = A[6 + j];
  }
  for (int i = 0; i < 3; ++i) {
    H[3 * i + 0] = B[3 * i + 0] * iK[0];
    H[3 * i + 1] = B[3 * i + 1] * iK[2];
    H[3 * i + 2] = B[3 * i + 0] * iK[1] + B[3 * i + 1] * iK[3] + B[3 * i + 2];
  }
}

/* Bilinear sample with zero border; texel i sits at coordinate i (the
 * reference's "+0.5" is CUDA's texel-centre convention,
 * kernel_functions.cu:1296-1301).  Coordinates are clamped into
 * [-3, W+2] first, which leaves every in-range sample unchanged and maps any
 * farther one (and NaN) to all-zero texels.  The fraction is
 * min(u - floor(u), 1 - 2^-24): u - floor(u) rounds up to 1.0 for tiny
 * negative u, the canonical fraction stays below 1 (the gfx950 v_fract_f32
 * definition). */
static inline float fract_canon(float u) { return fminf(u - floorf(u), 0.99999994f); }

static inline float bilinear_u8(const uint8_t* img, int W, int H, float u, float v) {
  u = fminf(fmaxf(u, -3.0f), (float)W + 2.0f);
  v = fminf(fmaxf(v, -3.0f), (float)H + 2.0f);
  const float x0f = floorf(u), y0f = floorf(v);
  const float fx = fract_canon(u), fy = fract_canon(v);
  const int x0 = (int)x0f, y0 = (int)y0f;
  float t[4];
  for (int q = 0; q < 4; ++q) {
    const int xx = x0 + (q & 1), yy = y0 + (q >> 1);
    t[q] = (xx < 0 || xx >= W || yy < 0 || yy >= H) ? 0.0f : (float)img[(size_t)yy * W + xx];
  }
  const float top = fmaf(fx, t[1] - t[0], t[0]);
  const float bot = fmaf(fx, t[3] - t[2], t[2]);
  return fmaf(fy, bot - top, top);
}

static inline float bilinear_f32(const float* img, int W, int H, float u, float v) {
  u = fminf(fmaxf(u, -3.0f), (float)W + 2.0f);
  v = fminf(fmaxf(v, -3.0f), (float)H + 2.0f);
  const float x0f = floorf(u), y0f = floorf(v);
  const float fx = u - x0f, fy = v - y0f;
  const int x0 = (int)x0f, y0 = (int)y0f;
  float t[4];
  for (int q = 0; q < 4; ++q) {
    const int xx = x0 + (q & 1), yy = y0 + (q >> 1);
    t[q] = (xx < 0 || xx >= W || yy < 0 || yy >= H) ? 0.0f : img[(size_t)yy * W + xx];
  }
  const float top = fmaf(fx, t[1] - t[0], t[0]);
  const float bot = fmaf(fx, t[3] - t[2], t[2]);
  return fmaf(fy, bot - top, top);
}

static inline void neumaier_add(float* s, float* c, float x) {
  const float t = *s + x;
  *c = *c + ((fabsf(*s) >= fabsf(x)) ? ((*s - t) + x) : ((x - t) + *s));
  *s = t;
}

static double ncc_f64(const pm_options* opt, const pm_view_desc* view, const float* poses, const float k8[8],
                      int row, int col, const double hyp[4], int src);
static int g_cost_fp64 = 0;
void pmo_set_cost_fp64(int on) { g_cost_fp64 = on != 0; }

/* ComputeNCCCost<R,S> (kernel_functions.cu:1245-1340). */
static float ncc_cost(const pm_options* opt, const pmo_view* v, int row, int col,
                      const hyp4* h, int src, pmo_counters* cnt) {
  const size_t p = (size_t)row * v->W + col;
  const float var_r = v->var_r[p];
  if (cnt) __atomic_fetch_add(&cnt->ncc_evaluations, 1, __ATOMIC_RELAXED);
  if (g_cost_fp64) {
    const double hd[4] = {h->n[0], h->n[1], h->n[2], h->d};
    const float k8[8] = {v->refK[0],    v->refK[1],    v->refK[2],    v->refK[3],
                         v->refInvK[0], v->refInvK[1], v->refInvK[2], v->refInvK[3]};
    return (float)ncc_f64(opt, v->desc, &v->pose[0][0], k8, row, col, hd, src);
  }
  float H[9];
  compose_homography(v->pose[src], v->refInvK, row, col, h, H);
  const float fcol = (float)col, frow = (float)row;
  /* centre check: (int) truncation toward zero, so (-1, 0) passes
   * (kernel_functions.cu:1264-1270); bounds are the source's own size. */
  const float zc = H[6] * fcol + H[7] * frow + H[8];
  const float izc = 1.0f / zc;
  const float xf = (H[0] * fcol + H[1] * frow + H[2]) * izc;
  const float yf = (H[3] * fcol + H[4] * frow + H[5]) * izc;
  const int Ws = v->srcW[src], Hs = v->srcH[src];
  if (!(xf > -1.0f && xf < (float)Ws && yf > -1.0f && yf < (float)Hs)) return KMAXCOST;
  /* var_r < threshold returns kMaxCost whatever the window holds
   * (kernel_functions.cu:1331), so the window walk is skipped. */
  if (var_r < VAR_THRESHOLD) return KMAXCOST;
  const int R = opt->window_radius, S = opt->window_step;
  const int n = v->nwin;
  const float c0 = (float)(col - R), r0 = (float)(row - R);
  float bx = H[0] * c0 + H[1] * r0 + H[2];
  float by = H[3] * c0 + H[4] * r0 + H[5];
  float bz = H[6] * c0 + H[7] * r0 + H[8];
  const float fs = (float)S;
  const float hxc = fs * H[0], hyc = fs * H[3], hzc = fs * H[6];
  const float hxr = fs * H[1], hyr = fs * H[4], hzr = fs * H[7];
  const uint8_t* img = v->src[src];
  const int c = (int)v->ref[p];
  const float mu_r = v->mu_r[p];
  const int shifted = (opt->accum_mode == PM_ACCUM_COMPENSATED) && (var_r < LOWTEX_VAR);
  float s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
  float e1 = 0.0f, e2 = 0.0f, e3 = 0.0f;
  int k = 0;
  for (int i = 0; i < n; ++i) {
    float x = bx, y = by, z = bz;
    const int rr = row - R + i * S;
    for (int j = 0; j < n; ++j) {
      const int cc = col - R + j * S;
      const float rz = 1.0f / z;
      const float s = bilinear_u8(img, Ws, Hs, x * rz, y * rz);
      const int ri = ref_px(v, rr, cc);
      const float r = (float)ri;
      const float w = v->sp[k++] * v->ct[abs(c - ri)];
      if (!shifted) {
        const float ws = w * s;
        s1 = s1 + ws;
        s2 = fmaf(ws, s, s2);
        s3 = fmaf(ws, r, s3);
      } else {
        const float d = s - mu_r;
        const float rd = r - mu_r;
        const float wd = w * d;
        neumaier_add(&s1, &e1, wd);
        neumaier_add(&s2, &e2, wd * d);
        neumaier_add(&s3, &e3, wd * rd);
      }
      x = x + hxc;
      y = y + hyc;
      z = z + hzc;
    }
    bx = bx + hxr;
    by = by + hyr;
    bz = bz + hzr;
  }
  if (cnt) __atomic_fetch_add(&cnt->ncc_samples, (uint64_t)n * n, __ATOMIC_RELAXED);
  const float inv_sw = v->inv_sw[p];
  float var_s, cov;
  if (!shifted) {
    const float mu_s = s1 * inv_sw;
    const float m2_s = s2 * inv_sw;
    const float m_rs = s3 * inv_sw;
    var_s = m2_s - mu_s * mu_s;
    cov = m_rs - mu_r * mu_s;
  } else {
    const float m1 = (s1 + e1) * inv_sw;
    const float m2 = (s2 + e2) * inv_sw;
    const float m3 = (s3 + e3) * inv_sw;
    var_s = m2 - m1 * m1;
    cov = m3 - v->err_r[p] * m1;
  }
  if (var_s < VAR_THRESHOLD) return KMAXCOST;
  const float val = 1.0f - cov / sqrtf(var_r * var_s);
  return fmaxf(0.0f, fminf(KMAXCOST, val));
}

/* ComputeGeomConsistencyCostShared (kernel_functions.cu:505-568); K is the
 * ref {fx, cx, fy, cy}, or {1/fx, -cx/fx, 1/fy, -cy/fy} where the reference
 * passes ref_invK (kernel_functions.cu:2924, compat_lo_geo_bug). */
static float geom_cost(const pm_options* opt, const pmo_view* v, int row, int col, float depth,
                       const float np[3], int src, const float K[4], pmo_counters* cnt) {
  const float* P = v->pose[src] + 16;
  const float* iP = v->pose[src] + 28;
  if (cnt) __atomic_fetch_add(&cnt->geo_evaluations, 1, __ATOMIC_RELAXED);
  const float fx = depth * np[0], fy = depth * np[1], fz = depth * np[2];
  const float inv_fz = 1.0f / (P[8] * fx + P[9] * fy + P[10] * fz + P[11]);
  float sc = inv_fz * (P[0] * fx + P[1] * fy + P[2] * fz + P[3]);
  float sr = inv_fz * (P[4] * fx + P[5] * fy + P[6] * fz + P[7]);
  const float sd = bilinear_f32(v->src_depth[src], v->srcW[src], v->srcH[src], sc, sr);
  if (sd == 0.0f) return opt->geom_consistency_max_cost;
  sc = sc * sd;
  sr = sr * sd;
  const float bx = iP[0] * sc + iP[1] * sr + iP[2] * sd + iP[3];
  const float by = iP[4] * sc + iP[5] * sr + iP[6] * sd + iP[7];
  const float bz = iP[8] * sc + iP[9] * sr + iP[10] * sd + iP[11];
  const float ibz = 1.0f / bz;
  const float bcol = ibz * K[0] * bx + K[1];
  const float brow = ibz * K[2] * by + K[3];
  const float dc = (float)col - bcol;
  const float dr = (float)row - brow;
  return fminf(opt->geom_consistency_max_cost, sqrtf(dc * dc + dr * dr));
}

/* ComputeMultiViewCostVector (kernel_functions.cu:1447-1495), priors off. */
static float mv_cost(const pm_options* opt, const pmo_view* v, int row, int col, const hyp4* h,
                     const float np[3], int src, const float K[4], pmo_counters* cnt) {
  float cost = ncc_cost(opt, v, row, col, h, src, cnt);
  if (opt->geom_consistency) {
    cost = cost + opt->geom_consistency_regularizer * geom_cost(opt, v, row, col, h->d, np, src, K, cnt);
  }
  return cost;
}

/* ------------------------------------------------------------------------ */
/* Random hypotheses */

/* GenerateRandomNormal (kernel_functions.cu:112-136): Marsaglia (1972),
 * flipped to face the camera.  The rejection loop is capped at 64 trials
 * (probability of reaching it < 1e-43). */
static void random_normal(pm_rng* g, const float ray[3], float n[3]) {
  float v1 = 0.0f, v2 = 0.0f, s = 2.0f;
  for (int t = 0; t < 64 && s >= 1.0f; ++t) {
    v1 = 2.0f * pm_rng_uniform(g) - 1.0f;
    v2 = 2.0f * pm_rng_uniform(g) - 1.0f;
    s = v1 * v1 + v2 * v2;
  }
  if (s >= 1.0f) { v1 = 0.0f; v2 = 0.0f; s = 0.0f; }
  const float s_norm = 2.0f * sqrtf(1.0f - s);
  n[0] = v1 * s_norm;
  n[1] = v2 * s_norm;
  n[2] = 1.0f - 2.0f * s;
  if (n[0] * ray[0] + n[1] * ray[1] + n[2] * ray[2] > 0.0f) {
    n[0] = -n[0];
    n[1] = -n[1];
    n[2] = -n[2];
  }
}

/* GenerateRandomDepth (kernel_functions.cu:76-80). */
static inline float random_depth(pm_rng* g, float dmin, float dmax) {
  return pm_rng_uniform(g) * (dmax - dmin) + dmin;
}

/* PerturbNormal, 5-argument form (kernel_functions.cu:211-225) with
 * Normal2Dir / Dir2Normal (utils.h:94-106). */
static void perturb_normal(pm_rng* g, const float n[3], float a1, float a2, float out[3]) {
  const float y = pm_atan2f(sqrtf(n[0] * n[0] + n[1] * n[1]), n[2]);
  const float x = pm_atan2f(n[1], n[0]);
  const float nx = a1 * (2.0f * pm_rng_uniform(g) - 1.0f) + x;
  const float ny = a2 * (2.0f * pm_rng_uniform(g) - 1.0f) + y;
  const float siny = pm_sinf(ny);
  out[0] = pm_cosf(nx) * siny;
  out[1] = pm_sinf(nx) * siny;
  out[2] = pm_cosf(ny);
}

/* CorrectNormal (kernel_functions.cu:227-306), including the reference's
 * rotation matrix as written: entries 4,5 copy entry 3 and entries 7,8 copy
 * entry 6 of (1 - cos phi) Omega^2. */
static void correct_normal(float n[3], const float vd[3], float inv_view_len) {
  const float cosl = n[0] * vd[0] + n[1] * vd[1] + n[2] * vd[2];
  if (!(cosl >= 0.0f)) return;
  float c[3];
  c[0] = n[1] * vd[2] - n[2] * vd[1];
  c[1] = n[2] * vd[0] - n[0] * vd[2];
  c[2] = n[0] * vd[1] - n[1] * vd[0];
  const float a = cosl * inv_view_len;
  float phi;
  if (!(a <= 1.0f)) {
    phi = -0.001f; /* acosf(a > 1) is NaN and min(NaN, -0.001) is -0.001 */
  } else {
    phi = pm_acosf(a) - PM_HALF_PI_F;
    phi = fminf(phi * 1.01f, -0.001f);
  }
  const float inv_len = 1.0f / sqrtf(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  c[0] = c[0] * inv_len;
  c[1] = c[1] * inv_len;
  c[2] = c[2] * inv_len;
  const float o[9] = {0.0f, -c[2], c[1], c[2], 0.0f, -c[0], -c[1], c[0], 0.0f};
  const float sp = pm_sinf(phi);
  const float cp = 1.0f - pm_cosf(phi);
  float oo[9];
  oo[0] = cp * (o[0] * o[0] + o[1] * o[3] + o[2] * o[6]);
  oo[1] = cp * (o[0] * o[1] + o[1] * o[4] + o[2] * o[7]);
  oo[2] = cp * (o[0] * o[2] + o[1] * o[5] + o[2] * o[8]);
  oo[3] = cp * (o[3] * o[0] + o[4] * o[3] + o[5] * o[6]);
  oo[4] = oo[3];
  oo[5] = oo[3];
  oo[6] = cp * (o[6] * o[0] + o[7] * o[3] + o[8] * o[6]);
  oo[7] = oo[6];
  oo[8] = oo[6];
  float Rm[9];
  for (int i = 0; i < 9; ++i) {
    const float id = (i == 0 || i == 4 || i == 8) ? 1.0f : 0.0f;
    Rm[i] = id + o[i] * sp + oo[i];
  }
  const float r0 = Rm[0] * n[0] + Rm[1] * n[1] + Rm[2] * n[2];
  const float r1 = Rm[3] * n[0] + Rm[4] * n[1] + Rm[5] * n[2];
  const float r2 = Rm[6] * n[0] + Rm[7] * n[1] + Rm[8] * n[2];
  n[0] = r0;
  n[1] = r1;
  n[2] = r2;
}

/* PropagateDepth, ray form (kernel_functions.cu:409-429). */
static float propagate_depth(float d1, const float n1[3], const float iK[4], int row1, int col1,
                             const float np2[3]) {
  const float x1 = d1 * (iK[0] * (float)col1 + iK[1]);
  const float y1 = d1 * (iK[2] * (float)row1 + iK[3]);
  const float z1 = d1;
  const float d = n1[0] * x1 + n1[1] * y1 + n1[2] * z1;
  const float t = d / (n1[0] * np2[0] + n1[1] * np2[1] + n1[2] * np2[2]);
  if (fabsf(t) < 1e-5f) return d1;
  return t;
}

/* ------------------------------------------------------------------------ */
/* Per-pixel passes */

typedef struct pass_ctx {
  const pm_options* opt;
  const pmo_view* v;
  pmo_state* st;
  float th_mc;
  uint32_t pass;
  pmo_counters* cnt;
} pass_ctx;

static inline void load_normal(const pmo_state* st, size_t plane, size_t p, float n[3]) {
  n[0] = st->normal[p];
  n[1] = st->normal[plane + p];
  n[2] = st->normal[2 * plane + p];
}

static inline void store_normal(pmo_state* st, size_t plane, size_t p, const float n[3]) {
  st->normal[p] = n[0];
  st->normal[plane + p] = n[1];
  st->normal[2 * plane + p] = n[2];
}

/* InitConfMapKernel (kernel_functions.cu:1497-1601) for one hypothesis:
 * costs of all sources, ascending stable bubble sort, conf = sum of the thk
 * smallest / mweights[thk], view_sel bits of those views. */
static float init_conf_eval(const pass_ctx* x, int row, int col, const hyp4* h, float* costs,
                            uint32_t* sel) {
  const pmo_view* v = x->v;
  const int N = v->N;
  const float np[3] = {v->refInvK[0] * (float)col + v->refInvK[1],
                       v->refInvK[2] * (float)row + v->refInvK[3], 1.0f};
  float sorted[PM_MAX_NUM_SRC];
  int order[PM_MAX_NUM_SRC];
  for (int i = 0; i < N; ++i) {
    costs[i] = mv_cost(x->opt, v, row, col, h, np, i, v->refK, x->cnt);
    sorted[i] = costs[i];
    order[i] = i;
  }
  for (int a = 0; a < N - 1; ++a) {
    for (int b = 0; b < N - 1 - a; ++b) {
      if (sorted[b] > sorted[b + 1]) {
        const float t = sorted[b];
        sorted[b] = sorted[b + 1];
        sorted[b + 1] = t;
        const int o = order[b];
        order[b] = order[b + 1];
        order[b + 1] = o;
      }
    }
  }
  const int thk = x->opt->thk < N ? x->opt->thk : N;
  float conf = 0.0f;
  uint32_t bits = 0;
  for (int i = 0; i < thk; ++i) {
    conf = conf + sorted[i];
    bits |= 1u << order[i];
  }
  *sel = bits;
  return conf / kMWeights[thk];
}

/* InitWorkspace (patch_match_cuda.cu:582-698): random depth
 * (FillWithRandomNumbers, gpu_mat.h:114-130) and normal (InitNormalMap,
 * kernel_functions.cu:689-701), or the given maps repaired by InitDepthMap
 * (kernel_functions.cu:703-737); then InitConfMapKernel.  With
 * init_samples = K > 1 the best of K random draws (lowest conf) is kept. */
static void init_pixel(const pass_ctx* x, int row, int col) {
  const pmo_view* v = x->v;
  pmo_state* st = x->st;
  const size_t plane = (size_t)v->W * v->H;
  const size_t p = (size_t)row * v->W + col;
  const float ray[3] = {v->refInvK[0] * (float)col + v->refInvK[1],
                        v->refInvK[2] * (float)row + v->refInvK[3], 1.0f};
  pm_rng g;
  pm_rng_init(&g, x->opt->seed, v->view_id, (uint32_t)p, x->pass);
  float costs[PM_MAX_NUM_SRC], best_costs[PM_MAX_NUM_SRC];
  hyp4 h, best;
  uint32_t sel = 0, best_sel = 0;
  float best_conf = 0.0f;
  if (x->opt->init_depth_random) {
    const int K = x->opt->init_samples > 0 ? x->opt->init_samples : 1;
    for (int s = 0; s < K; ++s) {
      h.d = random_depth(&g, v->depth_min, v->depth_max);
      random_normal(&g, ray, h.n);
      const float conf = init_conf_eval(x, row, col, &h, costs, &sel);
      if (s == 0 || conf < best_conf) {
        best_conf = conf;
        best = h;
        best_sel = sel;
        memcpy(best_costs, costs, sizeof(float) * v->N);
      }
    }
  } else {
    h.d = st->depth[p];
    load_normal(st, plane, p, h.n);
    if (h.d <= 0.0f) {
      h.d = random_depth(&g, v->depth_min, v->depth_max);
      random_normal(&g, ray, h.n);
    }
    if (h.n[0] * ray[0] + h.n[1] * ray[1] + h.n[2] * ray[2] >= 0.0f) {
      random_normal(&g, ray, h.n);
    }
    best_conf = init_conf_eval(x, row, col, &h, best_costs, &best_sel);
    best = h;
  }
  st->depth[p] = best.d;
  store_normal(st, plane, p, best.n);
  st->conf[p] = best_conf;
  st->view_sel[p] = best_sel;
  for (int i = 0; i < v->N; ++i) st->cost[(size_t)i * plane + p] = best_costs[i];
}

/* SweepFromCheckerBoardOpt (kernel_functions.cu:2136-2560), priors off and
 * propagate_depth off. */
/* Iteration iter's good-matching threshold (patch_match_cuda.cu:778-779,
 * 853): init_ncc_matching_cost * exp(-iter^2 / alpha). */
float pmo_th_mc(const pm_options* opt, int iter) {
  return (float)((double)opt->init_ncc_matching_cost * exp((double)(-(float)(iter * iter) / opt->alpha)));
}

/* LocalOptimization's choice among its 5 trials (kernel_functions.cu:
 * 2860-2900): trial t's cost is the mean of M[t][i] over the selected
 * sources (sel); trials are taken in order when their mean is below the
 * running conf, so the least mean wins and the first of equal means stays.
 * live masks the trials evaluated in full.  Returns the trial (-1: none)
 * and its mean in *conf_out (conf otherwise). */
int pmo_lo_decide(const float* M, uint32_t sel, int N, uint32_t live, float conf, float* conf_out) {
  int best = -1;
  for (int t = 0; t < 5; ++t) {
    if (!((live >> t) & 1u)) continue;
    float mconf = 0.0f, mw = 0.0f;
    for (int i = 0; i < N; ++i) {
      if (!((sel >> i) & 1u)) continue;
      mconf = mconf + M[t * PM_MAX_NUM_SRC + i];
      mw = mw + 1.0f;
    }
    mconf = mconf / mw;
    if (mconf < conf) {
      conf = mconf;
      best = t;
    }
  }
  *conf_out = conf;
  return best;
}

/* The refinement hypotheses of a sweep pixel (kernel_functions.cu:
 * 2449-2486), drawn in order from the pixel's stream g: ref[0] the winner
 * (n, d); ref[1] (PerturbNormal(n), d); ref[2] (n, U[(1 - rho) d,
 * (1 + rho) d]); ref[3] (random normal facing the camera, d); ref[4] (n,
 * U[depth_min, depth_max]).  Perturbed and random normals are
 * CorrectNormal-ed.  depth_max here is the pass's 1.2 x depth_max. */
static void refine_hyps(const pm_options* opt, pm_rng* g, float depth, const float normal[3], const float vd[3],
                        float inv_view_len, float depth_min, float depth_max, hyp4 ref[5]) {
  const float angle1 = fminf(opt->random_angle1_range, PM_PI_F);
  const float angle2 = fminf(opt->random_angle2_range, PM_HALF_PI_F);
  float nd;
  {
    const float dlo = (1.0f - opt->random_depth_ratio) * depth;
    const float dhi = (1.0f + opt->random_depth_ratio) * depth;
    nd = random_depth(g, dlo, dhi);
  }
  float nn[3];
  perturb_normal(g, normal, angle1, angle2, nn);
  correct_normal(nn, vd, inv_view_len);
  const float rd = random_depth(g, depth_min, depth_max);
  float rn[3];
  random_normal(g, vd, rn);
  correct_normal(rn, vd, inv_view_len);
  ref[0].d = depth;
  memcpy(ref[0].n, normal, 3 * sizeof(float));
  ref[1].d = depth;
  memcpy(ref[1].n, nn, sizeof(nn));
  ref[2].d = nd;
  memcpy(ref[2].n, normal, 3 * sizeof(float));
  ref[3].d = depth;
  memcpy(ref[3].n, rn, sizeof(rn));
  ref[4].d = rd;
  memcpy(ref[4].n, normal, 3 * sizeof(float));
}

/* LocalOptimization's trial hypotheses (kernel_functions.cu:2736-2860): the
 * depth range [low, up] and the normalised mean normal over the pixel and
 * its in-image 4-neighbours, then 5 trials (perturbed mean normal by
 * +-4 / +-2 degrees, uniform depth in [low, up]) drawn in order from g. */
static void lo_trials(const float* depth, const float* normal, int W, int Hh, int row, int col, pm_rng* g,
                      hyp4 out[5]) {
  const size_t plane = (size_t)W * Hh;
  const size_t p = (size_t)row * W + col;
  /* DEG2RAD(4.0) and DEG2RAD(2.0) are double products (kernel_functions.h:39) */
  const float angle1 = (float)(4.0 * 0.0174532925199432);
  const float angle2 = (float)(2.0 * 0.0174532925199432);
  float low = depth[p], up = depth[p];
  float m0 = normal[p], m1 = normal[plane + p], m2 = normal[2 * plane + p];
  for (int i = 0; i < 4; ++i) {
    const int u = col + kNeighborOffs[i][1], w = row + kNeighborOffs[i][0];
    if (u < 0 || u >= W || w < 0 || w >= Hh) continue;
    const size_t q = (size_t)w * W + u;
    const float ndp = depth[q];
    m0 = m0 + normal[q];
    m1 = m1 + normal[plane + q];
    m2 = m2 + normal[2 * plane + q];
    if (ndp <= low) low = ndp;
    if (ndp > up) up = ndp;
  }
  float mn[3] = {m0, m1, m2};
  const float inv_norm = 1.0f / sqrtf(mn[0] * mn[0] + mn[1] * mn[1] + mn[2] * mn[2]);
  mn[0] = mn[0] * inv_norm;
  mn[1] = mn[1] * inv_norm;
  mn[2] = mn[2] * inv_norm;
  for (int t = 0; t < 5; ++t) {
    out[t].d = random_depth(g, low, up);
    perturb_normal(g, mn, angle1, angle2, out[t].n);
  }
}

/* The refinement hypotheses of sweep pixel (row, col) after its num_samples
 * view-selection draws: out 5 rows {nx, ny, nz, d}. */
int pmo_refine_hypotheses(const pm_options* opt, uint32_t view_id, uint32_t pass, int W, int row, int col,
                          const float* iK, float depth, const float* normal, float depth_min, float depth_max,
                          float* out) {
  pm_rng g;
  pm_rng_init(&g, opt->seed, view_id, (uint32_t)(row * W + col), pass);
  for (int smp = 0; smp < opt->num_samples; ++smp) (void)pm_rng_uniform(&g);
  const float vd[3] = {iK[0] * (float)col + iK[1], iK[2] * (float)row + iK[3], 1.0f};
  const float inv_view_len = 1.0f / sqrtf(vd[0] * vd[0] + vd[1] * vd[1] + vd[2] * vd[2]);
  refine_hyps(opt, &g, depth, normal, vd, inv_view_len, depth_min, (float)(1.2 * (double)depth_max),
              (hyp4*)out);
  return 0;
}

/* LocalOptimization's 5 trials of pixel (row, col): out 5 rows {nx, ny, nz, d}. */
int pmo_lo_trials(int W, int Hh, const float* depth, const float* normal, int row, int col, uint64_t seed,
                  uint32_t view_id, uint32_t pass, float* out) {
  pm_rng g;
  pm_rng_init(&g, seed, view_id, (uint32_t)(row * W + col), pass);
  lo_trials(depth, normal, W, Hh, row, col, &g, (hyp4*)out);
  return 0;
}

/* Candidate hypotheses of one sweep pixel (kernel_functions.cu:2232-2327):
 * the pixel's own (n, d); when grad_ok, up to two 4-neighbours by ascending
 * conf (bubble sort); then the least-conf sample of each of the 8
 * directional groups (the first minimum wins).  Each neighbour's normal is
 * CorrectNormal-ed and its plane propagated to the pixel (PropagateDepth).
 * conf, depth: W*H row-major; normal: 3 planes of W*H.  out: up to
 * (1 + 2 + 8) rows {nx, ny, nz, d}.  Returns the number of rows. */
int pmo_gather_candidates(int W, int Hh, const float* conf, const float* depth, const float* normal, int grad_ok,
                          const float* iK, int row, int col, float* out) {
  const size_t plane = (size_t)W * Hh;
  const size_t p = (size_t)row * W + col;
  const float vd[3] = {iK[0] * (float)col + iK[1], iK[2] * (float)row + iK[3], 1.0f};
  const float inv_view_len = 1.0f / sqrtf(vd[0] * vd[0] + vd[1] * vd[1] + vd[2] * vd[2]);
  const float np[3] = {vd[0], vd[1], 1.0f};
  hyp4* hyps = (hyp4*)out;
  int nh = 0;
  hyps[nh].d = depth[p];
  hyps[nh].n[0] = normal[p];
  hyps[nh].n[1] = normal[plane + p];
  hyps[nh].n[2] = normal[2 * plane + p];
  nh++;

  /* up to two 4-neighbours by ascending conf, gated by the gradient */
  if (grad_ok) {
    float nconf[4];
    int order[4];
    int nvalid = 0;
    for (int i = 0; i < 4; ++i) {
      order[i] = i;
      const int u = col + kNeighborOffs[i][1], w = row + kNeighborOffs[i][0];
      if (u < 0 || u >= W || w < 0 || w >= Hh) {
        nconf[i] = FLT_MAX;
        continue;
      }
      nvalid++;
      nconf[i] = conf[(size_t)w * W + u];
    }
    for (int a = 0; a < 3; ++a) {
      for (int b = 0; b < 3 - a; ++b) {
        if (nconf[b] > nconf[b + 1]) {
          const float t = nconf[b];
          nconf[b] = nconf[b + 1];
          nconf[b + 1] = t;
          const int o = order[b];
          order[b] = order[b + 1];
          order[b + 1] = o;
        }
      }
    }
    if (nvalid > 2) nvalid = 2;
    for (int i = 0; i < nvalid; ++i) {
      const int idx = order[i];
      const int u = col + kNeighborOffs[idx][1], w = row + kNeighborOffs[idx][0];
      const size_t q = (size_t)w * W + u;
      float nn[3];
      nn[0] = normal[q];
      nn[1] = normal[plane + q];
      nn[2] = normal[2 * plane + q];
      correct_normal(nn, vd, inv_view_len);
      hyps[nh].d = propagate_depth(depth[q], nn, iK, w, u, np);
      hyps[nh].n[0] = nn[0];
      hyps[nh].n[1] = nn[1];
      hyps[nh].n[2] = nn[2];
      nh++;
    }
  }
  /* best-conf sample of each of the 8 directional groups */
  for (int gi = 0; gi < K_NUM_SAMPS; ++gi) {
    float best_conf = FLT_MAX;
    int br = 0, bc = 0;
    for (int f = 0; f < K_NUM_EACH_SAMP; ++f) {
      const int nr = row + kDirSamples[gi][2 * f];
      const int ncn = col + kDirSamples[gi][2 * f + 1];
      if (nr < 0 || nr >= Hh || ncn < 0 || ncn >= W) continue;
      const float c = conf[(size_t)nr * W + ncn];
      if (best_conf > c) {
        best_conf = c;
        br = nr;
        bc = ncn;
      }
    }
    if (best_conf < FLT_MAX) {
      const size_t q = (size_t)br * W + bc;
      float nn[3];
      nn[0] = normal[q];
      nn[1] = normal[plane + q];
      nn[2] = normal[2 * plane + q];
      correct_normal(nn, vd, inv_view_len);
      hyps[nh].d = propagate_depth(depth[q], nn, iK, br, bc, np);
      hyps[nh].n[0] = nn[0];
      hyps[nh].n[1] = nn[1];
      hyps[nh].n[2] = nn[2];
      nh++;
    }
  }

  return nh;
}

/* View selection (kernel_functions.cu:2364-2426), TransformPDFToCDF
 * (:672-686) and FindMinCost (:649-659) for one pixel.  M holds nh rows of
 * PM_MAX_NUM_SRC costs (row k = hypothesis k, column i = source i; row 0 the
 * current hypothesis); draws holds num_samples uniforms in (0, 1], the
 * pixel's first draws.  A source is flagged when at least num_good_hypothesis
 * of its costs are below th_mc and at most num_bad_hypothesis above
 * max_ncc_matching_cost; its probability is sum_k exp(-c^2 / (2 beta^2)),
 * zeroed when unflagged and normalised to a CDF.  Each draw u - FLT_EPSILON
 * picks the first source whose CDF exceeds it; picked sources get weight 1
 * (random_optimization) or a count, normalised.  Returns the hypothesis of
 * least weighted cost ('<=': the last minimum wins), with the weights and
 * the view_sel mask, or -1 when no source is flagged or the weights vanish
 * (the pixel keeps its state). */
int pmo_select_views(const pm_options* opt, float th_mc, int nh, int N, const float* M, const float* draws,
                     float* wts, uint32_t* sel_out) {
  const float inv_beta2 = 0.5f / (opt->beta * opt->beta);
  float prob[PM_MAX_NUM_SRC];
  uint32_t flags = 0;
  for (int i = 0; i < N; ++i) {
    float pr = 0.0f;
    int n1 = 0, n2 = 0;
    for (int k = 0; k < nh; ++k) {
      const float c = M[k * PM_MAX_NUM_SRC + i];
      pr = pr + pm_expf(-(c * c) * inv_beta2);
      n1 += (c < th_mc);
      n2 += (c > opt->max_ncc_matching_cost);
    }
    prob[i] = pr;
    if (n1 >= opt->num_good_hypothesis && n2 <= opt->num_bad_hypothesis) flags |= 1u << i;
  }
  if (flags == 0) return -1;
  for (int i = 0; i < N; ++i) {
    if (!((flags >> i) & 1u)) prob[i] = 0.0f;
  }
  float psum = 0.0f;
  for (int i = 0; i < N; ++i) psum = psum + prob[i];
  const float inv_psum = 1.0f / psum;
  float cum = 0.0f;
  for (int i = 0; i < N; ++i) {
    cum = cum + prob[i] * inv_psum;
    prob[i] = cum;
  }
  for (int i = 0; i < N; ++i) wts[i] = 0.0f;
  for (int smp = 0; smp < opt->num_samples; ++smp) {
    const float rp = draws[smp] - PM_FLT_EPSILON;
    for (int i = 0; i < N; ++i) {
      if (prob[i] > rp) {
        if (opt->random_optimization) wts[i] = 1.0f;
        else wts[i] = wts[i] + 1.0f;
        break;
      }
    }
  }
  float wnorm = 0.0f;
  for (int i = 0; i < N; ++i) wnorm = wnorm + wts[i];
  if (wnorm <= 1e-5f) return -1;
  const float inv_wnorm = 1.0f / wnorm;
  uint32_t sel = 0;
  for (int i = 0; i < N; ++i) {
    wts[i] = wts[i] * inv_wnorm;
    if (wts[i] > 0.0f) sel |= 1u << i;
  }
  *sel_out = sel;
  int best = -1;
  float minc = FLT_MAX;
  for (int k = 0; k < nh; ++k) {
    float c = 0.0f;
    for (int i = 0; i < N; ++i) c = c + wts[i] * M[k * PM_MAX_NUM_SRC + i];
    if (c <= minc) {
      minc = c;
      best = k;
    }
  }
  return best < 0 ? 0 : best;
}

static void sweep_pixel(const pass_ctx* x, int row, int col) {
  const pm_options* opt = x->opt;
  const pmo_view* v = x->v;
  pmo_state* st = x->st;
  const int W = v->W, Hh = v->H, N = v->N;
  const size_t plane = (size_t)W * Hh;
  const size_t p = (size_t)row * W + col;
  const float* iK = v->refInvK;

  float depth = st->depth[p];
  float normal[3];
  load_normal(st, plane, p, normal);

  const float depth_min = v->depth_min;
  const float depth_max = (float)(1.2 * (double)v->depth_max);
  const float vd[3] = {iK[0] * (float)col + iK[1], iK[2] * (float)row + iK[3], 1.0f};
  const float inv_view_len = 1.0f / sqrtf(vd[0] * vd[0] + vd[1] * vd[1] + vd[2] * vd[2]);
  const float np[3] = {vd[0], vd[1], 1.0f};

  hyp4 hyps[MAX_HYP];
  const int nh = pmo_gather_candidates(W, Hh, st->conf, st->depth, st->normal, v->grad_ok[p], iK, row, col,
                                       &hyps[0].n[0]);

  /* cost matrix: row 0 is the stored cost of the current hypothesis */
  float M[MAX_HYP][PM_MAX_NUM_SRC];
  for (int i = 0; i < N; ++i) M[0][i] = st->cost[(size_t)i * plane + p];
  for (int k = 1; k < nh; ++k) {
    for (int i = 0; i < N; ++i) M[k][i] = mv_cost(opt, v, row, col, &hyps[k], np, i, v->refK, x->cnt);
  }

  /* view selection and weighted argmin (pmo_select_views) */
  pm_rng g;
  pm_rng_init(&g, opt->seed, v->view_id, (uint32_t)p, x->pass);
  float draws[opt->num_samples > 0 ? opt->num_samples : 1];
  for (int smp = 0; smp < opt->num_samples; ++smp) draws[smp] = pm_rng_uniform(&g);
  float wts[PM_MAX_NUM_SRC];
  uint32_t sel = 0;
  int best = pmo_select_views(opt, x->th_mc, nh, N, &M[0][0], draws, wts, &sel);
  if (best == -1) return;
  st->view_sel[p] = sel;
  if (best < 0) best = 0;
  depth = hyps[best].d;
  normal[0] = hyps[best].n[0];
  normal[1] = hyps[best].n[1];
  normal[2] = hyps[best].n[2];

  /* refinement (kernel_functions.cu:2449-2534) */
  hyp4 ref[5];
  refine_hyps(opt, &g, depth, normal, vd, inv_view_len, depth_min, depth_max, ref);

  float RM[5][PM_MAX_NUM_SRC];
  float rc[5];
  rc[0] = 0.0f;
  for (int i = 0; i < N; ++i) {
    RM[0][i] = M[best][i];
    rc[0] = rc[0] + wts[i] * RM[0][i];
  }
  for (int k = 1; k < 5; ++k) {
    for (int i = 0; i < N; ++i) RM[k][i] = mv_cost(opt, v, row, col, &ref[k], np, i, v->refK, x->cnt);
  }
  for (int k = 1; k < 5; ++k) {
    float c = 0.0f;
    for (int i = 0; i < N; ++i) c = c + wts[i] * RM[k][i];
    rc[k] = c;
  }
  int b2 = -1;
  float minc = FLT_MAX;
  for (int k = 0; k < 5; ++k) {
    if (rc[k] <= minc) {
      minc = rc[k];
      b2 = k;
    }
  }
  if (b2 < 0) return;
  st->conf[p] = rc[b2];
  st->depth[p] = ref[b2].d;
  store_normal(st, plane, p, ref[b2].n);
  for (int i = 0; i < N; ++i) st->cost[(size_t)i * plane + p] = RM[b2][i];
}

/* LocalOptimization (kernel_functions.cu:2736-2931). */
static void local_opt_pixel(const pass_ctx* x, int row, int col) {
  const pm_options* opt = x->opt;
  const pmo_view* v = x->v;
  pmo_state* st = x->st;
  const int W = v->W, Hh = v->H, N = v->N;
  const size_t plane = (size_t)W * Hh;
  const size_t p = (size_t)row * W + col;
  const float* iK = v->refInvK;

  float conf = st->conf[p];
  const uint32_t sel = st->view_sel[p];
  const float np[3] = {iK[0] * (float)col + iK[1], iK[2] * (float)row + iK[3], 1.0f};

  if (sel == 0) return; /* m_weight <= 0 on the first trial */
  pm_rng g;
  pm_rng_init(&g, opt->seed, v->view_id, (uint32_t)p, x->pass);
  hyp4 trials[5];
  lo_trials(st->depth, st->normal, W, Hh, row, col, &g, trials);
  float M[5][PM_MAX_NUM_SRC];
  int best = -1;
  hyp4 bh;
  /* A trial whose running sum already gives a mean >= the pixel's initial
   * conf cannot be taken (costs are >= 0; float sums of non-negative terms
   * never decrease), so its remaining evaluations are skipped.  Not in the
   * reference; the maps are unchanged, and the device skips exactly these
   * evaluations, so the evaluation counts agree. */
  const float mw_sel = (float)__builtin_popcount(sel);
  const float conf0 = conf;
  const int prune = !opt->geom_consistency || opt->geom_consistency_regularizer >= 0.0f;
  uint32_t live = 0;
  for (int t = 0; t < 5; ++t) {
    const hyp4 h = trials[t];
    float mconf = 0.0f;
    int ruled_out = 0;
    for (int i = 0; i < N; ++i) {
      if (!((sel >> i) & 1u)) continue;
      const float c = mv_cost(opt, v, row, col, &h, np, i, v->refK, x->cnt);
      M[t][i] = c;
      mconf = mconf + c;
      if (prune && mconf / mw_sel >= conf0) {
        ruled_out = 1;
        break;
      }
    }
    if (!ruled_out) live |= 1u << t;
  }
  best = pmo_lo_decide(&M[0][0], sel, N, live, conf, &conf);
  if (best < 0) return;
  bh = trials[best];
  st->conf[p] = conf;
  st->depth[p] = bh.d;
  store_normal(st, plane, p, bh.n);
  const float* Kgeo = opt->compat_lo_geo_bug ? v->refInvK : v->refK;
  for (int i = 0; i < N; ++i) {
    float c;
    if ((sel >> i) & 1u) {
      c = M[best][i];
    } else {
      c = mv_cost(opt, v, row, col, &bh, np, i, Kgeo, x->cnt);
    }
    st->cost[(size_t)i * plane + p] = c;
  }
}

/* FilterDepthMap (kernel_functions.cu:3282-3354) for one pixel. */
static void filter_pixel(const pass_ctx* x, int row, int col) {
  const pm_options* opt = x->opt;
  const pmo_view* v = x->v;
  pmo_state* st = x->st;
  const int W = v->W, Hh = v->H;
  const size_t plane = (size_t)W * Hh;
  const size_t p = (size_t)row * W + col;
  const float* iK = v->refInvK;
  const float depth = st->depth[p];
  if (depth <= 0.0f) return;
  float n[3];
  load_normal(st, plane, p, n);
  const float conf = st->conf[p];
  const float pt[3] = {depth * (iK[0] * (float)col + iK[1]), depth * (iK[2] * (float)row + iK[3]), depth};
  int good = 0;
  for (int i = 0; i < 4; ++i) {
    const int u = col + kNeighborOffs[i][1], w = row + kNeighborOffs[i][0];
    if (u < 0 || u >= W || w < 0 || w >= Hh) continue;
    const size_t q = (size_t)w * W + u;
    const float nd = st->depth[q];
    float nn[3];
    load_normal(st, plane, q, nn);
    const float nc = st->conf[q];
    const float npt[3] = {nd * (iK[0] * (float)u + iK[1]), nd * (iK[2] * (float)w + iK[3]), nd};
    const float e0 = pt[0] - npt[0], e1 = pt[1] - npt[1], e2 = pt[2] - npt[2];
    const float d1 = fabsf(e0 * n[0] + e1 * n[1] + e2 * n[2]);
    const float d2 = fabsf(e0 * nn[0] + e1 * nn[1] + e2 * nn[2]);
    if ((conf < opt->conf_threshold && d1 / depth < opt->depth_diff_threshold) ||
        (nc < opt->conf_threshold && d2 / nd < opt->depth_diff_threshold)) {
      good++;
    }
  }
  if (good < 2) {
    st->depth[p] = 0.0f;
    st->normal[p] = 0.0f;
    st->normal[plane + p] = 0.0f;
    st->normal[2 * plane + p] = 0.0f;
    st->conf[p] = 0.0f;
  }
}

typedef void (*pixel_fn)(const pass_ctx*, int, int);

/* Apply fn to the pixels of one colour ((row + col) % 2 == color) in rows
 * [r0, r1).  Pixels of one colour read only the other colour's state
 * (every propagation offset has odd Manhattan parity), so the order does
 * not matter. */
static void run_color(const pass_ctx* x, pixel_fn fn, int color, int r0, int r1) {
  const int W = x->v->W;
  const long half = (W + 1) / 2;
  const long total = (long)(r1 - r0) * half;
  long idx;
#pragma omp parallel for schedule(dynamic, 64)
  for (idx = 0; idx < total; ++idx) {
    const int row = r0 + (int)(idx / half);
    const int col = 2 * (int)(idx % half) + ((row + color) & 1);
    if (col < W) fn(x, row, col);
  }
}

static void run_all(const pass_ctx* x, pixel_fn fn, int r0, int r1) {
  const int W = x->v->W;
  const long total = (long)(r1 - r0) * W;
  long idx;
#pragma omp parallel for schedule(dynamic, 64)
  for (idx = 0; idx < total; ++idx) fn(x, r0 + (int)(idx / W), (int)(idx % W));
}

/* ------------------------------------------------------------------------ */
/* view preparation */

static int prepare_view(const pm_options* opt, const pm_view_desc* d, pmo_view* v, uint8_t** ref_copy,
                        uint8_t** src_copy) {
  memset(v, 0, sizeof(*v));
  v->W = d->ref.width;
  v->H = d->ref.height;
  v->N = d->num_src;
  v->view_id = d->view_id;
  v->desc = d;
  v->depth_min = d->depth_min;
  v->depth_max = d->depth_max;
  if (v->N < 1 || v->N > PM_MAX_NUM_SRC || v->W <= 0 || v->H <= 0) return -1;
  *ref_copy = (uint8_t*)malloc((size_t)v->W * v->H);
  for (int r = 0; r < v->H; ++r) memcpy(*ref_copy + (size_t)r * v->W, d->ref.data + (size_t)r * d->ref.stride, v->W);
  v->ref = *ref_copy;
  for (int i = 0; i < v->N; ++i) {
    const pm_image* s = &d->src[i];
    src_copy[i] = (uint8_t*)malloc((size_t)s->width * s->height);
    for (int r = 0; r < s->height; ++r)
      memcpy(src_copy[i] + (size_t)r * s->width, s->data + (size_t)r * s->stride, s->width);
    v->src[i] = src_copy[i];
    v->srcW[i] = s->width;
    v->srcH[i] = s->height;
    v->src_depth[i] = d->src_depth_maps ? d->src_depth_maps[i] : NULL;
  }
  float k8[8];
  float* poses = (float*)malloc(sizeof(float) * PM_POSE_STRIDE * v->N);
  pmo_pack_poses(d, poses, k8);
  for (int i = 0; i < v->N; ++i) memcpy(v->pose[i], poses + (size_t)i * PM_POSE_STRIDE, sizeof(float) * PM_POSE_STRIDE);
  free(poses);
  memcpy(v->refK, k8, sizeof(float) * 4);
  memcpy(v->refInvK, k8 + 4, sizeof(float) * 4);
  build_weight_tables(opt, &v->nwin, &v->sp, v->ct);
  const size_t plane = (size_t)v->W * v->H;
  v->mu_r = (float*)malloc(sizeof(float) * plane);
  v->var_r = (float*)malloc(sizeof(float) * plane);
  v->inv_sw = (float*)malloc(sizeof(float) * plane);
  v->err_r = (float*)malloc(sizeof(float) * plane);
  v->grad_ok = (uint8_t*)malloc(plane);
  long r;
#pragma omp parallel for schedule(dynamic, 4)
  for (r = 0; r < v->H; ++r) {
    for (int c = 0; c < v->W; ++c) {
      const size_t p = (size_t)r * v->W + c;
      ref_moments_px(opt, v, (int)r, c, &v->mu_r[p], &v->var_r[p], &v->inv_sw[p], &v->err_r[p], &v->grad_ok[p]);
    }
  }
  return 0;
}

static void free_view(pmo_view* v, uint8_t* ref_copy, uint8_t** src_copy) {
  free(ref_copy);
  for (int i = 0; i < v->N; ++i) free(src_copy[i]);
  free(v->sp);
  free(v->mu_r);
  free(v->var_r);
  free(v->inv_sw);
  free(v->err_r);
  free(v->grad_ok);
}

int pmo_ref_moments(const pm_options* opt, const uint8_t* img, int W, int H, float* mu, float* var,
                    float* inv_sw, uint8_t* grad_ok) {
  pmo_view v;
  memset(&v, 0, sizeof(v));
  v.W = W;
  v.H = H;
  v.ref = img;
  build_weight_tables(opt, &v.nwin, &v.sp, v.ct);
  for (int r = 0; r < H; ++r) {
    for (int c = 0; c < W; ++c) {
      const size_t p = (size_t)r * W + c;
      float err;
      ref_moments_px(opt, &v, r, c, &mu[p], &var[p], &inv_sw[p], &err, &grad_ok[p]);
    }
  }
  free(v.sp);
  return 0;
}

float pmo_ncc_cost(const pm_options* opt, const pm_view_desc* view, int row, int col, const float hyp[4],
                   int src) {
  pmo_view v;
  uint8_t* rc;
  uint8_t* sc[PM_MAX_NUM_SRC];
  if (prepare_view(opt, view, &v, &rc, sc) != 0) return NAN;
  hyp4 h;
  h.n[0] = hyp[0];
  h.n[1] = hyp[1];
  h.n[2] = hyp[2];
  h.d = hyp[3];
  const float c = ncc_cost(opt, &v, row, col, &h, src, NULL);
  free_view(&v, rc, sc);
  return c;
}

float pmo_geom_cost(const pm_options* opt, const pm_view_desc* view, int row, int col, float depth, int src) {
  pmo_view v;
  uint8_t* rc;
  uint8_t* sc[PM_MAX_NUM_SRC];
  if (prepare_view(opt, view, &v, &rc, sc) != 0) return NAN;
  const float np[3] = {v.refInvK[0] * (float)col + v.refInvK[1], v.refInvK[2] * (float)row + v.refInvK[3], 1.0f};
  const float c = geom_cost(opt, &v, row, col, depth, np, src, v.refK, NULL);
  free_view(&v, rc, sc);
  return c;
}

/* FP64 restatement of ComputeNCCCost: the same definition (bilateral
 * weights, zero-border bilinear source, centre check, variance cut) with
 * every quantity in double and the homography applied per sample without
 * incremental accumulation.  The precision reference for the FP32 paths. */
static double ncc_f64(const pm_options* opt, const pm_view_desc* view, const float* poses, const float k8[8],
                      int row, int col, const double hyp[4], int src) {
  const float* pose = poses + (size_t)src * PM_POSE_STRIDE;
  const double K[4] = {pose[0], pose[1], pose[2], pose[3]};
  double Rr[9], Tt[3];
  for (int i = 0; i < 9; ++i) Rr[i] = pose[4 + i];
  for (int i = 0; i < 3; ++i) Tt[i] = pose[13 + i];
  const double iK[4] = {k8[4], k8[5], k8[6], k8[7]};
  const double n0 = hyp[0], n1 = hyp[1], n2 = hyp[2], dep = hyp[3];
  const double dist = dep * (n0 * (iK[0] * col + iK[1]) + n1 * (iK[2] * row + iK[3]) + n2);
  double A[9];
  for (int i = 0; i < 3; ++i) {
    A[3 * i + 0] = Rr[3 * i + 0] + n0 / dist * Tt[i];
    A[3 * i + 1] = Rr[3 * i + 1] + n1 / dist * Tt[i];
    A[3 * i + 2] = Rr[3 * i + 2] + n2 / dist * Tt[i];
  }
  double B[9], H[9];
  for (int j = 0; j < 3; ++j) {
    B[j] = K[0] * A[j] + K[1] * A[6 + j];
    B[3 + j] = K[2] * A[3 + j] + K[3] * A[6 + j];
    B[6 + j] = A[6 + j];
  }
  for (int i = 0; i < 3; ++i) {
    H[3 * i + 0] = B[3 * i + 0] * iK[0];
    H[3 * i + 1] = B[3 * i + 1] * iK[2];
    H[3 * i + 2] = B[3 * i + 0] * iK[1] + B[3 * i + 1] * iK[3] + B[3 * i + 2];
  }
  const pm_image* si = &view->src[src];
  const pm_image* ri = &view->ref;
  const int Ws = si->width, Hs = si->height, W = ri->width, Hh = ri->height;
  const double zc = H[6] * col + H[7] * row + H[8];
  const double xc = (H[0] * col + H[1] * row + H[2]) / zc;
  const double yc = (H[3] * col + H[4] * row + H[5]) / zc;
  if (!(xc > -1.0 && xc < Ws && yc > -1.0 && yc < Hs)) return 2.0;
  const int R = opt->window_radius, S = opt->window_step;
  const double ss = opt->sigma_spatial, sc = opt->sigma_color;
  const double sn = 1.0 / (2.0 * ss * ss), cn = 1.0 / (2.0 * sc * sc);
  const int c = ri->data[(size_t)row * ri->stride + col];
  double sw = 0, sr = 0, srr = 0, ss1 = 0, ss2 = 0, srs = 0;
  for (int dr = -R; dr <= R; dr += S) {
    for (int dc = -R; dc <= R; dc += S) {
      const int rr = row + dr, cc = col + dc;
      const int rv = (rr < 0 || rr >= Hh || cc < 0 || cc >= W) ? 0 : ri->data[(size_t)rr * ri->stride + cc];
      const double cd = (double)(c - rv) / 255.0;
      const double w = exp(-(double)(dr * dr + dc * dc) * sn) * exp(-cd * cd * cn);
      const double z = H[6] * cc + H[7] * rr + H[8];
      const double u = (H[0] * cc + H[1] * rr + H[2]) / z;
      const double v = (H[3] * cc + H[4] * rr + H[5]) / z;
      double s = 0.0;
      if (isfinite(u) && isfinite(v) && u > -2.0 && v > -2.0 && u < Ws + 1.0 && v < Hs + 1.0) {
        const double x0 = floor(u), y0 = floor(v);
        const double fx = u - x0, fy = v - y0;
        double t[4];
        for (int q = 0; q < 4; ++q) {
          const int xx = (int)x0 + (q & 1), yy = (int)y0 + (q >> 1);
          t[q] = (xx < 0 || xx >= Ws || yy < 0 || yy >= Hs) ? 0.0 : si->data[(size_t)yy * si->stride + xx];
        }
        s = (1 - fy) * ((1 - fx) * t[0] + fx * t[1]) + fy * ((1 - fx) * t[2] + fx * t[3]);
      }
      sw += w;
      sr += w * rv;
      srr += w * rv * rv;
      ss1 += w * s;
      ss2 += w * s * s;
      srs += w * rv * s;
    }
  }
  const double mr = sr / sw, ms = ss1 / sw;
  const double vr = srr / sw - mr * mr, vs = ss2 / sw - ms * ms;
  const double thr = 1e-5 * 255.0 * 255.0;
  if (vr < thr || vs < thr) return 2.0;
  const double cov = srs / sw - mr * ms;
  double val = 1.0 - cov / sqrt(vr * vs);
  if (val < 0) val = 0;
  if (val > 2) val = 2;
  return val;
}

double pmo_ncc_cost_f64(const pm_options* opt, const pm_view_desc* view, int row, int col, const double hyp[4],
                        int src) {
  float k8[8];
  float* poses = (float*)malloc(sizeof(float) * PM_POSE_STRIDE * view->num_src);
  pmo_pack_poses(view, poses, k8);
  const double c = ncc_f64(opt, view, poses, k8, row, col, hyp, src);
  free(poses);
  return c;
}

/* Per-pixel, per-source NCC costs of the hypothesis maps (depth W*H,
 * normal 3 planar W*H): the FP32 canonical cost under opt->accum_mode into
 * out32 and its FP64 restatement into out64 (both N planar W*H; NaN where
 * depth <= 0).  The precision study's reference. */
int pmo_cost_maps(const pm_options* opt, const pm_view_desc* view, const float* depth, const float* normal,
                  int num_threads, float* out32, double* out64) {
#ifdef _OPENMP
  if (num_threads > 0) omp_set_num_threads(num_threads);
#else
  (void)num_threads;
#endif
  pmo_view v;
  uint8_t* rc;
  uint8_t* sc[PM_MAX_NUM_SRC];
  if (prepare_view(opt, view, &v, &rc, sc) != 0) return -1;
  float k8[8];
  float* poses = (float*)malloc(sizeof(float) * PM_POSE_STRIDE * view->num_src);
  pmo_pack_poses(view, poses, k8);
  const int W = view->ref.width, H = view->ref.height, N = view->num_src;
  const long plane = (long)W * H;
#pragma omp parallel for schedule(dynamic, 64)
  for (long p = 0; p < plane; ++p) {
    const int row = (int)(p / W), col = (int)(p % W);
    const float d = depth[p];
    for (int i = 0; i < N; ++i) {
      if (!(d > 0.0f)) {
        out32[i * plane + p] = NAN;
        out64[i * plane + p] = NAN;
        continue;
      }
      hyp4 h;
      h.n[0] = normal[p];
      h.n[1] = normal[plane + p];
      h.n[2] = normal[2 * plane + p];
      h.d = d;
      const double hd[4] = {h.n[0], h.n[1], h.n[2], h.d};
      out32[i * plane + p] = ncc_cost(opt, &v, row, col, &h, i, NULL);
      out64[i * plane + p] = ncc_f64(opt, view, poses, k8, row, col, hd, i);
    }
  }
  free(poses);
  free_view(&v, rc, sc);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* RunFilter (patch_match_cuda.cu:1074-1108): the given depth map
 * (view->init_depth, InitWorkspaceFilter :701-712) through
 * FilterDepthMapsKernel (kernel_functions.cu:3356-3450) or, with dedup,
 * DeduplicDepthMapsKernel (:3452-3490). */
int pmo_run_filter(const pm_options* opt, const pm_view_desc* view, int dedup, float* depth_out, int num_threads) {
#ifdef _OPENMP
  if (num_threads > 0) omp_set_num_threads(num_threads);
#else
  (void)num_threads;
#endif
  if (!opt->geom_consistency || !view->init_depth || !view->src_depth_maps) return -1;
  pmo_view v;
  uint8_t* rc;
  uint8_t* sc[PM_MAX_NUM_SRC];
  if (prepare_view(opt, view, &v, &rc, sc) != 0) return -1;
  const int W = v.W, N = v.N;
  const long plane = (long)W * v.H;
  const int minc = opt->filter_min_num_consistent < N ? opt->filter_min_num_consistent : N;
  memcpy(depth_out, view->init_depth, sizeof(float) * (size_t)plane);
#pragma omp parallel for schedule(dynamic, 256)
  for (long p = 0; p < plane; ++p) {
    const int row = (int)(p / W), col = (int)(p % W);
    const float d = depth_out[p];
    if ((double)d < 1e-6) {
      depth_out[p] = 0.0f;
      continue;
    }
    const float np[3] = {v.refInvK[0] * (float)col + v.refInvK[1], v.refInvK[2] * (float)row + v.refInvK[3], 1.0f};
    int good = 0, explained = 0;
    for (int i = 0; i < N; ++i) {
      if (geom_cost(opt, &v, row, col, d, np, i, v.refK, NULL) <= opt->filter_geom_consistency_max_cost) {
        explained = 1;
        if (dedup) break;
        good++;
      }
    }
    if (dedup ? explained : good < minc) depth_out[p] = 0.0f;
  }
  free_view(&v, rc, sc);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* CrossFilterKernel (kernel_functions.cu:3561-3778) for one pixel of one
 * view, and CrossFilterCuda::Run (patch_match_cuda.cu:1493-1551) over the
 * views.  Sampling of the layered depth/normal textures at (x + 0.5, y + 0.5)
 * with point filtering and a zero border (patch_match_cuda.cu:1186-1196) is
 * texel (floor(x + 0.5), floor(y + 0.5)) inside the image, 0 outside. */
static void cross_decode(uint32_t code, float n[3]) {
  /* float q1 = iq1 / decode_factor - 1 (kernel_functions.cu:3597) */
  const float q1 = (float)(code >> 16) / 32767.0f - 1.0f;
  const float q2 = (float)(code & 0xffffu) / 32767.0f - 1.0f;
  const float s = q1 * q1 + q2 * q2;
  n[0] = (float)(1.0 - 2.0 * (double)s);
  const float r = sqrtf(1.0f - s);
  n[1] = (2.0f * q1) * r;
  n[2] = (2.0f * q2) * r;
}

static float cross_cos_deg(double deg) { return (float)cos((double)(float)(deg * 0.0174532925199432)); }

static float cross_pixel(const pm_cross_options* o, const pm_cross_image* images, const pm_cross_view* cv,
                         const float* poses, const float k8[8], int min_consistent, int row, int col) {
  const pm_cross_image* ref = &images[cv->ref];
  const int W = ref->width, H = ref->height;
  const float depth = ref->depth[(size_t)row * W + col];
  if ((double)depth < 1e-6 || col < 3 || col > W - 3 || row < 3 || row > H - 3) return 0.0f;
  float nr[3];
  cross_decode(ref->normal_codes[(size_t)row * W + col], nr);
  const float* iK = k8 + 4;
  const float fcol = (float)col, frow = (float)row;
  float vr[3] = {iK[0] * fcol + iK[1], iK[2] * frow + iK[3], 1.0f};
  const float nv = sqrtf(vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
  vr[0] = vr[0] / nv;
  vr[1] = vr[1] / nv;
  vr[2] = vr[2] / nv;
  if (nr[0] * vr[0] + nr[1] * vr[1] + nr[2] * vr[2] >= cross_cos_deg(100.0)) return 0.0f;
  const float fp[3] = {depth * (iK[0] * fcol + iK[1]), depth * (iK[2] * frow + iK[3]), depth};
  const float sqr_thr = o->depth_diff_threshold * o->depth_diff_threshold;
  const float max_cos = cross_cos_deg((double)o->max_normal_error);
  float depths[32];
  depths[0] = depth;
  int ng = 0;
  for (int i = 0; i < cv->num_src; ++i) {
    const float* pose = poses + (size_t)i * PM_POSE_STRIDE;
    const float* P = pose + 16;
    const float* iP = pose + 28;
    const float* R = pose + 4;
    const float inv_fz = 1.0f / (P[8] * fp[0] + P[9] * fp[1] + P[10] * fp[2] + P[11]);
    float sc = inv_fz * (P[0] * fp[0] + P[1] * fp[1] + P[2] * fp[2] + P[3]);
    float sr = inv_fz * (P[4] * fp[0] + P[5] * fp[1] + P[6] * fp[2] + P[7]);
    const pm_cross_image* si = &images[cv->src[i]];
    const float xf = sc + 0.5f, yf = sr + 0.5f;
    float sd = 0.0f;
    uint32_t scode = 0u;
    if (xf >= 0.0f && xf < (float)si->width && yf >= 0.0f && yf < (float)si->height) {
      const size_t q = (size_t)(int)floorf(yf) * si->width + (size_t)(int)floorf(xf);
      sd = si->depth[q];
      scode = si->normal_codes[q];
    }
    if ((double)sd < 1e-4) continue;
    sc = sc * sd;
    sr = sr * sd;
    const float bx = iP[0] * sc + iP[1] * sr + iP[2] * sd + iP[3];
    const float by = iP[4] * sc + iP[5] * sr + iP[6] * sd + iP[7];
    const float bz = iP[8] * sc + iP[9] * sr + iP[10] * sd + iP[11];
    const float inv_bz = 1.0f / bz;
    const float bcol = inv_bz * (k8[0] * bx + k8[1] * bz);
    const float brow = inv_bz * (k8[2] * by + k8[3] * bz);
    const float dc = fcol - bcol, dr = frow - brow;
    const float reproj = sqrtf(dc * dc + dr * dr);
    if (reproj > o->filter_geom_consistency_max_cost || (depth - bz) * (depth - bz) > depth * depth * sqr_thr)
      continue;
    float sn[3];
    cross_decode(scode, sn);
    const float r0 = R[0] * sn[0] + R[3] * sn[1] + R[6] * sn[2];
    const float r1 = R[1] * sn[0] + R[4] * sn[1] + R[7] * sn[2];
    const float r2 = R[2] * sn[0] + R[5] * sn[1] + R[8] * sn[2];
    if (nr[0] * r0 + nr[1] * r1 + nr[2] * r2 < max_cos) continue;
    ng++;
    depths[ng] = bz;
  }
  if (ng < min_consistent) return 0.0f;
  /* BubbleSort(depths, num_good_src_depth) (utils.h:206-220), then the
   * element at int(num_good_src_depth * 0.5) (kernel_functions.cu:3767-3771) */
  for (int a = 0; a < ng - 1; ++a)
    for (int b = 0; b < ng - 1 - a; ++b)
      if (depths[b] > depths[b + 1]) {
        const float t = depths[b];
        depths[b] = depths[b + 1];
        depths[b + 1] = t;
      }
  return depths[(int)(ng * 0.5)];
}

int pmo_cross_filter(const pm_cross_options* o, int num_images, const pm_cross_image* images, int num_views,
                     const pm_cross_view* views, int num_threads) {
#ifdef _OPENMP
  if (num_threads > 0) omp_set_num_threads(num_threads);
#else
  (void)num_threads;
#endif
  for (int k = 0; k < num_views; ++k) {
    const pm_cross_view* cv = &views[k];
    if (cv->ref < 0 || cv->ref >= num_images || cv->num_src < 1 || cv->num_src > 31) return -1;
    const pm_cross_image* ref = &images[cv->ref];
    pm_image srcs[31];
    pm_camera cams[31];
    memset(srcs, 0, sizeof(srcs));
    for (int s = 0; s < cv->num_src; ++s) {
      srcs[s].width = images[cv->src[s]].width;
      srcs[s].height = images[cv->src[s]].height;
      cams[s] = images[cv->src[s]].cam;
    }
    pm_view_desc vd;
    memset(&vd, 0, sizeof(vd));
    vd.num_src = cv->num_src;
    vd.ref_cam = ref->cam;
    vd.src = srcs;
    vd.src_cams = cams;
    float poses[31 * PM_POSE_STRIDE];
    float k8[8];
    pmo_pack_poses(&vd, poses, k8);
    const int minc = o->filter_min_num_consistent < cv->num_src ? o->filter_min_num_consistent : cv->num_src;
    const long plane = (long)ref->width * ref->height;
#pragma omp parallel for schedule(dynamic, 256)
    for (long p = 0; p < plane; ++p) {
      const int row = (int)(p / ref->width), col = (int)(p % ref->width);
      cv->refined[p] = cross_pixel(o, images, cv, poses, k8, minc, row, col);
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* CrossFilterCuda::DedupRun (patch_match_cuda.cu:1561-1612) */

/* DeduplicFilterKernel (kernel_functions.cu:3780-3916) at one pixel of view
 * cv: updates the ref's visibility word and, when kept, writes *delt. */
static void dedup_pixel(const pm_cross_options* o, const pm_dedup_image* images, const pm_dedup_view* cv,
                        uint32_t* const* vis, const float* poses, const float k8[8], int min_consistent, int row,
                        int col, float* delt) {
  const pm_dedup_image* ref = &images[cv->ref];
  const int W = ref->width, H = ref->height;
  const size_t p = (size_t)row * W + col;
  uint32_t* rv = vis[cv->ref];
  const float depth = ref->depth[p];
  if ((double)depth < 1e-6 || col < 3 || col > W - 3 || row < 3 || row > H - 3) {
    rv[p] = 0u;
    return;
  }
  /* visbility_value >> (sizeof(uint32_t) - 1) shifts by 3 (:3803); every
   * word the views write is 0 or has bit 31 set, so this equals a test of
   * bit 31 on every reachable state */
  if (!(rv[p] >> 3)) {
    rv[p] = 0u;
    return;
  }
  const float* iK = k8 + 4;
  const float fcol = (float)col, frow = (float)row;
  /* ComputePointAtDepth (kernel_functions.cu:308-315) */
  const float fp[3] = {depth * (iK[0] * fcol + iK[1]), depth * (iK[2] * frow + iK[3]), depth};
  const float sqr_thr = o->depth_diff_threshold * o->depth_diff_threshold;
  float deltas[32];
  int good_src[32];
  int ng = 0, nvis = 0;
  for (int i = 0; i < cv->num_src; ++i) {
    const float* pose = poses + (size_t)i * PM_POSE_STRIDE;
    const float* P = pose + 16;
    const float* iP = pose + 28;
    const float inv_fz = 1.0f / (P[8] * fp[0] + P[9] * fp[1] + P[10] * fp[2] + P[11]);
    float sc = inv_fz * (P[0] * fp[0] + P[1] * fp[1] + P[2] * fp[2] + P[3]);
    float sr = inv_fz * (P[4] * fp[0] + P[5] * fp[1] + P[6] * fp[2] + P[7]);
    const pm_dedup_image* si = &images[cv->src[i]];
    /* point-sampled layered texture with a zero border (InitDepthMaps,
     * :1139-1207): texel (floor(x + 0.5), floor(y + 0.5)) */
    const float xf = sc + 0.5f, yf = sr + 0.5f;
    if (!(xf >= 0.0f && xf < (float)si->width && yf >= 0.0f && yf < (float)si->height)) continue;
    const size_t q = (size_t)(int)floorf(yf) * si->width + (size_t)(int)floorf(xf);
    const float sd = si->depth[q];
    if (sd == 0.0f) continue;
    sc = sc * sd;
    sr = sr * sd;
    const float bx = iP[0] * sc + iP[1] * sr + iP[2] * sd + iP[3];
    const float by = iP[4] * sc + iP[5] * sr + iP[6] * sd + iP[7];
    const float bz = iP[8] * sc + iP[9] * sr + iP[10] * sd + iP[11];
    const float inv_bz = 1.0f / bz;
    const float bcol = inv_bz * (k8[0] * bx + k8[1] * bz);
    const float brow = inv_bz * (k8[2] * by + k8[3] * bz);
    const float dc = fcol - bcol, dr = frow - brow;
    const float reproj = sqrtf(dc * dc + dr * dr);
    if (reproj > o->filter_geom_consistency_max_cost || (depth - bz) * (depth - bz) > depth * depth * sqr_thr)
      continue;
    deltas[ng] = fabsf(bz - depth);
    good_src[ng] = i;
    ng++;
    /* the source's word at the sampled texel: (size_t)(x + 0.5) equals
     * floor(x + 0.5) here, x + 0.5 being >= 0 (:3893) */
    if ((vis[cv->src[i]][q] & 0x80000000u) && si->ref_flag) nvis++;
  }
  if (ng < min_consistent || nvis > 0) {
    rv[p] = 0u;
    return;
  }
  uint32_t bits = rv[p];
  for (int k = 0; k < ng; ++k) bits |= 1u << (31 - 1 - good_src[k]); /* MAX_NUM_CROSS_SRC - 1 - idx (:3907) */
  rv[p] = bits;
  if (ng == 0) return; /* the reference reads an unset entry; this build leaves 0 */
  /* BubbleSort(delt_depths, num_good_src_depth - 1) sorts all but the last
   * entry; the median index is int((num_good_src_depth - 1) * 0.5) (:3912-3914) */
  const int len = ng - 1;
  for (int a = 0; a < len - 1; ++a)
    for (int b = 0; b < len - 1 - a; ++b)
      if (deltas[b] > deltas[b + 1]) {
        const float t = deltas[b];
        deltas[b] = deltas[b + 1];
        deltas[b + 1] = t;
      }
  const int nth = (int)((float)(ng - 1) * 0.5f);
  if (delt) *delt = deltas[nth] / (o->depth_diff_threshold * depth);
}

int pmo_cross_dedup(const pm_cross_options* o, int num_images, const pm_dedup_image* images, int num_views,
                    const pm_dedup_view* views, int num_threads) {
#ifdef _OPENMP
  if (num_threads > 0) omp_set_num_threads(num_threads);
#else
  (void)num_threads;
#endif
  if (num_images < 1 || num_views < 0) return -1;
  uint32_t** vis = (uint32_t**)calloc((size_t)num_images, sizeof(uint32_t*));
  if (!vis) return -1;
  int status = 0;
  /* FillGpuMat (kernel_functions.cu:3930-3941) */
  for (int i = 0; i < num_images && status == 0; ++i) {
    const size_t n = (size_t)images[i].width * images[i].height;
    vis[i] = (uint32_t*)malloc(n * sizeof(uint32_t) + 4);
    if (!vis[i]) {
      status = -1;
      break;
    }
    for (size_t k = 0; k < n; ++k) vis[i][k] = 0x80000000u;
  }
  for (int k = 0; k < num_views && status == 0; ++k) {
    const pm_dedup_view* cv = &views[k];
    if (cv->ref < 0 || cv->ref >= num_images || cv->num_src < 0 || cv->num_src > 31) {
      status = -1;
      break;
    }
    const pm_dedup_image* ref = &images[cv->ref];
    const long plane = (long)ref->width * ref->height;
    if (cv->delt)
      for (long p = 0; p < plane; ++p) cv->delt[p] = 0.0f; /* FillWithScalar(0.0f) (:1595) */
    if (cv->num_src < 1) continue; /* FindSrcImageLists gave < 2 entries (:1586-1588) */
    pm_image srcs[31];
    pm_camera cams[31];
    memset(srcs, 0, sizeof(srcs));
    for (int s = 0; s < cv->num_src; ++s) {
      if (cv->src[s] < 0 || cv->src[s] >= num_images || cv->src[s] == cv->ref) status = -1;
      if (status) break;
      srcs[s].width = images[cv->src[s]].width;
      srcs[s].height = images[cv->src[s]].height;
      cams[s] = images[cv->src[s]].cam;
    }
    if (status) break;
    pm_view_desc vd;
    memset(&vd, 0, sizeof(vd));
    vd.num_src = cv->num_src;
    vd.ref_cam = ref->cam;
    vd.src = srcs;
    vd.src_cams = cams;
    float poses[31 * PM_POSE_STRIDE];
    float k8[8];
    pmo_pack_poses(&vd, poses, k8);
    const int minc = o->filter_min_num_consistent < cv->num_src ? o->filter_min_num_consistent : cv->num_src;
    /* a view writes only its ref's words and reads its sources' (none is the
     * ref), so its pixels are independent */
    long p;
#pragma omp parallel for schedule(dynamic, 256)
    for (p = 0; p < plane; ++p) {
      const int row = (int)(p / ref->width), col = (int)(p % ref->width);
      dedup_pixel(o, images, cv, vis, poses, k8, minc, row, col, cv->delt ? &cv->delt[p] : NULL);
    }
  }
  if (status == 0) {
    for (int i = 0; i < num_images; ++i) {
      const size_t n = (size_t)images[i].width * images[i].height;
      if (images[i].visibility) memcpy(images[i].visibility, vis[i], n * sizeof(uint32_t));
      if (images[i].dedup_depth) /* GetDeduplicDepthMapsLayer (kernel_functions.cu:3943-3959) */
        for (size_t k = 0; k < n; ++k) images[i].dedup_depth[k] = (vis[i][k] & 0x80000000u) ? images[i].depth[k] : 0.0f;
    }
  }
  for (int i = 0; i < num_images; ++i) free(vis[i]);
  free(vis);
  return status;
}

/* ------------------------------------------------------------------------ */
/* ComputeCurvatureMap: kernel_functions.cu:1603-1651 */

/* float -> unsigned short of the reference's GpuMat<unsigned short>::Set:
 * toward zero; NaN and negatives give 0, values above 65535 saturate. */
static uint16_t store_u16(float x) {
  if (x != x || x <= 0.0f) return 0;
  if (x >= 65535.0f) return 65535;
  return (uint16_t)x;
}

/* utils.h:206-221, ascending */
static void bubble_sort(float* vals, int length) {
  for (int i = 0; i < length - 1; ++i)
    for (int j = 0; j < length - 1 - i; ++j)
      if (vals[j] > vals[j + 1]) {
        const float t = vals[j];
        vals[j] = vals[j + 1];
        vals[j + 1] = t;
      }
}

/* filter_kernel_x / filter_kernel_y, kernel_functions.cu:56-60 */
static const float kScharrX[3][3] = {{-3, 0, 3}, {-10, 0, 10}, {-3, 0, 3}};
static const float kScharrY[3][3] = {{-3, -10, -3}, {0, 0, 0}, {3, 10, 3}};

int pmo_curvature_map(const uint8_t* ref, int W, int H, const float* normal, uint16_t* out) {
  if (!ref || !normal || !out || W <= 0 || H <= 0) return -1;
  const long plane = (long)W * H;
#pragma omp parallel for schedule(static)
  for (int row = 0; row < H; ++row) {
    for (int col = 0; col < W; ++col) {
      /* tex2D with border addressing and point filtering on the 8-bit image
       * read as normalised float: v / 255 inside, 0 outside */
      float conv_x = 0, conv_y = 0;
      for (int dr = -1; dr <= 1; ++dr) {
        for (int dc = -1; dc <= 1; ++dc) {
          const int x = col + dc, y = row + dr;
          const float color = (x < 0 || x >= W || y < 0 || y >= H) ? 0.0f : (float)ref[(long)y * W + x] / 255.0f;
          conv_x += color * kScharrX[dr + 1][dc + 1];
          conv_y += color * kScharrY[dr + 1][dc + 1];
        }
      }
      const float grad = sqrtf(conv_x * conv_x + conv_y * conv_y);

      const int min_x = col - 2 > 0 ? col - 2 : 0;
      const int max_x = col + 2 < W - 1 ? col + 2 : W - 1;
      const int min_y = row - 2 > 0 ? row - 2 : 0;
      const int max_y = row + 2 < H - 1 ? row + 2 : H - 1;
      const long p = (long)row * W + col;
      const float n[3] = {normal[p], normal[plane + p], normal[2 * plane + p]};
      int counter = 0;
      float angles[24];
      for (int y = min_y; y <= max_y; ++y) {
        for (int x = min_x; x <= max_x; ++x) {
          if (y == row && x == col) continue;
          const long q = (long)y * W + x;
          const float dr = n[0] * normal[q] + n[1] * normal[plane + q] + n[2] * normal[2 * plane + q];
          /* acos is NaN outside [-1, 1] */
          angles[counter++] = (dr >= -1.0f && dr <= 1.0f) ? pm_acosf(dr) : NAN;
        }
      }
      const float g = grad * 10000.0f;
      out[p] = store_u16(g < 32767.0f ? g : 32767.0f);
      if (counter > 0) {
        bubble_sort(angles, counter);
        out[plane + p] = store_u16(angles[counter / 2] * 10000.0f);
      } else {
        out[plane + p] = store_u16((float)(M_PI * 10000.0));
      }
    }
  }
  return 0;
}

/* ComputeGeomGradientMap: kernel_functions.cu:3095-3120.  Pixels with
 * depth <= 0 are not written there (uninitialised memory); they read 0. */
static void geom_gradient_map(const float* depth, int W, int H, float* grad) {
#pragma omp parallel for schedule(static)
  for (int row = 0; row < H; ++row) {
    for (int col = 0; col < W; ++col) {
      const float d = depth[(long)row * W + col];
      if (d <= 0) {
        grad[(long)row * W + col] = 0.0f;
        continue;
      }
      float conv_x = 0, conv_y = 0;
      for (int dr = -1; dr <= 1; ++dr) {
        for (int dc = -1; dc <= 1; ++dc) {
          const int nrow = row + dr, ncol = col + dc;
          if (nrow < 0 || nrow >= H || ncol < 0 || ncol >= W) continue;
          const float diff_depth = fabsf(depth[(long)nrow * W + ncol] - d) / d;
          conv_x += diff_depth * kScharrX[dr + 1][dc + 1];
          conv_y += diff_depth * kScharrY[dr + 1][dc + 1];
        }
      }
      grad[(long)row * W + col] = sqrtf(conv_x * conv_x + conv_y * conv_y);
    }
  }
}

/* MedianFilter: kernel_functions.cu:3122-3244, for one pixel. */
static void median_pixel(float* depth, float* normal, const float* grad_map, int row, int col, int W, int H,
                         float diff_thres) {
  const long plane = (long)W * H;
  const float grad_thres = 1.4f * diff_thres;
  float depths[17], nx[17], ny[17], nz[17];
  int index = 0;
  const long c = (long)row * W + col;
  const float grad = grad_map[c];
  const float cdepth = depth[c];
  const float inv_cdepth = 1.0f / cdepth;
  depths[index] = cdepth;
  nx[index] = normal[c];
  ny[index] = normal[plane + c];
  nz[index] = normal[2 * plane + c];
  index++;
  /* the cross at distances 1 and 3: up, down, left, right; then the knight
   * moves (dirs, :3214) */
  int taps[16][2];
  int nt = 0;
  for (int i = 0; i < 2; ++i) {
    const int k = (i << 1) + 1;
    taps[nt][0] = -k, taps[nt][1] = 0, ++nt;
    taps[nt][0] = k, taps[nt][1] = 0, ++nt;
    taps[nt][0] = 0, taps[nt][1] = -k, ++nt;
    taps[nt][0] = 0, taps[nt][1] = k, ++nt;
  }
  static const int dirs[8][2] = {{-2, 1}, {-1, 2}, {1, 2}, {2, 1}, {2, -1}, {1, -2}, {-1, -2}, {-2, -1}};
  for (int i = 0; i < 8; ++i) taps[nt][0] = dirs[i][0], taps[nt][1] = dirs[i][1], ++nt;
  for (int t = 0; t < nt; ++t) {
    const int nrow = row + taps[t][0], ncol = col + taps[t][1];
    if (nrow < 0 || nrow >= H || ncol < 0 || ncol >= W) continue;
    const long q = (long)nrow * W + ncol;
    depths[index] = depth[q];
    const float diff_depth = fabsf(depths[index] - cdepth) * inv_cdepth;
    if (grad < grad_thres || (grad >= grad_thres && diff_depth < diff_thres)) {
      nx[index] = normal[q];
      ny[index] = normal[plane + q];
      nz[index] = normal[2 * plane + q];
      index++;
    }
  }
  bubble_sort(depths, index);
  bubble_sort(nx, index);
  bubble_sort(ny, index);
  bubble_sort(nz, index);
  const int nth = index >> 1;
  const float inv_norm = 1.0f / sqrtf(nx[nth] * nx[nth] + ny[nth] * ny[nth] + nz[nth] * nz[nth]);
  depth[c] = depths[nth];
  normal[c] = nx[nth] * inv_norm;
  normal[plane + c] = ny[nth] * inv_norm;
  normal[2 * plane + c] = nz[nth] * inv_norm;
}

/* Run()'s median filter (patch_match_cuda.cu:991-1008): the gradient map,
 * then BlackPixelFilter ((row + col) even) and RedPixelFilter (odd), in
 * place.  Every tap of a pixel has the other colour. */
int pmo_median_filter(float* depth, float* normal, int W, int H, float diff_thres) {
  if (!depth || !normal || W <= 0 || H <= 0) return -1;
  float* grad = (float*)malloc(sizeof(float) * (size_t)W * H);
  if (!grad) return -1;
  geom_gradient_map(depth, W, H, grad);
  for (int color = 0; color < 2; ++color) {
#pragma omp parallel for schedule(static)
    for (int row = 0; row < H; ++row)
      for (int col = (row + color) & 1; col < W; col += 2) median_pixel(depth, normal, grad, row, col, W, H, diff_thres);
  }
  free(grad);
  return 0;
}

/* MaskDepthMapKernel: kernel_functions.cu:3080-3093 (the last row and
 * column are left as they are) */
int pmo_mask_depth(const uint8_t* mask, int W, int H, float* depth) {
  if (!mask || !depth || W <= 0 || H <= 0) return -1;
  for (int row = 0; row < H - 1; ++row)
    for (int col = 0; col < W - 1; ++col)
      if (mask[(long)row * W + col] == 0) depth[(long)row * W + col] = 0.0f;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Run(): patch_match_cuda.cu:748-1057 */

/* One phase of Run() over rows [r0, r1) (the init over [i0, i1)).  The
 * threshold and the random-stream pass follow the phase's iteration. */
static void run_phase(pass_ctx* x, int32_t phase, int r0, int r1, int i0, int i1) {
  const int kind = PMO_PHASE_KIND(phase), iter = PMO_PHASE_ITER(phase), color = PMO_PHASE_COLOR(phase);
  const uint32_t goff = x->opt->geom_consistency ? PM_PASS_GEOM_OFFSET : 0u;
  if (kind == PMO_PHASE_INIT) {
    x->th_mc = x->opt->init_ncc_matching_cost;
    x->pass = PM_PASS_INIT + goff;
    run_all(x, init_pixel, i0, i1);
  } else if (kind == PMO_PHASE_SWEEP) {
    x->th_mc = pmo_th_mc(x->opt, iter);
    x->pass = PM_PASS_SWEEP(iter, color) + goff;
    run_color(x, sweep_pixel, color, r0, r1);
  } else {
    x->th_mc = pmo_th_mc(x->opt, iter);
    x->pass = PM_PASS_LOCAL(iter, color) + goff;
    run_color(x, local_opt_pixel, color, r0, r1);
  }
}

/* Run()'s phase list (patch_match_cuda.cu:748-880): InitWorkspace +
 * InitConfMapKernel, then per iteration BlackPixelUpdateOpt,
 * RedPixelUpdateOpt and, from the fourth iteration (iter > 2) with
 * local_optimization, BlackPixelLocalOptimization and
 * RedPixelLocalOptimization.  Black pixels are those with (row + col) even:
 * BlackPixelUpdateOpt (kernel_functions.cu:2657-2662) maps thread x of a
 * block of even width (THREADS_PER_BLOCK = 32) to row 2y when x is even and
 * 2y + 1 when x is odd; black is colour 0 here.  Returns the count. */
static int run_schedule(const pm_options* opt, int32_t* phases) {
  int n = 0;
  phases[n++] = PMO_PHASE(PMO_PHASE_INIT, 0, 0);
  for (int iter = 0; iter < opt->num_iterations; ++iter) {
    phases[n++] = PMO_PHASE(PMO_PHASE_SWEEP, iter, 0);
    phases[n++] = PMO_PHASE(PMO_PHASE_SWEEP, iter, 1);
    if (opt->local_optimization && iter > 2) {
      phases[n++] = PMO_PHASE(PMO_PHASE_LOCAL, iter, 0);
      phases[n++] = PMO_PHASE(PMO_PHASE_LOCAL, iter, 1);
    }
  }
  return n;
}

/* The body shared by pmo_run_view and pmo_run_phases: state set up as
 * InitWorkspace does, the given phases, then (post != 0) the steps after the
 * iterations in Run()'s order. */
static int run_core(const pm_options* opt, const pm_view_desc* view, const pm_outputs* out, const int32_t* phases,
                    int num_phases, int post, int row_begin, int row_end, int num_threads, pmo_counters* counters) {
#ifdef _OPENMP
  if (num_threads > 0) omp_set_num_threads(num_threads);
#else
  (void)num_threads;
#endif
  pmo_view v;
  uint8_t* rc;
  uint8_t* sc[PM_MAX_NUM_SRC];
  if (opt->thk < 1 || opt->thk > 4) return -1;
  for (int k = 0; k < num_phases; ++k) {
    const int kind = PMO_PHASE_KIND(phases[k]);
    if (kind < PMO_PHASE_INIT || kind > PMO_PHASE_LOCAL || PMO_PHASE_ITER(phases[k]) >= 8192) return -1;
  }
  if (prepare_view(opt, view, &v, &rc, sc) != 0) return -1;
  if (opt->geom_consistency && !view->src_depth_maps) {
    free_view(&v, rc, sc);
    return -1;
  }
  const size_t plane = (size_t)v.W * v.H;
  pmo_state st;
  st.depth = (float*)calloc(plane, sizeof(float));
  st.normal = (float*)calloc(3 * plane, sizeof(float));
  st.conf = (float*)calloc(plane, sizeof(float));
  st.cost = (float*)calloc(plane * v.N, sizeof(float));
  st.view_sel = (uint32_t*)calloc(plane, sizeof(uint32_t));
  if (!opt->init_depth_random) {
    if (!view->init_depth || !view->init_normal) {
      free_view(&v, rc, sc);
      return -1;
    }
    memcpy(st.depth, view->init_depth, sizeof(float) * plane);
    memcpy(st.normal, view->init_normal, sizeof(float) * 3 * plane);
  }
  int r0 = row_begin < 0 ? 0 : row_begin;
  int r1 = (row_end <= 0 || row_end > v.H) ? v.H : row_end;
  pmo_counters local;
  memset(&local, 0, sizeof(local));
  pass_ctx x;
  x.opt = opt;
  x.v = &v;
  x.st = &st;
  x.th_mc = opt->init_ncc_matching_cost;
  x.pass = 0;
  x.cnt = counters ? &local : NULL;
  /* A band run initialises the band plus a margin that covers every
   * propagation offset (21 rows) and the window radius. */
  const int margin = 21 + opt->window_radius;
  const int i0 = (r0 > 0 || r1 < v.H) ? (r0 - margin < 0 ? 0 : r0 - margin) : 0;
  const int i1 = (r0 > 0 || r1 < v.H) ? (r1 + margin > v.H ? v.H : r1 + margin) : v.H;
  for (int k = 0; k < num_phases; ++k) run_phase(&x, phases[k], r0, r1, i0, i1);
  int status = 0;
  if (post) {
    if (opt->geom_consistency && opt->est_curvature && out->curvature)
      pmo_curvature_map(v.ref, v.W, v.H, st.normal, out->curvature);
    if (opt->median_filter) pmo_median_filter(st.depth, st.normal, v.W, v.H, opt->depth_diff_threshold);
    if (opt->filter && opt->conf_filter) {
      for (int color = 0; color < 2; ++color) run_color(&x, filter_pixel, color, r0, r1);
    }
    if (view->ref_mask.data) {
      uint8_t* m = (uint8_t*)malloc(plane);
      if (m) {
        for (int r = 0; r < v.H; ++r)
          memcpy(m + (size_t)r * v.W, view->ref_mask.data + (size_t)r * view->ref_mask.stride, v.W);
        pmo_mask_depth(m, v.W, v.H, st.depth);
        free(m);
      } else {
        status = -1;
      }
    }
  }
  if (counters) *counters = local;
  if (out->depth) memcpy(out->depth, st.depth, sizeof(float) * plane);
  if (out->normal) memcpy(out->normal, st.normal, sizeof(float) * 3 * plane);
  if (out->conf) memcpy(out->conf, st.conf, sizeof(float) * plane);
  if (out->cost) memcpy(out->cost, st.cost, sizeof(float) * plane * v.N);
  if (out->view_sel) memcpy(out->view_sel, st.view_sel, sizeof(uint32_t) * plane);
  free(st.depth);
  free(st.normal);
  free(st.conf);
  free(st.cost);
  free(st.view_sel);
  free_view(&v, rc, sc);
  return status;
}

int pmo_run_view(const pm_options* opt, const pm_view_desc* view, const pm_outputs* out, int row_begin,
                 int row_end, int num_threads, pmo_counters* counters) {
  if (opt->num_iterations < 0 || opt->num_iterations > 4096) return -1;
  int32_t* phases = (int32_t*)malloc(sizeof(int32_t) * (size_t)(1 + 4 * opt->num_iterations));
  if (!phases) return -1;
  const int n = run_schedule(opt, phases);
  const int rc = run_core(opt, view, out, phases, n, 1, row_begin, row_end, num_threads, counters);
  free(phases);
  return rc;
}

int pmo_run_phases(const pm_options* opt, const pm_view_desc* view, const pm_outputs* out, const int32_t* phases,
                   int num_phases, int num_threads) {
  if (num_phases < 0 || (num_phases > 0 && !phases)) return -1;
  return run_core(opt, view, out, phases, num_phases, 0, 0, 0, num_threads, NULL);
}

int pmo_schedule(const pm_options* opt, int32_t* phases, int capacity) {
  if (opt->num_iterations < 0 || capacity < 1 + 4 * opt->num_iterations) return -1;
  return run_schedule(opt, phases);
}

/* Geometric term of every pixel's depth against every source with the
 * reference K (N planar W*H, NaN where depth <= 0). */
int pmo_geom_cost_maps(const pm_options* opt, const pm_view_desc* view, const float* depth, int num_threads,
                       float* out) {
#ifdef _OPENMP
  if (num_threads > 0) omp_set_num_threads(num_threads);
#else
  (void)num_threads;
#endif
  if (!view->src_depth_maps) return -1;
  pmo_view v;
  uint8_t* rc;
  uint8_t* sc[PM_MAX_NUM_SRC];
  if (prepare_view(opt, view, &v, &rc, sc) != 0) return -1;
  const int W = v.W, N = v.N;
  const long plane = (long)W * v.H;
  long p;
#pragma omp parallel for schedule(dynamic, 64)
  for (p = 0; p < plane; ++p) {
    const int row = (int)(p / W), col = (int)(p % W);
    const float np[3] = {v.refInvK[0] * (float)col + v.refInvK[1], v.refInvK[2] * (float)row + v.refInvK[3], 1.0f};
    for (int i = 0; i < N; ++i)
      out[i * plane + p] = depth[p] > 0.0f ? geom_cost(opt, &v, row, col, depth[p], np, i, v.refK, NULL) : NAN;
  }
  free_view(&v, rc, sc);
  return 0;
}
