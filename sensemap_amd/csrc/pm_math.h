/* Deterministic FP32 primitives shared by the gfx950 kernels and the host.
 *
 * Every function here is built only from IEEE-754 basic operations (+, -, *,
 * correctly rounded / and sqrtf), fmaf, floorf and integer bit operations.
 * Compiled with -ffp-contract=off (hipcc for the device, gcc for the host)
 * they return bit-identical results on the CPU and on the GPU, which is what
 * lets a CPU transcription of the PatchMatch sweep be compared bitwise with
 * the HIP kernels.
 *
 * The reference (src/mvs/kernel_functions.cu) uses CUDA fast-math intrinsics
 * (__expf, __sinf, __cosf, rsqrtf) and cuRAND XORWOW.  Neither can be
 * reproduced, so this header supplies:
 *   - pm_expf:   Cody & Waite argument reduction (x = k ln2 + r, ln2 split in
 *                two parts) and a degree-7 Taylor polynomial for exp(r).
 *   - pm_sinf / pm_cosf: Cody & Waite reduction by pi/2 (three-part pi/2) and
 *                Taylor polynomials on [-pi/4, pi/4].
 *   - pm_atan2f: reduction to [0, 1], then atan(t) = pi/4 + atan((t-1)/(t+1))
 *                above tan(pi/8), and the Taylor series of atan below it.
 *   - pm_acosf:  acos(x) = atan2(sqrt((1-x)(1+x)), x).
 *   - Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
 *                as easy as 1, 2, 3", SC'11) as a counter-based generator, so
 *                every draw is a pure function of (seed, view, pixel, pass,
 *                draw index) and never of thread or batch layout.
 */
#ifndef SENSEMAP_PM_MATH_H_
#define SENSEMAP_PM_MATH_H_

#include <stdint.h>
#include <math.h>
#include <string.h>

#if defined(__HIPCC__)
#define PM_HD __host__ __device__ __forceinline__
#else
#define PM_HD static inline
#endif

#define PM_PI_F 3.14159265358979323846f
#define PM_HALF_PI_F 1.57079632679489661923f
#define PM_QUARTER_PI_F 0.78539816339744830962f
#define PM_FLT_EPSILON 1.19209290e-07f
#define PM_FLT_MAX 3.40282347e+38f

PM_HD uint32_t pm_f2u(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __float_as_uint(f);
#else
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
#endif
}

PM_HD float pm_u2f(uint32_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __uint_as_float(u);
#else
  float f;
  memcpy(&f, &u, 4);
  return f;
#endif
}

/* exp(x) in FP32.  Inputs below -87 return 0 and above 88 return +inf. */
PM_HD float pm_expf(float x) {
  if (!(x > -87.0f)) return 0.0f;
  if (x > 88.0f) return pm_u2f(0x7f800000u);
  const float kLog2e = 1.44269504088896341f;
  const float kLn2Hi = 0.693145751953125f;       /* 0x3f317200: 16 bits */
  const float kLn2Lo = 1.42860682030941723212e-6f;
  const float k = floorf(fmaf(x, kLog2e, 0.5f));
  float r = fmaf(-k, kLn2Hi, x);
  r = fmaf(-k, kLn2Lo, r);
  float p = 1.0f / 5040.0f;
  p = fmaf(p, r, 1.0f / 720.0f);
  p = fmaf(p, r, 1.0f / 120.0f);
  p = fmaf(p, r, 1.0f / 24.0f);
  p = fmaf(p, r, 1.0f / 6.0f);
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  const int ki = (int)k;                         /* in [-126, 128] */
  /* 2^k split in two factors so that k = 128 and k = -126 stay representable */
  const int k1 = ki / 2;
  const int k2 = ki - k1;
  const float s1 = pm_u2f((uint32_t)(k1 + 127) << 23);
  const float s2 = pm_u2f((uint32_t)(k2 + 127) << 23);
  return (p * s1) * s2;
}

/* Reduce x by multiples of pi/2: returns r in about [-pi/4, pi/4], *quad = j mod 4. */
PM_HD float pm_reduce_half_pi(float x, int* quad) {
  const float kTwoOverPi = 0.636619772367581343f;
  const float kP1 = 1.57079625129699707031f;     /* pi/2 split in three parts */
  const float kP2 = 7.54978941586159635335e-08f;
  const float kP3 = 5.39030252995776476554e-15f;
  const float j = floorf(fmaf(x, kTwoOverPi, 0.5f));
  float r = fmaf(-j, kP1, x);
  r = fmaf(-j, kP2, r);
  r = fmaf(-j, kP3, r);
  *quad = ((int)j) & 3;
  return r;
}

PM_HD float pm_sin_poly(float r) {
  const float r2 = r * r;
  float p = -1.0f / 39916800.0f;
  p = fmaf(p, r2, 1.0f / 362880.0f);
  p = fmaf(p, r2, -1.0f / 5040.0f);
  p = fmaf(p, r2, 1.0f / 120.0f);
  p = fmaf(p, r2, -1.0f / 6.0f);
  p = p * r2;
  return fmaf(p, r, r);
}

PM_HD float pm_cos_poly(float r) {
  const float r2 = r * r;
  float p = 1.0f / 479001600.0f;
  p = fmaf(p, r2, -1.0f / 3628800.0f);
  p = fmaf(p, r2, 1.0f / 40320.0f);
  p = fmaf(p, r2, -1.0f / 720.0f);
  p = fmaf(p, r2, 1.0f / 24.0f);
  p = fmaf(p, r2, -0.5f);
  return fmaf(p, r2, 1.0f);
}

PM_HD float pm_sinf(float x) {
  int q;
  const float r = pm_reduce_half_pi(x, &q);
  const float s = pm_sin_poly(r);
  const float c = pm_cos_poly(r);
  return (q == 0) ? s : (q == 1) ? c : (q == 2) ? -s : -c;
}

PM_HD float pm_cosf(float x) {
  int q;
  const float r = pm_reduce_half_pi(x, &q);
  const float s = pm_sin_poly(r);
  const float c = pm_cos_poly(r);
  return (q == 0) ? c : (q == 1) ? -s : (q == 2) ? -c : s;
}

/* atan(t) for t in [0, 1]. */
PM_HD float pm_atan01(float t) {
  const float kTanPi8 = 0.414213562373095049f;
  float base = 0.0f;
  if (t > kTanPi8) {
    t = (t - 1.0f) / (t + 1.0f);
    base = PM_QUARTER_PI_F;
  }
  const float t2 = t * t;
  float p = 1.0f / 17.0f;
  p = fmaf(p, t2, -1.0f / 15.0f);
  p = fmaf(p, t2, 1.0f / 13.0f);
  p = fmaf(p, t2, -1.0f / 11.0f);
  p = fmaf(p, t2, 1.0f / 9.0f);
  p = fmaf(p, t2, -1.0f / 7.0f);
  p = fmaf(p, t2, 1.0f / 5.0f);
  p = fmaf(p, t2, -1.0f / 3.0f);
  p = p * t2;
  return base + fmaf(p, t, t);
}

/* atan2(y, x); signed zeros are treated as +0 (atan2(0, -0) = 0). */
PM_HD float pm_atan2f(float y, float x) {
  const float ax = fabsf(x);
  const float ay = fabsf(y);
  const float mx = ax > ay ? ax : ay;
  const float mn = ax > ay ? ay : ax;
  if (!(mx > 0.0f)) return 0.0f;
  float r = pm_atan01(mn / mx);
  if (ay > ax) r = PM_HALF_PI_F - r;
  if (x < 0.0f) r = PM_PI_F - r;
  if (y < 0.0f) r = -r;
  return r;
}

/* acos(x) for x in [-1, 1]; callers handle |x| > 1 explicitly. */
PM_HD float pm_acosf(float x) {
  const float s = sqrtf((1.0f - x) * (1.0f + x));
  return pm_atan2f(s, x);
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 */

#define PM_PHILOX_M0 0xD2511F53u
#define PM_PHILOX_M1 0xCD9E8D57u
#define PM_PHILOX_W0 0x9E3779B9u
#define PM_PHILOX_W1 0xBB67AE85u

PM_HD uint32_t pm_mulhi32(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umulhi(a, b);
#else
  return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
#endif
}

PM_HD void pm_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; ++round) {
    const uint32_t hi0 = pm_mulhi32(PM_PHILOX_M0, c0);
    const uint32_t lo0 = PM_PHILOX_M0 * c0;
    const uint32_t hi1 = pm_mulhi32(PM_PHILOX_M1, c2);
    const uint32_t lo1 = PM_PHILOX_M1 * c2;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += PM_PHILOX_W0;
    k1 += PM_PHILOX_W1;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* One pixel's random stream within one pass.  Draw d of the stream is word
 * (d & 3) of Philox(ctr = {pixel, pass, d >> 2, view}, key = seed). */
typedef struct pm_rng {
  uint32_t key[2];
  uint32_t ctr[4];
  uint32_t buf[4];
  uint32_t draw;
} pm_rng;

PM_HD void pm_rng_init(pm_rng* g, uint64_t seed, uint32_t view, uint32_t pixel, uint32_t pass) {
  g->key[0] = (uint32_t)seed;
  g->key[1] = (uint32_t)(seed >> 32);
  g->ctr[0] = pixel;
  g->ctr[1] = pass;
  g->ctr[2] = 0;
  g->ctr[3] = view;
  g->draw = 0;
  g->buf[0] = g->buf[1] = g->buf[2] = g->buf[3] = 0;
}

PM_HD uint32_t pm_rng_u32(pm_rng* g) {
  const uint32_t lane = g->draw & 3u;
  if (lane == 0) {
    g->ctr[2] = g->draw >> 2;
    pm_philox4x32_10(g->ctr, g->key, g->buf);
  }
  g->draw++;
  return lane == 0 ? g->buf[0] : lane == 1 ? g->buf[1] : lane == 2 ? g->buf[2] : g->buf[3];
}

/* Uniform in (0, 1] with 24-bit resolution, matching the half-open range of
 * the reference's curand_uniform. */
PM_HD float pm_rng_uniform(pm_rng* g) {
  return (float)((pm_rng_u32(g) >> 8) + 1u) * 5.9604644775390625e-08f;
}

/* Pass identifiers of the random streams. */
#define PM_PASS_INIT 0u
#define PM_PASS_SWEEP(iter, color) (1u + 4u * (uint32_t)(iter) + (uint32_t)(color))
#define PM_PASS_LOCAL(iter, color) (3u + 4u * (uint32_t)(iter) + (uint32_t)(color))
#define PM_PASS_GEOM_OFFSET 0x10000u

#endif  /* SENSEMAP_PM_MATH_H_ */
