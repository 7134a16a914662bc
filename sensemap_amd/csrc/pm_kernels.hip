// pm_kernels.hip -- gfx950 kernels of the PatchMatch MVS engine.
//
// One thread per pixel of one checkerboard colour.  A 256-thread workgroup
// owns a 32x16 pixel tile (its 256 pixels of the colour); each wave64 covers
// 4 rows x 16 same-colour pixels, so the 64 windows of a wave overlap and
// their source footprints hit the same cache lines.  Per workgroup the
// reference tile (+ window halo; one 32-bit word per texel holding the
// intensity and 3x the intensity, see ncc_eval) and the 256-entry bilateral colour
// table are staged in LDS; the spatial weight table and the per-source pose
// blocks are read with uniform (scalar) loads because the source loop and
// the window position are uniform across the wave.  Source images live in
// HBM in a zero-padded "quad" layout: texel (x, y) holds the 2x2 bilinear
// footprint {(x,y), (x+1,y), (x,y+1), (x+1,y+1)}, so one 32-bit (u8) or
// 64-bit (f16) load feeds one bilinear sample with no bounds checks.
//
// Per-pixel arithmetic follows the canonical FP32 order of the CPU
// transcription (oracle/pm_oracle.c): -ffp-contract=off, explicit fmaf,
// correctly rounded division and sqrt, project math from pm_math.h.  With
// PM_RCP_EXACT the maps are bitwise equal to the transcription's.
//
// Reference semantics (src/mvs/kernel_functions.cu):
//   k_ref_moments   <- FilterKernel (gpu_mat_ref_image.cu:16-56) + ComputeGradientMap (:1655-1675)
//   k_init          <- FillWithRandomNumbers/InitNormalMap/InitDepthMap (:689-737) + InitConfMapKernel (:1497-1601)
//   k_sweep         <- Black/RedPixelUpdateOpt -> SweepFromCheckerBoardOpt (:2136-2560, :2634-2734)
//   k_local_opt     <- Black/RedPixelLocalOptimization -> LocalOptimization (:2736-3017)
//   k_filter        <- FilterDepthMap (:3282-3354)
//   k_curvature     <- ComputeCurvatureMap (:1603-1651)
//   k_mask          <- MaskDepthMapKernel (:3080-3093)
//   k_geom_grad     <- ComputeGeomGradientMap (:3095-3120)
//   k_median        <- BlackPixelFilter / RedPixelFilter -> MedianFilter (:3122-3280)
//   ncc_eval        <- ComputeNCCCost<R,S> (:1245-1340) with ComposeHomography (:431-503)
//   geom_eval       <- ComputeGeomConsistencyCostShared (:505-568)
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include <map>
#include <mutex>

#include "pm_device.h"
#include "pm_math.h"

namespace pm {
namespace {

// Pointers stored in ViewDev are generic; cast them to the global address
// space so loads and stores compile to global_* (or s_load for uniform
// addresses) instead of flat_*.
// A view descriptor as the colour-pass kernels read it: through the
// constant address space, so that the compiler may use scalar loads (the
// descriptor is written only by host copies between launches) and keeps
// its fields, the source texture bases among them, in scalar registers.
using VPtr = const __attribute__((address_space(4))) ViewDev*;
template <typename T>
using K4 = const __attribute__((address_space(4))) T;

template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* G(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

typedef float f4v __attribute__((ext_vector_type(4)));
typedef unsigned int u2v __attribute__((ext_vector_type(2)));

// 16-byte load of a float4 array element through the global address space.
__device__ __forceinline__ f4v ld_f4(const float4* p, int idx) {
  return *((const __attribute__((address_space(1))) f4v*)p + idx);
}

__constant__ int c_neighbor_offs[4][2] = {{0, -1}, {-1, 0}, {0, 1}, {1, 0}};
__constant__ int c_dir_samples[8][20] = {
    {-1, -2, 1, -2, -2, -3, 2, -3, -3, -4, 3, -4, -4, -5, 4, -5, -5, -6, 5, -6},
    {-1, 2, 1, 2, -2, 3, 2, 3, -3, 4, 3, 4, -4, 5, 4, 5, -5, 6, 5, 6},
    {-2, -1, -2, 1, -3, -2, -3, 2, -4, -3, -4, 3, -5, -4, -5, 4, -6, -5, -6, 5},
    {2, -1, 2, 1, 3, -2, 3, 2, 4, -3, 4, 3, 5, -4, 5, 4, 6, -5, 6, 5},
    {0, -3, 0, -5, 0, -7, 0, -9, 0, -11, 0, -13, 0, -15, 0, -17, 0, -19, 0, -21},
    {0, 3, 0, 5, 0, 7, 0, 9, 0, 11, 0, 13, 0, 15, 0, 17, 0, 19, 0, 21},
    {-3, 0, -5, 0, -7, 0, -9, 0, -11, 0, -13, 0, -15, 0, -17, 0, -19, 0, -21, 0},
    {3, 0, 5, 0, 7, 0, 9, 0, 11, 0, 13, 0, 15, 0, 17, 0, 19, 0, 21, 0}};
__constant__ float c_mweights[5] = {0.f, 1.f, 2.01f, 3.03f, 4.06f};

// ---------------------------------------------------------------------------
// per-lane pixel context

struct Pix {
  int row, col, lrow, lcol;
  int p;                 // row * W + col
  int c;                 // ref centre intensity (u8)
  float mu_r, var_r, inv_sw, err_r;
  float np0, np1;        // normalized point (x, y, 1)
  bool shifted;          // compensated moments for this pixel
};

// Workspace accessors (SoA over slots).
struct Ws {
  float* base;
  int slots, slot, N;  // base/slots: the launch's workspace; slot: this thread
  __device__ __attribute__((address_space(1))) float& hyp(int k, int c) const { return G(base)[(size_t)(4 * k + c) * slots + slot]; }
  __device__ __attribute__((address_space(1))) float& M(int k, int i) const {
    return G(base)[(size_t)(4 * (kMaxHyp + kNumRefine) + k * N + i) * slots + slot];
  }
  __device__ __attribute__((address_space(1))) float& RM(int k, int i) const {
    return G(base)[(size_t)(4 * (kMaxHyp + kNumRefine) + kMaxHyp * N + k * N + i) * slots + slot];
  }
  __device__ __attribute__((address_space(1))) float& Wt(int i) const {
    return G(base)[(size_t)(4 * (kMaxHyp + kNumRefine) + (kMaxHyp + kNumRefine) * N + i) * slots + slot];
  }
};

// Per-thread counts of one pixel's phase, 32 bits: even the largest window
// (65x65) x 32 sources x 16 evaluations is 2.2e6 samples per pixel and
// phase (init: init_samples <= 64, checked by pm_options_check).
struct Counters {
  uint32_t ncc_eval = 0, ncc_samples = 0, geo_eval = 0;
};

// Counter slots: [0..2] all kernels, [3..5] the sweep kernel alone.
__device__ __forceinline__ void flush_counters(VPtr v, const Counters& c, bool sweep = false) {
  if (v->counters) {
    atomicAdd(&v->counters[0], (unsigned long long)c.ncc_eval);
    atomicAdd(&v->counters[1], (unsigned long long)c.ncc_samples);
    atomicAdd(&v->counters[2], (unsigned long long)c.geo_eval);
    if (sweep) {
      atomicAdd(&v->counters[3], (unsigned long long)c.ncc_eval);
      atomicAdd(&v->counters[4], (unsigned long long)c.ncc_samples);
      atomicAdd(&v->counters[5], (unsigned long long)c.geo_eval);
    }
  }
}

// 1/z of a window sample.
// PM_RCP_EXACT: v_rcp_f32 refined by two Newton steps with FMA.  On gfx950
//   v_rcp_f32 is within one ulp of the correctly rounded reciprocal for every
//   significand (and scales exactly with exponent and sign), and two Newton
//   steps from any seed within one ulp land on the correctly rounded value,
//   so this equals 1.0f / z for every normal z, bit for bit, at about half
//   the instructions of the IEEE division sequence.
// PM_RCP_FAST: one Newton step (within one ulp, not always correctly rounded).
template <int RCP>
__device__ __forceinline__ float rcp_z(float z) {
  float r = __builtin_amdgcn_rcpf(z);
  float e = fmaf(-z, r, 1.0f);
  r = fmaf(r, e, r);
  if constexpr (RCP == PM_RCP_EXACT) {
    e = fmaf(-z, r, 1.0f);
    r = fmaf(r, e, r);
  }
  return r;
}

// The 2x2 footprint of source texel (x0, y0) in the quad layout: one
// 4-byte (u8) or 8-byte (f16) load.  tex is uniform across the wave (the
// source loop is uniform) and off is the 32-bit texel index, so the load
// uses the scalar-base + 32-bit-offset form.
template <int STORE>
struct Quad {
  using type = uint32_t;
};
template <>
struct Quad<PM_STORAGE_F16> {
  using type = u2v;
};

template <int STORE>
__device__ __forceinline__ typename Quad<STORE>::type load_quad(const void* tex, uint32_t boff) {
  const __attribute__((address_space(1))) char* b = G((const char*)tex);
  if constexpr (STORE == PM_STORAGE_U8) {
    return *(const __attribute__((address_space(1))) uint32_t*)(b + boff);
  } else {
    return *(const __attribute__((address_space(1))) u2v*)(b + boff);
  }
}

// Bilinear value of a quad at fraction (fx, fy):
// top = t00 + fx (t01 - t00), bot = t10 + fx (t11 - t10), top + fy (bot - top).
// The f16 quad holds (t00, t01 - t00, t10, t11 - t10), all exact, so the
// two horizontal lerps read the halves directly (v_fma_mix_f32) and give the
// same bits as the u8 path.
template <int STORE>
__device__ __forceinline__ float bilerp(typename Quad<STORE>::type q, float fx, float fy) {
  float top, bot;
  if constexpr (STORE == PM_STORAGE_U8) {
    const float t0 = (float)(q & 0xffu);
    const float t1 = (float)((q >> 8) & 0xffu);
    const float t2 = (float)((q >> 16) & 0xffu);
    const float t3 = (float)(q >> 24);
    top = fmaf(fx, t1 - t0, t0);
    bot = fmaf(fx, t3 - t2, t2);
  } else {
    const float t0 = __half2float(__ushort_as_half((unsigned short)(q.x & 0xffffu)));
    const float d01 = __half2float(__ushort_as_half((unsigned short)(q.x >> 16)));
    const float t2 = __half2float(__ushort_as_half((unsigned short)(q.y & 0xffffu)));
    const float d23 = __half2float(__ushort_as_half((unsigned short)(q.y >> 16)));
    top = fmaf(fx, d01, t0);
    bot = fmaf(fx, d23, t2);
  }
  return fmaf(fy, bot - top, top);
}

// floor(x) as int in one instruction (v_cvt_flr_i32_f32); the compiler
// emits v_floor_f32 + v_cvt_i32_f32 for (int)floorf(x).  Exact for the
// clamped coordinate range [-3, W+2].
__device__ __forceinline__ int cvt_flr(float x) {
  int r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

__device__ __forceinline__ void neumaier_add(float& s, float& c, float x) {
  const float t = s + x;
  c = c + ((fabsf(s) >= fabsf(x)) ? ((s - t) + x) : ((x - t) + s));
  s = t;
}

// Texel loads a window row keeps in flight (build knob).
#ifndef PM_PIPE_DEPTH
#define PM_PIPE_DEPTH 3
#endif

// ComputeNCCCost (kernel_functions.cu:1245-1340) for source i (uniform) and
// the lane's hypothesis.  tile points at the LDS ref tile entry of the
// window's top-left sample; tw is the tile row length.
template <int R_, int S_, int STORE, int RCP>
__device__ __forceinline__ float ncc_eval(VPtr v, int i, const Pix& px,
                                          const uint32_t* tile, int tw, const float* ct, const float* sp,
                                          float nx, float ny, float nz, float d, Counters& cnt) {
  const int R = R_ >= 0 ? R_ : v->radius;
  const int S = R_ >= 0 ? S_ : v->step;
  const int n = R_ >= 0 ? (2 * R_) / S_ + 1 : v->nwin;
  cnt.ncc_eval++;
  const __attribute__((address_space(1))) float* pose = G(v->poses) + i * PM_POSE_STRIDE;
  const float K0 = pose[0], K1 = pose[1], K2 = pose[2], K3 = pose[3];
  const __attribute__((address_space(1))) float* Rm = pose + 4;
  const __attribute__((address_space(1))) float* T = pose + 13;
  K4<float>* iK = v->refInvK;
  const float fcol = (float)px.col, frow = (float)px.row;
  const float ray0 = iK[0] * fcol + iK[1];
  const float ray1 = iK[2] * frow + iK[3];
  const float dist = d * (nx * ray0 + ny * ray1 + nz);
  const float inv_dist = 1.0f / dist;
  const float nd0 = inv_dist * nx, nd1 = inv_dist * ny, nd2 = inv_dist * nz;
  float A[9];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    A[3 * r + 0] = Rm[3 * r + 0] + nd0 * T[r];
    A[3 * r + 1] = Rm[3 * r + 1] + nd1 * T[r];
    A[3 * r + 2] = Rm[3 * r + 2] + nd2 * T[r];
  }
  float B[9];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    B[j] = K0 * A[j] + K1 * A[6 + j];
    B[3 + j] = K2 * A[3 + j] + K3 * A[6 + j];
    B[6 + j] = A[6 + j];
  }
  float H[9];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    H[3 * r + 0] = B[3 * r + 0] * iK[0];
    H[3 * r + 1] = B[3 * r + 1] * iK[2];
    H[3 * r + 2] = B[3 * r + 0] * iK[1] + B[3 * r + 1] * iK[3] + B[3 * r + 2];
  }
  const float zc = H[6] * fcol + H[7] * frow + H[8];
  const float izc = 1.0f / zc;
  const float xf = (H[0] * fcol + H[1] * frow + H[2]) * izc;
  const float yf = (H[3] * fcol + H[4] * frow + H[5]) * izc;
  K4<SrcDev>& s = v->src[i];
  const float Wsf = (float)s.width, Hsf = (float)s.height;
  if (!(xf > -1.0f && xf < Wsf && yf > -1.0f && yf < Hsf)) return kMaxCost;
  if (px.var_r < kVarThreshold) return kMaxCost;
  const float c0 = (float)(px.col - R), r0 = (float)(px.row - R);
  float bx = H[0] * c0 + H[1] * r0 + H[2];
  float by = H[3] * c0 + H[4] * r0 + H[5];
  float bz = H[6] * c0 + H[7] * r0 + H[8];
  const float fs = (float)S;
  const float hxc = fs * H[0], hyc = fs * H[3], hzc = fs * H[6];
  const float hxr = fs * H[1], hyr = fs * H[4], hzr = fs * H[7];
  const float umax = Wsf + 2.0f, vmax = Hsf + 2.0f;
  const int pitch = s.pitch;
  const void* tex = s.tex;
  const uint32_t c4 = ((uint32_t)px.c << 16) | (3u * (uint32_t)px.c);
  float s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
  float e1 = 0.0f, e2 = 0.0f, e3 = 0.0f;
  const float mu_r = px.mu_r;
  const int ipitch = pitch;
  constexpr uint32_t kQuadShift = STORE == PM_STORAGE_U8 ? 2u : 3u;  // bytes per quad: 4 or 8
  const uint32_t pad_bytes = (uint32_t)(kQuadPad * pitch + kQuadPad) << kQuadShift;
  // One window walk; SHIFTED selects the compensated moment update.  A
  // sample is issued (projection, fraction, texel load) ahead of its use:
  // with compile-time window shapes each row keeps kDepth loads in flight,
  // so a wave waits for memory once per row instead of once per sample.
  // Issue order, arithmetic and accumulation order are those of a plain
  // sample-by-sample walk.
  struct Smp {
    float fx, fy;
    typename Quad<STORE>::type q;
  };
  auto issue = [&](float x, float y, float z, auto clamp_tag) __attribute__((always_inline)) {
    const float rz = rcp_z<RCP>(z);
    float u = x * rz, w_ = y * rz;
    if constexpr (decltype(clamp_tag)::value) {
      // clamp to [-3, W+2]: v_med3_f32 equals fminf(fmaxf(u, -3), umax) for
      // every input, NaN and infinities included (a NaN goes to -3)
      u = __builtin_amdgcn_fmed3f(u, -3.0f, umax);
      w_ = __builtin_amdgcn_fmed3f(w_, -3.0f, vmax);
    }
    Smp sm;
    // fraction = min(u - floor(u), 1 - 2^-24): v_fract_f32
    sm.fx = __builtin_amdgcn_fractf(u);
    sm.fy = __builtin_amdgcn_fractf(w_);
    const int x0 = cvt_flr(u), y0 = cvt_flr(w_);
    // texel (x0 + 4, y0 + 4) of the padded image, as a byte offset: the
    // signed y0 * pitch + x0 (>= -3 pitch - 3) plus the uniform padding offset
    const int t = __mul24(y0, ipitch) + x0;
    const uint32_t boff = ((uint32_t)t << kQuadShift) + pad_bytes;
    sm.q = load_quad<STORE>(tex, boff);
    // keep the load where it is issued: without this the scheduler sinks
    // the f16 loads next to their use and the row's loads no longer overlap
    __builtin_amdgcn_sched_barrier(0);
    return sm;
  };
  auto walk = [&](auto shifted_tag, auto clamp_tag) __attribute__((always_inline)) {
    constexpr bool kShifted = decltype(shifted_tag)::value;
    for (int ir = 0; ir < n; ++ir) {
      const uint32_t* trow = tile + ir * S * tw;
      const float* sprow = sp + ir * n;
      auto consume = [&](int j, const Smp& sm) __attribute__((always_inline)) {
        const float sv = bilerp<STORE>(sm.q, sm.fx, sm.fy);
        // tile word (ri << 16) | 3 ri: byte 2 is the intensity; v_sad_u16
        // adds the absolute differences of both 16-bit halves, so against
        // the centre's (c << 16) | 3 c it gives 3|ri - c| + |ri - c| =
        // 4|ri - c|, the colour weight's LDS byte offset
        const uint32_t tw4 = trow[j * S];
        const float r = (float)((tw4 >> 16) & 0xffu);  // v_cvt_f32_ubyte2
        const float w = sprow[j] * *(const float*)((const char*)ct + __builtin_amdgcn_sad_u16(tw4, c4, 0u));
        if constexpr (!kShifted) {
          const float ws = w * sv;
          s1 = s1 + ws;
          s2 = fmaf(ws, sv, s2);
          s3 = fmaf(ws, r, s3);
        } else {
          const float dd = sv - mu_r;
          const float rd = r - mu_r;
          const float wd = w * dd;
          neumaier_add(s1, e1, wd);
          neumaier_add(s2, e2, wd * dd);
          neumaier_add(s3, e3, wd * rd);
        }
      };
      float x = bx, y = by, z = bz;  // the next sample to issue
      if constexpr (R_ >= 0) {
        constexpr int kN = R_ >= 0 ? (2 * R_) / S_ + 1 : 1;
        constexpr int kDepth = kN < PM_PIPE_DEPTH ? kN : PM_PIPE_DEPTH;
        Smp ring[kDepth];
#pragma unroll
        for (int j = 0; j < kDepth; ++j) {
          ring[j] = issue(x, y, z, clamp_tag);
          x = x + hxc;
          y = y + hyc;
          z = z + hzc;
        }
#pragma unroll
        for (int j = 0; j < kN; ++j) {
          const Smp cur = ring[j % kDepth];
          if (j + kDepth < kN) {
            ring[j % kDepth] = issue(x, y, z, clamp_tag);
            x = x + hxc;
            y = y + hyc;
            z = z + hzc;
          }
          consume(j, cur);
        }
      } else {
        for (int j = 0; j < n; ++j) {
          const Smp cur = issue(x, y, z, clamp_tag);
          x = x + hxc;
          y = y + hyc;
          z = z + hzc;
          consume(j, cur);
        }
      }
      bx = bx + hxr;
      by = by + hyr;
      bz = bz + hzr;
    }
  };
  // The clamps are identities when every sample lands inside [-3, W+2] x
  // [-3, H+2].  Under the homography the window maps onto the convex
  // quadrilateral of its 4 corner samples when z > 0 at all of them (z is
  // affine over the window, so its minimum is at a corner).  Corners at
  // least kInsideMargin px inside the range with z varying by at most 4x
  // bound the incremental walk's rounding (<= ~30 ulp of the homogeneous
  // coordinates, < 0.2 px for images up to 2^14 px wide) well below the
  // margin, so such a window needs no clamp.  The choice is wave-uniform:
  // one lane near a border keeps the clamped walk for the wave.
  bool inside;
  {
    constexpr float kInsideMargin = 0.5f;
    const float span = (float)(n - 1);
    const float ex = span * hxc, ey = span * hyc, ez = span * hzc;
    const float fx_ = span * hxr, fy_ = span * hyr, fz_ = span * hzr;
    const float cx[4] = {bx, bx + ex, bx + fx_, (bx + ex) + fx_};
    const float cy[4] = {by, by + ey, by + fy_, (by + ey) + fy_};
    const float cz[4] = {bz, bz + ez, bz + fz_, (bz + ez) + fz_};
    float zmin = cz[0], zmax = cz[0];
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      zmin = fminf(zmin, cz[k]);
      zmax = fmaxf(zmax, cz[k]);
      const float iz = __builtin_amdgcn_rcpf(cz[k]);  // a test with a margin: the approximation serves
      const float u = cx[k] * iz, w_ = cy[k] * iz;
      ok = ok && u >= -3.0f + kInsideMargin && u <= umax - kInsideMargin && w_ >= -3.0f + kInsideMargin &&
           w_ <= vmax - kInsideMargin;
    }
    inside = __all(ok && zmin > 0.0f && zmax <= 4.0f * zmin);
  }
  if (!px.shifted) {
    if (inside)
      walk(std::false_type{}, std::false_type{});
    else
      walk(std::false_type{}, std::true_type{});
  } else {
    walk(std::true_type{}, std::true_type{});
  }
  cnt.ncc_samples += (uint32_t)(n * n);
  const float inv_sw = px.inv_sw;
  float var_s, cov;
  if (!px.shifted) {
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
    cov = m3 - px.err_r * m1;
  }
  if (var_s < kVarThreshold) return kMaxCost;
  const float val = 1.0f - cov / sqrtf(px.var_r * var_s);
  return fmaxf(0.0f, fminf(kMaxCost, val));
}

// ComputeGeomConsistencyCostShared (kernel_functions.cu:505-568).
__device__ __forceinline__ float geom_eval(VPtr v, int i, const Pix& px, float depth,
                                           K4<float>* K, Counters& cnt) {
  cnt.geo_eval++;
  const __attribute__((address_space(1))) float* P = G(v->poses) + i * PM_POSE_STRIDE + 16;
  const __attribute__((address_space(1))) float* iP = G(v->poses) + i * PM_POSE_STRIDE + 28;
  const float fx = depth * px.np0, fy = depth * px.np1, fz = depth * 1.0f;
  const float inv_fz = 1.0f / (P[8] * fx + P[9] * fy + P[10] * fz + P[11]);
  float sc = inv_fz * (P[0] * fx + P[1] * fy + P[2] * fz + P[3]);
  float sr = inv_fz * (P[4] * fx + P[5] * fy + P[6] * fz + P[7]);
  K4<SrcDev>& s = v->src[i];
  float u = fminf(fmaxf(sc, -3.0f), (float)s.width + 2.0f);
  float w_ = fminf(fmaxf(sr, -3.0f), (float)s.height + 2.0f);
  const float x0f = floorf(u), y0f = floorf(w_);
  const float fxx = u - x0f, fyy = w_ - y0f;
  const f4v q = ld_f4(s.depth_q, ((int)y0f + kQuadPad) * s.pitch + ((int)x0f + kQuadPad));
  const float top = fmaf(fxx, q.y - q.x, q.x);
  const float bot = fmaf(fxx, q.w - q.z, q.z);
  const float sd = fmaf(fyy, bot - top, top);
  if (sd == 0.0f) return v->geom_max_cost;
  sc = sc * sd;
  sr = sr * sd;
  const float bx = iP[0] * sc + iP[1] * sr + iP[2] * sd + iP[3];
  const float by = iP[4] * sc + iP[5] * sr + iP[6] * sd + iP[7];
  const float bz = iP[8] * sc + iP[9] * sr + iP[10] * sd + iP[11];
  const float ibz = 1.0f / bz;
  const float bcol = ibz * K[0] * bx + K[1];
  const float brow = ibz * K[2] * by + K[3];
  const float dc = (float)px.col - bcol;
  const float dr = (float)px.row - brow;
  return fminf(v->geom_max_cost, sqrtf(dc * dc + dr * dr));
}

// ComputeMultiViewCostVector (kernel_functions.cu:1447-1495), priors off.
template <int R_, int S_, int STORE, int RCP>
__device__ __forceinline__ float mv_cost(VPtr v, int i, const Pix& px, const uint32_t* tile,
                                         int tw, const float* ct, const float* sp, float nx, float ny, float nz,
                                         float d, K4<float>* Kgeo, Counters& cnt) {
  float cost = ncc_eval<R_, S_, STORE, RCP>(v, i, px, tile, tw, ct, sp, nx, ny, nz, d, cnt);
  if (v->geom) cost = cost + v->geom_regularizer * geom_eval(v, i, px, d, Kgeo, cnt);
  return cost;
}

// ---------------------------------------------------------------------------
// random hypotheses (identical to the transcription)

__device__ __forceinline__ void random_normal(pm_rng* g, float r0, float r1, float n[3]) {
  float v1 = 0.0f, v2 = 0.0f, s = 2.0f;
  for (int t = 0; t < 64 && s >= 1.0f; ++t) {
    v1 = 2.0f * pm_rng_uniform(g) - 1.0f;
    v2 = 2.0f * pm_rng_uniform(g) - 1.0f;
    s = v1 * v1 + v2 * v2;
  }
  if (s >= 1.0f) {
    v1 = 0.0f;
    v2 = 0.0f;
    s = 0.0f;
  }
  const float s_norm = 2.0f * sqrtf(1.0f - s);
  n[0] = v1 * s_norm;
  n[1] = v2 * s_norm;
  n[2] = 1.0f - 2.0f * s;
  if (n[0] * r0 + n[1] * r1 + n[2] * 1.0f > 0.0f) {
    n[0] = -n[0];
    n[1] = -n[1];
    n[2] = -n[2];
  }
}

__device__ __forceinline__ float random_depth(pm_rng* g, float dmin, float dmax) {
  return pm_rng_uniform(g) * (dmax - dmin) + dmin;
}

__device__ __forceinline__ void perturb_normal(pm_rng* g, const float n[3], float a1, float a2, float out[3]) {
  const float y = pm_atan2f(sqrtf(n[0] * n[0] + n[1] * n[1]), n[2]);
  const float x = pm_atan2f(n[1], n[0]);
  const float nx = a1 * (2.0f * pm_rng_uniform(g) - 1.0f) + x;
  const float ny = a2 * (2.0f * pm_rng_uniform(g) - 1.0f) + y;
  const float siny = pm_sinf(ny);
  out[0] = pm_cosf(nx) * siny;
  out[1] = pm_sinf(nx) * siny;
  out[2] = pm_cosf(ny);
}

// CorrectNormal (kernel_functions.cu:227-306) with the reference's matrix.
__device__ __forceinline__ void correct_normal(float n[3], const float vd[3], float inv_view_len) {
  const float cosl = n[0] * vd[0] + n[1] * vd[1] + n[2] * vd[2];
  if (!(cosl >= 0.0f)) return;
  float c[3];
  c[0] = n[1] * vd[2] - n[2] * vd[1];
  c[1] = n[2] * vd[0] - n[0] * vd[2];
  c[2] = n[0] * vd[1] - n[1] * vd[0];
  const float a = cosl * inv_view_len;
  float phi;
  if (!(a <= 1.0f)) {
    phi = -0.001f;
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
  float Rr[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const float id = (k == 0 || k == 4 || k == 8) ? 1.0f : 0.0f;
    Rr[k] = id + o[k] * sp + oo[k];
  }
  const float r0 = Rr[0] * n[0] + Rr[1] * n[1] + Rr[2] * n[2];
  const float r1 = Rr[3] * n[0] + Rr[4] * n[1] + Rr[5] * n[2];
  const float r2 = Rr[6] * n[0] + Rr[7] * n[1] + Rr[8] * n[2];
  n[0] = r0;
  n[1] = r1;
  n[2] = r2;
}

__device__ __forceinline__ float propagate_depth(float d1, const float n1[3], K4<float>* iK, int row1, int col1,
                                                 float np0, float np1) {
  const float x1 = d1 * (iK[0] * (float)col1 + iK[1]);
  const float y1 = d1 * (iK[2] * (float)row1 + iK[3]);
  const float z1 = d1;
  const float dd = n1[0] * x1 + n1[1] * y1 + n1[2] * z1;
  const float t = dd / (n1[0] * np0 + n1[1] * np1 + n1[2] * 1.0f);
  if (fabsf(t) < 1e-5f) return d1;
  return t;
}

// ---------------------------------------------------------------------------
// workgroup set-up shared by the colour-pass kernels

template <int R_, int S_>
struct SpSize {
  static constexpr int kFloats = R_ >= 0 ? ((2 * R_) / S_ + 1) * ((2 * R_) / S_ + 1)
                                         : (2 * PM_MAX_WINDOW_RADIUS + 1) * (2 * PM_MAX_WINDOW_RADIUS + 1);
};

template <int R_>
struct TileSize {
  // The staged ref tile has one row pitch for every tile shape (the widest
  // tile plus the halo), so window offsets in LDS are compile-time
  // constants; rows: the tallest tile plus the halo.
  static constexpr int kR = R_ >= 0 ? R_ : PM_MAX_WINDOW_RADIUS;
  static constexpr int kPitch = (1 << kTileLog2WMax) + 2 * kR;
  static constexpr int kWords = ((kTilePixels >> kTileLog2WMin) + 2 * kR) * kPitch;
};

// Stage the colour table and the ref tile (zero outside the image) into
// LDS, then map the thread to its pixel.  Returns false for threads whose
// pixel lies outside the image (after the barrier).
// Persistent colour-pass kernels: a launch runs at most as many workgroups
// as the chip holds at once, and each workgroup loops claiming 32x16 tiles
// until the launch's tiles are exhausted.  The launch's tiles are split into
// 8 contiguous ranges, one per XCD (workgroups are dispatched round-robin
// over the XCDs, so block b starts on XCD b % 8); a workgroup takes tiles
// from its own XCD's range first, so the workgroups of one XCD cover compact
// bands of a view and share source footprints in that XCD's L2, and steals
// from the other ranges once its own is empty.
__device__ __forceinline__ int claim_tile(const LaunchTable* __restrict__ lt, int* next) {
  __syncthreads();  // the previous tile's LDS and *next are no longer read
  if (threadIdx.x == 0) {
    const __attribute__((address_space(1))) LaunchTable* g = G(lt);
    const int T = g->total_tiles;
    uint32_t* q = g->queue;
    const int xcd = blockIdx.x & 7;
    int tile = -1;
    for (int k = 0; k < 8 && tile < 0; ++k) {
      const int x = (xcd + k) & 7;
      const int lo = (int)(((long long)T * x) >> 3), hi = (int)(((long long)T * (x + 1)) >> 3);
      if (lo >= hi) continue;
      // a range already used up is skipped without an atomic: at the end of
      // a launch every workgroup scans all 8 ranges, and those atomics on 8
      // addresses serialise (counters only grow, so a stale read costs one
      // atomic, never a tile)
      if (__hip_atomic_load(q + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= (uint32_t)(hi - lo)) continue;
      const uint32_t c = atomicAdd(q + x, 1u);
      if (c < (uint32_t)(hi - lo)) tile = lo + (int)c;
    }
    *next = tile;
  }
  __syncthreads();
  // the same value in every lane: keep it (and everything resolved from it,
  // the view's parameters and pointers) in scalar registers
  return __builtin_amdgcn_readfirstlane(*next);
}

// The launch entry (view, colour, pass) a tile belongs to.
struct Ctx {
  VPtr v;
  int color;
  uint32_t pass;
  float th_mc;
  int local_block;  // tile index within the view
  int lw;           // log2 of the tile width
  float* ws;
  int ws_slots;
};

__device__ __forceinline__ Ctx resolve(const LaunchTable* __restrict__ lt, int t) {
  const __attribute__((address_space(1))) LaunchTable* g = G(lt);
  int lo = 0, hi = g->n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (g->e[mid].tile_start <= t) lo = mid;
    else hi = mid - 1;
  }
  Ctx c;
  c.v = (VPtr)g->e[lo].v;
  c.color = g->e[lo].color;
  c.pass = g->e[lo].pass;
  c.th_mc = g->e[lo].th_mc;
  c.local_block = t - g->e[lo].tile_start;
  c.lw = g->tile_lw;
  c.ws = g->ws;
  c.ws_slots = g->ws_slots;
  return c;
}

// Stage the colour table, the spatial weights and the ref tile (zero outside
// the image) into LDS, then map the thread to its pixel.  Returns false for
// threads whose pixel lies outside the image (after the barrier).
__device__ __forceinline__ bool setup_pass(const Ctx& cx, int R, int pitch, uint32_t* tile, float* ct, float* spl,
                                           Pix& px) {
  const VPtr v = cx.v;
  const int t = threadIdx.x;
  ct[t] = G(v->ct)[t];
  const int nsp = v->nwin * v->nwin;
  for (int e = t; e < nsp; e += kSweepThreads) spl[e] = G(v->sp)[e];
  const int W = v->W, H = v->H;
  const int lw = cx.lw;
  const int tiles_x = (W + (1 << lw) - 1) >> lw;
  const int ty = cx.local_block / tiles_x;
  const int tx = cx.local_block - ty * tiles_x;
  const int row0 = ty * (kTilePixels >> lw), col0 = tx << lw;
  const int tw = (1 << lw) + 2 * R;
  const int th = (kTilePixels >> lw) + 2 * R;
  for (int e = t; e < tw * th; e += kSweepThreads) {
    const int er = e / tw, ec = e - er * tw;
    const int rr = row0 - R + er, cc = col0 - R + ec;
    const uint32_t ri = (rr < 0 || rr >= H || cc < 0 || cc >= W) ? 0u : (uint32_t)G(v->ref)[rr * W + cc];
    tile[er * pitch + ec] = (ri << 16) | (3u * ri);
  }
  __syncthreads();
  px.lrow = t >> (lw - 1);
  px.row = row0 + px.lrow;
  px.lcol = 2 * (t & ((1 << (lw - 1)) - 1)) + ((px.row + cx.color) & 1);
  px.col = col0 + px.lcol;
  if (px.row >= H || px.col >= W) return false;
  px.p = px.row * W + px.col;
  const f4v m = ld_f4(v->refm, px.p);
  px.mu_r = m.x;
  px.var_r = m.y;
  px.inv_sw = m.z;
  px.err_r = m.w;
  px.c = (int)(tile[(px.lrow + R) * pitch + px.lcol + R] >> 16);
  px.shifted = (v->accum_mode == PM_ACCUM_COMPENSATED) && (px.var_r < kLowTexVar);
  px.np0 = v->refInvK[0] * (float)px.col + v->refInvK[1];
  px.np1 = v->refInvK[2] * (float)px.row + v->refInvK[3];
  return true;
}

// Workspace slot of this thread: slots belong to the resident workgroup, not
// to the tile, so the workspace is sized by the grid.
__device__ __forceinline__ int slot_index() { return blockIdx.x * kSweepThreads + threadIdx.x; }

// ---------------------------------------------------------------------------
// init: random (best of init_samples) or given depth/normal, then InitConf

template <int R_, int S_, int STORE, int RCP>
__device__ __forceinline__ void init_tile(const Ctx& cx, uint32_t* tile, float* ct, float* spl) {
  const VPtr v = cx.v;
  const uint32_t pass = cx.pass;
  const int R = R_ >= 0 ? R_ : v->radius;
  Pix px;
  const int tw = R_ >= 0 ? TileSize<R_>::kPitch : (1 << kTileLog2WMax) + 2 * R;
  if (!setup_pass(cx, R, tw, tile, ct, spl, px)) return;
  const uint32_t* wt = tile + px.lrow * tw + px.lcol;
  const int N = v->N;
  const int plane = v->W * v->H;
  Ws ws{cx.ws, cx.ws_slots, slot_index(), N};
  Counters cnt;
  pm_rng g;
  pm_rng_init(&g, v->seed, (uint32_t)v->view_id, (uint32_t)px.p, pass);
  const int thk = v->thk < N ? v->thk : N;
  float hd, hn[3];
  float best_conf = 0.0f;
  uint32_t best_sel = 0;
  float bd = 0.0f, bn[3] = {0.0f, 0.0f, 0.0f};
  const int K = v->init_random ? (v->init_samples > 0 ? v->init_samples : 1) : 1;
  for (int smp = 0; smp < K; ++smp) {
    if (v->init_random) {
      hd = random_depth(&g, v->depth_min, v->depth_max);
      random_normal(&g, px.np0, px.np1, hn);
    } else {
      hd = G(v->depth)[px.p];
      hn[0] = G(v->normal)[px.p];
      hn[1] = G(v->normal)[plane + px.p];
      hn[2] = G(v->normal)[2 * plane + px.p];
      if (hd <= 0.0f) {
        hd = random_depth(&g, v->depth_min, v->depth_max);
        random_normal(&g, px.np0, px.np1, hn);
      }
      if (hn[0] * px.np0 + hn[1] * px.np1 + hn[2] * 1.0f >= 0.0f) random_normal(&g, px.np0, px.np1, hn);
    }
    for (int i = 0; i < N; ++i) {
      ws.M(cx.color, i) = mv_cost<R_, S_, STORE, RCP>(v, i, px, wt, tw, ct, spl, hn[0], hn[1], hn[2], hd, v->refK, cnt);
    }
    // thk smallest costs in stable ascending order (= the reference's
    // bubble sort, kernel_functions.cu:1588-1597)
    float conf = 0.0f;
    uint32_t taken = 0;
    for (int j = 0; j < thk; ++j) {
      int bi = -1;
      float bv = 0.0f;
      for (int i = 0; i < N; ++i) {
        if ((taken >> i) & 1u) continue;
        const float cv = ws.M(cx.color, i);
        if (bi < 0 || cv < bv) {
          bi = i;
          bv = cv;
        }
      }
      taken |= 1u << bi;
      conf = conf + bv;
    }
    conf = conf / c_mweights[thk];
    if (smp == 0 || conf < best_conf) {
      best_conf = conf;
      best_sel = taken;
      bd = hd;
      bn[0] = hn[0];
      bn[1] = hn[1];
      bn[2] = hn[2];
      for (int i = 0; i < N; ++i) ws.RM(cx.color, i) = ws.M(cx.color, i);
    }
  }
  G(v->depth)[px.p] = bd;
  G(v->normal)[px.p] = bn[0];
  G(v->normal)[plane + px.p] = bn[1];
  G(v->normal)[2 * plane + px.p] = bn[2];
  G(v->conf)[px.p] = best_conf;
  G(v->view_sel)[px.p] = best_sel;
  for (int i = 0; i < N; ++i) G(v->cost)[i * plane + px.p] = ws.RM(cx.color, i);
  flush_counters(v, cnt);
}

template <int R_, int S_, int STORE, int RCP>
__global__ __launch_bounds__(kSweepThreads) void k_init(const LaunchTable* __restrict__ lt) {
  __shared__ uint32_t tile[TileSize<R_>::kWords];
  __shared__ float ct[256];
  __shared__ float spl[SpSize<R_, S_>::kFloats];
  __shared__ int next;
  for (int t = claim_tile(lt, &next); t >= 0; t = claim_tile(lt, &next))
    init_tile<R_, S_, STORE, RCP>(resolve(lt, t), tile, ct, spl);
}

// ---------------------------------------------------------------------------
// checkerboard sweep

// Optional occupancy target for the sweep kernel (waves per SIMD), a build
// knob for tuning register use against latency hiding.
#ifdef PM_SWEEP_WAVES_PER_EU
#define PM_SWEEP_OCC __attribute__((amdgpu_waves_per_eu(PM_SWEEP_WAVES_PER_EU, 8)))
#else
#define PM_SWEEP_OCC
#endif

template <int R_, int S_, int STORE, int RCP>
__device__ __forceinline__ void sweep_tile(const Ctx& cx, uint32_t* tile, float* ct, float* spl) {
  const VPtr v = cx.v;
  const uint32_t pass = cx.pass;
  const float th_mc = cx.th_mc;
  const int R = R_ >= 0 ? R_ : v->radius;
  Pix px;
  const int tw = R_ >= 0 ? TileSize<R_>::kPitch : (1 << kTileLog2WMax) + 2 * R;
  if (!setup_pass(cx, R, tw, tile, ct, spl, px)) return;
  const uint32_t* wt = tile + px.lrow * tw + px.lcol;
  const int W = v->W, H = v->H, N = v->N;
  const int plane = W * H;
  const int p = px.p, row = px.row, col = px.col;
  K4<float>* iK = v->refInvK;
  Ws ws{cx.ws, cx.ws_slots, slot_index(), N};
  Counters cnt;

  const float vd[3] = {px.np0, px.np1, 1.0f};
  const float inv_view_len = 1.0f / sqrtf(vd[0] * vd[0] + vd[1] * vd[1] + vd[2] * vd[2]);

  // candidates
  int nh = 0;
  ws.hyp(0, 0) = G(v->normal)[p];
  ws.hyp(0, 1) = G(v->normal)[plane + p];
  ws.hyp(0, 2) = G(v->normal)[2 * plane + p];
  ws.hyp(0, 3) = G(v->depth)[p];
  nh = 1;
  if (G(v->grad_ok)[p]) {
    float nconf[4];
    int order[4];
    int nvalid = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      order[k] = k;
      const int u = col + c_neighbor_offs[k][1], w = row + c_neighbor_offs[k][0];
      if (u < 0 || u >= W || w < 0 || w >= H) {
        nconf[k] = PM_FLT_MAX;
        continue;
      }
      nvalid++;
      nconf[k] = G(v->conf)[w * W + u];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
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
    for (int k = 0; k < nvalid; ++k) {
      const int idx = order[k];
      const int u = col + c_neighbor_offs[idx][1], w = row + c_neighbor_offs[idx][0];
      const int q = w * W + u;
      float nn[3] = {G(v->normal)[q], G(v->normal)[plane + q], G(v->normal)[2 * plane + q]};
      correct_normal(nn, vd, inv_view_len);
      const float dd = propagate_depth(G(v->depth)[q], nn, iK, w, u, px.np0, px.np1);
      ws.hyp(nh, 0) = nn[0];
      ws.hyp(nh, 1) = nn[1];
      ws.hyp(nh, 2) = nn[2];
      ws.hyp(nh, 3) = dd;
      nh++;
    }
  }
  for (int gi = 0; gi < 8; ++gi) {
    float best_conf = PM_FLT_MAX;
    int br = 0, bc = 0;
#pragma unroll
    for (int f = 0; f < 10; ++f) {
      const int nr = row + c_dir_samples[gi][2 * f];
      const int ncn = col + c_dir_samples[gi][2 * f + 1];
      if (nr < 0 || nr >= H || ncn < 0 || ncn >= W) continue;
      const float cv = G(v->conf)[nr * W + ncn];
      if (best_conf > cv) {
        best_conf = cv;
        br = nr;
        bc = ncn;
      }
    }
    if (best_conf < PM_FLT_MAX) {
      const int q = br * W + bc;
      float nn[3] = {G(v->normal)[q], G(v->normal)[plane + q], G(v->normal)[2 * plane + q]};
      correct_normal(nn, vd, inv_view_len);
      const float dd = propagate_depth(G(v->depth)[q], nn, iK, br, bc, px.np0, px.np1);
      ws.hyp(nh, 0) = nn[0];
      ws.hyp(nh, 1) = nn[1];
      ws.hyp(nh, 2) = nn[2];
      ws.hyp(nh, 3) = dd;
      nh++;
    }
  }

  // cost matrix, source-major so consecutive evaluations share a source
  for (int i = 0; i < N; ++i) {
    ws.M(0, i) = G(v->cost)[i * plane + p];
    for (int k = 1; k < kMaxHyp; ++k) {
      if (k < nh) {
        ws.M(k, i) = mv_cost<R_, S_, STORE, RCP>(v, i, px, wt, tw, ct, spl, ws.hyp(k, 0), ws.hyp(k, 1), ws.hyp(k, 2),
                                                 ws.hyp(k, 3), v->refK, cnt);
      }
    }
  }

  // view selection (kernel_functions.cu:2364-2426)
  uint32_t flags = 0;
  for (int i = 0; i < N; ++i) {
    float pr = 0.0f;
    int n1 = 0, n2 = 0;
    for (int k = 0; k < nh; ++k) {
      const float cv = ws.M(k, i);
      pr = pr + pm_expf(-(cv * cv) * v->inv_beta2);
      n1 += (cv < th_mc);
      n2 += (cv > v->max_ncc_matching_cost);
    }
    ws.Wt(i) = pr;
    if (n1 >= v->num_good && n2 <= v->num_bad) flags |= 1u << i;
  }
  if (flags == 0) {
    flush_counters(v, cnt, true);
    return;
  }
  float psum = 0.0f;
  for (int i = 0; i < N; ++i) {
    const float pr = ((flags >> i) & 1u) ? ws.Wt(i) : 0.0f;
    psum = psum + pr;
  }
  const float inv_psum = 1.0f / psum;
  float cum = 0.0f;
  for (int i = 0; i < N; ++i) {
    const float pr = ((flags >> i) & 1u) ? ws.Wt(i) : 0.0f;
    cum = cum + pr * inv_psum;
    ws.Wt(i) = cum;
    ws.RM(0, i) = 0.0f;
  }
  pm_rng g;
  pm_rng_init(&g, v->seed, (uint32_t)v->view_id, (uint32_t)p, pass);
  for (int smp = 0; smp < v->num_samples; ++smp) {
    const float rp = pm_rng_uniform(&g) - PM_FLT_EPSILON;
    for (int i = 0; i < N; ++i) {
      if (ws.Wt(i) > rp) {
        ws.RM(0, i) = v->random_optimization ? 1.0f : ws.RM(0, i) + 1.0f;
        break;
      }
    }
  }
  float wnorm = 0.0f;
  for (int i = 0; i < N; ++i) wnorm = wnorm + ws.RM(0, i);
  if (wnorm <= 1e-5f) {
    flush_counters(v, cnt, true);
    return;
  }
  const float inv_wnorm = 1.0f / wnorm;
  uint32_t sel = 0;
  for (int i = 0; i < N; ++i) {
    const float wv = ws.RM(0, i) * inv_wnorm;
    ws.Wt(i) = wv;
    if (wv > 0.0f) sel |= 1u << i;
  }
  int best = -1;
  float minc = PM_FLT_MAX;
  for (int k = 0; k < nh; ++k) {
    float cc = 0.0f;
    for (int i = 0; i < N; ++i) cc = cc + ws.Wt(i) * ws.M(k, i);
    if (cc <= minc) {
      minc = cc;
      best = k;
    }
  }
  G(v->view_sel)[p] = sel;
  if (best < 0) best = 0;
  float depth = ws.hyp(best, 3);
  float normal[3] = {ws.hyp(best, 0), ws.hyp(best, 1), ws.hyp(best, 2)};

  // refinement hypotheses (kernel_functions.cu:2449-2486)
  const float dlo = (1.0f - v->random_depth_ratio) * depth;
  const float dhi = (1.0f + v->random_depth_ratio) * depth;
  const float nd = random_depth(&g, dlo, dhi);
  float nn[3];
  perturb_normal(&g, normal, v->angle1, v->angle2, nn);
  correct_normal(nn, vd, inv_view_len);
  const float rd = random_depth(&g, v->depth_min, v->depth_max_eff);
  float rn[3];
  random_normal(&g, vd[0], vd[1], rn);
  correct_normal(rn, vd, inv_view_len);
  const int hb = kMaxHyp;  // refinement hypotheses stored after the candidates
  ws.hyp(hb + 0, 0) = normal[0];
  ws.hyp(hb + 0, 1) = normal[1];
  ws.hyp(hb + 0, 2) = normal[2];
  ws.hyp(hb + 0, 3) = depth;
  ws.hyp(hb + 1, 0) = nn[0];
  ws.hyp(hb + 1, 1) = nn[1];
  ws.hyp(hb + 1, 2) = nn[2];
  ws.hyp(hb + 1, 3) = depth;
  ws.hyp(hb + 2, 0) = normal[0];
  ws.hyp(hb + 2, 1) = normal[1];
  ws.hyp(hb + 2, 2) = normal[2];
  ws.hyp(hb + 2, 3) = nd;
  ws.hyp(hb + 3, 0) = rn[0];
  ws.hyp(hb + 3, 1) = rn[1];
  ws.hyp(hb + 3, 2) = rn[2];
  ws.hyp(hb + 3, 3) = depth;
  ws.hyp(hb + 4, 0) = normal[0];
  ws.hyp(hb + 4, 1) = normal[1];
  ws.hyp(hb + 4, 2) = normal[2];
  ws.hyp(hb + 4, 3) = rd;

  float rc0 = 0.0f;
  for (int i = 0; i < N; ++i) {
    const float m = ws.M(best, i);
    ws.RM(0, i) = m;
    rc0 = rc0 + ws.Wt(i) * m;
  }
  for (int i = 0; i < N; ++i) {
    for (int k = 1; k < kNumRefine; ++k) {
      ws.RM(k, i) = mv_cost<R_, S_, STORE, RCP>(v, i, px, wt, tw, ct, spl, ws.hyp(hb + k, 0), ws.hyp(hb + k, 1),
                                                ws.hyp(hb + k, 2), ws.hyp(hb + k, 3), v->refK, cnt);
    }
  }
  int b2 = -1;
  minc = PM_FLT_MAX;
  for (int k = 0; k < kNumRefine; ++k) {
    float cc = rc0;
    if (k > 0) {
      cc = 0.0f;
      for (int i = 0; i < N; ++i) cc = cc + ws.Wt(i) * ws.RM(k, i);
    }
    if (cc <= minc) {
      minc = cc;
      b2 = k;
    }
  }
  if (b2 < 0) {  // every refined cost NaN: the reference writes nothing
    flush_counters(v, cnt, true);
    return;
  }
  G(v->conf)[p] = minc;
  G(v->depth)[p] = ws.hyp(hb + b2, 3);
  G(v->normal)[p] = ws.hyp(hb + b2, 0);
  G(v->normal)[plane + p] = ws.hyp(hb + b2, 1);
  G(v->normal)[2 * plane + p] = ws.hyp(hb + b2, 2);
  for (int i = 0; i < N; ++i) G(v->cost)[i * plane + p] = ws.RM(b2, i);
  flush_counters(v, cnt, true);
}

template <int R_, int S_, int STORE, int RCP>
__global__ __launch_bounds__(kSweepThreads) PM_SWEEP_OCC void k_sweep(const LaunchTable* __restrict__ lt) {
  __shared__ uint32_t tile[TileSize<R_>::kWords];
  __shared__ float ct[256];
  __shared__ float spl[SpSize<R_, S_>::kFloats];
  __shared__ int next;
  for (int t = claim_tile(lt, &next); t >= 0; t = claim_tile(lt, &next))
    sweep_tile<R_, S_, STORE, RCP>(resolve(lt, t), tile, ct, spl);
}

// ---------------------------------------------------------------------------
// local optimisation

template <int R_, int S_, int STORE, int RCP>
__device__ __forceinline__ void local_opt_tile(const Ctx& cx, uint32_t* tile, float* ct, float* spl) {
  const VPtr v = cx.v;
  const uint32_t pass = cx.pass;
  const int R = R_ >= 0 ? R_ : v->radius;
  Pix px;
  const int tw = R_ >= 0 ? TileSize<R_>::kPitch : (1 << kTileLog2WMax) + 2 * R;
  if (!setup_pass(cx, R, tw, tile, ct, spl, px)) return;
  const uint32_t* wt = tile + px.lrow * tw + px.lcol;
  const int W = v->W, H = v->H, N = v->N;
  const int plane = W * H;
  const int p = px.p, row = px.row, col = px.col;
  Ws ws{cx.ws, cx.ws_slots, slot_index(), N};
  Counters cnt;

  float conf = G(v->conf)[p];
  const float depth = G(v->depth)[p];
  const float normal[3] = {G(v->normal)[p], G(v->normal)[plane + p], G(v->normal)[2 * plane + p]};
  const uint32_t sel = G(v->view_sel)[p];
  const float angle1 = (float)(4.0 * 0.0174532925199432);
  const float angle2 = (float)(2.0 * 0.0174532925199432);
  float low = depth, up = depth;
  float m0 = normal[0], m1 = normal[1], m2 = normal[2];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int u = col + c_neighbor_offs[k][1], w = row + c_neighbor_offs[k][0];
    if (u < 0 || u >= W || w < 0 || w >= H) continue;
    const int q = w * W + u;
    const float ndp = G(v->depth)[q];
    m0 = m0 + G(v->normal)[q];
    m1 = m1 + G(v->normal)[plane + q];
    m2 = m2 + G(v->normal)[2 * plane + q];
    if (ndp <= low) low = ndp;
    if (ndp > up) up = ndp;
  }
  float mn[3] = {m0, m1, m2};
  const float inv_norm = 1.0f / sqrtf(mn[0] * mn[0] + mn[1] * mn[1] + mn[2] * mn[2]);
  mn[0] = mn[0] * inv_norm;
  mn[1] = mn[1] * inv_norm;
  mn[2] = mn[2] * inv_norm;
  if (sel == 0) return;
  pm_rng g;
  pm_rng_init(&g, v->seed, (uint32_t)v->view_id, (uint32_t)p, pass);
  for (int t = 0; t < 5; ++t) {
    float hn[3];
    const float hd = random_depth(&g, low, up);
    perturb_normal(&g, mn, angle1, angle2, hn);
    ws.hyp(t, 0) = hn[0];
    ws.hyp(t, 1) = hn[1];
    ws.hyp(t, 2) = hn[2];
    ws.hyp(t, 3) = hd;
  }
  // A trial's mean over the selected sources is its running sum / mw.
  // Costs are >= 0 (NCC in [0, 2], the geometric term scaled by a
  // non-negative regulariser) and a float sum of non-negative terms never
  // decreases, so once sum / mw >= the pixel's conf the trial cannot be
  // taken: its remaining evaluations are skipped (a wave skips a walk when
  // every lane has ruled the trial out).  The maps are those of the full
  // evaluation; the transcription prunes the same evaluations.
  const float mw = (float)__popc(sel);
  const float conf0 = conf;
  const bool prune = !v->geom || v->geom_regularizer >= 0.0f;
  float part0 = 0.0f, part1 = 0.0f, part2 = 0.0f, part3 = 0.0f, part4 = 0.0f;
  uint32_t live = 0x1fu;
  for (int i = 0; i < N; ++i) {
    if (!((sel >> i) & 1u)) continue;
#pragma unroll 1
    for (int t = 0; t < 5; ++t) {
      if (!((live >> t) & 1u)) continue;
      const float m = mv_cost<R_, S_, STORE, RCP>(v, i, px, wt, tw, ct, spl, ws.hyp(t, 0), ws.hyp(t, 1),
                                                  ws.hyp(t, 2), ws.hyp(t, 3), v->refK, cnt);
      ws.M(t, i) = m;
      float sum;
      if (t == 0) sum = part0 = part0 + m;
      else if (t == 1) sum = part1 = part1 + m;
      else if (t == 2) sum = part2 = part2 + m;
      else if (t == 3) sum = part3 = part3 + m;
      else sum = part4 = part4 + m;
      if (prune && sum / mw >= conf0) live &= ~(1u << t);
    }
  }
  int best = -1;
  const float parts[5] = {part0, part1, part2, part3, part4};
#pragma unroll
  for (int t = 0; t < 5; ++t) {
    if (!((live >> t) & 1u)) continue;
    const float mconf = parts[t] / mw;
    if (mconf < conf) {
      conf = mconf;
      best = t;
    }
  }
  if (best < 0) {
    flush_counters(v, cnt);
    return;
  }
  const float bd = ws.hyp(best, 3);
  const float bn[3] = {ws.hyp(best, 0), ws.hyp(best, 1), ws.hyp(best, 2)};
  G(v->conf)[p] = conf;
  G(v->depth)[p] = bd;
  G(v->normal)[p] = bn[0];
  G(v->normal)[plane + p] = bn[1];
  G(v->normal)[2 * plane + p] = bn[2];
  K4<float>* Kgeo = v->compat_lo_geo_bug ? v->refInvK : v->refK;
  for (int i = 0; i < N; ++i) {
    float cv;
    if ((sel >> i) & 1u) {
      cv = ws.M(best, i);
    } else {
      cv = mv_cost<R_, S_, STORE, RCP>(v, i, px, wt, tw, ct, spl, bn[0], bn[1], bn[2], bd, Kgeo, cnt);
    }
    G(v->cost)[i * plane + p] = cv;
  }
  flush_counters(v, cnt);
}

template <int R_, int S_, int STORE, int RCP>
__global__ __launch_bounds__(kSweepThreads) void k_local_opt(const LaunchTable* __restrict__ lt) {
  __shared__ uint32_t tile[TileSize<R_>::kWords];
  __shared__ float ct[256];
  __shared__ float spl[SpSize<R_, S_>::kFloats];
  __shared__ int next;
  for (int t = claim_tile(lt, &next); t >= 0; t = claim_tile(lt, &next))
    local_opt_tile<R_, S_, STORE, RCP>(resolve(lt, t), tile, ct, spl);
}

// Tiles of a launch of an element-wise phase, assigned by workgroup index
// (no LDS, no barriers).  The work per tile is a few loads, so the tile
// queue's atomics (all on one cache line) would cost more than the tiles.
template <typename F>
__device__ __forceinline__ void for_each_tile(const LaunchTable* __restrict__ lt, F&& f) {
  const int T = __builtin_amdgcn_readfirstlane(G(lt)->total_tiles);
  for (int t = blockIdx.x; t < T; t += gridDim.x) f(t);
}

// ---------------------------------------------------------------------------
// FilterDepthMap (kernel_functions.cu:3282-3354)

__device__ __forceinline__ void filter_tile(const Ctx& cx) {
  const VPtr v = cx.v;
  const int t = threadIdx.x;
  const int W = v->W, H = v->H;
  const int lw = cx.lw;
  const int tiles_x = (W + (1 << lw) - 1) >> lw;
  const int ty = cx.local_block / tiles_x;
  const int tx = cx.local_block - ty * tiles_x;
  const int row = ty * (kTilePixels >> lw) + (t >> (lw - 1));
  const int col = (tx << lw) + 2 * (t & ((1 << (lw - 1)) - 1)) + ((row + cx.color) & 1);
  if (row >= H || col >= W) return;
  const int plane = W * H;
  const int p = row * W + col;
  K4<float>* iK = v->refInvK;
  const float depth = G(v->depth)[p];
  if (depth <= 0.0f) return;
  const float n[3] = {G(v->normal)[p], G(v->normal)[plane + p], G(v->normal)[2 * plane + p]};
  const float conf = G(v->conf)[p];
  const float pt[3] = {depth * (iK[0] * (float)col + iK[1]), depth * (iK[2] * (float)row + iK[3]), depth};
  int good = 0;
  for (int k = 0; k < 4; ++k) {
    const int u = col + c_neighbor_offs[k][1], w = row + c_neighbor_offs[k][0];
    if (u < 0 || u >= W || w < 0 || w >= H) continue;
    const int q = w * W + u;
    const float nd = G(v->depth)[q];
    const float nn[3] = {G(v->normal)[q], G(v->normal)[plane + q], G(v->normal)[2 * plane + q]};
    const float nc = G(v->conf)[q];
    const float npt[3] = {nd * (iK[0] * (float)u + iK[1]), nd * (iK[2] * (float)w + iK[3]), nd};
    const float e0 = pt[0] - npt[0], e1 = pt[1] - npt[1], e2 = pt[2] - npt[2];
    const float d1 = fabsf(e0 * n[0] + e1 * n[1] + e2 * n[2]);
    const float d2 = fabsf(e0 * nn[0] + e1 * nn[1] + e2 * nn[2]);
    if ((conf < v->conf_threshold && d1 / depth < v->depth_diff_threshold) ||
        (nc < v->conf_threshold && d2 / nd < v->depth_diff_threshold)) {
      good++;
    }
  }
  if (good < 2) {
    G(v->depth)[p] = 0.0f;
    G(v->normal)[p] = 0.0f;
    G(v->normal)[plane + p] = 0.0f;
    G(v->normal)[2 * plane + p] = 0.0f;
    G(v->conf)[p] = 0.0f;
  }
}

__global__ __launch_bounds__(kSweepThreads) void k_filter(const LaunchTable* __restrict__ lt) {
  for_each_tile(lt, [&](int t) { filter_tile(resolve(lt, t)); });
}

// Pixel of a colour-pass tile (the mapping of filter_tile); false outside the view.
__device__ __forceinline__ bool tile_pixel(const Ctx& cx, int& row, int& col) {
  const int W = cx.v->W;
  const int lw = cx.lw;
  const int t = threadIdx.x;
  const int tiles_x = (W + (1 << lw) - 1) >> lw;
  const int ty = cx.local_block / tiles_x;
  const int tx = cx.local_block - ty * tiles_x;
  row = ty * (kTilePixels >> lw) + (t >> (lw - 1));
  col = (tx << lw) + 2 * (t & ((1 << (lw - 1)) - 1)) + ((row + cx.color) & 1);
  return row < cx.v->H && col < W;
}

// Ascending bubble sort of the first n of K values (the reference's
// BubbleSort, utils.h:206-221), unrolled over K slots so the array stays in
// registers; the compare-swaps past n are switched off, which leaves the
// same sequence of swaps as the loop over n.
template <int K>
__device__ __forceinline__ void bubble_sort_regs(float (&a)[K], int n) {
#pragma unroll
  for (int i = 0; i < K - 1; ++i) {
#pragma unroll
    for (int j = 0; j < K - 1 - i; ++j) {
      if (j < n - 1 - i && a[j] > a[j + 1]) {
        const float t = a[j];
        a[j] = a[j + 1];
        a[j + 1] = t;
      }
    }
  }
}

template <int K>
__device__ __forceinline__ float pick_reg(const float (&a)[K], int k) {
  float r = 0.0f;
#pragma unroll
  for (int j = 0; j < K; ++j) r = j == k ? a[j] : r;
  return r;
}

// float -> unsigned short as the reference's stores convert: toward zero,
// NaN and negatives to 0, saturating at 65535.
__device__ __forceinline__ uint16_t sat_u16(float x) {
  if (!(x > 0.0f)) return 0;
  if (x >= 65535.0f) return 65535;
  return (uint16_t)(uint32_t)x;
}

// acos as the reference calls it: NaN outside [-1, 1] (a dot product of two
// unit normals can round above 1).
__device__ __forceinline__ float acos_or_nan(float x) { return fabsf(x) <= 1.0f ? pm_acosf(x) : __builtin_nanf(""); }

// ComputeCurvatureMap (kernel_functions.cu:1603-1651).  Slice 0: the Scharr
// gradient magnitude of the reference image (intensities / 255, zero
// outside the image, filter_kernel_x/y at :56-60) x 10000, capped at 32767.
// Slice 1: 10000 x the median of the angles between the pixel's normal and
// the normals of its 5x5 neighbourhood clipped to the image, taken as the
// reference does (angles in window order, bubble sort, element count / 2),
// so NaN angles keep their positions as they do there.  The compacted angle
// list and the sort are unrolled over the 24 slots with selects, so they
// stay in registers.
__device__ __forceinline__ void curvature_tile(const Ctx& cx) {
  int row, col;
  if (!tile_pixel(cx, row, col)) return;
  const VPtr v = cx.v;
  const int W = v->W, H = v->H;
  const int plane = W * H;
  const int p = row * W + col;
  constexpr float kx[3][3] = {{-3.0f, 0.0f, 3.0f}, {-10.0f, 0.0f, 10.0f}, {-3.0f, 0.0f, 3.0f}};
  constexpr float ky[3][3] = {{-3.0f, -10.0f, -3.0f}, {0.0f, 0.0f, 0.0f}, {3.0f, 10.0f, 3.0f}};
  float gx = 0.0f, gy = 0.0f;
#pragma unroll
  for (int dr = -1; dr <= 1; ++dr) {
#pragma unroll
    for (int dc = -1; dc <= 1; ++dc) {
      const int u = col + dc, w = row + dr;
      const float color = (u >= 0 && u < W && w >= 0 && w < H) ? (float)G(v->ref)[w * W + u] / 255.0f : 0.0f;
      gx = gx + color * kx[dr + 1][dc + 1];
      gy = gy + color * ky[dr + 1][dc + 1];
    }
  }
  const float grad = sqrtf(gx * gx + gy * gy);

  const float n0 = G(v->normal)[p], n1 = G(v->normal)[plane + p], n2 = G(v->normal)[2 * plane + p];
  float a[24];
#pragma unroll
  for (int k = 0; k < 24; ++k) a[k] = 0.0f;
  int cnt = 0;
#pragma unroll
  for (int dy = -2; dy <= 2; ++dy) {
#pragma unroll
    for (int dx = -2; dx <= 2; ++dx) {
      if (dy == 0 && dx == 0) continue;
      const int y = row + dy, x = col + dx;
      if (y < 0 || y >= H || x < 0 || x >= W) continue;
      const int q = y * W + x;
      const float d = n0 * G(v->normal)[q] + n1 * G(v->normal)[plane + q] + n2 * G(v->normal)[2 * plane + q];
      const float ang = acos_or_nan(d);
#pragma unroll
      for (int j = 0; j < 24; ++j) a[j] = j == cnt ? ang : a[j];
      ++cnt;
    }
  }
  bubble_sort_regs(a, cnt);
  const float med = pick_reg(a, cnt >> 1);
  G(v->curv)[p] = sat_u16(fminf(grad * 10000.0f, 32767.0f));
  G(v->curv)[plane + p] = cnt > 0 ? sat_u16(med * 10000.0f) : (uint16_t)31415;  // M_PI * 10000
}

__global__ __launch_bounds__(kSweepThreads) void k_curvature(const LaunchTable* __restrict__ lt) {
  for_each_tile(lt, [&](int t) { curvature_tile(resolve(lt, t)); });
}

// MaskDepthMapKernel (kernel_functions.cu:3080-3093): zero the depth where
// the mask is 0.  The reference leaves the last row and column as they are.
__device__ __forceinline__ void mask_tile(const Ctx& cx) {
  int row, col;
  if (!tile_pixel(cx, row, col)) return;
  const VPtr v = cx.v;
  if (row >= v->H - 1 || col >= v->W - 1) return;
  const int p = row * v->W + col;
  if (G(v->mask)[p] == 0) G(v->depth)[p] = 0.0f;
}

__global__ __launch_bounds__(kSweepThreads) void k_mask(const LaunchTable* __restrict__ lt) {
  for_each_tile(lt, [&](int t) { mask_tile(resolve(lt, t)); });
}

// ComputeGeomGradientMap (kernel_functions.cu:3095-3120): Scharr gradient of
// |d_q - d_p| / d_p over the in-image 3x3 neighbours (outside ones are
// skipped, not zero).  The reference leaves pixels with d_p <= 0 unwritten
// (uninitialised memory); here they read 0.
__device__ __forceinline__ void geom_grad_tile(const Ctx& cx) {
  int row, col;
  if (!tile_pixel(cx, row, col)) return;
  const VPtr v = cx.v;
  const int W = v->W, H = v->H;
  const int p = row * W + col;
  const float depth = G(v->depth)[p];
  if (depth <= 0.0f) {
    G(v->geom_grad)[p] = 0.0f;
    return;
  }
  constexpr float kx[3][3] = {{-3.0f, 0.0f, 3.0f}, {-10.0f, 0.0f, 10.0f}, {-3.0f, 0.0f, 3.0f}};
  constexpr float ky[3][3] = {{-3.0f, -10.0f, -3.0f}, {0.0f, 0.0f, 0.0f}, {3.0f, 10.0f, 3.0f}};
  float gx = 0.0f, gy = 0.0f;
#pragma unroll
  for (int dr = -1; dr <= 1; ++dr) {
#pragma unroll
    for (int dc = -1; dc <= 1; ++dc) {
      const int w = row + dr, u = col + dc;
      if (w < 0 || w >= H || u < 0 || u >= W) continue;
      const float diff = fabsf(G(v->depth)[w * W + u] - depth) / depth;
      gx = gx + diff * kx[dr + 1][dc + 1];
      gy = gy + diff * ky[dr + 1][dc + 1];
    }
  }
  G(v->geom_grad)[p] = sqrtf(gx * gx + gy * gy);
}

__global__ __launch_bounds__(kSweepThreads) void k_geom_grad(const LaunchTable* __restrict__ lt) {
  for_each_tile(lt, [&](int t) { geom_grad_tile(resolve(lt, t)); });
}

// MedianFilter (kernel_functions.cu:3122-3244) for the pixels of one colour
// (BlackPixelFilter: (row + col) even, then RedPixelFilter: odd).  Taps:
// the pixel, the cross at distances 1 and 3 (up, down, left, right), then
// the 8 knight moves, in the reference's order; a neighbour counts when the
// pixel's gradient is below 1.4 x depth_diff_threshold or its relative depth
// difference is below depth_diff_threshold.  Depth and each normal
// component take their own median (element count / 2 after the sort); the
// normal is then normalised.  Every tap of a colour-c pixel is of colour
// 1 - c, so a pass reads nothing it writes.
__device__ __forceinline__ void median_tile(const Ctx& cx) {
  int row, col;
  if (!tile_pixel(cx, row, col)) return;
  const VPtr v = cx.v;
  const int W = v->W, H = v->H;
  const int plane = W * H;
  const int p = row * W + col;
  const float thr = v->depth_diff_threshold;
  const float gthr = 1.4f * thr;
  const float grad = G(v->geom_grad)[p];
  const float cd = G(v->depth)[p];
  const float inv = 1.0f / cd;
  constexpr int kTaps = 17;
  constexpr int offs[kTaps - 1][2] = {{-1, 0}, {1, 0},  {0, -1}, {0, 1},  {-3, 0}, {3, 0},   {0, -3},  {0, 3},
                                      {-2, 1}, {-1, 2}, {1, 2},  {2, 1},  {2, -1}, {1, -2}, {-1, -2}, {-2, -1}};
  float dd[kTaps], nx[kTaps], ny[kTaps], nz[kTaps];
#pragma unroll
  for (int k = 0; k < kTaps; ++k) dd[k] = nx[k] = ny[k] = nz[k] = 0.0f;
  dd[0] = cd;
  nx[0] = G(v->normal)[p];
  ny[0] = G(v->normal)[plane + p];
  nz[0] = G(v->normal)[2 * plane + p];
  int cnt = 1;
#pragma unroll
  for (int k = 0; k < kTaps - 1; ++k) {
    const int w = row + offs[k][0], u = col + offs[k][1];
    if (w < 0 || w >= H || u < 0 || u >= W) continue;
    const int q = w * W + u;
    const float d = G(v->depth)[q];
    const float diff = fabsf(d - cd) * inv;
    if (grad < gthr || (grad >= gthr && diff < thr)) {
      const float a = G(v->normal)[q], b = G(v->normal)[plane + q], c = G(v->normal)[2 * plane + q];
#pragma unroll
      for (int j = 1; j < kTaps; ++j) {
        const bool at = j == cnt;
        dd[j] = at ? d : dd[j];
        nx[j] = at ? a : nx[j];
        ny[j] = at ? b : ny[j];
        nz[j] = at ? c : nz[j];
      }
      ++cnt;
    }
  }
  bubble_sort_regs(dd, cnt);
  bubble_sort_regs(nx, cnt);
  bubble_sort_regs(ny, cnt);
  bubble_sort_regs(nz, cnt);
  const int nth = cnt >> 1;
  const float mx = pick_reg(nx, nth), my = pick_reg(ny, nth), mz = pick_reg(nz, nth);
  const float inv_norm = 1.0f / sqrtf(mx * mx + my * my + mz * mz);
  G(v->depth)[p] = pick_reg(dd, nth);
  G(v->normal)[p] = mx * inv_norm;
  G(v->normal)[plane + p] = my * inv_norm;
  G(v->normal)[2 * plane + p] = mz * inv_norm;
}

__global__ __launch_bounds__(kSweepThreads) void k_median(const LaunchTable* __restrict__ lt) {
  for_each_tile(lt, [&](int t) { median_tile(resolve(lt, t)); });
}

// RunFilter (patch_match_cuda.cu:1074-1108): FilterDepthMapsKernel
// (kernel_functions.cu:3356-3450) or, with dedup, DeduplicDepthMapsKernel
// (:3452-3490).  The geometric-consistency cost is the sweep's geometric
// term (same bilinear source depth and back-projection).  One thread per
// pixel; each pixel reads and writes only its own depth.
__global__ __launch_bounds__(256) void k_run_filter(const ViewDev* vd, int dedup, float filter_max_cost,
                                                    int min_consistent) {
  const VPtr v = (VPtr)vd;
  const int col = blockIdx.x * 16 + (int)(threadIdx.x & 15u);
  const int row = blockIdx.y * 16 + (int)(threadIdx.x >> 4);
  if (col >= v->W || row >= v->H) return;
  const int p = row * v->W + col;
  __attribute__((address_space(1))) float* depth = G(v->depth);
  const float d = depth[p];
  if ((double)d < 1e-6) {
    depth[p] = 0.0f;
    return;
  }
  Pix px;
  px.row = row;
  px.col = col;
  px.np0 = v->refInvK[0] * (float)col + v->refInvK[1];
  px.np1 = v->refInvK[2] * (float)row + v->refInvK[3];
  Counters cnt;
  int good = 0;
  for (int i = 0; i < v->N; ++i) {
    if (geom_eval(v, i, px, d, v->refK, cnt) <= filter_max_cost) {
      if (dedup) {
        depth[p] = 0.0f;
        return;
      }
      good++;
    }
  }
  if (!dedup && good < min_consistent) depth[p] = 0.0f;
}

__global__ void k_zero_u32(uint32_t* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------------------
// preprocessing

__global__ void k_quads_u8(const uint8_t* __restrict__ img, int W, int H, int stride, uint32_t* __restrict__ out) {
  const int pw = W + 2 * kQuadPad, ph = H + 2 * kQuadPad;
  const int px = blockIdx.x * blockDim.x + threadIdx.x;
  const int py = blockIdx.y * blockDim.y + threadIdx.y;
  if (px >= pw || py >= ph) return;
  const int x = px - kQuadPad, y = py - kQuadPad;
  uint32_t q = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int xx = x + (k & 1), yy = y + (k >> 1);
    const uint32_t t = (xx < 0 || xx >= W || yy < 0 || yy >= H) ? 0u : (uint32_t)img[yy * stride + xx];
    q |= t << (8 * k);
  }
  out[py * pw + px] = q;
}

__global__ void k_quads_f16(const uint8_t* __restrict__ img, int W, int H, int stride, uint2* __restrict__ out) {
  const int pw = W + 2 * kQuadPad, ph = H + 2 * kQuadPad;
  const int px = blockIdx.x * blockDim.x + threadIdx.x;
  const int py = blockIdx.y * blockDim.y + threadIdx.y;
  if (px >= pw || py >= ph) return;
  const int x = px - kQuadPad, y = py - kQuadPad;
  float t[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int xx = x + (k & 1), yy = y + (k >> 1);
    t[k] = (xx < 0 || xx >= W || yy < 0 || yy >= H) ? 0.0f : (float)img[yy * stride + xx];
  }
  // (t00, t01 - t00, t10, t11 - t10): integers of magnitude <= 255, exact in
  // f16, so each horizontal lerp is one mixed-precision FMA on the halves
  const unsigned short h0 = __half_as_ushort(__float2half(t[0]));
  const unsigned short h1 = __half_as_ushort(__float2half(t[1] - t[0]));
  const unsigned short h2 = __half_as_ushort(__float2half(t[2]));
  const unsigned short h3 = __half_as_ushort(__float2half(t[3] - t[2]));
  out[py * pw + px] = make_uint2((uint32_t)h0 | ((uint32_t)h1 << 16), (uint32_t)h2 | ((uint32_t)h3 << 16));
}

__global__ void k_depth_quads(const float* __restrict__ d, int W, int H, float4* __restrict__ out) {
  const int pw = W + 2 * kQuadPad, ph = H + 2 * kQuadPad;
  const int px = blockIdx.x * blockDim.x + threadIdx.x;
  const int py = blockIdx.y * blockDim.y + threadIdx.y;
  if (px >= pw || py >= ph) return;
  const int x = px - kQuadPad, y = py - kQuadPad;
  float t[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int xx = x + (k & 1), yy = y + (k >> 1);
    t[k] = (xx < 0 || xx >= W || yy < 0 || yy >= H) ? 0.0f : d[yy * W + xx];
  }
  out[py * pw + px] = make_float4(t[0], t[1], t[2], t[3]);
}

// FilterKernel + ComputeGradientMap
__global__ void k_ref_moments(const ViewDev* __restrict__ v, int accum_mode, float4* __restrict__ refm,
                              uint8_t* __restrict__ grad_ok) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = blockIdx.y * blockDim.y + threadIdx.y;
  const int W = v->W, H = v->H;
  if (col >= W || row >= H) return;
  const int R = v->radius, S = v->step;
  const uint8_t* ref = v->ref;
  auto px_at = [&](int r, int c) -> int { return (r < 0 || r >= H || c < 0 || c >= W) ? 0 : (int)ref[r * W + c]; };
  const int c = px_at(row, col);
  float sw = 0.0f, sr = 0.0f, srr = 0.0f;
  int k = 0;
  for (int dr = -R; dr <= R; dr += S) {
    for (int dc = -R; dc <= R; dc += S) {
      const int ri = px_at(row + dr, col + dc);
      const float r = (float)ri;
      const float w = G(v->sp)[k++] * G(v->ct)[abs(c - ri)];
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
  if (accum_mode == PM_ACCUM_COMPENSATED) {
    float s1 = 0.0f, s2 = 0.0f;
    k = 0;
    for (int dr = -R; dr <= R; dr += S) {
      for (int dc = -R; dc <= R; dc += S) {
        const int ri = px_at(row + dr, col + dc);
        const float d = (float)ri - mu;
        const float w = G(v->sp)[k++] * G(v->ct)[abs(c - ri)];
        const float wd = w * d;
        s1 = s1 + wd;
        s2 = fmaf(wd, d, s2);
      }
    }
    err = s1 / sw;
    var = s2 / sw - err * err;
  }
  refm[row * W + col] = make_float4(mu, var, 1.0f / sw, err);
  const float kx[3][3] = {{-3, 0, 3}, {-10, 0, 10}, {-3, 0, 3}};
  const float ky[3][3] = {{-3, -10, -3}, {0, 0, 0}, {3, 10, 3}};
  float gx = 0.0f, gy = 0.0f;
#pragma unroll
  for (int dr = -1; dr <= 1; ++dr) {
#pragma unroll
    for (int dc = -1; dc <= 1; ++dc) {
      const float color = (float)px_at(row + dr, col + dc) / 255.0f;
      gx = gx + color * kx[dr + 1][dc + 1];
      gy = gy + color * ky[dr + 1][dc + 1];
    }
  }
  const float grad = sqrtf(gx * gx + gy * gy);
  grad_ok[row * W + col] = grad >= 1.0f ? 1 : 0;
}

// ---------------------------------------------------------------------------
// dispatch over the compiled (radius, step) shapes

template <int R, int S, int ST, int RC>
struct SweepK {
  static void* fn() { return (void*)&k_sweep<R, S, ST, RC>; }
};
template <int R, int S, int ST, int RC>
struct InitK {
  static void* fn() { return (void*)&k_init<R, S, ST, RC>; }
};
template <int R, int S, int ST, int RC>
struct LocalK {
  static void* fn() { return (void*)&k_local_opt<R, S, ST, RC>; }
};

template <template <int, int, int, int> class Kn, int R, int S>
void* pick_sr(int storage, int rcp) {
  if (storage == PM_STORAGE_U8) {
    return rcp == PM_RCP_EXACT ? Kn<R, S, PM_STORAGE_U8, PM_RCP_EXACT>::fn() : Kn<R, S, PM_STORAGE_U8, PM_RCP_FAST>::fn();
  }
  return rcp == PM_RCP_EXACT ? Kn<R, S, PM_STORAGE_F16, PM_RCP_EXACT>::fn() : Kn<R, S, PM_STORAGE_F16, PM_RCP_FAST>::fn();
}

template <template <int, int, int, int> class Kn>
void* pick_kernel(const LaunchCfg& cfg) {
  const int r = cfg.radius, st = cfg.step;
  if (r == 5 && st == 2) return pick_sr<Kn, 5, 2>(cfg.storage, cfg.rcp);
  if (r == 7 && st == 1) return pick_sr<Kn, 7, 1>(cfg.storage, cfg.rcp);
  if (r == 5 && st == 1) return pick_sr<Kn, 5, 1>(cfg.storage, cfg.rcp);
  if (r == 7 && st == 2) return pick_sr<Kn, 7, 2>(cfg.storage, cfg.rcp);
  return pick_sr<Kn, -1, -1>(cfg.storage, cfg.rcp);
}

}  // namespace


hipError_t launch_build_quads_u8(const uint8_t* img, int W, int H, int stride, uint32_t* out, hipStream_t s) {
  if (W <= 0 || H <= 0 || stride < W) return hipErrorInvalidValue;
  const dim3 blk(32, 8), grd((W + 2 * kQuadPad + 31) / 32, (H + 2 * kQuadPad + 7) / 8);
  hipLaunchKernelGGL(k_quads_u8, grd, blk, 0, s, img, W, H, stride, out);
  return hipGetLastError();
}

hipError_t launch_build_quads_f16(const uint8_t* img, int W, int H, int stride, uint2* out, hipStream_t s) {
  if (W <= 0 || H <= 0 || stride < W) return hipErrorInvalidValue;
  const dim3 blk(32, 8), grd((W + 2 * kQuadPad + 31) / 32, (H + 2 * kQuadPad + 7) / 8);
  hipLaunchKernelGGL(k_quads_f16, grd, blk, 0, s, img, W, H, stride, out);
  return hipGetLastError();
}

hipError_t launch_build_depth_quads(const float* depth, int W, int H, float4* out, hipStream_t s) {
  if (W <= 0 || H <= 0) return hipErrorInvalidValue;
  const dim3 blk(32, 8), grd((W + 2 * kQuadPad + 31) / 32, (H + 2 * kQuadPad + 7) / 8);
  hipLaunchKernelGGL(k_depth_quads, grd, blk, 0, s, depth, W, H, out);
  return hipGetLastError();
}

hipError_t launch_ref_moments(const ViewDev* v, int W, int H, int accum_mode, float4* refm, uint8_t* grad_ok,
                              hipStream_t s) {
  if (W <= 0 || H <= 0) return hipErrorInvalidValue;
  const dim3 blk(32, 8), grd((W + 31) / 32, (H + 7) / 8);
  hipLaunchKernelGGL(k_ref_moments, grd, blk, 0, s, v, accum_mode, refm, grad_ok);
  return hipGetLastError();
}

namespace {

void* phase_kernel(PhaseKind kind, const LaunchCfg& cfg) {
  switch (kind) {
    case kPhaseInit: return pick_kernel<InitK>(cfg);
    case kPhaseSweep: return pick_kernel<SweepK>(cfg);
    case kPhaseLocal: return pick_kernel<LocalK>(cfg);
    case kPhaseFilter: return (void*)&k_filter;
    case kPhaseCurvature: return (void*)&k_curvature;
    case kPhaseMask: return (void*)&k_mask;
    case kPhaseGeomGrad: return (void*)&k_geom_grad;
    case kPhaseMedian: return (void*)&k_median;
  }
  return nullptr;
}

}  // namespace

namespace {
std::mutex g_occ_mu;
std::map<void*, int> g_occ;
std::map<int, int> g_cus;
}  // namespace

int phase_occupancy(PhaseKind kind, const LaunchCfg& cfg) {
  void* fn = phase_kernel(kind, cfg);
  if (!fn) return 1;
  std::lock_guard<std::mutex> lk(g_occ_mu);
  auto it = g_occ.find(fn);
  if (it == g_occ.end()) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, kSweepThreads, 0) != hipSuccess || n <= 0) n = 1;
    it = g_occ.emplace(fn, n).first;
  }
  return it->second;
}

int phase_grid(PhaseKind kind, const LaunchCfg& cfg, int tiles, int blocks_per_cu) {
  if (tiles <= 0) return 0;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  int per_cu = phase_occupancy(kind, cfg), n_cu = 0;
  {
    std::lock_guard<std::mutex> lk(g_occ_mu);
    std::map<int, int>& cus = g_cus;
    auto ic = cus.find(dev);
    if (ic == cus.end()) {
      int n = 0;
      if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 1;
      ic = cus.emplace(dev, n).first;
    }
    n_cu = ic->second;
  }
  if (per_cu > kMaxBlocksPerCU) per_cu = kMaxBlocksPerCU;
  if (blocks_per_cu > 0 && blocks_per_cu < per_cu) per_cu = blocks_per_cu;
  const long long g = (long long)n_cu * per_cu;
  return (int)(g < tiles ? g : tiles);
}

hipError_t launch_phase(PhaseKind kind, const LaunchCfg& cfg, const LaunchTable* table, int grid, hipStream_t s) {
  if (grid <= 0 || !table) return hipErrorInvalidValue;
  void* fn = phase_kernel(kind, cfg);
  if (!fn) return hipErrorInvalidValue;
  void* args[] = {(void*)&table};
  return hipLaunchKernel(fn, dim3((unsigned)grid), dim3(kSweepThreads), args, 0, s);
}

hipError_t launch_run_filter(const ViewDev* v, int W, int H, int dedup, float filter_max_cost, int min_consistent,
                             hipStream_t s) {
  if (W <= 0 || H <= 0 || !v) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_run_filter, dim3((unsigned)((W + 15) / 16), (unsigned)((H + 15) / 16)), dim3(256), 0, s, v, dedup,
                     filter_max_cost, min_consistent);
  return hipGetLastError();
}

hipError_t launch_zero_u32(uint32_t* p, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (!p) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

}  // namespace pm
