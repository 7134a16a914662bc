// pm_io.hip -- the reference's on-disk map formats, so the engine's outputs
// drop into the rest of the MVS pipeline unchanged, and the normal-map
// encoding on the device (4 bytes per pixel over the link instead of 12).
//
//   Mat<T>::Write / Read (src/util/mat.h:178-215): text header "W&H&D&",
//   then width*height*depth values little-endian, slice-major.
//   NormalMap::Write / Read (src/mvs/normal_map.cc:121-178): the 3 planar
//   float slices become one uint32 per pixel, two 16-bit coordinates of the
//   unit normal.
#include "pm_host.h"

#include <thread>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace pmh;

namespace {

constexpr float kNormalFactor = 32767.0f;                 // decode_factor (normal_map.cc:8)
constexpr float kNormalInvFactor = 1.0f / kNormalFactor;  // decode_inv_factor (normal_map.cc:9)

// NormalMap::Encode for one pixel (normal_map.cc:160-174).  The reference
// mixes float and double, but every double step there is exact (a halving,
// or the product of two floats) and is rounded to float once, so these FP32
// operations give the same bits.  std::max / std::min are kept as written:
// a NaN coordinate clamps to the upper bound.
__host__ __device__ inline uint32_t encode_normal(float n0, float n1, float n2) {
  const float s = (1.0f - n0) * 0.5f;
  const float m = 1.0f / sqrtf(1.0f - s);
  float q1 = (n1 * 0.5f) * m;
  float q2 = (n2 * 0.5f) * m;
  q1 = (q1 < -0.999999f) ? -0.999999f : q1;  // std::max(q1, -0.999999f)
  q2 = (q2 < -0.999999f) ? -0.999999f : q2;
  q1 = (q1 < 0.999999f) ? q1 : 0.999999f;    // std::min(0.999999f, q1)
  q2 = (q2 < 0.999999f) ? q2 : 0.999999f;
  const uint32_t iq1 = (uint32_t)((q1 + 1.0f) * kNormalFactor);
  const uint32_t iq2 = (uint32_t)((q2 + 1.0f) * kNormalFactor);
  return (iq1 << 16) | iq2;
}

// NormalMap::Decode for one pixel (normal_map.cc:140-151).
inline void decode_normal(uint32_t code, float n[3]) {
  const float q1 = (float)(code >> 16) * kNormalInvFactor - 1.0f;
  const float q2 = (float)(code & 0xffffu) * kNormalInvFactor - 1.0f;
  const float a = q1 * q1, b = q2 * q2;
  const float s = a + b;
  const float r = std::sqrt(1.0f - s);
  n[0] = (float)(1.0 - 2.0 * (double)s);
  n[1] = (2.0f * q1) * r;
  n[2] = (2.0f * q2) * r;
}

__global__ void k_encode_normals(const float* __restrict__ normal, int64_t plane, uint32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= plane) return;
  out[i] = encode_normal(normal[i], normal[plane + i], normal[2 * plane + i]);
}

bool elem_size_ok(int32_t e) { return e == 1 || e == 2 || e == 4 || e == 8; }

pm_status check_dims(int32_t w, int32_t h, int32_t d) {
  if (w <= 0 || h <= 0 || d <= 0) return fail(PM_ERR_INVALID_ARG, "map dimensions must be positive");
  if ((int64_t)w * h * d > ((int64_t)1 << 32)) return fail(PM_ERR_INVALID_ARG, "map too large");
  return PM_OK;
}

// Parse "W&H&D&" (with the whitespace skipping of operator>>); returns the
// offset of the first data byte or -1.
long parse_header(FILE* f, int32_t* w, int32_t* h, int32_t* d) {
  long long v[3];
  for (int k = 0; k < 3; ++k) {
    char sep = 0;
    if (std::fscanf(f, " %lld", &v[k]) != 1) return -1;
    if (std::fscanf(f, "%c", &sep) != 1 || sep != '&') return -1;
    if (v[k] <= 0 || v[k] > INT32_MAX) return -1;
  }
  *w = (int32_t)v[0];
  *h = (int32_t)v[1];
  *d = (int32_t)v[2];
  return std::ftell(f);
}

}  // namespace

PM_API pm_status pm_mat_write(const char* path, int32_t width, int32_t height, int32_t depth, int32_t elem_size,
                              const void* data) {
  if (!path || !data) return fail(PM_ERR_INVALID_ARG, "null argument");
  if (!elem_size_ok(elem_size)) return fail(PM_ERR_INVALID_ARG, "elem_size must be 1, 2, 4 or 8");
  pm_status st = check_dims(width, height, depth);
  if (st != PM_OK) return st;
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(PM_ERR_INVALID_ARG, std::string("cannot write ") + path);
  const size_t n = (size_t)width * height * depth;
  // host byte order is little-endian on every target of this library
  const bool ok = std::fprintf(f, "%d&%d&%d&", width, height, depth) > 0 &&
                  std::fwrite(data, (size_t)elem_size, n, f) == n;
  const bool closed = std::fclose(f) == 0;
  if (!ok || !closed) return fail(PM_ERR_INVALID_ARG, std::string("short write to ") + path);
  return PM_OK;
}

PM_API pm_status pm_mat_read_header(const char* path, int32_t* width, int32_t* height, int32_t* depth) {
  if (!path || !width || !height || !depth) return fail(PM_ERR_INVALID_ARG, "null argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(PM_ERR_INVALID_ARG, std::string("cannot read ") + path);
  const long off = parse_header(f, width, height, depth);
  std::fclose(f);
  if (off < 0) return fail(PM_ERR_INVALID_ARG, std::string("malformed header in ") + path);
  return PM_OK;
}

PM_API pm_status pm_mat_read(const char* path, int32_t width, int32_t height, int32_t depth, int32_t elem_size,
                             void* data) {
  if (!path || !data) return fail(PM_ERR_INVALID_ARG, "null argument");
  if (!elem_size_ok(elem_size)) return fail(PM_ERR_INVALID_ARG, "elem_size must be 1, 2, 4 or 8");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(PM_ERR_INVALID_ARG, std::string("cannot read ") + path);
  int32_t w = 0, h = 0, d = 0;
  const long off = parse_header(f, &w, &h, &d);
  if (off < 0 || w != width || h != height || d != depth) {
    std::fclose(f);
    return fail(PM_ERR_INVALID_ARG, std::string("header of ") + path + " does not match the requested size");
  }
  const size_t n = (size_t)w * h * d;
  const bool ok = std::fread(data, (size_t)elem_size, n, f) == n;
  std::fclose(f);
  if (!ok) return fail(PM_ERR_INVALID_ARG, std::string("file too short: ") + path);
  return PM_OK;
}

namespace {
// f(begin, end) over [0, n) in contiguous ranges on up to 16 threads (large n)
template <class F>
void parallel_ranges(int64_t n, F f) {
  const int64_t min_per = int64_t(1) << 16;
  const int64_t hw = std::max<int64_t>(1, (int64_t)std::thread::hardware_concurrency());
  const int64_t parts = std::max<int64_t>(1, std::min<int64_t>({16, hw, n / min_per}));
  if (parts == 1) {
    f(0, n);
    return;
  }
  const int64_t per = (n + parts - 1) / parts;
  std::vector<std::thread> th;
  for (int64_t p = 1; p < parts; ++p) th.emplace_back(f, p * per, std::min(n, (p + 1) * per));
  f(0, std::min(n, per));
  for (std::thread& t : th) t.join();
}
}  // namespace

PM_API pm_status pm_normal_encode(int64_t pixels, const float* normal, uint32_t* codes) {
  if (pixels <= 0 || !normal || !codes) return fail(PM_ERR_INVALID_ARG, "invalid arguments");
  parallel_ranges(pixels, [=](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) codes[i] = encode_normal(normal[i], normal[pixels + i], normal[2 * pixels + i]);
  });
  return PM_OK;
}

PM_API pm_status pm_normal_decode(int64_t pixels, const uint32_t* codes, float* normal) {
  if (pixels <= 0 || !normal || !codes) return fail(PM_ERR_INVALID_ARG, "invalid arguments");
  parallel_ranges(pixels, [=](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      float n[3];
      decode_normal(codes[i], n);
      normal[i] = n[0];
      normal[pixels + i] = n[1];
      normal[2 * pixels + i] = n[2];
    }
  });
  return PM_OK;
}

namespace pmh {
hipError_t encode_normals_async(const float* normal, int64_t plane, uint32_t* out, hipStream_t s) {
  if (plane <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_encode_normals, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, s, normal, plane, out);
  return hipGetLastError();
}
}  // namespace pmh

PM_API pm_status pm_problem_get_normal_codes(const pm_problem* p, uint32_t* out) {
  if (!p || !out) return fail(PM_ERR_INVALID_ARG, "null argument");
  PM_HIP(hipSetDevice(p->engine->device));
  hipStream_t s = p->engine->stream;
  const int64_t plane = (int64_t)p->W * p->H;
  uint32_t* d = nullptr;
  PM_HIP(hipMalloc(&d, (size_t)plane * sizeof(uint32_t)));
  hipLaunchKernelGGL(k_encode_normals, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, s, p->host.normal, plane,
                     d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, (size_t)plane * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(PM_ERR_DEVICE, std::string("normal encoding: ") + hipGetErrorString(e));
  return PM_OK;
}

PM_API pm_status pm_problem_write_maps(const pm_problem* p, const char* depth_path, const char* normal_path) {
  if (!p) return fail(PM_ERR_INVALID_ARG, "null problem");
  const int64_t plane = (int64_t)p->W * p->H;
  pm_status st = PM_OK;
  if (depth_path) {
    std::vector<float> depth((size_t)plane);
    if ((st = pm_problem_get_depth_map(p, depth.data())) != PM_OK) return st;
    if ((st = pm_mat_write(depth_path, p->W, p->H, 1, 4, depth.data())) != PM_OK) return st;
  }
  if (normal_path) {
    std::vector<uint32_t> codes((size_t)plane);
    if ((st = pm_problem_get_normal_codes(p, codes.data())) != PM_OK) return st;
    if ((st = pm_mat_write(normal_path, p->W, p->H, 1, 4, codes.data())) != PM_OK) return st;
  }
  return PM_OK;
}
