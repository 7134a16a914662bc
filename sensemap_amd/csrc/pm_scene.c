/* pm_scene.c -- seeded synthetic multi-view scenes with exact ground truth.
 *
 * Renders a triangle soup into a pinhole camera with a z-buffer.  For every
 * pixel centre (pixel i at coordinate i, the convention of the matching
 * kernels) the ray is intersected with each covering triangle in FP64
 * (Moller-Trumbore), which gives the exact z-depth and the face normal in the
 * camera frame.  Intensities come from seeded 3D value noise evaluated at the
 * world-space hit point, so every view of a surface point sees the same
 * texture; optional Gaussian sensor noise is added before rounding to u8.
 * Pixels that see no surface get intensity 0 and depth 0.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

static double lattice(int64_t x, int64_t y, int64_t z, uint64_t seed) {
  uint64_t h = splitmix64(seed ^ (uint64_t)x);
  h = splitmix64(h ^ (uint64_t)y);
  h = splitmix64(h ^ (uint64_t)z);
  return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

static double fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }

/* Value noise in [-1, 1] with quintic interpolation between lattice values. */
static double value_noise(double x, double y, double z, uint64_t seed) {
  const double fx = floor(x), fy = floor(y), fz = floor(z);
  const int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
  const double tx = fade(x - fx), ty = fade(y - fy), tz = fade(z - fz);
  double c[8];
  for (int q = 0; q < 8; ++q) c[q] = lattice(ix + (q & 1), iy + ((q >> 1) & 1), iz + (q >> 2), seed);
  const double x00 = c[0] + tx * (c[1] - c[0]);
  const double x10 = c[2] + tx * (c[3] - c[2]);
  const double x01 = c[4] + tx * (c[5] - c[4]);
  const double x11 = c[6] + tx * (c[7] - c[6]);
  const double y0 = x00 + ty * (x10 - x00);
  const double y1 = x01 + ty * (x11 - x01);
  return y0 + tz * (y1 - y0);
}

/* Two octaves, normalised to [-1, 1]. */
double scn_texture(double x, double y, double z, double spacing, uint64_t seed) {
  const double a = value_noise(x / spacing, y / spacing, z / spacing, seed);
  const double b = value_noise(2.0 * x / spacing, 2.0 * y / spacing, 2.0 * z / spacing, seed ^ 0x5bd1e995ull);
  return (a + 0.5 * b) / 1.5;
}

static double gauss(uint64_t seed, uint64_t idx) {
  const uint64_t h1 = splitmix64(seed ^ (idx * 2 + 1));
  const uint64_t h2 = splitmix64(h1 ^ 0xD1B54A32D192ED03ull);
  const double u1 = ((double)(h1 >> 11) + 1.0) / 9007199254740993.0;
  const double u2 = (double)(h2 >> 11) / 9007199254740992.0;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

typedef struct scn_params {
  double K[4];          /* fx, cx, fy, cy */
  double R[9];          /* world -> camera, row-major */
  double T[3];
  int32_t width, height;
  double tex_spacing;   /* world units per noise lattice cell */
  double tex_amplitude; /* grey levels */
  double tex_offset;    /* mean grey level */
  uint64_t tex_seed;
  double noise_sigma;   /* grey levels, 0 = none */
  uint64_t noise_seed;
} scn_params;

/* tris: ntris * 9 doubles (three world-space vertices).  Outputs (any may be
 * NULL): image u8 W*H, depth f32 W*H (0 where nothing is hit), normal f32
 * 3*W*H planar in the camera frame facing the camera, tri_id int32 W*H. */
int scn_render(const double* tris, int32_t ntris, const scn_params* p, uint8_t* image, float* depth,
               float* normal, int32_t* tri_id) {
  const int W = p->width, H = p->height;
  const double fx = p->K[0], cx = p->K[1], fy = p->K[2], cy = p->K[3];
  double* cam = (double*)malloc(sizeof(double) * 9 * (size_t)ntris);
  int* bb = (int*)malloc(sizeof(int) * 4 * (size_t)ntris);
  const double znear = 1e-3;
  for (int t = 0; t < ntris; ++t) {
    double* C3 = cam + 9 * (size_t)t;
    for (int k = 0; k < 3; ++k) {
      const double* X = tris + 9 * (size_t)t + 3 * k;
      double* Cv = C3 + 3 * k;
      for (int i = 0; i < 3; ++i) Cv[i] = p->R[3 * i] * X[0] + p->R[3 * i + 1] * X[1] + p->R[3 * i + 2] * X[2] + p->T[i];
    }
    /* clip the triangle against z >= znear (Sutherland-Hodgman) for its
     * screen bounding box; the per-pixel test uses the exact triangle */
    double poly[4][3];
    int np = 0;
    for (int k = 0; k < 3; ++k) {
      const double* a = C3 + 3 * k;
      const double* b = C3 + 3 * ((k + 1) % 3);
      const int ain = a[2] >= znear, bin = b[2] >= znear;
      if (ain) {
        poly[np][0] = a[0];
        poly[np][1] = a[1];
        poly[np][2] = a[2];
        np++;
      }
      if (ain != bin && np < 4) {
        const double s = (znear - a[2]) / (b[2] - a[2]);
        poly[np][0] = a[0] + s * (b[0] - a[0]);
        poly[np][1] = a[1] + s * (b[1] - a[1]);
        poly[np][2] = znear;
        np++;
      }
    }
    double umin = 1e300, umax = -1e300, vmin = 1e300, vmax = -1e300;
    for (int k = 0; k < np; ++k) {
      const double u = fx * poly[k][0] / poly[k][2] + cx, v = fy * poly[k][1] / poly[k][2] + cy;
      if (u < umin) umin = u;
      if (u > umax) umax = u;
      if (v < vmin) vmin = v;
      if (v > vmax) vmax = v;
    }
    int* b = bb + 4 * (size_t)t;
    if (np == 0 || umax < -1 || vmax < -1 || umin > W || vmin > H) {
      b[0] = 1;
      b[1] = 0;
      b[2] = 1;
      b[3] = 0;
      continue;
    }
    b[0] = umin < 0 ? 0 : (int)ceil(umin);
    b[1] = umax > W - 1 ? W - 1 : (int)floor(umax);
    b[2] = vmin < 0 ? 0 : (int)ceil(vmin);
    b[3] = vmax > H - 1 ? H - 1 : (int)floor(vmax);
  }
  const double eps = 1e-9;
  long row;
#pragma omp parallel for schedule(dynamic, 4)
  for (row = 0; row < H; ++row) {
    double* zb = (double*)malloc(sizeof(double) * W);
    int* id = (int*)malloc(sizeof(int) * W);
    for (int c = 0; c < W; ++c) {
      zb[c] = 1e300;
      id[c] = -1;
    }
    const double dy = ((double)row - cy) / fy;
    for (int t = 0; t < ntris; ++t) {
      const int* b = bb + 4 * (size_t)t;
      if (row < b[2] || row > b[3] || b[0] > b[1]) continue;
      const double* A = cam + 9 * (size_t)t;
      const double* B = A + 3;
      const double* C = A + 6;
      const double e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
      const double e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
      const double tv[3] = {-A[0], -A[1], -A[2]};
      const double q[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2], tv[0] * e1[1] - tv[1] * e1[0]};
      const double tq = e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2];
      for (int col = b[0]; col <= b[1]; ++col) {
        const double d[3] = {((double)col - cx) / fx, dy, 1.0};
        const double pv[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
        const double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
        if (fabs(det) < 1e-14) continue;
        const double inv = 1.0 / det;
        const double uu = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * inv;
        if (uu < -eps || uu > 1.0 + eps) continue;
        const double vv = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
        if (vv < -eps || uu + vv > 1.0 + eps) continue;
        const double z = tq * inv;
        if (z > 1e-6 && z < zb[col]) {
          zb[col] = z;
          id[col] = t;
        }
      }
    }
    for (int col = 0; col < W; ++col) {
      const size_t pidx = (size_t)row * W + col;
      const int t = id[col];
      if (tri_id) tri_id[pidx] = t;
      if (t < 0) {
        if (image) image[pidx] = 0;
        if (depth) depth[pidx] = 0.0f;
        if (normal) {
          normal[pidx] = 0.0f;
          normal[(size_t)W * H + pidx] = 0.0f;
          normal[2 * (size_t)W * H + pidx] = 0.0f;
        }
        continue;
      }
      const double z = zb[col];
      const double d[3] = {((double)col - cx) / fx, dy, 1.0};
      if (depth) depth[pidx] = (float)z;
      if (normal) {
        const double* A = cam + 9 * (size_t)t;
        const double e1[3] = {A[3] - A[0], A[4] - A[1], A[5] - A[2]};
        const double e2[3] = {A[6] - A[0], A[7] - A[1], A[8] - A[2]};
        double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        double s = 1.0 / len;
        if (n[0] * d[0] + n[1] * d[1] + n[2] * d[2] > 0) s = -s;
        normal[pidx] = (float)(n[0] * s);
        normal[(size_t)W * H + pidx] = (float)(n[1] * s);
        normal[2 * (size_t)W * H + pidx] = (float)(n[2] * s);
      }
      if (image) {
        /* world point X = R^T (z d - T) */
        const double pc[3] = {z * d[0] - p->T[0], z * d[1] - p->T[1], z * d[2] - p->T[2]};
        double X[3];
        for (int i = 0; i < 3; ++i) X[i] = p->R[i] * pc[0] + p->R[3 + i] * pc[1] + p->R[6 + i] * pc[2];
        double val = p->tex_offset + p->tex_amplitude * scn_texture(X[0], X[1], X[2], p->tex_spacing, p->tex_seed);
        if (p->noise_sigma > 0) val += p->noise_sigma * gauss(p->noise_seed, pidx);
        double r = floor(val + 0.5);
        if (r < 0) r = 0;
        if (r > 255) r = 255;
        image[pidx] = (uint8_t)r;
      }
    }
    free(zb);
    free(id);
  }
  free(cam);
  free(bb);
  return 0;
}
