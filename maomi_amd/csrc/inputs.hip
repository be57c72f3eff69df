// Input generation for the optical stage (reference maomi/optics.py:120-197, 384-415 and
// maomi/_optical_math.py:20-67): the spectral Gaussian-process index fields and the
// depth-dependent collection disk with hemoglobin absorption.
//
//   grf_spectrum_kernel   envelope(x,y,z) * (re + i im), written at its ifftshifted index
//   grf_finish_kernel     scale * Re(g at the ifftshifted index) + mu, the partial first/last
//                         slab thickness and the vessel occupancy mix, in one pass
//   zsum_kernel           sums of z runs (vessel occupancy per slab)
//   slab_prefix_kernel    per row: slab projection (sum over z) and its prefix sum along y
//   disk_accumulate_kernel  normalised-disk convolution ('same' placement) from the row prefix
//                         sums, one run of columns per disk row
//   hemo_mask_kernel      10 ** (-total / vres * hemoabs)
//
// The FFT between the two field kernels is the caller's (rocFFT through torch).  This file is
// built with -ffp-contract=off so that the envelope and the occupancy mix round exactly as the
// reference's float32 tensor expressions do.
#include "common.hpp"
#include "maomi_amd.h"

namespace {

inline unsigned grid_for(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return (unsigned)(b < 65535 ? (b > 0 ? b : 1) : 65535);
}

// Envelope: sqrt(sum_i prod_l[i] * (kx_i kx_i...) ), with per scale i the 1-D factors
// ker[i] = [kx (X) | ky (Y) | kz (Z)] = exp(-linspace(-pi/2, pi/2)^2 * l^2) and
// coef[i] = (prod(l_i), w_i^2).  The operation order is the reference's:
// ((prod_l * (kx*ky*kz)^2) * w^2), accumulated over scales from zero.
// Source voxel j goes to index (j + n - n/2) mod n per axis (ifftshift).
__global__ void grf_spectrum_kernel(const float* __restrict__ re, const float* __restrict__ im,
                                    int64_t X, int64_t Y, int64_t Z,
                                    const float* __restrict__ ker,
                                    const float* __restrict__ coef, int K,
                                    float2* __restrict__ spec) {
  const int64_t tot = X * Y * Z, per = X + Y + Z;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t x = e / (Y * Z), y = (e / Z) % Y, z = e % Z;
    float acc = 0.0f;
    for (int i = 0; i < K; ++i) {
      const float* k = ker + i * per;
      float t = k[x] * k[X + y];
      t = t * k[X + Y + z];
      const float t2 = t * t;
      float a = coef[2 * i] * t2;
      a = a * coef[2 * i + 1];
      acc = acc + a;
    }
    const float env = sqrtf(acc);
    const int64_t dx = (x + X - X / 2) % X, dy = (y + Y - Y / 2) % Y, dz = (z + Z - Z / 2) % Z;
    spec[(dx * Y + dy) * Z + dz] = make_float2(re[e] * env, im[e] * env);
  }
}

// out[j] = scale * Re(g[(j + n/2) mod n]) + mu (g == NULL: zero field), times first_z on the
// plane z = 0 and last_z on z = Z - 1; with occupancy, out = occ + out * (1 - occ / occ_div).
__global__ void grf_finish_kernel(const float2* __restrict__ g, int64_t X, int64_t Y,
                                  int64_t Z, float scale, float mu, float first_z,
                                  float last_z, const float* __restrict__ occ, float occ_div,
                                  float* __restrict__ out) {
  const int64_t tot = X * Y * Z;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t x = e / (Y * Z), y = (e / Z) % Y, z = e % Z;
    float v = 0.0f;
    if (g) {
      const int64_t sx = (x + X / 2) % X, sy = (y + Y / 2) % Y, sz = (z + Z / 2) % Z;
      v = scale * g[(sx * Y + sy) * Z + sz].x;
      v = v + mu;
    }
    if (z == 0) v = v * first_z;
    if (z == Z - 1) v = v * last_z;
    if (occ) {
      const float o = occ[e];
      const float f = 1.0f - o / occ_div;
      v = o + v * f;
    }
    out[e] = v;
  }
}

// out[(x * Y + y) * Zo + zo] = sum_{z < L} src[x][y][z], accumulated in FP64 and rounded once.
__global__ void zsum_kernel(const float* __restrict__ src, int64_t X, int64_t Y, int64_t L,
                            float* __restrict__ out, int64_t Zo, int64_t zo) {
  const int64_t tot = X * Y;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const float* p = src + e * L;
    double s = 0.0;
    for (int64_t z = 0; z < L; ++z) s += p[z];
    out[e * Zo + zo] = (float)s;
  }
}

// One block per row x: v(y) = float32(sum_{z < L} slab[x][y][z]) and
// pre[x][m] = sum_{y < m} v(y) for m = 0..Y (FP64).  Each thread owns a contiguous run of y.
__global__ void __launch_bounds__(256) slab_prefix_kernel(const float* __restrict__ slab,
                                                          int64_t Y, int64_t L,
                                                          double* __restrict__ pre) {
  __shared__ double part[256];
  const int64_t x = blockIdx.x;
  const int t = threadIdx.x;
  const int64_t chunk = (Y + 255) / 256;
  const int64_t y0 = min<int64_t>(Y, t * chunk), y1 = min<int64_t>(Y, y0 + chunk);
  const float* row = slab + x * Y * L;
  double* prow = pre + x * (Y + 1);
  double local = 0.0;
  for (int64_t y = y0; y < y1; ++y) {
    double s = 0.0;
    for (int64_t z = 0; z < L; ++z) s += row[y * L + z];
    const double v = (double)(float)s;
    prow[y + 1] = v;
    local += v;
  }
  part[t] = local;
  __syncthreads();
  // inclusive scan of the 256 run totals (Hillis-Steele)
  for (int off = 1; off < 256; off <<= 1) {
    const double add = t >= off ? part[t - off] : 0.0;
    __syncthreads();
    part[t] += add;
    __syncthreads();
  }
  double run = t ? part[t - 1] : 0.0;
  if (t == 0) prow[0] = 0.0;
  for (int64_t y = y0; y < y1; ++y) {
    run += prow[y + 1];
    prow[y + 1] = run;
  }
}

// total[i][j] += w * sum_k sum_{l in run k} v[fi + c0 - k][fj + c1 - l] with (fi, fj) =
// (i + ox, j + oy), v zero outside X x Y; runs[2k], runs[2k+1] = first and last column of disk
// row k.  This is scipy.signal.fftconvolve(v, disk, 'same') restricted to the output window.
__global__ void disk_accumulate_kernel(const double* __restrict__ pre, int64_t X, int64_t Y,
                                       const int32_t* __restrict__ runs, int K0, int c0, int c1,
                                       double w, int64_t ox, int64_t oy, int64_t Xo, int64_t Yo,
                                       double* __restrict__ total) {
  const int64_t tot = Xo * Yo;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t fi = e / Yo + ox, fj = e % Yo + oy;
    double acc = 0.0;
    for (int k = 0; k < K0; ++k) {
      const int64_t r = fi + c0 - k;
      if (r < 0 || r >= X) continue;
      int64_t lo = fj + c1 - runs[2 * k + 1], hi = fj + c1 - runs[2 * k];
      lo = lo < 0 ? 0 : lo;
      hi = hi > Y - 1 ? Y - 1 : hi;
      if (lo <= hi) acc += pre[r * (Y + 1) + hi + 1] - pre[r * (Y + 1) + lo];
    }
    total[e] += w * acc;
  }
}

// out = 10 ** ((-float32(total) / vres) * hemoabs), float32 as the reference's NumPy expression
__global__ void hemo_mask_kernel(const double* __restrict__ total, int64_t n, float vres,
                                 float hemoabs, float* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const float t = (float)total[e];
    float a = -t / vres;
    a = a * hemoabs;
    out[e] = powf(10.0f, a);
  }
}

}  // namespace

MAOMI_API int maomi_grf_spectrum(const float* noise_re, const float* noise_im, int64_t X,
                                 int64_t Y, int64_t Z, const float* kernels, const float* coef,
                                 int32_t K, void* spec, maomi_stream_t stream) {
  if (!noise_re || !noise_im || !kernels || !coef || !spec || X < 1 || Y < 1 || Z < 1 ||
      K < 1) {
    maomi_set_error("maomi_grf_spectrum: bad arguments");
    return MAOMI_EINVAL;
  }
  hipLaunchKernelGGL(grf_spectrum_kernel, dim3(grid_for(X * Y * Z)), dim3(256), 0,
                     (hipStream_t)stream, noise_re, noise_im, X, Y, Z, kernels, coef, (int)K,
                     (float2*)spec);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_grf_finish(const void* field, int64_t X, int64_t Y, int64_t Z, float scale,
                               float mu, float first_z, float last_z, const float* occupancy,
                               float occupancy_div, float* out, maomi_stream_t stream) {
  if (!out || X < 1 || Y < 1 || Z < 1 || (occupancy && !(occupancy_div > 0.0f))) {
    maomi_set_error("maomi_grf_finish: bad arguments");
    return MAOMI_EINVAL;
  }
  hipLaunchKernelGGL(grf_finish_kernel, dim3(grid_for(X * Y * Z)), dim3(256), 0,
                     (hipStream_t)stream, (const float2*)field, X, Y, Z, scale, mu, first_z,
                     last_z, occupancy, occupancy_div, out);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_zsum(const float* src, int64_t X, int64_t Y, int64_t L, float* out,
                         int64_t Zo, int64_t zo, maomi_stream_t stream) {
  if (!src || !out || X < 1 || Y < 1 || L < 1 || Zo < 1 || zo < 0 || zo >= Zo) {
    maomi_set_error("maomi_zsum: bad arguments");
    return MAOMI_EINVAL;
  }
  hipLaunchKernelGGL(zsum_kernel, dim3(grid_for(X * Y)), dim3(256), 0, (hipStream_t)stream,
                     src, X, Y, L, out, Zo, zo);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_slab_prefix(const float* slab, int64_t X, int64_t Y, int64_t L,
                                double* prefix, maomi_stream_t stream) {
  if (!slab || !prefix || X < 1 || X > 2147483647 || Y < 1 || L < 1) {
    maomi_set_error("maomi_slab_prefix: bad arguments");
    return MAOMI_EINVAL;
  }
  hipLaunchKernelGGL(slab_prefix_kernel, dim3((unsigned)X), dim3(256), 0, (hipStream_t)stream,
                     slab, Y, L, prefix);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_disk_accumulate(const double* prefix, int64_t X, int64_t Y,
                                    const int32_t* runs, int32_t K0, int32_t c0, int32_t c1,
                                    double weight, int64_t ox, int64_t oy, int64_t Xo,
                                    int64_t Yo, double* total, maomi_stream_t stream) {
  if (!prefix || !runs || !total || X < 1 || Y < 1 || K0 < 1 || c0 < 0 || c1 < 0 ||
      c0 >= K0 || ox < 0 || oy < 0 || Xo < 1 || Yo < 1 || ox + Xo > X || oy + Yo > Y) {
    maomi_set_error("maomi_disk_accumulate: bad arguments");
    return MAOMI_EINVAL;
  }
  hipLaunchKernelGGL(disk_accumulate_kernel, dim3(grid_for(Xo * Yo)), dim3(256), 0,
                     (hipStream_t)stream, prefix, X, Y, runs, (int)K0, (int)c0, (int)c1, weight,
                     ox, oy, Xo, Yo, total);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_hemo_mask(const double* total, int64_t n, float vres, float hemoabs,
                              float* out, maomi_stream_t stream) {
  if (!total || !out || n < 1 || !(vres > 0.0f)) {
    maomi_set_error("maomi_hemo_mask: bad arguments");
    return MAOMI_EINVAL;
  }
  hipLaunchKernelGGL(hemo_mask_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream,
                     total, n, vres, hemoabs, out);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}
