// Input generation for the optical stage (reference maomi/optics.py:120-197, 384-415 and
// maomi/_optical_math.py:20-67): the spectral Gaussian-process index fields and the
// depth-dependent collection disk with hemoglobin absorption.
//
//   grf_spectrum_kernel   envelope(x,y,z) * (re + i im), written at its ifftshifted index
//   grf_finish_kernel     scale * Re(g at the ifftshifted index) + mu, the partial first/last
//                         slab thickness and the vessel occupancy mix, in one pass
//   zsum_kernel           sums of z runs (vessel occupancy per slab)
//   slab_prefix_kernel    per collection slab and row: slab projection (sum over z) and its
//                         prefix sum along y
//   disk_kernel           normalised-disk convolution ('same' placement) from the row prefix
//                         sums, one run of columns per disk row, every slab in one launch
//   hemo_mask_kernel      slab sum and 10 ** (-total / vres * hemoabs)
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

// Envelope: sqrt(sum_i (prod_l[i] * (kx ky kz)^2) * w_i^2), with per scale i the 1-D factors
// ker[i] = [kx (X) | ky (Y) | kz (Z)] = exp(-linspace(-pi/2, pi/2)^2 * l^2) and
// coef[i] = (prod(l_i), w_i^2).  The operation order is the reference's:
// ((prod_l * (kx*ky*kz)^2) * w^2), accumulated over scales from zero.
// Source voxel j goes to index (j + n - n/2) mod n per axis (ifftshift).  I is the index type
// (32-bit when the volume has fewer than 2^31 voxels, which makes the index divisions cheap).
template <typename I>
__global__ void grf_spectrum_kernel(const float* __restrict__ re, const float* __restrict__ im,
                                    I X, I Y, I Z, const float* __restrict__ ker,
                                    const float* __restrict__ coef, int K,
                                    float2* __restrict__ spec) {
  const I tot = X * Y * Z, per = X + Y + Z;
  for (I e = blockIdx.x * (I)blockDim.x + threadIdx.x; e < tot; e += (I)gridDim.x * blockDim.x) {
    const I x = e / (Y * Z), y = (e / Z) % Y, z = e % Z;
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
    const I dx = (x + X - X / 2) % X, dy = (y + Y - Y / 2) % Y, dz = (z + Z - Z / 2) % Z;
    spec[(dx * Y + dy) * Z + dz] = make_float2(re[e] * env, im[e] * env);
  }
}

// out[j] = scale * Re(g[(j + n/2) mod n]) + mu (g == NULL: zero field), times first_z on the
// plane z = 0 and last_z on z = Z - 1; with occupancy, out = occ + out * (1 - occ / occ_div).
template <typename I>
__global__ void grf_finish_kernel(const float2* __restrict__ g, I X, I Y, I Z, float scale,
                                  float mu, float first_z, float last_z,
                                  const float* __restrict__ occ, float occ_div,
                                  float* __restrict__ out) {
  const I tot = X * Y * Z;
  for (I e = blockIdx.x * (I)blockDim.x + threadIdx.x; e < tot; e += (I)gridDim.x * blockDim.x) {
    const I x = e / (Y * Z), y = (e / Z) % Y, z = e % Z;
    float v = 0.0f;
    if (g) {
      const I sx = (x + X / 2) % X, sy = (y + Y / 2) % Y, sz = (z + Z / 2) % Z;
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

// For every run r = (z0, z1) = (runs[rs * r], runs[rs * r + 1]) (runs sorted, not overlapping,
// clamped to [0, Z)): out[r * so_run + (x * Y + y) * so_xy] = sum_{z0 <= z < z1}
// src[x * sx + y * sy + z], summed in FP64 and rounded once.  A block takes 256 consecutive
// (x, y) columns and stages ZC planes of them at a time in LDS with loads in which consecutive
// lanes read consecutive z of one column; each thread then walks its own column through the
// runs.  The source is read once whatever the runs.  T is float (occupancy fractions) or
// uint8_t (boolean occupancy, as stored).
template <typename T, int ZC, int PAD>
__global__ void __launch_bounds__(256) zsum_kernel(const T* __restrict__ src, int64_t X,
                                                   int64_t Y, int Z, int64_t sx, int64_t sy,
                                                   const int32_t* __restrict__ runs, int nruns,
                                                   int rs, float* __restrict__ out,
                                                   int64_t so_run, int64_t so_xy) {
  __shared__ T tile[256 * (ZC + PAD)];
  __shared__ int64_t colbase[256];
  const int t = threadIdx.x;
  const int64_t tot = X * Y;
  const int zmin = min(max(runs[0], 0), Z);
  const int zmax = max(zmin, min(runs[rs * (nruns - 1) + 1], Z));
  for (int64_t e0 = (int64_t)blockIdx.x * 256; e0 < tot; e0 += (int64_t)gridDim.x * 256) {
    const int64_t e = e0 + t;
    __syncthreads();  // the previous tile is consumed
    colbase[t] = e < tot ? (e / Y) * sx + (e % Y) * sy : -1;
    int r = 0;
    int rz0 = runs[0], rz1 = runs[1];
    double acc = 0.0;
    for (int zc = zmin; zc < zmax; zc += ZC) {
      const int len = min(ZC, zmax - zc);
      __syncthreads();  // colbase written / previous chunk consumed
      // ZC loads per thread, unrolled so that many are in flight at once
#pragma unroll 16
      for (int q = t; q < 256 * ZC; q += 256) {
        const int c = q / ZC, j = q % ZC;
        const int64_t base = colbase[c];
        tile[c * (ZC + PAD) + j] = (base >= 0 && j < len) ? src[base + zc + j] : (T)0;
      }
      __syncthreads();
      if (e < tot) {
        for (int j = 0; j < len; ++j) {
          const int z = zc + j;
          while (r < nruns && z >= rz1) {
            out[r * so_run + e * so_xy] = (float)acc;
            acc = 0.0;
            if (++r < nruns) {
              rz0 = runs[rs * r];
              rz1 = runs[rs * r + 1];
            }
          }
          if (r < nruns && z >= rz0) acc += (double)tile[t * (ZC + PAD) + j];
        }
      }
    }
    if (e < tot) {
      for (; r < nruns; ++r) {
        out[r * so_run + e * so_xy] = (float)acc;
        acc = 0.0;
      }
    }
  }
}

// Per collection slab s (blockIdx.y) and row x (blockIdx.x), from the slab projections
// proj[s][x][y] (float32): pre[s][x][m] = sum_{y < m} proj[s][x][y] for m = 0..Y (FP64).  Each
// thread owns a contiguous run of y.
__global__ void __launch_bounds__(256) slab_prefix_kernel(const float* __restrict__ proj,
                                                          int64_t X, int64_t Y,
                                                          double* __restrict__ pre) {
  __shared__ double part[256];
  const int64_t x = blockIdx.x;
  const int64_t s = blockIdx.y;
  const int t = threadIdx.x;
  const int64_t chunk = (Y + 255) / 256;
  const int64_t y0 = min<int64_t>(Y, t * chunk), y1 = min<int64_t>(Y, y0 + chunk);
  const float* row = proj + (s * X + x) * Y;
  double* prow = pre + (s * X + x) * (Y + 1);
  double local = 0.0;
  for (int64_t y = y0; y < y1; ++y) local += (double)row[y];
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
    run += (double)row[y];
    prow[y + 1] = run;
  }
}

// part[s][i][j] = w_s * sum_k sum_{l in run k} v_s[fi + c0 - k][fj + c1 - l] with
// (fi, fj) = (i + ox, j + oy), v_s zero outside X x Y, runs[off + 2k], runs[off + 2k + 1] = first
// and last column of disk row k: scipy.signal.fftconvolve(v_s, disk_s, 'same') on the output
// window.  Slab s is blockIdx.y, so every slab's disk runs in the same launch.
template <typename I>
__global__ void disk_kernel(const double* __restrict__ pre, I X, I Y,
                            const int32_t* __restrict__ slabs, const double* __restrict__ wts,
                            const int32_t* __restrict__ runs_all, I ox, I oy, I Xo, I Yo,
                            double* __restrict__ part) {
  const int s = blockIdx.y;
  const int32_t* sd = slabs + 6 * s;
  const int32_t* runs = runs_all + sd[2];
  const int K0 = sd[3];
  const I c0 = sd[4], c1 = sd[5];
  const double w = wts[s];
  const double* ps = pre + (int64_t)s * X * (Y + 1);
  const I tot = Xo * Yo;
  for (I e = blockIdx.x * (I)blockDim.x + threadIdx.x; e < tot; e += (I)gridDim.x * blockDim.x) {
    const I fi = e / Yo + ox, fj = e % Yo + oy;
    double acc = 0.0;
    for (int k = 0; k < K0; ++k) {
      const I r = fi + c0 - k;
      if (r < 0 || r >= X) continue;
      I lo = fj + c1 - runs[2 * k + 1], hi = fj + c1 - runs[2 * k];
      lo = lo < 0 ? 0 : lo;
      hi = hi > Y - 1 ? Y - 1 : hi;
      if (lo <= hi) acc += ps[r * (Y + 1) + hi + 1] - ps[r * (Y + 1) + lo];
    }
    part[(int64_t)s * tot + e] = w * acc;
  }
}

// out = 10 ** ((-float32(sum over slabs of part) / vres) * hemoabs), float32 as the reference's
// NumPy expression; the slab sum runs in slab order in FP64.
__global__ void hemo_mask_kernel(const double* __restrict__ part, int ns, int64_t n, float vres,
                                 float hemoabs, float* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    double total = 0.0;
    for (int s = 0; s < ns; ++s) total += part[(int64_t)s * n + e];
    const float t = (float)total;
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
  if (X * Y * Z + (int64_t)65535 * 256 < 2147483647LL)
    hipLaunchKernelGGL(grf_spectrum_kernel<int32_t>, dim3(grid_for(X * Y * Z)), dim3(256), 0,
                       (hipStream_t)stream, noise_re, noise_im, (int32_t)X, (int32_t)Y,
                       (int32_t)Z, kernels, coef, (int)K, (float2*)spec);
  else
    hipLaunchKernelGGL(grf_spectrum_kernel<int64_t>, dim3(grid_for(X * Y * Z)), dim3(256), 0,
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
  if (X * Y * Z + (int64_t)65535 * 256 < 2147483647LL)
    hipLaunchKernelGGL(grf_finish_kernel<int32_t>, dim3(grid_for(X * Y * Z)), dim3(256), 0,
                       (hipStream_t)stream, (const float2*)field, (int32_t)X, (int32_t)Y,
                       (int32_t)Z, scale, mu, first_z, last_z, occupancy, occupancy_div, out);
  else
    hipLaunchKernelGGL(grf_finish_kernel<int64_t>, dim3(grid_for(X * Y * Z)), dim3(256), 0,
                       (hipStream_t)stream, (const float2*)field, X, Y, Z, scale, mu, first_z,
                       last_z, occupancy, occupancy_div, out);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_zsum(const void* src, int32_t elem_bytes, int64_t X, int64_t Y,
                         int64_t Z, int64_t stride_x, int64_t stride_y, const int32_t* runs,
                         int32_t nruns, float* out, int64_t out_stride_run,
                         int64_t out_stride_xy, maomi_stream_t stream) {
  if (!src || !runs || !out || (elem_bytes != 1 && elem_bytes != 4) || X < 1 || Y < 1 ||
      Z < 1 || Z > 2147483647 || stride_y < Z || stride_x < Y * stride_y || nruns < 1 ||
      out_stride_run < 1 || out_stride_xy < 1) {
    maomi_set_error("maomi_zsum: bad arguments");
    return MAOMI_EINVAL;
  }
  if (elem_bytes == 1)
    hipLaunchKernelGGL((zsum_kernel<uint8_t, 64, 4>), dim3(grid_for(X * Y)), dim3(256), 0,
                       (hipStream_t)stream, (const uint8_t*)src, X, Y, (int)Z, stride_x,
                       stride_y, runs, (int)nruns, 2, out, out_stride_run, out_stride_xy);
  else
    hipLaunchKernelGGL((zsum_kernel<float, 32, 1>), dim3(grid_for(X * Y)), dim3(256), 0,
                       (hipStream_t)stream, (const float*)src, X, Y, (int)Z, stride_x, stride_y,
                       runs, (int)nruns, 2, out, out_stride_run, out_stride_xy);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int64_t maomi_collection_mask_workspace(int64_t X, int64_t Y, int64_t Xo, int64_t Yo,
                                                  int32_t ns) {
  if (X < 1 || Y < 1 || Xo < 1 || Yo < 1 || ns < 0) return -1;
  return (int64_t)ns * ((X * Y + 1) / 2 + X * (Y + 1) + Xo * Yo);
}

MAOMI_API int maomi_collection_mask(const void* vessels, int32_t elem_bytes, int64_t X,
                                    int64_t Y, int64_t Z, int64_t stride_x, int64_t stride_y,
                                    const int32_t* slabs, const double* weights,
                                    const int32_t* runs, int32_t ns, int64_t ox, int64_t oy,
                                    int64_t Xo, int64_t Yo, float vres, float hemoabs,
                                    double* workspace, float* out, maomi_stream_t stream) {
  if (!out || (ns > 0 && (!vessels || !slabs || !weights || !runs || !workspace)) ||
      (elem_bytes != 1 && elem_bytes != 4) || X < 1 || X > 2147483647 || Y < 1 || Z < 1 ||
      Z > 2147483647 || stride_y < Z || stride_x < Y * stride_y || ns < 0 || ns > 65535 ||
      ox < 0 || oy < 0 || Xo < 1 || Yo < 1 || ox + Xo > X || oy + Yo > Y || !(vres > 0.0f)) {
    maomi_set_error("maomi_collection_mask: bad arguments");
    return MAOMI_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  double* pre = workspace + (int64_t)ns * ((X * Y + 1) / 2);
  double* part = pre + (int64_t)ns * X * (Y + 1);
  if (ns > 0) {
    float* proj = (float*)workspace;
    // slab projections: the z runs are the (z0, z1) pairs at the head of each slab's entry
    if (elem_bytes == 1)
      hipLaunchKernelGGL((zsum_kernel<uint8_t, 64, 4>), dim3(grid_for(X * Y)), dim3(256), 0, st,
                         (const uint8_t*)vessels, X, Y, (int)Z, stride_x, stride_y, slabs,
                         (int)ns, 6, proj, X * Y, (int64_t)1);
    else
      hipLaunchKernelGGL((zsum_kernel<float, 32, 1>), dim3(grid_for(X * Y)), dim3(256), 0, st,
                         (const float*)vessels, X, Y, (int)Z, stride_x, stride_y, slabs,
                         (int)ns, 6, proj, X * Y, (int64_t)1);
    MAOMI_CHECK_LAUNCH();
    hipLaunchKernelGGL(slab_prefix_kernel, dim3((unsigned)X, (unsigned)ns), dim3(256), 0, st,
                       proj, X, Y, pre);
    MAOMI_CHECK_LAUNCH();
    const dim3 gd(grid_for(Xo * Yo), (unsigned)ns);
    if (X * (Y + 1) + (int64_t)65535 * 256 < 2147483647LL)
      hipLaunchKernelGGL(disk_kernel<int32_t>, gd, dim3(256), 0, st, pre, (int32_t)X,
                         (int32_t)Y, slabs, weights, runs, (int32_t)ox, (int32_t)oy,
                         (int32_t)Xo, (int32_t)Yo, part);
    else
      hipLaunchKernelGGL(disk_kernel<int64_t>, gd, dim3(256), 0, st, pre, X, Y, slabs, weights,
                         runs, ox, oy, Xo, Yo, part);
    MAOMI_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(hemo_mask_kernel, dim3(grid_for(Xo * Yo)), dim3(256), 0, st, part,
                     (int)ns, Xo * Yo, vres, hemoabs, out);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}
