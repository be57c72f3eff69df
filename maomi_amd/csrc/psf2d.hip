// Kernels of the 2-D PSF contracts (reference maomi/psf.py:154-203), FP64.
//
// maomi_resample_linear_f64: explicit linear intensity resampling of a kernel at given sample
//   coordinates, zero outside the source support (scipy.ndimage.map_coordinates, order 1,
//   mode "constant", as resample_psf calls it): a sample with any coordinate outside
//   [0, size - 1] is 0, every other sample is the bilinear blend of its 2 x 2 neighbours.
// maomi_pupil_embed_f64: the pupil amp exp(2 pi i OPD) (zero where the OPD is not finite or the
//   amplitude is 0), embedded in an n x n grid with its discrete origin at n // 2 and written
//   already ifftshifted, so that a plain FFT follows.
// maomi_power_shift_f64: |F|^2 of the FFT, fftshifted back to the centred layout.
#include "common.hpp"
#include "maomi_amd.h"

namespace {

__global__ void resample_linear_kernel(const double* __restrict__ src, int64_t sx, int64_t sy,
                                       const double* __restrict__ cx, int64_t nx,
                                       const double* __restrict__ cy, int64_t ny,
                                       double* __restrict__ dst) {
  const int64_t tot = nx * ny;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / ny, j = e - i * ny;
    const double x = cx[i], y = cy[j];
    double v = 0.0;
    if (x >= 0.0 && x <= (double)(sx - 1) && y >= 0.0 && y <= (double)(sy - 1)) {
      const int64_t x0 = (int64_t)floor(x), y0 = (int64_t)floor(y);
      const double fx = x - (double)x0, fy = y - (double)y0;
      // the upper neighbour at the last sample carries weight 0: clamp its index
      const int64_t x1 = x0 + 1 < sx ? x0 + 1 : x0, y1 = y0 + 1 < sy ? y0 + 1 : y0;
      const double a00 = src[x0 * sy + y0], a01 = src[x0 * sy + y1];
      const double a10 = src[x1 * sy + y0], a11 = src[x1 * sy + y1];
      v = (1.0 - fx) * ((1.0 - fy) * a00 + fy * a01) + fx * ((1.0 - fy) * a10 + fy * a11);
    }
    dst[e] = v;
  }
}

__global__ void pupil_embed_kernel(const double* __restrict__ opd, const double* __restrict__ amp,
                                   int64_t px, int64_t py, int64_t n, double2* __restrict__ out) {
  const int64_t tot = n * n;
  const double two_pi = 6.283185307179586;  // 2 pi rounded to double, as 2 * numpy.pi
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k1 = e / n, k2 = e - k1 * n;
    // ifftshift places centred index m at (m - n // 2) mod n; the pupil sample (i, j) sits at
    // m = n // 2 - p // 2 + i, so output k holds sample i = (k + p // 2) mod n when i < p
    const int64_t i = (k1 + px / 2) % n, j = (k2 + py / 2) % n;
    double2 v = make_double2(0.0, 0.0);
    if (i < px && j < py) {
      const double o = opd[i * py + j];
      const double a = amp ? amp[i * py + j] : 1.0;
      if (isfinite(o) && a > 0.0) {
        double s, c;
        sincos(two_pi * o, &s, &c);
        v = make_double2(a * c, a * s);
      }
    }
    out[e] = v;
  }
}

__global__ void power_shift_kernel(const double2* __restrict__ F, int64_t n,
                                   double* __restrict__ out) {
  const int64_t tot = n * n, h = n / 2;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k1 = e / n, k2 = e - k1 * n;
    const double2 f = F[((k1 - h + n) % n) * n + (k2 - h + n) % n];  // fftshift
    out[e] = f.x * f.x + f.y * f.y;
  }
}

unsigned blocks_for(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return (unsigned)(b < 65535 ? (b > 0 ? b : 1) : 65535);
}

}  // namespace

MAOMI_API int maomi_resample_linear_f64(const double* src, int64_t sx, int64_t sy,
                                        const double* cx, int64_t nx, const double* cy,
                                        int64_t ny, double* dst, maomi_stream_t stream) {
  if (!src || !cx || !cy || !dst || sx < 1 || sy < 1 || nx < 1 || ny < 1) {
    maomi_set_error("maomi_resample_linear_f64: bad arguments");
    return MAOMI_EINVAL;
  }
  hipLaunchKernelGGL(resample_linear_kernel, dim3(blocks_for(nx * ny)), dim3(256), 0,
                     (hipStream_t)stream, src, sx, sy, cx, nx, cy, ny, dst);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_pupil_embed_f64(const double* opd, const double* amp, int64_t px, int64_t py,
                                    int64_t n, void* out, maomi_stream_t stream) {
  if (!opd || !out || px < 1 || py < 1 || n < px || n < py) {
    maomi_set_error("maomi_pupil_embed_f64: bad arguments (n must cover the pupil grid)");
    return MAOMI_EINVAL;
  }
  hipLaunchKernelGGL(pupil_embed_kernel, dim3(blocks_for(n * n)), dim3(256), 0,
                     (hipStream_t)stream, opd, amp, px, py, n, (double2*)out);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_power_shift_f64(const void* F, int64_t n, double* out,
                                    maomi_stream_t stream) {
  if (!F || !out || n < 1) {
    maomi_set_error("maomi_power_shift_f64: bad arguments");
    return MAOMI_EINVAL;
  }
  hipLaunchKernelGGL(power_shift_kernel, dim3(blocks_for(n * n)), dim3(256), 0,
                     (hipStream_t)stream, (const double2*)F, n, out);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}
