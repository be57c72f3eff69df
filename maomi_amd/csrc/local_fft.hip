// Kernels of the local-PSF FFT plan (RANA acquisition; reference maomi/rana/_fft.py:8-99):
// gather of weighted, zero-padded source tiles, the product of tile and kernel spectra
// (with the sum over z of a plane projection), and the weighted scatter of tile responses
// into the output.  All arrays are FP64; spectra are complex128 (double2), as torch lays
// them out.  The FFTs themselves are the caller's (rocFFT through torch).
#include "common.hpp"
#include "maomi_amd.h"

namespace {

inline unsigned grid_for(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return (unsigned)(b < 65535 ? (b > 0 ? b : 1) : 65535);
}

// tiles[b][x][y][z] (Lx x Ly x Lz) = source[sx0 + x][sy0 + y][z] * w[x][y] inside the
// tile's nx x ny x Z source window, zero in the padding.
__global__ void gather_kernel(const double* __restrict__ src, int64_t Y, int64_t Z,
                              const int32_t* __restrict__ desc, const double* __restrict__ wts,
                              double* __restrict__ tiles, int64_t B, int64_t Lx, int64_t Ly,
                              int64_t Lz) {
  const int64_t per = Lx * Ly * Lz, tot = B * per;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = e / per, r = e % per;
    const int64_t x = r / (Ly * Lz), y = (r / Lz) % Ly, z = r % Lz;
    const int32_t* d = desc + b * MAOMI_LFFT_GATHER_DESC;
    double v = 0.0;
    if (x < d[2] && y < d[3] && z < Z) {
      v = src[((int64_t)(d[0] + x) * Y + (d[1] + y)) * Z + z];
      if (d[4] >= 0) v *= wts[d[4] + x * d[3] + y];
    }
    tiles[e] = v;
  }
}

// spec[i] *= kspec[i] (complex)
__global__ void product_kernel(double2* __restrict__ spec, const double2* __restrict__ kspec,
                               int64_t n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const double2 a = spec[e], k = kspec[e];
    spec[e] = make_double2(a.x * k.x - a.y * k.y, a.x * k.y + a.y * k.x);
  }
}

// out[i] = sum_z spec[i][z] kspec[i][z] (complex; z innermost, length nz)
__global__ void product_sum_kernel(const double2* __restrict__ spec,
                                   const double2* __restrict__ kspec, double2* __restrict__ out,
                                   int64_t n, int64_t nz) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    double re = 0.0, im = 0.0;
    for (int64_t z = 0; z < nz; ++z) {
      const double2 a = spec[e * nz + z], k = kspec[e * nz + z];
      re += a.x * k.x - a.y * k.y;
      im += a.x * k.y + a.y * k.x;
    }
    out[e] = make_double2(re, im);
  }
}

// out[tx0 + x][ty0 + y][z] += full[b][rx0 + x][ry0 + y][rz0 + z] * w[x][y] for the tile's
// tx x ty x Zo target window (Zo = 1 and no z axis for a plane projection).
__global__ void scatter_kernel(const double* __restrict__ full, int64_t Lx, int64_t Ly,
                               int64_t Lz, const int32_t* __restrict__ desc,
                               const double* __restrict__ wts, double* __restrict__ out,
                               int64_t Y, int64_t Zo, int64_t B, int64_t Tx, int64_t Ty) {
  const int64_t per = Tx * Ty * Zo, tot = B * per;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = e / per, r = e % per;
    const int64_t x = r / (Ty * Zo), y = (r / Zo) % Ty, z = r % Zo;
    const int32_t* d = desc + b * MAOMI_LFFT_SCATTER_DESC;
    if (x >= d[2] || y >= d[3]) continue;
    double v = full[((b * Lx + d[4] + x) * Ly + d[5] + y) * Lz + d[6] + z];
    if (d[7] >= 0) v *= wts[d[7] + x * d[3] + y];
    atomicAdd(out + ((int64_t)(d[0] + x) * Y + (d[1] + y)) * Zo + z, v);
  }
}

}  // namespace

MAOMI_API int maomi_lfft_gather(const double* source, int64_t X, int64_t Y, int64_t Z,
                                const int32_t* desc, const double* weights, double* tiles,
                                int64_t B, int64_t Lx, int64_t Ly, int64_t Lz,
                                maomi_stream_t stream) {
  if (!source || !desc || !tiles || X < 1 || Y < 1 || Z < 1 || B < 1 || Lx < 1 || Ly < 1 ||
      Lz < Z) {
    maomi_set_error("maomi_lfft_gather: bad arguments");
    return MAOMI_EINVAL;
  }
  hipLaunchKernelGGL(gather_kernel, dim3(grid_for(B * Lx * Ly * Lz)), dim3(256), 0,
                     (hipStream_t)stream, source, Y, Z, desc, weights, tiles, B, Lx, Ly, Lz);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_lfft_product(void* spec, const void* kspec, void* out, int64_t n,
                                 int64_t nz, maomi_stream_t stream) {
  if (!spec || !kspec || n < 1 || (out && nz < 1)) {
    maomi_set_error("maomi_lfft_product: bad arguments");
    return MAOMI_EINVAL;
  }
  if (out)
    hipLaunchKernelGGL(product_sum_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream,
                       (const double2*)spec, (const double2*)kspec, (double2*)out, n, nz);
  else
    hipLaunchKernelGGL(product_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream,
                       (double2*)spec, (const double2*)kspec, n);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_lfft_scatter(const double* full, int64_t B, int64_t Lx, int64_t Ly,
                                 int64_t Lz, const int32_t* desc, const double* weights,
                                 double* out, int64_t X, int64_t Y, int64_t Zo, int64_t Tx,
                                 int64_t Ty, maomi_stream_t stream) {
  if (!full || !desc || !out || B < 1 || Lx < 1 || Ly < 1 || Lz < Zo || X < 1 || Y < 1 ||
      Zo < 1 || Tx < 1 || Ty < 1 || Tx > Lx || Ty > Ly) {
    maomi_set_error("maomi_lfft_scatter: bad arguments");
    return MAOMI_EINVAL;
  }
  hipLaunchKernelGGL(scatter_kernel, dim3(grid_for(B * Tx * Ty * Zo)), dim3(256), 0,
                     (hipStream_t)stream, full, Lx, Ly, Lz, desc, weights, out, Y, Zo, B, Tx, Ty);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}
