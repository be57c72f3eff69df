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
                              const double* __restrict__ zw, double* __restrict__ tiles, int64_t B,
                              int64_t Lx, int64_t Ly, int64_t Lz) {
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
      if (zw) v *= zw[z];
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
                               const double* __restrict__ wts, const double* __restrict__ zw,
                               double* __restrict__ out, int64_t Y, int64_t Zo, int64_t B,
                               int64_t Tx, int64_t Ty) {
  const int64_t per = Tx * Ty * Zo, tot = B * per;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = e / per, r = e % per;
    const int64_t x = r / (Ty * Zo), y = (r / Zo) % Ty, z = r % Zo;
    const int32_t* d = desc + b * MAOMI_LFFT_SCATTER_DESC;
    if (x >= d[2] || y >= d[3]) continue;
    double v = full[((b * Lx + d[4] + x) * Ly + d[5] + y) * Lz + d[6] + z];
    if (d[7] >= 0) v *= wts[d[7] + x * d[3] + y];
    if (zw) v *= zw[z];
    atomicAdd(out + ((int64_t)(d[0] + x) * Y + (d[1] + y)) * Zo + z, v);
  }
}

// out[i] = sum_f fw[f] spec[i] kspec[f][i] (nz = 0), or with the z sum of a plane projection,
// out[i] = sum_f fw[f] sum_z spec[i][z] kspec[f][i][z] (z innermost, length nz): the fields
// of a depth bank whose weights are one number each share one inverse transform.
__global__ void product_fields_kernel(const double2* __restrict__ spec,
                                      const double2* __restrict__ kspec, int64_t kstride,
                                      const double* __restrict__ fw, int F,
                                      double2* __restrict__ out, int64_t n, int64_t nz) {
  const int64_t m = nz > 0 ? nz : 1;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    double re = 0.0, im = 0.0;
    for (int f = 0; f < F; ++f) {
      const double w = fw ? fw[f] : 1.0;
      if (w == 0.0) continue;
      const double2* k = kspec + f * kstride + e * m;
      const double2* a = spec + e * m;
      double fr = 0.0, fi = 0.0;
      for (int64_t z = 0; z < m; ++z) {
        fr += a[z].x * k[z].x - a[z].y * k[z].y;
        fi += a[z].x * k[z].y + a[z].y * k[z].x;
      }
      re += w * fr;
      im += w * fi;
    }
    out[e] = make_double2(re, im);
  }
}

// Smooth node weights of the acquisition operator at pixel centres origin + (i + 1/2) pitch:
// the bracketing nodes k, k + 1 (clamped to the grid) share the position by a smoothstep of its
// fraction between them; out is n_nodes x n_pos, zero off the two brackets.
__global__ void node_weights_kernel(const double* __restrict__ nodes, int n_nodes, double origin,
                                    double pitch, int64_t n_pos, double* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_pos;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double c = origin + ((double)i + 0.5) * pitch;
    int lo = 0, hi = n_nodes;  // first node above c
    while (lo < hi) {
      const int mid = (lo + hi) / 2;
      if (nodes[mid] <= c) lo = mid + 1; else hi = mid;
    }
    int k = lo - 1;
    k = k < 0 ? 0 : (k > n_nodes - 2 ? n_nodes - 2 : k);
    double t = (c - nodes[k]) / (nodes[k + 1] - nodes[k]);
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    t = t * t * (3.0 - 2.0 * t);
    for (int j = 0; j < n_nodes; ++j) out[(int64_t)j * n_pos + i] = 0.0;
    out[(int64_t)k * n_pos + i] = 1.0 - t;
    out[(int64_t)(k + 1) * n_pos + i] = t;
  }
}

// Linear depth weights: field f's hat function over the depth nodes at each position z[i]
// (clamped to the end nodes); out is F x nz and each column sums to one.
__global__ void depth_weights_kernel(const double* __restrict__ depths, int F,
                                     const double* __restrict__ z, int64_t nz,
                                     double* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nz;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double x = z[i];
    for (int f = 0; f < F; ++f) out[(int64_t)f * nz + i] = 0.0;
    if (x <= depths[0]) {
      out[i] = 1.0;
    } else if (x >= depths[F - 1]) {
      out[(int64_t)(F - 1) * nz + i] = 1.0;
    } else {
      int k = 0;
      while (k + 2 < F && depths[k + 1] <= x) ++k;
      const double t = (x - depths[k]) / (depths[k + 1] - depths[k]);
      out[(int64_t)k * nz + i] = 1.0 - t;
      out[(int64_t)(k + 1) * nz + i] = t;
    }
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
                     (hipStream_t)stream, source, Y, Z, desc, weights, nullptr, tiles, B, Lx,
                     Ly, Lz);
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
                     (hipStream_t)stream, full, Lx, Ly, Lz, desc, weights, nullptr, out, Y, Zo,
                     B, Tx, Ty);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_lfft_gather_z(const double* source, int64_t X, int64_t Y, int64_t Z,
                                  const int32_t* desc, const double* weights,
                                  const double* zweights, double* tiles, int64_t B, int64_t Lx,
                                  int64_t Ly, int64_t Lz, maomi_stream_t stream) {
  if (!source || !desc || !tiles || X < 1 || Y < 1 || Z < 1 || B < 1 || Lx < 1 || Ly < 1 ||
      Lz < Z) {
    maomi_set_error("maomi_lfft_gather_z: bad arguments");
    return MAOMI_EINVAL;
  }
  hipLaunchKernelGGL(gather_kernel, dim3(grid_for(B * Lx * Ly * Lz)), dim3(256), 0,
                     (hipStream_t)stream, source, Y, Z, desc, weights, zweights, tiles, B, Lx, Ly,
                     Lz);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_lfft_scatter_z(const double* full, int64_t B, int64_t Lx, int64_t Ly,
                                   int64_t Lz, const int32_t* desc, const double* weights,
                                   const double* zweights, double* out, int64_t X, int64_t Y,
                                   int64_t Zo, int64_t Tx, int64_t Ty, maomi_stream_t stream) {
  if (!full || !desc || !out || B < 1 || Lx < 1 || Ly < 1 || Lz < Zo || X < 1 || Y < 1 ||
      Zo < 1 || Tx < 1 || Ty < 1 || Tx > Lx || Ty > Ly) {
    maomi_set_error("maomi_lfft_scatter_z: bad arguments");
    return MAOMI_EINVAL;
  }
  hipLaunchKernelGGL(scatter_kernel, dim3(grid_for(B * Tx * Ty * Zo)), dim3(256), 0,
                     (hipStream_t)stream, full, Lx, Ly, Lz, desc, weights, zweights, out, Y, Zo, B,
                     Tx, Ty);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_lfft_product_fields(const void* spec, const void* kspec, int64_t kstride,
                                        const double* fweights, int64_t F, void* out, int64_t n,
                                        int64_t nz, maomi_stream_t stream) {
  if (!spec || !kspec || !out || F < 1 || n < 1 || nz < 0 || kstride < n * (nz > 0 ? nz : 1)) {
    maomi_set_error("maomi_lfft_product_fields: bad arguments");
    return MAOMI_EINVAL;
  }
  hipLaunchKernelGGL(product_fields_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream,
                     (const double2*)spec, (const double2*)kspec, kstride, fweights, (int)F,
                     (double2*)out, n, nz);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_rana_node_weights(const double* nodes, int64_t n_nodes, double origin,
                                      double pitch, int64_t n_pos, double* out,
                                      maomi_stream_t stream) {
  if (!nodes || !out || n_nodes < 2 || n_pos < 1 || !(pitch > 0.0)) {
    maomi_set_error("maomi_rana_node_weights: bad arguments (>= 2 nodes, pitch > 0)");
    return MAOMI_EINVAL;
  }
  hipLaunchKernelGGL(node_weights_kernel, dim3(grid_for(n_pos)), dim3(256), 0,
                     (hipStream_t)stream, nodes, (int)n_nodes, origin, pitch, n_pos, out);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_rana_depth_weights(const double* depths, int64_t F, const double* z,
                                       int64_t nz, double* out, maomi_stream_t stream) {
  if (!depths || !z || !out || F < 2 || nz < 1) {
    maomi_set_error("maomi_rana_depth_weights: bad arguments (>= 2 depth nodes)");
    return MAOMI_EINVAL;
  }
  hipLaunchKernelGGL(depth_weights_kernel, dim3(grid_for(nz)), dim3(256), 0,
                     (hipStream_t)stream, depths, (int)F, z, nz, out);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}
