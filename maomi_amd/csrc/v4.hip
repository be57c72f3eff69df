// FP64 MATLAB-v4 (biharmonic spline) mask assembly on the GPU.
//
// Replaces green_apply (maomi/backends/cuda/interpolation.cu:3-19) and the
// CPU Green-matrix build of v4_interpolate (maomi/numerics.py:122-129):
//   G(p,q) = d^2 (ln d - 1) = r2 (0.5 ln r2 - 1),  G = 0 at d = 0,
//   out(q, c) = sum_i G(|q - p_i|) coef(i, c).
// Points and coefficients are staged through LDS in tiles; each thread owns
// one query and keeps C accumulators in registers.
#include "common.hpp"
#include "maomi_amd.h"

namespace {

constexpr int kTile = 256;
constexpr int kMaxC = 8;

__device__ __forceinline__ double green_r2(double r2) {
  return r2 > 0.0 ? r2 * (0.5 * log(r2) - 1.0) : 0.0;
}

// Points [p0, p1) of the sum go to block row blockIdx.y: p0 = blockIdx.y * pchunk.  With one
// row the block writes out(q, c); with several each row writes its partial sum to
// out + blockIdx.y * Q * C and reduce_chunks_kernel adds the rows in order.
template <int C>
__global__ __launch_bounds__(256) void green_apply_kernel(const double* __restrict__ p,
                                                          const double* __restrict__ coef,
                                                          const double* __restrict__ q,
                                                          double* __restrict__ out, int64_t P,
                                                          int64_t Q, int64_t pchunk) {
  __shared__ double px[kTile], py[kTile], pc[kTile * C];
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const bool live = t < Q;
  const double x = live ? q[2 * t] : 0.0, y = live ? q[2 * t + 1] : 0.0;
  const int64_t p0 = (int64_t)blockIdx.y * pchunk, p1 = min(P, p0 + pchunk);
  out += (int64_t)blockIdx.y * Q * C;
  double acc[C];
#pragma unroll
  for (int j = 0; j < C; ++j) acc[j] = 0.0;
  for (int64_t base = p0; base < p1; base += kTile) {
    const int n = (int)min((int64_t)kTile, p1 - base);
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      px[i] = p[2 * (base + i)];
      py[i] = p[2 * (base + i) + 1];
#pragma unroll
      for (int j = 0; j < C; ++j) pc[i * C + j] = coef[(base + i) * C + j];
    }
    __syncthreads();
    for (int i = 0; i < n; ++i) {
      const double dx = x - px[i], dy = y - py[i];
      const double g = green_r2(dx * dx + dy * dy);
#pragma unroll
      for (int j = 0; j < C; ++j) acc[j] += g * pc[i * C + j];
    }
  }
  if (live) {
#pragma unroll
    for (int j = 0; j < C; ++j) out[t * C + j] = acc[j];
  }
}

// out[e] = sum over rows r < R of part[r][e], in row order
__global__ void reduce_chunks_kernel(const double* __restrict__ part, int R, int64_t n,
                                     double* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int r = 0; r < R; ++r) s += part[(int64_t)r * n + e];
    out[e] = s;
  }
}

__global__ void green_matrix_kernel(const double* __restrict__ p, double* __restrict__ G,
                                    int64_t P) {
  const int64_t tot = P * P;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < tot;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = k / P, j = k % P;
    const double dx = p[2 * i] - p[2 * j], dy = p[2 * i + 1] - p[2 * j + 1];
    // d = sqrt(r2); g = d*d*(log d - 1) as the reference writes it
    const double d = sqrt(dx * dx + dy * dy);
    G[k] = d == 0.0 ? 0.0 : d * d * (log(d) - 1.0);
  }
}

}  // namespace

MAOMI_API int maomi_v4_apply_split_f64(const double* points, const double* coef,
                                       const double* queries, double* out, int64_t P, int64_t Q,
                                       int64_t C, double* partial, int64_t rows,
                                       maomi_stream_t stream);

MAOMI_API int maomi_v4_green_matrix_f64(const double* points, int64_t P, double* G,
                                        maomi_stream_t stream) {
  if (!points || !G || P < 1) {
    maomi_set_error("maomi_v4_green_matrix_f64: bad arguments");
    return MAOMI_EINVAL;
  }
  const int64_t tot = P * P;
  const unsigned blocks = (unsigned)((tot + 255) / 256 < 65535 ? (tot + 255) / 256 : 65535);
  hipLaunchKernelGGL(green_matrix_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, points,
                     G, P);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_v4_apply_f64(const double* points, const double* coef, const double* queries,
                                 double* out, int64_t P, int64_t Q, int64_t C,
                                 maomi_stream_t stream) {
  if (!points || !coef || !queries || !out || P < 1 || Q < 0 || C < 1 || C > kMaxC) {
    maomi_set_error("maomi_v4_apply_f64: bad arguments (1 <= C <= %d)", kMaxC);
    return MAOMI_EINVAL;
  }
  if (Q == 0) return MAOMI_OK;
  return maomi_v4_apply_split_f64(points, coef, queries, out, P, Q, C, nullptr, 1, stream);
}

MAOMI_API int64_t maomi_v4_apply_rows(int64_t P, int64_t Q) {
  // enough point rows for ~8 waves per SIMD on 256 CUs, each row at least 4 tiles of points
  const int64_t qblocks = (Q + 255) / 256;
  int64_t rows = (8 * 1024 + qblocks - 1) / (qblocks > 0 ? qblocks : 1);
  const int64_t cap = (P + 4 * kTile - 1) / (4 * kTile);
  rows = rows < cap ? rows : cap;
  return rows < 1 ? 1 : (rows > 65535 ? 65535 : rows);
}

MAOMI_API int maomi_v4_apply_split_f64(const double* points, const double* coef,
                                       const double* queries, double* out, int64_t P, int64_t Q,
                                       int64_t C, double* partial, int64_t rows,
                                       maomi_stream_t stream) {
  if (!points || !coef || !queries || !out || P < 1 || Q < 0 || C < 1 || C > kMaxC ||
      rows < 1 || rows > 65535 || (rows > 1 && !partial)) {
    maomi_set_error("maomi_v4_apply_split_f64: bad arguments (1 <= C <= %d)", kMaxC);
    return MAOMI_EINVAL;
  }
  if (Q == 0) return MAOMI_OK;
  const int64_t pchunk = (P + rows - 1) / rows;
  const dim3 grid((unsigned)((Q + 255) / 256), (unsigned)rows), block(256);
  double* dst = rows > 1 ? partial : out;
  hipStream_t st = (hipStream_t)stream;
  switch (C) {
#define CASE(c) \
  case c: hipLaunchKernelGGL(green_apply_kernel<c>, grid, block, 0, st, points, coef, queries, dst, P, Q, pchunk); break;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
  }
  MAOMI_CHECK_LAUNCH();
  if (rows > 1) {
    const int64_t n = Q * C;
    const unsigned b = (unsigned)((n + 255) / 256 < 65535 ? (n + 255) / 256 : 65535);
    hipLaunchKernelGGL(reduce_chunks_kernel, dim3(b), dim3(256), 0, st, partial, (int)rows, n,
                       out);
    MAOMI_CHECK_LAUNCH();
  }
  return MAOMI_OK;
}
