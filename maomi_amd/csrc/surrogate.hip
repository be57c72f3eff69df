// Kernels of the learned PSF surrogate: BF16 MFMA GEMM with FP32 accumulation
// (forward and both backward products of every convolution), im2col / col2im,
// bias + leaky-ReLU forward and backward, nearest 2x upsampling, the relative
// squared-error loss and an Adam step on FP32 master weights with BF16 copies.
//
// Activations are channel-major: [C][B][H][W], so a convolution is
//   Y[co][n] = sum_k W[co][k] col[k][n],  k = (ci, ky, kx),  n = (b, y, x),
// dcol = W^T dY and dW = dY col^T, each one call of the GEMM below.
#include "common.hpp"
#include "maomi_amd.h"
#include <hip/hip_bf16.h>

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TM = 64, TN = 64, TK = 32;  // block tile
constexpr int GT = 256;                   // threads per GEMM block (4 waves, 2 x 2 of 32 x 32)

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }

// C[m][n] = alpha * sum_k A(m, k) B(k, n) + beta * C[m][n]
// A(m, k) = A[m * sam + k * sak], B(k, n) = B[k * sbk + n * sbn], C row stride ldc.
__global__ __launch_bounds__(GT) void gemm_bf16_kernel(const __bf16* __restrict__ A,
                                                       const __bf16* __restrict__ B,
                                                       float* __restrict__ C, int M, int N, int K,
                                                       int64_t sam, int64_t sak, int64_t sbk,
                                                       int64_t sbn, int64_t ldc, float alpha,
                                                       float beta) {
  __shared__ __bf16 As[TM][TK + 8];
  __shared__ __bf16 Bs[TN][TK + 8];  // stored transposed: Bs[n][k]
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += TK) {
    // 64 x 32 elements of each operand, 8 per thread
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int e = tid + r * GT;
      const int mm = e / TK, kk = e % TK;
      const int gm = m0 + mm, gk = k0 + kk;
      As[mm][kk] = (gm < M && gk < K) ? A[gm * sam + gk * sak] : to_bf16(0.f);
      const int nn = e % TN, kb = e / TN;
      const int gn = n0 + nn, gkb = k0 + kb;
      Bs[nn][kb] = (gn < N && gkb < K) ? B[gkb * sbk + gn * sbn] : to_bf16(0.f);
    }
    __syncthreads();
    const int row = lane & 15, kq = (lane >> 4) * 8;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      bf16x8 a;
#pragma unroll
      for (int t = 0; t < 8; ++t) a[t] = As[wm + 16 * i + row][kq + t];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        bf16x8 b;
#pragma unroll
        for (int t = 0; t < 8; ++t) b[t] = Bs[wn + 16 * j + row][kq + t];
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // D[4 (lane / 16) + t][lane % 16] of each 16 x 16 tile
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int gm = m0 + wm + 16 * i + 4 * (lane >> 4) + t;
        const int gn = n0 + wn + 16 * j + (lane & 15);
        if (gm < M && gn < N) {
          float* c = C + gm * ldc + gn;
          *c = alpha * acc[i][j][t] + (beta != 0.f ? beta * *c : 0.f);
        }
      }
}

// Tiled GEMM for operands with one unit stride each (every product of the network):
//   AK: A(m, k) = A[m lda + k] (k contiguous), else A(m, k) = A[k lda + m] (m contiguous);
//   BK: B(k, n) = B[n ldb + k] (k contiguous), else B(k, n) = B[k ldb + n] (n contiguous).
// BM x BN block tile, K in slabs of 32, 256 threads = 2 x 2 waves of (BM/2) x (BN/2).
// Global loads are 8-element (16-byte when VEC) runs along the contiguous direction; the
// next slab is fetched into registers while the MFMAs consume the current one from LDS.
// LDS rows hold 32 k + 8 pad (80 bytes): the 16 rows one MFMA fragment read touches fall
// in distinct 16-byte bank groups.
union Vec8 {
  uint4 v;
  unsigned short h[8];
};

template <bool VEC>
__device__ __forceinline__ uint4 load_run(const unsigned short* __restrict__ p, int valid) {
  Vec8 r;
  if (VEC && valid >= 8) {
    r.v = *reinterpret_cast<const uint4*>(p);
  } else {
#pragma unroll
    for (int t = 0; t < 8; ++t) r.h[t] = t < valid ? p[t] : (unsigned short)0;
  }
  return r.v;
}

constexpr int SK = 32, SKP = SK + 8;

template <int BM, int BN, bool AK, bool BK, bool VEC>
__global__ __launch_bounds__(256) void gemm_tiled_kernel(
    const unsigned short* __restrict__ A, const unsigned short* __restrict__ B,
    float* __restrict__ C, int M, int N, int K, int64_t lda, int64_t ldb, int64_t ldc,
    float alpha, float beta) {
  constexpr int RA = BM * SK / 8 / 256, RB = BN * SK / 8 / 256;
  constexpr int FM = BM / 32, FN = BN / 32;
  __shared__ __attribute__((aligned(16))) unsigned short As[BM][SKP];
  __shared__ __attribute__((aligned(16))) unsigned short Bs[BN][SKP];
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * (BM / 2), wn = (wave & 1) * (BN / 2);
  uint4 ra[RA], rb[RB];

  auto fetch = [&](int k0) {
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      const int v = tid + r * 256;
      if (AK) {
        const int row = v / (SK / 8), kv = (v % (SK / 8)) * 8;
        const int gm = m0 + row, gk = k0 + kv;
        ra[r] = load_run<VEC>(A + (int64_t)gm * lda + gk, gm < M ? K - gk : 0);
      } else {
        const int k = v / (BM / 8), mv = (v % (BM / 8)) * 8;
        const int gm = m0 + mv, gk = k0 + k;
        ra[r] = load_run<VEC>(A + (int64_t)gk * lda + gm, gk < K ? M - gm : 0);
      }
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int v = tid + r * 256;
      if (BK) {
        const int col = v / (SK / 8), kv = (v % (SK / 8)) * 8;
        const int gn = n0 + col, gk = k0 + kv;
        rb[r] = load_run<VEC>(B + (int64_t)gn * ldb + gk, gn < N ? K - gk : 0);
      } else {
        const int k = v / (BN / 8), nv = (v % (BN / 8)) * 8;
        const int gn = n0 + nv, gk = k0 + k;
        rb[r] = load_run<VEC>(B + (int64_t)gk * ldb + gn, gk < K ? N - gn : 0);
      }
    }
  };
  auto stage = [&]() {
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      const int v = tid + r * 256;
      if (AK) {
        *reinterpret_cast<uint4*>(&As[v / (SK / 8)][(v % (SK / 8)) * 8]) = ra[r];
      } else {
        Vec8 u;
        u.v = ra[r];
        const int k = v / (BM / 8), mv = (v % (BM / 8)) * 8;
#pragma unroll
        for (int t = 0; t < 8; ++t) As[mv + t][k] = u.h[t];
      }
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int v = tid + r * 256;
      if (BK) {
        *reinterpret_cast<uint4*>(&Bs[v / (SK / 8)][(v % (SK / 8)) * 8]) = rb[r];
      } else {
        Vec8 u;
        u.v = rb[r];
        const int k = v / (BN / 8), nv = (v % (BN / 8)) * 8;
#pragma unroll
        for (int t = 0; t < 8; ++t) Bs[nv + t][k] = u.h[t];
      }
    }
  };

  f32x4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  fetch(0);
  const int row = lane & 15, kq = (lane >> 4) * 8;
  for (int k0 = 0; k0 < K; k0 += SK) {
    stage();
    __syncthreads();
    if (k0 + SK < K) fetch(k0 + SK);
    bf16x8 af[FM], bfr[FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
      af[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&As[wm + 16 * i + row][kq]));
#pragma unroll
    for (int j = 0; j < FN; ++j)
      bfr[j] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&Bs[wn + 16 * j + row][kq]));
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int gm = m0 + wm + 16 * i + 4 * (lane >> 4) + t;
        const int gn = n0 + wn + 16 * j + (lane & 15);
        if (gm < M && gn < N) {
          float* c = C + (int64_t)gm * ldc + gn;
          *c = alpha * acc[i][j][t] + (beta != 0.f ? beta * *c : 0.f);
        }
      }
}

template <int BM, bool AK, bool BK, bool VEC>
void launch_tiled(const void* A, const void* B, float* C, int64_t M, int64_t N, int64_t K,
                  int64_t lda, int64_t ldb, int64_t ldc, float alpha, float beta,
                  hipStream_t stream) {
  constexpr int BN = 128;
  const dim3 grid((unsigned)((N + BN - 1) / BN), (unsigned)((M + BM - 1) / BM));
  hipLaunchKernelGGL((gemm_tiled_kernel<BM, BN, AK, BK, VEC>), grid, dim3(256), 0, stream,
                     (const unsigned short*)A, (const unsigned short*)B, C, (int)M, (int)N,
                     (int)K, lda, ldb, ldc, alpha, beta);
}

template <int BM, bool AK, bool BK>
void launch_tiled_vec(bool vec, const void* A, const void* B, float* C, int64_t M, int64_t N,
                      int64_t K, int64_t lda, int64_t ldb, int64_t ldc, float alpha, float beta,
                      hipStream_t stream) {
  if (vec)
    launch_tiled<BM, AK, BK, true>(A, B, C, M, N, K, lda, ldb, ldc, alpha, beta, stream);
  else
    launch_tiled<BM, AK, BK, false>(A, B, C, M, N, K, lda, ldb, ldc, alpha, beta, stream);
}

template <int BM>
void launch_tiled_layout(bool ak, bool bk, bool vec, const void* A, const void* B, float* C,
                         int64_t M, int64_t N, int64_t K, int64_t lda, int64_t ldb, int64_t ldc,
                         float alpha, float beta, hipStream_t stream) {
  if (ak && bk)
    launch_tiled_vec<BM, true, true>(vec, A, B, C, M, N, K, lda, ldb, ldc, alpha, beta, stream);
  else if (ak)
    launch_tiled_vec<BM, true, false>(vec, A, B, C, M, N, K, lda, ldb, ldc, alpha, beta, stream);
  else if (bk)
    launch_tiled_vec<BM, false, true>(vec, A, B, C, M, N, K, lda, ldb, ldc, alpha, beta, stream);
  else
    launch_tiled_vec<BM, false, false>(vec, A, B, C, M, N, K, lda, ldb, ldc, alpha, beta, stream);
}

// col[(c*k + ky)*k + kx][n], n = (b*Ho + oy)*Wo + ox, from X[c][b][iy][ix] (float -> bf16)
__global__ void im2col_kernel(const float* __restrict__ X, __bf16* __restrict__ col, int C, int Bn,
                              int H, int W, int k, int s, int p, int Ho, int Wo) {
  const int64_t Ncol = (int64_t)Bn * Ho * Wo;
  const int64_t tot = (int64_t)C * k * k * Ncol;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t n = e % Ncol;
    const int r = (int)(e / Ncol);
    const int kx = r % k, ky = (r / k) % k, c = r / (k * k);
    const int ox = (int)(n % Wo), oy = (int)((n / Wo) % Ho), b = (int)(n / ((int64_t)Wo * Ho));
    const int iy = s * oy + ky - p, ix = s * ox + kx - p;
    float v = 0.f;
    if (iy >= 0 && iy < H && ix >= 0 && ix < W) v = X[(((int64_t)c * Bn + b) * H + iy) * W + ix];
    col[e] = to_bf16(v);
  }
}

// dX[c][b][iy][ix] = sum over (ky, kx, oy, ox) with iy = s oy + ky - p, ix = s ox + kx - p
__global__ void col2im_kernel(const float* __restrict__ dcol, float* __restrict__ dX, int C, int Bn,
                              int H, int W, int k, int s, int p, int Ho, int Wo) {
  const int64_t tot = (int64_t)C * Bn * H * W;
  const int64_t Ncol = (int64_t)Bn * Ho * Wo;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int ix = (int)(e % W), iy = (int)((e / W) % H);
    const int b = (int)((e / ((int64_t)W * H)) % Bn), c = (int)(e / ((int64_t)W * H * Bn));
    float acc = 0.f;
    for (int ky = 0; ky < k; ++ky) {
      const int ty = iy + p - ky;
      if (ty < 0 || ty % s) continue;
      const int oy = ty / s;
      if (oy >= Ho) continue;
      for (int kx = 0; kx < k; ++kx) {
        const int tx = ix + p - kx;
        if (tx < 0 || tx % s) continue;
        const int ox = tx / s;
        if (ox >= Wo) continue;
        acc += dcol[(int64_t)((c * k + ky) * k + kx) * Ncol + ((int64_t)b * Ho + oy) * Wo + ox];
      }
    }
    dX[e] = acc;
  }
}

// Y[c][n] += bias[c]; leaky ReLU with negative slope `slope` (slope = 1: identity)
__global__ void bias_act_kernel(float* __restrict__ Y, const float* __restrict__ bias, int C,
                                int64_t n, float slope) {
  const int64_t tot = (int64_t)C * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const float v = Y[e] + bias[e / n];
    Y[e] = v > 0.f ? v : slope * v;
  }
}

// dZ = dY * act'(Y) in place (Y is the activated output; sign is preserved), and
// db[c] = sum_n dZ[c][n] (one block per channel).
__global__ void act_bwd_kernel(float* __restrict__ dY, const float* __restrict__ Y, int64_t n,
                               float slope, float* __restrict__ db) {
  __shared__ double red[16];
  const int c = blockIdx.x;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const int64_t e = (int64_t)c * n + i;
    const float g = Y[e] > 0.f ? dY[e] : slope * dY[e];
    dY[e] = g;
    acc += g;
  }
  const double t = block_sum_double(acc, red);
  if (threadIdx.x == 0) db[c] = (float)t;
}

// out[c][b][i] = in[c][b][i] + T[c][i]
__global__ void plane_bias_kernel(const float* __restrict__ in, const float* __restrict__ T,
                                  float* __restrict__ out, int64_t B, int64_t HW, int64_t total) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = e / (B * HW), i = e % HW;
    out[e] = in[e] + T[c * HW + i];
  }
}

// gT[c][i] = scale * sum_b dY[c][b][i]; consecutive threads take consecutive i (coalesced)
__global__ void plane_bias_grad_kernel(const float* __restrict__ dY, float* __restrict__ gT,
                                       int64_t B, int64_t HW, int64_t total, float scale) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = e / HW, i = e % HW;
    const float* src = dY + c * B * HW + i;
    double acc = 0.0;
    for (int64_t b = 0; b < B; ++b) acc += src[b * HW];
    gT[e] = (float)(acc * scale);
  }
}

__global__ void to_bf16_kernel(const float* __restrict__ x, __bf16* __restrict__ y, int64_t n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    y[e] = to_bf16(x[e]);
}

// nearest 2x: out[c][b][y][x] = in[c][b][y/2][x/2]
__global__ void up2_fwd_kernel(const float* __restrict__ in, float* __restrict__ out, int64_t CB,
                               int H, int W) {
  const int64_t tot = CB * 4 * H * W;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(e % (2 * W)), y = (int)((e / (2 * W)) % (2 * H));
    const int64_t cb = e / (4 * (int64_t)H * W);
    out[e] = in[(cb * H + y / 2) * W + x / 2];
  }
}

__global__ void up2_bwd_kernel(const float* __restrict__ dout, float* __restrict__ din, int64_t CB,
                               int H, int W) {
  const int64_t tot = CB * H * W;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(e % W), y = (int)((e / W) % H);
    const int64_t cb = e / ((int64_t)H * W);
    const float* d = dout + (cb * 2 * H + 2 * y) * (2 * W) + 2 * x;
    din[e] = d[0] + d[1] + d[2 * W] + d[2 * W + 1];
  }
}

// Per sample b (columns b*HW .. of channel-major [C][B][HW] tensors):
//   L_b = sum (y - t)^2 / sum t^2;  grad = 2 (y - t) / sum t^2 / B.  loss[b] written.
__global__ void rel_sq_loss_kernel(const float* __restrict__ y, const float* __restrict__ t,
                                   float* __restrict__ grad, float* __restrict__ loss, int C,
                                   int Bn, int HW) {
  __shared__ double red[16];
  const int b = blockIdx.x;
  double num = 0.0, den = 0.0;
  for (int i = threadIdx.x; i < C * HW; i += blockDim.x) {
    const int c = i / HW, q = i % HW;
    const int64_t e = ((int64_t)c * Bn + b) * HW + q;
    const double d = (double)y[e] - t[e];
    num += d * d;
    den += (double)t[e] * t[e];
  }
  num = block_sum_double(num, red);
  den = block_sum_double(den, red);
  const float inv = den > 0.0 ? (float)(1.0 / den) : 0.f;
  for (int i = threadIdx.x; i < C * HW; i += blockDim.x) {
    const int c = i / HW, q = i % HW;
    const int64_t e = ((int64_t)c * Bn + b) * HW + q;
    if (grad) grad[e] = 2.f * (y[e] - t[e]) * inv / (float)Bn;
  }
  if (threadIdx.x == 0) loss[b] = (float)(den > 0.0 ? num / den : 0.0);
}

// Energy-normalised squared error, pass 1: per sample b, work[b] = |y_b - t_b|^2 and
// work[Bn + b] = |t_b|^2.
__global__ void sq_terms_kernel(const float* __restrict__ y, const float* __restrict__ t,
                                double* __restrict__ work, int C, int Bn, int HW) {
  __shared__ double red[16];
  const int b = blockIdx.x;
  double num = 0.0, den = 0.0;
  for (int i = threadIdx.x; i < C * HW; i += blockDim.x) {
    const int c = i / HW, q = i % HW;
    const int64_t e = ((int64_t)c * Bn + b) * HW + q;
    const double d = (double)y[e] - t[e];
    num += d * d;
    den += (double)t[e] * t[e];
  }
  num = block_sum_double(num, red);
  den = block_sum_double(den, red);
  if (threadIdx.x == 0) {
    work[b] = num;
    work[Bn + b] = den;
  }
}

// Pass 2: E = mean_b |t_b|^2 (each block sums the Bn terms itself); loss[b] = work[b] / E and
// grad = 2 (y - t) / (E Bn), the gradient of the mean loss.
__global__ void energy_loss_kernel(const float* __restrict__ y, const float* __restrict__ t,
                                   const double* __restrict__ work, float* __restrict__ grad,
                                   float* __restrict__ loss, int Bn, int64_t total) {
  __shared__ double red[16];
  double s = 0.0;
  for (int b = threadIdx.x; b < Bn; b += blockDim.x) s += work[Bn + b];
  const double E = block_sum_double(s, red) / Bn;
  const float inv = E > 0.0 ? (float)(1.0 / E) : 0.f;
  if (blockIdx.x == 0)
    for (int b = threadIdx.x; b < Bn; b += blockDim.x) loss[b] = (float)(work[b] * inv);
  if (!grad) return;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x)
    grad[e] = 2.f * (y[e] - t[e]) * inv / (float)Bn;
}

// Adam on FP32 master weights; refreshes the BF16 copy.
__global__ void adam_kernel(float* __restrict__ w, __bf16* __restrict__ wb, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, int64_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2, float wd) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const float gr = g[e];
    const float mm = b1 * m[e] + (1.f - b1) * gr;
    const float vv = b2 * v[e] + (1.f - b2) * gr * gr;
    m[e] = mm;
    v[e] = vv;
    // decoupled decay (AdamW): the decay term does not pass through the moment estimates
    const float upd = lr * ((mm / c1) / (sqrtf(vv / c2) + eps) + wd * w[e]);
    w[e] -= upd;
    if (wb) wb[e] = to_bf16(w[e]);
  }
}

inline unsigned blocks_for(int64_t n, int bs = 256) {
  const int64_t b = (n + bs - 1) / bs;
  return (unsigned)(b < 16384 ? (b > 0 ? b : 1) : 16384);
}

}  // namespace

MAOMI_API int maomi_gemm_bf16(const void* A, const void* B, float* C, int64_t M, int64_t N,
                              int64_t K, int64_t sam, int64_t sak, int64_t sbk, int64_t sbn,
                              int64_t ldc, float alpha, float beta, maomi_stream_t stream) {
  if (!A || !B || !C || M < 1 || N < 1 || K < 1 || M > (1ll << 30) || N > (1ll << 30) ||
      K > (1ll << 30)) {
    maomi_set_error("maomi_gemm_bf16: bad arguments");
    return MAOMI_EINVAL;
  }
  if ((M + TM - 1) / TM > 65535) { maomi_set_error("maomi_gemm_bf16: M too large"); return MAOMI_EINVAL; }
  const bool ak = sak == 1, am = sam == 1, bk = sbk == 1, bn = sbn == 1;
  if ((ak || am) && (bk || bn)) {
    // one unit stride per operand: the tiled kernel; its leading dimensions
    const int64_t lda = ak ? sam : sak, ldb = bk ? sbn : sbk;
    const bool vec = ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) && lda % 8 == 0 &&
                     ldb % 8 == 0;
    if (M > 64)
      launch_tiled_layout<128>(ak, bk, vec, A, B, C, M, N, K, lda, ldb, ldc, alpha, beta,
                               (hipStream_t)stream);
    else
      launch_tiled_layout<64>(ak, bk, vec, A, B, C, M, N, K, lda, ldb, ldc, alpha, beta,
                              (hipStream_t)stream);
  } else {
    const dim3 grid((unsigned)((N + TN - 1) / TN), (unsigned)((M + TM - 1) / TM));
    hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(GT), 0, (hipStream_t)stream,
                       (const __bf16*)A, (const __bf16*)B, C, (int)M, (int)N, (int)K, sam, sak,
                       sbk, sbn, ldc, alpha, beta);
  }
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_im2col_bf16(const float* X, void* col, int64_t C, int64_t Bn, int64_t H,
                                int64_t W, int64_t k, int64_t s, int64_t p, maomi_stream_t stream) {
  if (!X || !col || C < 1 || Bn < 1 || H < 1 || W < 1 || k < 1 || s < 1 || p < 0) {
    maomi_set_error("maomi_im2col_bf16: bad arguments");
    return MAOMI_EINVAL;
  }
  const int64_t Ho = (H + 2 * p - k) / s + 1, Wo = (W + 2 * p - k) / s + 1;
  hipLaunchKernelGGL(im2col_kernel, dim3(blocks_for(C * k * k * Bn * Ho * Wo)), dim3(256), 0,
                     (hipStream_t)stream, X, (__bf16*)col, (int)C, (int)Bn, (int)H, (int)W, (int)k,
                     (int)s, (int)p, (int)Ho, (int)Wo);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_col2im(const float* dcol, float* dX, int64_t C, int64_t Bn, int64_t H,
                           int64_t W, int64_t k, int64_t s, int64_t p, maomi_stream_t stream) {
  if (!dcol || !dX || C < 1 || Bn < 1 || H < 1 || W < 1 || k < 1 || s < 1 || p < 0) {
    maomi_set_error("maomi_col2im: bad arguments");
    return MAOMI_EINVAL;
  }
  const int64_t Ho = (H + 2 * p - k) / s + 1, Wo = (W + 2 * p - k) / s + 1;
  hipLaunchKernelGGL(col2im_kernel, dim3(blocks_for(C * Bn * H * W)), dim3(256), 0,
                     (hipStream_t)stream, dcol, dX, (int)C, (int)Bn, (int)H, (int)W, (int)k, (int)s,
                     (int)p, (int)Ho, (int)Wo);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_bias_act(float* Y, const float* bias, int64_t C, int64_t n, float slope,
                             maomi_stream_t stream) {
  if (!Y || !bias || C < 1 || n < 1) { maomi_set_error("maomi_bias_act: bad arguments"); return MAOMI_EINVAL; }
  hipLaunchKernelGGL(bias_act_kernel, dim3(blocks_for(C * n)), dim3(256), 0, (hipStream_t)stream,
                     Y, bias, (int)C, n, slope);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_act_bwd(float* dY, const float* Y, int64_t C, int64_t n, float slope,
                            float* db, maomi_stream_t stream) {
  if (!dY || !Y || !db || C < 1 || n < 1) { maomi_set_error("maomi_act_bwd: bad arguments"); return MAOMI_EINVAL; }
  hipLaunchKernelGGL(act_bwd_kernel, dim3((unsigned)C), dim3(256), 0, (hipStream_t)stream, dY, Y, n,
                     slope, db);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_plane_bias(const float* in, const float* T, float* out, int64_t C, int64_t B,
                               int64_t HW, maomi_stream_t stream) {
  if (!in || !T || !out || C < 1 || B < 1 || HW < 1) { maomi_set_error("maomi_plane_bias: bad arguments"); return MAOMI_EINVAL; }
  hipLaunchKernelGGL(plane_bias_kernel, dim3(blocks_for(C * B * HW)), dim3(256), 0,
                     (hipStream_t)stream, in, T, out, B, HW, C * B * HW);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_plane_bias_grad(const float* dY, float* gT, int64_t C, int64_t B, int64_t HW,
                                    float scale, maomi_stream_t stream) {
  if (!dY || !gT || C < 1 || B < 1 || HW < 1) { maomi_set_error("maomi_plane_bias_grad: bad arguments"); return MAOMI_EINVAL; }
  hipLaunchKernelGGL(plane_bias_grad_kernel, dim3(blocks_for(C * HW)), dim3(256), 0,
                     (hipStream_t)stream, dY, gT, B, HW, C * HW, scale);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_to_bf16(const float* x, void* y, int64_t n, maomi_stream_t stream) {
  if (!x || !y || n < 1) { maomi_set_error("maomi_to_bf16: bad arguments"); return MAOMI_EINVAL; }
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, x,
                     (__bf16*)y, n);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_upsample2(const float* in, float* out, int64_t CB, int64_t H, int64_t W,
                              int64_t backward, maomi_stream_t stream) {
  if (!in || !out || CB < 1 || H < 1 || W < 1) { maomi_set_error("maomi_upsample2: bad arguments"); return MAOMI_EINVAL; }
  if (backward)
    hipLaunchKernelGGL(up2_bwd_kernel, dim3(blocks_for(CB * H * W)), dim3(256), 0,
                       (hipStream_t)stream, in, out, CB, (int)H, (int)W);
  else
    hipLaunchKernelGGL(up2_fwd_kernel, dim3(blocks_for(CB * 4 * H * W)), dim3(256), 0,
                       (hipStream_t)stream, in, out, CB, (int)H, (int)W);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_rel_sq_loss(const float* y, const float* t, float* grad, float* loss, int64_t C,
                                int64_t Bn, int64_t HW, maomi_stream_t stream) {
  if (!y || !t || !loss || C < 1 || Bn < 1 || HW < 1) { maomi_set_error("maomi_rel_sq_loss: bad arguments"); return MAOMI_EINVAL; }
  hipLaunchKernelGGL(rel_sq_loss_kernel, dim3((unsigned)Bn), dim3(256), 0, (hipStream_t)stream, y,
                     t, grad, loss, (int)C, (int)Bn, (int)HW);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_energy_sq_loss(const float* y, const float* t, float* grad, float* loss,
                                   double* work, int64_t C, int64_t Bn, int64_t HW,
                                   maomi_stream_t stream) {
  if (!y || !t || !loss || !work || C < 1 || Bn < 1 || HW < 1) { maomi_set_error("maomi_energy_sq_loss: bad arguments"); return MAOMI_EINVAL; }
  hipLaunchKernelGGL(sq_terms_kernel, dim3((unsigned)Bn), dim3(256), 0, (hipStream_t)stream, y, t,
                     work, (int)C, (int)Bn, (int)HW);
  MAOMI_CHECK_LAUNCH();
  const int64_t total = C * Bn * HW;
  hipLaunchKernelGGL(energy_loss_kernel, dim3(grad ? blocks_for(total) : 1u), dim3(256), 0,
                     (hipStream_t)stream, y, t, work, grad, loss, (int)Bn, total);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_adam(float* w, void* wb, const float* g, float* m, float* v, int64_t n,
                         float lr, float b1, float b2, float eps, int64_t step, float wd,
                         maomi_stream_t stream) {
  if (!w || !g || !m || !v || n < 1 || step < 1) { maomi_set_error("maomi_adam: bad arguments"); return MAOMI_EINVAL; }
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, w,
                     (__bf16*)wb, g, m, v, n, lr, b1, b2, eps, c1, c2, wd);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

// ---------------------------------------------------------------- input patches
namespace {
// out[c][p][y][x], c < zbins: mean over the pool x pool x (z-bin) voxels of
// (n - n0) * inv_std in the site window; channel zbins holds the site's depth scalar.
// Column sums of (n - n0) over each z bin: zsum[c][x][y] (FP64), one thread per column, the
// column's voxels read once in order.
__global__ void zbin_kernel(const float* __restrict__ index, int64_t XY, int Zv, int zbins,
                            double n0, double* __restrict__ zsum) {
  for (int64_t col = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; col < XY;
       col += (int64_t)gridDim.x * blockDim.x) {
    const float* v = index + col * Zv;
    int z = 0;
    for (int c = 0; c < zbins; ++c) {
      const int z1 = (int)((int64_t)(c + 1) * Zv / zbins);
      double acc = 0.0;
      for (; z < z1; ++z) acc += (double)v[z] - n0;
      zsum[(int64_t)c * XY + col] = acc;
    }
  }
}

// out[c][p][yy][xx]: the pool x pool block sum of zsum[c] at the site's window cell, over the
// block's voxel count and the index std; the depth channel after the bins.
__global__ void patch_kernel(const double* __restrict__ zsum, int64_t X, int64_t Y, int Zv,
                             const int32_t* __restrict__ wx, const int32_t* __restrict__ wy,
                             const float* __restrict__ depth, int P, int side, int pool, int zbins,
                             float inv_std, float* __restrict__ out) {
  const int S = side / pool;
  const int64_t plane = (int64_t)P * S * S;
  const int64_t tot = (int64_t)zbins * plane;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int xx = (int)(e % S), yy = (int)((e / S) % S);
    const int p = (int)((e / ((int64_t)S * S)) % P), c = (int)(e / plane);
    const int z0 = (int)((int64_t)c * Zv / zbins), z1 = (int)((int64_t)(c + 1) * Zv / zbins);
    const double* zc = zsum + (int64_t)c * X * Y;
    double acc = 0.0;
    for (int i = 0; i < pool; ++i) {
      const double* row = zc + (wx[p] + (int64_t)yy * pool + i) * Y + wy[p] + (int64_t)xx * pool;
      for (int j = 0; j < pool; ++j) acc += row[j];
    }
    out[e] = (float)(acc / ((double)pool * pool * (z1 - z0)) * inv_std);
    if (c == 0) out[(int64_t)zbins * plane + (int64_t)p * S * S + (int64_t)yy * S + xx] = depth[p];
  }
}
}  // namespace

MAOMI_API int maomi_patch_features(const float* index, int64_t X, int64_t Y, int64_t Zv,
                                   const int32_t* wx, const int32_t* wy, const float* depth,
                                   int64_t P, int64_t side, int64_t pool, int64_t zbins, double n0,
                                   double inv_std, double* zsum, float* out,
                                   maomi_stream_t stream) {
  if (!index || !wx || !wy || !depth || !zsum || !out || P < 1 || side < pool || pool < 1 ||
      zbins < 1 || zbins > Zv || side % pool) {
    maomi_set_error("maomi_patch_features: bad arguments");
    return MAOMI_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(zbin_kernel, dim3(blocks_for(X * Y)), dim3(256), 0, st, index, X * Y,
                     (int)Zv, (int)zbins, n0, zsum);
  MAOMI_CHECK_LAUNCH();
  const int64_t S = side / pool;
  hipLaunchKernelGGL(patch_kernel, dim3(blocks_for(zbins * P * S * S)), dim3(256), 0, st, zsum,
                     X, Y, (int)Zv, wx, wy, depth, (int)P, (int)side, (int)pool, (int)zbins,
                     (float)inv_std, out);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}
