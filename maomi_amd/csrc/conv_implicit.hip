// Implicit-GEMM convolution for the surrogate's inference (no counterpart in the reference).
//
// Activations are channel-last BF16, X[b][y][x][c] with C a multiple of 8, and the weights are
// W[co][ky][kx][c] BF16, so that
//   Y[n][co] = act(bias[co] + sum_k X(n, k) W[co][k]),  n = (b, oy, ox),  k = (ky, kx, c),
//   X(n, k) = X[b][iy >> up][ix >> up][c],  iy = s oy - p + ky,  ix = s ox - p + kx,
// zero where (iy, ix) falls outside the (optionally 2x nearest-upsampled) input.  Eight
// consecutive k of one tap are eight consecutive channels, so each operand run is one 16-byte
// load straight from the activation tensor: no im2col buffer is written or read back, the
// upsampling of the decoder is a shift of the read coordinates, and bias plus leaky ReLU are
// applied to the accumulators before the single store (BF16 channel-last for the next layer,
// or FP32 channel-major [co][b][oy][ox] for the network's output planes).
//
// Tiles: BM = 128 output positions x BN output channels per 256-thread block (2 x 2 waves of
// 64 x BN/2), K in slabs of 32 with the next slab's runs loaded into registers while the MFMAs
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation) consume the current one from LDS.  Each
// thread's output positions are fixed for the whole K loop, so their (b, oy, ox) are decoded
// once; the tap and channel of its k run advance incrementally from slab to slab.
#include "common.hpp"
#include "maomi_amd.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int CBM = 128, CSK = 32, CSKP = CSK + 8, CT = 256;

struct ConvGeom {
  int N, K, C, Hs, Ws, up, KW, s, p, Ho, Wo, Co;
};

// One thread's k-run position: channel ci of tap (ky, kx), advanced by 32 k per slab.
struct TapRun {
  int ci, kx, ky;
  __device__ __forceinline__ void advance(const ConvGeom& g) {
    ci += CSK;
    while (ci >= g.C) {
      ci -= g.C;
      if (++kx == g.KW) {
        kx = 0;
        ++ky;
      }
    }
  }
};

template <int BN, bool OUT_BF16>
__global__ __launch_bounds__(CT) void conv_implicit_kernel(const uint4* __restrict__ Xv,
                                                           const unsigned short* __restrict__ Wt,
                                                           const float* __restrict__ bias,
                                                           void* __restrict__ Y, ConvGeom g,
                                                           float slope) {
  constexpr int RA = CBM * CSK / 8 / CT, RB = BN * CSK / 8 / CT > 0 ? BN * CSK / 8 / CT : 1;
  constexpr int FM = CBM / 32, FN = BN / 32;
  __shared__ __attribute__((aligned(16))) unsigned short As[CBM][CSKP];
  __shared__ __attribute__((aligned(16))) unsigned short Bs[BN][CSKP];
  const int m0 = blockIdx.x * CBM, n0 = blockIdx.y * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * (CBM / 2), wn = (wave & 1) * (BN / 2);
  const int kv = (tid % (CSK / 8)) * 8;  // this thread's k offset inside a slab (A and B)
  const unsigned short* Xs = reinterpret_cast<const unsigned short*>(Xv);

  // output positions of this thread's A runs, decoded once
  int rb[RA], ry[RA], rx[RA];
  bool rlive[RA];
#pragma unroll
  for (int r = 0; r < RA; ++r) {
    const int n = m0 + (tid + r * CT) / (CSK / 8);
    rlive[r] = n < g.N;
    const int nn = rlive[r] ? n : 0;
    rx[r] = nn % g.Wo;
    ry[r] = (nn / g.Wo) % g.Ho;
    rb[r] = nn / (g.Wo * g.Ho);
  }
  TapRun tap{kv % g.C, (kv / g.C) % g.KW, kv / g.C / g.KW};
  const int Hi = g.Hs << g.up, Wi = g.Ws << g.up;
  uint4 ra[RA], rbv[RB];

  auto fetch = [&](int k0) {
    const bool kin = k0 + kv < g.K;
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      const int iy = ry[r] * g.s - g.p + tap.ky, ix = rx[r] * g.s - g.p + tap.kx;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (kin && rlive[r] && iy >= 0 && iy < Hi && ix >= 0 && ix < Wi) {
        const int64_t e = (((int64_t)rb[r] * g.Hs + (iy >> g.up)) * g.Ws + (ix >> g.up)) * g.C +
                          tap.ci;
        v = *reinterpret_cast<const uint4*>(Xs + e);
      }
      ra[r] = v;
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int v = tid + r * CT;
      uint4 w = make_uint4(0u, 0u, 0u, 0u);
      if (v < BN * CSK / 8) {
        const int co = n0 + v / (CSK / 8), gk = k0 + kv;
        if (co < g.Co && gk < g.K) w = *reinterpret_cast<const uint4*>(Wt + (int64_t)co * g.K + gk);
      }
      rbv[r] = w;
    }
  };
  auto stage = [&]() {
#pragma unroll
    for (int r = 0; r < RA; ++r)
      *reinterpret_cast<uint4*>(&As[(tid + r * CT) / (CSK / 8)][kv]) = ra[r];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int v = tid + r * CT;
      if (v < BN * CSK / 8) *reinterpret_cast<uint4*>(&Bs[v / (CSK / 8)][kv]) = rbv[r];
    }
  };

  f32x4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  fetch(0);
  const int row = lane & 15, kq = (lane >> 4) * 8;
  for (int k0 = 0; k0 < g.K; k0 += CSK) {
    stage();
    __syncthreads();
    if (k0 + CSK < g.K) {
      tap.advance(g);
      fetch(k0 + CSK);
    }
    bf16x8 af[FM], bfr[FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
      af[i] = __builtin_bit_cast(bf16x8,
                                 *reinterpret_cast<const uint4*>(&As[wm + 16 * i + row][kq]));
#pragma unroll
    for (int j = 0; j < FN; ++j)
      bfr[j] = __builtin_bit_cast(bf16x8,
                                  *reinterpret_cast<const uint4*>(&Bs[wn + 16 * j + row][kq]));
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // D[4 (lane / 16) + t][lane % 16]: output position m0 + wm + 16 i + 4 (lane / 16) + t,
  // channel n0 + wn + 16 j + lane % 16
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int co = n0 + wn + 16 * j + (lane & 15);
      if (co >= g.Co) continue;
      const float b = bias[co];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int n = m0 + wm + 16 * i + 4 * (lane >> 4) + t;
        if (n >= g.N) continue;
        float v = acc[i][j][t] + b;
        v = v > 0.f ? v : slope * v;
        if (OUT_BF16)
          reinterpret_cast<__bf16*>(Y)[(int64_t)n * g.Co + co] = (__bf16)v;
        else
          reinterpret_cast<float*>(Y)[(int64_t)co * g.N + n] = v;
      }
    }
}

// [C][B][H][W] FP32 -> [B][H][W][Cp] BF16, channels C..Cp-1 zero
__global__ void to_channel_last_kernel(const float* __restrict__ in, int C, int Bn, int HW, int Cp,
                                       __bf16* __restrict__ out) {
  const int64_t tot = (int64_t)Bn * HW * Cp;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % Cp);
    const int64_t bp = e / Cp;  // b * HW + pixel
    const int64_t b = bp / HW, px = bp - b * HW;
    out[e] = c < C ? (__bf16)in[((int64_t)c * Bn + b) * HW + px] : (__bf16)0.f;
  }
}

template <int BN, bool OUT_BF16>
void launch_conv(const void* X, const void* W, const float* bias, void* Y, const ConvGeom& g,
                 float slope, hipStream_t st) {
  const dim3 grid((unsigned)((g.N + CBM - 1) / CBM), (unsigned)((g.Co + BN - 1) / BN));
  hipLaunchKernelGGL((conv_implicit_kernel<BN, OUT_BF16>), grid, dim3(CT), 0, st,
                     (const uint4*)X, (const unsigned short*)W, bias, Y, g, slope);
}

}  // namespace

MAOMI_API int maomi_conv_bf16_fwd(const void* X, int64_t Bn, int64_t Hs, int64_t Ws, int64_t C,
                                  int64_t upsample, const void* W, const float* bias, int64_t Co,
                                  int64_t k, int64_t s, int64_t p, float slope, void* Y,
                                  int64_t out_bf16, maomi_stream_t stream) {
  if (!X || !W || !bias || !Y || Bn < 1 || Hs < 1 || Ws < 1 || C < 8 || C % 8 || Co < 1 ||
      k < 1 || s < 1 || p < 0 || (upsample != 0 && upsample != 1)) {
    maomi_set_error("maomi_conv_bf16_fwd: bad arguments (C a multiple of 8, upsample 0 or 1)");
    return MAOMI_EINVAL;
  }
  const int64_t Hi = Hs << upsample, Wi = Ws << upsample;
  const int64_t Ho = (Hi + 2 * p - k) / s + 1, Wo = (Wi + 2 * p - k) / s + 1;
  const int64_t N = Bn * Ho * Wo, K = k * k * C;
  if (Ho < 1 || Wo < 1 || N > INT32_MAX || K > INT32_MAX || Bn * Hs * Ws * C > INT32_MAX * 8LL) {
    maomi_set_error("maomi_conv_bf16_fwd: output or operand sizes out of range");
    return MAOMI_EINVAL;
  }
  const ConvGeom g{(int)N, (int)K, (int)C, (int)Hs, (int)Ws, (int)upsample, (int)k, (int)s,
                   (int)p, (int)Ho, (int)Wo, (int)Co};
  hipStream_t st = (hipStream_t)stream;
  const int bn = Co % 128 == 0 ? 128 : (Co % 64 == 0 || Co > 32 ? 64 : 32);
  if (out_bf16) {
    if (bn == 128) launch_conv<128, true>(X, W, bias, Y, g, slope, st);
    else if (bn == 64) launch_conv<64, true>(X, W, bias, Y, g, slope, st);
    else launch_conv<32, true>(X, W, bias, Y, g, slope, st);
  } else {
    if (bn == 128) launch_conv<128, false>(X, W, bias, Y, g, slope, st);
    else if (bn == 64) launch_conv<64, false>(X, W, bias, Y, g, slope, st);
    else launch_conv<32, false>(X, W, bias, Y, g, slope, st);
  }
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_to_channel_last_bf16(const float* in, int64_t C, int64_t Bn, int64_t HW,
                                         int64_t Cp, void* out, maomi_stream_t stream) {
  if (!in || !out || C < 1 || Bn < 1 || HW < 1 || Cp < C) {
    maomi_set_error("maomi_to_channel_last_bf16: bad arguments");
    return MAOMI_EINVAL;
  }
  const int64_t tot = Bn * HW * Cp;
  const unsigned blocks = (unsigned)((tot + 255) / 256 < 65535 ? (tot + 255) / 256 : 65535);
  hipLaunchKernelGGL(to_channel_last_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, in,
                     (int)C, (int)Bn, (int)HW, (int)Cp, (__bf16*)out);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}
