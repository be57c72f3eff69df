// In-register radix-2 FFTs for power-of-two sizes up to 32, plus the
// two-level line decomposition used by the LDS-tiled 2-D passes.
//
// A line of length N = N1 * N2 is transformed in two register stages joined
// by one LDS exchange:
//   stage 1, job n1 in [0,N1): y[n1][k2] = W_N^{n1 k2} sum_n2 x[n1 + N1 n2] W_N2^{n2 k2}
//   stage 2, job k2 in [0,N2): X[k2 + N2 k1] = sum_n1 y[n1][k2] W_N1^{n1 k1}
// After stage 2 a job holds X at natural-order positions k2 + N2 k1, which is
// exactly the stage-1 input of a second transform with the factors swapped
// (N1' = N2, N2' = N1).  A forward/inverse pair along one axis therefore
// costs two LDS exchanges and one global read and write per element.
#pragma once
#include "common.hpp"

namespace fftk {

// cos(2 pi k / 32), sin(2 pi k / 32) for k in [0, 16), from double precision.
__device__ constexpr float kCos32[16] = {
    1.0f, 0.9807852506637573f, 0.9238795042037964f, 0.8314695954322815f,
    0.7071067690849304f, 0.5555702447891235f, 0.3826834261417389f, 0.19509032368659973f,
    0.0f, -0.19509032368659973f, -0.3826834261417389f, -0.5555702447891235f,
    -0.7071067690849304f, -0.8314695954322815f, -0.9238795042037964f, -0.9807852506637573f};
__device__ constexpr float kSin32[16] = {
    0.0f, 0.19509032368659973f, 0.3826834261417389f, 0.5555702447891235f,
    0.7071067690849304f, 0.8314695954322815f, 0.9238795042037964f, 0.9807852506637573f,
    1.0f, 0.9807852506637573f, 0.9238795042037964f, 0.8314695954322815f,
    0.7071067690849304f, 0.5555702447891235f, 0.3826834261417389f, 0.19509032368659973f};

__host__ __device__ constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }
__host__ __device__ constexpr int bitrev(int i, int bits) {
  int r = 0;
  for (int b = 0; b < bits; ++b) r = (r << 1) | ((i >> b) & 1);
  return r;
}

// In-place DFT of R points held in registers, natural order in and out.
// INV selects exp(+2 pi i nk/R); no normalisation.
template <int R, bool INV>
__device__ __forceinline__ void fft(float2 (&v)[R]) {
  static_assert(R >= 2 && R <= 32 && (R & (R - 1)) == 0, "R must be a power of two <= 32");
  constexpr int BITS = ilog2c(R);
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int j = bitrev(i, BITS);
    if (j > i) {
      const float2 t = v[i];
      v[i] = v[j];
      v[j] = t;
    }
  }
#pragma unroll
  for (int h = 1; h < R; h <<= 1) {
#pragma unroll
    for (int k = 0; k < R; k += 2 * h) {
#pragma unroll
      for (int j = 0; j < h; ++j) {
        const float2 a = v[k + j];
        const float2 b = v[k + j + h];
        float2 t;
        if (j == 0) {
          t = b;
        } else if (2 * j == h) {
          t = INV ? make_float2(-b.y, b.x) : make_float2(b.y, -b.x);
        } else {
          const int e = j * (32 / (2 * h));
          const float c = kCos32[e];
          const float s = INV ? kSin32[e] : -kSin32[e];
          t = make_float2(b.x * c - b.y * s, b.x * s + b.y * c);
        }
        v[k + j] = cadd(a, t);
        v[k + j + h] = csub(a, t);
      }
    }
  }
}

template <int N> struct Split;
template <> struct Split<64> { static constexpr int N1 = 8, N2 = 8; };
template <> struct Split<128> { static constexpr int N1 = 8, N2 = 16; };
template <> struct Split<256> { static constexpr int N1 = 16, N2 = 16; };
template <> struct Split<512> { static constexpr int N1 = 16, N2 = 32; };

__host__ __device__ constexpr int cmax(int a, int b) { return a > b ? a : b; }

// Tile geometry for a block of NT threads transforming L lines of length N.
// N = 512 (16 x 32) sizes the tile by the smaller factor, so that its 32-point stages fill the
// block; the 16-point stages then take two jobs per thread.
template <int N, int NT>
struct Tile {
  static constexpr int N1 = Split<N>::N1, N2 = Split<N>::N2;
  static constexpr int JMAX = cmax(N1, N2), JMIN = N1 < N2 ? N1 : N2;
  static constexpr int L = N >= 512 ? NT / JMIN : NT / JMAX;  // lines per tile
  static constexpr int LS = cmax(N1 * (N2 + 1), N2 * (N1 + 1)) | 1;  // odd line stride
  static constexpr int LDS2 = L * LS;                     // float2 slots of exchange buffer
  static_assert(N % L == 0, "tile must divide the field");
};

}  // namespace fftk
