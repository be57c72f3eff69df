// In-register radix-4 FFTs for power-of-two sizes up to 32, plus the two-level line
// decomposition used by the LDS-tiled 2-D passes.
//
// fft<R> is a decimation-in-time radix-4 recursion: R = 4 M points split into the four
// decimated sequences x[p + 4 m], each transformed by fft<M>, then combined as
//   a  = Y0 + w^2k Y2,  b  = Y0 - w^2k Y2,  c' = Y1 + w^2k Y3,  d' = Y1 - w^2k Y3,
//   X[k] = a + w^k c',  X[k + 2M] = a - w^k c',  X[k + M] = b + w^(k+M) d',  X[k + 3M] = b - w^(k+M) d'
// with w = exp(-+2 pi i / R) (w^M = -+i), i.e. four butterflies u +- t v per k.  A butterfly
// whose twiddle t = (c, s) is not a power of i runs in fused multiply-adds only, scaled by
// the larger of |c| and |s| (Linzer and Feig, "Modified FFTs for fused multiply-add
// architectures", Math. Comp. 60 (1993)): with r = s / c,
//   v' = (v.x - r v.y, v.y + r v.x),  u +- t v = u +- c v'
// (the roles of c and s exchanged when |s| > |c|) costs 6 instructions instead of 8 (a
// complex product and two complex additions).  A 16-point transform takes 148 instructions
// against 168 for the radix-2 form with the same trivial twiddles skipped; 32 points take 388
// against 456.  fft<32> runs the radix-2 form (fft_radix2 below); the sites kernels' 32-point
// lines (the 512 side) are transformed by lane pairs of 16-point transforms instead.
//
// Line decomposition, N = N1 * N2, two register stages joined by one LDS exchange:
//   stage 1, job n1 in [0,N1): y[n1][k2] = W_N^{n1 k2} sum_n2 x[n1 + N1 n2] W_N2^{n2 k2}
//   stage 2, job k2 in [0,N2): X[k2 + N2 k1] = sum_n1 y[n1][k2] W_N1^{n1 k1}
// After stage 2 a job holds X at natural-order positions k2 + N2 k1, which is exactly the
// stage-1 input of a second transform with the factors swapped (N1' = N2, N2' = N1).  A
// forward/inverse pair along one axis therefore costs two LDS exchanges and one global read
// and write per element.
#pragma once
#include "common.hpp"

namespace fftk {

// Twiddles exp(2 pi i e / 64) in the fused form of the butterflies: {m, r, form} with
// form 1: t = m (1, r) (|cos| >= |sin|, r = tan), form 2: t = m (r, 1) (r = cot); form 0 marks
// the powers of i, which need no multiplication.  The float pairs were chosen, within 4.5e-8
// rad of the exact angle, so that |t|^2 - 1 of a twiddle and of its conjugate (entry 64 - e,
// the same twiddle in the other direction) nearly cancel (sum below 3e-8): a forward and
// inverse transform then leave the energy of a field unchanged to first order, where plain
// rounding of (m, r) or of (cos, sin) biases every round trip of a split-step loop the same
// way (a 32-point round trip repeated 240 times drifted by -1.6e-5 in energy).
struct LfTwiddle {
  float m, r;
  int form;
};
constexpr LfTwiddle kLF[64] = {
    {0.0f, 0.0f, 0},
    {0.9951847195625305f, 0.09849143773317337f, 1},
    {0.9807852506637573f, 0.19891239702701569f, 1},
    {0.9569403529167175f, 0.3033466637134552f, 1},
    {0.9238795042037964f, 0.4142135977745056f, 1},
    {0.8819212913513184f, 0.5345110893249512f, 1},
    {0.8314695954322815f, 0.6681786775588989f, 1},
    {0.7730104327201843f, 0.8206788301467896f, 1},
    {0.7071067690849304f, 1.0f, 1},
    {0.7730104327201843f, 0.8206788301467896f, 2},
    {0.8314695954322815f, 0.6681786775588989f, 2},
    {0.8819212913513184f, 0.5345110893249512f, 2},
    {0.9238795042037964f, 0.4142135977745056f, 2},
    {0.9569403529167175f, 0.3033466637134552f, 2},
    {0.9807852506637573f, 0.19891239702701569f, 2},
    {0.9951847195625305f, 0.09849143773317337f, 2},
    {0.0f, 0.0f, 0},
    {0.9951847195625305f, -0.09849143773317337f, 2},
    {0.9807852506637573f, -0.19891239702701569f, 2},
    {0.9569403529167175f, -0.3033466637134552f, 2},
    {0.9238795042037964f, -0.4142135977745056f, 2},
    {0.8819212913513184f, -0.5345110893249512f, 2},
    {0.8314695954322815f, -0.6681786775588989f, 2},
    {0.7730104327201843f, -0.8206788301467896f, 2},
    {-0.7071067690849304f, -1.0f, 1},
    {-0.7730104327201843f, -0.8206788301467896f, 1},
    {-0.8314695954322815f, -0.6681786775588989f, 1},
    {-0.8819212913513184f, -0.5345110893249512f, 1},
    {-0.9238795042037964f, -0.4142135977745056f, 1},
    {-0.9569403529167175f, -0.3033466637134552f, 1},
    {-0.9807852506637573f, -0.19891239702701569f, 1},
    {-0.9951847195625305f, -0.09849143773317337f, 1},
    {0.0f, 0.0f, 0},
    {-0.9951847195625305f, 0.09849143773317337f, 1},
    {-0.9807852506637573f, 0.19891239702701569f, 1},
    {-0.9569403529167175f, 0.3033466637134552f, 1},
    {-0.9238795042037964f, 0.4142135977745056f, 1},
    {-0.8819212913513184f, 0.5345110893249512f, 1},
    {-0.8314695954322815f, 0.6681786775588989f, 1},
    {-0.7730104327201843f, 0.8206788301467896f, 1},
    {-0.7071067690849304f, 1.0f, 1},
    {-0.7730104327201843f, 0.8206788301467896f, 2},
    {-0.8314695954322815f, 0.6681786775588989f, 2},
    {-0.8819212913513184f, 0.5345110893249512f, 2},
    {-0.9238795042037964f, 0.4142135977745056f, 2},
    {-0.9569403529167175f, 0.3033466637134552f, 2},
    {-0.9807852506637573f, 0.19891239702701569f, 2},
    {-0.9951847195625305f, 0.09849143773317337f, 2},
    {0.0f, 0.0f, 0},
    {-0.9951847195625305f, -0.09849143773317337f, 2},
    {-0.9807852506637573f, -0.19891239702701569f, 2},
    {-0.9569403529167175f, -0.3033466637134552f, 2},
    {-0.9238795042037964f, -0.4142135977745056f, 2},
    {-0.8819212913513184f, -0.5345110893249512f, 2},
    {-0.8314695954322815f, -0.6681786775588989f, 2},
    {-0.7730104327201843f, -0.8206788301467896f, 2},
    {0.7071067690849304f, -1.0f, 1},
    {0.7730104327201843f, -0.8206788301467896f, 1},
    {0.8314695954322815f, -0.6681786775588989f, 1},
    {0.8819212913513184f, -0.5345110893249512f, 1},
    {0.9238795042037964f, -0.4142135977745056f, 1},
    {0.9569403529167175f, -0.3033466637134552f, 1},
    {0.9807852506637573f, -0.19891239702701569f, 1},
    {0.9951847195625305f, -0.09849143773317337f, 1}};

// u + t v and u - t v for t = exp(-+2 pi i E / R) (forward: minus), R dividing 64.
template <int R, int E, bool INV>
__host__ __device__ __forceinline__ void bfly(float2& u, float2& v) {
  static_assert(64 % R == 0, "twiddle sizes divide 64");
  constexpr int e = ((E % R + R) % R) * (64 / R);
  constexpr int ed = INV ? e : (64 - e) % 64;  // angle as a positive multiple of 2 pi / 64
  const float2 a = u, b = v;
  if constexpr (ed == 0) {
    u = make_float2(a.x + b.x, a.y + b.y);
    v = make_float2(a.x - b.x, a.y - b.y);
  } else if constexpr (ed == 32) {
    u = make_float2(a.x - b.x, a.y - b.y);
    v = make_float2(a.x + b.x, a.y + b.y);
  } else if constexpr (ed == 16) {  // t = i: t v = (-v.y, v.x)
    u = make_float2(a.x - b.y, a.y + b.x);
    v = make_float2(a.x + b.y, a.y - b.x);
  } else if constexpr (ed == 48) {  // t = -i: t v = (v.y, -v.x)
    u = make_float2(a.x + b.y, a.y - b.x);
    v = make_float2(a.x - b.y, a.y + b.x);
  } else {
    constexpr LfTwiddle t = kLF[ed];
    constexpr float m = t.m, r = t.r;
    float px, py;
    if constexpr (t.form == 1) {  // t v = m (v.x - r v.y, v.y + r v.x)
      px = fmaf(-r, b.y, b.x);
      py = fmaf(r, b.x, b.y);
    } else {  // t v = m (r v.x - v.y, v.x + r v.y)
      px = fmaf(r, b.x, -b.y);
      py = fmaf(r, b.y, b.x);
    }
    u = make_float2(fmaf(m, px, a.x), fmaf(m, py, a.y));
    v = make_float2(fmaf(-m, px, a.x), fmaf(-m, py, a.y));
  }
}

template <int R, bool INV>
__host__ __device__ __forceinline__ void fft(float2 (&v)[R]);

// cos(2 pi k / 32), sin(2 pi k / 32) for k in [0, 16), rounded from double precision.
constexpr float kCos32[16] = {
    1.0f, 0.9807852506637573f, 0.9238795042037964f, 0.8314695954322815f,
    0.7071067690849304f, 0.5555702447891235f, 0.3826834261417389f, 0.19509032368659973f,
    0.0f, -0.19509032368659973f, -0.3826834261417389f, -0.5555702447891235f,
    -0.7071067690849304f, -0.8314695954322815f, -0.9238795042037964f, -0.9807852506637573f};
constexpr float kSin32[16] = {
    0.0f, 0.19509032368659973f, 0.3826834261417389f, 0.5555702447891235f,
    0.7071067690849304f, 0.8314695954322815f, 0.9238795042037964f, 0.9807852506637573f,
    1.0f, 0.9807852506637573f, 0.9238795042037964f, 0.8314695954322815f,
    0.7071067690849304f, 0.5555702447891235f, 0.3826834261417389f, 0.19509032368659973f};

__host__ __device__ constexpr int bitrev(int i, int bits) {
  int r = 0;
  for (int b = 0; b < bits; ++b) r = (r << 1) | ((i >> b) & 1);
  return r;
}

// Radix-2 decimation in time with complex products by the rounded (cos, sin) pairs, for the
// 32-point lines.  The fused radix-4 form has the same accuracy per transform, but the moduli
// of its float twiddle pairs err low on balance: 240 forward/inverse round trips of 32 points
// lose 8.5e-6 of the energy against 3.6e-6 here, and a 60-slice field of side 512 (two round
// trips per slice) drifts measurably from the reference's float32 path.
template <int R, bool INV>
__host__ __device__ __forceinline__ void fft_radix2(float2 (&v)[R]) {
  constexpr int BITS = R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : R == 16 ? 4 : 5;
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
        v[k + j] = make_float2(a.x + t.x, a.y + t.y);
        v[k + j + h] = make_float2(a.x - t.x, a.y - t.y);
      }
    }
  }
}

// Radix-4 combination of the four decimated M-point transforms for output index k and the
// following ones.
template <int M, int K, bool INV>
__host__ __device__ __forceinline__ void combine4(float2 (&y)[4][M], float2 (&v)[4 * M]) {
  if constexpr (K < M) {
    constexpr int R = 4 * M;
    float2 a = y[0][K], b = y[2][K], c = y[1][K], d = y[3][K];
    bfly<R, 2 * K, INV>(a, b);  // a = Y0 + w^2k Y2, b = Y0 - w^2k Y2
    bfly<R, 2 * K, INV>(c, d);  // c = Y1 + w^2k Y3, d = Y1 - w^2k Y3
    bfly<R, K, INV>(a, c);      // X[k], X[k + 2M]
    bfly<R, K + M, INV>(b, d);  // X[k + M], X[k + 3M]
    v[K] = a;
    v[K + 2 * M] = c;
    v[K + M] = b;
    v[K + 3 * M] = d;
    combine4<M, K + 1, INV>(y, v);
  }
}

// In-place DFT of R points held in registers, natural order in and out.
// INV selects exp(+2 pi i nk/R); no normalisation.
template <int R, bool INV>
__host__ __device__ __forceinline__ void fft(float2 (&v)[R]) {
  static_assert(R >= 1 && R <= 32 && (R & (R - 1)) == 0, "R must be a power of two <= 32");
  if constexpr (R == 2) {
    bfly<2, 0, INV>(v[0], v[1]);
  } else if constexpr (R == 32) {
    fft_radix2<R, INV>(v);
  } else if constexpr (R >= 4) {
    constexpr int M = R / 4;
    float2 y[4][M];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
#pragma unroll
      for (int m = 0; m < M; ++m) y[p][m] = v[p + 4 * m];
      if constexpr (M > 1) fft<M, INV>(y[p]);
    }
    combine4<M, 0, INV>(y, v);
  }
}

template <int N> struct Split;
template <> struct Split<64> { static constexpr int N1 = 8, N2 = 8; };
template <> struct Split<128> { static constexpr int N1 = 8, N2 = 16; };
template <> struct Split<256> { static constexpr int N1 = 16, N2 = 16; };
template <> struct Split<512> { static constexpr int N1 = 32, N2 = 16; };

__host__ __device__ constexpr int cmax(int a, int b) { return a > b ? a : b; }
__host__ __device__ constexpr int cmin(int a, int b) { return a < b ? a : b; }

// Tile geometry for a block of NT threads transforming L lines of length N.  Up to 256 points
// L is sized by the smaller factor, so that both stages have at least NT jobs; the stage with
// the larger factor then gives each thread J1 (stage 1: N1 jobs per line, N2-point transforms)
// or J2 (stage 2: N2 jobs per line, N1-point transforms) jobs.  The 512 side (32 x 16) sizes L
// by the larger factor instead: stage 1 has NT jobs and stage 2 NT / 2 (A2 < NT: the first
// A2 / 64 waves transform the 32-point lines, the others go on to the next barrier), and the
// tile is half as tall, so that two exchange buffers fit in LDS.
template <int N, int NT>
struct Tile {
  static constexpr int N1 = Split<N>::N1, N2 = Split<N>::N2;
  static constexpr int L = NT / (N >= 512 ? cmax(N1, N2) : cmin(N1, N2));  // lines per tile
  static constexpr int A2 = L * N2;                        // stage-2 jobs of a tile
  static constexpr int J1 = L * N1 / NT, J2 = cmax(1, A2 / NT);  // jobs per thread, 1 or 2
  // Exchange layouts of a line: stage 1 writes y[n1][k2] at S1(n1) + k2, stage 2 reads it and
  // writes element j of job k2 at k2 KS2 + C2(j).  The 512 side's 32-point lines are
  // transformed by lane pairs, each lane holding one half (n1 or j below or above N1 / 2); its
  // layouts keep the two halves in two planes HP float2 apart, HP = 8 mod 16, so that the two
  // lanes of a pair reach different LDS banks.
  static constexpr bool SPLIT = N >= 512;
  static constexpr int H1 = N1 / 2;
  static constexpr int HP = SPLIT ? H1 * (N2 + 1) + 8 : 0;
  static constexpr int KS2 = SPLIT ? H1 + 1 : N1 + 1;
  __host__ __device__ static constexpr int S1(int n1) {
    return SPLIT ? (n1 >= H1 ? HP : 0) + (n1 % H1) * (N2 + 1) : n1 * (N2 + 1);
  }
  __host__ __device__ static constexpr int C2(int j) {
    return SPLIT ? (j >= H1 ? HP : 0) + j % H1 : j;
  }
  static constexpr int LS =  // odd line stride
      (SPLIT ? HP + cmax(H1 * (N2 + 1), N2 * KS2) : cmax(N1 * (N2 + 1), N2 * (N1 + 1))) | 1;
  static constexpr int LDS2 = L * LS;                     // float2 slots of exchange buffer
  static_assert(N % L == 0, "tile must divide the field");
  static_assert(J1 * NT == L * N1, "whole stage-1 jobs per thread");
  static_assert(J2 * NT == A2 || (A2 < NT && A2 % 64 == 0), "stage 2: whole jobs or waves");
};

}  // namespace fftk
