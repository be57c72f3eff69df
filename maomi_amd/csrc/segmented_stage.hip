// Reference-geometry optical stage: per site, three split-step segments
//   A (n x n, phase A, saves planes >= cut cropped to n2: top tail),
//   B (n2 x n2, phase B, saves focal planes cropped to the focal window),
//   C (n x n, phase C, the B field zero-padded back to n: bottom tail),
// with the focal |u|^(2 order) resample, the trapezoid tail accumulations,
// their profiles and weights, all inside one persistent launch (one block per
// site at a time, field in a block-private workspace).
//
// Reference: OpticalPlan.evaluate (maomi/backends/cuda/batched.py:170-235),
// _stage/_screens (148-168), gather_screens and resample_power
// (optical_kernels.cu:3-38), the tail block of optics.propagate
// (maomi/optics.py:318-340).  Field sides are arbitrary even sizes: every line
// transform is a two-level direct DFT (N = N1 N2) run from LDS.
#include "common.hpp"
#include "fft_reg.hpp"
#include "maomi_amd.h"
#include "screen_trig.hpp"

// Dynamic LDS of the segmented kernel; the line transforms address it by offset.
extern __shared__ float2 seg_lds[];

namespace {

constexpr int NT = 256;
// Tile loads are issued in batches of kBatch elements per thread, each batch loaded into
// registers before any of it is stored, so a batch waits out one memory latency instead of
// one per element (larger batches run the kernel out of registers).
constexpr int kBatch = 8;

// Workgroup barrier that orders LDS only: global loads and stores stay in flight.  Full
// __syncthreads() is kept where one thread reads global workspace another thread wrote.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

struct Seg {
  const float* phase;    // [Z][H][W] phase in radians
  int64_t H, W, Z;       // volume dims; Z = screens
  int n, steps;          // field side, propagation steps
  int n1, n2f;           // N = n1 * n2f for the line DFT
  const float2* q;       // q tables, n entries each (FFT order, 1/n folded in)
  const int32_t* qidx;   // [steps] table index per step
  const float* edge;     // [n] edge window
  const float2* tw;      // [n] W_n^j
};

struct SegArgs {
  Seg seg[3];
  const float2* base;
  const int32_t* site_x;
  const int32_t* site_y;
  const int32_t* order_list;
  int64_t P;
  int cut, tails, order;
  int n, n2, coff;           // coff = (n - n2) / 2
  int fx, fy, foff_x, foff_y;
  int X, Y, TT;              // focal output X x Y, tail output TT x TT
  const float* mfx;          // [X][fx]
  const float* mfy;          // [Y][fy]
  const float* mt;           // [TT][n2]
  const int32_t* bfx;        // [X][2] band (first column, span) of mfx rows, or null
  const int32_t* bfy;        // [Y][2]
  const int32_t* bt;         // [TT][2]
  const float* ax_focus;     // [b]
  const float* ax_top;       // [a + 1 - cut]
  const float* ax_bot;       // [c + 1]
  float ss2;                 // ss^2 applied to tail intensities
  float focal_scale;         // ss^4 r (1e6 lambda)^1.5 / (pi nr)
  float tail_scale;          // ss^2 applied after the tail resample
  double prop_um;
  float* psf;                // [P][X][Y][b]
  float* top_conv;           // [P][TT][TT]
  float* top_prof;           // [P][a+1-cut]
  double* top_w;             // [P]
  float* bot_conv;
  float* bot_prof;           // [P][c+1]
  double* bot_w;
  float2* work;              // per block: 2 n^2 complex + n2^2 + fx fy + X max(fy,n2) floats
  int64_t work_stride;       // float2 elements per block
  int L;                     // lines per LDS tile
  int nb;                    // line stride of the scratch buffer B (>= n, see fast_stage1)
};

// acc + x * w as four FMAs (a multiply-add chain per output of the direct DFTs)
__device__ __forceinline__ float2 cmac(float2 acc, float2 x, float2 w) {
  acc.x = fmaf(x.x, w.x, acc.x);
  acc.x = fmaf(-x.y, w.y, acc.x);
  acc.y = fmaf(x.x, w.y, acc.y);
  acc.y = fmaf(x.y, w.x, acc.y);
  return acc;
}

__device__ __forceinline__ float2 twid(const float2* T, int j, bool inv) {
  const float2 w = T[j];
  return inv ? make_float2(w.x, -w.y) : w;
}

// Transform nl lines of length N held in A (natural order) into A, via B.
__device__ void line_dft(float2* A, float2* B, int nl, int N, int N1, int N2, const float2* T,
                         bool inv) {
  const int s1 = N / N2, s2 = N / N1;
  for (int it = threadIdx.x; it < nl * N; it += NT) {
    const int l = it / N, r = it % N, n1 = r / N2, k2 = r % N2;
    const float2* a = A + l * N + n1;
    float2 acc = make_float2(0.f, 0.f);
    for (int m = 0; m < N2; ++m) acc = cadd(acc, cmul(a[N1 * m], twid(T, ((m * k2) % N2) * s1, inv)));
    B[l * N + n1 * N2 + k2] = cmul(acc, twid(T, (n1 * k2) % N, inv));
  }
  lds_barrier();
  for (int it = threadIdx.x; it < nl * N; it += NT) {
    const int l = it / N, k = it % N, k2 = k % N2, k1 = k / N2;
    const float2* b = B + l * N + k2;
    float2 acc = make_float2(0.f, 0.f);
    for (int m = 0; m < N1; ++m) acc = cadd(acc, cmul(b[m * N2], twid(T, ((m * k1) % N1) * s2, inv)));
    A[l * N + k] = acc;
  }
  lds_barrier();
}

__host__ __device__ constexpr int smallest_factor(int n, int f = 2) {
  return f * f > n ? n : (n % f == 0 ? f : smallest_factor(n, f + 1));
}

// Twiddle tables of the small DFTs hold W_R^c for c <= R / 2 only: W_R^(R - c) is conj W_R^c,
// which halves the registers a table takes (its sign folds into the products).
template <int R>
__device__ __forceinline__ float2 wtab(const float2 (&w)[R / 2 + 1], int i) {
  return i <= R / 2 ? w[i] : make_float2(w[R - i].x, -w[R - i].y);
}

// D-point DFT Y[k] = sum_m y[m] W^(m k) with W^j = W_(D SW)^(j SW) from the half table w (the
// transform's sign already in w).  The terms m and D - m share W^(m k) and its conjugate, so with S_m = y_m + y_(D-m) and
// D_m = y_m - y_(D-m) (m < D/2),
//   Y[k]     = y_0 + sum_m S_m c_mk + i sum_m D_m s_mk  (+ y_(D/2) (-1)^k for even D),
//   Y[D - k] = the same with s_mk negated,
// so each pair of outputs costs four real h-term sums (h = (D - 1) / 2): 66 operations for 7
// points against 168 for the direct sums.
template <int D, int SW>
__device__ __forceinline__ void sym_dft(const float2 (&y)[D], const float2 (&w)[D * SW / 2 + 1],
                                        float2 (&Y)[D]) {
  if constexpr (D == 1) {
    Y[0] = y[0];
  } else if constexpr (D == 2) {
    Y[0] = cadd(y[0], y[1]);
    Y[1] = csub(y[0], y[1]);
  } else {
    constexpr int H = (D - 1) / 2;
    constexpr bool EVEN = D % 2 == 0;
    float2 S[H + 1], Df[H + 1];
    float2 y0 = y[0], yh = EVEN ? y[D / 2] : make_float2(0.f, 0.f);
    float2 X0 = y0, XH = y0;
#pragma unroll
    for (int m = 1; m <= H; ++m) {
      S[m] = cadd(y[m], y[D - m]);
      Df[m] = csub(y[m], y[D - m]);
      X0 = cadd(X0, S[m]);
      if constexpr (EVEN) XH = (m & 1) ? csub(XH, S[m]) : cadd(XH, S[m]);
    }
    if constexpr (EVEN) {
      X0 = cadd(X0, yh);
      XH = ((D / 2) & 1) ? csub(XH, yh) : cadd(XH, yh);
      Y[D / 2] = XH;
    }
    Y[0] = X0;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      float ar = y0.x, ai = y0.y, br = 0.f, bi = 0.f;
      if constexpr (EVEN) {
        ar = (k & 1) ? ar - yh.x : ar + yh.x;
        ai = (k & 1) ? ai - yh.y : ai + yh.y;
      }
#pragma unroll
      for (int m = 1; m <= H; ++m) {
        const float2 t = wtab<D * SW>(w, ((m * k) % D) * SW);
        ar = fmaf(S[m].x, t.x, ar);
        ai = fmaf(S[m].y, t.x, ai);
        br = fmaf(Df[m].y, t.y, br);
        bi = fmaf(Df[m].x, t.y, bi);
      }
      Y[k] = make_float2(ar - br, ai + bi);
      Y[D - k] = make_float2(ar + br, ai - bi);
    }
  }
}

// X[k] = sum_m x[m] W_R^(m k) with w[c] = W_R^c.  A composite R = P Q (P its smallest prime
// factor) runs as P-point DFTs over m = Q n1 + n2, the twiddle W_R^(n2 k1), then Q-point DFTs,
// X[k1 + P k2]; every level is a sym_dft.
template <int R>
__device__ __forceinline__ void small_dft(const float2 (&x)[R], const float2 (&w)[R / 2 + 1],
                                          float2 (&X)[R]) {
  constexpr int P = smallest_factor(R), Q = R / P;
  if constexpr (P == R) {
    sym_dft<R, 1>(x, w, X);
  } else {
    float2 y[R];  // y[n2 P + k1]
#pragma unroll
    for (int n2 = 0; n2 < Q; ++n2) {
      float2 t[P], u[P];
#pragma unroll
      for (int n1 = 0; n1 < P; ++n1) t[n1] = x[Q * n1 + n2];
      sym_dft<P, Q>(t, w, u);
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1)
        y[n2 * P + k1] = (n2 * k1) % R == 0 ? u[k1] : cmul(u[k1], wtab<R>(w, (n2 * k1) % R));
    }
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
      float2 t[Q], u[Q];
#pragma unroll
      for (int n2 = 0; n2 < Q; ++n2) t[n2] = y[n2 * P + k1];
      sym_dft<Q, P>(t, w, u);
#pragma unroll
      for (int k2 = 0; k2 < Q; ++k2) X[k1 + P * k2] = u[k2];
    }
  }
}

// Fast line transform for N = R1 * R2 (the split of fast_split: R2 <= 32, R1 <= 64): stage 1
// is an R2-point DFT per (line, n1) job (a register FFT when R2 is a power of two, else the
// factorised small_dft) followed by the twiddle W_N^(n1 k2); stage 2 an R1-point DFT per
// (line, k2) job in registers (R1 <= 32) or per output from LDS (wide_stage2).
// X[k2 + R2 k1] lands in natural order.  Lines live at LDS offsets a_off (in and out, line
// stride N) and b_off (scratch); the twiddle table W_N^j at t_off, and just below it, ending
// at t_off, the row twiddles W_N^(n1 k2) at n1 (R2 + 1) + k2 (the segment loads both).  In
// the scratch buffer,
// element (n1, k2) of a line sits at n1 (R2 + 1) + k2 and lines are pad_stride(N) apart: the
// padding keeps the stage-1 stores of neighbouring jobs (R2 apart otherwise) off the same
// LDS banks.
__host__ __device__ inline int pad_stride(int N, int R2) { return N + N / R2; }

// R2 of the fast split of N, 0 when there is none (the generic two-level DFT runs):
// the power-of-two part (<= 16) when the rest is at most 32 points; else the largest factor
// R2 <= 24 leaving at most 32; else the power-of-two part when the rest is at most 64.
__host__ __device__ inline int fast_split(int N) {
  int r2 = 1;
  while (r2 < 16 && N % (2 * r2) == 0) r2 *= 2;
  if (r2 >= 2 && N / r2 <= 32) return r2;
  for (int f = 24; f >= 2; --f)
    if (N % f == 0 && N / f <= 32) return f;
  return (r2 >= 2 && N / r2 <= 64) ? r2 : 0;
}

template <int R2, bool INV>
__device__ __noinline__ void fast_stage1(int a_off, int b_off, int t_off, int nl, int N) {
  constexpr bool POW2 = (R2 & (R2 - 1)) == 0;
  const int R1 = N / R2, NB = pad_stride(N, R2);
  const float2* T = seg_lds + t_off;
  float2 wr[POW2 ? 1 : R2 / 2 + 1];  // W_R2^c = W_N^(c R1), c <= R2 / 2, for the factorised DFT
  if constexpr (!POW2) {
#pragma unroll
    for (int c = 0; c <= R2 / 2; ++c) wr[c] = twid(T, c * R1, INV);
  }
  for (int job = threadIdx.x; job < nl * R1; job += NT) {
    const int l = job / R1, n1 = job - l * R1;
    const float2* a = seg_lds + a_off + l * N + n1;
    float2 v[R2], w[R2];
#pragma unroll
    for (int m = 0; m < R2; ++m) v[m] = a[R1 * m];
    // W_N^(n1 k2) from row n1 of the row twiddles (consecutive jobs on different banks); read
    // before any store
    const float2* tr = T - NB + n1 * (R2 + 1);
#pragma unroll
    for (int k2 = 1; k2 < R2; ++k2) w[k2] = twid(tr, k2, INV);
    if constexpr (POW2) {
      fftk::fft<R2, INV>(v);
    } else {
      float2 X[R2];
      small_dft<R2>(v, wr, X);
#pragma unroll
      for (int k = 0; k < R2; ++k) v[k] = X[k];
    }
    float2* b = seg_lds + b_off + l * NB + n1 * (R2 + 1);
    b[0] = v[0];
#pragma unroll
    for (int k2 = 1; k2 < R2; ++k2) b[k2] = cmul(v[k2], w[k2]);
  }
}

template <int R1, bool INV>
__device__ __noinline__ void fast_stage2(int a_off, int b_off, int t_off, int nl, int N) {
  const int R2 = N / R1, NB = pad_stride(N, R2);
  float2 w[R1 / 2 + 1];  // W_R1^c = W_N^(c R2), c <= R1 / 2
#pragma unroll
  for (int c = 0; c <= R1 / 2; ++c) w[c] = twid(seg_lds + t_off, c * R2, INV);
  for (int job = threadIdx.x; job < nl * R2; job += NT) {
    const int l = job / R2, k2 = job - l * R2;
    const float2* b = seg_lds + b_off + l * NB + k2;
    float2 x[R1];
#pragma unroll
    for (int m = 0; m < R1; ++m) x[m] = b[m * (R2 + 1)];
    float2* a = seg_lds + a_off + l * N + k2;
    float2 X[R1];
    small_dft<R1>(x, w, X);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) a[k1 * R2] = X[k1];
  }
}

// Fused middle of a half step (R1 <= 16): per (line, k2) job, in place in the scratch layout,
// the forward R1-point DFT over n1, the 1-D Fresnel factor q[k2 + R2 k1], the inverse
// R1-point DFT (the forward one with its outputs read in reverse order, IDFT[n] = DFT[-n]) and
// the inverse transform's twiddle conj W_N^(n1 k2).  A job reads and writes only its own
// column k2 of a line, so no barrier separates its reads from its writes.  With the inverse
// transform's factors swapped (n = n1 + R1 n2, k = k2 + R2 k1), what remains of it is an
// R2-point transform per (line, n1) (fused_last).
template <int R1>
__device__ __noinline__ void fused_mid(int b_off, int t_off, int q_off, int nl, int N) {
  const int R2 = N / R1, NB = pad_stride(N, R2);
  const float2* T = seg_lds + t_off;
  const float2* Q = seg_lds + q_off;
  float2 w[R1 / 2 + 1];  // W_R1^c = W_N^(c R2), c <= R1 / 2, forward
#pragma unroll
  for (int c = 0; c <= R1 / 2; ++c) w[c] = T[c * R2];
  const float2* TR = T - NB;  // row twiddles W_N^(n1 k2) at TR[n1 (R2 + 1) + k2]
  for (int job = threadIdx.x; job < nl * R2; job += NT) {
    const int l = job / R2, k2 = job - l * R2;
    float2* b = seg_lds + b_off + l * NB + k2;
    float2 x[R1], X[R1];
#pragma unroll
    for (int m = 0; m < R1; ++m) x[m] = b[m * (R2 + 1)];
    small_dft<R1>(x, w, X);
    // (scheduling fences: the factor and twiddle reads are not hoisted above the transforms,
    // where their registers pushed the larger sizes into spills)
    if constexpr (R1 > 12) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) x[k1] = cmul(X[k1], Q[k2 + R2 * k1]);
    if constexpr (R1 > 12) __builtin_amdgcn_sched_barrier(0);
    small_dft<R1>(x, w, X);
    if constexpr (R1 > 12) __builtin_amdgcn_sched_barrier(0);
    b[0] = X[0];
#pragma unroll
    for (int n1 = 1; n1 < R1; ++n1)
      b[n1 * (R2 + 1)] = cmul(X[R1 - n1], twid(TR + n1 * (R2 + 1), k2, true));
  }
}

// Last stage of the fused half step: per (line, n1) job, the inverse R2-point transform over
// k2 of scratch row n1, giving x[n1 + R1 n2] in natural order in A.
template <int R2>
__device__ __noinline__ void fused_last(int a_off, int b_off, int t_off, int nl, int N) {
  constexpr bool POW2 = (R2 & (R2 - 1)) == 0;
  const int R1 = N / R2, NB = pad_stride(N, R2);
  const float2* T = seg_lds + t_off;
  float2 wr[POW2 ? 1 : R2 / 2 + 1];  // W_R2^-c = conj W_N^(c R1), c <= R2 / 2
  if constexpr (!POW2) {
#pragma unroll
    for (int c = 0; c <= R2 / 2; ++c) wr[c] = twid(T, c * R1, true);
  }
  for (int job = threadIdx.x; job < nl * R1; job += NT) {
    const int l = job / R1, n1 = job - l * R1;
    const float2* b = seg_lds + b_off + l * NB + n1 * (R2 + 1);
    float2 v[R2];
#pragma unroll
    for (int m = 0; m < R2; ++m) v[m] = b[m];
    if constexpr (POW2) {
      fftk::fft<R2, true>(v);
    } else {
      float2 X[R2];
      small_dft<R2>(v, wr, X);
#pragma unroll
      for (int k = 0; k < R2; ++k) v[k] = X[k];
    }
    float2* a = seg_lds + a_off + l * N + n1;
#pragma unroll
    for (int m = 0; m < R2; ++m) a[R1 * m] = v[m];
  }
}

// Stage 2 for R1 > 16: one output X[k2 + R2 k1] per work item, inputs and twiddles W_R1^(m k1)
// read from LDS with the exponent advanced incrementally.
template <bool INV>
__device__ __noinline__ void wide_stage2(int a_off, int b_off, int t_off, int nl, int N, int R1) {
  const int R2 = N / R1, NB = pad_stride(N, R2);
  const float2* T = seg_lds + t_off;
  for (int it = threadIdx.x; it < nl * N; it += NT) {
    const int l = it / N, k = it - l * N, k1 = k / R2, k2 = k - k1 * R2;
    const float2* b = seg_lds + b_off + l * NB + k2;
    float2 acc = b[0];
    int e = 0;  // m k1 mod R1
    for (int m = 1; m < R1; ++m) {
      e += k1;
      if (e >= R1) e -= R1;
      acc = cmac(acc, b[m * (R2 + 1)], twid(T, e * R2, INV));
    }
    seg_lds[a_off + l * N + k] = acc;
  }
}

__device__ __forceinline__ int fast_r2(int N) { return fast_split(N); }

#define MAOMI_SEG_SIZES(X)                                                                 \
  X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17)  \
  X(18) X(19) X(20) X(21) X(22) X(23) X(24)
#define MAOMI_SEG_SIZES_R1(X) MAOMI_SEG_SIZES(X) X(25) X(26) X(27) X(28) X(29) X(30) X(31) X(32)

// Line splits fall into three classes: 0, a power of two <= 16 times at most 16 (the common
// field sides); 1, both factors at most 16; 2, the rest.  Each class has a kernel of its own
// (CLS) that contains only the dispatch of classes up to CLS, so that the registers the
// larger instantiations clobber do not weigh on the call sites of the common sizes.
__host__ __device__ inline int split_class(int N, int r2) {
  const int r1 = N / r2;
  if (r2 <= 16 && (r2 & (r2 - 1)) == 0 && r1 <= 16) return 0;
  return (r2 <= 16 && r1 <= 16) ? 1 : 2;
}

template <bool INV, int CLS>
__device__ void fast_line_transform(int a_off, int b_off, int t_off, int nl, int N, int r2) {
  const int r1 = N / r2, c = split_class(N, r2);
  if (CLS == 0 || c == 0) {
    switch (r2) {
      case 2: fast_stage1<2, INV>(a_off, b_off, t_off, nl, N); break;
      case 4: fast_stage1<4, INV>(a_off, b_off, t_off, nl, N); break;
      case 8: fast_stage1<8, INV>(a_off, b_off, t_off, nl, N); break;
      default: fast_stage1<16, INV>(a_off, b_off, t_off, nl, N); break;
    }
    lds_barrier();
    switch (r1) {
#define MAOMI_R1(R) case R: fast_stage2<R, INV>(a_off, b_off, t_off, nl, N); break;
      MAOMI_R1(1) MAOMI_R1(2) MAOMI_R1(3) MAOMI_R1(4) MAOMI_R1(5) MAOMI_R1(6) MAOMI_R1(7)
      MAOMI_R1(8) MAOMI_R1(9) MAOMI_R1(10) MAOMI_R1(11) MAOMI_R1(12) MAOMI_R1(13) MAOMI_R1(14)
      MAOMI_R1(15) MAOMI_R1(16)
#undef MAOMI_R1
      default: break;
    }
  } else if (CLS == 1 || c == 1) {
    if constexpr (CLS >= 1) {
      switch (r2) {
#define MAOMI_R2(R) case R: fast_stage1<R, INV>(a_off, b_off, t_off, nl, N); break;
        MAOMI_R2(2) MAOMI_R2(3) MAOMI_R2(4) MAOMI_R2(5) MAOMI_R2(6) MAOMI_R2(7) MAOMI_R2(8)
        MAOMI_R2(9) MAOMI_R2(10) MAOMI_R2(11) MAOMI_R2(12) MAOMI_R2(13) MAOMI_R2(14)
        MAOMI_R2(15) MAOMI_R2(16)
#undef MAOMI_R2
        default: break;
      }
      lds_barrier();
      switch (r1) {
#define MAOMI_R1(R) case R: fast_stage2<R, INV>(a_off, b_off, t_off, nl, N); break;
        MAOMI_R1(1) MAOMI_R1(2) MAOMI_R1(3) MAOMI_R1(4) MAOMI_R1(5) MAOMI_R1(6) MAOMI_R1(7)
        MAOMI_R1(8) MAOMI_R1(9) MAOMI_R1(10) MAOMI_R1(11) MAOMI_R1(12) MAOMI_R1(13) MAOMI_R1(14)
        MAOMI_R1(15) MAOMI_R1(16)
#undef MAOMI_R1
        default: break;
      }
    }
  } else if constexpr (CLS == 2) {
    switch (r2) {
#define MAOMI_R2(R) case R: fast_stage1<R, INV>(a_off, b_off, t_off, nl, N); break;
      MAOMI_SEG_SIZES(MAOMI_R2)
#undef MAOMI_R2
      default: break;  // fast_split never returns another size
    }
    lds_barrier();
    switch (r1) {
#define MAOMI_R1(R) case R: fast_stage2<R, INV>(a_off, b_off, t_off, nl, N); break;
      MAOMI_R1(1) MAOMI_SEG_SIZES_R1(MAOMI_R1)
#undef MAOMI_R1
      default: wide_stage2<INV>(a_off, b_off, t_off, nl, N, r1); break;
    }
  }
  lds_barrier();
}

// One half step of nl lines through the fused stages: forward stage 1 (A -> B), fused_mid
// (B in place) and fused_last (B -> A), two barriers fewer and two LDS passes fewer than a
// forward transform, a separate Fresnel-factor pass and an inverse transform.  For splits of
// classes 0 and 1 (R1 <= 16); returns false for the others.
template <int CLS>
__device__ bool fused_half_step(float2* A, float2* B, int nl, int N, const float2* T,
                                const float2* q) {
  const int r2 = fast_r2(N);
  if (!r2) return false;
  const int r1 = N / r2, c = split_class(N, r2);
  if (c > CLS || c > 1) return false;
  const int a_off = (int)(A - seg_lds), b_off = (int)(B - seg_lds), t_off = (int)(T - seg_lds),
            q_off = (int)(q - seg_lds);
  if (CLS == 0 || c == 0) {
    switch (r2) {
      case 2: fast_stage1<2, false>(a_off, b_off, t_off, nl, N); break;
      case 4: fast_stage1<4, false>(a_off, b_off, t_off, nl, N); break;
      case 8: fast_stage1<8, false>(a_off, b_off, t_off, nl, N); break;
      default: fast_stage1<16, false>(a_off, b_off, t_off, nl, N); break;
    }
  } else if constexpr (CLS >= 1) {
    switch (r2) {
#define MAOMI_R2(R) case R: fast_stage1<R, false>(a_off, b_off, t_off, nl, N); break;
      MAOMI_R2(2) MAOMI_R2(3) MAOMI_R2(4) MAOMI_R2(5) MAOMI_R2(6) MAOMI_R2(7) MAOMI_R2(8)
      MAOMI_R2(9) MAOMI_R2(10) MAOMI_R2(11) MAOMI_R2(12) MAOMI_R2(13) MAOMI_R2(14)
      MAOMI_R2(15) MAOMI_R2(16)
#undef MAOMI_R2
      default: break;
    }
  }
  lds_barrier();
  switch (r1) {
#define MAOMI_R1(R) case R: fused_mid<R>(b_off, t_off, q_off, nl, N); break;
    MAOMI_R1(1) MAOMI_R1(2) MAOMI_R1(3) MAOMI_R1(4) MAOMI_R1(5) MAOMI_R1(6) MAOMI_R1(7)
    MAOMI_R1(8) MAOMI_R1(9) MAOMI_R1(10) MAOMI_R1(11) MAOMI_R1(12) MAOMI_R1(13) MAOMI_R1(14)
    MAOMI_R1(15) MAOMI_R1(16)
#undef MAOMI_R1
    default: break;
  }
  lds_barrier();
  if (CLS == 0 || c == 0) {
    switch (r2) {
      case 2: fused_last<2>(a_off, b_off, t_off, nl, N); break;
      case 4: fused_last<4>(a_off, b_off, t_off, nl, N); break;
      case 8: fused_last<8>(a_off, b_off, t_off, nl, N); break;
      default: fused_last<16>(a_off, b_off, t_off, nl, N); break;
    }
  } else if constexpr (CLS >= 1) {
    switch (r2) {
#define MAOMI_R2(R) case R: fused_last<R>(a_off, b_off, t_off, nl, N); break;
      MAOMI_R2(2) MAOMI_R2(3) MAOMI_R2(4) MAOMI_R2(5) MAOMI_R2(6) MAOMI_R2(7) MAOMI_R2(8)
      MAOMI_R2(9) MAOMI_R2(10) MAOMI_R2(11) MAOMI_R2(12) MAOMI_R2(13) MAOMI_R2(14)
      MAOMI_R2(15) MAOMI_R2(16)
#undef MAOMI_R2
      default: break;
    }
  }
  lds_barrier();
  return true;
}

// Transform nl lines of A through B with the fast split when N has one, else line_dft.
template <int CLS>
__device__ __forceinline__ void line_transform(float2* A, float2* B, int nl, int N, int N1,
                                               int N2, const float2* T, bool inv) {
  const int r2 = fast_r2(N);
  if (r2) {
    const int a_off = (int)(A - seg_lds), b_off = (int)(B - seg_lds), t_off = (int)(T - seg_lds);
    if (inv)
      fast_line_transform<true, CLS>(a_off, b_off, t_off, nl, N, r2);
    else
      fast_line_transform<false, CLS>(a_off, b_off, t_off, nl, N, r2);
  } else {
    line_dft(A, B, nl, N, N1, N2, T, inv);
  }
}

// out[x][y] = scale * sum_i mx[x][i] sum_j my[y][j] in[i][j]  (in: R x C, tmp: X x C).
// With bands (bx[x] = first nonzero column of row x of mx and the span holding all its
// nonzeros; likewise by), only the span is summed: the same nonzero terms in the same order,
// so the result equals the full-row sum.
__device__ void resample2d(const float* in, int R, int C, const float* mx, const int32_t* bx,
                           int X, const float* my, const int32_t* by, int Y, float* tmp,
                           float* out, int64_t ostride, float scale) {
  for (int it = threadIdx.x; it < X * C; it += NT) {
    const int x = it / C, j = it % C;
    const int i0 = bx ? bx[2 * x] : 0, i1 = bx ? i0 + bx[2 * x + 1] : R;
    float acc = 0.f;
    for (int i = i0; i < i1; ++i) acc += mx[x * R + i] * in[i * C + j];
    tmp[it] = acc;
  }
  __syncthreads();
  for (int it = threadIdx.x; it < X * Y; it += NT) {
    const int x = it / Y, y = it % Y;
    const int j0 = by ? by[2 * y] : 0, j1 = by ? j0 + by[2 * y + 1] : C;
    float acc = 0.f;
    for (int j = j0; j < j1; ++j) acc += my[y * C + j] * tmp[x * C + j];
    out[(int64_t)it * ostride] = scale * acc;
  }
  __syncthreads();
}

__device__ __forceinline__ float powk(float p, int order) {
  return order == 1 ? p : order == 2 ? p * p : p * p * p;
}

// Tile lines of the field into A (line-major, A[l N + j] = element j of line l) and back.
// Rows: line l is row r0 + l.  Columns: line l is column r0 + l, read and written with the line
// index fastest so that consecutive threads touch consecutive addresses.  Loads go in batches
// of kBatch per thread, each batch loaded into registers before any of it is stored.
__device__ __forceinline__ void load_tile(float2* A, const float2* fld, bool rows, int r0, int nl,
                                          int N) {
  if (rows) {
    const float2* f = fld + (int64_t)r0 * N;
    for (int b0 = threadIdx.x; b0 < nl * N; b0 += kBatch * NT) {
      float2 ld[kBatch];
#pragma unroll
      for (int k = 0; k < kBatch; ++k)
        if (b0 + k * NT < nl * N) ld[k] = f[b0 + k * NT];
#pragma unroll
      for (int k = 0; k < kBatch; ++k)
        if (b0 + k * NT < nl * N) A[b0 + k * NT] = ld[k];
    }
  } else {
    const int di = NT / nl, dc = NT - di * nl;
    int i = threadIdx.x / nl, l = threadIdx.x - (threadIdx.x / nl) * nl;
    for (int b0 = threadIdx.x; b0 < nl * N; b0 += kBatch * NT) {
      float2 ld[kBatch];
      int dst[kBatch];
#pragma unroll
      for (int k = 0; k < kBatch; ++k) {
        if (b0 + k * NT < nl * N) {
          ld[k] = fld[(int64_t)i * N + r0 + l];
          dst[k] = l * N + i;
        }
        i += di;
        l += dc;
        if (l >= nl) {
          l -= nl;
          ++i;
        }
      }
#pragma unroll
      for (int k = 0; k < kBatch; ++k)
        if (b0 + k * NT < nl * N) A[dst[k]] = ld[k];
    }
  }
}

__device__ __forceinline__ void store_tile(const float2* A, float2* fld, bool rows, int r0, int nl,
                                           int N) {
  if (rows) {
    for (int it = threadIdx.x; it < nl * N; it += NT) fld[(int64_t)r0 * N + it] = A[it];
  } else {
    const int di = NT / nl, dc = NT - di * nl;
    int i = threadIdx.x / nl, l = threadIdx.x - (threadIdx.x / nl) * nl;
    for (int it = threadIdx.x; it < nl * N; it += NT, i += di, l += dc) {
      if (l >= nl) {
        l -= nl;
        ++i;
      }
      fld[(int64_t)i * N + r0 + l] = A[l * N + i];
    }
  }
}

// One half step of nl lines along their axis: forward transform, the 1-D Fresnel factor
// q[k] (1/N folded in) at frequency k, inverse transform.
template <int CLS>
__device__ __forceinline__ void half_step(float2* A, float2* B, int nl, int N, int N1, int N2,
                                          const float2* T, const float2* q) {
  if (fused_half_step<CLS>(A, B, nl, N, T, q)) return;
  line_transform<CLS>(A, B, nl, N, N1, N2, T, false);
  const int dl = NT / N, dk = NT - dl * N;
  for (int it = threadIdx.x, k = threadIdx.x - (threadIdx.x / N) * N; it < nl * N;
       it += NT, k += dk) {
    if (k >= N) k -= N;
    A[it] = cmul(A[it], q[k]);
  }
  lds_barrier();
  line_transform<CLS>(A, B, nl, N, N1, N2, T, true);
}

// One segment of one site.  src: first-plane source (complex), with index shift
// (sx, sy) and side src_n; dst (may be null): receives the last plane (n x n).
//
// The Fresnel step is separable, Q(kx, ky) = q(kx) q(ky), so a step is Px Py with Px (along x)
// and Py (along y) commuting, and the planes u_{s+1} = Px Py T_s u_s (T_s: screen times edge
// window) are produced by one pass per plane, alternating rows and columns: pass s finishes
// the half step of slice s - 1 along its axis (the field is then u_s), does plane s's pointwise
// work (screen, tails, focal staging, the segment's last plane) and starts slice s's half step
// along the same axis.  Pass 0 starts from the source, pass S only finishes.  Each pass moves
// the field through the workspace once: S + 1 round trips for S steps.
template <int CLS>
__device__ __forceinline__ void run_segment(const SegArgs& a, int k, int64_t ox, int64_t oy, const float2* src,
                            int src_n, int sx, int sy, float2* fld, float2* dst, float* tail,
                            float* stage, float* tmp, int site, float2* A, float2* B, float2* T,
                            float2* Qp, float2* Qn, float* E, double* red) {
  const Seg& g = a.seg[k];
  const int N = g.n, N1 = g.n1, N2 = g.n2f, S = g.steps;
  // lines per tile for this side: as many as the line buffers (sized for the largest side)
  // hold, so that smaller fields keep the block's threads busy
  const int r2s = fast_r2(N), NBs = r2s ? pad_stride(N, r2s) : N;
  const int L = max(1, min(a.L * a.n / N, a.L * a.nb / NBs));
  for (int i = threadIdx.x; i < N; i += NT) { T[i] = g.tw[i]; E[i] = g.edge[i]; }
  if (r2s) {  // row twiddles W_N^(n1 k2) at TR[n1 (R2 + 1) + k2], TR = T - pad_stride(N, R2)
    float2* TR = T - NBs;
    for (int i = threadIdx.x; i < N; i += NT) {
      const int n1 = i / r2s, k2 = i - n1 * r2s;
      TR[n1 * (r2s + 1) + k2] = g.tw[n1 * k2];
    }
  }
  __syncthreads();
  const bool top = k == 0, bot = k == 2, foc = k == 1;
  const int first_saved = top ? a.cut : 0;
  const int nsaved = top ? S + 1 - a.cut : S + 1;
  double wsum = 0.0;
  if ((top || bot) && a.tails) {
    for (int i = threadIdx.x; i < a.n2 * a.n2; i += NT) tail[i] = 0.f;
    __syncthreads();  // other threads accumulate into these pixels
  }
  // element it = l * N + j of a tile, stepped by NT without divisions
  const int tl0 = threadIdx.x / N, ty0 = threadIdx.x - tl0 * N;
  const int dl = NT / N, dy = NT - dl * N;
  for (int s = 0; s <= S; ++s) {
    const bool first = s == 0, last = s == S;
    const bool rows = (s & 1) == 0;  // pass s runs along rows for even s, columns for odd s
    const bool save_tail = a.tails && (top || bot) && s >= first_saved;
    const bool save_foc = foc && s < S;
    const float trap = (s == first_saved || s == S) ? 0.5f : 1.0f;
    const float ax = save_tail ? (top ? a.ax_top[s - a.cut] : a.ax_bot[s])
                               : (save_foc ? a.ax_focus[s] : 1.f);
    double prof = 0.0;
    // q of the step that this pass finishes (s - 1) and of the step it starts (s)
    for (int i = threadIdx.x; i < N; i += NT) {
      if (!first) Qp[i] = g.q[(int64_t)g.qidx[s - 1] * N + i];
      if (!last) Qn[i] = g.q[(int64_t)g.qidx[s] * N + i];
    }
    __syncthreads();
    for (int r0 = 0; r0 < N; r0 += L) {
      const int nl = min(L, N - r0);
      if (first) {  // the segment's source, cropped or zero-padded (rows: s = 0 is a row pass)
        for (int it = threadIdx.x; it < nl * N; it += NT) {
          const int l = it / N, y = it % N;
          const int xs = r0 + l + sx, ys = y + sy;
          A[l * N + y] = (xs >= 0 && xs < src_n && ys >= 0 && ys < src_n)
                             ? src[(int64_t)xs * src_n + ys] : make_float2(0.f, 0.f);
        }
        lds_barrier();
      } else {
        load_tile(A, fld, rows, r0, nl, N);
        lds_barrier();
        half_step<CLS>(A, B, nl, N, N1, N2, T, Qp);
      }
      // the tile's phases, batched the same way, wait in B (free between the transforms) at
      // the indices of their elements
      float* PH = (float*)B;
      if (!last) {
        const float* ph_s = g.phase + (int64_t)s * g.H * g.W;
        if (rows) {  // each thread reads back only its own elements: no barrier needed
          int l = tl0, y = ty0;
          for (int b0 = threadIdx.x; b0 < nl * N; b0 += kBatch * NT) {
            float ph[kBatch];
#pragma unroll
            for (int k = 0; k < kBatch; ++k) {
              if (b0 + k * NT < nl * N) ph[k] = ph_s[(ox + r0 + l) * g.W + oy + y];
              l += dl;
              y += dy;
              if (y >= N) {
                y -= N;
                ++l;
              }
            }
#pragma unroll
            for (int k = 0; k < kBatch; ++k)
              if (b0 + k * NT < nl * N) PH[b0 + k * NT] = ph[k];
          }
        } else {  // columns: read with the line index fastest, stored at the element's index
          const int di = NT / nl, dc = NT - di * nl;
          int i = threadIdx.x / nl, l = threadIdx.x - (threadIdx.x / nl) * nl;
          for (int b0 = threadIdx.x; b0 < nl * N; b0 += kBatch * NT) {
            float ph[kBatch];
            int at[kBatch];
#pragma unroll
            for (int k = 0; k < kBatch; ++k) {
              if (b0 + k * NT < nl * N) {
                ph[k] = ph_s[(ox + i) * g.W + oy + r0 + l];
                at[k] = l * N + i;
              }
              i += di;
              l += dc;
              if (l >= nl) {
                l -= nl;
                ++i;
              }
            }
#pragma unroll
            for (int k = 0; k < kBatch; ++k)
              if (b0 + k * NT < nl * N) PH[at[k]] = ph[k];
          }
          lds_barrier();
        }
      }
      int l = tl0, j = ty0;
      for (int it = threadIdx.x; it < nl * N; it += NT, l += dl, j += dy) {
        if (j >= N) {
          j -= N;
          ++l;
        }
        const int x = rows ? r0 + l : j, y = rows ? j : r0 + l;
        float2 v = A[it];
        if (!last) {
          const float phi = PH[it];
          float sn, cs;
          trig::sincos_screen(phi, &sn, &cs);  // |phi| < 2^17, checked by the host
          const float sg = E[x] * E[y];
          v = cmul(v, make_float2(cs * sg, sn * sg));
        }
        if (save_tail) {
          const int cx = x - a.coff, cy = y - a.coff;
          if (cx >= 0 && cx < a.n2 && cy >= 0 && cy < a.n2) {
            const float I = a.ss2 * powk(v.x * v.x + v.y * v.y, a.order) * ax;
            prof += (double)I;
            tail[cx * a.n2 + cy] += trap * I;
          }
        }
        if (save_foc) {
          const int cx = x - a.foff_x, cy = y - a.foff_y;
          if (cx >= 0 && cx < a.fx && cy >= 0 && cy < a.fy)
            stage[cx * a.fy + cy] = powk(v.x * v.x + v.y * v.y, a.order);
        }
        if (last && dst) dst[(int64_t)x * N + y] = v;
        A[it] = v;
      }
      lds_barrier();
      if (!last) {
        half_step<CLS>(A, B, nl, N, N1, N2, T, Qn);
        store_tile(A, fld, rows, r0, nl, N);
      }
      lds_barrier();  // A is refilled by the next tile
    }
    if (save_tail) {
      const double p = block_sum_double(prof, red);
      if (threadIdx.x == 0) (top ? a.top_prof : a.bot_prof)[(int64_t)site * nsaved + s - first_saved] = (float)p;
      wsum += (double)trap * p;
    }
    if (save_foc) {
      __syncthreads();
      const int b = a.seg[1].steps;
      resample2d(stage, a.fx, a.fy, a.mfx, a.bfx, a.X, a.mfy, a.bfy, a.Y, tmp,
                 a.psf + (int64_t)site * a.X * a.Y * b + s, b, a.focal_scale * ax);
    }
    __syncthreads();  // this pass's field stores are read by other threads in the next pass
  }
  if ((top || bot) && a.tails) {
    __syncthreads();
    float* conv = (top ? a.top_conv : a.bot_conv) + (int64_t)site * a.TT * a.TT;
    resample2d(tail, a.n2, a.n2, a.mt, a.bt, a.TT, a.mt, a.bt, a.TT, tmp, conv, 1,
               a.tail_scale);
    if (threadIdx.x == 0) (top ? a.top_w : a.bot_w)[site] = wsum * a.prop_um;
  }
}

template <int CLS>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2))) void
segmented_sites_kernel(SegArgs a) {
  const int nmax = a.n;
  float2* A = seg_lds;
  float2* B = A + a.L * nmax;
  float2* T = B + a.L * a.nb + a.nb;  // after the row-twiddle region (see run_segment)
  float2* Qp = T + nmax;
  float2* Qn = Qp + nmax;
  float* E = (float*)(Qn + nmax);
  double* red = (double*)(E + nmax + (nmax & 1));
  float2* w = a.work + (int64_t)blockIdx.x * a.work_stride;
  float2* fld = w;
  float2* xfer = w + (int64_t)a.n * a.n;          // A's last plane (n x n)
  float2* xfer2 = xfer + (int64_t)a.n * a.n;      // B's last plane (n2 x n2)
  float* tail = (float*)(xfer2 + (int64_t)a.n2 * a.n2);
  float* stage = tail + a.n2 * a.n2;
  float* tmp = stage + a.fx * a.fy;
  for (int64_t i = blockIdx.x; i < a.P; i += gridDim.x) {
    const int site = a.order_list[i];
    const int64_t ox = a.site_x[site], oy = a.site_y[site];
    // A: base -> planes >= cut (top tail), last plane kept whole in xfer
    run_segment<CLS>(a, 0, ox, oy, a.base, a.n, 0, 0, fld, xfer, tail, stage, tmp, site, A, B, T, Qp, Qn, E,
                red);
    // B: crop of A's last plane -> focal planes; last plane kept in xfer (n2 x n2)
    run_segment<CLS>(a, 1, ox, oy, xfer, a.n, a.coff, a.coff, fld, a.tails ? xfer2 : nullptr, tail,
                stage, tmp, site, A, B, T, Qp, Qn, E, red);
    if (a.tails)  // C: B's last plane padded back to n -> bottom tail
      run_segment<CLS>(a, 2, ox, oy, xfer2, a.n2, -a.coff, -a.coff, fld, nullptr, tail, stage, tmp, site,
                  A, B, T, Qp, Qn, E, red);
  }
}

// Per-site FP64 sums of psf[P][K] (one block per site) ...
__global__ void psf_site_sum_kernel(const float* __restrict__ psf, int64_t K,
                                    double* __restrict__ energy) {
  __shared__ double red[NT / 64];
  const float* row = psf + (int64_t)blockIdx.x * K;
  double acc = 0.0;
  for (int64_t k = threadIdx.x; k < K; k += NT) acc += (double)row[k];
  const double t = block_sum_double(acc, red);
  if (threadIdx.x == 0) energy[blockIdx.x] = t;
}

// ... and per-voxel FP64 sums over sites of |x|, x and min(x, 0).  A block takes kVox
// consecutive voxels; its NT / kVox site lanes each sum every (NT / kVox)-th site of one
// voxel (consecutive threads read consecutive voxels of one site), and the lanes' partial
// sums are added in lane order, so the result does not depend on scheduling.
constexpr int kVox = 16;

__global__ __launch_bounds__(NT) void psf_voxel_sums_kernel(const float* __restrict__ psf,
                                                            int64_t P, int64_t K,
                                                            double* __restrict__ sabs,
                                                            double* __restrict__ ssum,
                                                            double* __restrict__ sneg) {
  constexpr int LANES = NT / kVox;
  __shared__ double part[3][LANES][kVox];
  const int v = threadIdx.x % kVox, j = threadIdx.x / kVox;
  for (int64_t k0 = (int64_t)blockIdx.x * kVox; k0 < K; k0 += (int64_t)gridDim.x * kVox) {
    const int64_t k = k0 + v;
    double a = 0.0, s = 0.0, n = 0.0;
    if (k < K) {
#pragma unroll 4
      for (int64_t p = j; p < P; p += LANES) {
        const double x = psf[p * K + k];
        a += fabs(x);
        s += x;
        n += x < 0.0 ? x : 0.0;
      }
    }
    part[0][j][v] = a;
    part[1][j][v] = s;
    part[2][j][v] = n;
    __syncthreads();
    if (threadIdx.x < 3 * kVox) {
      const int c = threadIdx.x / kVox, w = threadIdx.x % kVox;
      double t = 0.0;
      for (int l = 0; l < LANES; ++l) t += part[c][l][w];
      if (k0 + w < K) (c == 0 ? sabs : c == 1 ? ssum : sneg)[k0 + w] = t;
    }
    __syncthreads();
  }
}

}  // namespace

MAOMI_API int64_t maomi_segmented_workspace_elems(int64_t n, int64_t n2, int64_t fx, int64_t fy,
                                                  int64_t X, int64_t TT) {
  // field + A's last plane + B's last plane (complex), then tail accumulator,
  // focal staging and the resample intermediate (floats)
  const int64_t tmp = X * fy > TT * n2 ? X * fy : TT * n2;
  const int64_t floats = n2 * n2 + fx * fy + tmp + 64;
  return 2 * n * n + n2 * n2 + (floats + 1) / 2;
}

MAOMI_API int maomi_segmented_sites(const maomi_segmented_desc* d, maomi_stream_t stream) {
  if (!d) { maomi_set_error("maomi_segmented_sites: null descriptor"); return MAOMI_EINVAL; }
  if (d->P <= 0) return MAOMI_OK;
  const int64_t n = d->n, n2 = d->n2;
  if (n < 2 || n2 < 2 || n2 > n || (n - n2) % 2 || n > 1024 || d->work == nullptr ||
      !d->base || !d->site_x || !d->site_y || !d->order_list || !d->psf) {
    maomi_set_error("maomi_segmented_sites: bad sizes or null pointers (n %lld, n2 %lld)",
                    (long long)n, (long long)n2);
    return MAOMI_EINVAL;
  }
  if (d->order < 1 || d->order > 3 || d->fx > n2 || d->fy > n2 || d->X < 1 || d->Y < 1) {
    maomi_set_error("maomi_segmented_sites: bad order or focal window");
    return MAOMI_EINVAL;
  }
  SegArgs a;
  for (int k = 0; k < 3; ++k) {
    const maomi_segment& s = d->seg[k];
    a.seg[k].phase = s.phase; a.seg[k].H = s.H; a.seg[k].W = s.W; a.seg[k].Z = s.Z;
    a.seg[k].n = (int)s.n; a.seg[k].steps = (int)s.steps; a.seg[k].n1 = (int)s.n1;
    a.seg[k].n2f = (int)s.n2f; a.seg[k].q = (const float2*)s.q; a.seg[k].qidx = s.qidx;
    a.seg[k].edge = s.edge; a.seg[k].tw = (const float2*)s.tw;
    const bool used = k < 2 || d->tails;
    if (used && (!s.phase || !s.q || !s.qidx || !s.edge || !s.tw || s.n1 * s.n2f != s.n ||
                 s.steps < 1 || s.Z < s.steps || s.n != (k == 1 ? n2 : n))) {
      maomi_set_error("maomi_segmented_sites: segment %d is inconsistent", k);
      return MAOMI_EINVAL;
    }
  }
  a.base = (const float2*)d->base; a.site_x = d->site_x; a.site_y = d->site_y;
  a.order_list = d->order_list; a.P = d->P; a.cut = (int)d->cut; a.tails = (int)d->tails;
  a.order = (int)d->order; a.n = (int)n; a.n2 = (int)n2; a.coff = (int)((n - n2) / 2);
  a.fx = (int)d->fx; a.fy = (int)d->fy; a.foff_x = (int)((n2 - d->fx) / 2);
  a.foff_y = (int)((n2 - d->fy) / 2); a.X = (int)d->X; a.Y = (int)d->Y; a.TT = (int)d->TT;
  a.mfx = d->mfx; a.mfy = d->mfy; a.mt = d->mt; a.ax_focus = d->ax_focus; a.ax_top = d->ax_top;
  a.bfx = d->mfx_band; a.bfy = d->mfy_band; a.bt = d->mt_band;
  a.ax_bot = d->ax_bot; a.ss2 = (float)d->ss2; a.focal_scale = (float)d->focal_scale;
  a.tail_scale = (float)d->tail_scale; a.prop_um = d->prop_um;
  a.psf = d->psf; a.top_conv = d->top_conv; a.top_prof = d->top_prof; a.top_w = d->top_w;
  a.bot_conv = d->bot_conv; a.bot_prof = d->bot_prof; a.bot_w = d->bot_w;
  a.work = (float2*)d->work; a.work_stride = d->work_stride;
  if (a.work_stride < maomi_segmented_workspace_elems(n, n2, d->fx, d->fy, d->X, d->TT)) {
    maomi_set_error("maomi_segmented_sites: work_stride too small");
    return MAOMI_EINVAL;
  }
  // two line buffers of L x n complex plus tables, kept under 60 KiB of LDS
  // scratch line stride: the padded layout of the fast split (fast_r2 on the device)
  a.nb = (int)n;
  for (const int64_t side : {n, n2}) {
    const int r2 = fast_split((int)side);
    if (r2 && pad_stride((int)side, r2) > a.nb) a.nb = pad_stride((int)side, r2);
  }
  // line buffers A (L x n) and B (L x nb) complex plus tables (row twiddles, twiddles, two
  // Fresnel factors, edge window), kept under 60 KiB of LDS
  a.L = (int)((61440 - 24 * n - 8 * a.nb - 4 * n - 96) / (8 * n + 8 * a.nb));
  if (a.L < 1) a.L = 1;
  const size_t shm = ((size_t)a.L * (n + a.nb) + a.nb + 3 * n) * sizeof(float2) +
                     (n + 2) * sizeof(float) + (NT / 64) * sizeof(double) + 16;
  int64_t grid = d->grid > 0 ? d->grid : d->P;
  if (grid > d->P) grid = d->P;
  // the kernel of the largest line-split class among the two field sides
  int cls = 0;
  for (const int64_t side : {n, n2}) {
    const int r2 = fast_split((int)side);
    const int c = r2 ? split_class((int)side, r2) : 2;
    if (c > cls) cls = c;
  }
  if (cls == 0)
    hipLaunchKernelGGL(segmented_sites_kernel<0>, dim3((unsigned)grid), dim3(NT), shm,
                       (hipStream_t)stream, a);
  else if (cls == 1)
    hipLaunchKernelGGL(segmented_sites_kernel<1>, dim3((unsigned)grid), dim3(NT), shm,
                       (hipStream_t)stream, a);
  else
    hipLaunchKernelGGL(segmented_sites_kernel<2>, dim3((unsigned)grid), dim3(NT), shm,
                       (hipStream_t)stream, a);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_psf_stats(const float* psf, int64_t P, int64_t K, double* energy,
                              double* sum_abs, double* sum, double* sum_neg,
                              maomi_stream_t stream) {
  if (!psf || !energy || !sum_abs || !sum || !sum_neg || P < 1 || K < 1 || P > (1ll << 31)) {
    maomi_set_error("maomi_psf_stats: bad arguments");
    return MAOMI_EINVAL;
  }
  hipLaunchKernelGGL(psf_site_sum_kernel, dim3((unsigned)P), dim3(NT), 0, (hipStream_t)stream,
                     psf, K, energy);
  MAOMI_CHECK_LAUNCH();
  const int64_t blocks = (K + kVox - 1) / kVox;
  hipLaunchKernelGGL(psf_voxel_sums_kernel, dim3((unsigned)(blocks < 65535 ? blocks : 65535)),
                     dim3(NT), 0, (hipStream_t)stream, psf, P, K, sum_abs, sum, sum_neg);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}
