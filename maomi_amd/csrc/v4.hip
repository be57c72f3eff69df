// FP64 MATLAB-v4 (biharmonic spline) mask assembly on the GPU.
//
// Replaces green_apply (maomi/backends/cuda/interpolation.cu:3-19) and the
// CPU Green-matrix build of v4_interpolate (maomi/numerics.py:122-129):
//   G(p,q) = d^2 (ln d - 1) = r2 (0.5 ln r2 - 1),  G = 0 at d = 0,
//   out(q, c) = sum_i G(|q - p_i|) coef(i, c).
// Points and coefficients are staged through LDS in tiles; each thread owns
// one query and keeps C accumulators in registers.
#include <mutex>

#include "common.hpp"
#include "maomi_amd.h"

namespace {

constexpr int kTile = 256;
constexpr int kMaxC = 8;

// Natural logarithm of a finite x > 0 in FP64, within 2 ulp: x = 2^e m with m in
// [sqrt(1/2), sqrt(2)), log x = e ln2 + 2 atanh(s), s = (m - 1) / (m + 1), |s| <= 0.1716, with
// atanh(s) = s + s^3/3 + ... + s^23/23 (remainder below 1e-18 relative).  ln 2 is split into
// a 32-bit head, so e * kLn2Hi is exact, and the rounded remainder.  About a third of the
// instructions of the library log, which carries a double-double result.
__device__ __forceinline__ double log_pos(double x) {
  constexpr double kLn2Hi = 0x1.62e42fee00000p-1, kLn2Lo = 0x1.a39ef35793c76p-33;
  int e;
  double m = frexp(x, &e);
  const bool low = m < 0.70710678118654752440;
  m = low ? 2.0 * m : m;
  e = low ? e - 1 : e;
  const double s = (m - 1.0) / (m + 1.0), t = s * s;
  double p = 1.0 / 23.0;
  p = fma(t, p, 1.0 / 21.0);
  p = fma(t, p, 1.0 / 19.0);
  p = fma(t, p, 1.0 / 17.0);
  p = fma(t, p, 1.0 / 15.0);
  p = fma(t, p, 1.0 / 13.0);
  p = fma(t, p, 1.0 / 11.0);
  p = fma(t, p, 1.0 / 9.0);
  p = fma(t, p, 1.0 / 7.0);
  p = fma(t, p, 1.0 / 5.0);
  p = fma(t, p, 1.0 / 3.0);
  const double s2 = 2.0 * s;
  const double lm = fma(s2 * t, p, s2);
  const double ed = (double)e;
  return fma(ed, kLn2Hi, fma(ed, kLn2Lo, lm));
}

// G as a function of r2 = d^2: r2 (ln(r2) / 2 - 1), 0 at r2 = 0; no branch.
__device__ __forceinline__ double green_r2(double r2) {
  const double g = r2 * fma(0.5, log_pos(r2 > 0.0 ? r2 : 1.0), -1.0);
  return r2 > 0.0 ? g : 0.0;
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
#pragma unroll 4
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

// G on a table, for points on a complete nx x ny tensor grid and queries on a complete
// mx x my tensor grid: every |dx| and |dy| is one of U distances du, so G(dx, dy) =
// T[u * U + v] with u, v the indices of |dx| and |dy| (green_table_kernel fills T with the
// same green_r2).  Query q = (a, b) with a = q / my, b = q % my; ix[a * nx + j] indexes
// |qx_a - px_j|, iy[b * ny + k] indexes |qy_b - py_k|; coef is in grid order, point (j, k) at
// j * ny + k.  A pair costs an index read from LDS, one gather of T and C fused multiply-adds
// instead of a logarithm.  Point columns [j0, j1) go to block row blockIdx.y, as in
// green_apply_kernel.
__global__ void green_table_kernel(const double* __restrict__ du, int64_t U,
                                   double* __restrict__ T) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < U * U;
       e += (int64_t)gridDim.x * blockDim.x) {
    const double a = du[e / U], b = du[e % U];
    T[e] = green_r2(a * a + b * b);
  }
}

constexpr int kGridMaxNy = 128;  // point rows staged in LDS per query (16-bit indices)

// yl: ny x 256 16-bit indices in dynamic LDS, this block's queries' rows of iy.  ROW: the
// block's 256 queries share the query row a (my a multiple of 256), so each point column's
// row of T (U doubles, after yl in LDS) is staged once and read from LDS instead of L2.
template <int C, bool ROW>
__global__ __launch_bounds__(256) void green_grid_kernel(
    const double* __restrict__ T, int64_t U, const int* __restrict__ ix,
    const int* __restrict__ iy, const double* __restrict__ coef, double* __restrict__ out,
    int nx, int ny, int64_t mx, int64_t my, int jchunk) {
  extern __shared__ unsigned short yl[];
  const int tid = threadIdx.x;
  const int64_t Q = mx * my, q = blockIdx.x * (int64_t)blockDim.x + tid;
  const bool live = q < Q;
  const int64_t qq = live ? q : Q - 1;
  const int64_t a = qq / my, b = qq % my;
  for (int k = 0; k < ny; ++k) yl[k * 256 + tid] = (unsigned short)iy[b * ny + k];
  __syncthreads();
  const int j0 = blockIdx.y * jchunk, j1 = min(nx, j0 + jchunk);
  out += (int64_t)blockIdx.y * Q * C;
  double* tl = reinterpret_cast<double*>(yl + ny * 256);  // ny * 512 bytes: 8-byte aligned
  double acc[C];
#pragma unroll
  for (int c = 0; c < C; ++c) acc[c] = 0.0;
  for (int j = j0; j < j1; ++j) {
    const double* trow = T + (int64_t)ix[a * nx + j] * U;
    if (ROW) {
      __syncthreads();  // the previous row's reads are done
      for (int e = tid; e < U; e += 256) tl[e] = trow[e];
      __syncthreads();
      trow = tl;
    }
    const double* cj = coef + (int64_t)j * ny * C;
#pragma unroll 8
    for (int k = 0; k < ny; ++k) {
      const double g = trow[yl[k * 256 + tid]];
#pragma unroll
      for (int c = 0; c < C; ++c) acc[c] = fma(g, cj[k * C + c], acc[c]);
    }
  }
  if (live) {
#pragma unroll
    for (int c = 0; c < C; ++c) out[q * C + c] = acc[c];
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


// ---------------------------------------------------------------------------------------------
// FP64 Cholesky factorisation A = L L^T of the v4 system's positive-definite block (row-major,
// lower triangle, in place), right-looking in steps of kNb columns with a one-step lookahead
// on the diagonal block.  L11 of step k (the factored diagonal block, zero above the diagonal,
// identity past the matrix) lives in side buffer k % 2, transposed: buf[c * kNb + i] = L11[i][c].
// Per step:
//   chol_panel_kernel: solves the rows of the panel below the diagonal block, x L11^T = a,
//     with L11 from the side buffer (a wave per 64 rows, a row per lane);
//   chol_update_kernel: block 0 copies L11 into the diagonal block of A; the others subtract
//     L21 L21^T from the lower 64 x 64 tiles of the trailing matrix, and the block of the first
//     tile (the next diagonal block) then factors it into the other side buffer, so that the
//     sequential factorisation overlaps the rest of the update.
// chol_diag_kernel factors the first diagonal block.  Only the lower triangle is read or
// written.  The first non-positive (or NaN) pivot sets *info to its 1-based column, as
// LAPACK's potrf does.
//
// Both sequential parts run on one wave with no workgroup barrier: a barrier per column costs
// about a thousand cycles, 64 of them a step.  The columns go in four groups of 16; a group
// takes its own triangle column by column, then its update of the later columns as one
// rank-16 update, and the lane's row is then shifted down 16 registers, so that one loop body
// serves the four groups.

constexpr int kNb = 64;
constexpr int kCholThreads = 256;

struct CholShared {
  double lt[kNb][kNb];  // lt[c][i] = L11[i][c] (zero above the diagonal)
  double rd[kNb];       // 1 / L11[j][j]
  double col[16][kNb];  // the current group's columns as broadcast, col[JJ][c] = S[c][16 jg + JJ]
};

// Workgroup barrier that orders LDS only: global loads stay in flight across it.
__device__ __forceinline__ void chol_lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Orders one wave's LDS accesses for the compiler.  A wave's LDS operations are carried out in
// issue order, so a store read back by other lanes of the same wave needs no barrier.
__device__ __forceinline__ void chol_wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Lane l's value of a wave-uniform lane index.
__device__ __forceinline__ double readlane_f64(double v, int l) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(unsigned)u, l);
  const int hi = __builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                          (unsigned long long)(unsigned)lo));
}

// x[16 + 16 Q + jj] -= m[LL] B[LL][jj] for LL, jj < 16, B's rows kNb doubles apart in LDS (a
// broadcast operand: every lane reads the same words).  LL runs in order, and row LL + 1 of B
// is requested before the products of row LL, so that one wave keeps the LDS latency hidden.
template <int Q>
__device__ __forceinline__ void chol_rank16(double (&x)[kNb], const double (&mul)[16],
                                            const double* __restrict__ B) {
  double cur[16], nxt[16];
#pragma unroll
  for (int jj = 0; jj < 16; ++jj) cur[jj] = B[jj];
#pragma unroll
  for (int LL = 0; LL < 16; ++LL) {
    if (LL < 15) {
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) nxt[jj] = B[(LL + 1) * kNb + jj];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int jj = 0; jj < 16; ++jj) x[16 + 16 * Q + jj] = fma(-mul[LL], cur[jj], x[16 + 16 * Q + jj]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int jj = 0; jj < 16; ++jj) cur[jj] = nxt[jj];
  }
}

// Step JJ of column group jg (column j = 16 jg + JJ) of the one-wave factorisation.  Lane i
// holds row i, columns 16 jg .. 63 in s[0 .. 63 - 16 jg].  The pivot S[j][j] and the group's
// rows of column j, S[16 jg + cc][j], are read from their lanes (v_readlane: no LDS round trip
// on the chain of dependent steps); the lane's multiplier l_ij = S[i][j] / S[j][j] updates the
// group's later columns at once, S[i][c] -= l_ij S[c][j], and is kept for the columns past
// the group, whose rows of column j go through LDS for that update.  Rows above j hold zeros
// in column j, so their multipliers are zero.  The pivot is kept by lane j.
template <int JJ>
struct CholWaveSteps {
  __device__ __forceinline__ static void run(double (&s)[kNb], double (&l)[16], double& piv,
                                             int i, int jg, double (*col)[kNb]) {
    col[JJ][i] = s[JJ];
    const double d = readlane_f64(s[JJ], 16 * jg + JJ);
    piv = i == 16 * jg + JJ ? d : piv;
    l[JJ] = s[JJ] / d;
#pragma unroll
    for (int cc = JJ + 1; cc < 16; ++cc)
      s[cc] = fma(-l[JJ], readlane_f64(s[JJ], 16 * jg + cc), s[cc]);
    CholWaveSteps<JJ + 1>::run(s, l, piv, i, jg, col);
  }
};
template <>
struct CholWaveSteps<16> {
  __device__ __forceinline__ static void run(double (&)[kNb], double (&)[16], double&, int, int,
                                             double (*)[kNb]) {}
};

// Factor the b x b block at D (b <= kNb) into sm.lt and sm.rd on wave 0; rows and columns past
// b are the identity.  Called by the whole workgroup (the other waves wait at the barrier).
__device__ __forceinline__ void chol_factor_diag(const double* __restrict__ D, int64_t lda, int b,
                                                 CholShared& sm, int* info, int64_t k0,
                                                 bool report) {
  const int t = threadIdx.x;
  if (t < 64) {
    const int i = t;
    // every lane loads from the block's lower triangle (clamped indices) and then selects: a
    // load under a condition is issued in a branch of its own that waits for it
    const int ib = min(i, b - 1);
    const double* drow = D + (int64_t)ib * lda;
    double s[kNb];
#pragma unroll
    for (int c = 0; c < kNb; ++c) {
      const double v = drow[min(c, ib)];
      s[c] = (i < b && c <= i) ? v : (c == i ? 1.0 : 0.0);
    }
    double piv = 1.0;
#pragma unroll 1
    for (int jg = 0; jg < 4; ++jg) {
      double l[16];
      chol_wave_fence();  // the previous group's reads of col come before this group's stores
      CholWaveSteps<0>::run(s, l, piv, i, jg, sm.col);
      chol_wave_fence();  // the group's columns are in LDS before the trailing update reads them
      // the group's update of columns 16 (jg + 1 + Q) .. + 15, its 16 steps in column order
      if (jg < 3) chol_rank16<0>(s, l, &sm.col[0][16 * jg + 16]);
      if (jg < 2) chol_rank16<1>(s, l, &sm.col[0][16 * jg + 32]);
      if (jg < 1) chol_rank16<2>(s, l, &sm.col[0][16 * jg + 48]);
#pragma unroll
      for (int cc = 0; cc < 16; ++cc) sm.lt[16 * jg + cc][i] = s[cc];  // unscaled S[i][c]
#pragma unroll
      for (int c = 0; c < kNb - 16; ++c) s[c] = s[c + 16];
    }
    const unsigned long long bad = __ballot(i < b && !(piv > 0.0));
    if (report && bad != 0ull && i == 0) atomicCAS(info, 0, (int)(k0 + __ffsll(bad)));
    const double d = sqrt(piv);
    sm.rd[i] = 1.0 / d;
    chol_wave_fence();
#pragma unroll 8
    for (int c = 0; c < kNb; ++c) {
      const double v = sm.lt[c][i];
      sm.lt[c][i] = c < i ? v * sm.rd[c] : (c == i ? d : 0.0);
    }
  }
  chol_lds_barrier();
}

// L11 of the factored block (sm.lt) to a side buffer, transposed as sm.lt.
__device__ __forceinline__ void chol_store_l11(const CholShared& sm, double* __restrict__ l11t) {
  const double* src = &sm.lt[0][0];
  for (int e = threadIdx.x; e < kNb * kNb; e += kCholThreads) l11t[e] = src[e];
}

__global__ __launch_bounds__(kCholThreads) void chol_diag_kernel(const double* __restrict__ A,
                                                                 int64_t lda, int64_t m,
                                                                 int64_t k,
                                                                 double* __restrict__ l11t,
                                                                 int* info) {
  __shared__ __align__(16) CholShared sm;
  chol_factor_diag(A + k * lda + k, lda, (int)min((int64_t)kNb, m - k), sm, info, k, true);
  chol_store_l11(sm, l11t);
}

// Panel rows k + kNb .. m - 1, one wave per 64 rows (launched only when there are some): lane =
// row, its 64 columns in registers, L11 in LDS.  The rows are independent, so x L11^T = a is
// solved per lane without any exchange: column l is x[l] /= L11[l][l], then x[j] -= x[l]
// L11[j][l] for j > l, in column groups of 16 as in the factorisation (the group's triangle,
// then its update of the later columns; the same operations in the same order per column).
__global__ __launch_bounds__(64) void chol_panel_kernel(double* __restrict__ A, int64_t lda,
                                                        int64_t m, int64_t k,
                                                        const double* __restrict__ l11t) {
  __shared__ __align__(16) double lt[kNb][kNb];  // lt[c][i] = L11[i][c]
  __shared__ double rd[kNb];                     // 1 / L11[j][j]
  const int lane = threadIdx.x;
  const int64_t r = k + kNb + (int64_t)blockIdx.x * 64 + lane;
  const bool live = r < m;
  double* row = A + (live ? r : k) * lda + k;
  // every load in flight at once: the row, then L11 (staged to LDS when it has arrived)
  double x[kNb];  // a lane past m works on row k (finite values) and stores nothing
#pragma unroll
  for (int jj = 0; jj < kNb; ++jj) x[jj] = row[jj];
  {
    double lv[kNb];
#pragma unroll
    for (int c = 0; c < kNb; ++c) lv[c] = l11t[c * kNb + lane];
    const double dg = l11t[lane * kNb + lane];
#pragma unroll
    for (int c = 0; c < kNb; ++c) lt[c][lane] = lv[c];
    rd[lane] = 1.0 / dg;
  }
  chol_wave_fence();
#pragma unroll 1
  for (int lg = 0; lg < 4; ++lg) {
    const double* L = &lt[16 * lg][16 * lg];  // L[LL * kNb + jj] = L11[16 lg + jj][16 lg + LL]
#pragma unroll
    for (int LL = 0; LL < 16; ++LL) {
      x[LL] *= rd[16 * lg + LL];
#pragma unroll
      for (int jj = LL + 1; jj < 16; ++jj) x[jj] = fma(-x[LL], L[LL * kNb + jj], x[jj]);
    }
    double xl[16];
#pragma unroll
    for (int LL = 0; LL < 16; ++LL) xl[LL] = x[LL];
    if (lg < 3) chol_rank16<0>(x, xl, L + 16);
    if (lg < 2) chol_rank16<1>(x, xl, L + 32);
    if (lg < 1) chol_rank16<2>(x, xl, L + 48);
    if (live) {
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) row[16 * lg + jj] = x[jj];
    }
#pragma unroll
    for (int jj = 0; jj < kNb - 16; ++jj) x[jj] = x[jj + 16];
  }
}

__global__ __launch_bounds__(kCholThreads) void chol_update_kernel(double* __restrict__ A,
                                                                   int64_t lda, int64_t m,
                                                                   int64_t k,
                                                                   const double* __restrict__ l11t,
                                                                   double* __restrict__ l11t_next,
                                                                   int* info) {
  constexpr int LS = kNb + 1;  // odd row stride: the transposed stores spread over the banks
  // the two transposed panel tiles; the first tile's block reuses the space to factor
  __shared__ __align__(16) double smem[2 * kNb * LS];
  static_assert(sizeof(CholShared) <= sizeof(double) * 2 * kNb * LS, "factor scratch fits");
  double(*pi)[LS] = reinterpret_cast<double(*)[LS]>(smem);
  double(*pj)[LS] = reinterpret_cast<double(*)[LS]>(smem + kNb * LS);
  const int t = threadIdx.x;
  if (blockIdx.x == 0) {
    const int b = (int)min((int64_t)kNb, m - k);
    double v[kNb * kNb / kCholThreads];
#pragma unroll
    for (int n = 0; n < kNb * kNb / kCholThreads; ++n) v[n] = l11t[t + n * kCholThreads];
#pragma unroll
    for (int n = 0; n < kNb * kNb / kCholThreads; ++n) {
      const int e = t + n * kCholThreads, c = e >> 6, i = e & 63;
      if (i < b && c <= i) A[(k + i) * lda + k + c] = v[n];
    }
    return;
  }
  const int64_t id = blockIdx.x - 1;
  int64_t I = (int64_t)((sqrt(8.0 * (double)id + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= id) ++I;
  while (I * (I + 1) / 2 > id) --I;
  const int64_t J = id - I * (I + 1) / 2;
  const int64_t k1 = k + kNb, r0 = k1 + kNb * I, c0 = k1 + kNb * J;
  const int tx = t & 15, ty = t >> 4;
  // panel rows r0.. and c0.. (columns k .. k + 63), loaded in one batch, stored transposed
  double va[16], vb[16];
#pragma unroll
  for (int n = 0; n < 16; ++n) {
    const int rr = (t >> 6) + 4 * n, kk = t & 63;
    // clamped rows, loaded unconditionally (a load under a condition gets a branch of its own
    // that waits for it): a row past m only reaches outputs past m, which are not stored
    va[n] = A[min(r0 + rr, m - 1) * lda + k + kk];
    vb[n] = A[min(c0 + rr, m - 1) * lda + k + kk];
  }
  // the tile being updated, requested before the product
  double cv[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int64_t rw = r0 + 4 * ty + u;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t c = c0 + 4 * tx + v;
      const int64_t rwc = min(rw, m - 1);  // in the lower triangle; stored only if c <= rw < m
      cv[u][v] = A[rwc * lda + min(c, rwc)];
    }
  }
#pragma unroll
  for (int n = 0; n < 16; ++n) {
    const int rr = (t >> 6) + 4 * n, kk = t & 63;
    pi[kk][rr] = va[n];
    pj[kk][rr] = vb[n];
  }
  __syncthreads();
  double acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
#pragma unroll 4
  for (int kk = 0; kk < kNb; ++kk) {
    double av[4], bv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      av[u] = pi[kk][4 * ty + u];
      bv[u] = pj[kk][4 * tx + u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[u][v] = fma(av[u], bv[v], acc[u][v]);
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int64_t rw = r0 + 4 * ty + u;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t c = c0 + 4 * tx + v;
      if (rw < m && c <= rw) A[rw * lda + c] = cv[u][v] - acc[u][v];
    }
  }
  if (id != 0) return;
  // first tile = next diagonal block: factor it now (its stores above are visible to the
  // whole block after the barrier; the LDS of the product is reused)
  __syncthreads();
  CholShared& sm = *reinterpret_cast<CholShared*>(smem);
  chol_factor_diag(A + k1 * lda + k1, lda, (int)min((int64_t)kNb, m - k1), sm, info, k1, true);
  chol_store_l11(sm, l11t_next);
}

__global__ void chol_info_reset_kernel(int* info) {
  if (threadIdx.x == 0) info[0] = 0;
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

MAOMI_API int maomi_v4_apply_grid_f64(const double* du, int64_t U, double* table,
                                      const int32_t* ix, const int32_t* iy, const double* coef,
                                      double* out, int64_t nx, int64_t ny, int64_t mx,
                                      int64_t my, int64_t C, double* partial, int64_t rows,
                                      maomi_stream_t stream) {
  if (!du || !table || !ix || !iy || !coef || !out || U < 1 || U > 65536 || nx < 1 || ny < 1 ||
      ny > kGridMaxNy || nx > INT32_MAX / (ny > 0 ? ny : 1) || mx < 1 || my < 1 || C < 1 ||
      C > kMaxC || rows < 1 || rows > 65535 || (rows > 1 && !partial) ||
      (mx * my + 255) / 256 > INT32_MAX) {
    maomi_set_error("maomi_v4_apply_grid_f64: bad arguments (1 <= C <= %d, ny <= %d, U <= 65536)",
                    kMaxC, kGridMaxNy);
    return MAOMI_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  const int64_t tot = U * U;
  const unsigned tb = (unsigned)((tot + 255) / 256 < 65535 ? (tot + 255) / 256 : 65535);
  hipLaunchKernelGGL(green_table_kernel, dim3(tb), dim3(256), 0, st, du, U, table);
  MAOMI_CHECK_LAUNCH();
  const int64_t Q = mx * my;
  const int64_t rr = rows < nx ? rows : nx;  // at least one point column per block row
  const int jchunk = (int)((nx + rr - 1) / rr);
  const dim3 grid((unsigned)((Q + 255) / 256), (unsigned)((nx + jchunk - 1) / jchunk));
  double* dst = grid.y > 1 ? partial : out;
  // the table row in LDS when every block's queries share one query row and it fits
  const size_t idx_lds = (size_t)ny * 256 * sizeof(unsigned short);
  const bool row = my % 256 == 0 && idx_lds + (size_t)U * sizeof(double) <= 65536;
  const size_t lds = idx_lds + (row ? (size_t)U * sizeof(double) : 0);
  switch (C * 2 + (row ? 1 : 0)) {
#define CASE(c) \
  case 2 * c: hipLaunchKernelGGL((green_grid_kernel<c, false>), grid, dim3(256), lds, st, table, U, ix, iy, coef, dst, (int)nx, (int)ny, mx, my, jchunk); break; \
  case 2 * c + 1: hipLaunchKernelGGL((green_grid_kernel<c, true>), grid, dim3(256), lds, st, table, U, ix, iy, coef, dst, (int)nx, (int)ny, mx, my, jchunk); break;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
  }
  MAOMI_CHECK_LAUNCH();
  if (grid.y > 1) {
    const int64_t n = Q * C;
    const unsigned b = (unsigned)((n + 255) / 256 < 65535 ? (n + 255) / 256 : 65535);
    hipLaunchKernelGGL(reduce_chunks_kernel, dim3(b), dim3(256), 0, st, partial, (int)grid.y, n,
                       out);
    MAOMI_CHECK_LAUNCH();
  }
  return MAOMI_OK;
}

MAOMI_API int64_t maomi_cholesky_f64_workspace(void) {
  return 2 * (int64_t)kNb * kNb * (int64_t)sizeof(double);
}

MAOMI_API int maomi_cholesky_f64(double* A, int64_t lda, int64_t m, double* workspace, int* info,
                                 maomi_stream_t stream) {
  if (!info || m < 0 || lda < m || lda < 1 || (m > 0 && (!A || !workspace))) {
    maomi_set_error("maomi_cholesky_f64: bad arguments");
    return MAOMI_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(chol_info_reset_kernel, dim3(1), dim3(64), 0, st, info);
  MAOMI_CHECK_LAUNCH();
  if (m == 0) return MAOMI_OK;
  double* buf[2] = {workspace, workspace + kNb * kNb};
  hipLaunchKernelGGL(chol_diag_kernel, dim3(1), dim3(kCholThreads), 0, st, A, lda, m,
                     (int64_t)0, buf[0], info);
  MAOMI_CHECK_LAUNCH();
  for (int64_t k = 0, step = 0; k < m; k += kNb, ++step) {
    const int64_t b = m - k < kNb ? m - k : kNb, mp = m - k - b;
    double* cur = buf[step & 1];
    double* next = buf[(step + 1) & 1];
    if (mp > 0) {
      hipLaunchKernelGGL(chol_panel_kernel, dim3((unsigned)((mp + 63) / 64)), dim3(64), 0, st, A,
                         lda, m, k, cur);
      MAOMI_CHECK_LAUNCH();
    }
    const int64_t T = (mp + kNb - 1) / kNb;
    hipLaunchKernelGGL(chol_update_kernel, dim3((unsigned)(1 + T * (T + 1) / 2)),
                       dim3(kCholThreads), 0, st, A, lda, m, k, cur, next, info);
    MAOMI_CHECK_LAUNCH();
  }
  return MAOMI_OK;
}

namespace {

// ---------------------------------------------------------------------------------------------
// FP64 finish of the v4 solve G w = v from the factored projected system (maomi_v4_finish_f64).
// With Q = I - V T V^T (compact WY form of the QR reflectors of P = [1 x y]) and A = Q^T G Q,
// held in G: A11 = A[:3, :3], A12 = A[:3, 3:], and L (A22 = L L^T) in the lower triangle of
// A[3:, 3:].  Then, with y = Q^T v split into y1 (3 rows) and y2:
//   [Z z] = A22^-1 [A12^T y2]                 (v4_rhs_kernel, the two triangular sweeps)
//   a = (A11 - A12 Z)^-1 (y1 - A12 z)         (3 x 3 Schur complement, pivoted elimination)
//   x = [a; z - Z a],  w = Q x                (v4_schur_kernel)
// The triangular sweeps go in panels of kTs rows, one launch per panel: every block subtracts
// the last solved panel's contribution from 64 trailing rows, and block 0 (which holds the next
// panel's rows) then solves that panel's diagonal block, a wave per group of right-hand-side
// columns with the rows on the lanes.  R = 3 + C right-hand sides, row-major in the workspace.
constexpr int kTs = 64;
constexpr int kTsThreads = 256;  // 4 waves: wave g holds columns g, g + 4, g + 8, g + 12
constexpr int kTsCols = 4;       // columns per thread
constexpr int kMaxRhs = kTsThreads / 64 * kTsCols;  // 16
constexpr int kFinThreads = 1024;

struct TrsShared {
  double tile[kTs][kTs + 1];  // L tile of the update (stride 65 doubles against bank conflicts)
  double diag[kTs][kTs + 1];  // the next diagonal block (block 0)
  double pan[kTs][kMaxRhs];   // solved rows of the last panel
};


// Sweep step of L Y = B (forward) or L^T X = Y (backward) on B (m x R, row-major).
// Forward, panel s = step: blocks update rows r0 = (s + 1 + k) kTs ...; backward, panel p =
// step: rows r0 = (p - 1 - k) kTs ...  With diag_only, block 0 solves the diagonal block of
// panel `step` itself (the first panel of the sweep) and nothing is updated.
// Block 0's diagonal solve keeps its lane's row (forward) or column (backward) of the block in
// registers and the reciprocal of its diagonal entry: one step of the substitution is a lane
// broadcast, a multiply and a fused multiply-add.  Rows past m are padded with the identity.
template <bool FWD>
__global__ __launch_bounds__(kTsThreads) void trs_step_kernel(const double* __restrict__ L,
                                                              int64_t ldl, int64_t m,
                                                              int64_t step, int diag_only,
                                                              double* __restrict__ B, int R) {
  __shared__ TrsShared sm;
  const int tid = threadIdx.x, lane = tid & 63, g = tid >> 6;
  int64_t r0;
  if (diag_only)
    r0 = step * kTs;
  else
    r0 = FWD ? (step + 1 + (int64_t)blockIdx.x) * kTs : (step - 1 - (int64_t)blockIdx.x) * kTs;
  const int rows = (int)(m - r0 < kTs ? m - r0 : kTs);  // >= 1 by the grid's construction
  const int kc = (R - g + 3) / 4;                       // columns of this wave (g + 4 k < R)
  // Tiles are staged with every load of a thread in flight at once (unrolled into registers,
  // then stored to LDS): a loop that waits for each load before the next one is issued spends
  // 16 memory latencies per tile.  So does a load under a condition (it gets a branch of its
  // own that waits for it): the loads use clamped, in-bounds indices and the values are
  // selected afterwards.
  constexpr int kPer = kTs * kTs / kTsThreads;  // tile elements per thread
  double dv[kPer];
  if (blockIdx.x == 0) {  // requested first: its loads overlap the update's
#pragma unroll
    for (int it = 0; it < kPer; ++it) {
      const int e = tid + it * kTsThreads, i = e / kTs, j = e % kTs;
      const int ic = min(i, rows - 1);
      dv[it] = L[(r0 + ic) * ldl + r0 + min(j, ic)];  // padded when staged
    }
  }
  double b[kTsCols];
#pragma unroll
  for (int k = 0; k < kTsCols; ++k) {
    const int c = g + 4 * k;
    // clamped (finite) values for lanes past the block and columns past R: those are never
    // stored, and in the diagonal solve they meet only the zero entries of the padded block
    b[k] = B[(r0 + min(lane, rows - 1)) * R + min(c, R - 1)];
  }
  if (!diag_only) {
    const int64_t p0 = step * kTs;                               // rows of the solved panel
    const int bp = (int)(m - p0 < kTs ? m - p0 : kTs);
    double tv[kPer];
    constexpr int kPan = kTs * kMaxRhs / kTsThreads;
    double pv[kPan];
#pragma unroll
    for (int it = 0; it < kPer; ++it) {
      const int e = tid + it * kTsThreads, i = e / kTs, j = e % kTs;
      // rows past m are clamped: forward, they reach only this block's rows past m; backward,
      // they meet the zero rows of the staged panel
      if (FWD)  // tile[i][j] = L[r0 + i][p0 + j]: this block's rows, the panel's columns
        tv[it] = L[(r0 + min(i, rows - 1)) * ldl + p0 + j];
      else      // tile[j][i] = L[p0 + j][r0 + i]: the panel's rows, this block's columns
        tv[it] = L[(p0 + min(i, bp - 1)) * ldl + r0 + j];
    }
#pragma unroll
    for (int it = 0; it < kPan; ++it) {
      const int e = tid + it * kTsThreads, j = e / kMaxRhs, c = e % kMaxRhs;
      pv[it] = B[(p0 + min(j, bp - 1)) * R + min(c, R - 1)];
    }
#pragma unroll
    for (int it = 0; it < kPer; ++it) {
      const int e = tid + it * kTsThreads;
      sm.tile[e / kTs][e % kTs] = tv[it];
    }
#pragma unroll
    for (int it = 0; it < kPan; ++it) {
      const int e = tid + it * kTsThreads, j = e / kMaxRhs, c = e % kMaxRhs;
      sm.pan[j][c] = (j < bp && c < R) ? pv[it] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int j = 0; j < kTs; ++j) {
      const double l = FWD ? sm.tile[lane][j] : sm.tile[j][lane];
#pragma unroll
      for (int k = 0; k < kTsCols; ++k)
        if (k < kc) b[k] -= l * sm.pan[j][g + 4 * k];
    }
  }
  if (blockIdx.x == 0) {
#pragma unroll
    for (int it = 0; it < kPer; ++it) {
      const int e = tid + it * kTsThreads, i = e / kTs, j = e % kTs;
      sm.diag[i][j] = (i < rows && j <= i) ? dv[it] : (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();  // the diagonal block is staged
    double d[kTs];    // forward: d[j] = D[lane][j]; backward: d[j] = D[j][lane]
#pragma unroll
    for (int j = 0; j < kTs; ++j) d[j] = FWD ? sm.diag[lane][j] : sm.diag[j][lane];
    const double inv = 1.0 / sm.diag[lane][lane];
#pragma unroll
    for (int k = 0; k < kTsCols; ++k) {
      if (k >= kc) continue;  // uniform in the wave
      double x = b[k];
      if (FWD) {
#pragma unroll
        for (int j = 0; j < kTs; ++j) {  // y_j = b_j / D_jj, then b_i -= D_ij y_j for i > j
          const double yj = readlane_f64(x, j) * readlane_f64(inv, j);
          const double t = fma(-d[j], yj, x);
          x = lane > j ? t : x;
          x = lane == j ? yj : x;
        }
      } else {
#pragma unroll
        for (int j = kTs - 1; j >= 0; --j) {  // x_j = b_j / D_jj, then b_i -= D_ji x_j, i < j
          const double xj = readlane_f64(x, j) * readlane_f64(inv, j);
          const double t = fma(-d[j], xj, x);
          x = lane < j ? t : x;
          x = lane == j ? xj : x;
        }
      }
      b[k] = x;
    }
  }
#pragma unroll
  for (int k = 0; k < kTsCols; ++k) {
    const int c = g + 4 * k;
    if (lane < rows && c < R) B[(r0 + lane) * R + c] = b[k];
  }
}

// Block-wide sums of n <= 64 per-thread values in place (every thread receives every sum).
__device__ void block_sums(double* v, int n, double* red /* [kFinThreads / 64][64] */,
                           double* out /* [64] */) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int i = 0; i < n; ++i) {
    double x = v[i];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) red[wave * 64 + i] = x;
  }
  __syncthreads();
  if (tid < n) {
    double t = 0.0;
    for (int w = 0; w < kFinThreads / 64; ++w) t += red[w * 64 + tid];
    out[tid] = t;
  }
  __syncthreads();
  for (int i = 0; i < n; ++i) v[i] = out[i];
  __syncthreads();
}

// y = Q^T v = v - V (T^T (V^T v)); y1 = y[:3] -> y1 out; B[i] = [A12^T[i], y[3 + i]].
__global__ __launch_bounds__(kFinThreads) void v4_rhs_kernel(
    const double* __restrict__ G, int64_t ldg, int64_t n, const double* __restrict__ V,
    const double* __restrict__ T, const double* __restrict__ v, int C, double* __restrict__ B,
    double* __restrict__ y1) {
  __shared__ double red[kFinThreads];
  __shared__ double out[64];
  const int tid = threadIdx.x, R = 3 + C;
  double u[3 * kMaxRhs];
  for (int e = 0; e < 3 * C; ++e) u[e] = 0.0;
  for (int64_t i = tid; i < n; i += kFinThreads)
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < C; ++c) u[k * C + c] += V[i * 3 + k] * v[i * C + c];
  block_sums(u, 3 * C, red, out);
  double t[3 * kMaxRhs];  // T^T u
  for (int k = 0; k < 3; ++k)
    for (int c = 0; c < C; ++c)
      t[k * C + c] = T[0 * 3 + k] * u[0 * C + c] + T[1 * 3 + k] * u[1 * C + c] +
                     T[2 * 3 + k] * u[2 * C + c];
  for (int64_t i = tid; i < n; i += kFinThreads) {
    for (int c = 0; c < C; ++c) {
      const double y = v[i * C + c] - (V[i * 3 + 0] * t[0 * C + c] + V[i * 3 + 1] * t[1 * C + c] +
                                       V[i * 3 + 2] * t[2 * C + c]);
      if (i < 3) y1[i * C + c] = y;
      else B[(i - 3) * R + 3 + c] = y;
    }
    if (i >= 3)
      for (int k = 0; k < 3; ++k) B[(i - 3) * R + k] = G[k * ldg + i];
  }
}

// a = (A11 - A12 Z)^-1 (y1 - A12 z); x = [a; z - Z a]; w = x - V (T (V^T x)).
__global__ __launch_bounds__(kFinThreads) void v4_schur_kernel(
    const double* __restrict__ G, int64_t ldg, int64_t n, const double* __restrict__ V,
    const double* __restrict__ T, const double* __restrict__ X, const double* __restrict__ y1,
    int C, double* __restrict__ w) {
  __shared__ double red[kFinThreads];
  __shared__ double out[64];
  __shared__ double av[3 * kMaxRhs];
  const int tid = threadIdx.x, R = 3 + C;
  const int64_t m = n - 3;
  // S = A12 Z (9 sums) and r = A12 z (3 C sums)
  double s[9 + 3 * kMaxRhs];
  const int ns = 9 + 3 * C;
  for (int e = 0; e < ns; ++e) s[e] = 0.0;
  for (int64_t i = tid; i < m; i += kFinThreads)
    for (int k = 0; k < 3; ++k) {
      const double a12 = G[k * ldg + 3 + i];
      for (int l = 0; l < R; ++l) s[k * R + l] += a12 * X[i * R + l];
    }
  block_sums(s, ns, red, out);  // s[k R + l]: (A12 X)[k][l]
  if (tid == 0) {
    // S a = r with partial pivoting; S = A11 - A12 Z, r = y1 - A12 z
    double S[3][3], r[3][kMaxRhs];
    for (int k = 0; k < 3; ++k) {
      for (int l = 0; l < 3; ++l) S[k][l] = G[k * ldg + l] - s[k * R + l];
      for (int c = 0; c < C; ++c) r[k][c] = y1[k * C + c] - s[k * R + 3 + c];
    }
    for (int col = 0; col < 3; ++col) {
      int piv = col;
      for (int k = col + 1; k < 3; ++k)
        if (fabs(S[k][col]) > fabs(S[piv][col])) piv = k;
      if (piv != col) {
        for (int l = 0; l < 3; ++l) { const double t = S[col][l]; S[col][l] = S[piv][l]; S[piv][l] = t; }
        for (int c = 0; c < C; ++c) { const double t = r[col][c]; r[col][c] = r[piv][c]; r[piv][c] = t; }
      }
      for (int k = col + 1; k < 3; ++k) {
        const double f = S[k][col] / S[col][col];
        for (int l = col; l < 3; ++l) S[k][l] -= f * S[col][l];
        for (int c = 0; c < C; ++c) r[k][c] -= f * r[col][c];
      }
    }
    for (int k = 2; k >= 0; --k)
      for (int c = 0; c < C; ++c) {
        double t = r[k][c];
        for (int l = k + 1; l < 3; ++l) t -= S[k][l] * av[l * C + c];
        av[k * C + c] = t / S[k][k];
      }
  }
  __syncthreads();
  // x into w, and u = V^T x
  double u[3 * kMaxRhs];
  for (int e = 0; e < 3 * C; ++e) u[e] = 0.0;
  for (int64_t i = tid; i < n; i += kFinThreads)
    for (int c = 0; c < C; ++c) {
      double x;
      if (i < 3) {
        x = av[i * C + c];
      } else {
        const double* Xi = X + (i - 3) * R;
        x = Xi[3 + c] - (Xi[0] * av[0 * C + c] + Xi[1] * av[1 * C + c] + Xi[2] * av[2 * C + c]);
      }
      w[i * C + c] = x;
      for (int k = 0; k < 3; ++k) u[k * C + c] += V[i * 3 + k] * x;
    }
  block_sums(u, 3 * C, red, out);
  double t[3 * kMaxRhs];  // T u
  for (int k = 0; k < 3; ++k)
    for (int c = 0; c < C; ++c)
      t[k * C + c] = T[k * 3 + 0] * u[0 * C + c] + T[k * 3 + 1] * u[1 * C + c] +
                     T[k * 3 + 2] * u[2 * C + c];
  for (int64_t i = tid; i < n; i += kFinThreads)  // each thread rereads the x it wrote
    for (int c = 0; c < C; ++c)
      w[i * C + c] -= V[i * 3 + 0] * t[0 * C + c] + V[i * 3 + 1] * t[1 * C + c] +
                      V[i * 3 + 2] * t[2 * C + c];
}

}  // namespace

MAOMI_API int64_t maomi_v4_finish_f64_workspace(int64_t n, int64_t C) {
  if (n <= 3 || C < 1 || C > MAOMI_V4_FINISH_MAX_C) return -1;
  return ((n - 3) * (3 + C) + 3 * C) * (int64_t)sizeof(double);
}

MAOMI_API int maomi_v4_finish_f64(const double* G, int64_t ldg, int64_t n, const double* V,
                                  const double* T, const double* v, int64_t C, double* w,
                                  double* workspace, maomi_stream_t stream) {
  if (!G || !V || !T || !v || !w || !workspace || n <= 3 || ldg < n || C < 1 ||
      C > MAOMI_V4_FINISH_MAX_C) {
    maomi_set_error("maomi_v4_finish_f64: bad arguments (n > 3, ldg >= n, 1 <= C <= %d)",
                    MAOMI_V4_FINISH_MAX_C);
    return MAOMI_EINVAL;
  }
  static_assert(3 + MAOMI_V4_FINISH_MAX_C <= kMaxRhs, "right-hand sides per sweep");
  hipStream_t st = (hipStream_t)stream;
  const int64_t m = n - 3;
  const int R = (int)(3 + C);
  double* B = workspace;
  double* y1 = workspace + m * R;
  const double* L = G + 3 * ldg + 3;
  hipLaunchKernelGGL(v4_rhs_kernel, dim3(1), dim3(kFinThreads), 0, st, G, ldg, n, V, T, v,
                     (int)C, B, y1);
  MAOMI_CHECK_LAUNCH();
  const int64_t P = (m + kTs - 1) / kTs;
  // forward sweep: panel 0's diagonal block, then one launch per panel s < P - 1
  hipLaunchKernelGGL(trs_step_kernel<true>, dim3(1), dim3(kTsThreads), 0, st, L, ldg, m,
                     (int64_t)0, 1, B, R);
  MAOMI_CHECK_LAUNCH();
  for (int64_t s = 0; s + 1 < P; ++s) {
    const unsigned blocks = (unsigned)((m - (s + 1) * kTs + kTs - 1) / kTs);
    hipLaunchKernelGGL(trs_step_kernel<true>, dim3(blocks), dim3(kTsThreads), 0, st, L, ldg, m,
                       s, 0, B, R);
    MAOMI_CHECK_LAUNCH();
  }
  // backward sweep: the last panel's diagonal block, then one launch per panel p >= 1
  hipLaunchKernelGGL(trs_step_kernel<false>, dim3(1), dim3(kTsThreads), 0, st, L, ldg, m,
                     P - 1, 1, B, R);
  MAOMI_CHECK_LAUNCH();
  for (int64_t p = P - 1; p >= 1; --p) {
    hipLaunchKernelGGL(trs_step_kernel<false>, dim3((unsigned)p), dim3(kTsThreads), 0, st, L,
                       ldg, m, p, 0, B, R);
    MAOMI_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(v4_schur_kernel, dim3(1), dim3(kFinThreads), 0, st, G, ldg, n, V, T, B, y1,
                     (int)C, w);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

// ---------------------------------------------------------------------------------------------
// maomi_v4_apply_split_f64 without a caller-supplied partial-sum buffer, for callers that cannot
// allocate device memory.  The rows * Q * C partial sums live in `scratch`, a file-scope device
// array of kV4StaticPartials doubles (32 MiB) that each GPU holds until the process ends.  Host
// code never passes the array's name to a kernel: it is a host-side stand-in.  Its device
// address is looked up once per device (hipGetSymbolAddress, with that device current) on the
// first call there and kept in g_scratch.  Every partial sum is written before it is read, so
// nothing needs clearing between calls; calls that overlap on one GPU share the array and race.
namespace {

constexpr int64_t kV4StaticPartials = MAOMI_V4_STATIC_PARTIALS;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kV4StaticPartials];

std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// Device address of this GPU's scratch array, or nullptr (message set) on failure.
double* device_scratch() {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    maomi_set_error("maomi_v4_apply_split_f64_noalloc: no current device (%s)",
                    hipGetErrorString(e));
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      maomi_set_error("maomi_v4_apply_split_f64_noalloc: hipGetSymbolAddress failed on device "
                      "%d (%s)", dev, hipGetErrorString(e));
      return nullptr;
    }
    g_scratch[dev] = static_cast<double*>(ptr);
  }
  return g_scratch[dev];
}

}  // namespace

MAOMI_API int maomi_v4_apply_split_f64_noalloc(const double* points, const double* coef,
                                               const double* queries, double* out, int64_t P,
                                               int64_t Q, int64_t C, int64_t rows,
                                               maomi_stream_t stream) {
  if (!points || !coef || !queries || !out || P < 1 || Q < 0 || C < 1 || C > kMaxC ||
      rows < 1 || rows > 65535) {
    maomi_set_error("maomi_v4_apply_split_f64_noalloc: bad arguments (1 <= C <= %d)", kMaxC);
    return MAOMI_EINVAL;
  }
  if (rows > 1 && Q > kV4StaticPartials / (rows * C)) {
    maomi_set_error("maomi_v4_apply_split_f64_noalloc: rows * Q * C = %lld partial sums exceed "
                    "the static buffer of %lld; use fewer rows or split the queries",
                    (long long)(rows * C) * (long long)Q, (long long)kV4StaticPartials);
    return MAOMI_EINVAL;
  }
  if (Q == 0) return MAOMI_OK;
  double* part = nullptr;
  if (rows > 1) {
    part = device_scratch();
    if (!part) return MAOMI_EINVAL;
  }
  return maomi_v4_apply_split_f64(points, coef, queries, out, P, Q, C, part, rows, stream);
}
