// Persistent per-site split-step propagation for the plain optical stage.
//
// One block owns one site at a time and runs all of its slices: the field lives in a
// block-private workspace (L2/MALL resident) between passes over LDS tiles of lines.  The
// Fresnel step is separable and its two axis factors commute, so one pass along an axis
// finishes the previous slice's step along that axis, applies the plane's screen exp(i phi)
// times the edge window (read straight from the phase volume) and its |u|^4 trapezoid term,
// and starts the next slice's step along that axis (forward FFT, 1-D Fresnel factor with 1/N
// folded in, inverse FFT); the passes alternate between rows and columns.  A site of S slices
// takes S + 1 passes, each reading and writing the field once.  Reference behaviour:
// maomi/numerics.py:156-210 (fresnel, constant pitch), maomi/backends/cuda/batched.py:157-168
// (_stage), maomi/optics.py:327-340 (trapezoid over saved planes).
#include "fft_reg.hpp"
#include "maomi_amd.h"
#include "screen_trig.hpp"

namespace {

// Workgroup barrier that orders LDS only: global loads and stores stay in flight.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// A middle pass requests the next tile's field lines right after the current tile's first
// stage, so that their latency overlaps the rest of the tile (config 2: roofline 0.437 -> 0.461).
constexpr bool kPrefetch = true;

// A pass requests a tile's screen phases and mask pixels at the start of the tile; a middle
// pass requests those of the next tile right after the current tile's pointwise stage.
constexpr bool kHoist = true;
constexpr bool kHoistAhead = true;  // config 2: roofline 0.476 -> 0.492

// Raw buffer access for the per-site arrays: a resource over [p, p + bytes) held in scalar
// registers, a 32-bit byte offset per lane in one VGPR and a wave-uniform byte offset (an SGPR
// or the instruction's immediate), so that the 16 or 32 strided elements of a line job cost no
// 64-bit address registers each.  Loads outside the range return zero; stores are dropped.
using Rsrc = __amdgpu_buffer_rsrc_t;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
constexpr int kRawBufferWord3 = 0x00020000;  // gfx9 raw buffer, 32-bit data format
// The base and size are wave-uniform by construction (per-site or per-block values); they are
// read from the first lane so that the resource is known to live in scalar registers (a
// resource the compiler takes for divergent costs a waterfall loop around every access).
__device__ __forceinline__ Rsrc make_rsrc(const void* p, uint32_t bytes) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  void* base = reinterpret_cast<void*>(((uint64_t)hi << 32) | lo);
  return __builtin_amdgcn_make_buffer_rsrc(base, 0, (int)__builtin_amdgcn_readfirstlane(bytes),
                                           kRawBufferWord3);
}
// Default cache policy: the non-temporal hint on the field or the mask traffic measured 10-20%
// slower on config 2 (their round trips are served from the caches between passes).
__device__ __forceinline__ float2 ld2(Rsrc r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ float ld1(Rsrc r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ float ldh(Rsrc r, uint32_t voff, uint32_t soff) {
  const unsigned short b = __builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0);
  return __half2float(__ushort_as_half(b));
}
__device__ __forceinline__ void st2(Rsrc r, uint32_t voff, uint32_t soff, float2 v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
}
__device__ __forceinline__ void st1(Rsrc r, uint32_t voff, uint32_t soff, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
}

// The value of the other lane of an even/odd lane pair (DPP quad permutation 1, 0, 3, 2).
__device__ __forceinline__ float pair_swap(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, false));
}

// With two jobs per thread, the scheduler would interleave them across a whole tile (one basic
// block) and hold both jobs' loads and transforms at once; this keeps each job's code together.
template <int J>
__device__ __forceinline__ void job_fence() {
  if constexpr (J > 1) __builtin_amdgcn_sched_barrier(0);
}

__device__ __forceinline__ void screen_sincos(float phi, float* s, float* c) {
  trig::sincos_screen(phi, s, c);  // |phi| < 2^17, checked on the host
}

template <int N> struct ClassIdx;
template <> struct ClassIdx<64> { static constexpr int v = 0; };
template <> struct ClassIdx<128> { static constexpr int v = 1; };
template <> struct ClassIdx<256> { static constexpr int v = 2; };
template <> struct ClassIdx<512> { static constexpr int v = 3; };

// Block shapes: NT threads transform the L lines of an LDS tile (fftk::Tile).
// Exchange-buffer slots for field sides up to NMAX: two buffers per side (the passes alternate
// between them, see line_pass).
template <int NT, int NMAX>
constexpr int lds_slots() {
  int m = fftk::cmax(fftk::Tile<64, NT>::LDS2, fftk::Tile<128, NT>::LDS2);
  if (NMAX >= 256) m = fftk::cmax(m, fftk::Tile<256, NT>::LDS2);
  if (NMAX >= 512) m = fftk::cmax(m, fftk::Tile<512, NT>::LDS2);
  return 2 * m;
}
constexpr int kNmax = 512;

struct SitesArgs {
  const float2* bases;
  const int64_t* base_off;
  const float* phase;
  const __half* phase16;
  const float2* screen;
  const float* slice_scale;
  const float* edge;
  const float2* q1d;
  const float2* twid;
  const int32_t* site_x;
  const int32_t* site_y;
  const int32_t* site_n;
  const int32_t* site_s;
  const int32_t* order;
  float* mask;
  const int64_t* mask_off;
  double* energy;
  float2* scratch;
  int64_t scratch_stride;
  int64_t P, H, W, Z;
  int64_t cedge[5], cq[5], ctw[5];
  float* psf_out;  // optional [psf_nz][P][psf_w][psf_w] planes |u|^4 around focus
  int psf_z0, psf_nz, psf_c0, psf_w;
  int* queue;      // dynamic work counter, or nullptr for static striding
  // In-launch audit of surrogate predictions (focal-plane launches): the first n_audit
  // entries of the order list are the audited sites.
  const float* audit_pred;     // [psf_nz][n_audit][psf_w][psf_w], or nullptr
  const int32_t* audit_slot;   // [P] slot of an audited site, -1 otherwise
  int* audit_state;            // [2] audited sites done, audited sites above tol
  float* audit_rel;            // [n_audit] relative L2 of each audited site
  int32_t* site_done;          // [P] 1 once a site is propagated
  int n_audit, audit_max_fail;
  float audit_tol;
};

// True once every audited site is done and no more than audit_max_fail of them missed the
// tolerance: the predictions are trusted and blocks stop taking unaudited sites.
__device__ __forceinline__ bool audit_passed(const SitesArgs& a) {
  const int done = __hip_atomic_load(&a.audit_state[0], __ATOMIC_ACQUIRE,
                                     __HIP_MEMORY_SCOPE_AGENT);
  const int fail = __hip_atomic_load(&a.audit_state[1], __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
  return done == a.n_audit && fail <= a.audit_max_fail;
}

// Site loop of the uniform kernels' dynamic form: take_site returns the index into the order
// list of the block's next site, or a value >= P when the block is done.  With a work counter
// each block takes the next entry as it finishes a site (greedy list scheduling); every block
// leaves once the counter passes P.  Without one, block b takes entries b, b + grid, ...
// Either way, an audited launch stops taking unaudited sites once the audit has passed.
__device__ __forceinline__ int64_t take_site(const SitesArgs& a, int64_t prev, int* next) {
  const int64_t step = prev < 0 ? (int64_t)blockIdx.x : prev + gridDim.x;
  if (threadIdx.x == 0) {
    int64_t i = a.queue ? (int64_t)atomicAdd(a.queue, 1) : step;
    // an unaudited site after a passed audit is left to the surrogate
    if (a.audit_pred && i >= a.n_audit && i < a.P && audit_passed(a)) i = a.P;
    *next = (int)(i < a.P ? i : a.P);
  }
  __syncthreads();
  // the same value in every lane: readfirstlane keeps it, and the site addresses derived from
  // it, in scalar registers
  const int64_t i = __builtin_amdgcn_readfirstlane(*next);
  __syncthreads();  // every thread has read `next` before thread 0 takes the following one
  return i;
}

__global__ void queue_reset_kernel(int* queue) {
  if (threadIdx.x == 0) *queue = 0;
}

// Zeroes the audit counters and the per-site done flags before an audited launch.
__global__ void audit_reset_kernel(int* state, int32_t* done, int64_t P) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < 2) state[i] = 0;
  for (int64_t j = i; j < P; j += (int64_t)gridDim.x * blockDim.x) done[j] = 0;
}

// Relative L2 of an audited site's propagated focal planes against its prediction; counts a
// miss of the tolerance, then the site as done (release: a block that sees every audited site
// done also sees every miss).
template <class SM>
__device__ void audit_site(const SitesArgs& a, int site, int slot, int tid, int nt, SM& sm) {
  const int64_t plane = (int64_t)a.psf_w * a.psf_w;
  double num = 0.0, den = 0.0;
  for (int64_t e = tid; e < a.psf_nz * plane; e += nt) {
    const int64_t z = e / plane, k = e - z * plane;
    const float t = a.psf_out[(z * a.P + site) * plane + k];
    const float p = a.audit_pred[(z * a.n_audit + slot) * plane + k];
    num += (double)(p - t) * (p - t);
    den += (double)t * t;
  }
  num = block_sum_double(num, sm.red);
  den = block_sum_double(den, sm.red);
  if (tid == 0) {
    const float rel = den > 0.0 ? (float)sqrt(num / den) : INFINITY;
    a.audit_rel[slot] = rel;
    if (!(rel <= a.audit_tol))
      __hip_atomic_fetch_add(&a.audit_state[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(&a.audit_state[0], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Tables of a block.  With room for the 512 side (TAB = 512) they also hold what its lane
// pairs read: lane 1 of a pair reads its own copies of the twiddles and Fresnel factors, 8
// slots further along the LDS banks than lane 0's, so that the two lanes of a pair do not
// meet on a bank.
template <int SLOTS, int TAB, int NT>
struct SmemT {
  static constexpr bool kPair = TAB >= 512;
  // the uniform 256 side: W256^(a b) at twr[17 a + b] (a, b < 16), the inter-stage twiddles of
  // a job read at immediate offsets from its row, neighbouring jobs' rows on different banks
  static constexpr bool kRowTw = TAB == 256;
  // the 512 side: W512^(k2 (n + 16 h)) at tw[kTwR + h kH + 16 k2 + n] (k2, n < 16), and
  // q[k2 + 16 (2 k + h)] at q[h kH + k2 + 16 k]
  static constexpr int kTwR = TAB + 8, kH = TAB / 2 + 8;
  float2 x[SLOTS];                          // line exchange buffer(s)
  float2 tw[kPair ? kTwR + kH + TAB / 2 : TAB];  // W_N^j
  float2 q[kPair ? TAB + 8 : TAB];          // 1-D Fresnel factor (FFT order, times 1/N)
  float2 twr[kRowTw ? 16 * 17 : 1];
  float eg[TAB];                            // 1-D edge window
  double red[NT / 64];
};
// Mixed launches: room for every field side up to NMAX.
template <int NT, int NMAX>
using Smem = SmemT<lds_slots<NT, NMAX>(), kNmax, NT>;
// Uniform launches: exchange buffers and tables of the one field side N, so that two
// 256-thread blocks of the 256 side fit one CU (2 x 75 KiB of the 160 KiB).
template <int N, int NT>
using SmemU = SmemT<2 * fftk::Tile<N, NT>::LDS2, N, NT>;

// Launch modes, compile-time so that each kernel carries only its own address arithmetic:
// FP32 screens; FP16 screens with a per-slice scale; FP32 screens plus the focal-window
// planes of |u|^4 (a.psf_out); complex64 screens exp(i phi) computed once per volume voxel
// (the same sin/cos as the per-site path, so results are identical; the screen read doubles
// but the per-pixel sin/cos, shared by every site whose window covers the pixel, is gone).
enum Mode { kScreens32 = 0, kScreens16 = 1, kScreens32Psf = 2, kScreensC64 = 3, kScreens32Mixed = 4 };

// Per-site state shared by the passes.
struct SiteCtx {
  int site;
  int64_t ox, oy;
  const float2* base;
  float* mask;
  float2* fld;
};

// Passes of a site.  The Fresnel step is separable, P = Px Py, with Px (along x) and Py (along
// y) commuting, so the planes u_{p+1} = P T_p u_p (T_p: screen times edge window of plane p)
// can be produced by one pass per plane, alternating the axis:
//   pass 0 (x): T_0 and mask(0), then Px;
//   pass p (y for odd p, x for even p): the other half step of slice p-1 (the field is then
//     u_p), T_p and mask(p), then this axis' half step of slice p;
//   pass S: the remaining half step of slice S-1, mask(S).
// Pass p + 1 starts with the half step that pass p did not do, and the two halves of a slice
// commute, so pass p + 1 holds u_{p+1}.  Every pass moves the field through memory once: S
// slices cost S + 1 field round trips and 4 S line transforms.
enum PassKind { kHead = 0, kMid = 1, kTail = 2, kOnly = 3 };
enum Axis { kAxisX = 0, kAxisY = 1 };  // the coordinate that varies along a line

// A line pass over tiles of L lines.  A tile runs: (A) stage 1 of the forward transform,
// global -> xa; (B) stage 2, the 1-D Fresnel factor and stage 1 of the inverse, xa -> xb;
// then for a middle pass (C) stage 2 of the inverse, the pointwise work of plane p and stage 1
// of the forward transform, xb -> xa, and (B) again; finally stage 2 of the inverse,
// xb -> global.  Stage-1 and final jobs hold the same elements: offsets n1 + N1 m, m < N2,
// along line l1.  A stage writes the exchange buffer that the previous stage did not read, so
// a tile needs one barrier per stage boundary.
// Barriers inside a pass only order LDS; each thread's mask pixels and the phase volume need
// no cross-thread ordering.
template <int N, int NT, int MODE, int AXIS, int KIND, class SM>
__device__ __forceinline__ void line_pass(const SitesArgs& a, const SiteCtx& c, int p, int tid,
                                          SM& sm, double& esum) {
  constexpr bool half = MODE == kScreens16, PSF = MODE == kScreens32Psf;
  constexpr bool cplx = MODE == kScreensC64;
  constexpr bool FIRST = KIND == kHead || KIND == kOnly, LAST = KIND == kTail || KIND == kOnly;
  constexpr bool ROWS = AXIS == kAxisY;  // lines are rows x, elements along y
  using T = fftk::Tile<N, NT>;
  constexpr int N1 = T::N1, N2 = T::N2, L = T::L, LS = T::LS, J1 = T::J1, J2 = T::J2;
  // a tile's screen phases and mask pixels are requested at its start, and a middle pass
  // requests the next tile's field lines after its first stage (one stage-1 job per thread;
  // the focal-plane launches keep those registers for the output addressing)
  constexpr bool lean = MODE == kScreens32Mixed;
  constexpr bool hoist = kHoist && J1 == 1;
  // (FP16 screens: the converted phases held a tile ahead cost more registers than they save
  // in latency, 88.0 against 85.5 ms on config 4)
  // (the 512 side: neither the next tile's lines nor its phases and mask pixels are held
  // across the 32-point stage, whose lines take twice the registers; with the lines held, the
  // uniform 512 launch of 576 sites ran 109.6 against 98.9 ms)
  constexpr bool ahead =
      kHoistAhead && hoist && KIND == kMid && !PSF && !lean && !half && N <= 256;
  constexpr bool prefetch = kPrefetch && KIND == kMid && J1 == 1 && N <= 256 && !PSF && !lean;
  float2* const xa = sm.x;
  float2* const xb = sm.x + T::LDS2;
  // the inter-stage twiddle W_N^(a b) of job a, element b
  auto tw_at = [&](int a, int b) {
    if constexpr (SM::kRowTw && N == 256) return sm.twr[17 * a + b];
    else return sm.tw[a * b];
  };
  // element m of a stage-1 job lies at (x0 + DX m, y0 + DY m); GS is its index stride in the
  // field, the mask and the entry field, PS in a slice of the phase volume
  constexpr int DX = ROWS ? 0 : N1, DY = ROWS ? N1 : 0;
  constexpr int64_t GS = (int64_t)DX * N + DY;
  constexpr int64_t TS = ROWS ? (int64_t)L * N : L;  // field index step from tile to tile
  const int64_t PS = (int64_t)DX * a.W + DY;
  // byte strides between a job's elements: field and entry field, mask, screen slice
  constexpr uint32_t GS8 = (uint32_t)GS * 8, GS4 = (uint32_t)GS * 4;
  constexpr int SE = cplx ? 8 : half ? 2 : 4;  // bytes per screen element
  const uint32_t PSB = (uint32_t)PS * SE;
  const Rsrc r_fld = make_rsrc(c.fld, (uint32_t)N * N * 8);
  const Rsrc r_base = make_rsrc(c.base, (uint32_t)N * N * 8);
  const Rsrc r_mask = make_rsrc(c.mask, (uint32_t)N * N * 4);
  const int64_t slice = LAST ? 0 : (int64_t)p * a.H * a.W;  // plane p's screen
  const Rsrc r_scr = make_rsrc(cplx ? (const void*)(a.screen + slice)
                               : half ? (const void*)(a.phase16 + slice)
                                      : (const void*)(a.phase + slice),
                               (uint32_t)(a.H * a.W * SE));
  const float wgt = (FIRST || LAST) ? 0.5f : 1.0f;
  const float sscale = half && !LAST ? a.slice_scale[p] : 1.0f;
  // stage-1 job jj of this thread: line l1 of the tile, offset n1 (rows: consecutive lanes take
  // consecutive offsets; columns: consecutive lines, for coalesced global access).  With two
  // jobs per thread, job jj covers lines [jj L / J1, (jj + 1) L / J1).
  constexpr int LJ = L / J1;  // lines of one job
  auto job1 = [&](int jj, int& l1, int& n1) {
    const int jb = tid + jj * NT;
    l1 = ROWS ? jb / N1 : jj * LJ + tid % LJ;
    n1 = ROWS ? jb % N1 : tid / LJ;
  };
  float2 vpre[prefetch ? N2 : 1];  // the next tile's stage-1 input

  // stage B: xin -> xout (stage 2 of the forward transform, Fresnel factor, stage 1 of the
  // inverse); job jj covers lines [jj NT / N2, (jj + 1) NT / N2) when a thread has two.
  // The uniform 256 side holds its stage-B job's Fresnel factors in registers for the whole
  // pass: a thread's k2 is the same in every tile.
  constexpr bool qreg = SM::kRowTw && N == 256 && J2 == 1 && !PSF;
  float2 qk[qreg ? N1 : 1];
  if constexpr (qreg) {
    const int k2 = ROWS ? tid % N2 : tid / L;
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) qk[k1] = sm.q[k2 + N2 * k1];
  }
  auto stage_b = [&](const float2* xin, float2* xout) {
    if constexpr (N1 == 32) {
      // 32-point lines (the 512 side): a lane pair per line, lane h holding the half
      // y[m + 16 h], m < 16.  Forward, decimation in frequency: the pair exchange gives
      // y[m] + y[m + 16] on lane 0 and y[m + 16] - y[m] on lane 1, which multiplies by
      // -W32^m; 16-point transforms then give X[2k + h].  After the Fresnel factor, the
      // inverse (decimation in time): 16-point transforms, lane 1 multiplies by -conj W32^n,
      // and the pair exchange gives x[n] = E + O' on lane 0 and x[n + 16] = E - O' on lane 1.
      static_assert(2 * T::A2 == NT, "one lane pair per 32-point line");
      const int h = tid & 1, jb = tid >> 1;
      const int l = ROWS ? jb / N2 : jb % L;
      const int k2 = ROWS ? jb % N2 : jb / L;
      const float sf = h ? -1.0f : 1.0f;  // forward exchange: own + sf * partner
      // the pair twiddles, 1 on lane 0 and -W32^m (forward) or -conj W32^n (inverse) on lane 1,
      // as hf (-c, +-s) + (1 - hf, 0) with hf = h: exact on both lanes, no table reads
      const float hf = (float)h, omh = 1.0f - hf;
      auto pair_tw = [&](int m, bool inv) {
        const float c = fftk::kCos32[m], sn = fftk::kSin32[m];
        return make_float2(fmaf(hf, -c, omh), hf * (inv ? -sn : sn));
      };
      const float2* qh = sm.q + h * SM::kH + k2;  // q[k2 + 16 (2 k + h)] at qh[16 k]
      const float2* twh = sm.tw + SM::kTwR + h * SM::kH + 16 * k2;  // W^(k2 (n + 16 h)) at twh[n]
      // (loads, table reads, exchanges and products in separate loops: with each exchange
      // right after its element's load, the loads were issued one at a time)
      float2 v[16];
      const float2* src = xin + l * LS + k2 + h * T::HP;  // y[m + 16 h][k2], m < 16
#pragma unroll
      for (int m = 0; m < 16; ++m) v[m] = src[m * (N2 + 1)];
#pragma unroll
      for (int m = 0; m < 16; ++m)
        v[m] = make_float2(fmaf(sf, pair_swap(v[m].x), v[m].x), fmaf(sf, pair_swap(v[m].y), v[m].y));
#pragma unroll
      for (int m = 0; m < 16; ++m) v[m] = cmul(v[m], pair_tw(m, false));
      fftk::fft<16, false>(v);
#pragma unroll
      for (int k = 0; k < 16; ++k) v[k] = cmul(v[k], qh[16 * k]);
      fftk::fft<16, true>(v);
#pragma unroll
      for (int n = 0; n < 16; ++n) v[n] = cmul(v[n], pair_tw(n, true));
#pragma unroll
      for (int n = 0; n < 16; ++n)
        v[n] = make_float2(fmaf(-sf, pair_swap(v[n].x), v[n].x),
                           fmaf(-sf, pair_swap(v[n].y), v[n].y));
      float2* dst = xout + l * LS + k2 * T::KS2 + h * T::HP;  // element n + 16 h
#pragma unroll
      for (int n = 0; n < 16; ++n) dst[n] = cmulc(v[n], twh[n]);
    } else {
#pragma unroll
      for (int jj = 0; jj < J2; ++jj) {
        const int jb = tid + jj * NT;
        const int l = (ROWS || J2 > 1) ? jb / N2 : jb % L;
        const int k2 = (ROWS || J2 > 1) ? jb % N2 : jb / L;
        float2 v[N1];
        const float2* src = xin + l * LS + k2;
#pragma unroll
        for (int n = 0; n < N1; ++n) v[n] = src[n * (N2 + 1)];
        fftk::fft<N1, false>(v);
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) v[k1] = cmul(v[k1], qreg ? qk[k1] : sm.q[k2 + N2 * k1]);
        fftk::fft<N1, true>(v);
#pragma unroll
        for (int j = 1; j < N1; ++j) v[j] = cmulc(v[j], tw_at(k2, j));
        float2* dst = xout + l * LS + k2 * (N1 + 1);
#pragma unroll
        for (int j = 0; j < N1; ++j) dst[j] = v[j];
        job_fence<J2>();
      }
    }
  };

  // plane p's screen phases and mask pixels of the (single) stage-1 job of tile tt
  float phn[ahead && !cplx ? N2 : 1], mkn[ahead ? N2 : 1];
  float2 scn[ahead && cplx ? N2 : 1];
  auto load_pointwise = [&](int tt, float* ph, float2* sc, float* mk) {
    int l1, n1;
    job1(0, l1, n1);
    const int line = tt * L + l1;
    const int x0 = ROWS ? line : n1, y0 = ROWS ? n1 : line;
    const uint32_t g0 = (uint32_t)(x0 * N + y0);
    const uint32_t ph0 = (uint32_t)((c.ox + x0) * a.W + c.oy + y0) * SE;
#pragma unroll
    for (int m = 0; m < N2; ++m) {
      if (!LAST) {
        if (cplx)
          sc[m] = ld2(r_scr, ph0, PSB * m);
        else
          ph[m] = half ? ldh(r_scr, ph0, PSB * m) * sscale : ld1(r_scr, ph0, PSB * m);
      }
      if (!FIRST) mk[m] = ld1(r_mask, g0 * 4, GS4 * m);
    }
  };

#pragma unroll 1
  for (int t = 0; t < N / L; ++t) {
    // requested at the start of the tile (or during the previous tile) so that their latency
    // overlaps the transforms before the pointwise stage
    float phv[cplx ? 1 : N2], mkv[N2];
    float2 scv[cplx ? N2 : 1];
    if (ahead && t > 0) {
#pragma unroll
      for (int m = 0; m < N2; ++m) {
        if (cplx) scv[m] = scn[m]; else phv[m] = phn[m];
        mkv[m] = mkn[m];
      }
    } else if (hoist) {
      load_pointwise(t, phv, scv, mkv);
    }
    // pointwise work of plane p on stage-1 job (l1, n1): screen times edge window (planes before
    // the last), |u|^4 into the mask with the trapezoid weight, focal-plane output
    // FP16 screens: a job whose phases (held in registers) are all below pi/2 in magnitude, on
    // every lane of the wave, skips the sin/cos range reduction (bit-identical there, 8 VALU
    // operations per pixel fewer); a physical screen of a few tenths of a radian takes this
    // path almost always.  (The FP32 kernel has no registers left for the second copy of the
    // pointwise stage: it spilled 22 and ran 0.8% slower.)
    auto pointwise = [&](int l1, int n1, float2 (&v)[N2]) {
      const int line = t * L + l1;
      const int x0 = ROWS ? line : n1, y0 = ROWS ? n1 : line;
      const uint32_t g0 = (uint32_t)(x0 * N + y0);
      const uint32_t ph0 = (uint32_t)((c.ox + x0) * a.W + c.oy + y0) * SE;
      const float efix = sm.eg[ROWS ? x0 : y0];
      constexpr bool split = !LAST && !cplx && hoist && half;
      bool small = false;
      if constexpr (split) {
        bool in = true;
#pragma unroll
        for (int m = 0; m < N2; ++m) in = in && fabsf(phv[m]) <= trig::kNoReduction;
        small = __all(in);
      }
      auto body = [&](auto reduce) {
#pragma unroll
      for (int m = 0; m < N2; ++m) {
        const int x = x0 + DX * m, y = y0 + DY * m;
        if (!LAST) {
          float sn, cs;
          if (cplx) {
            const float2 e = hoist ? scv[m] : ld2(r_scr, ph0, PSB * m);
            cs = e.x;
            sn = e.y;
          } else {
            const float phi = hoist ? phv[m]
                              : half ? ldh(r_scr, ph0, PSB * m) * sscale
                                     : ld1(r_scr, ph0, PSB * m);
            if constexpr (decltype(reduce)::value)
              screen_sincos(phi, &sn, &cs);
            else
              trig::sincos_reduced(phi, &sn, &cs);
          }
          const float sg = efix * sm.eg[ROWS ? y : x];
          v[m] = cmul(v[m], make_float2(cs * sg, sn * sg));
        }
        const float pw = v[m].x * v[m].x + v[m].y * v[m].y;
        if (PSF && p >= a.psf_z0 && p < a.psf_z0 + a.psf_nz) {
          const int cx = x - a.psf_c0, cy = y - a.psf_c0;
          if (cx >= 0 && cx < a.psf_w && cy >= 0 && cy < a.psf_w)
            a.psf_out[(((int64_t)(p - a.psf_z0) * a.P + c.site) * a.psf_w + cx) * a.psf_w + cy] =
                pw * pw;
        }
        const float mm = wgt * (pw * pw);
        const float acc = FIRST ? mm : (hoist ? mkv[m] : ld1(r_mask, g0 * 4, GS4 * m)) + mm;
        st1(r_mask, g0 * 4, GS4 * m, acc);
        if (LAST) esum += (double)acc;
      }
      };
      if (split && small)
        body(std::false_type{});
      else
        body(std::true_type{});
    };
    auto field_byte = [&](int l1, int n1) {  // byte offset of a stage-1 job's element 0
      const int line = t * L + l1;
      return (uint32_t)(ROWS ? line * N + n1 : n1 * N + line) * 8u;
    };

    if constexpr (KIND == kOnly) {  // no slices: the entry plane is first and last
#pragma unroll
      for (int jj = 0; jj < J1; ++jj) {
        int l1, n1;
        job1(jj, l1, n1);
        float2 v[N2];
        const uint32_t f0 = field_byte(l1, n1);
#pragma unroll
        for (int m = 0; m < N2; ++m) v[m] = ld2(r_base, f0, GS8 * m);
        pointwise(l1, n1, v);
        job_fence<J1>();
      }
      continue;
    }
    // (A) the pass's first half step: stage 1 of the forward transform
#pragma unroll
    for (int jj = 0; jj < J1; ++jj) {
      int l1, n1;
      job1(jj, l1, n1);
      float2 v[N2];
      const Rsrc r_in = FIRST ? r_base : r_fld;
      const uint32_t f0 = field_byte(l1, n1);
      if (prefetch && t > 0) {
#pragma unroll
        for (int m = 0; m < N2; ++m) v[m] = vpre[m];
      } else {
#pragma unroll
        for (int m = 0; m < N2; ++m) v[m] = ld2(r_in, f0, GS8 * m);
      }
      if (FIRST) pointwise(l1, n1, v);
      fftk::fft<N2, false>(v);
      float2* dst = xa + l1 * LS + T::S1(n1);
#pragma unroll
      for (int j = 0; j < N2; ++j) dst[j] = j == 0 ? v[j] : cmul(v[j], tw_at(n1, j));
      if (prefetch && t + 1 < N / L) {
#pragma unroll
        for (int m = 0; m < N2; ++m) vpre[m] = ld2(r_fld, f0 + (uint32_t)TS * 8, GS8 * m);
      }
      job_fence<J1>();
    }
    lds_barrier();
    stage_b(xa, xb);
    lds_barrier();
    if constexpr (KIND == kMid) {
      // (C) stage 2 of the inverse (the field is now u_p), plane p's pointwise work, stage 1 of
      // this axis' next half step
#pragma unroll
      for (int jj = 0; jj < J1; ++jj) {
        int l1, n1;
        job1(jj, l1, n1);
        float2 v[N2];
        const float2* src = xb + l1 * LS + T::C2(n1);
#pragma unroll
        for (int n = 0; n < N2; ++n) v[n] = src[n * T::KS2];
        fftk::fft<N2, true>(v);
        pointwise(l1, n1, v);
        if (ahead && t + 1 < N / L) load_pointwise(t + 1, phn, scn, mkn);
        fftk::fft<N2, false>(v);
#pragma unroll
        for (int j = 1; j < N2; ++j) v[j] = cmul(v[j], tw_at(n1, j));
        float2* dst = xa + l1 * LS + T::S1(n1);
#pragma unroll
        for (int j = 0; j < N2; ++j) dst[j] = v[j];
        job_fence<J1>();
      }
      lds_barrier();
      stage_b(xa, xb);
      lds_barrier();
    }
    // final stage: stage 2 of the inverse, then the field to memory or the last plane's mask
#pragma unroll
    for (int jj = 0; jj < J1; ++jj) {
      int l1, n1;
      job1(jj, l1, n1);
      float2 v[N2];
      const float2* src = xb + l1 * LS + T::C2(n1);
#pragma unroll
      for (int n = 0; n < N2; ++n) v[n] = src[n * T::KS2];
      fftk::fft<N2, true>(v);
      if (LAST) {
        pointwise(l1, n1, v);
      } else {
        const uint32_t f0 = field_byte(l1, n1);
#pragma unroll
        for (int m = 0; m < N2; ++m) st2(r_fld, f0, GS8 * m, v[m]);
      }
      job_fence<J1>();
    }
  }
  __syncthreads();  // this pass's field stores are read by other threads in the next pass
}

// Pass of plane p.  In a focal-plane launch only the passes of the output planes carry the
// focal-window stores (and their registers); the other passes are the FP32-screen passes.
template <int N, int NT, int MODE, int AXIS, int KIND, class SM>
__device__ __forceinline__ void plane_pass(const SitesArgs& a, const SiteCtx& c, int p, int tid,
                                           SM& sm, double& esum) {
  if constexpr (MODE == kScreens32Psf) {
    if (p >= a.psf_z0 && p < a.psf_z0 + a.psf_nz)
      line_pass<N, NT, kScreens32Psf, AXIS, KIND>(a, c, p, tid, sm, esum);
    else
      line_pass<N, NT, kScreens32, AXIS, KIND>(a, c, p, tid, sm, esum);
  } else {
    line_pass<N, NT, MODE, AXIS, KIND>(a, c, p, tid, sm, esum);
  }
}

template <int N, int NT, int MODE, class SM>
__device__ __forceinline__ void run_site(const SitesArgs& a, int site, SM& sm) {
  constexpr int C = ClassIdx<N>::v;
  // The thread index is made opaque once per site: values derived from it are then computed
  // inside this site's code path instead of being hoisted above the persistent site loop,
  // where the paths of every field side of a mixed launch would all hold registers at once.
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  for (int i = tid; i < N; i += NT) {
    const float2 t = a.twid[a.ctw[C] + i];
    const float2 q = a.q1d[a.cq[C] + i];
    sm.tw[i] = t;
    sm.eg[i] = a.edge[a.cedge[C] + i];
    if constexpr (SM::kRowTw && N == 256) {
      const int r = i >> 4, c = i & 15;
      sm.twr[17 * r + c] = a.twid[a.ctw[C] + r * c];
    }
    if constexpr (N == 512) {  // the lane pairs' tables
      const int k2 = i & 15, k1 = i >> 4;
      sm.q[(k1 & 1) * SM::kH + k2 + 16 * (k1 >> 1)] = q;
      const int h = i >> 8, r = (i >> 4) & 15;  // row r = k2 of the pair twiddles, column k2
      sm.tw[SM::kTwR + h * SM::kH + i % 256] = a.twid[a.ctw[C] + r * (k2 + 16 * h)];
    } else {
      sm.q[i] = q;
    }
  }
  __syncthreads();

  const int S = a.site_s[site];
  const SiteCtx c{site, a.site_x[site], a.site_y[site], a.bases + a.base_off[site],
                  a.mask + a.mask_off[site], a.scratch + (int64_t)blockIdx.x * a.scratch_stride};
  double esum = 0.0;
  if (S < 1) {
    plane_pass<N, NT, MODE, kAxisX, kOnly>(a, c, 0, tid, sm, esum);
  } else {
    plane_pass<N, NT, MODE, kAxisX, kHead>(a, c, 0, tid, sm, esum);
    for (int p = 1; p < S; ++p) {
      if (p & 1)
        plane_pass<N, NT, MODE, kAxisY, kMid>(a, c, p, tid, sm, esum);
      else
        plane_pass<N, NT, MODE, kAxisX, kMid>(a, c, p, tid, sm, esum);
    }
    if (S & 1)
      plane_pass<N, NT, MODE, kAxisY, kTail>(a, c, S, tid, sm, esum);
    else
      plane_pass<N, NT, MODE, kAxisX, kTail>(a, c, S, tid, sm, esum);
  }
  const double tot = block_sum_double(esum, sm.red);
  if (tid == 0) a.energy[site] = tot;
  if constexpr (MODE == kScreens32Psf) {
    if (a.audit_pred) {
      // this block's focal-plane stores are visible to it after the barriers above
      const int slot = a.audit_slot[site];
      if (slot >= 0) audit_site(a, site, slot, tid, NT, sm);
      if (tid == 0) a.site_done[site] = 1;
    }
  }
  __syncthreads();
}

// One field side for every site of the launch.
// DYN: sites from the work counter and/or an audited launch (take_site); otherwise the plain
// static-striding loop.  The schedule is a compile-time choice because the loop's form changes
// the code generated for the inlined site body: with take_site's barriers in the loop, the
// uniform 4225-site launch runs 112 ms against 103 ms for the plain loop, static either way.
// 256-thread blocks are held to two waves per SIMD (256 VGPRs), so that two blocks share a CU
// as one 512-thread block would, each with barriers of its own.
template <int N, int NT, int MODE, bool DYN>
__global__ __launch_bounds__(NT, 512 / NT) void sites_kernel(SitesArgs a) {
  __shared__ SmemU<N, NT> sm;
  if constexpr (DYN) {
    __shared__ int next;
    for (int64_t i = take_site(a, -1, &next); i < a.P; i = take_site(a, i, &next))
      run_site<N, NT, MODE>(a, a.order[i], sm);
  } else {
    for (int64_t i = blockIdx.x; i < a.P; i += gridDim.x) run_site<N, NT, MODE>(a, a.order[i], sm);
  }
}

// One field side's site body of the mixed launch, a function of its own: its registers are
// allocated for that side alone (one function holding all four sides spilled 480 registers'
// worth in the 512 side's passes), and none of them is held across the site loop.
template <int N, int NT, class SM>
__device__ __noinline__ void mixed_site(const SitesArgs& a, int site, SM& sm) {
  run_site<N, NT, kScreens32Mixed>(a, site, sm);
}

// Sites of mixed field sides in one launch (the order list puts large sites first), from both
// loops (static striding, work counter).  Kept out of line, the site bodies run the mixed
// 4225-site launch in 80 ms against 93-97 ms inlined into the loops.
template <int NT>
__global__ __launch_bounds__(NT) void sites_mixed_kernel(SitesArgs a) {
  __shared__ Smem<NT, kNmax> sm;
  __shared__ int next;
  auto run = [&](int site) {
    switch (a.site_n[site]) {
      case 64: mixed_site<64, NT>(a, site, sm); break;
      case 128: mixed_site<128, NT>(a, site, sm); break;
      case 256: mixed_site<256, NT>(a, site, sm); break;
      case 512: mixed_site<512, NT>(a, site, sm); break;
      default: break;  // rejected on the host
    }
  };
  if (!a.queue) {
    for (int64_t i = blockIdx.x; i < a.P; i += gridDim.x) run(a.order[i]);
    return;
  }
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(a.queue, 1);
    __syncthreads();
    const int64_t i = next;
    __syncthreads();  // every thread has read `next` before thread 0 takes the following one
    if (i >= a.P) return;
    run(a.order[i]);
  }
}

// ---------------------------------------------------------------- phase prep
// Screens from the index volume, FP64 accumulation over the vps voxels of each slab.
// A block owns `cols` consecutive (x, y) columns; their voxels are one contiguous range of
// cols * Zv floats, staged through LDS with coalesced loads (stride Zv + 1 against bank
// conflicts).  Each thread then sums its column slab by slab, and the block reduces every
// slice's max |phi| before a single atomic per slice.
constexpr int kPrepThreads = 256;
constexpr int kPrepLds = 16384;  // floats of staged voxels per block (64 KiB)

__global__ __launch_bounds__(kPrepThreads) void phase_prepare_kernel(
    const float* __restrict__ index, int64_t tot, int Zv, int vps, int cols, double n0,
    double k0dz, float* __restrict__ phase32, float* __restrict__ slice_max) {
  extern __shared__ float stage[];
  __shared__ float wmax[kPrepThreads / 64];
  const int Z = Zv / vps;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int64_t p0 = (int64_t)blockIdx.x * cols; p0 < tot; p0 += (int64_t)gridDim.x * cols) {
    const int nc = (int)(tot - p0 < cols ? tot - p0 : cols);
    const float* src = index + p0 * Zv;
    for (int e = tid; e < nc * Zv; e += kPrepThreads) stage[(e / Zv) * (Zv + 1) + e % Zv] = src[e];
    __syncthreads();
    for (int z = 0; z < Z; ++z) {
      float mag = 0.f;
      if (tid < nc) {
        double acc = 0.0;
        const float* col = stage + tid * (Zv + 1) + z * vps;
        for (int v = 0; v < vps; ++v) acc += (double)col[v] - n0;
        const float ph = (float)(k0dz * acc);
        phase32[(int64_t)z * tot + p0 + tid] = ph;
        mag = fabsf(ph);
      }
      if (slice_max) {  // uniform branch
        for (int off = 32; off > 0; off >>= 1) mag = fmaxf(mag, __shfl_down(mag, off, 64));
        if (lane == 0) wmax[wave] = mag;
        __syncthreads();
        if (tid == 0) {
          float m = wmax[0];
          for (int w = 1; w < kPrepThreads / 64; ++w) m = fmaxf(m, wmax[w]);
          atomicMax((int*)&slice_max[z], __float_as_int(m));
        }
        __syncthreads();
      }
    }
    __syncthreads();  // the stage buffer is refilled for the next column block
  }
}

__global__ void phase_to_half_kernel(const float* __restrict__ phase32, int64_t plane, int64_t Z,
                                     const float* __restrict__ slice_max,
                                     __half* __restrict__ out, float* __restrict__ scale) {
  const int64_t tot = plane * Z;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t z = i / plane;
    const float m = slice_max[z];
    const float inv = m > 0.f ? 1.0f / m : 0.f;
    out[i] = __float2half_rn(phase32[i] * inv);
    if (i % plane == 0) scale[z] = m;
  }
}

// screen[i] = (cos phase[i], sin phase[i]) with the sites kernels' own sin/cos
__global__ void screen_exp_kernel(const float* __restrict__ phase, int64_t n,
                                  float2* __restrict__ screen) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    float sn, cs;
    screen_sincos(phase[i], &sn, &cs);
    screen[i] = make_float2(cs, sn);
  }
}

__global__ void fill_float_kernel(float* p, int64_t n, float v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace

MAOMI_API int64_t maomi_sites_workspace_bytes(int64_t grid, int64_t nmax) {
  return grid * nmax * nmax * (int64_t)sizeof(float2);
}

MAOMI_API int64_t maomi_sites_default_grid(void) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  int per_cu = 0;
  const auto kernel = sites_kernel<256, 512, kScreens32, false>;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 512, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return (int64_t)cus * per_cu;
}

MAOMI_API int maomi_persistent_sites(const maomi_sites_desc* d, maomi_stream_t stream) {
  if (!d) { maomi_set_error("maomi_persistent_sites: null descriptor"); return MAOMI_EINVAL; }
  if (d->P <= 0) return MAOMI_OK;
  if (!d->bases || !d->base_off || !d->edge || !d->q1d || !d->twiddle || !d->site_x ||
      !d->site_y || !d->site_n || !d->site_s || !d->order || !d->mask || !d->mask_off ||
      !d->energy || !d->workspace) {
    maomi_set_error("maomi_persistent_sites: null pointer argument");
    return MAOMI_EINVAL;
  }
  if ((d->phase != nullptr) + (d->phase16 != nullptr) + (d->screen != nullptr) != 1) {
    maomi_set_error("maomi_persistent_sites: give exactly one of phase / phase16 / screen");
    return MAOMI_EINVAL;
  }
  if (d->phase16 && !d->slice_scale) {
    maomi_set_error("maomi_persistent_sites: phase16 needs slice_scale");
    return MAOMI_EINVAL;
  }
  if (d->scratch_stride < 64 * 64 || d->H <= 0 || d->W <= 0 || d->Z <= 0 ||
      d->P > (1LL << 30)) {  // site indices and the int32 work counter stay far from overflow
    maomi_set_error("maomi_persistent_sites: bad sizes");
    return MAOMI_EINVAL;
  }
  // the passes reach a screen slice and a site's arrays through 32-bit byte ranges
  if (d->H * d->W * (int64_t)sizeof(float2) >= (1LL << 32) ||
      d->scratch_stride * (int64_t)sizeof(float2) >= (1LL << 32)) {
    maomi_set_error("maomi_persistent_sites: a screen slice (H x W) or the per-block workspace "
                    "exceeds 4 GiB");
    return MAOMI_EINVAL;
  }
  if (d->uniform_n > 0 && d->scratch_stride < d->uniform_n * d->uniform_n) {
    maomi_set_error("maomi_persistent_sites: scratch_stride %lld is below uniform_n^2 = %lld",
                    (long long)d->scratch_stride, (long long)(d->uniform_n * d->uniform_n));
    return MAOMI_EINVAL;
  }
  SitesArgs a;
  a.bases = (const float2*)d->bases; a.base_off = d->base_off;
  a.phase = d->phase; a.phase16 = (const __half*)d->phase16; a.slice_scale = d->slice_scale;
  a.screen = (const float2*)d->screen;
  a.edge = d->edge; a.q1d = (const float2*)d->q1d; a.twid = (const float2*)d->twiddle;
  a.site_x = d->site_x; a.site_y = d->site_y; a.site_n = d->site_n; a.site_s = d->site_s;
  a.order = d->order; a.mask = d->mask; a.mask_off = d->mask_off; a.energy = d->energy;
  a.scratch = (float2*)d->workspace; a.scratch_stride = d->scratch_stride;
  a.P = d->P; a.H = d->H; a.W = d->W; a.Z = d->Z;
  for (int i = 0; i < 5; ++i) { a.cedge[i] = d->class_edge[i]; a.cq[i] = d->class_q[i]; a.ctw[i] = d->class_tw[i]; }
  a.psf_out = d->psf_out;
  a.psf_z0 = (int)d->psf_z0; a.psf_nz = (int)d->psf_nz; a.psf_c0 = (int)d->psf_c0;
  a.psf_w = (int)d->psf_w;
  a.queue = d->queue;
  a.audit_pred = nullptr;
  a.audit_slot = nullptr;
  a.audit_state = nullptr;
  a.audit_rel = nullptr;
  a.site_done = nullptr;
  a.n_audit = 0;
  a.audit_max_fail = 0;
  a.audit_tol = 0.f;
  if (const maomi_sites_audit* au = d->audit) {
    if (!d->psf_out || !au->pred || !au->slot || !au->state || !au->rel || !au->done ||
        au->n_audit < 1 || au->n_audit > d->P || au->max_fail < 0 || !(au->tol >= 0.0)) {
      maomi_set_error("maomi_persistent_sites: an audit needs focal-plane output, every audit "
                      "array and 1 <= n_audit <= P");
      return MAOMI_EINVAL;
    }
    a.audit_pred = au->pred;
    a.audit_slot = au->slot;
    a.audit_state = au->state;
    a.audit_rel = au->rel;
    a.site_done = au->done;
    a.n_audit = (int)au->n_audit;
    a.audit_max_fail = (int)au->max_fail;
    a.audit_tol = (float)au->tol;
  }
  if (a.psf_out && (a.psf_nz < 1 || a.psf_w < 1 || a.psf_z0 < 0 || a.psf_c0 < 0 ||
                    (d->uniform_n > 0 && d->psf_c0 + d->psf_w > d->uniform_n) ||
                    d->psf_z0 + d->psf_nz > d->Z + 1)) {
    maomi_set_error("maomi_persistent_sites: bad focal-plane output window");
    return MAOMI_EINVAL;
  }
  int64_t grid = d->grid > 0 ? d->grid : maomi_sites_default_grid();
  if (grid <= 0) grid = 256;
  if (grid > d->P) grid = d->P;
  const int nt = d->block_threads ? (int)d->block_threads : 512;
  if (nt != 512 && nt != 256) {
    maomi_set_error("maomi_persistent_sites: block_threads must be 0, 256 or 512");
    return MAOMI_EUNSUPPORTED;
  }
  const dim3 g((unsigned)grid), b(nt);
  hipStream_t st = (hipStream_t)stream;
  if (a.psf_out && (d->uniform_n == 0 || nt != 512 || !a.phase)) {
    maomi_set_error("maomi_persistent_sites: focal-plane output needs a uniform field side, "
                    "FP32 phase screens and block_threads 512");
    return MAOMI_EUNSUPPORTED;
  }
  if (a.screen && (d->uniform_n == 0 || nt != 512)) {
    maomi_set_error("maomi_persistent_sites: complex screens need a uniform field side and "
                    "block_threads 512");
    return MAOMI_EUNSUPPORTED;
  }
  if (a.phase16 && d->uniform_n == 0) {
    maomi_set_error("maomi_persistent_sites: FP16 screens need a uniform field side");
    return MAOMI_EUNSUPPORTED;
  }
  const int mode = a.psf_out ? kScreens32Psf
                   : a.phase16 ? kScreens16
                   : a.screen  ? kScreensC64
                               : kScreens32;
  const bool dyn = a.queue || a.audit_pred;
#define MAOMI_LAUNCH_D(NTV, MODEV, DV)                                                         \
  switch (d->uniform_n) {                                                                      \
    case 64: hipLaunchKernelGGL((sites_kernel<64, NTV, MODEV, DV>), g, b, 0, st, a); break;    \
    case 128: hipLaunchKernelGGL((sites_kernel<128, NTV, MODEV, DV>), g, b, 0, st, a); break;  \
    case 256: hipLaunchKernelGGL((sites_kernel<256, NTV, MODEV, DV>), g, b, 0, st, a); break;  \
    case 512: hipLaunchKernelGGL((sites_kernel<512, NTV, MODEV, DV>), g, b, 0, st, a); break;  \
    case 0: hipLaunchKernelGGL((sites_mixed_kernel<NTV>), g, b, 0, st, a); break;              \
    default:                                                                                   \
      maomi_set_error("maomi_persistent_sites: uniform_n must be 0, 64, 128, 256 or 512");     \
      return MAOMI_EUNSUPPORTED;                                                               \
  }
#define MAOMI_LAUNCH(NTV, MODEV)  \
  if (dyn) {                      \
    MAOMI_LAUNCH_D(NTV, MODEV, true) \
  } else {                        \
    MAOMI_LAUNCH_D(NTV, MODEV, false) \
  }
  if (d->uniform_n != 0 && d->uniform_n != 64 && d->uniform_n != 128 && d->uniform_n != 256 &&
      d->uniform_n != 512) {
    maomi_set_error("maomi_persistent_sites: uniform_n must be 0, 64, 128, 256 or 512");
    return MAOMI_EUNSUPPORTED;
  }
  if (a.queue) {  // every argument is checked: nothing is enqueued before this point
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, st, a.queue);
    MAOMI_CHECK_LAUNCH();
  }
  if (a.audit_pred) {
    const unsigned rb = (unsigned)((d->P + 255) / 256 < 1024 ? (d->P + 255) / 256 : 1024);
    hipLaunchKernelGGL(audit_reset_kernel, dim3(rb), dim3(256), 0, st, a.audit_state,
                       a.site_done, (int64_t)d->P);
    MAOMI_CHECK_LAUNCH();
  }
  if (mode == kScreens32Psf) {
    MAOMI_LAUNCH(512, kScreens32Psf)
  } else if (mode == kScreensC64) {
    MAOMI_LAUNCH(512, kScreensC64)
  } else if (mode == kScreens16) {
    if (nt == 512) {
      MAOMI_LAUNCH(512, kScreens16)
    } else {
      MAOMI_LAUNCH(256, kScreens16)
    }
  } else if (nt == 512) {
    MAOMI_LAUNCH(512, kScreens32)
  } else {
    MAOMI_LAUNCH(256, kScreens32)
  }
#undef MAOMI_LAUNCH
#undef MAOMI_LAUNCH_D
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_phase_prepare(const float* index, int64_t X, int64_t Y, int64_t Zv,
                                  int64_t vps, double n0, double k0dz, float* phase32,
                                  void* phase16, float* slice_scale, float* slice_max,
                                  maomi_stream_t stream) {
  if (!index || !phase32 || X <= 0 || Y <= 0 || Zv <= 0 || vps <= 0 || Zv % vps) {
    maomi_set_error("maomi_phase_prepare: bad arguments (phase32 is required as staging)");
    return MAOMI_EINVAL;
  }
  if (phase16 && (!slice_scale || !slice_max)) {
    maomi_set_error("maomi_phase_prepare: phase16 needs slice_scale and slice_max");
    return MAOMI_EINVAL;
  }
  const int64_t Z = Zv / vps;
  hipStream_t st = (hipStream_t)stream;
  if (slice_max) {
    hipLaunchKernelGGL(fill_float_kernel, dim3(1), dim3(64), 0, st, slice_max, Z, 0.0f);
    MAOMI_CHECK_LAUNCH();
  }
  const int64_t tot = X * Y;
  if (Zv > kPrepLds - 1) {
    maomi_set_error("maomi_phase_prepare: more than 16383 voxels per column");
    return MAOMI_EUNSUPPORTED;
  }
  // columns per block: as many as the 64 KiB stage holds, at most one per thread
  const int cols = (int)(kPrepLds / (Zv + 1) < kPrepThreads ? kPrepLds / (Zv + 1) : kPrepThreads);
  const int64_t nblk = (tot + cols - 1) / cols;
  const unsigned blocks = (unsigned)(nblk < 65535 ? nblk : 65535);
  hipLaunchKernelGGL(phase_prepare_kernel, dim3(blocks), dim3(kPrepThreads),
                     (size_t)cols * (Zv + 1) * sizeof(float), st, index, tot, (int)Zv, (int)vps,
                     cols, n0, k0dz, phase32, slice_max);
  MAOMI_CHECK_LAUNCH();
  if (phase16) {
    const int64_t n = tot * Z;
    const unsigned b2 = (unsigned)((n + 255) / 256 < 65535 ? (n + 255) / 256 : 65535);
    hipLaunchKernelGGL(phase_to_half_kernel, dim3(b2), dim3(256), 0, st, phase32, tot, Z,
                       slice_max, (__half*)phase16, slice_scale);
    MAOMI_CHECK_LAUNCH();
  }
  return MAOMI_OK;
}

MAOMI_API int maomi_screen_exp(const float* phase, int64_t n, void* screen,
                               maomi_stream_t stream) {
  if (!phase || !screen || n < 1) {
    maomi_set_error("maomi_screen_exp: bad arguments");
    return MAOMI_EINVAL;
  }
  const unsigned b = (unsigned)((n + 255) / 256 < 65535 ? (n + 255) / 256 : 65535);
  hipLaunchKernelGGL(screen_exp_kernel, dim3(b), dim3(256), 0, (hipStream_t)stream, phase, n,
                     (float2*)screen);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}
