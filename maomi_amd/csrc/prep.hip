// Once-per-run beam preparation: pupil field, the scaled Fresnel step that
// carries it to the top of the tissue, and its unit-norm scaling.  Built with
// -ffp-contract=off: the chirp arguments reach ~1e4-1e5 rad, where float32
// rounding of the argument is the dominant error, so every product and sum
// is rounded in the same order as the reference's float32 expression.
//
// Reference: maomi/_optical_math.py:121-134 (gaussian profile), 137-197
// (Zernike), 232-296 (generate_ba); maomi/numerics.py:156-210 (fresnel, one
// scaled step); maomi/backends/cuda/batched.py:43-51 (base beam, unit norm).
#include "common.hpp"
#include "maomi_amd.h"
#include <cmath>

namespace {

// ---------------------------------------------------------------- generic line FFT
// Lines of arbitrary length N = N1 * N2 (each factor up to N) transformed by a
// two-level direct DFT from LDS.  Used once per run for the base beam and by
// the segmented stage for sizes that are not powers of two.
__global__ void twiddle_table_kernel(float2* tw, int N) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < N; j += gridDim.x * blockDim.x) {
    const double a = 2.0 * M_PI * (double)j / (double)N;
    tw[j] = make_float2((float)cos(a), (float)(-sin(a)));
  }
}

// data: B fields of N x N; axis 1 = rows (contiguous), 0 = columns.
__global__ void fft_lines_kernel(float2* data, int N, int N1, int N2, int axis, int inverse,
                                 int L, const float2* __restrict__ tw) {
  extern __shared__ float2 sh[];
  float2* A = sh;
  float2* Bf = sh + (size_t)L * N;
  float2* T = Bf + (size_t)L * N;
  const int line0 = blockIdx.x * L;
  const int nl = min(L, N - line0);
  float2* f = data + (size_t)blockIdx.y * N * N;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const float2 w = tw[i];
    T[i] = inverse ? make_float2(w.x, -w.y) : w;
  }
  for (int it = threadIdx.x; it < nl * N; it += blockDim.x) {
    int l, i;
    if (axis == 1) { l = it / N; i = it % N; }
    else { i = it / nl; l = it % nl; }
    const size_t g = axis == 1 ? (size_t)(line0 + l) * N + i : (size_t)i * N + line0 + l;
    A[l * N + i] = f[g];
  }
  __syncthreads();
  const int s1 = N / N2, s2 = N / N1;  // W_N2 = W_N^{N1}, W_N1 = W_N^{N2}
  for (int it = threadIdx.x; it < nl * N; it += blockDim.x) {
    const int l = it / N, r = it % N, n1 = r / N2, k2 = r % N2;
    const float2* a = A + l * N + n1;
    float2 acc = make_float2(0.f, 0.f);
    for (int n2 = 0; n2 < N2; ++n2) acc = cadd(acc, cmul(a[N1 * n2], T[((n2 * k2) % N2) * s1]));
    Bf[l * N + n1 * N2 + k2] = cmul(acc, T[(n1 * k2) % N]);
  }
  __syncthreads();
  for (int it = threadIdx.x; it < nl * N; it += blockDim.x) {
    int l, k;
    if (axis == 1) { l = it / N; k = it % N; }
    else { k = it / nl; l = it % nl; }
    const int k2 = k % N2, k1 = k / N2;
    const float2* b = Bf + l * N + k2;
    float2 acc = make_float2(0.f, 0.f);
    for (int n1 = 0; n1 < N1; ++n1) acc = cadd(acc, cmul(b[n1 * N2], T[((n1 * k1) % N1) * s2]));
    const size_t g = axis == 1 ? (size_t)(line0 + l) * N + k : (size_t)k * N + line0 + l;
    f[g] = acc;
  }
}

// ---------------------------------------------------------------- pupil
// generate_ba, gaussian profile: U = exp(-(X^2+Y^2)/rad^2) [rho2 < objrad^2]
// exp(-i k rho2 / (2 fl)) with Y, X = meshgrid(y, x, 'ij'), x_j = (j - N/2) D1.
__global__ void pupil_kernel(float2* out, float* zmap, int N, float D1, float rad2, float obj2,
                             float ck, float objrad, const float* __restrict__ zw, int nz) {
  const int tot = N * N;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < tot; p += gridDim.x * blockDim.x) {
    const int i = p / N, j = p % N;
    const float X = (float)(j - N / 2) * D1;
    const float Y = (float)(i - N / 2) * D1;
    const float rho2 = X * X + Y * Y;
    float U = expf(-(X * X + Y * Y) / rad2);
    U = rho2 < obj2 ? U : 0.0f;
    const float th = ck * rho2;
    float s, c;
    sincosf(th, &s, &c);
    out[p] = make_float2(U * c, U * s);
    if (nz > 0) {
      const float x = X / objrad, y = Y / objrad;
      const float theta = atan2f(y, x);
      const float r = sqrtf(x * x + y * y);
      float ph = 0.0f;
      for (int q = 0; q < nz; ++q) {
        const float w = zw[q];
        if (w == 0.0f) continue;
        const int jn = q + 1;  // Noll index
        const int n = (int)ceil(sqrt(0.25 + 2.0 * jn) - 1.5);
        const int rem = jn - (n + 1) * n / 2;
        const int m = ((n + 1) % 2) * 2 * (rem / 2) + (n % 2) * (2 * ((rem + 1) / 2) - 1);
        float R = 0.0f;
        for (int sidx = 0; sidx <= (n - m) / 2; ++sidx) {
          double num = ((sidx % 2) ? -1.0 : 1.0) * tgamma((double)(n - sidx + 1));
          double den = tgamma((double)sidx + 1) * tgamma((double)((n + m) / 2 - sidx + 1)) *
                       tgamma((double)((n - m) / 2 - sidx + 1));
          const int e = n - 2 * sidx;
          float pw = 1.0f;
          for (int t = 0; t < e; ++t) pw *= r;
          R += (float)(num / den) * pw;
        }
        float Z;
        if (m == 0) Z = (float)sqrt((double)n + 1.0) * R;
        else if (jn % 2 == 0) Z = (float)sqrt(2.0 * (n + 1)) * R * cosf((float)m * theta);
        else Z = (float)sqrt(2.0 * (n + 1)) * R * sinf((float)m * theta);
        ph += w * Z;
      }
      zmap[p] = ph;
    }
  }
}

__global__ void sum_kernel(const float* __restrict__ v, int64_t n, double* out) {
  __shared__ double red[16];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    acc += (double)v[i];
  const double t = block_sum_double(acc, red);
  if (threadIdx.x == 0) atomicAdd(out, t);
}

__global__ void zernike_apply_kernel(float2* out, const float* __restrict__ zmap, int64_t n,
                                     const double* sum, float kf) {
  const float mean = (float)(*sum / (double)n);
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    const float arg = kf * (zmap[p] - mean);
    float s, c;
    sincosf(arg, &s, &c);
    out[p] = cmul(out[p], make_float2(c, s));
  }
}

// ---------------------------------------------------------------- scaled step
// Entry: u <- u * exp(i theta_in) * s0 with
//   theta_in = ((k T) (1 - sc)) * (1 / (2 dz)),  T = (x dx0)^2 + (y dx0)^2,
// the order Python evaluates 1j*k*T*(1-sc)/(2dz) on a complex64 tensor.
// Per-beam parameters of a batch of Fresnel steps with different distances: bp[3 b + 0..2] =
// (1 / (2 dz), -pi lambda' dz, k (sc - 1) / (2 sc dz)), each rounded to float as the scalar
// arguments are; null bp uses the scalar arguments for every beam.
__global__ void entry_kernel(float2* u, int B, int N, float dx0, float kf, float onemsc,
                             float inv2dz, const float* __restrict__ s0, float pre,
                             const float* __restrict__ bp) {
  const int64_t tot = (int64_t)N * N;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < tot * B;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = p % tot;
    const int i = (int)(q / N), j = (int)(q % N);
    const float x = (float)(i - N / 2), y = (float)(j - N / 2);
    float2 v = u[p];
    if (onemsc != 0.0f) {
      const float xa = x * dx0, ya = y * dx0;
      const float T = xa * xa + ya * ya;
      const float i2 = bp ? bp[3 * (p / tot)] : inv2dz;
      const float th = ((kf * T) * onemsc) * i2;
      float s, c;
      sincosf(th, &s, &c);
      v = cmul(v, make_float2(c, s));
    }
    if (s0) v = cscale(v, s0[q]);
    if (pre != 1.0f) v = cscale(v, pre);  // u / sc as a multiply by the rounded reciprocal
    u[p] = v;
  }
}

// Multiply by the Fresnel factor in FFT order (ifftshift of the centred grid)
// and 1/N^2: arg = (c * v) * (1/sc), v = (x/(N dx))^2 + (y/(N dx))^2.
__global__ void fresnel_q_kernel(float2* u, int B, int N, float cq, float hd, float invsc,
                                 int scaled, float norm, const float* __restrict__ bp) {
  const int64_t tot = (int64_t)N * N;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < tot * B;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = p % tot;
    const int ki = (int)(q / N), kj = (int)(q % N);
    const int ci = (ki + N / 2) % N, cj = (kj + N / 2) % N;  // centred index
    const float x = (float)(ci - N / 2), y = (float)(cj - N / 2);
    const float xa = x / hd, ya = y / hd;
    float arg = (bp ? bp[3 * (p / tot) + 1] : cq) * (xa * xa + ya * ya);
    if (scaled) arg = arg * invsc;
    float s, c;
    sincosf(arg, &s, &c);
    const float2 v = cmul(u[p], make_float2(c, s));
    u[p] = cscale(v, norm);
  }
}

// Exit: u <- u * s1 * exp(i theta_out), theta_out = cexit * ((x dx1)^2 + (y dx1)^2).
__global__ void exit_kernel(float2* u, int B, int N, float dx1, float cexit,
                            const float* __restrict__ s1, int chirp,
                            const float* __restrict__ bp) {
  const int64_t tot = (int64_t)N * N;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < tot * B;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = p % tot;
    const int i = (int)(q / N), j = (int)(q % N);
    float2 v = u[p];
    if (s1) v = cscale(v, s1[q]);
    if (chirp) {
      const float xa = (float)(i - N / 2) * dx1, ya = (float)(j - N / 2) * dx1;
      const float th = (bp ? bp[3 * (p / tot) + 2] : cexit) * (xa * xa + ya * ya);
      float s, c;
      sincosf(th, &s, &c);
      v = cmul(v, make_float2(c, s));
    }
    u[p] = v;
  }
}

__global__ void sumsq_kernel(const float2* __restrict__ u, int64_t n, double* out) {
  __shared__ double red[16];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double a = u[i].x, b = u[i].y;
    acc += a * a + b * b;
  }
  const double t = block_sum_double(acc, red);
  if (threadIdx.x == 0) atomicAdd(out, t);
}

// per-beam sums of squares (beam blockIdx.y) and the scaling by their inverse root
__global__ void sumsq_beams_kernel(const float2* __restrict__ u, int64_t n, double* out) {
  __shared__ double red[16];
  const float2* ub = u + (int64_t)blockIdx.y * n;
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double a = ub[i].x, b = ub[i].y;
    acc += a * a + b * b;
  }
  const double t = block_sum_double(acc, red);
  if (threadIdx.x == 0) atomicAdd(out + blockIdx.y, t);
}

__global__ void scale_beams_kernel(float2* u, int64_t n, const double* sumsq) {
  const float inv = 1.0f / (float)sqrt(sumsq[blockIdx.y]);
  float2* ub = u + (int64_t)blockIdx.y * n;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    ub[i] = cscale(ub[i], inv);
}

__global__ void scale_by_norm_kernel(float2* u, int64_t n, const double* sumsq) {
  const float inv = 1.0f / (float)sqrt(*sumsq);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    u[i] = cscale(u[i], inv);
}

__global__ void zero_double_kernel(double* p, int n) {
  if (threadIdx.x < n) p[threadIdx.x] = 0.0;
}

inline unsigned nblocks(int64_t n, int bs = 256) {
  const int64_t b = (n + bs - 1) / bs;
  return (unsigned)(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

}  // namespace

// Factor N = N1 * N2 with N1 <= N2 as balanced as possible.
static void split_size(int N, int* N1, int* N2) {
  int best = 1;
  for (int d = 1; d * d <= N; ++d)
    if (N % d == 0) best = d;
  *N1 = best;
  *N2 = N / best;
}

// In-place 2-D FFT of B fields of N x N (complex64); tw holds W_N^j.
int maomi_fft2_generic(float2* u, int B, int N, int inverse, const float2* tw, hipStream_t st) {
  int N1, N2;
  split_size(N, &N1, &N2);
  int L = 2048 / N;
  if (L < 1) L = 1;
  if (L > N) L = N;
  const size_t shm = ((size_t)2 * L * N + N) * sizeof(float2);
  const dim3 grid((unsigned)((N + L - 1) / L), (unsigned)B);
  for (int axis = 1; axis >= 0; --axis) {
    hipLaunchKernelGGL(fft_lines_kernel, grid, dim3(256), shm, st, u, N, N1, N2, axis, inverse, L,
                       tw);
    MAOMI_CHECK_LAUNCH();
  }
  return MAOMI_OK;
}

MAOMI_API int maomi_twiddles(void* tw, int64_t N, maomi_stream_t stream) {
  if (!tw || N < 2) { maomi_set_error("maomi_twiddles: bad arguments"); return MAOMI_EINVAL; }
  hipLaunchKernelGGL(twiddle_table_kernel, dim3(nblocks(N)), dim3(256), 0, (hipStream_t)stream,
                     (float2*)tw, (int)N);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

// ---------------------------------------------------------------- pupil families
// Ring amplitude of the Bessel pupils at radial offset dr (pupils.py:112-122).  p.width is
// already the profile's own width; this file is built without contraction, so each float
// operation rounds as the reference's float32 array arithmetic does.
__device__ inline float ring_amplitude(float dr, const maomi_pupil_params& p) {
  switch (p.profile) {
    case 0: {
      const float t = dr / p.width;
      return expf(-(t * t));
    }
    case 1:
      return (dr < p.width && dr > -p.width) ? 1.0f : 0.0f;
    default: {
      const float z = (p.profile == 2 ? dr : -dr) + p.offset;
      const float t = z / p.width;
      return z > 0.0f ? z * expf(-(t * t)) : 0.0f;
    }
  }
}

__global__ void pupil_field_kernel(float2* __restrict__ out, int n0, int64_t tot,
                                   const float* __restrict__ xs, const float* __restrict__ ys,
                                   const float* __restrict__ cols, maomi_pupil_params p) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n0), j = (int)(e - (int64_t)i * n0);
    const float x = xs[j], y = ys[i];
    const float y2 = y * y;
    const float rho2 = x * x + y2;
    const bool inside = rho2 < p.obj2;
    const bool pair = p.kind >= 2;
    const float u = pair ? x - p.shift : x;  // the pair beams' amplitude is shifted along x
    const float r2 = pair ? u * u + y2 : rho2;
    float amp;
    if (p.kind == 0 || p.kind == 2) {
      amp = inside ? expf(-r2 / p.rr) : 0.0f;
      if (p.cylinder) {  // generateCylindricalBA's support, in FP64 as the reference has it
        const double sx = (double)x / p.scale_x, sy = (double)y / p.scale_y;
        const double sc = sx * sx + sy * sy;
        if (!(sc < (double)p.obj2 && exp(-sc / (double)p.rr) > 0.0)) amp = 0.0f;
      }
    } else {
      amp = ring_amplitude(sqrtf(r2) - p.radius, p);
      if (p.kind == 3 && p.clip && !inside) amp = 0.0f;
    }
    float s, c;
    sincosf(p.focus * rho2, &s, &c);
    float re = amp * c, im = amp * s;
    if (pair) {  // times exp(i k sign phase(x))
      float ts, tc;
      sincosf(p.tilt * cols[j], &ts, &tc);
      const float r = re * tc - im * ts;
      im = re * ts + im * tc;
      re = r;
    }
    out[e] = make_float2(re, im);
  }
}

MAOMI_API int maomi_pupil_gaussian(void* out, int64_t N, double D1, double rad, double objrad,
                                   double k, double fl, const float* zern_abb, int64_t nz,
                                   float* zmap, double* dwork, maomi_stream_t stream) {
  if (!out || N < 2 || N > 8192 || (nz > 0 && (!zern_abb || !zmap || !dwork))) {
    maomi_set_error("maomi_pupil_gaussian: bad arguments");
    return MAOMI_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  const float rad2 = (float)(rad * rad), obj2 = (float)(objrad * objrad);
  const float ck = (float)(-k / (2.0 * fl));
  const int64_t n = N * N;
  hipLaunchKernelGGL(pupil_kernel, dim3(nblocks(n)), dim3(256), 0, st, (float2*)out, zmap, (int)N,
                     (float)D1, rad2, obj2, ck, (float)objrad, zern_abb, (int)nz);
  MAOMI_CHECK_LAUNCH();
  if (nz > 0) {
    hipLaunchKernelGGL(zero_double_kernel, dim3(1), dim3(64), 0, st, dwork, 1);
    hipLaunchKernelGGL(sum_kernel, dim3(nblocks(n)), dim3(256), 0, st, zmap, n, dwork);
    hipLaunchKernelGGL(zernike_apply_kernel, dim3(nblocks(n)), dim3(256), 0, st, (float2*)out,
                       zmap, n, dwork, (float)k);
    MAOMI_CHECK_LAUNCH();
  }
  return MAOMI_OK;
}

MAOMI_API int maomi_pupil_field(void* out, int64_t n0, int64_t n1, const float* xs,
                                const float* ys, const float* cols,
                                const maomi_pupil_params* p, maomi_stream_t stream) {
  if (!out || !xs || !ys || !p || n0 < 1 || n1 < 1 || n0 > 65536 || n1 > 65536 ||
      p->kind < 0 || p->kind > 3 || p->profile < 0 || p->profile > 3 ||
      (p->kind >= 2 && !cols) || (p->cylinder && (p->kind != 0 || !(p->scale_x > 0.0) ||
                                                  !(p->scale_y > 0.0)))) {
    maomi_set_error("maomi_pupil_field: bad arguments (n0 %lld, n1 %lld)", (long long)n0,
                    (long long)n1);
    return MAOMI_EINVAL;
  }
  const int64_t tot = n0 * n1;
  hipLaunchKernelGGL(pupil_field_kernel, dim3(nblocks(tot)), dim3(256), 0, (hipStream_t)stream,
                     (float2*)out, (int)n0, tot, xs, ys, cols, *p);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

static int fresnel_steps(void* u, int64_t B, int64_t N, double dx0, double dx1, double dz,
                         double wavelength, double nr, const float* screen0,
                         const float* screen1, const void* tw, const float* beam_params,
                         maomi_stream_t stream) {
  if (!u || !tw || B < 1 || N < 2 || N % 2 || !(dz > 0) || !(dx0 > 0) || !(dx1 > 0)) {
    maomi_set_error("maomi_fresnel_step: bad arguments (even N, positive dz/dx required)");
    return MAOMI_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  float2* f = (float2*)u;
  const double lam = wavelength / nr;
  const double k = 2.0 * M_PI / lam;
  const double sc = dx1 / dx0;
  const int64_t n = B * N * N;
  const float onemsc = (float)(1.0 - sc);
  const float inv2dz = 1.0f / (float)(2.0 * dz);
  const float invsc0 = 1.0f / (float)sc;
  hipLaunchKernelGGL(entry_kernel, dim3(nblocks(n)), dim3(256), 0, st, f, (int)B, (int)N,
                     (float)dx0, (float)k, onemsc, inv2dz, screen0, sc != 1.0 ? invsc0 : 1.0f,
                     beam_params);
  MAOMI_CHECK_LAUNCH();
  int rc = maomi_fft2_generic(f, (int)B, (int)N, 0, (const float2*)tw, st);
  if (rc) return rc;
  const float cq = (float)(-M_PI * lam * dz);
  const float hd = (float)(N * dx0);
  const int scaled = sc != 1.0;
  const float invsc = 1.0f / (float)sc;
  hipLaunchKernelGGL(fresnel_q_kernel, dim3(nblocks(n)), dim3(256), 0, st, f, (int)B, (int)N, cq,
                     hd, invsc, scaled, (float)(1.0 / ((double)N * (double)N)), beam_params);
  MAOMI_CHECK_LAUNCH();
  rc = maomi_fft2_generic(f, (int)B, (int)N, 1, (const float2*)tw, st);
  if (rc) return rc;
  const double cexit = k / 2.0 * (sc - 1.0) / (sc * dz);
  hipLaunchKernelGGL(exit_kernel, dim3(nblocks(n)), dim3(256), 0, st, f, (int)B, (int)N,
                     (float)dx1, (float)cexit, screen1, scaled, beam_params);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_fresnel_step(void* u, int64_t B, int64_t N, double dx0, double dx1,
                                 double dz, double wavelength, double nr, const float* screen0,
                                 const float* screen1, const void* tw, maomi_stream_t stream) {
  return fresnel_steps(u, B, N, dx0, dx1, dz, wavelength, nr, screen0, screen1, tw, nullptr,
                       stream);
}

MAOMI_API int maomi_fresnel_steps(void* u, int64_t B, int64_t N, double dx0, double dx1,
                                  const float* beam_params, double wavelength, double nr,
                                  const float* screen0, const float* screen1, const void* tw,
                                  double* dwork, maomi_stream_t stream) {
  if (!beam_params || B > 65535) {
    maomi_set_error("maomi_fresnel_steps: beam parameters required, at most 65535 beams");
    return MAOMI_EINVAL;
  }
  // dz = 1 stands in for the scalar distance: every distance-dependent factor is per beam
  int rc = fresnel_steps(u, B, N, dx0, dx1, 1.0, wavelength, nr, screen0, screen1, tw,
                         beam_params, stream);
  if (rc || !dwork) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = N * N;
  for (int64_t b0 = 0; b0 < B; b0 += 64) {  // zero the per-beam sums, 64 at a time
    hipLaunchKernelGGL(zero_double_kernel, dim3(1), dim3(64), 0, st, dwork + b0,
                       (int)(B - b0 < 64 ? B - b0 : 64));
  }
  const unsigned bx = (unsigned)((n + 255) / 256 < 64 ? (n + 255) / 256 : 64);
  hipLaunchKernelGGL(sumsq_beams_kernel, dim3(bx, (unsigned)B), dim3(256), 0, st,
                     (const float2*)u, n, dwork);
  hipLaunchKernelGGL(scale_beams_kernel, dim3(bx, (unsigned)B), dim3(256), 0, st, (float2*)u, n,
                     dwork);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_normalize_l2(void* u, int64_t n, double* dwork, maomi_stream_t stream) {
  if (!u || !dwork || n < 1) {
    maomi_set_error("maomi_normalize_l2: bad arguments");
    return MAOMI_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(zero_double_kernel, dim3(1), dim3(64), 0, st, dwork, 1);
  hipLaunchKernelGGL(sumsq_kernel, dim3(nblocks(n)), dim3(256), 0, st, (const float2*)u, n, dwork);
  hipLaunchKernelGGL(scale_by_norm_kernel, dim3(nblocks(n)), dim3(256), 0, st, (float2*)u, n,
                     dwork);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

// ---------------------------------------------------------------- per-size tables
namespace {
// e[i] = exp(-a^16), a = (i - N/2) / (0.47 N), float32 as edge_apodization; e2 = e e^T.
__global__ void edge_kernel(float* e1, float* e2, int N, float denom) {
  const int tot = N * N;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < tot; p += gridDim.x * blockDim.x) {
    const int i = p / N, j = p % N;
    const float ai = (float)(i - N / 2) / denom, aj = (float)(j - N / 2) / denom;
    const float ei = expf(-powf(ai, 16.0f)), ej = expf(-powf(aj, 16.0f));
    if (e2) e2[p] = ei * ej;
    if (e1 && i == 0) e1[j] = ej;
  }
}

// q[k] = exp(-i pi lambda' dz f_k^2) / N, f_k = m_k / (N dx), m_k the signed FFT index.
__global__ void q1d_kernel(float2* q, int N, double dx, double dz, double lam) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < N; k += gridDim.x * blockDim.x) {
    const int m = k < N / 2 ? k : k - N;
    const double f = (double)m / ((double)N * dx);
    const double a = -M_PI * lam * dz * f * f;
    q[k] = make_float2((float)(cos(a) / N), (float)(sin(a) / N));
  }
}
}  // namespace

MAOMI_API int maomi_edge_window(float* e1d, float* e2d, int64_t N, maomi_stream_t stream) {
  if ((!e1d && !e2d) || N < 2) { maomi_set_error("maomi_edge_window: bad arguments"); return MAOMI_EINVAL; }
  hipLaunchKernelGGL(edge_kernel, dim3(nblocks(N * N)), dim3(256), 0, (hipStream_t)stream, e1d, e2d,
                     (int)N, (float)(0.47 * (double)N));
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_fresnel_q1d(void* q, int64_t N, double dx, double dz, double lam_eff,
                                maomi_stream_t stream) {
  if (!q || N < 2 || N % 2) { maomi_set_error("maomi_fresnel_q1d: bad arguments"); return MAOMI_EINVAL; }
  hipLaunchKernelGGL(q1d_kernel, dim3(nblocks(N)), dim3(256), 0, (hipStream_t)stream, (float2*)q,
                     (int)N, dx, dz, lam_eff);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

// ---------------------------------------------------------------- phase resize
// resize_xy (maomi/numerics.py:15-39) of a raw (X, Y, Z) field by the supersampling
// factor, X then Y, from host-built tap tables, times the phase scale, written as
// (Z, Xo, Yo) screens (batched.py:52-60 without the complex exponential).
namespace {
__global__ void resize_x_kernel(const float* __restrict__ src, int64_t YZ, const int32_t* __restrict__ idx,
                                const float* __restrict__ w, int m, int K, float* __restrict__ tmp) {
  const int64_t tot = (int64_t)m * YZ;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < tot;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int xo = (int)(p / YZ);
    const int64_t r = p % YZ;
    float acc = 0.f;
    for (int k = 0; k < K; ++k) acc += src[(int64_t)idx[xo * K + k] * YZ + r] * w[xo * K + k];
    tmp[p] = acc;
  }
}
__global__ void resize_y_kernel(const float* __restrict__ tmp, int m0, int64_t Y, int64_t Z,
                                const int32_t* __restrict__ idx, const float* __restrict__ w, int m1,
                                int K, float scale, float* __restrict__ dst) {
  const int64_t tot = (int64_t)Z * m0 * m1;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < tot;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int yo = (int)(p % m1);
    const int xo = (int)((p / m1) % m0);
    const int64_t z = p / ((int64_t)m0 * m1);
    float acc = 0.f;
    for (int k = 0; k < K; ++k) acc += tmp[((int64_t)xo * Y + idx[yo * K + k]) * Z + z] * w[yo * K + k];
    dst[p] = acc * scale;
  }
}
}  // namespace

MAOMI_API int maomi_resize_phase(const float* src, int64_t X, int64_t Y, int64_t Z,
                                 const int32_t* idx0, const float* w0, int64_t m0, int64_t K0,
                                 const int32_t* idx1, const float* w1, int64_t m1, int64_t K1,
                                 double scale, float* tmp, float* dst, maomi_stream_t stream) {
  if (!src || !idx0 || !w0 || !idx1 || !w1 || !tmp || !dst || X < 1 || Y < 1 || Z < 1 || m0 < 1 ||
      m1 < 1 || K0 < 1 || K1 < 1) {
    maomi_set_error("maomi_resize_phase: bad arguments");
    return MAOMI_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(resize_x_kernel, dim3(nblocks(m0 * Y * Z)), dim3(256), 0, st, src, Y * Z, idx0,
                     w0, (int)m0, (int)K0, tmp);
  MAOMI_CHECK_LAUNCH();
  hipLaunchKernelGGL(resize_y_kernel, dim3(nblocks(Z * m0 * m1)), dim3(256), 0, st, tmp, (int)m0, Y,
                     Z, idx1, w1, (int)m1, (int)K1, (float)scale, dst);
  MAOMI_CHECK_LAUNCH();
  return MAOMI_OK;
}

MAOMI_API int maomi_fft2_c64(void* u, int64_t B, int64_t N, int64_t inverse, const void* tw,
                             maomi_stream_t stream) {
  if (!u || !tw || B < 1 || N < 2 || N > 4096) {
    maomi_set_error("maomi_fft2_c64: bad arguments");
    return MAOMI_EINVAL;
  }
  return maomi_fft2_generic((float2*)u, (int)B, (int)N, inverse ? 1 : 0, (const float2*)tw,
                            (hipStream_t)stream);
}
