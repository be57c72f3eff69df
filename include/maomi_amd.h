/*
 * maomi_amd.h -- C ABI of the MI355X-native MaoMi optical stage.
 *
 * All pointers are device pointers (owned by the caller, e.g. torch tensors)
 * unless noted.  Every function enqueues work on `stream` and returns 0 on
 * success, a negative MAOMI_E* code when an argument is rejected, or a
 * positive hipError_t from the runtime.  maomi_last_error() gives the text of
 * the last failure on the calling thread.  No function allocates device
 * memory; workspaces are passed in and their sizes are given by the
 * *_workspace_bytes queries.
 *
 * The reference (TomaKakuei/MaoMi 0.8.0b1) has no C ABI: its optical stage is
 * Python driving CuPy RawKernels and torch.fft.  Each entry point below names
 * the reference code it replaces.
 */
#ifndef MAOMI_AMD_H
#define MAOMI_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* maomi_stream_t; /* a hipStream_t (torch.cuda.current_stream().cuda_stream) */

#define MAOMI_OK 0
#define MAOMI_EINVAL (-1)
#define MAOMI_EUNSUPPORTED (-2)

/* Text of the last error on this thread (never NULL). */
const char* maomi_last_error(void);

/* Library version string. */
const char* maomi_version(void);

/* ---------------------------------------------------------------------------
 * Plain split-step stage: every site of a run in one persistent launch.
 *
 * Replaces the per-batch Python loop of OpticalPlan.evaluate / _stage /
 * _screens (maomi/backends/cuda/batched.py:148-201) together with the
 * gather_screens kernel (maomi/backends/cuda/optical_kernels.cu:3-12) and the
 * cuFFT fft2/ifft2 calls it drives, for a single constant-pitch segment
 * (numerics.py:156-210 with sc = 1).  Per site:
 *   u_0 = base;  for s < S: u_s <- u_s * exp(i phi_s(window)) * edge;
 *   M += trap_s |u_s|^4 ;  u_{s+1} = IFFT2(Q * FFT2(u_s));  M += 1/2 |u_S|^4
 * with trap_0 = 1/2.  Field sides N in {64,128,256,512}; sites may mix N and S.
 * ------------------------------------------------------------------------- */
typedef struct maomi_sites_desc {
  const void* bases;          /* complex64 base fields, concatenated                     */
  const int64_t* base_off;    /* [P] complex-element offset of each site's base (device) */
  const float* phase;         /* [Z][H][W] float32 phase (radians), or NULL              */
  const void* phase16;        /* [Z][H][W] float16 phase / slice_scale[z], or NULL       */
  const float* slice_scale;   /* [Z] per-slice scale of phase16                          */
  const float* edge;          /* concatenated 1-D edge windows (N floats per class)      */
  const void* q1d;            /* concatenated 1-D Fresnel factors, complex64, FFT order  */
  const void* twiddle;        /* concatenated W_N^j tables, complex64                    */
  const int32_t* site_x;      /* [P] window top-left row in the phase volume             */
  const int32_t* site_y;      /* [P] window top-left column                              */
  const int32_t* site_n;      /* [P] field side N                                        */
  const int32_t* site_s;      /* [P] number of screens S (<= Z)                          */
  const int32_t* order;       /* [P] processing order (a permutation of 0..P-1)          */
  float* mask;                /* concatenated per-site N*N float32 masks (output)        */
  const int64_t* mask_off;    /* [P] float offset of each site's mask                    */
  double* energy;             /* [P] FP64 sum of each mask (output)                      */
  void* workspace;            /* grid * scratch_stride complex64 elements                */
  int64_t scratch_stride;     /* complex elements per block (>= max N * N)               */
  int64_t P, H, W, Z;
  int64_t grid;               /* persistent blocks (0 = choose)                          */
  int64_t uniform_n;          /* N shared by every site, or 0 for a mixed-N launch       */
  int64_t block_threads;      /* 512 (32-line tiles at N=256) or 256 (16-line); 0 = 512  */
  int64_t class_edge[5];      /* offsets for N = 64,128,256,512,(unused)                 */
  int64_t class_q[5];
  int64_t class_tw[5];
  float* psf_out;             /* optional [psf_nz][P][psf_w][psf_w]: |u|^4 of planes      */
  int64_t psf_z0, psf_nz;     /*   psf_z0 .. psf_z0+psf_nz-1, crop [psf_c0, psf_c0+psf_w)^2 */
  int64_t psf_c0, psf_w;      /*   of each site window (training targets of the surrogate) */
  const void* screen;         /* [Z][H][W] complex64 exp(i phase) (maomi_screen_exp), or NULL;
                                 exactly one of phase / phase16 / screen is given            */
  int32_t* queue;             /* optional device int32 work counter: blocks take the next
                                 site of the order list as they finish one (dynamic
                                 scheduling; the launch resets it first).  NULL: block b takes
                                 entries b, b + grid, ... (static striding)                  */
  const struct maomi_sites_audit* audit;  /* optional in-launch audit (see below), or NULL */
} maomi_sites_desc;

/* In-launch audit of surrogate predictions (no counterpart in the reference).  Needs a
 * focal-plane launch (psf_out), with or without a work counter.  The first n_audit entries of
 * the order list are the audited sites; as each is propagated, its focal planes are compared
 * with its prediction (relative L2, FP64 sums) and counted as a miss above tol.  Once every audited site
 * is done with at most max_fail misses, blocks stop taking unaudited sites: those are left for
 * the surrogate (done[site] = 0).  With more misses every site is propagated, so the fallback
 * costs no second launch.  state and done are zeroed by the launch. */
typedef struct maomi_sites_audit {
  const float* pred;        /* [psf_nz][n_audit][psf_w][psf_w] predictions of audited sites   */
  const int32_t* slot;      /* [P] slot (0..n_audit-1) of each audited site, -1 otherwise     */
  int32_t* state;           /* [2] device counters: audited sites done, misses of tol        */
  float* rel;               /* [n_audit] relative L2 of each audited site (output)            */
  int32_t* done;            /* [P] 1 for every propagated site (output)                       */
  int64_t n_audit, max_fail;
  double tol;
} maomi_sites_audit;

int maomi_persistent_sites(const maomi_sites_desc* desc, maomi_stream_t stream);
/* Workspace bytes for `grid` blocks at largest side nmax. */
int64_t maomi_sites_workspace_bytes(int64_t grid, int64_t nmax);
/* Default persistent grid for this device (blocks resident at once). */
int64_t maomi_sites_default_grid(void);

/* Plan object of the sites launch.  Replaces the captured batches of OpticalPlan.prepare_graph
 * / evaluate_graph (maomi/backends/cuda/batched.py:103-146): one HIP graph holds the whole
 * run's launch (every site, mixed field sides and depths), and
 * the autotuner picks the grid (sites in flight), the block shape (256 or 512 threads; a tile of
 * the line transforms holds block_threads / min(N1, N2) lines of a field side N = N1 N2, e.g.
 * 32 or 16 lines of 256 = 16 x 16), the site ordering and the schedule (static
 * striding of the order list over the blocks, or the work counter) by timing captured
 * candidates.  The plan owns the orderings, the per-block workspace, the work counter and the
 * hipGraphExec; the descriptor's other buffers stay the caller's and are read at each replay
 * (its order, workspace, scratch_stride, grid, block_threads and queue are the plan's).
 * create reads site_n / site_s back to the host (synchronous).  FP32 phase screens only.
 * capture, tune and destroy must not overlap replays still running (they synchronise the
 * device before freeing). */
typedef struct maomi_sites_plan_t* maomi_sites_plan;
enum { MAOMI_ORDER_NATURAL = 0,  /* the caller's site order                                  */
       MAOMI_ORDER_LPT = 1,      /* longest first by S N^2 log2 N                            */
       MAOMI_ORDER_CLASS = 2 };  /* field side, then depth, both descending                  */
enum { MAOMI_SCHEDULE_STATIC = 0,  /* block b takes order entries b, b + grid, ...        */
       MAOMI_SCHEDULE_QUEUE = 1 };  /* each block takes the next entry from the counter    */
typedef struct maomi_sites_choice {
  int64_t grid, block_threads, order, schedule;
  double ms;                     /* best timed replay, or -1 for an untimed capture          */
} maomi_sites_choice;
int maomi_sites_plan_create(const maomi_sites_desc* desc, maomi_sites_plan* plan);
/* Capture one given choice without timing. */
int maomi_sites_plan_capture(maomi_sites_plan plan, int64_t grid, int64_t block_threads,
                             int64_t order, int64_t schedule);
/* Capture and time every (grid, threads, order, schedule) candidate (reps timed replays each,
 * after one warm-up) on `stream` and keep the fastest; results (optional) receives every
 * candidate in grid-major order (schedule innermost), best the winner. */
int maomi_sites_plan_tune(maomi_sites_plan plan, const int64_t* grids, int n_grids,
                          const int64_t* threads, int n_threads, const int64_t* orders,
                          int n_orders, const int64_t* schedules, int n_schedules, int reps,
                          maomi_stream_t stream, maomi_sites_choice* results,
                          maomi_sites_choice* best);
int maomi_sites_plan_replay(maomi_sites_plan plan, maomi_stream_t stream);
int maomi_sites_plan_choice(maomi_sites_plan plan, maomi_sites_choice* out);
int maomi_sites_plan_destroy(maomi_sites_plan plan);

/* ---------------------------------------------------------------------------
 * Phase preparation from a refractive-index volume.
 *
 * Replaces the per-run phase preparation of OpticalPlan.__init__
 * (batched.py:52-60: exp(1j*resize_xy(raw)*scale)) for index volumes given
 * per voxel: phase[z][x][y] = k0 * dz * sum_{v<vps} (n[x][y][z*vps+v] - n0),
 * accumulated in FP64.  `index` is [X][Y][Zv] float32.  Writes float32 phase
 * (phase32 != NULL) and/or float16 phase with per-slice max-abs scale
 * (phase16 != NULL; slice_scale receives max|phase_z| / 1, so stored values
 * are in [-1, 1]).  slice_max must hold Z floats of workspace.
 * ------------------------------------------------------------------------- */
int maomi_phase_prepare(const float* index, int64_t X, int64_t Y, int64_t Zv, int64_t vps,
                        double n0, double k0dz, float* phase32, void* phase16,
                        float* slice_scale, float* slice_max, maomi_stream_t stream);
/* screen[i] = (cos phase[i], sin phase[i]) for n phases (|phase| < 2^17), with the sin/cos of
 * the sites kernels: the complex screens of maomi_sites_desc.screen, computed once per volume
 * voxel instead of once per site and pixel (the screen product of numerics.py:200-205) */
int maomi_screen_exp(const float* phase, int64_t n, void* screen, maomi_stream_t stream);

/* ---------------------------------------------------------------------------
 * Reference-geometry stage: segments A / B / C per site in one persistent launch.
 *
 * Replaces OpticalPlan.evaluate (maomi/backends/cuda/batched.py:170-235) with
 * _stage (157-168), _screens/gather_screens (148-155; optical_kernels.cu:3-12),
 * ResamplePlan/resample_power (native.py:51-105; optical_kernels.cu:14-38) and
 * the tail reductions (batched.py:202-226); same math as optics.propagate
 * (maomi/optics.py:259-340).  Field sides n (A, C) and n2 (B) are any even
 * sizes <= 1024 with n - n2 even.
 * ------------------------------------------------------------------------- */
typedef struct maomi_segment {
  const float* phase;   /* [Z][H][W] phase (radians) of this segment's slabs       */
  int64_t H, W, Z;      /* volume dims, Z = number of screens (>= steps)           */
  int64_t n, steps;     /* field side, propagation steps                           */
  int64_t n1, n2f;      /* line factorisation n = n1 * n2f                         */
  const void* q;        /* complex64 1-D Fresnel tables, n entries each            */
  const int32_t* qidx;  /* [steps] table used by each step                         */
  const float* edge;    /* [n] edge window                                         */
  const void* tw;       /* [n] W_n^j                                               */
} maomi_segment;

typedef struct maomi_segmented_desc {
  maomi_segment seg[3];           /* A, B, C                                          */
  const void* base;               /* n x n complex64 base beam                        */
  const int32_t* site_x;          /* [P] window offsets (supersampled pixels)         */
  const int32_t* site_y;
  const int32_t* order_list;      /* [P] processing order                             */
  int64_t P, cut, tails, order;   /* sites, first top-tail plane, tails on/off, photon order */
  int64_t n, n2, fx, fy;          /* fields and focal window                          */
  int64_t X, Y, TT;               /* focal output X x Y, tail output TT x TT          */
  const float* mfx;               /* [X][fx] focal resample (translate + resize)      */
  const float* mfy;               /* [Y][fy]                                          */
  const float* mt;                /* [TT][n2] tail resample                           */
  const float* ax_focus;          /* [b] axial weights                                */
  const float* ax_top;            /* [a + 1 - cut]                                    */
  const float* ax_bot;            /* [c + 1]                                          */
  double ss2, focal_scale, tail_scale, prop_um;
  float* psf;                     /* [P][X][Y][b]                                     */
  float* top_conv;                /* [P][TT][TT]                                      */
  float* top_prof;                /* [P][a + 1 - cut]                                 */
  double* top_w;                  /* [P]                                              */
  float* bot_conv;                /* [P][TT][TT]                                      */
  float* bot_prof;                /* [P][c + 1]                                       */
  double* bot_w;                  /* [P]                                              */
  void* work;                     /* grid * work_stride complex64 elements            */
  int64_t work_stride;            /* >= maomi_segmented_workspace_elems(...)          */
  int64_t grid;                   /* persistent blocks (0 = one per site)             */
  const int32_t* mfx_band;        /* [X][2] first nonzero column and span of each row of */
  const int32_t* mfy_band;        /*   mfx, mfy, mt (banded resamples: only the span is  */
  const int32_t* mt_band;         /*   summed), or NULL to sum whole rows                */
} maomi_segmented_desc;

int maomi_segmented_sites(const maomi_segmented_desc* desc, maomi_stream_t stream);
/* Statistics of the per-site focal PSFs psf[P][K] for mask assembly (fixed.py:152-189, in FP64):
 * energy[p] = sum_k psf[p][k]; sum_abs[k], sum[k], sum_neg[k] = sums over sites of |x|, x and
 * min(x, 0) (the mean PSFs and the negative resampling mass). */
int maomi_psf_stats(const float* psf, int64_t P, int64_t K, double* energy, double* sum_abs,
                    double* sum, double* sum_neg, maomi_stream_t stream);
/* complex64 elements of workspace per persistent block; TT is the tail output side. */
int64_t maomi_segmented_workspace_elems(int64_t n, int64_t n2, int64_t fx, int64_t fy, int64_t X,
                                        int64_t TT);

/* ---------------------------------------------------------------------------
 * Beam preparation (once per run, inside the timed "plan" stage).
 * ------------------------------------------------------------------------- */

/* W_N^j = exp(-2 pi i j / N), j < N, complex64, computed in FP64.  Twiddles of
 * the FFTs that replace torch.fft.fft2/ifft2 (batched.py:163; numerics.py:200-203). */
int maomi_twiddles(void* tw, int64_t N, maomi_stream_t stream);

/* Gaussian pupil of generate_ba (maomi/_optical_math.py:121-134, 232-266) on an
 * N x N grid of pitch D1: exp(-rho2/rad^2) [rho2 < objrad^2] exp(-i k rho2/(2 fl)),
 * with the Zernike screen of apply_zernike (_optical_math.py:137-197) when nz > 0:
 * zern_abb[j-1] = weight_j * lambda (metres), Noll j = 1..nz.  zmap (N*N floats)
 * and dwork (1 double) are workspace when nz > 0. */
int maomi_pupil_gaussian(void* out, int64_t N, double D1, double rad, double objrad, double k,
                         double fl, const float* zern_abb, int64_t nz, float* zmap,
                         double* dwork, maomi_stream_t stream);

/* Alternative input pupils of maomi/pupils.py: make_pupil (pupils.py:7-34; kinds gaussian,
 * cylindrical, bessel) and one beam of make_vtwins_pupils (pupils.py:54-109), with the Bessel
 * ring profiles of _bessel_amplitude (pupils.py:112-122).  The grid is pupils.py:37-51's:
 * pixel (i, j) of an n1 x n0 row-major complex64 field has x = xs[j], y = ys[i].  Every scalar
 * below is the float32 value the reference's expression yields, computed by the caller; the
 * kernel does the per-pixel float32 arithmetic in the reference's operation order. */
typedef struct maomi_pupil_params {
  int32_t kind;      /* 0 gaussian (make_pupil gaussian/cylindrical), 1 bessel (make_pupil),
                        2 pair beam, gaussian, 3 pair beam, bessel */
  int32_t profile;   /* ring profile of kinds 1 and 3: 0 gaussian, 1 uniform, 2 posax, 3 negax */
  int32_t clip;      /* kind 3: zero outside the objective aperture */
  int32_t cylinder;  /* kind 0: multiply by the cylindrical support (scale_x, scale_y) */
  float rr;          /* kinds 0 and 2: Gaussian radius squared */
  float obj2;        /* objective aperture radius squared */
  float focus;       /* imaginary part of the focus exponent (-i k / (2 fl)) */
  float radius;      /* kinds 1 and 3: ring radius */
  float width;       /* ring width as each profile uses it: gaussian w / (2 sqrt(ln 2)),
                        uniform w / 2, posax and negax w / 1.133 */
  float offset;      /* posax and negax: their width / sqrt(2) */
  float shift;       /* kinds 2 and 3: sign * displacement along x */
  float tilt;        /* kinds 2 and 3: imaginary part of the tilt exponent (i k sign) */
  double scale_x, scale_y;  /* kind 0 with cylinder: the cylindrical scales */
} maomi_pupil_params;

/* The field out[n1][n0] (complex64).  cols[n0] is the pair beams' mean-removed tilt phase per
 * column (pupils.py:99-105; ignored by kinds 0 and 1, may be NULL there).  Replaces the NumPy
 * arithmetic of make_pupil / make_vtwins_pupils (one call per beam). */
int maomi_pupil_field(void* out, int64_t n0, int64_t n1, const float* xs, const float* ys,
                      const float* cols, const maomi_pupil_params* params,
                      maomi_stream_t stream);

/* One Fresnel step of numerics.fresnel (maomi/numerics.py:156-210) for B fields of
 * N x N (N even), in place: entry chirp and screen0, FFT, Q (scaled when dx1 != dx0),
 * inverse FFT, screen1, exit chirp.  This is the base-beam step of
 * OpticalPlan.__init__ (batched.py:43-50).  screen0/screen1 are N x N float32
 * (or NULL for 1); tw is maomi_twiddles(N). */
int maomi_fresnel_step(void* u, int64_t B, int64_t N, double dx0, double dx1, double dz,
                       double wavelength, double nr, const float* screen0, const float* screen1,
                       const void* tw, maomi_stream_t stream);

/* In-place unnormalised 2-D DFT of B square N x N complex64 fields (forward:
 * exp(-2 pi i jk/N); inverse: exp(+...)), the torch.fft.fft2/ifft2 of
 * numerics.fresnel (maomi/numerics.py:200-203) for any N factoring as N1*N2. */
int maomi_fft2_c64(void* u, int64_t B, int64_t N, int64_t inverse, const void* tw,
                   maomi_stream_t stream);

/* u <- u / ||u||_2 with the norm accumulated in FP64 (batched.py:51).  dwork: 1 double. */
int maomi_normalize_l2(void* u, int64_t n, double* dwork, maomi_stream_t stream);
/* B Fresnel steps of maomi_fresnel_step that differ only in their distance dz_b: beam_params
 * (device, B x 3 floats) = (1 / (2 dz_b), -pi lambda' dz_b, k (sc - 1) / (2 sc dz_b)), each
 * rounded to float from double with lambda' = wavelength / nr, k = 2 pi / lambda',
 * sc = dx1 / dx0; then, when dwork (B doubles) is given, each beam scaled to unit L2 norm
 * (the base beams of one field side at every depth, batched.py:43-51) */
int maomi_fresnel_steps(void* u, int64_t B, int64_t N, double dx0, double dx1,
                        const float* beam_params, double wavelength, double nr,
                        const float* screen0, const float* screen1, const void* tw,
                        double* dwork, maomi_stream_t stream);

/* Super-Gaussian edge window of edge_apodization (maomi/_propagation.py:59-61):
 * e[i] = exp(-a_i^16), a_i = (i - N/2)/(0.47 N); e1d (N) and/or e2d = e e^T (N*N). */
int maomi_edge_window(float* e1d, float* e2d, int64_t N, maomi_stream_t stream);

/* 1-D factor of the constant-pitch Fresnel transfer function in FFT order
 * (batched.py:68-82): q[k] = exp(-i pi lam_eff dz f_k^2) / N, f_k = m_k/(N dx). */
int maomi_fresnel_q1d(void* q, int64_t N, double dx, double dz, double lam_eff,
                      maomi_stream_t stream);

/* Phase screens of the reference geometry (batched.py:52-60, numerics.py:15-39):
 * dst[z][xo][yo] = scale * sum_k w1[yo][k] sum_j w0[xo][j] src[idx0[xo][j]][idx1[yo][k]][z]
 * for a raw (X, Y, Z) float32 field; idx/w are the imresize tap tables (m x K);
 * tmp holds m0 * Y * Z floats. */
int maomi_resize_phase(const float* src, int64_t X, int64_t Y, int64_t Z, const int32_t* idx0,
                       const float* w0, int64_t m0, int64_t K0, const int32_t* idx1,
                       const float* w1, int64_t m1, int64_t K1, double scale, float* tmp,
                       float* dst, maomi_stream_t stream);

/* ---------------------------------------------------------------------------
 * Mask assembly (the timed "aggregation" stage), FP64.
 * ------------------------------------------------------------------------- */

/* Green matrix of v4_interpolate (maomi/numerics.py:122-129): G[i][j] = d^2 (ln d - 1),
 * d = |p_i - p_j|, G = 0 at d = 0.  points: P x 2 doubles; G: P x P doubles. */
int maomi_v4_green_matrix_f64(const double* points, int64_t P, double* G, maomi_stream_t stream);

/* Replaces green_apply (maomi/backends/cuda/interpolation.cu:3-19):
 * out[t][c] = sum_i g(|q_t - p_i|^2) coef[i][c], g(r2) = r2 (ln(r2)/2 - 1), 1 <= C <= 8. */
int maomi_v4_apply_f64(const double* points, const double* coef, const double* queries,
                       double* out, int64_t P, int64_t Q, int64_t C, maomi_stream_t stream);
/* maomi_v4_apply_f64 with the point sum split over `rows` block rows for parallelism:
 * partial (rows * Q * C doubles, unused for rows = 1) holds the row sums, added in row order;
 * maomi_v4_apply_rows(P, Q) is a row count that fills the device */
int maomi_v4_apply_split_f64(const double* points, const double* coef, const double* queries,
                             double* out, int64_t P, int64_t Q, int64_t C, double* partial,
                             int64_t rows, maomi_stream_t stream);
int64_t maomi_v4_apply_rows(int64_t P, int64_t Q);
/* maomi_v4_apply_split_f64 for points on a complete nx x ny tensor grid and queries on a
 * complete mx x my tensor grid, with G from a table instead of a logarithm per pair (replaces
 * green_apply, maomi/backends/cuda/interpolation.cu:3-19, on such grids).  du: the U <= 65536
 * distinct axis distances |dx| and |dy| (device); table: U * U doubles of workspace, filled
 * here with G(du[u], du[v]); ix (mx * nx int32): index into du of |qx_a - px_j|; iy
 * (my * ny int32, ny <= 128): of |qy_b - py_k|; every index in [0, U).  coef: nx * ny * C
 * doubles, point (j, k) at row j * ny + k; out: (mx * my) x C, query q = (a, b) at
 * a * my + b.  The point columns are split over `rows` block rows (clamped to nx) whose
 * partial sums (rows * mx * my * C doubles) are added in row order. */
int maomi_v4_apply_grid_f64(const double* du, int64_t U, double* table, const int32_t* ix,
                            const int32_t* iy, const double* coef, double* out, int64_t nx,
                            int64_t ny, int64_t mx, int64_t my, int64_t C, double* partial,
                            int64_t rows, maomi_stream_t stream);
/* maomi_v4_apply_split_f64 that allocates nothing and takes no partial-sum buffer.  No
 * counterpart in the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  The partial sums live in a static device
 * array of MAOMI_V4_STATIC_PARTIALS doubles (32 MiB) that each GPU holds until the process
 * ends; rows * Q * C above that (for rows > 1) is rejected with MAOMI_EINVAL.  Calls that
 * overlap on one GPU share the array and race: order them on one stream. */
#define MAOMI_V4_STATIC_PARTIALS ((int64_t)4 * 1024 * 1024)
int maomi_v4_apply_split_f64_noalloc(const double* points, const double* coef,
                                     const double* queries, double* out, int64_t P, int64_t Q,
                                     int64_t C, int64_t rows, maomi_stream_t stream);
/* Part of the v4 solve that replaces np.linalg.solve (maomi/numerics.py:129): in-place FP64
 * Cholesky factorisation A = L L^T of the positive-definite block of the projected Green
 * matrix.  A is row-major with leading dimension lda >= m; only the lower triangle is read and
 * written (it receives L).  *info (a device int) becomes 0, or the 1-based column of the first
 * non-positive pivot.  workspace: maomi_cholesky_f64_workspace() bytes of device memory.
 * m = 0 only sets *info (A and workspace may be NULL). */
int maomi_cholesky_f64(double* A, int64_t lda, int64_t m, double* workspace, int* info,
                       maomi_stream_t stream);
int64_t maomi_cholesky_f64_workspace(void);
/* Finish of the v4 solve (replaces the rest of np.linalg.solve, maomi/numerics.py:129): w
 * (n x C, row-major) with G0 w = v for the Green matrix G0 of n > 3 points, from G holding the
 * projected system A = Q^T G0 Q (Q = I - V T V^T: V n x 3 row-major, T 3 x 3 upper triangular,
 * the compact WY form of the reflectors of the QR of [1 x y]) with maomi_cholesky_f64's factor
 * of A[3:, 3:] in its lower triangle (row-major, leading dimension ldg >= n).  Solves
 * A[3:, 3:] [Z z] = [A[:3, 3:]^T y2] for y = Q^T v by blocked forward and backward substitution,
 * the 3 x 3 Schur complement by pivoted elimination, and returns w = Q [a; z - Z a].  FP64.
 * v is n x C (1 <= C <= MAOMI_V4_FINISH_MAX_C); workspace: maomi_v4_finish_f64_workspace(n, C)
 * bytes of device memory.  G, V, T and v are not modified. */
#define MAOMI_V4_FINISH_MAX_C 13
int maomi_v4_finish_f64(const double* G, int64_t ldg, int64_t n, const double* V,
                        const double* T, const double* v, int64_t C, double* w,
                        double* workspace, maomi_stream_t stream);
/* Workspace bytes of maomi_v4_finish_f64, or -1 for sizes it rejects. */
int64_t maomi_v4_finish_f64_workspace(int64_t n, int64_t C);

/* ---------------------------------------------------------------------------
 * 2-D PSF contracts (maomi/psf.py), FP64.  maomi_amd.psf is the host side.
 * ------------------------------------------------------------------------- */

/* The interpolation of resample_psf (maomi/psf.py:154-176, scipy map_coordinates order 1,
 * mode "constant"): dst[i][j] = bilinear(src, cx[i], cy[j]), 0 where cx[i] or cy[j] lies
 * outside [0, size - 1].  src: sx x sy row-major; dst: nx x ny. */
int maomi_resample_linear_f64(const double* src, int64_t sx, int64_t sy, const double* cx,
                              int64_t nx, const double* cy, int64_t ny, double* dst,
                              maomi_stream_t stream);
/* The pupil of pupil_to_intensity (maomi/psf.py:179-203): amp exp(2 pi i opd) where opd is
 * finite and amp > 0 (amp NULL: ones), else 0, embedded at n // 2 - p // 2 of an n x n grid and
 * written ifftshifted (complex128, n x n), ready for a forward FFT. */
int maomi_pupil_embed_f64(const double* opd, const double* amp, int64_t px, int64_t py,
                          int64_t n, void* out, maomi_stream_t stream);
/* out = fftshift(|F|^2) for an n x n complex128 F (pupil_to_intensity's intensity before its
 * normalisation). */
int maomi_power_shift_f64(const void* F, int64_t n, double* out, maomi_stream_t stream);

/* ---------------------------------------------------------------------------
 * Learned PSF surrogate (no counterpart in the reference: the reference has no
 * ML).  BF16 MFMA training and inference kernels; activations are FP32
 * channel-major [C][B][H][W], operands of the GEMM are BF16.
 * ------------------------------------------------------------------------- */

/* C[m][n] = alpha sum_k A(m,k) B(k,n) + beta C[m][n] with A(m,k) = A[m sam + k sak],
 * B(k,n) = B[k sbk + n sbn] (BF16), C FP32 row stride ldc; v_mfma_f32_16x16x32_bf16. */
int maomi_gemm_bf16(const void* A, const void* B, float* C, int64_t M, int64_t N, int64_t K,
                    int64_t sam, int64_t sak, int64_t sbk, int64_t sbn, int64_t ldc, float alpha,
                    float beta, maomi_stream_t stream);
/* col[(c k + ky) k + kx][(b Ho + oy) Wo + ox] = X[c][b][s oy + ky - p][s ox + kx - p] (BF16) */
int maomi_im2col_bf16(const float* X, void* col, int64_t C, int64_t B, int64_t H, int64_t W,
                      int64_t k, int64_t s, int64_t p, maomi_stream_t stream);
/* adjoint of im2col: dX = sum of the dcol entries that read each input sample */
int maomi_col2im(const float* dcol, float* dX, int64_t C, int64_t B, int64_t H, int64_t W,
                 int64_t k, int64_t s, int64_t p, maomi_stream_t stream);
/* Y[c][n] = act(Y[c][n] + bias[c]), act = leaky ReLU with negative slope `slope` */
int maomi_bias_act(float* Y, const float* bias, int64_t C, int64_t n, float slope,
                   maomi_stream_t stream);
/* dY <- dY act'(Y) in place; db[c] = sum_n dY[c][n] */
int maomi_act_bwd(float* dY, const float* Y, int64_t C, int64_t n, float slope, float* db,
                  maomi_stream_t stream);
/* learned output planes: out[c][b][i] = in[c][b][i] + T[c][i] (in may equal out) */
int maomi_plane_bias(const float* in, const float* T, float* out, int64_t C, int64_t B, int64_t HW,
                     maomi_stream_t stream);
/* gT[c][i] = scale sum_b dY[c][b][i] (the planes' gradient; scale 1/B gives the batch mean) */
int maomi_plane_bias_grad(const float* dY, float* gT, int64_t C, int64_t B, int64_t HW,
                          float scale, maomi_stream_t stream);
int maomi_to_bf16(const float* x, void* y, int64_t n, maomi_stream_t stream);
/* Implicit-GEMM convolution for inference (BF16 MFMA, FP32 accumulation): X is channel-last
 * BF16 [B][Hs][Ws][C] (C a multiple of 8), read through a nearest 2x upsampling when
 * upsample = 1; W is BF16 [Co][k][k][C]; Y = leaky(bias + conv) with negative slope `slope`,
 * stored channel-last BF16 [B][Ho][Wo][Co] (out_bf16 = 1) or channel-major FP32
 * [Co][B][Ho][Wo] (out_bf16 = 0).  Stride s, zero padding p. */
int maomi_conv_bf16_fwd(const void* X, int64_t B, int64_t Hs, int64_t Ws, int64_t C,
                        int64_t upsample, const void* W, const float* bias, int64_t Co, int64_t k,
                        int64_t s, int64_t p, float slope, void* Y, int64_t out_bf16,
                        maomi_stream_t stream);
/* [C][B][HW] FP32 -> [B][HW][Cp] BF16, channels C..Cp-1 zero. */
int maomi_to_channel_last_bf16(const float* in, int64_t C, int64_t B, int64_t HW, int64_t Cp,
                               void* out, maomi_stream_t stream);
/* nearest 2x upsampling of CB planes of H x W (backward = 2 x 2 sums) */
int maomi_upsample2(const float* in, float* out, int64_t CB, int64_t H, int64_t W,
                    int64_t backward, maomi_stream_t stream);
/* per sample b: loss[b] = |y - t|^2 / |t|^2; grad = 2 (y - t) / |t|^2 / B (grad may be NULL) */
int maomi_rel_sq_loss(const float* y, const float* t, float* grad, float* loss, int64_t C,
                      int64_t B, int64_t HW, maomi_stream_t stream);
/* per sample b: loss[b] = |y - t|^2 / E with E = mean_b |t_b|^2 (the batch's mean target
 * energy); grad = 2 (y - t) / (E B) (may be NULL); work holds 2 B doubles */
int maomi_energy_sq_loss(const float* y, const float* t, float* grad, float* loss, double* work,
                         int64_t C, int64_t B, int64_t HW, maomi_stream_t stream);
/* Adam step (bias-corrected, decoupled weight decay wd) on FP32 w; wb (BF16 copy) may be NULL */
int maomi_adam(float* w, void* wb, const float* g, float* m, float* v, int64_t n, float lr,
               float b1, float b2, float eps, int64_t step, float wd, maomi_stream_t stream);

/* Network input per site p: out[c][p][y][x] (c < zbins) = mean of (n - n0) * inv_std over the
 * pool x pool x (Zv/zbins) voxels of the side x side window at (wx[p], wy[p]) of the
 * [X][Y][Zv] index volume; out[zbins][p][.][.] = depth[p]. */
int maomi_patch_features(const float* index, int64_t X, int64_t Y, int64_t Zv, const int32_t* wx,
                         const int32_t* wy, const float* depth, int64_t P, int64_t side,
                         int64_t pool, int64_t zbins, double n0, double inv_std, double* zsum,
                         float* out, maomi_stream_t stream);
/* (zsum: zbins * X * Y doubles of workspace: the volume's column sums per z bin, computed in
 * one pass before the per-site block sums) */

/* ---------------------------------------------------------------------------
 * Local-PSF FFT plan (replaces the array operations of LocalFFTPlan.__call__,
 * maomi/rana/_fft.py:76-95, and the torch backend it runs on, maomi/rana/_backends.py).
 * FP64 real arrays, complex128 spectra; the FFTs are the caller's.
 * ------------------------------------------------------------------------- */

/* per tile: source window start x, y, size nx, ny, weights offset (-1: unweighted) */
#define MAOMI_LFFT_GATHER_DESC 5
/* per tile: target start x, y, size tx, ty, response start x, y, z, weights offset (-1) */
#define MAOMI_LFFT_SCATTER_DESC 8

/* tiles[b] (Lx x Ly x Lz, zero padded) = source[X][Y][Z] window of tile b times its
 * nx x ny weights (rana/_fft.py:84-86 plus the zero padding rfftn would add) */
int maomi_lfft_gather(const double* source, int64_t X, int64_t Y, int64_t Z, const int32_t* desc,
                      const double* weights, double* tiles, int64_t B, int64_t Lx, int64_t Ly,
                      int64_t Lz, maomi_stream_t stream);
/* out == NULL: spec[i] *= kspec[i] over n complex values; otherwise
 * out[i] = sum_z spec[i][z] kspec[i][z] with z innermost of length nz (plane projection,
 * rana/_fft.py:88-90) */
int maomi_lfft_product(void* spec, const void* kspec, void* out, int64_t n, int64_t nz,
                       maomi_stream_t stream);
/* out[target of b] += full[b][response of b] times the tile's tx x ty weights, atomically;
 * out is X x Y x Zo (Zo = 1 for a plane), full is B x Lx x Ly x Lz (rana/_fft.py:92-94) */
int maomi_lfft_scatter(const double* full, int64_t B, int64_t Lx, int64_t Ly, int64_t Lz,
                       const int32_t* desc, const double* weights, double* out, int64_t X,
                       int64_t Y, int64_t Zo, int64_t Tx, int64_t Ty, maomi_stream_t stream);
/* The same gather and scatter with a per-z weight (zweights: Z, resp. Zo doubles, or NULL):
 * a depth bank's z weights applied on the fly to the source (detection) or to the responses
 * (excitation), rana/operator.py:127-166. */
int maomi_lfft_gather_z(const double* source, int64_t X, int64_t Y, int64_t Z, const int32_t* desc,
                        const double* weights, const double* zweights, double* tiles, int64_t B,
                        int64_t Lx, int64_t Ly, int64_t Lz, maomi_stream_t stream);
int maomi_lfft_scatter_z(const double* full, int64_t B, int64_t Lx, int64_t Ly, int64_t Lz,
                         const int32_t* desc, const double* weights, const double* zweights,
                         double* out, int64_t X, int64_t Y, int64_t Zo, int64_t Tx, int64_t Ty,
                         maomi_stream_t stream);
/* out[i] = sum_f fweights[f] spec[i] kspec[f][i] (nz = 0), or with the sum over z of a plane
 * projection (nz > 0, z innermost): several fields' spectra (kstride complex elements apart)
 * against one source spectrum, one inverse transform for all.  fweights NULL: ones. */
int maomi_lfft_product_fields(const void* spec, const void* kspec, int64_t kstride,
                              const double* fweights, int64_t F, void* out, int64_t n, int64_t nz,
                              maomi_stream_t stream);
/* Node weights of the acquisition operator (rana/operator.py:10-17) on the device: for pixel
 * centres origin + (i + 1/2) pitch, the smoothstep split between the bracketing nodes;
 * out: n_nodes x n_pos. */
int maomi_rana_node_weights(const double* nodes, int64_t n_nodes, double origin, double pitch,
                            int64_t n_pos, double* out, maomi_stream_t stream);
/* Depth weights of a DepthPSFBank (rana/operator.py:55-62): each field's linear hat function
 * over the depth nodes at the positions z (clamped at the ends); out: F x nz. */
int maomi_rana_depth_weights(const double* depths, int64_t F, const double* z, int64_t nz,
                             double* out, maomi_stream_t stream);

/* ---------------------------------------------------------------------------
 * Input generation: spectral Gaussian-process index fields and the collection
 * disk (maomi/_optical_math.py:20-67, maomi/optics.py:120-197 and 384-415).
 * ------------------------------------------------------------------------- */

/* spec (complex64, X x Y x Z) at the ifftshifted index of voxel j =
 * (noise_re[j] + i noise_im[j]) * sqrt(sum_i coef[2i] * (kx*ky*kz)^2 * coef[2i+1]);
 * kernels holds per scale i the 1-D factors [kx (X) | ky (Y) | kz (Z)] and coef = (prod l, w^2)
 * (_optical_math.py:26-57, the envelope and noise product before the inverse FFT) */
int maomi_grf_spectrum(const float* noise_re, const float* noise_im, int64_t X, int64_t Y,
                       int64_t Z, const float* kernels, const float* coef, int32_t K, void* spec,
                       maomi_stream_t stream);
/* out = scale * Re(field at the ifftshifted index) + mu (field NULL: zeros), times first_z on
 * z = 0 and last_z on z = Z - 1, then occupancy + out * (1 - occupancy / occupancy_div) when
 * occupancy is given (_optical_math.py:59-64; optics.py:163-166 and 191-192) */
int maomi_grf_finish(const void* field, int64_t X, int64_t Y, int64_t Z, float scale, float mu,
                     float first_z, float last_z, const float* occupancy, float occupancy_div,
                     float* out, maomi_stream_t stream);
/* For every run r (z0, z1) = (runs[2r], runs[2r + 1]), clamped to [0, Z):
 * out[r * out_stride_run + (x * Y + y) * out_stride_xy] = sum over z0 <= z < z1 of
 * src[x * stride_x + y * stride_y + z] (vessel slab occupancy, optics.py:184-190); src is
 * X x Y x Z with unit z stride, elements float32 (elem_bytes 4) or 0/1 bytes (elem_bytes 1);
 * runs is a device array; the source is read once for all runs */
int maomi_zsum(const void* src, int32_t elem_bytes, int64_t X, int64_t Y, int64_t Z,
               int64_t stride_x, int64_t stride_y, const int32_t* runs, int32_t nruns, float* out,
               int64_t out_stride_run, int64_t out_stride_xy, maomi_stream_t stream);
/* FP64 workspace elements of maomi_collection_mask: ns * ((X * Y + 1) / 2 + X * (Y + 1) + Xo * Yo) */
int64_t maomi_collection_mask_workspace(int64_t X, int64_t Y, int64_t Xo, int64_t Yo,
                                        int32_t ns);
/* Collection mask (optics.py:384-415): for each of ns slabs, slabs[6s..6s+5] = (z0, z1, runs
 * offset, K0, c0 = (K0 - 1) / 2, c1 = (K1 - 1) / 2), the projection of vessels[:, :, z0:z1]
 * (X x Y x Z, element strides stride_x, stride_y, 1; elements as maomi_zsum) is convolved with
 * the slab's disk, given as K0 row runs runs[off + 2k], runs[off + 2k + 1] (first and last
 * column) and the weight weights[s] = 1 / cells, with scipy 'same' placement, on the output
 * window rows ox.., columns oy.. (Xo x Yo); out (float32) = 10 ** (-total / vres * hemoabs) of
 * the slab sum.  slabs, weights and runs are device arrays. */
int maomi_collection_mask(const void* vessels, int32_t elem_bytes, int64_t X, int64_t Y,
                          int64_t Z, int64_t stride_x, int64_t stride_y, const int32_t* slabs,
                          const double* weights, const int32_t* runs, int32_t ns, int64_t ox,
                          int64_t oy, int64_t Xo, int64_t Yo, float vres, float hemoabs,
                          double* workspace, float* out, maomi_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* MAOMI_AMD_H */
