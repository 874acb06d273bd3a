/*
 * mosaic_hip.h -- C ABI of the gfx950 (MI355X) USFFT Radon transform and
 * Fresnel propagator that replace model_mosaic's cupy operators.
 *
 * Conventions
 *   - Every entry point that can fail returns an int status: MM_OK (0) or a
 *     negative MM_ERR_* code.  mm_last_error() returns the calling thread's
 *     message for the last failure.  No C++ exception crosses this ABI.
 *   - Device pointers are caller-owned, contiguous, on the device that is
 *     current when the plan was created.  Host pointers are read during the
 *     call only.
 *   - Work is enqueued on the HIP stream passed as `stream` (NULL = default
 *     stream) and is asynchronous with respect to the host.
 *   - A plan owns its workspace and its per-angle tables.  A plan is not
 *     thread-safe: one in-flight stream per plan.
 *   - float2 = complex64 (interleaved re, im), row-major layouts as stated.
 */
#ifndef MOSAIC_HIP_H
#define MOSAIC_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MM_OK 0
#define MM_ERR_ARG (-1)         /* bad argument, shape or size */
#define MM_ERR_OOM (-2)         /* device allocation failed */
#define MM_ERR_HIP (-3)         /* HIP runtime error */
#define MM_ERR_UNSUPPORTED (-4) /* size the kernels do not cover */

typedef struct mm_radon mm_radon;
typedef struct mm_fresnel mm_fresnel;
typedef struct mm_model mm_model;

/* Message for the last failure on the calling thread ("" if none). */
const char* mm_last_error(void);
/* ABI version (major*100 + minor). */
int mm_version(void);

/* ---------------------------------------------------------------------------
 * USFFT Radon transform.
 * Replaces processing/tomo.py:17 TomoReal (constructor :30, R :121, RT :153,
 * free :109) and the device work of processing/tomo_large.py:46 TomoLargeReal
 * (R :344, RT :582).  eps = 1e-3, mu = -ln(eps)/(2 n^2), m = 4 taps each side
 * (processing/tomo.py:39-42), 2x oversampled Fourier grid.
 *
 * mm_radon_create: n = 2^k in [16, 32768] or 3 * 2^k in [96, 24576], nz_max >= 1,
 *   (above 8192 the 2n-point row and column passes run as two n-point halves,
 *   split by frequency parity forward and by output parity inverse; above 16384
 *   as four n/2-point quarters and the n-point lines as two halves, through a
 *   second buffer of the n object rows; mm_radon_adj_fbp is not built there),
 *   ntheta >= 1 angles in radians (theta_host, host memory, ntheta doubles;
 *   the reference stores theta as float32 -- pass those values).
 *   flags: 0 or an OR of
 *     MM_RADON_FP64_TILES  scatter tiles accumulate in FP64;
 *     MM_RADON_KAHAN_TILES scatter tiles accumulate in FP32 with Kahan
 *                          compensation (not together with FP64_TILES);
 *     MM_RADON_LIST_TILES  scatter tiles accumulate in FP32 on the vector
 *                          ALUs from per-point tap lists;
 *     MM_RADON_GATHER_ZP1  gather tiles of width <= 32 stage one z slice per
 *                          round (default: two slices interleaved per patch
 *                          element; wider tiles always stage one);
 *     MM_RADON_GATHER_ROOT32 gather quadtree tiles of at most 32 x 32 (default
 *                          64 x 64 roots).  The gather flags change the speed,
 *                          not the result.
 *   Without one of the three tile flags the scatter runs as dense 4 x 4-patch
 *   products on the matrix cores (FP32 in, FP32 accumulate) when nz_max >= 16;
 *   those products take 16 slices at a time, so plans with nz_max < 16 use the
 *   FP32 per-point lists instead (as MM_RADON_LIST_TILES).
 */
#define MM_RADON_FP64_TILES 1u
#define MM_RADON_GATHER_ZP1 2u
#define MM_RADON_KAHAN_TILES 4u
#define MM_RADON_LIST_TILES 8u
#define MM_RADON_GATHER_ROOT32 16u
int mm_radon_create(mm_radon** plan, int n, int nz_max, int ntheta, const double* theta_host, unsigned flags);

/* R: obj (nz, n, n) float32 -> sino (ntheta, nz, n) float32, 1 <= nz <= nz_max.
 * Same operator as processing/tomo.py:121-151 (TomoReal.R). */
int mm_radon_fwd(mm_radon* plan, const float* obj_dev, float* sino_dev, int nz, void* stream);

/* R that allocates nothing (no counterpart in the reference): for callers that
 * cannot allocate device memory during the call (code under HIP graph capture,
 * tight serving loops).  Same operator, kernels and result as mm_radon_fwd; the
 * half spectrum and the representatives live in one static device array of
 * 128 MiB per GPU (held until the process ends) instead of the plan's
 * workspace.  Needs ceil256(8 nz (n/4 + 1) 8 n) + 8 ntheta nz (n/2 + 1) bytes
 * of the 128 MiB (n = 256 with 384 angles: up to 91 slices per call; n = 1024
 * with 1024 angles: 6; n = 2048 with 2048 angles: 1); larger calls return
 * MM_ERR_ARG.  The plan's device must be current.  Every call on
 * a GPU uses that GPU's one array: calls that overlap on one GPU (other
 * streams, other threads) race and must be ordered by the caller. */
int mm_radon_fwd_noalloc(mm_radon* plan, const float* obj_dev, float* sino_dev, int nz, void* stream);

/* RT: sino (ntheta, nz, n) float32 (sino_is_c64 = 0) or complex64 (= 1)
 *  -> obj (nz, n, n) of the same element type.
 * Same operator as processing/tomo.py:153-188 (TomoReal.RT): real part for a
 * float32 sinogram, the full complex result for a complex64 one. */
int mm_radon_adj(mm_radon* plan, const void* sino_dev, int sino_is_c64, void* obj_dev, int nz, void* stream);
/* RT of the FBP-filtered real sinogram in one pass over it: replaces
 * FBPFilter.filter (processing/fbp_filter.py:128-210) followed by RT
 * (processing/tomo.py:153-188), the pair step8_fbp.py:204-205 calls.  The
 * filter runs in the length-2n form of mm_fbp_filter_2n (G, A, B as there)
 * inside the line transform; sino_dev (ntheta, nz, n) float32 is not
 * modified; obj_dev (nz, n, n) float32. */
int mm_radon_adj_fbp(mm_radon* plan, const float* sino_dev, const float* G_dev, const float* A_dev,
                     const float* B_dev, float* obj_dev, int nz, void* stream);

/* Stage timing (no counterpart in the reference; for benchmarks).
 * mm_radon_profile(plan, 1) makes every later R / RT call record HIP events
 * on its stream around each kernel stage; mm_radon_stage_ms waits for the
 * events of the last R and RT calls and writes MM_RADON_NSTAGES durations
 * (ms, -1 if the stage has not run): R row rfft, R column FFT, R gather,
 * R line IFFT, RT line FFT, RT scatter (+ partial-tile sum), RT column IFFT,
 * RT row C2R. */
#define MM_RADON_NSTAGES 8
int mm_radon_profile(mm_radon* plan, int enable);
int mm_radon_stage_ms(mm_radon* plan, float* ms_out, int max_stages);

/* Launch knobs (no counterpart in the reference; used by the tile autotuner).
 * MM_RADON_TUNE_GATHER_XCD_GROUP / MM_RADON_TUNE_SCATTER_XCD_GROUP: number G
 * of neighbouring gather / scatter work items that run on one XCD (one L2)
 * before the next XCD takes the next G; 1 = plain round-robin order.
 * MM_RADON_TUNE_COL_SPLIT: 0 = the 2n-point column FFTs at 16384 points per
 * workgroup (4 columns at n = 2048: one workgroup per CU), 1 = half as many
 * columns per workgroup (two workgroups per CU where the LDS allows; the
 * workgroups of one 4-column block kept on one XCD).
 * MM_RADON_TUNE_GATHER_ROOT (64 or 32) and MM_RADON_TUNE_GATHER_ZP (2 or 1):
 * the gather's quadtree root tile and z slices per patch element, as the
 * MM_RADON_GATHER_* flags; setting one rebuilds the gather tables (host work,
 * device allocations and a device synchronisation: not while launches of the
 * plan are being captured into a graph).
 * Results do not depend on the knobs, only the speed does. */
#define MM_RADON_TUNE_GATHER_XCD_GROUP 1
#define MM_RADON_TUNE_SCATTER_XCD_GROUP 2
#define MM_RADON_TUNE_COL_SPLIT 3
#define MM_RADON_TUNE_GATHER_ROOT 4
#define MM_RADON_TUNE_GATHER_ZP 5
#define MM_RADON_TUNE_NKNOBS 6
int mm_radon_tune(mm_radon* plan, int knob, int value);
/* Current value of a knob (or MM_ERR_ARG). */
int mm_radon_get_tune(const mm_radon* plan, int knob);
/* Tile autotuner: times R (obj -> sino) for every gather launch shape (root
 * tile 64 / 32, two / one z slices per patch element) with every gather XCD
 * group size in {1, 2, 4, 8, 16} and keeps the fastest pair; for 512 <= n <=
 * 8192 then R (and RT) with the column split off and on; with back != NULL
 * last RT (sino -> back) for the scatter XCD group size in {4, 8, 16, 32, 64}
 * (default 64 from n = 2048 on, 32 below).  Synchronises the stream (not for use under graph capture).  obj is
 * only read; sino and back are overwritten. */
int mm_radon_autotune(mm_radon* plan, const float* obj_dev, float* sino_dev, float* back_dev, int nz, void* stream);

/* Host-only self-check (no counterpart in the reference; for tests on machines
 * without a GPU): builds the gather / scatter tables of an (n, theta) plan and
 * checks them against the kernels' index arithmetic, and that the scatter
 * lists, decoded as the kernel decodes them, give every written grid point
 * exactly the taps of the samples that land on it.  MM_OK or MM_ERR_ARG with
 * the first violation in mm_last_error(). */
int mm_radon_plan_check(int n, int ntheta, const double* theta_host);

/* Device bytes the plan holds (workspace + tables). */
size_t mm_radon_workspace_bytes(const mm_radon* plan);
void mm_radon_destroy(mm_radon* plan);

/* ---------------------------------------------------------------------------
 * Fresnel propagation.
 * Replaces processing/propagation.py:15 Propagation (constructor :18, D :47)
 * and processing/propagation_large.py:35 PropagationLarge.D (:202); DT is the
 * exact adjoint of D (the reference has no DT, propagation_large.py:15).
 *
 * mm_fresnel_create: frame (nz, n), nz 2^k in [8, 16384] or 3 * 2^k in [24, 12288],
 *   n 2^k in [8, 32768] or 3 * 2^k in [24, 24576]; a side above 8192 needs pad_mode 0
 *   and both sides 2^k >= 16 or 3 * 2^k >= 48 (it runs only on the frame-length DCT
 *   form; rows above 16384 as two n/2-point halves, plain D / DT only);
 *   ndist distances (dist_host, metres), wavelength and voxelsize in metres.
 *   pad_mode 0 = symmetric mirror padding (processing/kernels.py:530-550, the
 *   reference), 1 = zero padding.  The chirp exp(-i pi lambda z f^2) is built in
 *   FP64 with the phase reduced mod 2 pi, then rounded to complex64.
 */
int mm_fresnel_create(mm_fresnel** plan, int n, int nz, int ndist, double wavelength, double voxelsize,
                      const double* dist_host, int pad_mode);

/* D: psi (nframes, nz, n) complex64 -> out (nframes, nz, n) complex64 with
 * distance index j.  out may equal psi (in place). */
int mm_fresnel_fwd(mm_fresnel* plan, const void* psi_dev, void* out_dev, int nframes, int j, void* stream);

/* DT: adjoint of D, same shapes.  out may equal in. */
int mm_fresnel_adj(mm_fresnel* plan, const void* in_dev, void* out_dev, int nframes, int j, void* stream);

size_t mm_fresnel_workspace_bytes(const mm_fresnel* plan);
void mm_fresnel_destroy(mm_fresnel* plan);

/* ---------------------------------------------------------------------------
 * FBP filter.  Replaces processing/fbp_filter.py:128-210 (FBPFilter.filter):
 * every row of sino (rows x n float32, in place) is edge-replicated to
 * ne = 4 n, filtered with the real half-spectrum weights w (ne / 2 + 1
 * float32, FBPFilter.calc_filter, :48-91) and cropped back:
 * row <- crop(irfft(rfft(edge_pad(row)) * w, ne)).  4 n must be 2^k or 3 * 2^k
 * in [64, 16384]. */
int mm_fbp_filter(float* sino_dev, const float* w_dev, long long rows, int n, void* stream);
/* The same filter on transforms of length 2 n (no separate counterpart in the
 * reference; same result as mm_fbp_filter, half the transform length).  With
 * h = irfft(w, 4 n):  row[i] <- sum_j row[j] h[i - j] + row[0] A[i] + row[n-1] B[i]
 * (i, j in [0, n)); the host passes G = rfft(g) / (2 n) (n + 1 float32, real
 * because g is even), g = h on the lags |l| < n laid out mod 2 n, and A[i] /
 * B[i] (n float32 each) = the sums of h over the lags of the left / right edge
 * padding.  2 n must be 2^k or 3 * 2^k in [16, 32768] (above 16384 the
 * transforms run as two n-point halves). */
int mm_fbp_filter_2n(float* sino_dev, const float* G_dev, const float* A_dev, const float* B_dev, long long rows,
                     int n, void* stream);

/* ---------------------------------------------------------------------------
 * Paganin phase retrieval.  mm_paganin_create replaces Paganin.__init__
 * (processing/paganin.py:37-68: n, nz, ntheta frames, wavelength, voxelsize,
 * distance, alpha); mm_paganin_retrieve replaces Paganin.retrieve
 * (:70-92): every (nz, n) float32 frame of in_dev (nframes, nz, n) becomes
 * -log(max(irfft2(H rfft2(frame)), 1e-6)) in out_dev (same shape; may equal
 * in_dev), H = alpha / (wavelength distance (kx^2 + ky^2) / (4 pi) + alpha)
 * on the fftfreq grids of spacing voxelsize.  n and nz: 2^k or 3 * 2^k in
 * [16, 16384]; nframes <= nframes_max. */
typedef struct mm_paganin mm_paganin;
int mm_paganin_create(mm_paganin** plan, int n, int nz, int nframes_max, double wavelength, double voxelsize,
                      double distance, double alpha);
int mm_paganin_retrieve(mm_paganin* plan, const float* in_dev, float* out_dev, int nframes, void* stream);
size_t mm_paganin_workspace_bytes(const mm_paganin* plan);
void mm_paganin_destroy(mm_paganin* plan);

/* ---------------------------------------------------------------------------
 * Element-wise kernels of step3_propagation.py:94-110.
 * psi = exp(-p*inv_norm*inv_beta_ratio) * (cos(p*inv_norm) + i sin(p*inv_norm)).
 */
int mm_psi_from_proj(const float* proj_dev, void* psi_dev, long long count, float inv_norm, float inv_beta_ratio,
                     void* stream);
/* out = re^2 + im^2 (step3_propagation.py:105-110). */
int mm_abs2(const void* z_dev, float* out_dev, long long count, void* stream);

/* ---------------------------------------------------------------------------
 * Uneven mosaic batch: for every tile t (detector width n[t], nz[t] rows,
 * ntheta[t] angles theta_all[sum_{s<t} ntheta[s] ...]),
 *   out[t] = |D(psi(R delta[t]))|^2,  psi(p) = exp(-k beta_ratio p) exp(i k p),
 * at one Fresnel distance.  Replaces the tile loops of the reference pipeline:
 * step2_radon.py:147,191 (R per tile chunk) followed by
 * step3_propagation.py:175,223 (psi, D, |.|^2 per angle chunk, :94-110).
 * Tiles of equal (n, angles) share a forward-only Radon plan, tiles of equal
 * (n, nz) a Fresnel plan.  The tiles are split over B concurrent branches
 * of the graph (flags MM_MOSAIC_BRANCHES(B), B in [1, 8], default 2; longest
 * first, balancing the tiles' HBM bytes), each with its own workspaces (sized
 * for its largest tile) and stream; a branch runs its tiles in order.  Every n[t] must be a valid Radon width and every (nz[t], n[t])
 * a valid Fresnel frame (see mm_radon_create / mm_fresnel_create).  delta[t] (nz, n, n) float32, out[t] (ntheta, nz, n) float32,
 * device pointers.
 * mm_mosaic_run enqueues the whole batch as one HIP graph (captured on first
 * use and again whenever the pointers change) on the batch's own stream,
 * ordered after prior work on `stream` and before later work on it.
 * flags: MM_MOSAIC_NO_GRAPH enqueues the launches directly (same results);
 * MM_MOSAIC_AUTOTUNE runs mm_radon_autotune for every Radon plan on its
 * first tile's input at the first run (synchronous).  Results do not depend
 * on the branch count. */
typedef struct mm_mosaic mm_mosaic;
#define MM_MOSAIC_NO_GRAPH 1u
#define MM_MOSAIC_AUTOTUNE 2u
#define MM_MOSAIC_BRANCHES(b) ((unsigned)(b) << 8)
#define MM_MOSAIC_INFO_RADON_PLANS 0
#define MM_MOSAIC_INFO_FRESNEL_PLANS 1
#define MM_MOSAIC_INFO_CAPTURES 2
#define MM_MOSAIC_INFO_TILES 3
#define MM_MOSAIC_INFO_BRANCHES 4
int mm_mosaic_create(mm_mosaic** plan, int ntiles, const int* n, const int* nz, const int* ntheta,
                     const double* theta_all, double wavelength, double voxelsize, double distance, float k,
                     float beta_ratio, int pad_mode, unsigned flags);
int mm_mosaic_run(mm_mosaic* plan, const float* const* delta_dev, float* const* out_dev, void* stream);
int mm_mosaic_info(const mm_mosaic* plan, int what);
/* Launch knob (MM_RADON_TUNE_*) of tile t's Radon plan (after autotuning: the chosen value), or with
 * MM_MOSAIC_TILE_BRANCH the graph branch tile t runs on. */
#define MM_MOSAIC_TILE_BRANCH 100
int mm_mosaic_tile_knob(const mm_mosaic* plan, int tile, int knob);
size_t mm_mosaic_workspace_bytes(const mm_mosaic* plan);
void mm_mosaic_destroy(mm_mosaic* plan);

/* ---------------------------------------------------------------------------
 * Reference-structured comparator (the benchmark's plain one-slice path; not
 * used by the operators above): the per-sample 81-tap gather and atomic
 * scatter of processing/kernels.py:274-349 and :16-93 on one slice, reading /
 * adding taps directly in global memory.
 * mm_plain_gather: F (2n, n+1) complex64 rfft2 of the padded slice ->
 *   g (ntheta, n) complex64 (the value the reference's gather writes).
 * mm_plain_scatter: g (ntheta, n) complex64 -> G (2n, 2n) complex64, centred
 *   full grid, accumulated (the caller zeroes G). theta_dev: ntheta float32. */
int mm_plain_gather(const void* F_dev, void* g_dev, const float* theta_dev, int n, int ntheta, void* stream);
int mm_plain_scatter(const void* g_dev, void* G_dev, const float* theta_dev, int n, int ntheta, void* stream);

/* ---------------------------------------------------------------------------
 * Differentiable forward model (not in the reference):
 *   p = R delta, b = R beta  (ntheta, nz, n),
 *   psi = exp(i k p - k b), u = D_j psi, I = |u|^2,
 *   L = sum (I - data)^2  (accumulated in FP64).
 * Gradient: G = DT_j(4 (I - data) u), dL/dp = -k Im(conj(G) psi),
 *   dL/db = -k Re(conj(G) psi), dL/ddelta = RT(dL/dp), dL/dbeta = RT(dL/db).
 * The Fresnel plan's frame (nz, n) must equal (nz_call, radon n).
 * mm_model_fwd leaves p, b and u in the model's workspace and returns a token
 * for them (token may be NULL).  mm_model_bwd given that token, with no other
 * mm_model_fwd / mm_model_bwd on the model since and the same delta, beta,
 * data and j (the caller guarantees they are unchanged), starts from p, b, u
 * and recomputes only psi; with token 0 or a stale token it runs the forward
 * first.  psi is never stored.  With mirror padding (pad_mode 0) and frame
 * sides of 2^k >= 16 or 3 * 2^k >= 48 the element-wise steps run inside the
 * Fresnel passes.
 */
int mm_model_create(mm_model** model, mm_radon* radon, mm_fresnel* fresnel, int nz, float k);
/* loss_dev: one double on the device. */
int mm_model_fwd(mm_model* model, const float* delta_dev, const float* beta_dev, const float* data_dev, int j,
                 double* loss_dev, unsigned long long* token, void* stream);
/* grad_* (nz, n, n) float32, multiplied by gscale (the upstream gradient). */
int mm_model_bwd(mm_model* model, const float* delta_dev, const float* beta_dev, const float* data_dev, int j,
                 float gscale, unsigned long long token, float* grad_delta_dev, float* grad_beta_dev, void* stream);
size_t mm_model_workspace_bytes(const mm_model* model);
void mm_model_destroy(mm_model* model);

#ifdef __cplusplus
}
#endif

#endif /* MOSAIC_HIP_H */
