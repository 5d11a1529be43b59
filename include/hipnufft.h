/*
 * hipnufft -- C ABI of the MI355X (gfx950) NUFFT library.
 *
 * Every pointer argument named *_d is a device pointer (HBM) owned by the
 * caller; every size is in elements.  Complex data is interleaved complex64
 * (float re, float im).  All calls are asynchronous on the plan's stream and
 * return 0 on success or a negative code; hipnufft_last_error() gives the
 * message of the last failure on the calling thread.
 *
 * Conventions (identical to the reference's NDFT, nudft_numpy.py:38-53):
 *   - image arrays are C-ordered, axis 0 slowest;
 *   - sample coordinate d (radians, in [-pi, pi], float32) pairs with image
 *     axis d;
 *   - image index i along an axis of size N is the frequency k = i - N/2;
 *   - type 2 ("op") computes c_m = sum_k f_k exp(iflag * i k.x_m) with
 *     iflag = -1 by default; its adjoint is the type 1 with the opposite sign.
 *
 * Interfaces replaced (reference file:line):
 *   cufinufft Plan(...) / RawCufinufftPlan._make_plan  operators/interfaces/cufinufft.py:83-92
 *   plan._setpts                                        operators/interfaces/cufinufft.py:94-113
 *   Plan.execute (type 1 / type 2)                      operators/interfaces/cufinufft.py:127-133
 *   finufft Plan.execute / execute_adjoint              operators/interfaces/finufft.py:64-76
 *   grad plan (type 2, isign=+1)                        operators/interfaces/cufinufft.py:1227-1237
 *   spread/interp-only plan (Pipe)                      operators/interfaces/cufinufft.py:1311-1318
 *   SENSE op / adj / data consistency chunk loops       operators/interfaces/cufinufft.py:478-500,
 *                                                       :743-770, :1084-1121
 *   _coil_combine_kernel (fused into the adjoint)       operators/gpu_utils.py:12-24
 */
#ifndef HIPNUFFT_H
#define HIPNUFFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hipnufft_plan_s hipnufft_plan;

/* Error codes. */
#define HIPNUFFT_OK 0
#define HIPNUFFT_ERR_ARG -1     /* invalid argument (maps to ValueError) */
#define HIPNUFFT_ERR_HIP -2     /* HIP runtime failure */
#define HIPNUFFT_ERR_NOPTS -3   /* execute before setpts */
#define HIPNUFFT_ERR_ALLOC -4   /* device allocation failed */

/* Tunables.  Zero / negative fields mean "choose automatically". */
typedef struct hipnufft_opts {
  double upsampfac;     /* oversampling factor sigma (default 2.0)          */
  int spreadinterponly; /* 1: grid == image, no FFT, no deconvolution       */
  int tile[3];          /* bin / tile edge per axis (powers of two)          */
  int spread_group;     /* transforms accumulated together in one spread WG */
  int interp_group;     /* transforms gathered together in one interp WG
                           (<= 0 in 2D / 3D: all of a work item's
                           transforms in one WG, the region streamed
                           through LDS; >= 1: that many regions staged side
                           by side; 0 in 1D: 4)                             */
  int max_subproblem;   /* max points per interp work item (dense bins split;
                           0: 128 in 3D, 256 in 2D when streaming, else
                           1024)                                            */
  int deterministic;    /* reserved (spreading is always fixed-order)        */
  int spread_max;       /* max list entries per spread work item (dense
                           tiles are split and their partials summed;
                           0: 1024 in 3D; in 2D 256 up to 8 transforms
                           per plan, 1024 up to 64, else 4096; 4096 in 1D) */
  int spread_method;    /* HIPNUFFT_SPREAD_*: which spreading kernel         */
  int interp_span;      /* bins along the last axis staged together by one
                           interpolation workgroup (0: 2 in 3D, else 1)      */
} hipnufft_opts;

/* Spreading kernels.  CELLS, BLOCKS and COLUMNS sum each cell in the same
 * list order with the same products, so they give the same bits; MATRIX
 * sums the same entries in a fixed order on the matrix cores (the same
 * result on every run, equal to the others to FP32 rounding).  AUTO:
 * COLUMNS for 3D 8 x 8 x 16 tiles when spread_group >= 16 and a set has
 * >= 8 transforms, else MATRIX for 3D tiles with edges of 4, 8 or 16 cells
 * (4, 8 or 16 blocks of 4 x 4 x 4 cells) and for 2D tiles of at most 16
 * blocks of 4 x 16 cells, else CELLS. */
enum {
  HIPNUFFT_SPREAD_AUTO = 0,
  HIPNUFFT_SPREAD_CELLS = 1,   /* thread per (column, 4-cell segment)   */
  HIPNUFFT_SPREAD_BLOCKS = 2,  /* wave per 4^3 blocks, untouched skipped */
  HIPNUFFT_SPREAD_COLUMNS = 3, /* lane per 16-cell column, 16 transforms */
  HIPNUFFT_SPREAD_MATRIX = 4   /* 4^3 blocks on the matrix cores (FP32
                                  MFMA); same order every run, equal to the
                                  others to rounding, not bit for bit; 8
                                  transforms per workgroup, or 16 (same
                                  bits) with spread_group >= 16 */
};

/* Fill *opts with defaults.  Replaces the cufinufft Plan **kwargs
 * (gpu_method, gpu_sort, upsampfac, gpu_spreadinterponly;
 * cufinufft.py:38-46, :83-92). */
int hipnufft_default_opts(hipnufft_opts* opts);

/* Create a plan.  type is 1 or 2 (the plan executes both directions, like
 * finufft's Plan.execute / execute_adjoint, finufft.py:64-76); nmodes[ndim]
 * is the image shape in C order; iflag is the sign of the type-2 exponent
 * (-1 for op, +1 for the reference's gradient plan, cufinufft.py:1227-1237);
 * ntrans is the number of transforms the workspace holds (n_trans).
 * stream is a hipStream_t (NULL = default stream).
 * Replaces Plan(type, n_modes, n_trans, eps, ...) (cufinufft.py:83-92). */
int hipnufft_plan_create(int type, int ndim, const int64_t* nmodes, int iflag,
                         int ntrans, double eps, const hipnufft_opts* opts,
                         void* stream, hipnufft_plan** out);

/* Set the non-uniform points: M samples, coordinate arrays x_d[M] on the
 * device (radians, float32), one per image axis (unused ones NULL).  The
 * points are binned and stably sorted on the device; the plan keeps its own
 * sorted copy (the caller's buffers may be freed afterwards, unlike
 * cufinufft's borrowed pointers, cufinufft.py:105-113).
 * Replaces plan._setpts (cufinufft.py:94-113) / Plan.setpts (finufft.py:55-62). */
int hipnufft_setpts(hipnufft_plan* plan, int64_t M, const float* x0_d,
                    const float* x1_d, const float* x2_d);

/* Set K point sets (trajectories) at once: set k has Ms[k] samples (host
 * array), coordinates of all sets back to back in x_d[sum Ms].  All sets are
 * binned and sorted together (set-major keys), so one launch per stage covers
 * every set whatever its length.  With K > 1 the SENSE entry points take one
 * batch per set (B == K) and ntrans >= C; batch b uses set b, k-space arrays
 * are laid out [k][c][Ms[k]] back to back, smaps_d (if any) holds one map
 * per set and coil [k][c][prod(nmodes)], and hipnufft_execute /
 * execute_adjoint run one transform per set.  K == 1 is hipnufft_setpts.
 * No counterpart in the reference, which plans one trajectory per operator
 * (cufinufft.py:202-275); it replaces a Python loop over per-trajectory
 * operators. */
int hipnufft_setpts_sets(hipnufft_plan* plan, int nsets, const int64_t* Ms,
                         const float* x0_d, const float* x1_d,
                         const float* x2_d);

/* Type 2 over ntrans transforms: in_d[ntrans][prod(nmodes)] ->
 * out_d[ntrans][M].  Replaces Plan.execute for type 2 (cufinufft.py:131-133)
 * and finufft Plan.execute (finufft.py:71-76). */
int hipnufft_execute(hipnufft_plan* plan, const void* in_d, void* out_d);

/* Adjoint (type 1, sign -iflag): in_d[ntrans][M] -> out_d[ntrans][nmodes].
 * Replaces Plan.execute for type 1 (cufinufft.py:127-129) and finufft
 * Plan.execute_adjoint (finufft.py:64-69). */
int hipnufft_execute_adjoint(hipnufft_plan* plan, const void* in_d, void* out_d);

/* SENSE / multi-coil forward over B batches and C coils:
 *   ksp_d[b][c][M] = scale * F(smaps_d[c] * img_d[b])   if smaps_d != NULL
 *   ksp_d[b][c][M] = scale * F(img_d[b][c])             if smaps_d == NULL
 * Replaces MRICufiNUFFT._op_sense_device / _op_calibless_device and the
 * normalisation (cufinufft.py:389-405, :478-500, :609-620). */
int hipnufft_sense_op(hipnufft_plan* plan, const void* img_d,
                      const void* smaps_d, void* ksp_d, int B, int C,
                      float scale);

/* SENSE / multi-coil adjoint:
 *   img_d[b] = scale * sum_c conj(smaps_d[c]) * F^H(density_d * ksp_d[b][c])
 *   (img_d[b][c] = scale * F^H(density_d * ksp_d[b][c]) if smaps_d == NULL)
 * density_d may be NULL.  Replaces _adj_op_sense_device,
 * _adj_op_calibless_device and _coil_combine_kernel (cufinufft.py:743-770,
 * :883-902; gpu_utils.py:12-24). */
int hipnufft_sense_adj(hipnufft_plan* plan, const void* ksp_d,
                       const void* smaps_d, const float* density_d,
                       void* img_d, int B, int C, float scale);

/* Off-resonance corrected SENSE / multi-coil forward with L interpolators:
 *   ksp_d[b][c][m] = sum_l bt_d[m % NK][l] * scale * F(smaps_d[c] * cmap_d[l] * img_d[b])[m]
 * (img_d[b][c] instead of smaps_d[c] * img_d[b] when smaps_d == NULL).
 * cmap_d: the spatial interpolators C, complex64 [L][N]; bt_d: the temporal
 * interpolators B, complex64 [NK][L] (NK samples per shot, NK divides M).
 * The B*C*L transforms go through one batched chain (chunked by the plan's
 * ntrans); uses the plan's ntrans*M residual workspace.  Needs a plan with
 * one point set.  Replaces MRIFourierCorrected.op
 * (operators/off_resonance.py:230-262). */
int hipnufft_orc_op(hipnufft_plan* plan, const void* img_d, const void* smaps_d,
                    const void* cmap_d, const void* bt_d, int L, int64_t NK,
                    void* ksp_d, int B, int C, float scale);

/* Off-resonance corrected adjoint:
 *   img_d[b] = scale * sum_l conj(cmap_d[l]) * sum_c conj(smaps_d[c])
 *              * F^H(density_d * conj(bt_d[. % NK][l]) * ksp_d[b][c])
 * (img_d[b][c] without smaps; density_d may be NULL).  Replaces
 * MRIFourierCorrected.adj_op (operators/off_resonance.py:264-302). */
int hipnufft_orc_adj(hipnufft_plan* plan, const void* ksp_d, const void* smaps_d,
                     const void* cmap_d, const void* bt_d, int L, int64_t NK,
                     const float* density_d, void* img_d, int B, int C,
                     float scale);

/* Data consistency  grad = scale * A^H D (scale * A x - obs) with the same
 * SENSE / calibrationless conventions (D = diag(density_d), may be NULL); the
 * residual lives in a plan-owned ntrans*M buffer.  Replaces
 * _dc_sense_device / _dc_calibless_device (cufinufft.py:1084-1121,
 * :1153-1178). */
int hipnufft_sense_dc(hipnufft_plan* plan, const void* img_d,
                      const void* smaps_d, const void* obs_d,
                      const float* density_d, void* grad_d, int B, int C,
                      float scale);

/* Trajectory gradient of a (SENSE) type-2 transform by the analytic kernel
 * derivative: with y[b][c] = scale * T_iflag(S_c img[b]) (S_c = smaps_d[c],
 * conjugated when flags & 1; calibrationless when smaps_d == NULL, as in
 * hipnufft_sense_op),
 *   gtraj_d[m*ndim + d] = Re sum_{b,c} conj(dy_d[b][c][m]) * D_{m,d} y[b][c][m]
 * where D_{m,d} is the derivative with respect to coordinate d of sample m
 * (radians).  One type-2 chain per coil and no extra NUFFTs, replacing
 * the ndim extra transforms of _backward_op_samples / _backward_adj_samples
 * (autodiff.py:22-41, :69-86). */
int hipnufft_sense_grad_traj(hipnufft_plan* plan, const void* img_d,
                             const void* smaps_d, const void* dy_d, int B,
                             int C, float scale, int iflag, int flags,
                             float* gtraj_d);

/* FP64-accumulated complex dot over n complex64 entries:
 *   out_d[0] + i out_d[1] = sum_j (conj ? conj(a_j) : a_j) * b_j
 * (double precision output on the device), reduced in a fixed order so the
 * result is bit-reproducible.  ws_d is an optional caller workspace of
 * HIPNUFFT_DOT_WS doubles (NULL: the library allocates one per call, which a
 * HIP graph capture cannot contain).  Replaces the CG inner products
 * xp.dot / xp.linalg.norm (extras/optim.py:872-877). */
#define HIPNUFFT_DOT_WS 2048
int hipnufft_dot_c64(const void* a_d, const void* b_d, int64_t n, int conj,
                     double* out_d, double* ws_d, void* stream);

/* rows independent dots of the same kind over a_d[rows][n], b_d[rows][n]:
 * out_d[2r], out_d[2r+1] for row r, each reduced in the fixed order of
 * hipnufft_dot_c64 (so row r equals hipnufft_dot_c64 on that row bit for
 * bit).  ws_d: rows * HIPNUFFT_DOT_WS doubles or NULL; rows <= 65535.  The
 * per-problem inner products of a batched CG (one reconstruction per row);
 * the reference's CG reduces over the whole batch (extras/optim.py:872-877). */
int hipnufft_dot_c64_rows(const void* a_d, const void* b_d, int64_t rows,
                          int64_t n, int conj, double* out_d, double* ws_d,
                          void* stream);

/* hipnufft_dot_c64_rows that allocates nothing: no hipMalloc / hipFree and
 * no caller workspace; the partial sums live in a static device array of
 * the library (one per GPU, 1 MiB, kept until the process ends), so rows <=
 * HIPNUFFT_NOALLOC_DOT_ROWS (HIPNUFFT_ERR_ARG otherwise).  Same kernels and
 * summation order: each row equals hipnufft_dot_c64_rows bit for bit.  Every
 * call on one GPU shares that array, so calls that overlap on one GPU (two
 * streams at once) race; calls in one stream do not.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation). */
#define HIPNUFFT_NOALLOC_DOT_ROWS 64
int hipnufft_dot_c64_rows_noalloc(const void* a_d, const void* b_d,
                                  int64_t rows, int64_t n, int conj,
                                  double* out_d, void* stream);

/* Sorted-point introspection for bit-exact tests: copies the permutation
 * (sorted position -> original index, M int32), the bin offsets
 * (nkeys+1 int32) and the spread lists (entries int32 and offsets
 * nkeys+1 int32) into caller device buffers; any pointer may be NULL.
 * sizes[0..2] receive M, nkeys (point sets x tiles), number of list
 * entries.  Keys are set * tiles + tile. */
int hipnufft_get_sort(const hipnufft_plan* plan, int32_t* perm_d,
                      int32_t* bin_start_d, int32_t* list_d,
                      int32_t* list_start_d, int64_t* sizes);

/* Plan introspection: nf[3] fine grid (C order, leading 1s for ndim < 3),
 * kernel width, beta, tile[3], number of points. */
int hipnufft_plan_info(const hipnufft_plan* plan, int64_t* nf, int* width,
                       double* beta, int* tile, int64_t* M);

/* Bytes of device workspace the plan holds (fine grids, sorted points,
 * lists, tables).  Used for HBM-budget admission. */
int hipnufft_workspace_bytes(const hipnufft_plan* plan, size_t* bytes);

/* Per-stage kernel timing (no counterpart in the reference, whose only
 * instrumentation is the nvtx_mark decorator, operators/gpu_utils.py:52-70).
 * While enabled, HIP events are recorded on the plan stream around every
 * kernel launch; hipnufft_get_timing synchronises, returns the summed
 * milliseconds and launch counts per stage (HIPNUFFT_STAGE_* order) and
 * resets.  Stages: rows_fwd, cols_fwd, slabs_fwd, interp, spread, slabs_adj,
 * cols_adj, rows_adj, grad_traj. */
#define HIPNUFFT_NUM_STAGES 9
int hipnufft_set_timing(hipnufft_plan* plan, int enable);
int hipnufft_get_timing(hipnufft_plan* plan, double* ms, int64_t* launches,
                        int nstage);

/* ---- Toeplitz (Gram) operator ---------------------------------------------
 * G x = scale * A^H W A x for the type-2 transform A of the M points
 * x_d[M] (radians, device) with weights W = diag(weights_d) (NULL: ones),
 * applied as a circulant convolution on a 2N grid: pad, FFT, multiply by the
 * kernel spectrum, inverse FFT, crop -- no spreading or interpolation per
 * application.  The kernel is built once by one type-1 NUFFT of the weights
 * onto the 2N grid (at tolerance eps) and one FFT.  2N must be 2,3,5-smooth
 * on every axis.  With scale = 1/(prod(N) 2^ndim) and W the density,
 * G equals adj_op(op(.)) of the normalised operator.  Replaces
 * compute_toeplitz_kernel / apply_toeplitz_kernel and the gram_op family
 * (operators/toeplitz.py:35-269, :337-468; base.py:316-405). */
typedef struct hipnufft_gram_s hipnufft_gram;
int hipnufft_gram_create(int ndim, const int64_t* nmodes, int64_t M,
                         const float* x0_d, const float* x1_d,
                         const float* x2_d, const float* weights_d,
                         double eps, int iflag, float scale, int ntrans,
                         void* stream, hipnufft_gram** out);
/* SENSE Gram over B batches and C coils (ntrans bounds the workspace):
 *   out_d[b] = sum_c conj(smaps_d[c]) * G(smaps_d[c] * img_d[b])
 *   (out_d[b][c] = G(img_d[b][c]) if smaps_d == NULL). */
int hipnufft_gram_apply(hipnufft_gram* gram, const void* img_d,
                        const void* smaps_d, void* out_d, int B, int C);
int hipnufft_gram_set_stream(hipnufft_gram* gram, void* stream);
int hipnufft_gram_workspace_bytes(const hipnufft_gram* gram, size_t* bytes);
int hipnufft_gram_destroy(hipnufft_gram* gram);

/* ---- Stacked 2.5D (stack-of-stars) ---------------------------------------
 * The Cartesian last image axis z (length Z) of images [.][X*Y][Z] against
 * the NZ sampled planes of a 2D plan, laid out planes_d[t][i][X*Y] (the 2D
 * plan's T*NZ transforms).  F is the centred DFT along z, F u[k] =
 * sum_n u[n] exp(-2 pi i (n - Z/2)(k - Z/2) / Z) (integer Z/2; numpy's
 * fftshift(fft(ifftshift(u)))); any Z up to 2730 (2,3,5-smooth lengths by
 * Stockham passes, others by one direct DFT).  Global transform t0 + t is
 * batch b = (t0 + t) / C, coil c = (t0 + t) % C.  Replaces the z-axis part
 * of MRIStackedNUFFTGPU: _fftz / _ifftz (operators/stacked.py:467-489), the
 * z_index plane selection of _op_sense_device / _op_calibless_device
 * (:527-558, :639-662) and the zero-filled scatter + coil sum of
 * _adj_op_sense_device / _adj_op_calibless_device (:711-740, :767-790). */
typedef struct hipnufft_stack_s hipnufft_stack;
/* z_index[NZ] (host): the sampled plane positions in [0, Z) (a repeated
 * position is filled, in the adjoint, from its last plane, as the
 * reference's fancy-index assignment does). */
int hipnufft_stack_create(int64_t X, int64_t Y, int64_t Z, const int32_t* z_index, int NZ,
                          void* stream, hipnufft_stack** out);
/* planes_d[t][i] = scale * F(smaps_d[c] * img_d[b])[z_index[i]] for t < T
 * (smaps_d NULL: img_d holds one image per global transform). */
int hipnufft_stack_fwd(hipnufft_stack* st, const void* img_d, const void* smaps_d, int C,
                       int t0, int T, float scale, void* planes_d);
/* img_d[b] = scale * sum over the chunk's coils c of batch b of
 *   conj(smaps_d[c]) * F^H(u), u[z_index[i]] = planes_d[t][i], 0 elsewhere;
 * a batch whose coils started in an earlier chunk (t0 % C != 0) is added to.
 * smaps_d NULL: img_d[t0 + t] = scale * F^H(u) per transform. */
int hipnufft_stack_adj(hipnufft_stack* st, const void* planes_d, const void* smaps_d, int C,
                       int t0, int T, float scale, void* img_d);
int hipnufft_stack_set_stream(hipnufft_stack* st, void* stream);
int hipnufft_stack_destroy(hipnufft_stack* st);

/* ---- Coil-map estimation (low-frequency method) ------------------------------
 * The array passes of extras/smaps.py low_frequency (:219-306) around its
 * least-squares NUFFT solve.  Windows of the k-space centre
 * (_extract_kspace_center, :83-170) for samples traj_d[M][d] in cycles
 * ([-0.5, 0.5)) and thresholds threshold[d] (host):
 *   ELLIPSE  1 if sum_d (k_d / t_d)^2 <= 1, else 0;
 *   RECT     1 if |k_d| <= t_d on every axis, else 0 (the reference keeps
 *            only those samples: use the weights to select them);
 *   HANN     0.5 + 0.5 cos(pi |k| / t_0);  HAMMING with a0 = 0.53836.
 * ksp_out_d[c][m] = w_m * ksp_in_d[c][m] for C coils (ksp_in_d NULL: none),
 * weights_d[m] = w_m (NULL: not written). */
#define HIPNUFFT_WINDOW_ELLIPSE 0
#define HIPNUFFT_WINDOW_RECT 1
#define HIPNUFFT_WINDOW_HANN 2
#define HIPNUFFT_WINDOW_HAMMING 3
int hipnufft_kspace_window(const float* traj_d, int64_t M, int d, const float* threshold,
                           int kind, const void* ksp_in_d, void* ksp_out_d, int C,
                           float* weights_d, void* stream);
/* maps_d[c][n] /= sqrt(sum_c |maps_d[c][n]|^2) + eps (the SOS division,
 * smaps.py:300-306; the sum in FP64). */
int hipnufft_rss_normalize(void* maps_d, int C, int64_t N, float eps, void* stream);
/* |maps_d[c]| filtered by a Gaussian of sigma[q] cells along image axis q
 * (0: none), radius int(truncate * sigma + 0.5), "nearest" edges, the phase
 * kept -- skimage.filters.gaussian(|maps|, sigma=(0,) + sigma) * exp(i
 * angle(maps)) (smaps.py:293-298); a zero map value takes phase 0 (numpy's
 * angle of a signed zero is 0 or pi).  work_d: 2 * C * prod(shape) floats. */
int hipnufft_blur_magnitude(void* maps_d, int C, int ndim, const int64_t* shape,
                            const double* sigma, double truncate, float* work_d, void* stream);

/* Change the stream of subsequent calls. */
int hipnufft_set_stream(hipnufft_plan* plan, void* stream);

/* Free the plan and its workspace. */
int hipnufft_plan_destroy(hipnufft_plan* plan);

/* Message of the last error on this thread ("" if none). */
const char* hipnufft_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* HIPNUFFT_H */
