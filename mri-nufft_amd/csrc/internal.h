// Internal declarations of libhipnufft (gfx950).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "hipnufft.h"

namespace hn {

constexpr int kThreads = 256;     // threads per workgroup of every kernel here
constexpr int kMaxElemsFFT = 4096;  // complex elements per FFT workgroup (16 per thread)
constexpr int kMaxRadices = 16;

// Record an error message for hipnufft_last_error().
int set_error(int code, const std::string& msg);

// Failure injection for tests of the error paths, driven from the host:
// when the environment variable HIPNUFFT_INJECT_FAILURE equals ``site``, the
// call fails at that site as an allocation failure would.  No kernel is
// affected; unset, it costs one getenv per call.
inline bool injected_failure(const char* site) {
  const char* v = std::getenv("HIPNUFFT_INJECT_FAILURE");
  return v && std::strcmp(v, site) == 0;
}

#define HN_HIP(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      return ::hn::set_error(HIPNUFFT_ERR_HIP,                              \
                             std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define HN_LAUNCH_CHECK()                                                   \
  do {                                                                      \
    hipError_t e_ = hipGetLastError();                                      \
    if (e_ != hipSuccess)                                                   \
      return ::hn::set_error(HIPNUFFT_ERR_HIP,                              \
                             std::string("kernel launch: ") + hipGetErrorString(e_)); \
  } while (0)

// One FFT axis: length, radix plan, twiddle table, lines per workgroup.
struct FFTAxis {
  int n = 1;
  int L = 1;          // lines per workgroup (power of two)
  int nrad = 0;
  int rad[kMaxRadices] = {0};
  uint32_t ns_magic[kMaxRadices] = {0};  // ceil(2^32 / Ns) of each pass (Ns > 1)
  float2* tw = nullptr;  // tw[t] = exp(-2 pi i t / n), t < n
  int direct = 0;        // n is not 2,3,5-smooth: one direct DFT pass (z axis of stacks)
};

// Geometry of the points / tiles for spreading and interpolation.  Arrays
// are indexed by the D actual dimensions (0 = slowest).
//
// A plan holds nsets point sets (trajectories) of M_k points each, stored
// back to back: set k owns original indices and sorted positions
// [set_off[k], set_off[k+1]).  Bin / tile keys are set * nbins + tile.
struct PtsGeom {
  int D = 0;
  int nf[3] = {1, 1, 1};
  int tile[3] = {1, 1, 1};
  int nt[3] = {1, 1, 1};
  int w = 1;
  float beta = 0.f;
  int64_t M = 0;                                         // all sets
  int nbins = 1;                                         // tiles per set
  int nsets = 1;
  const int32_t* set_off = nullptr;                      // nsets + 1
  const int32_t* cell[3] = {nullptr, nullptr, nullptr};  // sorted order
  const float* off[3] = {nullptr, nullptr, nullptr};     // sorted order
  const int32_t* idx = nullptr;                          // sorted -> original
};

// The point sets one launch covers and how its transforms map onto them:
// sets [k0, k1), ts transforms per set; transform t_local of set k is
// chunk transform (k - k0) * ts + t_local.  A plan with one set shared by
// all transforms of a chunk uses {0, 1, T}.
//
// k-space (non-uniform) arrays are laid out per set as [ts][M_k], sets back
// to back: element (k, t_local, original index i) sits at
//   ts * set_off[k] + t_local * M_k + (i - set_off[k]).
// Sorted-order workspaces are [t_local][M] over all sets (position p).
struct SetRange {
  int k0 = 0, k1 = 1;
  int ts = 1;
};

// Kernel stages timed by the optional per-plan instrumentation.
enum Stage {
  kStageRowsFwd = 0,   // deconvolve / expand / pad / FFT along the contiguous axis
  kStageColsFwd,       // FFT along axis 1 (forward chain)
  kStageSlabsFwd,      // FFT along axis 0 (forward chain, 3D)
  kStageInterp,        // interpolation (type 2)
  kStageSpread,        // spreading (type 1)
  kStageSlabsAdj,      // FFT along axis 0 (adjoint chain, 3D)
  kStageColsAdj,       // FFT along axis 1 (adjoint chain)
  kStageRowsAdj,       // FFT / crop / deconvolve / coil combine (contiguous axis)
  kStageGradTraj,      // kernel-derivative interpolation
  kNumStages
};

struct EventPair {
  int stage;
  hipEvent_t start, stop;
};

}  // namespace hn

struct hipnufft_plan_s {
  int type = 2, ndim = 0, iflag = -1, ntrans = 1;
  int sign2 = -1;  // sign of the type-2 (forward) exponent
  double eps = 1e-6, upsampfac = 2.0;
  int sio = 0;
  int64_t N[3] = {1, 1, 1};   // modes, 3-padded C order (leading 1s)
  int64_t nf[3] = {1, 1, 1};  // fine grid, 3-padded
  int w = 1;
  double beta = 0.0;
  int tile[3] = {1, 1, 1};    // 3-padded
  int nt[3] = {1, 1, 1};      // tiles per axis, 3-padded
  int64_t nbins = 1;           // tiles per point set
  int spread_group = 1, interp_group = 1, max_sub = 1024;
  int spread_method = 0;  // HIPNUFFT_SPREAD_*
  int interp_span = 1;    // bins along the last axis per interpolation work item
  // point sets (trajectories): set_off[k] (host and device, nsets + 1) and the
  // first interp item / spread item / dense tile of each set (host, nsets + 1)
  int nsets = 1;
  std::vector<int64_t> set_off_h{0, 0};
  int32_t* set_off = nullptr;
  int64_t set_off_cap = 0;
  std::vector<int64_t> sub_of_set{0, 0}, item_of_set{0, 0}, dense_of_set{0, 0};
  hipStream_t stream = nullptr;

  float* dec[3] = {nullptr, nullptr, nullptr};  // deconvolution per padded axis
  hn::FFTAxis fft[3];

  // sorted points
  int64_t M = 0;
  int has_pts = 0;
  int32_t* pt_cell[3] = {nullptr, nullptr, nullptr};
  float* pt_off[3] = {nullptr, nullptr, nullptr};
  int32_t* pt_idx = nullptr;
  int32_t* bin_start = nullptr;   // nsets * nbins + 1
  int32_t* sub_tile = nullptr;    // interp work items
  int32_t* sub_begin = nullptr;
  int64_t nsub = 0;
  // per interp work item, per row of its staged region: the 16-cell chunks
  // along the last axis its points read (2D / 3D; mask_rows = 0: none)
  uint16_t* sub_mask = nullptr;
  int64_t sub_mask_cap = 0;
  int mask_rows = 0;
  int32_t* list_pts = nullptr;    // spread lists (sorted-point positions)
  int32_t* list_start = nullptr;  // nsets * nbins + 1
  int64_t nlist = 0;
  int64_t M_cap = 0, list_cap = 0, sub_cap = 0, keys_cap = 0;

  // spread work items: dense tiles are split into items of at most
  // spread_max entries whose partial tiles are summed in a fixed order
  int spread_max = 4096;
  int32_t* it_tile = nullptr;   // per item: key (set * nbins + tile)
  int32_t* it_begin = nullptr;
  int32_t* it_slot = nullptr;   // -1: the item writes the grid directly
  int64_t nitems = 0, items_cap = 0;
  int32_t* dn_tile = nullptr;   // per dense tile
  int32_t* dn_slot0 = nullptr;
  int32_t* dn_count = nullptr;
  int64_t ndense = 0, dense_cap = 0;
  int64_t nslots = 0;
  float2* part = nullptr;       // nslots * ntrans * tile cells
  int64_t part_cap = 0;

  // optional per-stage timing (hipnufft_set_timing)
  int timing = 0;
  std::vector<hn::EventPair> ev_used, ev_free;

  // workspace
  float2* grid = nullptr;   // ntrans * prod(nf)
  float2* kwork = nullptr;  // ntrans * M (data-consistency residual, sorted order)
  int64_t kwork_cap = 0;
  float2* ksorted = nullptr;  // ntrans * M (spreading input, sorted order)
  int64_t ksorted_cap = 0;
  size_t bytes = 0;
};

namespace hn {

// Records HIP events around one kernel launch on the plan stream when the
// plan's timing is enabled (no-op otherwise).
class StageScope {
 public:
  StageScope(const hipnufft_plan_s* p, int stage);
  ~StageScope();

 private:
  hipnufft_plan_s* p_;
  EventPair ev_{};
  bool on_ = false;
};

int64_t grid_size(const hipnufft_plan_s* p);
PtsGeom pts_geom(const hipnufft_plan_s* p);
// dynamic LDS bytes of the last-axis adjoint kernel (lines, twiddles, coil sum)
size_t rows_adj_lds(const FFTAxis& ax, int64_t m2);

// ---- sort.hip -------------------------------------------------------------
// Bins the points of nsets sets (set k: Ms[k] points, coordinates back to
// back in xd) and builds the sorted arrays, interp work items and spread
// lists of the plan.  Synchronises the plan stream.
int build_points(hipnufft_plan_s* p, int nsets, const int64_t* Ms, const float* const* xd);

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X (each
// with its own L2), so consecutive block ids land on different L2s.  This
// bijection on [0, n) gives XCD x the contiguous logical range
// [x*n/8, (x+1)*n/8): neighbouring tiles / column blocks, which share halo
// cells or cache lines, then run on the same XCD.
constexpr int kNumXCD = 8;
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t n) {
  const int64_t q = n / kNumXCD, r = n % kNumXCD;
  const int64_t x = b % kNumXCD, k = b / kNumXCD;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// (x, y, z) block coordinates after the XCD remap of the whole grid
struct BlockIdx3 {
  int x, y, z;
};
__device__ __forceinline__ BlockIdx3 remapped_block() {
  const int64_t nx = gridDim.x, ny = gridDim.y;
  const int64_t n = nx * ny * gridDim.z;
  const int64_t b = blockIdx.x + nx * (blockIdx.y + ny * (int64_t)blockIdx.z);
  const int64_t l = xcd_remap(b, n);
  BlockIdx3 r;
  r.x = (int)(l % nx);
  r.y = (int)((l / nx) % ny);
  r.z = (int)(l / (nx * ny));
  return r;
}

// device helper: the set owning position (original index or sorted
// position) i, by bisection of set_off[0..nsets]
__device__ __forceinline__ int find_set(const int32_t* set_off, int nsets, int64_t i) {
  int lo = 0, hi = nsets;  // set_off[lo] <= i < set_off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (set_off[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}

// ---- spread_interp.hip ---------------------------------------------------
// Interpolation of the chunk's transforms (SetRange r) from the grids.
// out / obs are k-space arrays laid out as SetRange describes.
// mode 0: out = scale * interp
// mode 1: out[t_local][p] = (scale * interp - obs) * (density ? density[i] : 1),
//         written in sorted order (input of launch_spread_sorted)
int launch_interp(const hipnufft_plan_s* p, const float2* grid,
                  int64_t grid_stride, SetRange r, float2* out, const float2* obs,
                  const float* density, float scale, int mode);
// Spreading of in * (density ? density[i] : 1) (k-space layout of SetRange)
// into the chunk's grids (every cell written once; no pre-zeroing).  The
// strengths are first gathered into the plan's sorted-order workspace.
int launch_spread(const hipnufft_plan_s* p, const float2* in,
                  const float* density, SetRange r, float2* grid,
                  int64_t grid_stride);
// Same, for strengths already in sorted-point order (density applied).
int launch_spread_sorted(const hipnufft_plan_s* p, const float2* in_sorted,
                         SetRange r, float2* grid, int64_t grid_stride);
// Trajectory gradient by the kernel derivative (type-2 direction); dy has
// the k-space layout of SetRange.
// The chunk masks of the interpolation work items' regions (sub_mask),
// built once per setpts from the sorted points; leaves mask_rows = 0 (and
// interpolation staging whole regions) where they do not apply.
int build_interp_masks(hipnufft_plan_s* p);
int launch_grad_traj(const hipnufft_plan_s* p, const float2* grid,
                     int64_t grid_stride, SetRange r, const float2* dy,
                     float scale, float* gtraj);

// ---- fft.hip -------------------------------------------------------------
// Forward chain image(s) -> fine grid (deconvolve, pad, all FFT passes).
// Transforms are ordered [image][coil][l] with L interpolator maps cmap[L][N]
// (L = 1 and no cmap for plain SENSE).  Transform tt = t0 + t of the chunk
// uses image tt / per_img (per_img = (smaps ? C : 1) * L), map
// smaps[smaps_all ? tt : (tt / L) % C] and cmap[tt % L].
int fwd_to_grid(const hipnufft_plan_s* p, const float2* img,
                const float2* smaps, int C, int t0, int T, float scale,
                int sign, int conj_smaps = 0, int smaps_all = 0,
                const float2* cmap = nullptr, int L = 1);
// Adjoint chain fine grid -> image(s) (all FFT passes, crop, deconvolve and
// the conjugate combine over the per_img transforms of each image).  Images
// only partly in the chunk [t0, t0+T) are accumulated into img.
int grid_to_img(const hipnufft_plan_s* p, float2* img, const float2* smaps,
                int C, int t0, int T, float scale, int sign, int smaps_all = 0,
                const float2* cmap = nullptr, int L = 1);

// ---- orc.hip -------------------------------------------------------------
// Off-resonance k-space passes for transforms [t0, t0+T) ordered
// [image][coil][l] (L interpolators, NK samples per shot, bt = B[NK][L]).
// combine: y[q] (+)= sum_l B[:, l] * kv[q L + l - t0] for each output row q
// present (kv: the chunk's interpolated k-space, original order); rows
// whose l range starts at 0 are overwritten, the others accumulated.
int orc_combine(const hipnufft_plan_s* p, const float2* kv, const float2* bt, int L, int64_t NK,
                int t0, int T, float2* y);
// gather: out[t] (sorted order) = y[tt / L] * conj(B[:, tt % L]) * density
int orc_gather(const hipnufft_plan_s* p, const float2* y, const float2* bt, int L, int64_t NK,
               const float* density, int t0, int T, float2* out);

// rows independent FP64 dots of a[rows][n] and b[rows][n] -> out[rows][2]
// grid[t] *= K (elementwise complex) for the T grids of the plan workspace
int mul_spectrum(const hipnufft_plan_s* p, const float2* K, int T);
// Toeplitz convolution of chunk transforms [t0, t0+T) (an FFT-only plan whose
// grid is 2N): expand / pad, FFT, multiply by K, inverse FFT, crop / combine
// (forward sign -1; K includes the inverse's 1/prod(2N))
int gram_chain(const hipnufft_plan_s* p, const float2* img, const float2* smaps, float2* out,
               const float2* K, int C, int t0, int T);

// FFT-only plan (no points): images of nmodes embedded, centred, in grids of
// nf >= nmodes per axis, unit deconvolution (the Toeplitz operator's FFTs)
int create_fft_plan(int ndim, const int64_t* nmodes, const int64_t* nf, int ntrans,
                    hipStream_t s, hipnufft_plan_s** out);
void free_plan(hipnufft_plan_s* p);

int dot_c64(const float2* a, const float2* b, int64_t rows, int64_t n, int conj,
            double* out, double* ws, hipStream_t s);

// Stacked 2.5D: the Cartesian last image axis (z, length Z = ax.n) of
// images [.][rows][Z] against the sampled planes [t][NZ][rows] of a 2D plan.
// Global transform t0 + t = b * C + c: image b, map c (smaps NULL: image
// t0 + t).  Forward: planes[t][i][r] = scale * F(z_index[i]) with F the
// centred unitary-free DFT along z of smaps[c] * img[b] (row r).  Adjoint:
// img[b] (+)= scale * sum_c conj(smaps[c]) * F^H(zero-filled planes), plane
// z_plane[k] (or none, -1) at position k.
struct StackGeom {
  int64_t rows = 0;
  int Z = 0, NZ = 0;
  const int32_t* z_index = nullptr;  // [NZ] device
  const int32_t* z_plane = nullptr;  // [Z] device
  FFTAxis ax;
};
int stack_fwd(const StackGeom& g, const float2* img, const float2* smaps, int C, int t0, int T,
              float scale, float2* planes, hipStream_t s);
int stack_adj(const StackGeom& g, const float2* planes, const float2* smaps, int C, int t0, int T,
              float scale, float2* img, hipStream_t s);

// ---- maps.hip ---------------------------------------------------------------
// Coil-map estimation helpers (values of the HIPNUFFT_WINDOW_* constants).
enum { kWindowEllipse = 0, kWindowRect = 1, kWindowHann = 2, kWindowHamming = 3 };
// out[c][m] = w_m * in[c][m] (in may be NULL: weights only), weights[m] = w_m
// (may be NULL) for the window of traj[M][d] (cycles) with thresholds thr[d]
int kspace_window(const float* traj, int64_t M, int d, const float* thr, int kind,
                  const float2* in, float2* out, int C, float* weights, hipStream_t s);
// maps[c][n] /= sqrt(sum_c |maps[c][n]|^2) + eps
int rss_normalize(float2* maps, int C, int64_t N, float eps, hipStream_t s);
// |maps| blurred by a separable Gaussian per axis (sigma[q], scipy's
// truncation and "nearest" edges), phase kept; work holds 2 C N floats
int blur_magnitude(float2* maps, int C, int ndim, const int64_t* shape, const double* sigma,
                   double truncate, float* work, hipStream_t s);

// Exclusive scan of n int32 (device), result in out (may alias in), total in
// *total (host, synchronises the stream) if total != nullptr.
int scan_exclusive(const int32_t* in, int32_t* out, int64_t n,
                   hipStream_t s, int64_t* total);

}  // namespace hn
