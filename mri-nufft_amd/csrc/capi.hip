// C ABI of libhipnufft: plan life cycle, points, transforms, SENSE entry
// points and FP64 reductions.  Declarations and the reference interfaces each
// entry point replaces are in include/hipnufft.h.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "internal.h"

// Partial sums of hipnufft_dot_c64_rows_noalloc: HIPNUFFT_NOALLOC_DOT_ROWS
// rows of HIPNUFFT_DOT_WS / 2 double2 each (1 MiB).  Every GPU the library
// runs on holds its own copy until the process ends; host code reaches it
// only through the per-device address resolved by noalloc_scratch().
static __device__ __align__(256) double2 scratch[HIPNUFFT_NOALLOC_DOT_ROWS * (HIPNUFFT_DOT_WS / 2)];

namespace hn {

thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

StageScope::StageScope(const hipnufft_plan_s* cp, int stage)
    : p_(const_cast<hipnufft_plan_s*>(cp)) {
  if (!p_->timing) return;
  if (p_->ev_free.empty()) {
    EventPair e{};
    if (hipEventCreate(&e.start) != hipSuccess || hipEventCreate(&e.stop) != hipSuccess) return;
    p_->ev_free.push_back(e);
  }
  ev_ = p_->ev_free.back();
  p_->ev_free.pop_back();
  ev_.stage = stage;
  on_ = hipEventRecord(ev_.start, p_->stream) == hipSuccess;
}

StageScope::~StageScope() {
  if (!on_) return;
  (void)hipEventRecord(ev_.stop, p_->stream);
  p_->ev_used.push_back(ev_);
}

namespace {

// ----------------------------------------------------------------------------
// host-side kernel parameters (published FINUFFT choices, see the oracle)
// ----------------------------------------------------------------------------
void kernel_params(double eps, double sigma, int* w, double* beta) {
  int ww;
  if (sigma == 2.0)
    ww = (int)std::ceil(-std::log10(eps / 10.0));
  else
    ww = (int)std::ceil(-std::log(eps) / (M_PI * std::sqrt(1.0 - 1.0 / sigma)));
  ww = std::min(std::max(ww, 2), 16);
  double bw;
  if (sigma == 2.0)
    bw = ww == 2 ? 2.20 : ww == 3 ? 2.26 : ww == 4 ? 2.38 : 2.30;
  else
    bw = 0.97 * M_PI * (1.0 - 1.0 / (2.0 * sigma));
  *w = ww;
  *beta = bw * ww;
}

int64_t next_smooth_even(int64_t n) {
  if (n < 2) n = 2;
  if (n & 1) ++n;
  for (;; n += 2) {
    int64_t m = n;
    for (int f : {2, 3, 5})
      while (m % f == 0) m /= f;
    if (m == 1) return n;
  }
}

// Gauss-Legendre nodes / weights on [-1, 1] (Newton on P_n)
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& wt) {
  x.assign(n, 0.0);
  wt.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-15) break;
    }
    x[i] = z;
    wt[i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

double es_kernel(double u, int w, double beta) {
  const double z = 2.0 * u / w;
  const double a = 1.0 - z * z;
  return a >= 0.0 ? std::exp(beta * (std::sqrt(a) - 1.0)) : 0.0;
}

// 1 / phihat(2 pi k / nf) for k = i - M/2, i < M
std::vector<float> deconv(int64_t M, int64_t nf, int w, double beta) {
  std::vector<double> x, wt;
  gauss_legendre(400, x, wt);
  const double half = 0.5 * w;
  std::vector<double> u(x.size()), f(x.size());
  for (size_t q = 0; q < x.size(); ++q) {
    u[q] = half * (x[q] + 1.0) / 2.0;
    f[q] = es_kernel(u[q], w, beta) * wt[q] * half / 2.0;
  }
  std::vector<float> out(M);
  for (int64_t i = 0; i < M; ++i) {
    const double xi = 2.0 * M_PI * (double)(i - M / 2) / (double)nf;
    double s = 0.0;
    for (size_t q = 0; q < x.size(); ++q) s += std::cos(xi * u[q]) * f[q];
    out[i] = (float)(1.0 / (2.0 * s));
  }
  return out;
}

// lines per FFT workgroup: n * L elements held 16 per thread, plus the
// twiddle table, within 64 KiB of LDS
int lines_per_wg(int64_t n, int max_elems) {
  int L = 16;
  while (L > 1 && (n * L > max_elems || (size_t)n * (L + 2) * sizeof(float2) > 65536)) L >>= 1;
  return L;
}

// allow_direct: a length with other prime factors gets one direct DFT pass
// (O(n^2) per line; the z axis of stacked transforms, whose length is the
// image's)
// strided: an axis transformed by k_fft_strided, which also runs 512-thread
// workgroups, so its lines per workgroup may hold 2 * kMaxElemsFFT elements
int make_axis(int64_t n, FFTAxis* ax, size_t* bytes, bool allow_direct = false,
              bool strided = false) {
  ax->n = (int)n;
  ax->L = lines_per_wg(n, strided ? 2 * kMaxElemsFFT : kMaxElemsFFT);
  if ((size_t)n * (ax->L + 2) * sizeof(float2) > 65536 || n > kMaxElemsFFT)
    return set_error(HIPNUFFT_ERR_ARG, "fine grid axis too long for the LDS FFT (max 2730)");
  int64_t m = n;
  ax->nrad = 0;
  ax->direct = 0;
  while (m % 8 == 0 && ax->nrad < kMaxRadices) { ax->rad[ax->nrad++] = 8; m /= 8; }
  if (m % 4 == 0) { ax->rad[ax->nrad++] = 4; m /= 4; }
  // a remaining factor 2 pairs with a 3 into one radix-6 pass (one LDS round
  // trip fewer): 384 = 8 * 8 * 6
  if (m % 6 == 0) { ax->rad[ax->nrad++] = 6; m /= 6; }
  if (m % 2 == 0) { ax->rad[ax->nrad++] = 2; m /= 2; }
  while (m % 3 == 0 && ax->nrad < kMaxRadices) { ax->rad[ax->nrad++] = 3; m /= 3; }
  while (m % 5 == 0 && ax->nrad < kMaxRadices) { ax->rad[ax->nrad++] = 5; m /= 5; }
  if (m != 1) {
    if (!allow_direct) return set_error(HIPNUFFT_ERR_ARG, "fine grid size is not 2,3,5-smooth");
    ax->nrad = 0;
    ax->direct = 1;
  }
  for (int i = 0, ns = 1; i < ax->nrad; ns *= ax->rad[i], ++i)
    ax->ns_magic[i] = ns > 1 ? (uint32_t)((((uint64_t)1 << 32) + ns - 1) / ns) : 0u;
  std::vector<float2> tw(n);
  for (int64_t t = 0; t < n; ++t) {
    const double a = -2.0 * M_PI * (double)t / (double)n;
    tw[t] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  HN_HIP(hipMalloc((void**)&ax->tw, sizeof(float2) * n));
  HN_HIP(hipMemcpy(ax->tw, tw.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
  *bytes += sizeof(float2) * n;
  return 0;
}

int default_tile(int ndim, int d) {
  if (ndim == 1) return 512;
  if (ndim == 2) return 32;
  return d == 2 ? 16 : 8;
}

int floor_pow2(int64_t x) {
  int p = 1;
  while ((int64_t)p * 2 <= x) p *= 2;
  return p;
}

}  // namespace

void free_plan(hipnufft_plan_s* p) {
  if (!p) return;
  for (auto* v : {&p->ev_used, &p->ev_free})
    for (auto& e : *v) {
      (void)hipEventDestroy(e.start);
      (void)hipEventDestroy(e.stop);
    }
  for (int d = 0; d < 3; ++d) {
    (void)hipFree(p->dec[d]);
    (void)hipFree(p->fft[d].tw);
    (void)hipFree(p->pt_cell[d]);
    (void)hipFree(p->pt_off[d]);
  }
  (void)hipFree(p->pt_idx);
  (void)hipFree(p->bin_start);
  (void)hipFree(p->set_off);
  (void)hipFree(p->sub_tile);
  (void)hipFree(p->sub_begin);
  (void)hipFree(p->sub_mask);
  (void)hipFree(p->list_pts);
  (void)hipFree(p->list_start);
  (void)hipFree(p->it_tile);
  (void)hipFree(p->it_begin);
  (void)hipFree(p->it_slot);
  (void)hipFree(p->dn_tile);
  (void)hipFree(p->dn_slot0);
  (void)hipFree(p->dn_count);
  (void)hipFree(p->part);
  (void)hipFree(p->grid);
  (void)hipFree(p->kwork);
  (void)hipFree(p->ksorted);
  delete p;
}

namespace {

int need_pts(const hipnufft_plan_s* p) {
  if (!p) return set_error(HIPNUFFT_ERR_ARG, "null plan");
  if (!p->has_pts) return set_error(HIPNUFFT_ERR_NOPTS, "setpts has not been called");
  return 0;
}

int64_t img_size(const hipnufft_plan_s* p) { return p->N[0] * p->N[1] * p->N[2]; }

// (re)allocate an ntrans * M complex workspace of the plan
int ensure_kbuf(hipnufft_plan_s* p, float2** buf, int64_t* cap, const char* what) {
  const int64_t need = (int64_t)p->ntrans * std::max<int64_t>(p->M, 1);
  if (*buf && *cap >= need) return 0;
  if (*buf) {
    HN_HIP(hipFree(*buf));
    p->bytes -= sizeof(float2) * *cap;
  }
  *buf = nullptr;
  if (hipMalloc((void**)buf, sizeof(float2) * need) != hipSuccess)
    return set_error(HIPNUFFT_ERR_ALLOC, std::string("cannot allocate the ") + what);
  *cap = need;
  p->bytes += sizeof(float2) * need;
  return 0;
}

int ensure_kwork(hipnufft_plan_s* p) {
  return ensure_kbuf(p, &p->kwork, &p->kwork_cap, "residual workspace");
}

int ensure_ksorted(hipnufft_plan_s* p) {
  return ensure_kbuf(p, &p->ksorted, &p->ksorted_cap, "sorted k-space workspace");
}

// One launch chunk of a SENSE call: chunk transforms [t0, t0 + T) of the
// B*C (batch-major), the point sets they use, and the element offset of the
// chunk's k-space in the caller's array.
struct Chunk {
  int t0, T;
  SetRange r;
  int64_t kofs;
};

// One point set: every batch shares it; chunks of whole batches when they
// fit ntrans (SENSE), otherwise slices of ntrans transforms.
// Several sets: batch b uses set b (B == nsets) with its own C coils; chunks
// of whole sets, and k-space is addressed per set ([k][c][M_k]).
int make_chunks(const hipnufft_plan_s* p, int B, int C, bool sense, std::vector<Chunk>& out) {
  out.clear();
  const int TT = B * C;
  if (p->nsets == 1) {
    int T = p->ntrans;
    if (sense && T >= C) T = (T / C) * C;
    T = std::max(T, 1);
    for (int t0 = 0; t0 < TT; t0 += T) {
      const int Tc = std::min(T, TT - t0);
      out.push_back(Chunk{t0, Tc, SetRange{0, 1, Tc}, (int64_t)t0 * p->M});
    }
    return 0;
  }
  if (B != p->nsets)
    return set_error(HIPNUFFT_ERR_ARG, "a plan with several point sets needs one batch per set");
  if (C > p->ntrans)
    return set_error(HIPNUFFT_ERR_ARG, "a plan with several point sets needs ntrans >= coils");
  const int spc = p->ntrans / C;  // sets per chunk
  for (int k0 = 0; k0 < p->nsets; k0 += spc) {
    const int k1 = std::min(p->nsets, k0 + spc);
    out.push_back(Chunk{k0 * C, (k1 - k0) * C, SetRange{k0, k1, C}, 0});
  }
  return 0;
}

__global__ void k_zero_f32(float* x, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.f;
}

// ----------------------------------------------------------------------------
// FP64 dot: fixed-order two-stage reduction (bit-reproducible)
// ----------------------------------------------------------------------------
constexpr int kDotBlocks = HIPNUFFT_DOT_WS / 2;

// row blockIdx.y of a[rows][n], b[rows][n]; partials part[row][gridDim.x]
__global__ __launch_bounds__(kThreads) void k_dot_partial(const float2* a, const float2* b,
                                                          int64_t n, int conj, double2* part) {
  __shared__ double sr[kThreads], si[kThreads];
  a += (int64_t)blockIdx.y * n;
  b += (int64_t)blockIdx.y * n;
  part += (int64_t)blockIdx.y * gridDim.x;
  double re = 0.0, im = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads) {
    const float2 x = a[i], y = b[i];
    const double xr = x.x, xi = conj ? -(double)x.y : (double)x.y;
    re += xr * y.x - xi * y.y;
    im += xr * y.y + xi * y.x;
  }
  sr[threadIdx.x] = re;
  si[threadIdx.x] = im;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sr[threadIdx.x] += sr[threadIdx.x + s];
      si[threadIdx.x] += si[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = make_double2(sr[0], si[0]);
}

// row blockIdx.x: out[2 row], out[2 row + 1]
__global__ __launch_bounds__(kThreads) void k_dot_final(const double2* part, int nb, double* out) {
  __shared__ double sr[kThreads], si[kThreads];
  part += (int64_t)blockIdx.x * nb;
  out += 2 * (int64_t)blockIdx.x;
  double re = 0.0, im = 0.0;
  for (int i = threadIdx.x; i < nb; i += kThreads) {
    re += part[i].x;
    im += part[i].y;
  }
  sr[threadIdx.x] = re;
  si[threadIdx.x] = im;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sr[threadIdx.x] += sr[threadIdx.x + s];
      si[threadIdx.x] += si[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sr[0];
    out[1] = si[0];
  }
}

}  // namespace

int dot_c64(const float2* a, const float2* b, int64_t rows, int64_t n, int conj, double* out,
            double* ws, hipStream_t s) {
  if (rows <= 0) return 0;
  if (rows > 65535) return set_error(HIPNUFFT_ERR_ARG, "too many dot rows (max 65535)");
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(kDotBlocks, (n + kThreads - 1) / kThreads));
  double2* part = reinterpret_cast<double2*>(ws);
  if (!part) HN_HIP(hipMallocAsync((void**)&part, sizeof(double2) * nb * rows, s));
  hipLaunchKernelGGL(k_dot_partial, dim3(nb, (unsigned)rows), dim3(kThreads), 0, s, a, b, n, conj,
                     part);
  HN_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_dot_final, dim3((unsigned)rows), dim3(kThreads), 0, s, part, nb, out);
  HN_LAUNCH_CHECK();
  if (!ws) HN_HIP(hipFreeAsync(part, s));
  return 0;
}

namespace {

std::mutex g_scratch_mu;
std::vector<double2*> g_scratch_addr;  // device address of `scratch`, per device

// Device address of `scratch` on the device of stream s, resolved once per
// device with that device current (a symbol's host-side name is not a
// device address).
int noalloc_scratch(hipStream_t s, double2** out) {
  int dev = 0;
  HN_HIP(hipGetDevice(&dev));
  if (s) HN_HIP(hipStreamGetDevice(s, &dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_addr.size() <= dev) g_scratch_addr.resize(dev + 1, nullptr);
  if (!g_scratch_addr[dev]) {
    int cur = 0;
    HN_HIP(hipGetDevice(&cur));
    if (cur != dev) HN_HIP(hipSetDevice(dev));
    void* addr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
    if (cur != dev) HN_HIP(hipSetDevice(cur));
    if (e != hipSuccess || !addr)
      return set_error(HIPNUFFT_ERR_HIP,
                       std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_addr[dev] = (double2*)addr;
  }
  *out = g_scratch_addr[dev];
  return 0;
}

}  // namespace

int create_fft_plan(int ndim, const int64_t* nmodes, const int64_t* nf, int ntrans, hipStream_t s,
                    hipnufft_plan_s** out) {
  *out = nullptr;
  hipnufft_plan_s* p = new (std::nothrow) hipnufft_plan_s();
  if (!p) return set_error(HIPNUFFT_ERR_ALLOC, "host allocation failed");
  p->ndim = ndim;
  p->ntrans = ntrans;
  p->stream = s;
  p->sign2 = -1;
  const int off = 3 - ndim;
  int rc = 0;
  for (int d = 0; d < ndim; ++d) {
    if (nmodes[d] < 1 || nf[d] < nmodes[d] || (nf[d] & 1)) rc = set_error(HIPNUFFT_ERR_ARG, "invalid FFT grid");
    p->N[off + d] = nmodes[d];
    p->nf[off + d] = nf[d];
  }
  if (!rc && p->nf[0] * p->nf[1] * p->nf[2] >= ((int64_t)1 << 30))
    rc = set_error(HIPNUFFT_ERR_ARG, "grid too large (at most 2^30 cells)");
  for (int d = 0; d < 3 && !rc; ++d) {
    std::vector<float> dv(p->N[d], 1.0f);
    if (hipMalloc((void**)&p->dec[d], sizeof(float) * dv.size()) != hipSuccess ||
        hipMemcpy(p->dec[d], dv.data(), sizeof(float) * dv.size(), hipMemcpyHostToDevice) !=
            hipSuccess)
      rc = set_error(HIPNUFFT_ERR_ALLOC, "cannot allocate deconvolution tables");
    p->bytes += sizeof(float) * dv.size();
  }
  for (int d = off; d < 3 && !rc; ++d)
    rc = make_axis(p->nf[d], &p->fft[d], &p->bytes, false, d < 2);
  while (!rc && p->fft[2].L > 1 && rows_adj_lds(p->fft[2], p->N[2]) > 65536) p->fft[2].L >>= 1;
  if (!rc && rows_adj_lds(p->fft[2], p->N[2]) > 65536)
    rc = set_error(HIPNUFFT_ERR_ARG, "grid too long along the last axis (max 2048)");
  if (!rc) {
    const size_t g = sizeof(float2) * (size_t)grid_size(p) * (size_t)ntrans;
    if (hipMalloc((void**)&p->grid, g) != hipSuccess)
      rc = set_error(HIPNUFFT_ERR_ALLOC, "cannot allocate the grid workspace");
    p->bytes += g;
  }
  if (rc) {
    free_plan(p);
    return rc;
  }
  *out = p;
  return 0;
}

}  // namespace hn

// Toeplitz (Gram) operator: the FFT plan of the images in a 2N grid and the
// spectrum of the circulant embedding of the point-spread kernel.
struct hipnufft_gram_s {
  hipnufft_plan_s* fft = nullptr;
  float2* K = nullptr;
  size_t bytes = 0;
};

struct hipnufft_stack_s {
  hn::StackGeom g;
  int32_t* z_index = nullptr;  // device
  int32_t* z_plane = nullptr;  // device
  hipStream_t stream = nullptr;
};

using namespace hn;

namespace {

__global__ void k_cplx_weights(const float* w, int64_t M, float2* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < M) out[i] = make_float2(w ? w[i] : 1.0f, 0.0f);
}

}  // namespace

extern "C" {

const char* hipnufft_last_error(void) { return g_last_error.c_str(); }

int hipnufft_default_opts(hipnufft_opts* o) {
  if (!o) return set_error(HIPNUFFT_ERR_ARG, "null opts");
  o->upsampfac = 2.0;
  o->spreadinterponly = 0;
  o->tile[0] = o->tile[1] = o->tile[2] = 0;
  o->spread_group = 0;
  o->interp_group = 0;
  o->max_subproblem = 0;
  o->deterministic = 0;
  o->spread_max = 0;
  o->spread_method = HIPNUFFT_SPREAD_AUTO;
  o->interp_span = 0;
  return 0;
}

int hipnufft_plan_create(int type, int ndim, const int64_t* nmodes, int iflag, int ntrans,
                         double eps, const hipnufft_opts* opts, void* stream,
                         hipnufft_plan** out) {
  g_last_error.clear();
  if (!out) return set_error(HIPNUFFT_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (type != 1 && type != 2) return set_error(HIPNUFFT_ERR_ARG, "type must be 1 or 2");
  if (ndim < 1 || ndim > 3) return set_error(HIPNUFFT_ERR_ARG, "ndim must be 1, 2 or 3");
  if (!nmodes) return set_error(HIPNUFFT_ERR_ARG, "null nmodes");
  if (ntrans < 1) return set_error(HIPNUFFT_ERR_ARG, "ntrans must be >= 1");
  if (!(eps > 0.0)) return set_error(HIPNUFFT_ERR_ARG, "eps must be > 0");
  if (iflag != 1 && iflag != -1) return set_error(HIPNUFFT_ERR_ARG, "iflag must be +1 or -1");
  hipnufft_opts o;
  hipnufft_default_opts(&o);
  if (opts) o = *opts;
  if (!(o.upsampfac > 1.0)) o.upsampfac = 2.0;

  hipnufft_plan_s* p = new (std::nothrow) hipnufft_plan_s();
  if (!p) return set_error(HIPNUFFT_ERR_ALLOC, "host allocation failed");
  p->type = type;
  p->ndim = ndim;
  p->iflag = iflag;
  p->sign2 = (type == 2) ? iflag : -iflag;
  p->ntrans = ntrans;
  p->eps = eps;
  p->upsampfac = o.upsampfac;
  p->sio = o.spreadinterponly ? 1 : 0;
  p->stream = (hipStream_t)stream;
  kernel_params(eps, o.upsampfac, &p->w, &p->beta);
  const int off = 3 - ndim;
  for (int d = 0; d < ndim; ++d) {
    if (nmodes[d] < 1 || nmodes[d] > (1 << 20)) {
      free_plan(p);
      return set_error(HIPNUFFT_ERR_ARG, "invalid mode count");
    }
    p->N[off + d] = nmodes[d];
    int64_t nf;
    if (p->sio) {
      nf = nmodes[d];
      if (nf < p->w) {
        free_plan(p);
        return set_error(HIPNUFFT_ERR_ARG, "spread/interp-only grid smaller than the kernel width");
      }
    } else {
      nf = (int64_t)(o.upsampfac * (double)nmodes[d]);
      nf = next_smooth_even(std::max<int64_t>(nf, 2 * p->w));
    }
    p->nf[off + d] = nf;
    if (p->nf[0] * p->nf[1] * p->nf[2] >= ((int64_t)1 << 30)) {
      free_plan(p);
      return set_error(HIPNUFFT_ERR_ARG, "fine grid too large (at most 2^30 cells per transform)");
    }
    int t = o.tile[d] > 0 ? o.tile[d] : default_tile(ndim, d);
    t = floor_pow2(std::max(t, 1));
    t = std::min(t, floor_pow2(nf));
    p->tile[off + d] = t;
    p->nt[off + d] = (int)((nf + t - 1) / t);
  }
  p->nbins = (int64_t)p->nt[0] * p->nt[1] * p->nt[2];
  if (p->nbins >= ((int64_t)1 << 30)) {
    free_plan(p);
    return set_error(HIPNUFFT_ERR_ARG, "too many tiles");
  }
  p->spread_group = o.spread_group > 0 ? o.spread_group : 0;  // 0: chosen per launch
  // 2D / 3D: 0 = every transform of a work item in one workgroup (items of
  // up to 128 points in 3D: more, smaller items measured faster at 16 coils
  // than 256 or 512; 256 in 2D: the stacked 128^3 slices' type 2 took 2.29
  // ms against 2.34 at 128, radial 256^2 at 8 and 64 coils the same within
  // 2%); an explicit group picks the kernel that stages that many
  // transforms side by side (items of 1024 points)
  p->interp_group = o.interp_group > 0 ? o.interp_group : ndim == 1 ? 4 : 0;
  p->max_sub = o.max_subproblem > 0 ? o.max_subproblem
               : ndim == 3              ? 128
               : ndim == 2 && p->interp_group == 0 ? 256 : 1024;
  // dense tiles (a radial k-space centre holds ~80 K list entries per tile)
  // split, so that their work items do not form a tail: 1024 entries in 3D.
  // In 2D (matrix-core spreading) the best
  // split grows with the transforms, whose partial tiles the reduction sums:
  // op+adj of radial 256^2, 128 x 512 samples, 8 coils 0.18 ms at 256, 0.23
  // at 1024; 256 x 512 at 32 coils 0.69 ms at 256, 0.58 at 1024; 1024
  // stacked planes 4.23 ms at 512, 3.76 at 4096, 6.25 at 8192
  const int split2d = ntrans <= 8 ? 256 : ntrans <= 64 ? 1024 : 4096;
  p->spread_max = o.spread_max > 0 ? o.spread_max : (ndim == 3 ? 1024 : ndim == 2 ? split2d : 4096);
  if (o.spread_method < HIPNUFFT_SPREAD_AUTO || o.spread_method > HIPNUFFT_SPREAD_MATRIX) {
    free_plan(p);
    return set_error(HIPNUFFT_ERR_ARG, "unknown spread_method");
  }
  p->spread_method = o.spread_method;
  p->interp_span = o.interp_span > 0 ? o.interp_span : (ndim == 3 ? 2 : 1);
  p->interp_span = (int)std::min<int64_t>(p->interp_span, p->nt[2]);

  int rc = 0;
  // deconvolution tables (padded axes hold a single 1)
  for (int d = 0; d < 3 && !rc; ++d) {
    std::vector<float> dv(p->N[d], 1.0f);
    if (!p->sio && d >= off) dv = deconv(p->N[d], p->nf[d], p->w, p->beta);
    if (hipMalloc((void**)&p->dec[d], sizeof(float) * dv.size()) != hipSuccess ||
        hipMemcpy(p->dec[d], dv.data(), sizeof(float) * dv.size(), hipMemcpyHostToDevice) !=
            hipSuccess)
      rc = set_error(HIPNUFFT_ERR_ALLOC, "cannot allocate deconvolution tables");
    p->bytes += sizeof(float) * dv.size();
  }
  if (!rc && !p->sio) {
    for (int d = off; d < 3 && !rc; ++d)
    rc = make_axis(p->nf[d], &p->fft[d], &p->bytes, false, d < 2);
    // the fastest axis' adjoint kernel also holds its coil sum in LDS
    while (!rc && p->fft[2].L > 1 && rows_adj_lds(p->fft[2], p->N[2]) > 65536) p->fft[2].L >>= 1;
    if (!rc && rows_adj_lds(p->fft[2], p->N[2]) > 65536)
      rc = set_error(HIPNUFFT_ERR_ARG, "fine grid too long along the last axis (max 2048)");
    if (!rc) {
      const size_t g = sizeof(float2) * (size_t)grid_size(p) * (size_t)ntrans;
      if (hipMalloc((void**)&p->grid, g) != hipSuccess)
        rc = set_error(HIPNUFFT_ERR_ALLOC, "cannot allocate the fine-grid workspace");
      p->bytes += g;
    }
  }
  if (rc) {
    free_plan(p);
    return rc;
  }
  *out = p;
  return 0;
}

int hipnufft_setpts(hipnufft_plan* p, int64_t M, const float* x0, const float* x1,
                    const float* x2) {
  g_last_error.clear();
  if (!p) return set_error(HIPNUFFT_ERR_ARG, "null plan");
  if (M < 0) return set_error(HIPNUFFT_ERR_ARG, "negative number of points");
  const float* xd[3] = {x0, x1, x2};
  return build_points(p, 1, &M, xd);
}

int hipnufft_setpts_sets(hipnufft_plan* p, int nsets, const int64_t* Ms, const float* x0,
                         const float* x1, const float* x2) {
  g_last_error.clear();
  if (!p) return set_error(HIPNUFFT_ERR_ARG, "null plan");
  if (nsets < 1 || !Ms) return set_error(HIPNUFFT_ERR_ARG, "need at least one point set");
  if (nsets > 1 && p->sio)
    return set_error(HIPNUFFT_ERR_ARG, "spread/interp-only plans take a single point set");
  const float* xd[3] = {x0, x1, x2};
  return build_points(p, nsets, Ms, xd);
}

int hipnufft_sense_op(hipnufft_plan* p, const void* img, const void* smaps, void* ksp, int B,
                      int C, float scale) {
  int rc = need_pts(p);
  if (rc) return rc;
  if (B < 1 || C < 1) return set_error(HIPNUFFT_ERR_ARG, "B and C must be >= 1");
  if (!img || !ksp) return set_error(HIPNUFFT_ERR_ARG, "null data pointer");
  const float2* x = (const float2*)img;
  const float2* s = (const float2*)smaps;
  float2* y = (float2*)ksp;
  if (p->sio) {
    if (s) return set_error(HIPNUFFT_ERR_ARG, "spread/interp-only plans take no smaps");
    const int TT = B * C;
    for (int t0 = 0; t0 < TT; t0 += p->ntrans) {
      const int Tc = std::min(p->ntrans, TT - t0);
      if ((rc = launch_interp(p, x + (int64_t)t0 * img_size(p), img_size(p), SetRange{0, 1, Tc},
                              y + (int64_t)t0 * p->M, nullptr, nullptr, scale, 0)))
        return rc;
    }
    return 0;
  }
  std::vector<Chunk> chunks;
  if ((rc = make_chunks(p, B, C, s != nullptr, chunks))) return rc;
  const int sall = p->nsets > 1;
  for (const Chunk& c : chunks) {
    if ((rc = fwd_to_grid(p, x, s, C, c.t0, c.T, 1.0f, p->sign2, 0, sall))) return rc;
    if ((rc = launch_interp(p, p->grid, grid_size(p), c.r, y + c.kofs, nullptr, nullptr, scale, 0)))
      return rc;
  }
  return 0;
}

int hipnufft_sense_adj(hipnufft_plan* p, const void* ksp, const void* smaps, const float* density,
                       void* img, int B, int C, float scale) {
  int rc = need_pts(p);
  if (rc) return rc;
  if (B < 1 || C < 1) return set_error(HIPNUFFT_ERR_ARG, "B and C must be >= 1");
  if (!img || !ksp) return set_error(HIPNUFFT_ERR_ARG, "null data pointer");
  const float2* y = (const float2*)ksp;
  const float2* s = (const float2*)smaps;
  float2* x = (float2*)img;
  if ((rc = ensure_ksorted(p))) return rc;
  if (p->sio) {
    if (s) return set_error(HIPNUFFT_ERR_ARG, "spread/interp-only plans take no smaps");
    if (scale != 1.0f) return set_error(HIPNUFFT_ERR_ARG, "spread/interp-only adjoint has no scale");
    const int TT = B * C;
    for (int t0 = 0; t0 < TT; t0 += p->ntrans) {
      const int Tc = std::min(p->ntrans, TT - t0);
      if ((rc = launch_spread(p, y + (int64_t)t0 * p->M, density, SetRange{0, 1, Tc},
                              x + (int64_t)t0 * img_size(p), img_size(p))))
        return rc;
    }
    return 0;
  }
  std::vector<Chunk> chunks;
  if ((rc = make_chunks(p, B, C, s != nullptr, chunks))) return rc;
  const int sall = p->nsets > 1;
  for (const Chunk& c : chunks) {
    if ((rc = launch_spread(p, y + c.kofs, density, c.r, p->grid, grid_size(p)))) return rc;
    if ((rc = grid_to_img(p, x, s, C, c.t0, c.T, scale, -p->sign2, sall))) return rc;
  }
  return 0;
}

int hipnufft_sense_dc(hipnufft_plan* p, const void* img, const void* smaps, const void* obs,
                      const float* density, void* grad, int B, int C, float scale) {
  int rc = need_pts(p);
  if (rc) return rc;
  if (B < 1 || C < 1) return set_error(HIPNUFFT_ERR_ARG, "B and C must be >= 1");
  if (!img || !obs || !grad) return set_error(HIPNUFFT_ERR_ARG, "null data pointer");
  if (p->sio) return set_error(HIPNUFFT_ERR_ARG, "data consistency needs a full NUFFT plan");
  if ((rc = ensure_kwork(p))) return rc;
  const float2* x = (const float2*)img;
  const float2* s = (const float2*)smaps;
  const float2* y = (const float2*)obs;
  float2* g = (float2*)grad;
  std::vector<Chunk> chunks;
  if ((rc = make_chunks(p, B, C, s != nullptr, chunks))) return rc;
  const int sall = p->nsets > 1;
  for (const Chunk& c : chunks) {
    if ((rc = fwd_to_grid(p, x, s, C, c.t0, c.T, 1.0f, p->sign2, 0, sall))) return rc;
    if ((rc = launch_interp(p, p->grid, grid_size(p), c.r, p->kwork, y + c.kofs, density, scale,
                            1)))
      return rc;
    if ((rc = launch_spread_sorted(p, p->kwork, c.r, p->grid, grid_size(p)))) return rc;
    if ((rc = grid_to_img(p, g, s, C, c.t0, c.T, scale, -p->sign2, sall))) return rc;
  }
  return 0;
}

}  // extern "C"

namespace {

// the checks and chunks shared by the two off-resonance entry points: the
// transforms are [image][coil][l], per_img = (smaps ? C : 1) * L of them per
// output image, chunks aligned to whole images when the workspace holds one
int orc_setup(hipnufft_plan_s* p, const void* cmap, const void* bt, int L, int64_t NK, int B,
              int C, bool sense, std::vector<Chunk>& chunks) {
  int rc = need_pts(p);
  if (rc) return rc;
  if (p->sio) return set_error(HIPNUFFT_ERR_ARG, "the off-resonance operator needs a full NUFFT plan");
  if (p->nsets != 1)
    return set_error(HIPNUFFT_ERR_ARG, "the off-resonance operator needs a plan with one point set");
  if (B < 1 || C < 1) return set_error(HIPNUFFT_ERR_ARG, "B and C must be >= 1");
  if (L < 1 || NK < 1) return set_error(HIPNUFFT_ERR_ARG, "L and NK must be >= 1");
  if (!cmap || !bt) return set_error(HIPNUFFT_ERR_ARG, "null interpolator pointer");
  if (p->M % NK != 0)
    return set_error(HIPNUFFT_ERR_ARG, "the samples per shot NK must divide the number of samples");
  if ((int64_t)B * C * L > (int64_t)1 << 30) return set_error(HIPNUFFT_ERR_ARG, "too many transforms");
  return make_chunks(p, sense ? B : B * C, (sense ? C : 1) * L, true, chunks);
}

}  // namespace

extern "C" {

int hipnufft_orc_op(hipnufft_plan* p, const void* img, const void* smaps, const void* cmap,
                    const void* bt, int L, int64_t NK, void* ksp, int B, int C, float scale) {
  if (!p) return set_error(HIPNUFFT_ERR_ARG, "null plan");
  if (!img || !ksp) return set_error(HIPNUFFT_ERR_ARG, "null data pointer");
  std::vector<Chunk> chunks;
  int rc = orc_setup(p, cmap, bt, L, NK, B, C, smaps != nullptr, chunks);
  if (rc) return rc;
  if ((rc = ensure_kwork(p))) return rc;
  const float2* x = (const float2*)img;
  const float2* s = (const float2*)smaps;
  const float2* cm = (const float2*)cmap;
  const float2* b = (const float2*)bt;
  for (const Chunk& c : chunks) {
    if ((rc = fwd_to_grid(p, x, s, C, c.t0, c.T, 1.0f, p->sign2, 0, 0, cm, L))) return rc;
    if ((rc = launch_interp(p, p->grid, grid_size(p), SetRange{0, 1, c.T}, p->kwork, nullptr,
                            nullptr, scale, 0)))
      return rc;
    if ((rc = orc_combine(p, p->kwork, b, L, NK, c.t0, c.T, (float2*)ksp))) return rc;
  }
  return 0;
}

int hipnufft_orc_adj(hipnufft_plan* p, const void* ksp, const void* smaps, const void* cmap,
                     const void* bt, int L, int64_t NK, const float* density, void* img, int B,
                     int C, float scale) {
  if (!p) return set_error(HIPNUFFT_ERR_ARG, "null plan");
  if (!img || !ksp) return set_error(HIPNUFFT_ERR_ARG, "null data pointer");
  std::vector<Chunk> chunks;
  int rc = orc_setup(p, cmap, bt, L, NK, B, C, smaps != nullptr, chunks);
  if (rc) return rc;
  if ((rc = ensure_ksorted(p))) return rc;
  const float2* s = (const float2*)smaps;
  const float2* cm = (const float2*)cmap;
  const float2* b = (const float2*)bt;
  for (const Chunk& c : chunks) {
    if ((rc = orc_gather(p, (const float2*)ksp, b, L, NK, density, c.t0, c.T, p->ksorted)))
      return rc;
    if ((rc = launch_spread_sorted(p, p->ksorted, SetRange{0, 1, c.T}, p->grid, grid_size(p))))
      return rc;
    if ((rc = grid_to_img(p, (float2*)img, s, C, c.t0, c.T, scale, -p->sign2, 0, cm, L)))
      return rc;
  }
  return 0;
}

int hipnufft_execute(hipnufft_plan* p, const void* in, void* out) {
  if (!p) return set_error(HIPNUFFT_ERR_ARG, "null plan");
  const int B = p->nsets > 1 ? p->nsets : p->ntrans;
  if (p->type == 2) return hipnufft_sense_op(p, in, nullptr, out, B, 1, 1.0f);
  return hipnufft_sense_adj(p, in, nullptr, nullptr, out, B, 1, 1.0f);
}

int hipnufft_execute_adjoint(hipnufft_plan* p, const void* in, void* out) {
  if (!p) return set_error(HIPNUFFT_ERR_ARG, "null plan");
  const int B = p->nsets > 1 ? p->nsets : p->ntrans;
  if (p->type == 2) return hipnufft_sense_adj(p, in, nullptr, nullptr, out, B, 1, 1.0f);
  return hipnufft_sense_op(p, in, nullptr, out, B, 1, 1.0f);
}

int hipnufft_sense_grad_traj(hipnufft_plan* p, const void* img, const void* smaps, const void* dy,
                             int B, int C, float scale, int iflag, int flags, float* gtraj) {
  int rc = need_pts(p);
  if (rc) return rc;
  if (B < 1 || C < 1) return set_error(HIPNUFFT_ERR_ARG, "B and C must be >= 1");
  if (!img || !dy || !gtraj) return set_error(HIPNUFFT_ERR_ARG, "null data pointer");
  if (p->sio) return set_error(HIPNUFFT_ERR_ARG, "trajectory gradient needs a full NUFFT plan");
  if (iflag != 1 && iflag != -1) return set_error(HIPNUFFT_ERR_ARG, "iflag must be +1 or -1");
  const float2* x = (const float2*)img;
  const float2* s = (const float2*)smaps;
  const float2* d = (const float2*)dy;
  std::vector<Chunk> chunks;
  if ((rc = make_chunks(p, B, C, s != nullptr, chunks))) return rc;
  const int64_t n = p->M * p->ndim;
  if (n > 0) {
    hipLaunchKernelGGL(k_zero_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->stream,
                       gtraj, n);
    HN_LAUNCH_CHECK();
  }
  const int sall = p->nsets > 1;
  for (const Chunk& c : chunks) {
    if ((rc = fwd_to_grid(p, x, s, C, c.t0, c.T, 1.0f, iflag, flags & 1, sall))) return rc;
    if ((rc = launch_grad_traj(p, p->grid, grid_size(p), c.r, d + c.kofs, scale, gtraj))) return rc;
  }
  return 0;
}

int hipnufft_dot_c64(const void* a, const void* b, int64_t n, int conj, double* out, double* ws,
                     void* stream) {
  if (!a || !b || !out) return set_error(HIPNUFFT_ERR_ARG, "null pointer");
  if (n < 0) return set_error(HIPNUFFT_ERR_ARG, "negative length");
  return dot_c64((const float2*)a, (const float2*)b, 1, n, conj, out, ws, (hipStream_t)stream);
}

int hipnufft_dot_c64_rows(const void* a, const void* b, int64_t rows, int64_t n, int conj,
                          double* out, double* ws, void* stream) {
  if (!a || !b || !out) return set_error(HIPNUFFT_ERR_ARG, "null pointer");
  if (n < 0 || rows < 0) return set_error(HIPNUFFT_ERR_ARG, "negative size");
  return dot_c64((const float2*)a, (const float2*)b, rows, n, conj, out, ws, (hipStream_t)stream);
}

int hipnufft_dot_c64_rows_noalloc(const void* a, const void* b, int64_t rows, int64_t n,
                                  int conj, double* out, void* stream) {
  if (!a || !b || !out) return set_error(HIPNUFFT_ERR_ARG, "null pointer");
  if (n < 0 || rows < 0) return set_error(HIPNUFFT_ERR_ARG, "negative size");
  if (rows > HIPNUFFT_NOALLOC_DOT_ROWS)
    return set_error(HIPNUFFT_ERR_ARG, "dot_c64_rows_noalloc: at most " +
                                           std::to_string(HIPNUFFT_NOALLOC_DOT_ROWS) + " rows");
  double2* ws = nullptr;
  int rc = noalloc_scratch((hipStream_t)stream, &ws);
  if (rc) return rc;
  return dot_c64((const float2*)a, (const float2*)b, rows, n, conj, out, (double*)ws,
                 (hipStream_t)stream);
}

int hipnufft_get_sort(const hipnufft_plan* p, int32_t* perm, int32_t* bin_start, int32_t* list,
                      int32_t* list_start, int64_t* sizes) {
  int rc = need_pts(p);
  if (rc) return rc;
  hipStream_t s = p->stream;
  const int64_t nkeys = p->nbins * p->nsets;
  if (perm && p->M)
    HN_HIP(hipMemcpyAsync(perm, p->pt_idx, sizeof(int32_t) * p->M, hipMemcpyDeviceToDevice, s));
  if (bin_start)
    HN_HIP(hipMemcpyAsync(bin_start, p->bin_start, sizeof(int32_t) * (nkeys + 1),
                          hipMemcpyDeviceToDevice, s));
  if (list && p->nlist)
    HN_HIP(hipMemcpyAsync(list, p->list_pts, sizeof(int32_t) * p->nlist, hipMemcpyDeviceToDevice, s));
  if (list_start)
    HN_HIP(hipMemcpyAsync(list_start, p->list_start, sizeof(int32_t) * (nkeys + 1),
                          hipMemcpyDeviceToDevice, s));
  if (sizes) {
    sizes[0] = p->M;
    sizes[1] = nkeys;
    sizes[2] = p->nlist;
  }
  return 0;
}

int hipnufft_plan_info(const hipnufft_plan* p, int64_t* nf, int* width, double* beta, int* tile,
                       int64_t* M) {
  if (!p) return set_error(HIPNUFFT_ERR_ARG, "null plan");
  for (int d = 0; d < 3; ++d) {
    if (nf) nf[d] = p->nf[d];
    if (tile) tile[d] = p->tile[d];
  }
  if (width) *width = p->w;
  if (beta) *beta = p->beta;
  if (M) *M = p->M;
  return 0;
}

int hipnufft_workspace_bytes(const hipnufft_plan* p, size_t* bytes) {
  if (!p || !bytes) return set_error(HIPNUFFT_ERR_ARG, "null pointer");
  size_t b = p->bytes;
  b += p->M_cap * (sizeof(int32_t) + p->ndim * (sizeof(int32_t) + sizeof(float)));
  b += (p->list_cap + 2 * p->sub_cap + 2 * p->keys_cap + p->set_off_cap) * sizeof(int32_t);
  b += (3 * p->items_cap + 3 * p->dense_cap) * sizeof(int32_t) + p->part_cap * sizeof(float2);
  *bytes = b;
  return 0;
}

int hipnufft_set_timing(hipnufft_plan* p, int enable) {
  if (!p) return set_error(HIPNUFFT_ERR_ARG, "null plan");
  p->timing = enable ? 1 : 0;
  return 0;
}

int hipnufft_get_timing(hipnufft_plan* p, double* ms, int64_t* launches, int nstage) {
  if (!p) return set_error(HIPNUFFT_ERR_ARG, "null plan");
  HN_HIP(hipStreamSynchronize(p->stream));
  for (int i = 0; i < nstage; ++i) {
    if (ms) ms[i] = 0.0;
    if (launches) launches[i] = 0;
  }
  for (auto& e : p->ev_used) {
    float t = 0.f;
    HN_HIP(hipEventElapsedTime(&t, e.start, e.stop));
    if (e.stage < nstage) {
      if (ms) ms[e.stage] += t;
      if (launches) launches[e.stage] += 1;
    }
    p->ev_free.push_back(e);
  }
  p->ev_used.clear();
  return 0;
}

int hipnufft_set_stream(hipnufft_plan* p, void* stream) {
  if (!p) return set_error(HIPNUFFT_ERR_ARG, "null plan");
  p->stream = (hipStream_t)stream;
  return 0;
}

int hipnufft_plan_destroy(hipnufft_plan* p) {
  if (!p) return 0;
  (void)hipStreamSynchronize(p->stream);
  free_plan(p);
  return 0;
}

int hipnufft_gram_create(int ndim, const int64_t* nmodes, int64_t M, const float* x0,
                         const float* x1, const float* x2, const float* weights, double eps,
                         int iflag, float scale, int ntrans, void* stream, hipnufft_gram** out) {
  g_last_error.clear();
  if (!out) return set_error(HIPNUFFT_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (ndim < 1 || ndim > 3 || !nmodes) return set_error(HIPNUFFT_ERR_ARG, "ndim must be 1, 2 or 3");
  if (ntrans < 1) return set_error(HIPNUFFT_ERR_ARG, "ntrans must be >= 1");
  if (M < 0) return set_error(HIPNUFFT_ERR_ARG, "negative number of points");
  hipStream_t s = (hipStream_t)stream;
  int64_t n2[3], G2 = 1;
  for (int d = 0; d < ndim; ++d) {
    if (nmodes[d] < 1) return set_error(HIPNUFFT_ERR_ARG, "invalid mode count");
    n2[d] = 2 * nmodes[d];
    if (next_smooth_even(n2[d]) != n2[d])
      return set_error(HIPNUFFT_ERR_ARG, "Toeplitz operator needs 2N to be 2,3,5-smooth on every axis");
    G2 *= n2[d];
  }
  // 1. kernel kappa(l) = sum_m w_m exp(-iflag i omega_m . l), lags l in [-N, N):
  //    one type-1 NUFFT of the weights onto the 2N image grid
  hipnufft_plan_s* nu = nullptr;
  int rc = hipnufft_plan_create(2, ndim, n2, iflag, 1, eps, nullptr, stream, &nu);
  if (rc) return rc;
  float2 *w = nullptr, *kappa = nullptr;
  rc = hipnufft_setpts(nu, M, x0, x1, x2);
  if (!rc && (hipMalloc((void**)&w, sizeof(float2) * std::max<int64_t>(M, 1)) != hipSuccess ||
              hipMalloc((void**)&kappa, sizeof(float2) * G2) != hipSuccess))
    rc = set_error(HIPNUFFT_ERR_ALLOC, "cannot allocate the Toeplitz kernel");
  if (!rc && M > 0) {
    hipLaunchKernelGGL(k_cplx_weights, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, s, weights,
                       M, w);
    if (hipGetLastError() != hipSuccess) rc = set_error(HIPNUFFT_ERR_HIP, "kernel launch failed");
  }
  if (!rc) rc = hipnufft_sense_adj(nu, w, nullptr, nullptr, kappa, 1, 1, 1.0f);
  // 2. spectrum of the circulant: the centred placement of the 2N FFT plan
  //    puts lag l at position l mod 2N, so its forward chain gives
  //    K = FFT(circulant) directly; the inverse's 1/prod(2N) and the
  //    operator scale are folded in
  hipnufft_plan_s* kp = nullptr;
  if (!rc) rc = create_fft_plan(ndim, n2, n2, 1, s, &kp);
  if (!rc) rc = fwd_to_grid(kp, kappa, nullptr, 1, 0, 1, (float)((double)scale / (double)G2), -1);
  hipnufft_gram_s* g = nullptr;
  if (!rc) {
    g = new (std::nothrow) hipnufft_gram_s();
    if (!g) rc = set_error(HIPNUFFT_ERR_ALLOC, "host allocation failed");
  }
  if (!rc) {
    g->K = kp->grid;  // take the spectrum
    kp->grid = nullptr;
    g->bytes = sizeof(float2) * G2;
    rc = create_fft_plan(ndim, nmodes, n2, ntrans, s, &g->fft);
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(w);
  (void)hipFree(kappa);
  if (kp) free_plan(kp);
  if (nu) free_plan(nu);
  if (rc) {
    if (g) {
      (void)hipFree(g->K);
      if (g->fft) free_plan(g->fft);
      delete g;
    }
    return rc;
  }
  g->bytes += g->fft->bytes;
  *out = g;
  return 0;
}

int hipnufft_gram_apply(hipnufft_gram* g, const void* img, const void* smaps, void* out, int B,
                        int C) {
  if (!g || !g->fft) return set_error(HIPNUFFT_ERR_ARG, "null Toeplitz operator");
  if (B < 1 || C < 1) return set_error(HIPNUFFT_ERR_ARG, "B and C must be >= 1");
  if (!img || !out) return set_error(HIPNUFFT_ERR_ARG, "null data pointer");
  hipnufft_plan_s* p = g->fft;
  const float2* x = (const float2*)img;
  const float2* sm = (const float2*)smaps;
  std::vector<Chunk> chunks;
  int rc = make_chunks(p, B, C, sm != nullptr, chunks);
  if (rc) return rc;
  for (const Chunk& c : chunks)
    if ((rc = gram_chain(p, x, sm, (float2*)out, g->K, C, c.t0, c.T))) return rc;
  return 0;
}

int hipnufft_gram_set_stream(hipnufft_gram* g, void* stream) {
  if (!g || !g->fft) return set_error(HIPNUFFT_ERR_ARG, "null Toeplitz operator");
  g->fft->stream = (hipStream_t)stream;
  return 0;
}

int hipnufft_gram_workspace_bytes(const hipnufft_gram* g, size_t* bytes) {
  if (!g || !bytes) return set_error(HIPNUFFT_ERR_ARG, "null pointer");
  *bytes = g->bytes;
  return 0;
}

int hipnufft_stack_create(int64_t X, int64_t Y, int64_t Z, const int32_t* z_index, int NZ,
                          void* stream, hipnufft_stack** out) {
  if (!out || (NZ > 0 && !z_index)) return set_error(HIPNUFFT_ERR_ARG, "null pointer");
  *out = nullptr;
  if (X < 1 || Y < 1 || Z < 1 || NZ < 0 || X * Y >= ((int64_t)1 << 31))
    return set_error(HIPNUFFT_ERR_ARG, "invalid stack shape");
  std::vector<int32_t> zp(Z, -1);
  for (int i = 0; i < NZ; ++i) {
    if (z_index[i] < 0 || z_index[i] >= Z) return set_error(HIPNUFFT_ERR_ARG, "z_index out of range");
    zp[z_index[i]] = i;  // a repeated position keeps its last plane
  }
  auto* st = new hipnufft_stack_s();
  st->stream = (hipStream_t)stream;
  st->g.rows = X * Y;
  st->g.Z = (int)Z;
  st->g.NZ = NZ;
  size_t bytes = 0;
  int rc = make_axis(Z, &st->g.ax, &bytes, true);
  if (!rc && NZ > 0 && hipMalloc((void**)&st->z_index, sizeof(int32_t) * NZ) != hipSuccess)
    rc = set_error(HIPNUFFT_ERR_ALLOC, "stack: out of device memory");
  if (!rc && hipMalloc((void**)&st->z_plane, sizeof(int32_t) * Z) != hipSuccess)
    rc = set_error(HIPNUFFT_ERR_ALLOC, "stack: out of device memory");
  if (!rc && NZ > 0 &&
      hipMemcpy(st->z_index, z_index, sizeof(int32_t) * NZ, hipMemcpyHostToDevice) != hipSuccess)
    rc = set_error(HIPNUFFT_ERR_HIP, "stack: copy failed");
  if (!rc && hipMemcpy(st->z_plane, zp.data(), sizeof(int32_t) * Z, hipMemcpyHostToDevice) !=
                 hipSuccess)
    rc = set_error(HIPNUFFT_ERR_HIP, "stack: copy failed");
  if (rc) {
    hipnufft_stack_destroy(st);
    return rc;
  }
  st->g.z_index = st->z_index;
  st->g.z_plane = st->z_plane;
  *out = st;
  return 0;
}

int hipnufft_stack_fwd(hipnufft_stack* st, const void* img_d, const void* smaps_d, int C, int t0,
                       int T, float scale, void* planes_d) {
  if (!st || !img_d || !planes_d) return set_error(HIPNUFFT_ERR_ARG, "null pointer");
  if (C < 1 || t0 < 0 || T < 0) return set_error(HIPNUFFT_ERR_ARG, "invalid chunk");
  if (st->g.NZ == 0) return 0;
  return stack_fwd(st->g, (const float2*)img_d, (const float2*)smaps_d, C, t0, T, scale,
                   (float2*)planes_d, st->stream);
}

int hipnufft_stack_adj(hipnufft_stack* st, const void* planes_d, const void* smaps_d, int C,
                       int t0, int T, float scale, void* img_d) {
  if (!st || !img_d || (!planes_d && st->g.NZ > 0))
    return set_error(HIPNUFFT_ERR_ARG, "null pointer");
  if (C < 1 || t0 < 0 || T < 0) return set_error(HIPNUFFT_ERR_ARG, "invalid chunk");
  return stack_adj(st->g, (const float2*)planes_d, (const float2*)smaps_d, C, t0, T, scale,
                   (float2*)img_d, st->stream);
}

int hipnufft_stack_set_stream(hipnufft_stack* st, void* stream) {
  if (!st) return set_error(HIPNUFFT_ERR_ARG, "null stack");
  st->stream = (hipStream_t)stream;
  return 0;
}

int hipnufft_stack_destroy(hipnufft_stack* st) {
  if (!st) return 0;
  (void)hipStreamSynchronize(st->stream);
  (void)hipFree(st->g.ax.tw);
  (void)hipFree(st->z_index);
  (void)hipFree(st->z_plane);
  delete st;
  return 0;
}

int hipnufft_kspace_window(const float* traj_d, int64_t M, int d, const float* threshold,
                           int kind, const void* ksp_in_d, void* ksp_out_d, int C,
                           float* weights_d, void* stream) {
  if (!traj_d || !threshold || (ksp_in_d && !ksp_out_d))
    return set_error(HIPNUFFT_ERR_ARG, "null pointer");
  if (M < 0 || d < 1 || d > 3 || C < 0 || kind < HIPNUFFT_WINDOW_ELLIPSE ||
      kind > HIPNUFFT_WINDOW_HAMMING)
    return set_error(HIPNUFFT_ERR_ARG, "invalid window arguments");
  for (int q = 0; q < d; ++q)
    if (!(threshold[q] > 0.0f)) return set_error(HIPNUFFT_ERR_ARG, "window threshold must be > 0");
  return kspace_window(traj_d, M, d, threshold, kind, (const float2*)ksp_in_d,
                       (float2*)ksp_out_d, C, weights_d, (hipStream_t)stream);
}

int hipnufft_rss_normalize(void* maps_d, int C, int64_t N, float eps, void* stream) {
  if (!maps_d) return set_error(HIPNUFFT_ERR_ARG, "null pointer");
  if (C < 1 || N < 0) return set_error(HIPNUFFT_ERR_ARG, "invalid map shape");
  return rss_normalize((float2*)maps_d, C, N, eps, (hipStream_t)stream);
}

int hipnufft_blur_magnitude(void* maps_d, int C, int ndim, const int64_t* shape,
                            const double* sigma, double truncate, float* work_d, void* stream) {
  if (!maps_d || !shape || !sigma || !work_d) return set_error(HIPNUFFT_ERR_ARG, "null pointer");
  if (C < 1 || ndim < 1 || ndim > 3 || !(truncate > 0.0))
    return set_error(HIPNUFFT_ERR_ARG, "invalid blur arguments");
  for (int q = 0; q < ndim; ++q)
    if (shape[q] < 1 || sigma[q] < 0.0) return set_error(HIPNUFFT_ERR_ARG, "invalid blur arguments");
  return blur_magnitude((float2*)maps_d, C, ndim, shape, sigma, truncate, work_d,
                        (hipStream_t)stream);
}

int hipnufft_gram_destroy(hipnufft_gram* g) {
  if (!g) return 0;
  if (g->fft) {
    (void)hipStreamSynchronize(g->fft->stream);
    free_plan(g->fft);
  }
  (void)hipFree(g->K);
  delete g;
  return 0;
}

}  // extern "C"
