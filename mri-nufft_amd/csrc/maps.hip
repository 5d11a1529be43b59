// Coil-map estimation helpers (the low-frequency method): k-space windows,
// root-sum-of-squares normalisation and a separable Gaussian blur of the
// map magnitudes.  All are elementwise / stencil passes over HBM.
//
// Replaces the array work of extras/smaps.py: _extract_kspace_center
// (:83-170, the ellipse / Hann / Hamming / rect windows), the skimage
// Gaussian of |maps| (:293-298) and the SOS normalisation (:300-306).

#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "internal.h"

namespace hn {
namespace {

// weight of sample m: traj[m][d] in [-0.5, 0.5), thr[d] > 0
__device__ __forceinline__ float window_weight(const float* k, int d, const float* thr, int kind) {
  if (kind == kWindowEllipse) {
    float s = 0.f;
    for (int a = 0; a < d; ++a) s += (k[a] * k[a]) / (thr[a] * thr[a]);
    return s <= 1.0f ? 1.0f : 0.0f;
  }
  if (kind == kWindowRect) {
    bool in = true;
    for (int a = 0; a < d; ++a) in = in && fabsf(k[a]) - thr[a] <= 0.0f;
    return in ? 1.0f : 0.0f;
  }
  // Hann (a0 = 0.5) or Hamming (a0 = 0.53836) of the radius, as a cosine of
  // pi r / thr[0] for every sample
  float r2 = 0.f;
  for (int a = 0; a < d; ++a) r2 += k[a] * k[a];
  const float a0 = kind == kWindowHann ? 0.5f : 0.53836f;
  return a0 + (1.0f - a0) * cosf(3.14159265358979f * sqrtf(r2) / thr[0]);
}

struct WindowArgs {
  const float* traj;
  int64_t M;
  int d;
  float thr[3];
  int kind;
  const float2* in;
  float2* out;
  float* weights;
  int C;
};

__global__ __launch_bounds__(kThreads) void k_window(WindowArgs a) {
  const int64_t m = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (m >= a.M) return;
  float k[3];
  for (int q = 0; q < a.d; ++q) k[q] = a.traj[m * a.d + q];
  const float w = window_weight(k, a.d, a.thr, a.kind);
  if (a.weights) a.weights[m] = w;
  if (a.in) {
    for (int c = 0; c < a.C; ++c) {
      const float2 v = a.in[(int64_t)c * a.M + m];
      a.out[(int64_t)c * a.M + m] = make_float2(w * v.x, w * v.y);
    }
  }
}

// maps[c][n] /= sqrt(sum_c |maps[c][n]|^2) + eps (the sum in FP64)
__global__ __launch_bounds__(kThreads) void k_rss_normalize(float2* maps, int C, int64_t N,
                                                           float eps) {
  const int64_t n = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (n >= N) return;
  double s = 0.0;
  for (int c = 0; c < C; ++c) {
    const float2 v = maps[(int64_t)c * N + n];
    s += (double)v.x * v.x + (double)v.y * v.y;
  }
  const float r = (float)sqrt(s) + eps;
  for (int c = 0; c < C; ++c) {
    float2 v = maps[(int64_t)c * N + n];
    v.x /= r;
    v.y /= r;
    maps[(int64_t)c * N + n] = v;
  }
}

__global__ __launch_bounds__(kThreads) void k_magnitude(const float2* maps, float* mag, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const float2 v = maps[i];
  mag[i] = sqrtf(v.x * v.x + v.y * v.y);
}

// out[.., i, ..] = sum_j w[j] in[.., clamp(i + j - r), ..] along one axis of
// a [outer][len][inner] view ("nearest" edges)
__global__ __launch_bounds__(kThreads) void k_blur_axis(const float* in, float* out, int64_t outer,
                                                       int len, int64_t inner, const float* w,
                                                       int r) {
  const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int64_t total = outer * len * inner;
  if (e >= total) return;
  const int64_t o = e / (len * inner);
  const int64_t rem = e - o * len * inner;
  const int i = (int)(rem / inner);
  const int64_t q = rem - (int64_t)i * inner;
  const float* line = in + o * len * inner + q;
  float s = 0.f;
  for (int j = -r; j <= r; ++j) {
    int x = i + j;
    x = x < 0 ? 0 : (x >= len ? len - 1 : x);
    s += w[j + r] * line[(int64_t)x * inner];
  }
  out[e] = s;
}

// maps <- mag * maps / |maps| (phase kept; a zero map takes phase 0)
__global__ __launch_bounds__(kThreads) void k_apply_magnitude(float2* maps, const float* mag,
                                                             int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const float2 v = maps[i];
  const float a = sqrtf(v.x * v.x + v.y * v.y);
  maps[i] = a > 0.f ? make_float2(mag[i] * v.x / a, mag[i] * v.y / a) : make_float2(mag[i], 0.f);
}

unsigned blocks_of(int64_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

}  // namespace

int kspace_window(const float* traj, int64_t M, int d, const float* thr, int kind,
                  const float2* in, float2* out, int C, float* weights, hipStream_t s) {
  if (M <= 0) return 0;
  WindowArgs a{};
  a.traj = traj;
  a.M = M;
  a.d = d;
  for (int q = 0; q < d; ++q) a.thr[q] = thr[q];
  a.kind = kind;
  a.in = in;
  a.out = out;
  a.weights = weights;
  a.C = C;
  hipLaunchKernelGGL(k_window, dim3(blocks_of(M)), dim3(kThreads), 0, s, a);
  HN_LAUNCH_CHECK();
  return 0;
}

int rss_normalize(float2* maps, int C, int64_t N, float eps, hipStream_t s) {
  if (N <= 0) return 0;
  hipLaunchKernelGGL(k_rss_normalize, dim3(blocks_of(N)), dim3(kThreads), 0, s, maps, C, N, eps);
  HN_LAUNCH_CHECK();
  return 0;
}

int blur_magnitude(float2* maps, int C, int ndim, const int64_t* shape, const double* sigma,
                   double truncate, float* work, hipStream_t s) {
  int64_t N = 1;
  for (int q = 0; q < ndim; ++q) N *= shape[q];
  const int64_t n = (int64_t)C * N;
  if (n <= 0) return 0;
  float* mag = work;
  float* tmp = work + n;
  hipLaunchKernelGGL(k_magnitude, dim3(blocks_of(n)), dim3(kThreads), 0, s, maps, mag, n);
  HN_LAUNCH_CHECK();
  int64_t inner = N;
  for (int q = 0; q < ndim; ++q) {
    const int len = (int)shape[q];
    inner /= len;
    if (sigma[q] <= 0.0) continue;
    // kernel of scipy.ndimage.gaussian_filter1d: radius int(truncate * sd + 0.5)
    const int r = (int)(truncate * sigma[q] + 0.5);
    std::vector<float> w(2 * r + 1);
    double tot = 0.0;
    std::vector<double> wd(2 * r + 1);
    for (int j = -r; j <= r; ++j) tot += (wd[j + r] = std::exp(-0.5 * j * j / (sigma[q] * sigma[q])));
    for (int j = 0; j <= 2 * r; ++j) w[j] = (float)(wd[j] / tot);
    float* wdev = nullptr;
    HN_HIP(hipMallocAsync((void**)&wdev, sizeof(float) * w.size(), s));
    HN_HIP(hipMemcpyAsync(wdev, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice, s));
    const int64_t outer = (int64_t)C * (N / (inner * len));
    hipLaunchKernelGGL(k_blur_axis, dim3(blocks_of(n)), dim3(kThreads), 0, s, mag, tmp, outer,
                       len, inner, wdev, r);
    HN_LAUNCH_CHECK();
    HN_HIP(hipStreamSynchronize(s));  // the host weights must outlive the copy
    HN_HIP(hipFreeAsync(wdev, s));
    std::swap(mag, tmp);
  }
  hipLaunchKernelGGL(k_apply_magnitude, dim3(blocks_of(n)), dim3(kThreads), 0, s, maps, mag, n);
  HN_LAUNCH_CHECK();
  return 0;
}

}  // namespace hn
