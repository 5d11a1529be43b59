// Off-resonance corrected NUFFT: the k-space side of
//   y[c] = sum_l B[:, l] * F(S_c C_l x)          (operators/off_resonance.py:230-262)
//   x    = sum_l conj(C_l) sum_c conj(S_c) F^H(conj(B[:, l]) * y[c])   (:264-302)
// The L spatial interpolators C_l are L extra transforms per coil of one
// batched SENSE chain (ordered [image][coil][l], see fwd_to_grid); these two
// kernels apply the time interpolators B[NK][L] (sample m uses row m % NK,
// NK samples per shot) on the way out of the forward interpolation and on
// the way into the adjoint spreading.  Both are one pass over HBM.

#include <hip/hip_runtime.h>

#include "internal.h"

namespace hn {
namespace {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// y[q][m] = (l range starts at 0 ? 0 : y[q][m]) + sum_l B[m % NK][l] kv[q L + l - t0][m]
// for the output rows q (= image * C + coil) present in transforms [t0, t0 + T)
__global__ __launch_bounds__(kThreads) void k_orc_combine(const float2* kv, const float2* bt,
                                                         int L, int64_t NK, int64_t M, int t0,
                                                         int T, int q0, float2* y) {
  const int64_t m = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (m >= M) return;
  const int q = q0 + (int)blockIdx.y;
  const int lb = max(t0, q * L) - q * L;
  const int le = min(t0 + T, (q + 1) * L) - q * L;
  const float2* b = bt + (m % NK) * L;
  float2 s = make_float2(0.f, 0.f);
  for (int l = lb; l < le; ++l) s = cadd(s, cmul(b[l], kv[(int64_t)(q * L + l - t0) * M + m]));
  float2* dst = y + (int64_t)q * M + m;
  if (lb > 0) s = cadd(*dst, s);
  *dst = s;
}

// out[t][j] = y[tt / L][o] * conj(B[o % NK][tt % L]) * density[o] with
// o = idx[j] and tt = t0 + t: the spreading input in sorted-point order
__global__ __launch_bounds__(kThreads) void k_orc_gather(const int32_t* idx, int64_t M,
                                                        const float2* y, const float2* bt, int L,
                                                        int64_t NK, const float* density, int t0,
                                                        float2* out) {
  const int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (j >= M) return;
  const int tt = t0 + (int)blockIdx.y;
  const int32_t o = idx[j];
  float2 v = y[(int64_t)(tt / L) * M + o];
  float2 b = bt[(o % NK) * L + tt % L];
  b.y = -b.y;
  v = cmul(v, b);
  if (density) v = cscale(v, density[o]);
  out[(int64_t)blockIdx.y * M + j] = v;
}

unsigned blocks_of(int64_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

}  // namespace

int orc_combine(const hipnufft_plan_s* p, const float2* kv, const float2* bt, int L, int64_t NK,
                int t0, int T, float2* y) {
  if (T <= 0 || p->M <= 0) return 0;
  const int q0 = t0 / L, q1 = (t0 + T - 1) / L;
  hipLaunchKernelGGL(k_orc_combine, dim3(blocks_of(p->M), (unsigned)(q1 - q0 + 1)),
                     dim3(kThreads), 0, p->stream, kv, bt, L, NK, p->M, t0, T, q0, y);
  HN_LAUNCH_CHECK();
  return 0;
}

int orc_gather(const hipnufft_plan_s* p, const float2* y, const float2* bt, int L, int64_t NK,
               const float* density, int t0, int T, float2* out) {
  if (T <= 0 || p->M <= 0) return 0;
  hipLaunchKernelGGL(k_orc_gather, dim3(blocks_of(p->M), (unsigned)T), dim3(kThreads), 0,
                     p->stream, p->pt_idx, p->M, y, bt, L, NK, density, t0, out);
  HN_LAUNCH_CHECK();
  return 0;
}

}  // namespace hn
