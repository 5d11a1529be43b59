// Batched Stockham FFTs on the oversampled grid, fused with deconvolution,
// zero-padding / cropping and the SENSE coil expand / conjugate combine.
//
// Replaces the FFT, deapodisation, pad and crop stages inside cufinufft's
// Plan.execute (operators/interfaces/cufinufft.py:127-133), the SENSE
// broadcast multiply (cufinufft.py:497) and the coil-combine kernel
// (operators/gpu_utils.py:12-24, operators/toeplitz.py:299-321).
//
// Layout: the fine grid of transform t is grid[t][nf0][nf1][nf2] (C order,
// axis 2 contiguous; 2D / 1D grids use leading axes of size 1).  Image index
// i along an axis of size M is frequency k = i - M/2 and lives at fine-grid
// position (k mod nf).  Every pass works in place on whole lines held in LDS:
// a workgroup owns L lines (L adjacent columns for the strided axes, so
// global accesses are L*8-byte contiguous segments), loads only the positions
// that are non-zero (pruned input), runs all Stockham passes in LDS and
// stores only the positions later stages read (pruned output).

#include <hip/hip_runtime.h>

#include "internal.h"

namespace hn {
namespace {


__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// a * (S i)
template <int S>
__device__ __forceinline__ float2 mul_si(float2 a) {
  return S > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// y_k = sum_j v_j exp(S 2 pi i j k / R)
template <int S>
__device__ __forceinline__ void dft2(float2* v) {
  const float2 t = v[0];
  v[0] = cadd(t, v[1]);
  v[1] = csub(t, v[1]);
}

template <int S>
__device__ __forceinline__ void dft4(float2& v0, float2& v1, float2& v2, float2& v3) {
  const float2 a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3);
  const float2 a3 = mul_si<S>(csub(v1, v3));
  v0 = cadd(a0, a2);
  v2 = csub(a0, a2);
  v1 = cadd(a1, a3);
  v3 = csub(a1, a3);
}

template <int S>
__device__ __forceinline__ void dft4(float2* v) {
  dft4<S>(v[0], v[1], v[2], v[3]);
}

template <int S>
__device__ __forceinline__ void dft8(float2* v) {
  float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<S>(e0, e1, e2, e3);
  dft4<S>(o0, o1, o2, o3);
  const float h = 0.70710678118654752f;
  const float2 w1 = cscale(cadd(o1, mul_si<S>(o1)), h);   // W8^1 o1
  const float2 w2 = mul_si<S>(o2);                        // W8^2 o2
  const float2 w3 = cscale(csub(mul_si<S>(o3), o3), h);   // W8^3 o3
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, w1);
  v[5] = csub(e1, w1);
  v[2] = cadd(e2, w2);
  v[6] = csub(e2, w2);
  v[3] = cadd(e3, w3);
  v[7] = csub(e3, w3);
}

template <int S>
__device__ __forceinline__ void dft3(float2* v) {
  const float c = 0.86602540378443865f;
  const float2 t = cadd(v[1], v[2]);
  const float2 s = cscale(mul_si<S>(csub(v[1], v[2])), c);
  const float2 m = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
  v[0] = cadd(v[0], t);
  v[1] = cadd(m, s);
  v[2] = csub(m, s);
}

template <int S>
__device__ __forceinline__ void dft5(float2* v) {
  const float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;
  const float s1 = 0.95105651629515357f, s2 = 0.58778525229247313f;
  const float2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
  const float2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
  const float2 x0 = v[0];
  const float2 p1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
  const float2 p2 = make_float2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
  const float2 q1 = mul_si<S>(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
  const float2 q2 = mul_si<S>(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
  v[0] = make_float2(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
  v[1] = cadd(p1, q1);
  v[4] = csub(p1, q1);
  v[2] = cadd(p2, q2);
  v[3] = csub(p2, q2);
}

// 6 = 2 x 3 by the prime-factor mapping (no inner twiddles): inputs
// (3 n1 + 2 n2) mod 6 go through two 3-point DFTs over n2, then 2-point
// DFTs over n1, and (k1, k2) lands at (3 k1 + 4 k2) mod 6
template <int S>
__device__ __forceinline__ void dft6(float2* v) {
  float2 a[3] = {v[0], v[2], v[4]};
  float2 b[3] = {v[3], v[5], v[1]};
  dft3<S>(a);
  dft3<S>(b);
  v[0] = cadd(a[0], b[0]);
  v[3] = csub(a[0], b[0]);
  v[4] = cadd(a[1], b[1]);
  v[1] = csub(a[1], b[1]);
  v[2] = cadd(a[2], b[2]);
  v[5] = csub(a[2], b[2]);
}

template <int R, int S>
__device__ __forceinline__ void dftR(float2* v) {
  if constexpr (R == 2) dft2<S>(v);
  if constexpr (R == 3) dft3<S>(v);
  if constexpr (R == 4) dft4<S>(v);
  if constexpr (R == 5) dft5<S>(v);
  if constexpr (R == 6) dft6<S>(v);
  if constexpr (R == 8) dft8<S>(v);
}

// One Stockham radix-R pass over L interleaved lines of length n held in LDS
// at buf[pos * LP + line].  Every thread first reads all its butterflies into
// registers, then (after a barrier) writes the results: a single LDS buffer.
template <int R, int S, int EPT, int NT = kThreads>
__device__ __forceinline__ void stockham_pass(float2* buf, int n, int lgL, int LP, int Ns,
                                              uint32_t ns_magic, const float2* tw) {
  constexpr int MAXB = (EPT + R - 1) / R;
  const int L = 1 << lgL;
  // uniform (scalar) quantities; every LDS address below is one 24-bit
  // multiply per butterfly plus uniform strides per radix leg
  const int nr = n / R;
  const int nbf = nr << lgL;
  const int tstep = n / (Ns * R);
  const int in_stride = nr * LP;
  const int out_stride = Ns * LP;
  const int NsR = Ns * R;
  float2 v[MAXB][R];
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    const int b = threadIdx.x + i * NT;
    if (b < nbf) {
      const int l = b & (L - 1), j = b >> lgL;
      const int a0 = __mul24(j, LP) + l;
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = buf[a0 + r * in_stride];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    const int b = threadIdx.x + i * NT;
    if (b < nbf) {
      const int l = b & (L - 1), j = b >> lgL;
      // q = j / Ns by multiply-high (exact for j < 2^16); Ns = 1 in the first pass
      const int q = Ns == 1 ? j : (int)__umulhi((uint32_t)j, ns_magic);
      const int k = j - __mul24(q, Ns);
      if (Ns > 1) {  // the first pass has unit twiddles
        const int kt = __mul24(k, tstep);
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[r * kt];
          if (S > 0) w.y = -w.y;
          v[i][r] = cmul(v[i][r], w);
        }
      }
      dftR<R, S>(v[i]);
      const int o0 = __mul24(__mul24(q, NsR) + k, LP) + l;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[o0 + r * out_stride] = v[i][r];
    }
  }
  __syncthreads();
}

// The same Stockham pass with the line length N, the lines 2^LGL and the
// pass's Ns known at compile time: strides fold into the LDS instruction
// offsets, j / Ns is a shift for the power-of-two Ns of an 8 * 8 * R plan,
// and the butterfly-range test disappears when NT divides the butterflies.
template <int R, int S, int EPT, int NT, int N, int LGL, int NS>
__device__ __forceinline__ void stockham_pass_ct(float2* buf, const float2* tw) {
  constexpr int MAXB = (EPT + R - 1) / R;
  constexpr int L = 1 << LGL, LP = L + 1;
  constexpr int nr = N / R;
  constexpr int nbf = nr << LGL;
  constexpr int tstep = N / (NS * R);
  constexpr int in_stride = nr * LP;
  constexpr int out_stride = NS * LP;
  constexpr int NsR = NS * R;
  static_assert(nbf <= MAXB * NT, "butterflies exceed the threads' slots");
  float2 v[MAXB][R];
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    const int b = threadIdx.x + i * NT;
    if (nbf % NT == 0 || b < nbf) {
      const int l = b & (L - 1), j = b >> LGL;
      const int a0 = j * LP + l;
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = buf[a0 + r * in_stride];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < MAXB; ++i) {
    const int b = threadIdx.x + i * NT;
    if (nbf % NT == 0 || b < nbf) {
      const int l = b & (L - 1), j = b >> LGL;
      const int q = j / NS;
      const int k = j - q * NS;
      if constexpr (NS > 1) {  // the first pass has unit twiddles
        const int kt = k * tstep;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[r * kt];
          if (S > 0) w.y = -w.y;
          v[i][r] = cmul(v[i][r], w);
        }
      }
      dftR<R, S>(v[i]);
      const int o0 = (q * NsR + k) * LP + l;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[o0 + r * out_stride] = v[i][r];
    }
  }
  __syncthreads();
}

// The p-th radix of the plan of an n-point line: 8s, then at most one each of
// 4, 6 and 2, then 3s and 5s (0 past the end) -- the rule of the host's axis
// set-up, so that a compile-time plan can be checked against the runtime one
__host__ __device__ constexpr int plan_radix(int n, int p) {
  int rad[16] = {};
  int k = 0, m = n;
  while (m % 8 == 0 && k < 16) { rad[k++] = 8; m /= 8; }
  if (m % 4 == 0) { rad[k++] = 4; m /= 4; }
  if (m % 6 == 0) { rad[k++] = 6; m /= 6; }
  if (m % 2 == 0) { rad[k++] = 2; m /= 2; }
  while (m % 3 == 0 && k < 16) { rad[k++] = 3; m /= 3; }
  while (m % 5 == 0 && k < 16) { rad[k++] = 5; m /= 5; }
  return p < k ? rad[p] : 0;
}

__host__ __device__ constexpr int plan_ns(int n, int p) {
  int ns = 1;
  for (int i = 0; i < p; ++i) ns *= plan_radix(n, i);
  return ns;
}

template <int S, int EPT, int NT, int N, int LGL, int P>
__device__ __forceinline__ void fft_lines_ct_from(float2* buf, const float2* twl) {
  constexpr int R = plan_radix(N, P);
  if constexpr (R != 0) {
    stockham_pass_ct<R, S, EPT, NT, N, LGL, plan_ns(N, P)>(buf, twl);
    fft_lines_ct_from<S, EPT, NT, N, LGL, P + 1>(buf, twl);
  }
}

// All passes of N-point lines on 2^LGL lines, every size fixed at compile time
template <int S, int EPT, int NT, int N, int LGL>
__device__ __forceinline__ void fft_lines_ct(float2* buf, const float2* twl) {
  fft_lines_ct_from<S, EPT, NT, N, LGL, 0>(buf, twl);
}

// All passes of the FFT of the L lines in buf; twl is the twiddle table
// staged in LDS by stage_twiddles (visible after the caller's barrier).
template <int S, int EPT, int NT = kThreads>
__device__ void fft_lines(float2* buf, const float2* twl, const FFTAxis& ax, int lgL, int LP) {
  int Ns = 1;
  for (int p = 0; p < ax.nrad; ++p) {
    const int R = ax.rad[p];
    const uint32_t mg = ax.ns_magic[p];
    switch (R) {
      case 2: stockham_pass<2, S, EPT, NT>(buf, ax.n, lgL, LP, Ns, mg, twl); break;
      case 3: stockham_pass<3, S, EPT, NT>(buf, ax.n, lgL, LP, Ns, mg, twl); break;
      case 4: stockham_pass<4, S, EPT, NT>(buf, ax.n, lgL, LP, Ns, mg, twl); break;
      case 5: stockham_pass<5, S, EPT, NT>(buf, ax.n, lgL, LP, Ns, mg, twl); break;
      case 6: stockham_pass<6, S, EPT, NT>(buf, ax.n, lgL, LP, Ns, mg, twl); break;
      default: stockham_pass<8, S, EPT, NT>(buf, ax.n, lgL, LP, Ns, mg, twl); break;
    }
    Ns *= R;
  }
}

// twiddles go after the n * LP data elements of the dynamic LDS buffer
template <int NT = kThreads>
__device__ __forceinline__ float2* stage_twiddles(float2* buf, const FFTAxis& ax, int LP) {
  float2* twl = buf + ax.n * LP;
  for (int i = threadIdx.x; i < ax.n; i += NT) twl[i] = ax.tw[i];
  return twl;
}

__device__ __forceinline__ int kept_pos(int i, int M, int nf) {
  int p = i - M / 2;
  return p < 0 ? p + nf : p;
}

// ---------------------------------------------------------------------------
// fastest axis, fused with the image side
// ---------------------------------------------------------------------------
struct RowsArgs {
  const float2* img;
  float2* img_out;
  const float2* smaps;
  int C, t0, T;
  int conj_smaps;  // forward: multiply by conj(smap) instead of smap
  int smaps_all;   // smaps hold one map per transform ([B][C]) instead of [C]
  // optional second factor cmap[L][N]: transforms are ordered [B][C][L] and
  // transform tt also multiplies by cmap[tt % L] (off-resonance interpolators)
  const float2* cmap;
  int L;
  float2* grid;
  int64_t gstride;
  int M0, M1, M2;
  int nf0, nf1, nf2;
  const float* dec0;
  const float* dec1;
  const float* dec2;
  float scale;
  FFTAxis ax;
  int lgL;
};

// image rows -> deconvolve * scale (* smap) -> pad -> FFT -> grid rows
template <int S, int EPT, int FIX = 0, int FLGL = 0>
__global__ __launch_bounds__(kThreads) void k_rows_fwd(RowsArgs a) {
  extern __shared__ float2 buf[];
  // FIX > 0: FIX-point rows, 2^FLGL per workgroup, passes fixed at compile time
  if constexpr (FIX > 0) a.lgL = FLGL;
  const int L = 1 << a.lgL, LP = L + 1;
  const float2* twl = stage_twiddles(buf, a.ax, LP);
  const int rows = a.M0 * a.M1;
  // transforms fastest in the remapped order: the workgroups that read one
  // block of image rows (all coils of an image) run together on one XCD, so
  // the image comes from HBM once rather than once per coil (192^3, 16
  // coils: 0.809 -> 0.788 ms)
  const BlockIdx3 bi = remapped_block();
  const int lin = bi.x + (int)gridDim.x * bi.y;
  const int t = lin % (int)gridDim.y;
  const int r0 = (lin / (int)gridDim.y) * L;
  const int tt = a.t0 + t;
  const int64_t N = (int64_t)rows * a.M2;
  const int n = a.nf2;
  // line-major thread mapping: thread (l, x0) handles positions x0 + j * TPL
  // of row r0 + l (n <= TPL * EPT), so it keeps one row base and one scale
  const int TPL = kThreads >> a.lgL;
  const int l = threadIdx.x / TPL, x0 = threadIdx.x - l * TPL;
  const int r = r0 + l;
  const bool rok = r < rows;
  const int i0 = rok ? r / a.M1 : 0, i1 = rok ? r - i0 * a.M1 : 0;
  const int per_img = (a.smaps ? a.C : 1) * a.L;  // transforms of one image
  const int rl = rok ? r : 0;  // lanes past the last row read row 0 and drop it
  const float2* src = a.img + (int64_t)(tt / per_img) * N + (int64_t)rl * a.M2;
  const float2* sm =
      a.smaps ? a.smaps + (int64_t)(a.smaps_all ? tt : (tt / a.L) % a.C) * N + (int64_t)rl * a.M2
              : nullptr;
  const float2* cm = a.cmap ? a.cmap + (int64_t)(tt % a.L) * N + (int64_t)rl * a.M2 : nullptr;
  const float rs = rok ? a.dec0[i0] * a.dec1[i1] * a.scale : 0.f;
  const int M2 = a.M2, h2 = M2 / 2;

  // loads first, then the arithmetic and LDS stores
  float2 v[EPT], sv[EPT];
  if constexpr (FIX > 0) {
    // all in flight together: no branch and no arithmetic on a loaded value
    // between them (positions past the row end read its last element and
    // are dropped).  The lengths fixed at compile time have the registers
    // for it within the workgroups per CU their LDS allows.
#pragma unroll
    for (int j = 0; j < EPT; ++j) v[j] = src[min(x0 + j * TPL, M2 - 1)];
    if (sm) {
#pragma unroll
      for (int j = 0; j < EPT; ++j) sv[j] = sm[min(x0 + j * TPL, M2 - 1)];
    } else {
#pragma unroll
      for (int j = 0; j < EPT; ++j) sv[j] = make_float2(1.f, 0.f);
    }
    if (cm) {
#pragma unroll
      for (int j = 0; j < EPT; ++j) v[j] = cmul(v[j], cm[min(x0 + j * TPL, M2 - 1)]);
    }
#pragma unroll
    for (int j = 0; j < EPT; ++j)
      if (!rok) v[j] = make_float2(0.f, 0.f);
  } else {
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      const int x = x0 + j * TPL;
      v[j] = make_float2(0.f, 0.f);
      sv[j] = make_float2(1.f, 0.f);
      if (rok && x < M2) {
        v[j] = src[x];
        if (sm) sv[j] = sm[x];
        if (cm) v[j] = cmul(v[j], cm[x]);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < EPT; ++j) {
    const int x = x0 + j * TPL;
    if (x < M2) {
      float2 s2 = sv[j];
      if (a.conj_smaps) s2.y = -s2.y;
      const float2 w = cscale(cmul(v[j], s2), rs * a.dec2[x]);
      buf[(x < h2 ? x - h2 + n : x - h2) * LP + l] = w;
    }
  }
  const int gap0 = M2 - h2, gap = n - M2;
#pragma unroll
  for (int j = 0; j < EPT; ++j) {
    const int e = x0 + j * TPL;
    if (e < gap) buf[(gap0 + e) * LP + l] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  if constexpr (FIX > 0)
    fft_lines_ct<S, EPT, kThreads, FIX, FLGL>(buf, twl);
  else
    fft_lines<S, EPT>(buf, twl, a.ax, a.lgL, LP);
  if (!rok) return;
  float2* g = a.grid + (int64_t)t * a.gstride +
              ((int64_t)kept_pos(i0, a.M0, a.nf0) * a.nf1 + kept_pos(i1, a.M1, a.nf1)) * n;
#pragma unroll
  for (int j = 0; j < EPT; ++j) {
    const int x = x0 + j * TPL;
    if (x < n) g[x] = buf[x * LP + l];
  }
}

// grid rows -> FFT -> crop -> deconvolve * scale (* conj smap, summed over the
// chunk's coils of one batch) -> image rows
template <int S, int EPT, int FIX = 0, int FLGL = 0>
__global__ __launch_bounds__(kThreads) void k_rows_adj(RowsArgs a) {
  extern __shared__ float2 buf[];
  // FIX > 0: FIX-point rows, 2^FLGL per workgroup, passes fixed at compile time
  if constexpr (FIX > 0) a.lgL = FLGL;
  const int L = 1 << a.lgL, LP = L + 1;
  const float2* twl = stage_twiddles(buf, a.ax, LP);
  const int rows = a.M0 * a.M1;
  const BlockIdx3 bi = remapped_block();
  const int r0 = bi.x * L;
  const int n = a.nf2;
  const int64_t N = (int64_t)rows * a.M2;
  // one output image per image present in the chunk, the sum of its
  // per_img consecutive transforms (coils and / or interpolators)
  const int per_img = (a.smaps ? a.C : 1) * a.L;
  const int out_img = a.t0 / per_img + bi.y;
  const int tb = max(a.t0, out_img * per_img);
  const int te = min(a.t0 + a.T, (out_img + 1) * per_img);
  const bool add_existing = tb % per_img != 0;
  // line-major thread mapping: thread (l, x0) handles positions x0 + j * TPL
  // of line l, so it keeps one row base across the coil loop
  const int TPL = kThreads >> a.lgL;  // threads per line (n <= TPL * EPT)
  const int l = threadIdx.x / TPL, x0 = threadIdx.x - l * TPL;
  const int r = r0 + l;
  const bool rok = r < rows;
  int64_t grow = 0;
  if (rok) {
    const int i0 = r / a.M1, i1 = r - i0 * a.M1;
    grow = ((int64_t)kept_pos(i0, a.M0, a.nf0) * a.nf1 + kept_pos(i1, a.M1, a.nf1)) * n;
  }
  // coil sum of the cropped lines in LDS after the line data and twiddles;
  // each entry is only touched by the thread that owns it
  float2* accl = buf + (size_t)n * (LP + 1) + (size_t)l * a.M2;

  for (int tt = tb; tt < te; ++tt) {
    const float2* g = a.grid + (int64_t)(tt - a.t0) * a.gstride + grow;
    {
      float2 v[EPT];
#pragma unroll
      for (int j = 0; j < EPT; ++j) {
        const int x = x0 + j * TPL;
        v[j] = (rok && x < n) ? g[x] : make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int j = 0; j < EPT; ++j) {
        const int x = x0 + j * TPL;
        if (x < n) buf[x * LP + l] = v[j];
      }
    }
    __syncthreads();
    // the FFT's addressing is the same for every coil; hiding the line
    // geometry from the optimiser keeps it from hoisting all of it out of the
    // coil loop into registers
    if constexpr (FIX > 0) {
      fft_lines_ct<S, EPT, kThreads, FIX, FLGL>(buf, twl);
    } else {
      int lg = a.lgL, lp = LP;
      asm volatile("" : "+s"(lg), "+s"(lp));
      fft_lines<S, EPT>(buf, twl, a.ax, lg, lp);
    }
    const int rl = rok ? r : 0;  // lanes past the last row read row 0's maps and drop them
    const float2* sm =
        a.smaps ? a.smaps + (int64_t)(a.smaps_all ? tt : (tt / a.L) % a.C) * N + (int64_t)rl * a.M2
                : nullptr;
    const float2* cm = a.cmap ? a.cmap + (int64_t)(tt % a.L) * N + (int64_t)rl * a.M2 : nullptr;
    // the coil's map values, all loads in flight together (compile-time
    // lengths; the others load each where it is used)
    float2 sv[EPT];
    if constexpr (FIX > 0) {
      if (sm) {
#pragma unroll
        for (int j = 0; j < EPT; ++j) sv[j] = sm[min(x0 + j * TPL, a.M2 - 1)];
      }
    }
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      const int x = x0 + j * TPL;
      if (x < a.M2) {
        float2 v = make_float2(0.f, 0.f);
        if (rok) {
          v = buf[kept_pos(x, a.M2, n) * LP + l];
          if (sm) {
            float2 s2 = FIX > 0 ? sv[j] : sm[x];
            v = cmul(v, make_float2(s2.x, -s2.y));
          }
          if (cm) {
            float2 cv = cm[x];
            cv.y = -cv.y;
            v = cmul(v, cv);
          }
        }
        accl[x] = (tt == tb) ? v : cadd(accl[x], v);
      }
    }
    __syncthreads();
  }
  if (!rok) return;
  const int i0 = r / a.M1, i1 = r - i0 * a.M1;
  const float rs = a.dec0[i0] * a.dec1[i1] * a.scale;
  float2* dst = a.img_out + (int64_t)out_img * N + (int64_t)r * a.M2;
  for (int x = x0; x < a.M2; x += TPL) {
    float2 v = cscale(accl[x], rs * a.dec2[x]);
    if (add_existing) v = cadd(v, dst[x]);
    dst[x] = v;
  }
}

// ---------------------------------------------------------------------------
// stacked 2.5D: the Cartesian z axis (the image's last, contiguous axis)
// ---------------------------------------------------------------------------

// Direct DFT of the L lines of n elements in buf (any n): each output is a
// sum over the whole line with twiddle index (j k) mod n advanced by k.
template <int S, int EPT>
__device__ void dft_lines_direct(float2* buf, const float2* twl, int n, int lgL, int LP) {
  const int L = 1 << lgL, E = kThreads >> lgL;
  const int l = threadIdx.x & (L - 1), e0 = threadIdx.x >> lgL;
  float2 out[EPT];
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int k = e0 + q * E;
    float2 acc = make_float2(0.f, 0.f);
    if (k < n) {
      int t = 0;
      for (int j = 0; j < n; ++j) {
        float2 w = twl[t];
        if (S > 0) w.y = -w.y;
        const float2 v = buf[j * LP + l];
        acc.x += v.x * w.x - v.y * w.y;
        acc.y += v.x * w.y + v.y * w.x;
        t += k;
        if (t >= n) t -= n;
      }
    }
    out[q] = acc;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int k = e0 + q * E;
    if (k < n) buf[k * LP + l] = out[q];
  }
  __syncthreads();
}

template <int S, int EPT>
__device__ __forceinline__ void fft_lines_any(float2* buf, const float2* twl, const FFTAxis& ax,
                                              int lgL, int LP) {
  if (ax.direct)
    dft_lines_direct<S, EPT>(buf, twl, ax.n, lgL, LP);
  else
    fft_lines<S, EPT>(buf, twl, ax, lgL, LP);
}

struct StackArgs {
  const float2* img;
  const float2* smaps;
  float2* planes;         // forward output
  const float2* planes_in;  // adjoint input
  float2* img_out;
  const int32_t* z_index;
  const int32_t* z_plane;
  int C, t0, T;
  int NZ;
  int64_t rows;
  float scale;
  FFTAxis ax;
  int lgL;
};

// rows r0 .. r0 + L of transform t: (smap *) image rows -> centred DFT along
// z -> the NZ sampled positions, written plane-major (L adjacent rows of a
// plane are one contiguous run)
template <int EPT>
__global__ __launch_bounds__(kThreads) void k_stack_fwd(StackArgs a) {
  extern __shared__ float2 buf[];
  const int L = 1 << a.lgL, LP = L + 1;
  const float2* twl = stage_twiddles(buf, a.ax, LP);
  const int Z = a.ax.n, h = Z / 2;
  const int64_t r0 = (int64_t)blockIdx.x * L;
  const int t = blockIdx.y, tt = a.t0 + t;
  // line-major loads: thread (l, x0) holds positions x0 + j * TPL of row r0 + l
  const int TPL = kThreads >> a.lgL;
  {
    const int l = threadIdx.x / TPL, x0 = threadIdx.x - l * TPL;
    const int64_t r = r0 + l;
    const bool rok = r < a.rows;
    const int64_t N = a.rows * Z;
    const float2* src = a.img + (int64_t)(a.smaps ? tt / a.C : tt) * N + r * Z;
    const float2* sm = a.smaps ? a.smaps + (int64_t)(tt % a.C) * N + r * Z : nullptr;
    float2 v[EPT], sv[EPT];
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      const int x = x0 + j * TPL;
      v[j] = make_float2(0.f, 0.f);
      sv[j] = make_float2(1.f, 0.f);
      if (rok && x < Z) {
        v[j] = src[x];
        if (sm) sv[j] = sm[x];
      }
    }
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      const int x = x0 + j * TPL;
      // centred input: position x sits at (x - Z/2) mod Z of the DFT line
      if (x < Z) buf[(x < h ? x - h + Z : x - h) * LP + l] = cmul(v[j], sv[j]);
    }
  }
  __syncthreads();
  fft_lines_any<-1, EPT>(buf, twl, a.ax, a.lgL, LP);
  // plane-major stores: thread (l, i0) writes planes i0 + q * (kThreads / L)
  const int l = threadIdx.x & (L - 1), i0 = threadIdx.x >> a.lgL;
  const int64_t r = r0 + l;
  if (r >= a.rows) return;
  float2* dst = a.planes + (int64_t)t * a.NZ * a.rows + r;
  for (int i = i0; i < a.NZ; i += TPL) {
    const int k = a.z_index[i];  // centred output k is DFT index (k - Z/2) mod Z
    dst[(int64_t)i * a.rows] = cscale(buf[(k < h ? k - h + Z : k - h) * LP + l], a.scale);
  }
}

// rows r0 .. r0 + L of the output image(s) of block y: per transform of the
// chunk, zero-filled z line from the planes -> centred inverse DFT -> conj
// smap, summed over the transforms of one image in registers
template <int EPT>
__global__ __launch_bounds__(kThreads) void k_stack_adj(StackArgs a) {
  extern __shared__ float2 buf[];
  const int L = 1 << a.lgL, LP = L + 1;
  const float2* twl = stage_twiddles(buf, a.ax, LP);
  const int Z = a.ax.n, h = Z / 2;
  const int64_t r0 = (int64_t)blockIdx.x * L;
  int tb, te, out_img;
  if (a.smaps) {
    const int b = a.t0 / a.C + blockIdx.y;
    out_img = b;
    tb = max(a.t0, b * a.C);
    te = min(a.t0 + a.T, (b + 1) * a.C);
  } else {
    out_img = a.t0 + blockIdx.y;
    tb = out_img;
    te = out_img + 1;
  }
  const bool add_existing = a.smaps && (tb % a.C != 0);
  const int TPL = kThreads >> a.lgL;
  const int lq = threadIdx.x & (L - 1), q0 = threadIdx.x >> a.lgL;  // gather mapping
  const int lr = threadIdx.x / TPL, x0 = threadIdx.x - lr * TPL;    // row mapping
  const int64_t rq = r0 + lq, rr = r0 + lr;
  const int64_t N = a.rows * Z;
  float2 acc[EPT];
#pragma unroll
  for (int j = 0; j < EPT; ++j) acc[j] = make_float2(0.f, 0.f);
  for (int tt = tb; tt < te; ++tt) {
    const float2* pl = a.planes_in + (int64_t)(tt - a.t0) * a.NZ * a.rows + rq;
    // DFT position j holds centred position (j + Z/2) mod Z
    for (int j = q0; j < Z; j += TPL) {
      const int k = j + h < Z ? j + h : j + h - Z;
      const int i = a.z_plane[k];
      buf[j * LP + lq] = (i >= 0 && rq < a.rows) ? pl[(int64_t)i * a.rows] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft_lines_any<1, EPT>(buf, twl, a.ax, a.lgL, LP);
    const float2* sm = a.smaps ? a.smaps + (int64_t)(tt % a.C) * N + rr * Z : nullptr;
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      const int x = x0 + j * TPL;
      if (x < Z && rr < a.rows) {
        float2 v = cscale(buf[(x < h ? x - h + Z : x - h) * LP + lr], a.scale);
        if (sm) {
          float2 sv = sm[x];
          sv.y = -sv.y;
          v = cmul(v, sv);
        }
        acc[j] = cadd(acc[j], v);
      }
    }
    __syncthreads();
  }
  if (rr >= a.rows) return;
  float2* dst = a.img_out + (int64_t)out_img * N + rr * Z;
#pragma unroll
  for (int j = 0; j < EPT; ++j) {
    const int x = x0 + j * TPL;
    if (x < Z) dst[x] = add_existing ? cadd(acc[j], dst[x]) : acc[j];
  }
}

// ---------------------------------------------------------------------------
// strided axes (0 and 1), in place
// ---------------------------------------------------------------------------
struct StridedArgs {
  float2* data;
  int64_t tstride;
  int n;
  int64_t lstride;
  int nouter, outer_M, outer_nf;  // outer_M > 0: outer index i maps to kept_pos
  int64_t ostride;
  int nx;
  int in_M, out_M;  // 0: full line, > 0: kept positions only
  FFTAxis ax;
  int lgL;
  const float2* K;  // CONV: spectrum (one grid) applied between the FFTs
};

// In-place FFT of L adjacent lines along a strided axis.  CONV: forward FFT,
// multiply by the spectrum K, inverse FFT, all in LDS (the middle of the
// Toeplitz convolution, one grid round trip instead of three).
template <int S, bool CONV, int EPT, int NT, int FIX = 0, int FLGL = 0>
__global__ __launch_bounds__(NT) void k_fft_strided(StridedArgs a) {
  extern __shared__ float2 buf[];
  // FIX > 0: FIX-point lines, 2^FLGL per workgroup, passes fixed at compile time
  if constexpr (FIX > 0) a.lgL = FLGL;
  const int L = 1 << a.lgL, LP = L + 1;
  const float2* twl = stage_twiddles<NT>(buf, a.ax, LP);
  const BlockIdx3 bi = remapped_block();
  const int x0 = bi.x * L;
  const int oi = bi.y;
  const int t = bi.z;
  const int opos = a.outer_M ? kept_pos(oi, a.outer_M, a.outer_nf) : oi;
  // thread -> (line l, positions e0 + i*E); a grid holds < 2^31 cells, so
  // position offsets are 32-bit and advance by uniform steps per slot
  const int E = NT >> a.lgL;
  const int l = threadIdx.x & (L - 1), e0 = threadIdx.x >> a.lgL;
  const bool lok = x0 + l < a.nx;
  float2* base = a.data + (int64_t)t * a.tstride + (int64_t)opos * a.ostride + x0 + l;
  const int ls = (int)a.lstride;
  const int n = a.n;
  const int nls = n * ls, nLP = n * LP;
  const int nin = a.in_M ? a.in_M : n;
  const int hin = a.in_M ? a.in_M / 2 : 0;  // kept position of e: e - h (+ n if e < h)
  const int gofs = (e0 - hin) * ls, lofs = (e0 - hin) * LP + l;
  // all of a thread's loads are issued before the first LDS store, so they
  // overlap instead of paying one memory latency each (n * L <= EPT * NT)
  float2 v[EPT];
#pragma unroll
  for (int i = 0; i < EPT; ++i) {
    const int e = e0 + i * E;
    v[i] = make_float2(0.f, 0.f);
    if (e < nin && lok) v[i] = base[gofs + i * (E * ls) + (e < hin ? nls : 0)];
  }
#pragma unroll
  for (int i = 0; i < EPT; ++i) {
    const int e = e0 + i * E;
    if (e < nin) buf[lofs + i * (E * LP) + (e < hin ? nLP : 0)] = v[i];
  }
  if (a.in_M) {  // zero the positions between the kept halves
    const int gap0 = a.in_M - hin, gap = n - a.in_M;
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int e = e0 + i * E;
      if (e < gap) buf[(gap0 + e) * LP + l] = make_float2(0.f, 0.f);
    }
  }
  __syncthreads();
  if constexpr (FIX > 0)
    fft_lines_ct<S, EPT, NT, FIX, FLGL>(buf, twl);
  else
    fft_lines<S, EPT, NT>(buf, twl, a.ax, a.lgL, LP);
  if constexpr (CONV) {
    const float2* kb = a.K + (int64_t)opos * a.ostride + x0 + l;
    const int kofs = e0 * ls;
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int e = e0 + i * E;
      v[i] = make_float2(0.f, 0.f);
      if (e < n && lok) v[i] = kb[kofs + i * (E * ls)];
    }
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int e = e0 + i * E;
      if (e < n) {
        const int q = e * LP + l;
        buf[q] = cmul(buf[q], v[i]);
      }
    }
    __syncthreads();
    if constexpr (FIX > 0)
      fft_lines_ct<-S, EPT, NT, FIX, FLGL>(buf, twl);
    else
      fft_lines<-S, EPT, NT>(buf, twl, a.ax, a.lgL, LP);
  }
  const int nout = a.out_M ? a.out_M : n;
  const int hout = a.out_M ? a.out_M / 2 : 0;
  const int gofs2 = (e0 - hout) * ls, lofs2 = (e0 - hout) * LP + l;
  if (lok) {
#pragma unroll
    for (int i = 0; i < EPT; ++i) {
      const int e = e0 + i * E;
      if (e < nout)
        base[gofs2 + i * (E * ls) + (e < hout ? nls : 0)] =
            buf[lofs2 + i * (E * LP) + (e < hout ? nLP : 0)];
    }
  }
}

int lg2(int x) {
  int l = 0;
  while ((1 << l) < x) ++l;
  return l;
}

size_t fft_lds(const FFTAxis& ax) { return sizeof(float2) * (size_t)ax.n * (ax.L + 2); }

// the axis has n points on L lines and the runtime radix plan is the
// compile-time one of fft_lines_ct
bool fixed_plan(const FFTAxis& ax, int n, int L) {
  if (ax.n != n || ax.L != L || ax.direct) return false;
  int p = 0;
  for (; p < ax.nrad; ++p)
    if (ax.rad[p] != plan_radix(n, p)) return false;
  return plan_radix(n, p) == 0;
}

// complex elements per thread for the L lines of n elements of an axis:
// the smallest of 8, 12, 16 with n * L <= EPT * nt (fewer registers,
// more resident waves for the shorter lines)
int ept_of(const FFTAxis& ax, int nt = kThreads) {
  const int need = (ax.n * ax.L + nt - 1) / nt;
  return need <= 8 ? 8 : need <= 12 ? 12 : 16;
}

#define HN_EPT_SWITCH(ept, ...)             \
  switch (ept) {                            \
    case 8: { constexpr int EE = 8; __VA_ARGS__ } break;   \
    case 12: { constexpr int EE = 12; __VA_ARGS__ } break; \
    default: { constexpr int EE = 16; __VA_ARGS__ } break; \
  }

}  // namespace

// k_rows_adj also keeps the coil sum of its L cropped lines in LDS
size_t rows_adj_lds(const FFTAxis& ax, int64_t m2) {
  return fft_lds(ax) + sizeof(float2) * (size_t)m2 * ax.L;
}

namespace {

int strided(const hipnufft_plan_s* p, int axis, int T, int in_M, int out_M, int outer_M,
            int sign, int stage, const float2* K = nullptr) {
  const FFTAxis& ax = p->fft[axis];
  StridedArgs a{};
  a.data = p->grid;
  a.tstride = grid_size(p);
  a.n = (int)p->nf[axis];
  a.nx = (int)p->nf[2];
  a.in_M = in_M;
  a.out_M = out_M;
  a.ax = ax;
  a.lgL = lg2(ax.L);
  if (axis == 1) {
    a.lstride = p->nf[2];
    a.ostride = p->nf[1] * p->nf[2];
    a.outer_nf = (int)p->nf[0];
  } else {
    a.lstride = p->nf[1] * p->nf[2];
    a.ostride = p->nf[2];
    a.outer_nf = (int)p->nf[1];
  }
  a.outer_M = outer_M;
  a.nouter = outer_M ? outer_M : a.outer_nf;
  a.K = K;
  const dim3 grd((unsigned)((a.nx + ax.L - 1) / ax.L), (unsigned)a.nouter, (unsigned)T);
  StageScope timer(p, stage);
  const size_t lds = fft_lds(ax);
  // strided axes may hold up to 2 * kMaxElemsFFT elements (16 lines of up to
  // 512): those run 512-thread workgroups
// compile-time plans: (points, log2 lines, threads, elements per thread)
#define HN_STRIDED_FIXED(NN, LG, NTT, EPTT)                                                    \
  if (fixed_plan(ax, NN, 1 << LG)) {                                                             \
    if (K) {                                                                                     \
      if (sign < 0)                                                                              \
        hipLaunchKernelGGL((k_fft_strided<-1, true, EPTT, NTT, NN, LG>), grd, dim3(NTT), lds,    \
                           p->stream, a);                                                        \
      else                                                                                       \
        hipLaunchKernelGGL((k_fft_strided<1, true, EPTT, NTT, NN, LG>), grd, dim3(NTT), lds,     \
                           p->stream, a);                                                        \
    } else if (sign < 0) {                                                                       \
      hipLaunchKernelGGL((k_fft_strided<-1, false, EPTT, NTT, NN, LG>), grd, dim3(NTT), lds,     \
                         p->stream, a);                                                          \
    } else {                                                                                     \
      hipLaunchKernelGGL((k_fft_strided<1, false, EPTT, NTT, NN, LG>), grd, dim3(NTT), lds,      \
                         p->stream, a);                                                          \
    }                                                                                            \
    HN_LAUNCH_CHECK();                                                                           \
    return 0;                                                                                    \
  }
  HN_STRIDED_FIXED(384, 4, 2 * kThreads, 12)
  HN_STRIDED_FIXED(256, 4, kThreads, 16)
  HN_STRIDED_FIXED(512, 3, kThreads, 16)
  HN_STRIDED_FIXED(768, 3, 2 * kThreads, 12)
#undef HN_STRIDED_FIXED
  if (ax.n * ax.L > kMaxElemsFFT) {
    HN_EPT_SWITCH(ept_of(ax, 2 * kThreads), {
      constexpr int NT2 = 2 * kThreads;
      if (K) {
        if (sign < 0)
          hipLaunchKernelGGL((k_fft_strided<-1, true, EE, NT2>), grd, dim3(NT2), lds, p->stream, a);
        else
          hipLaunchKernelGGL((k_fft_strided<1, true, EE, NT2>), grd, dim3(NT2), lds, p->stream, a);
      } else if (sign < 0) {
        hipLaunchKernelGGL((k_fft_strided<-1, false, EE, NT2>), grd, dim3(NT2), lds, p->stream, a);
      } else {
        hipLaunchKernelGGL((k_fft_strided<1, false, EE, NT2>), grd, dim3(NT2), lds, p->stream, a);
      }
    });
  } else {
    HN_EPT_SWITCH(ept_of(ax), {
      if (K) {
        if (sign < 0)
          hipLaunchKernelGGL((k_fft_strided<-1, true, EE, kThreads>), grd, dim3(kThreads), lds,
                             p->stream, a);
        else
          hipLaunchKernelGGL((k_fft_strided<1, true, EE, kThreads>), grd, dim3(kThreads), lds,
                             p->stream, a);
      } else if (sign < 0) {
        hipLaunchKernelGGL((k_fft_strided<-1, false, EE, kThreads>), grd, dim3(kThreads), lds,
                           p->stream, a);
      } else {
        hipLaunchKernelGGL((k_fft_strided<1, false, EE, kThreads>), grd, dim3(kThreads), lds,
                           p->stream, a);
      }
    });
  }
  HN_LAUNCH_CHECK();
  return 0;
}

RowsArgs rows_args(const hipnufft_plan_s* p) {
  RowsArgs a{};
  a.grid = p->grid;
  a.gstride = grid_size(p);
  a.M0 = (int)p->N[0];
  a.M1 = (int)p->N[1];
  a.M2 = (int)p->N[2];
  a.nf0 = (int)p->nf[0];
  a.nf1 = (int)p->nf[1];
  a.nf2 = (int)p->nf[2];
  a.dec0 = p->dec[0];
  a.dec1 = p->dec[1];
  a.dec2 = p->dec[2];
  a.ax = p->fft[2];
  a.lgL = lg2(p->fft[2].L);
  a.L = 1;
  return a;
}

}  // namespace

int64_t grid_size(const hipnufft_plan_s* p) { return p->nf[0] * p->nf[1] * p->nf[2]; }

namespace {

// grid[t][i] *= K[i] for the T grids of a chunk (the Toeplitz spectrum)
__global__ __launch_bounds__(kThreads) void k_mul_spectrum(float2* grid, const float2* K, int64_t G) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= G) return;
  float2* g = grid + (int64_t)blockIdx.y * G;
  g[i] = cmul(g[i], K[i]);
}

}  // namespace

int mul_spectrum(const hipnufft_plan_s* p, const float2* K, int T) {
  const int64_t G = grid_size(p);
  if (T <= 0 || G <= 0) return 0;
  const dim3 grd((unsigned)((G + kThreads - 1) / kThreads), (unsigned)T);
  hipLaunchKernelGGL(k_mul_spectrum, grd, dim3(kThreads), 0, p->stream, p->grid, K, G);
  HN_LAUNCH_CHECK();
  return 0;
}

namespace {

int rows_fwd(const hipnufft_plan_s* p, const float2* img, const float2* smaps, int C, int t0, int T,
             float scale, int sign, int conj_smaps, int smaps_all, const float2* cmap = nullptr,
             int nmaps = 1) {
  RowsArgs a = rows_args(p);
  a.conj_smaps = conj_smaps;
  a.smaps_all = smaps_all;
  a.cmap = cmap;
  a.L = nmaps;
  a.img = img;
  a.smaps = smaps;
  a.C = C;
  a.t0 = t0;
  a.T = T;
  a.scale = scale;
  const int rows = a.M0 * a.M1;
  const int L = p->fft[2].L;
  const dim3 grd((unsigned)((rows + L - 1) / L), (unsigned)T);
  const size_t lds = fft_lds(p->fft[2]);
  {
    StageScope timer(p, kStageRowsFwd);
    const FFTAxis& ax = p->fft[2];
    if (fixed_plan(ax, 384, 8)) {
      if (sign < 0)
        hipLaunchKernelGGL((k_rows_fwd<-1, 12, 384, 3>), grd, dim3(kThreads), lds, p->stream, a);
      else
        hipLaunchKernelGGL((k_rows_fwd<1, 12, 384, 3>), grd, dim3(kThreads), lds, p->stream, a);
    } else if (fixed_plan(ax, 256, 16)) {
      if (sign < 0)
        hipLaunchKernelGGL((k_rows_fwd<-1, 16, 256, 4>), grd, dim3(kThreads), lds, p->stream, a);
      else
        hipLaunchKernelGGL((k_rows_fwd<1, 16, 256, 4>), grd, dim3(kThreads), lds, p->stream, a);
    } else if (fixed_plan(ax, 512, 8)) {
      if (sign < 0)
        hipLaunchKernelGGL((k_rows_fwd<-1, 16, 512, 3>), grd, dim3(kThreads), lds, p->stream, a);
      else
        hipLaunchKernelGGL((k_rows_fwd<1, 16, 512, 3>), grd, dim3(kThreads), lds, p->stream, a);
    } else if (fixed_plan(ax, 768, 4)) {
      if (sign < 0)
        hipLaunchKernelGGL((k_rows_fwd<-1, 12, 768, 2>), grd, dim3(kThreads), lds, p->stream, a);
      else
        hipLaunchKernelGGL((k_rows_fwd<1, 12, 768, 2>), grd, dim3(kThreads), lds, p->stream, a);
    } else {
      HN_EPT_SWITCH(ept_of(p->fft[2]), {
        if (sign < 0)
          hipLaunchKernelGGL((k_rows_fwd<-1, EE>), grd, dim3(kThreads), lds, p->stream, a);
        else
          hipLaunchKernelGGL((k_rows_fwd<1, EE>), grd, dim3(kThreads), lds, p->stream, a);
      });
    }
    HN_LAUNCH_CHECK();
  }
  return 0;
}

int rows_adj(const hipnufft_plan_s* p, float2* img, const float2* smaps, int C, int t0, int T,
             float scale, int sign, int smaps_all, const float2* cmap = nullptr, int nmaps = 1) {
  RowsArgs a = rows_args(p);
  a.img_out = img;
  a.smaps = smaps;
  a.smaps_all = smaps_all;
  a.cmap = cmap;
  a.L = nmaps;
  a.C = C;
  a.t0 = t0;
  a.T = T;
  a.scale = scale;
  const int rows = a.M0 * a.M1;
  const int L = p->fft[2].L;
  const int per_img = (smaps ? C : 1) * nmaps;
  const int nimg = (t0 + T - 1) / per_img - t0 / per_img + 1;
  const dim3 grd((unsigned)((rows + L - 1) / L), (unsigned)nimg);
  const size_t lds = rows_adj_lds(p->fft[2], a.M2);
  StageScope timer(p, kStageRowsAdj);
  const FFTAxis& ax = p->fft[2];
  if (fixed_plan(ax, 384, 8)) {
    if (sign < 0)
      hipLaunchKernelGGL((k_rows_adj<-1, 12, 384, 3>), grd, dim3(kThreads), lds, p->stream, a);
    else
      hipLaunchKernelGGL((k_rows_adj<1, 12, 384, 3>), grd, dim3(kThreads), lds, p->stream, a);
  } else if (fixed_plan(ax, 256, 16)) {
    if (sign < 0)
      hipLaunchKernelGGL((k_rows_adj<-1, 16, 256, 4>), grd, dim3(kThreads), lds, p->stream, a);
    else
      hipLaunchKernelGGL((k_rows_adj<1, 16, 256, 4>), grd, dim3(kThreads), lds, p->stream, a);
  } else if (fixed_plan(ax, 512, 8)) {
    if (sign < 0)
      hipLaunchKernelGGL((k_rows_adj<-1, 16, 512, 3>), grd, dim3(kThreads), lds, p->stream, a);
    else
      hipLaunchKernelGGL((k_rows_adj<1, 16, 512, 3>), grd, dim3(kThreads), lds, p->stream, a);
  } else if (fixed_plan(ax, 768, 4)) {
    if (sign < 0)
      hipLaunchKernelGGL((k_rows_adj<-1, 12, 768, 2>), grd, dim3(kThreads), lds, p->stream, a);
    else
      hipLaunchKernelGGL((k_rows_adj<1, 12, 768, 2>), grd, dim3(kThreads), lds, p->stream, a);
  } else {
    HN_EPT_SWITCH(ept_of(p->fft[2]), {
      if (sign < 0)
        hipLaunchKernelGGL((k_rows_adj<-1, EE>), grd, dim3(kThreads), lds, p->stream, a);
      else
        hipLaunchKernelGGL((k_rows_adj<1, EE>), grd, dim3(kThreads), lds, p->stream, a);
    });
  }
  HN_LAUNCH_CHECK();
  return 0;
}

}  // namespace

int fwd_to_grid(const hipnufft_plan_s* p, const float2* img, const float2* smaps, int C, int t0,
                int T, float scale, int sign, int conj_smaps, int smaps_all, const float2* cmap,
                int L) {
  if (T <= 0) return 0;
  int rc = rows_fwd(p, img, smaps, C, t0, T, scale, sign, conj_smaps, smaps_all, cmap, L);
  if (rc) return rc;
  if (p->ndim >= 2 &&
      (rc = strided(p, 1, T, (int)p->N[1], 0, (int)p->N[0], sign, kStageColsFwd)))
    return rc;
  if (p->ndim == 3 && (rc = strided(p, 0, T, (int)p->N[0], 0, 0, sign, kStageSlabsFwd)))
    return rc;
  return 0;
}

int grid_to_img(const hipnufft_plan_s* p, float2* img, const float2* smaps, int C, int t0, int T,
                float scale, int sign, int smaps_all, const float2* cmap, int L) {
  if (T <= 0) return 0;
  int rc;
  if (p->ndim == 3 && (rc = strided(p, 0, T, 0, (int)p->N[0], 0, sign, kStageSlabsAdj)))
    return rc;
  if (p->ndim >= 2 &&
      (rc = strided(p, 1, T, 0, (int)p->N[1], (int)p->N[0], sign, kStageColsAdj)))
    return rc;
  return rows_adj(p, img, smaps, C, t0, T, scale, sign, smaps_all, cmap, L);
}

int gram_chain(const hipnufft_plan_s* p, const float2* img, const float2* smaps, float2* out,
               const float2* K, int C, int t0, int T) {
  if (T <= 0) return 0;
  int rc;
  if (p->ndim == 1) {  // the only FFT axis is the contiguous one: unfused
    if ((rc = fwd_to_grid(p, img, smaps, C, t0, T, 1.0f, -1))) return rc;
    if ((rc = mul_spectrum(p, K, T))) return rc;
    return grid_to_img(p, out, smaps, C, t0, T, 1.0f, 1);
  }
  if ((rc = rows_fwd(p, img, smaps, C, t0, T, 1.0f, -1, 0, 0))) return rc;
  if (p->ndim == 3) {
    if ((rc = strided(p, 1, T, (int)p->N[1], 0, (int)p->N[0], -1, kStageColsFwd))) return rc;
    // axis 0: forward FFT, spectrum, inverse FFT in one pass
    if ((rc = strided(p, 0, T, (int)p->N[0], (int)p->N[0], 0, -1, kStageSlabsFwd, K))) return rc;
    if ((rc = strided(p, 1, T, 0, (int)p->N[1], (int)p->N[0], 1, kStageColsAdj))) return rc;
  } else {
    if ((rc = strided(p, 1, T, (int)p->N[1], (int)p->N[1], (int)p->N[0], -1, kStageColsFwd, K)))
      return rc;
  }
  return rows_adj(p, out, smaps, C, t0, T, 1.0f, 1, 0);
}

int stack_fwd(const StackGeom& g, const float2* img, const float2* smaps, int C, int t0, int T,
              float scale, float2* planes, hipStream_t s) {
  if (T <= 0 || g.rows <= 0) return 0;
  StackArgs a{};
  a.img = img;
  a.smaps = smaps;
  a.planes = planes;
  a.z_index = g.z_index;
  a.C = C;
  a.t0 = t0;
  a.T = T;
  a.NZ = g.NZ;
  a.rows = g.rows;
  a.scale = scale;
  a.ax = g.ax;
  a.lgL = lg2(g.ax.L);
  const dim3 grd((unsigned)((g.rows + g.ax.L - 1) / g.ax.L), (unsigned)T);
  HN_EPT_SWITCH(ept_of(g.ax), {
    hipLaunchKernelGGL((k_stack_fwd<EE>), grd, dim3(kThreads), fft_lds(g.ax), s, a);
  });
  HN_LAUNCH_CHECK();
  return 0;
}

int stack_adj(const StackGeom& g, const float2* planes, const float2* smaps, int C, int t0, int T,
              float scale, float2* img, hipStream_t s) {
  if (T <= 0 || g.rows <= 0) return 0;
  StackArgs a{};
  a.planes_in = planes;
  a.smaps = smaps;
  a.img_out = img;
  a.z_plane = g.z_plane;
  a.C = C;
  a.t0 = t0;
  a.T = T;
  a.NZ = g.NZ;
  a.rows = g.rows;
  a.scale = scale;
  a.ax = g.ax;
  a.lgL = lg2(g.ax.L);
  const int nimg = smaps ? ((t0 + T - 1) / C - t0 / C + 1) : T;
  const dim3 grd((unsigned)((g.rows + g.ax.L - 1) / g.ax.L), (unsigned)nimg);
  HN_EPT_SWITCH(ept_of(g.ax), {
    hipLaunchKernelGGL((k_stack_adj<EE>), grd, dim3(kThreads), fft_lds(g.ax), s, a);
  });
  HN_LAUNCH_CHECK();
  return 0;
}

}  // namespace hn
