// Bit-exact binning and stable sort of the non-uniform points (setpts).
//
// Replaces the bin-sort that cufinufft performs inside plan._setpts
// (operators/interfaces/cufinufft.py:94-113, gpu_sort=1, cufinufft.py:38-46).
// Every rule here is integer or a single IEEE double operation, evaluated with
// floating-point contraction disabled, so the permutation equals the CPU
// oracle's numpy.argsort(kind="stable") bit for bit.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "internal.h"

namespace hn {
namespace {

// Stream-ordered temporaries, freed (hipFreeAsync on the owning stream) on
// every exit of the scope that holds them, error returns included.
class StreamTemps {
 public:
  explicit StreamTemps(hipStream_t s) : s_(s) {}
  StreamTemps(const StreamTemps&) = delete;
  StreamTemps& operator=(const StreamTemps&) = delete;
  ~StreamTemps() { (void)release(); }

  template <typename T>
  hipError_t alloc(T** ptr, int64_t n) {
    *ptr = nullptr;
    const size_t bytes = sizeof(T) * (size_t)(n > 0 ? n : 1);
    hipError_t e = hipMallocAsync((void**)ptr, bytes, s_);
    if (e == hipSuccess) held_.push_back((void*)*ptr);
    return e;
  }

  // frees everything held so far; returns the first error
  hipError_t release() {
    hipError_t first = hipSuccess;
    for (void* q : held_) {
      hipError_t e = hipFreeAsync(q, s_);
      if (first == hipSuccess) first = e;
    }
    held_.clear();
    return first;
  }

 private:
  hipStream_t s_;
  std::vector<void*> held_;
};


constexpr int kScanIPT = 8;                 // items per thread in scans
constexpr int kScanChunk = kThreads * kScanIPT;
constexpr int kSortIPT = 16;                // items per thread in radix passes
constexpr int kSortChunk = kThreads * kSortIPT;
constexpr int kWaves = kThreads / 64;

// ---------------------------------------------------------------------------
// exclusive scan
// ---------------------------------------------------------------------------
__device__ int block_exclusive_scan(int v, int* wsum, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int w = 0; w < kWaves; ++w) {
      int t = wsum[w];
      wsum[w] = run;
      run += t;
    }
    *total = run;
  }
  __syncthreads();
  return x - v + wsum[wave];
}

__global__ __launch_bounds__(kThreads) void k_scan_blocks(const int32_t* in, int32_t* out,
                                                          int32_t* bsum, int64_t n) {
  __shared__ int wsum[kWaves];
  __shared__ int total;
  const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanIPT;
  int v[kScanIPT];
  int s = 0;
#pragma unroll
  for (int i = 0; i < kScanIPT; ++i) {
    v[i] = (base + i < n) ? in[base + i] : 0;
    s += v[i];
  }
  int ex = block_exclusive_scan(s, wsum, &total);
#pragma unroll
  for (int i = 0; i < kScanIPT; ++i) {
    if (base + i < n) out[base + i] = ex;
    ex += v[i];
  }
  if (threadIdx.x == 0 && bsum) bsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(kThreads) void k_scan_add(int32_t* out, const int32_t* boff, int64_t n) {
  const int64_t base = (int64_t)blockIdx.x * kScanChunk;
  const int add = boff[blockIdx.x];
  for (int i = threadIdx.x; i < kScanChunk; i += kThreads)
    if (base + i < n) out[base + i] += add;
}

// ---------------------------------------------------------------------------
// point folding / binning
// ---------------------------------------------------------------------------
struct FoldArgs {
  const float* x[3];
  double scale[3];
  double nf[3];
  double inv_tile[3];
  int nt[3];
  int nfi[3];
  double halfw;
  int64_t M;
  const int32_t* set_off;  // point sets: key = set * nbins + tile
  int nsets;
  uint32_t nbins;
  uint32_t* key;
  int32_t* val;
  int32_t* cell[3];
  float* off[3];
};

__device__ inline double fold_coord(float x, double scale, double nf) {
  double g = (double)x * scale;
  if (g < 0.0) g = g + nf;
  if (g < 0.0) g = g + nf;
  if (g >= nf) g = g - nf;
  if (g >= nf) g = g - nf;
  if (!(g >= 0.0 && g < nf)) g = 0.0;  // NaN / far out of range: defined, in bounds
  return g;
}

template <int D>
__global__ __launch_bounds__(kThreads) void k_fold_bin(FoldArgs a) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= a.M) return;
  uint32_t key = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double g = fold_coord(a.x[d][i], a.scale[d], a.nf[d]);
    int b = (int)floor(g * a.inv_tile[d]);
    if (b >= a.nt[d]) b = a.nt[d] - 1;
    key = key * (uint32_t)a.nt[d] + (uint32_t)b;
    const double s = ceil(g - a.halfw);
    int c0 = (int)s;
    a.off[d][i] = (float)(g - s);
    if (c0 < 0) c0 += a.nfi[d];
    if (c0 >= a.nfi[d]) c0 -= a.nfi[d];
    a.cell[d][i] = c0;
  }
  if (a.nsets > 1) key += (uint32_t)find_set(a.set_off, a.nsets, i) * a.nbins;
  a.key[i] = key;
  a.val[i] = (int32_t)i;
}

// ---------------------------------------------------------------------------
// stable LSD radix sort, 8-bit digits
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_digit_hist(const uint32_t* keys, int64_t n, int shift,
                                                         int32_t* hist, int nblk) {
  __shared__ int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortChunk;
  for (int r = 0; r < kSortIPT; ++r) {
    const int64_t i = base + r * kThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kThreads) void k_digit_scatter(const uint32_t* kin, const int32_t* vin,
                                                            uint32_t* kout, int32_t* vout,
                                                            int64_t n, int shift,
                                                            const int32_t* offs, int nblk) {
  __shared__ int base[256];
  __shared__ int wcnt[kWaves][256];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  base[threadIdx.x] = offs[(int64_t)threadIdx.x * nblk + blockIdx.x];
  const int64_t b0 = (int64_t)blockIdx.x * kSortChunk;
  for (int r = 0; r < kSortIPT; ++r) {
    const int64_t i = b0 + r * kThreads + threadIdx.x;
    const bool valid = i < n;
    const uint32_t key = valid ? kin[i] : 0u;
    const int dig = (int)((key >> shift) & 255u);
    for (int k = threadIdx.x; k < kWaves * 256; k += kThreads) (&wcnt[0][0])[k] = 0;
    __syncthreads();
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (dig >> b) & 1;
      const unsigned long long m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const int lrank = __popcll(peers & lt);
    if (valid && lrank == 0) wcnt[wave][dig] = __popcll(peers);
    __syncthreads();
    {
      int run = base[threadIdx.x];
      for (int w = 0; w < kWaves; ++w) {
        const int c = wcnt[w][threadIdx.x];
        wcnt[w][threadIdx.x] = run;
        run += c;
      }
      base[threadIdx.x] = run;
    }
    __syncthreads();
    if (valid) {
      const int pos = wcnt[wave][dig] + lrank;
      kout[pos] = key;
      vout[pos] = vin[i];
    }
    __syncthreads();
  }
}

// Sort (keys, vals) stably by key < 2^bits.  Result ends in (k0, v0).
int radix_sort(uint32_t* k0, int32_t* v0, uint32_t* k1, int32_t* v1, int64_t n,
               int bits, hipStream_t s) {
  if (n <= 1 || bits <= 0) return 0;
  const int nblk = (int)((n + kSortChunk - 1) / kSortChunk);
  int32_t* hist = nullptr;
  StreamTemps tmp(s);
  HN_HIP(tmp.alloc(&hist, 256 * (int64_t)nblk));
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(k_digit_hist, dim3(nblk), dim3(kThreads), 0, s, k0, n, shift, hist, nblk);
    HN_LAUNCH_CHECK();
    int rc = scan_exclusive(hist, hist, 256LL * nblk, s, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL(k_digit_scatter, dim3(nblk), dim3(kThreads), 0, s, k0, v0, k1, v1, n,
                       shift, hist, nblk);
    HN_LAUNCH_CHECK();
    std::swap(k0, k1);
    std::swap(v0, v1);
  }
  HN_HIP(tmp.release());
  // an odd number of passes leaves the result in the second buffer: copy back
  const int passes = (bits + 7) / 8;
  if (passes & 1) {
    HN_HIP(hipMemcpyAsync(k1, k0, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
    HN_HIP(hipMemcpyAsync(v1, v0, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
  }
  return 0;
}

// start[b] = first sorted position with key >= b, b in [0, nbins]
__global__ __launch_bounds__(kThreads) void k_bin_start(const uint32_t* skeys, int64_t n, int64_t nbins,
                                                        int32_t* start) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i > n) return;
  const int64_t prev = (i == 0) ? -1 : (int64_t)skeys[i - 1];
  const int64_t cur = (i == n) ? nbins : (int64_t)skeys[i];
  for (int64_t b = prev + 1; b <= cur; ++b) start[b] = (int32_t)i;
}

struct PermArgs {
  const int32_t* perm;
  int64_t M;
  const int32_t* cin[3];
  const float* oin[3];
  int32_t* cout[3];
  float* oout[3];
  int32_t* idx;
};

template <int D>
__global__ __launch_bounds__(kThreads) void k_permute(PermArgs a) {
  const int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (p >= a.M) return;
  const int32_t o = a.perm[p];
  a.idx[p] = o;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    a.cout[d][p] = a.cin[d][o];
    a.oout[d][p] = a.oin[d][o];
  }
}

// Interpolation work items: a lead bin (last-axis tile index a multiple of
// span) owns the points of its span consecutive bins along the last axis
// (adjacent keys, so one contiguous range), split into items of at most S
// points; the other bins own none.
__device__ __forceinline__ int span_points(const int32_t* bs, int64_t b, uint32_t nbins, int nt2,
                                           int span) {
  const int t2 = (int)((uint64_t)b % nbins) % nt2;
  if (t2 % span) return 0;
  const int m = min(span, nt2 - t2);
  return bs[b + m] - bs[b];
}

__global__ __launch_bounds__(kThreads) void k_sub_count(const int32_t* bs, int64_t nkeys,
                                                        uint32_t nbins, int nt2, int span, int S,
                                                        int32_t* cnt) {
  const int64_t b = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (b >= nkeys) return;
  const int n = span_points(bs, b, nbins, nt2, span);
  cnt[b] = (n + S - 1) / S;
}

__global__ __launch_bounds__(kThreads) void k_sub_emit(const int32_t* bs, const int32_t* soff,
                                                       int64_t nkeys, uint32_t nbins, int nt2,
                                                       int span, int S, int32_t* st,
                                                       int32_t* sb) {
  const int64_t b = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (b >= nkeys) return;
  const int n = span_points(bs, b, nbins, nt2, span);
  const int k = (n + S - 1) / S;
  for (int s = 0; s < k; ++s) {
    st[soff[b] + s] = (int32_t)b;
    sb[soff[b] + s] = bs[b] + s * S;
  }
}

// spread work items per tile: 1 for light tiles (also empty ones: every tile
// is written), ceil(E / S) for dense ones; dense tiles also get a flag
__global__ __launch_bounds__(kThreads) void k_item_count(const int32_t* ls, int64_t nbins, int S,
                                                         int32_t* cnt, int32_t* dense_cnt,
                                                         int32_t* dense_flag) {
  const int64_t b = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (b >= nbins) return;
  const int e = ls[b + 1] - ls[b];
  const int k = e <= S ? 1 : (e + S - 1) / S;
  cnt[b] = k;
  dense_cnt[b] = k > 1 ? k : 0;
  dense_flag[b] = k > 1 ? 1 : 0;
}

__global__ __launch_bounds__(kThreads) void k_item_emit(const int32_t* ls, int64_t nbins, int S,
                                                        const int32_t* item_off,
                                                        const int32_t* slot_off,
                                                        const int32_t* dense_off,
                                                        int32_t* it_tile, int32_t* it_begin,
                                                        int32_t* it_slot, int32_t* dn_tile,
                                                        int32_t* dn_slot0, int32_t* dn_count) {
  const int64_t b = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (b >= nbins) return;
  const int e = ls[b + 1] - ls[b];
  const int k = e <= S ? 1 : (e + S - 1) / S;
  for (int s = 0; s < k; ++s) {
    const int i = item_off[b] + s;
    it_tile[i] = (int32_t)b;
    it_begin[i] = ls[b] + s * S;
    it_slot[i] = k > 1 ? slot_off[b] + s : -1;
  }
  if (k > 1) {
    const int j = dense_off[b];
    dn_tile[j] = (int32_t)b;
    dn_slot0[j] = slot_off[b];
    dn_count[j] = k;
  }
}

// distinct tiles touched along one axis, in order of first appearance
__device__ inline int axis_tiles(int c0, int w, int nf, int tile, int* t) {
  int n = 0;
  for (int j = 0; j < w; ++j) {
    int c = c0 + j;
    if (c >= nf) c -= nf;
    const int tt = c / tile;
    bool seen = false;
    for (int q = 0; q < n; ++q) seen |= (t[q] == tt);
    if (!seen && n < 4) t[n++] = tt;  // tile >= w/2 bounds this at 3
  }
  return n;
}

struct ListArgs {
  const int32_t* cell[3];
  int nf[3];
  int tile[3];
  int nt[3];
  int w;
  int64_t M;
  const int32_t* set_off;
  int nsets;
  uint32_t nbins;
  const int32_t* eoff;
  int32_t* cnt;
  uint32_t* lkey;
  int32_t* lval;
};

template <int D>
__global__ __launch_bounds__(kThreads) void k_list_count(ListArgs a) {
  const int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (p >= a.M) return;
  int total = 1;
  for (int d = 0; d < D; ++d) {
    int t[4];
    total *= axis_tiles(a.cell[d][p], a.w, a.nf[d], a.tile[d], t);
  }
  a.cnt[p] = total;
}

template <int D>
__global__ __launch_bounds__(kThreads) void k_list_emit(ListArgs a) {
  const int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (p >= a.M) return;
  int t[3][4];
  int n[3] = {1, 1, 1};
  for (int d = 0; d < D; ++d) n[d] = axis_tiles(a.cell[d][p], a.w, a.nf[d], a.tile[d], t[d]);
  int e = a.eoff[p];
  // sorted positions of set k are [set_off[k], set_off[k+1]) like its
  // original indices (the set is the key's most significant part)
  const uint32_t kb = a.nsets > 1 ? (uint32_t)find_set(a.set_off, a.nsets, p) * a.nbins : 0u;
  for (int i0 = 0; i0 < n[0]; ++i0)
    for (int i1 = 0; i1 < (D > 1 ? n[1] : 1); ++i1)
      for (int i2 = 0; i2 < (D > 2 ? n[2] : 1); ++i2) {
        uint32_t key = (uint32_t)t[0][i0];
        if (D > 1) key = key * (uint32_t)a.nt[1] + (uint32_t)t[1][i1];
        if (D > 2) key = key * (uint32_t)a.nt[2] + (uint32_t)t[2][i2];
        a.lkey[e] = kb + key;
        a.lval[e] = (int32_t)p;
        ++e;
      }
}

int bits_for(int64_t nbins) {
  int b = 0;
  while (((int64_t)1 << b) < nbins) ++b;
  return b;
}

template <typename T>
int grow(T** ptr, int64_t* cap, int64_t need) {
  if (need <= *cap && *ptr) return 0;
  if (*ptr) HN_HIP(hipFree(*ptr));
  *ptr = nullptr;
  const int64_t n = need > 0 ? need : 1;
  HN_HIP(hipMalloc((void**)ptr, sizeof(T) * n));
  *cap = n;
  return 0;
}

}  // namespace

int scan_exclusive(const int32_t* in, int32_t* out, int64_t n, hipStream_t s, int64_t* total) {
  if (n <= 0) {
    if (total) *total = 0;
    return 0;
  }
  const int64_t nblk = (n + kScanChunk - 1) / kScanChunk;
  // the last input element is read before the (possibly in-place) scan
  int32_t last_in = 0;
  if (total)
    HN_HIP(hipMemcpyAsync(&last_in, in + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  StreamTemps tmp(s);
  int32_t* bsum = nullptr;
  HN_HIP(tmp.alloc(&bsum, nblk + 1));
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nblk), dim3(kThreads), 0, s, in, out, bsum, n);
  HN_LAUNCH_CHECK();
  if (nblk > 1) {
    int rc = scan_exclusive(bsum, bsum, nblk, s, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nblk), dim3(kThreads), 0, s, out, bsum, n);
    HN_LAUNCH_CHECK();
  } else {
    int32_t z = 0;
    HN_HIP(hipMemcpyAsync(bsum, &z, sizeof(int32_t), hipMemcpyHostToDevice, s));
  }
  if (total) {
    int32_t last_out = 0;
    HN_HIP(hipMemcpyAsync(&last_out, out + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HN_HIP(hipStreamSynchronize(s));
    *total = (int64_t)last_in + last_out;
  }
  HN_HIP(tmp.release());
  return 0;
}

// Spread work items from the list offsets (see k_item_count) and the
// partial-tile scratch for the dense tiles.
// dst[k] = src[k * stride] for k < n (the first scanned entry of each set)
__global__ void k_pick(const int32_t* src, int64_t stride, int n, int32_t* dst) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = src[(int64_t)k * stride];
}

// out[k] = exclusive-scan value at the first key of set k, out[nsets] = total
static int per_set_offsets(const hipnufft_plan_s* p, const int32_t* scanned, int64_t total,
                           std::vector<int64_t>& out) {
  const int K = p->nsets;
  out.assign(K + 1, 0);
  out[K] = total;
  if (K == 1) return 0;
  StreamTemps tmp(p->stream);
  int32_t* d = nullptr;
  HN_HIP(tmp.alloc(&d, K));
  hipLaunchKernelGGL(k_pick, dim3((unsigned)((K + 255) / 256)), dim3(256), 0, p->stream, scanned,
                     p->nbins, K, d);
  HN_LAUNCH_CHECK();
  std::vector<int32_t> h(K);
  HN_HIP(hipMemcpyAsync(h.data(), d, sizeof(int32_t) * K, hipMemcpyDeviceToHost, p->stream));
  HN_HIP(hipStreamSynchronize(p->stream));
  HN_HIP(tmp.release());
  for (int k = 0; k < K; ++k) out[k] = h[k];
  return 0;
}

static int build_spread_items(hipnufft_plan_s* p) {
  hipStream_t s = p->stream;
  const int64_t nkeys = p->nbins * p->nsets;
  const int S = p->spread_max;
  int rc;
  StreamTemps tmp(s);
  int32_t *cnt = nullptr, *dcnt = nullptr, *dflag = nullptr;
  HN_HIP(tmp.alloc(&cnt, nkeys));
  HN_HIP(tmp.alloc(&dcnt, nkeys));
  HN_HIP(tmp.alloc(&dflag, nkeys));
  const unsigned g = (unsigned)((nkeys + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(k_item_count, dim3(g), dim3(kThreads), 0, s, p->list_start, nkeys, S, cnt,
                     dcnt, dflag);
  HN_LAUNCH_CHECK();
  int64_t nitems = 0, nslots = 0, ndense = 0;
  if ((rc = scan_exclusive(cnt, cnt, nkeys, s, &nitems))) return rc;
  if ((rc = scan_exclusive(dcnt, dcnt, nkeys, s, &nslots))) return rc;
  if ((rc = scan_exclusive(dflag, dflag, nkeys, s, &ndense))) return rc;
  if ((rc = per_set_offsets(p, cnt, nitems, p->item_of_set))) return rc;
  if ((rc = per_set_offsets(p, dflag, ndense, p->dense_of_set))) return rc;
  if (nitems > p->items_cap || !p->it_tile) {
    int64_t c0 = 0, c1 = 0, c2 = 0;
    if ((rc = grow(&p->it_tile, &c0, nitems))) return rc;
    if ((rc = grow(&p->it_begin, &c1, nitems))) return rc;
    if ((rc = grow(&p->it_slot, &c2, nitems))) return rc;
    p->items_cap = c0;
  }
  if (ndense > p->dense_cap || !p->dn_tile) {
    int64_t c0 = 0, c1 = 0, c2 = 0;
    if ((rc = grow(&p->dn_tile, &c0, ndense))) return rc;
    if ((rc = grow(&p->dn_slot0, &c1, ndense))) return rc;
    if ((rc = grow(&p->dn_count, &c2, ndense))) return rc;
    p->dense_cap = c0;
  }
  int64_t tcells = 1;
  for (int d = 0; d < 3; ++d) tcells *= p->tile[d];
  const int64_t part_need = nslots * (int64_t)p->ntrans * tcells;
  if (part_need > p->part_cap) {
    if (p->part) HN_HIP(hipFree(p->part));
    p->part = nullptr;
    p->part_cap = 0;
    if (hipMalloc((void**)&p->part, sizeof(float2) * part_need) != hipSuccess)
      return set_error(HIPNUFFT_ERR_ALLOC, "cannot allocate the spreading scratch");
    p->part_cap = part_need;
  }
  hipLaunchKernelGGL(k_item_emit, dim3(g), dim3(kThreads), 0, s, p->list_start, nkeys, S, cnt,
                     dcnt, dflag, p->it_tile, p->it_begin, p->it_slot, p->dn_tile, p->dn_slot0,
                     p->dn_count);
  HN_LAUNCH_CHECK();
  p->nitems = nitems;
  p->nslots = nslots;
  p->ndense = ndense;
  HN_HIP(tmp.release());
  return 0;
}

int build_points(hipnufft_plan_s* p, int nsets, const int64_t* Ms, const float* const* xd) {
  const int D = p->ndim;
  const int o = 3 - D;  // first real axis in the padded arrays
  hipStream_t s = p->stream;
  int rc;
  if (nsets < 1 || !Ms) return set_error(HIPNUFFT_ERR_ARG, "need at least one point set");
  std::vector<int64_t> soff(nsets + 1, 0);
  for (int k = 0; k < nsets; ++k) {
    if (Ms[k] < 0) return set_error(HIPNUFFT_ERR_ARG, "negative number of points");
    soff[k + 1] = soff[k] + Ms[k];
  }
  const int64_t M = soff[nsets];
  if (M > INT32_MAX / 2) return set_error(HIPNUFFT_ERR_ARG, "too many points (max 2^30)");
  if (p->nbins * nsets >= ((int64_t)1 << 31))
    return set_error(HIPNUFFT_ERR_ARG, "too many point sets for this grid (sets x tiles >= 2^31)");
  for (int d = 0; d < D; ++d)
    if (!xd[d] && M > 0) return set_error(HIPNUFFT_ERR_ARG, "missing coordinate array");

  // From here on the plan's point arrays are rebuilt: the plan is unusable
  // (every transform returns HIPNUFFT_ERR_NOPTS) until the whole build has
  // succeeded, so a failure part-way never leaves stale or half-written
  // indices behind a usable plan.
  p->has_pts = 0;
  StreamTemps tmp(s);

  // persistent sorted arrays (one capacity for all of them)
  if (M > p->M_cap || !p->pt_idx) {
    for (int d = 0; d < 3; ++d) {
      if (p->pt_cell[d]) HN_HIP(hipFree(p->pt_cell[d]));
      if (p->pt_off[d]) HN_HIP(hipFree(p->pt_off[d]));
      p->pt_cell[d] = nullptr;
      p->pt_off[d] = nullptr;
    }
    if (p->pt_idx) HN_HIP(hipFree(p->pt_idx));
    p->pt_idx = nullptr;
    p->M_cap = 0;
    const int64_t cap = M > 0 ? M : 1;
    for (int d = 0; d < D; ++d) {
      HN_HIP(hipMalloc((void**)&p->pt_cell[d], sizeof(int32_t) * cap));
      HN_HIP(hipMalloc((void**)&p->pt_off[d], sizeof(float) * cap));
    }
    HN_HIP(hipMalloc((void**)&p->pt_idx, sizeof(int32_t) * cap));
    p->M_cap = cap;
  }
  // point-set offsets (host and device)
  if (nsets + 1 > p->set_off_cap || !p->set_off) {
    if (p->set_off) HN_HIP(hipFree(p->set_off));
    p->set_off = nullptr;
    p->set_off_cap = 0;
    HN_HIP(hipMalloc((void**)&p->set_off, sizeof(int32_t) * (nsets + 1)));
    p->set_off_cap = nsets + 1;
  }
  {
    std::vector<int32_t> so(soff.begin(), soff.end());
    HN_HIP(hipMemcpyAsync(p->set_off, so.data(), sizeof(int32_t) * (nsets + 1),
                          hipMemcpyHostToDevice, s));
    HN_HIP(hipStreamSynchronize(s));  // so is a stack buffer
  }
  p->nsets = nsets;
  p->set_off_h = soff;
  p->M = M;

  const int64_t nkeys = p->nbins * nsets;
  if (nkeys + 1 > p->keys_cap || !p->bin_start) {
    if (p->bin_start) HN_HIP(hipFree(p->bin_start));
    if (p->list_start) HN_HIP(hipFree(p->list_start));
    p->bin_start = p->list_start = nullptr;
    p->keys_cap = 0;
    HN_HIP(hipMalloc((void**)&p->bin_start, sizeof(int32_t) * (nkeys + 1)));
    HN_HIP(hipMalloc((void**)&p->list_start, sizeof(int32_t) * (nkeys + 1)));
    p->keys_cap = nkeys + 1;
  }

  if (M == 0) {
    HN_HIP(hipMemsetAsync(p->bin_start, 0, sizeof(int32_t) * (nkeys + 1), s));
    HN_HIP(hipMemsetAsync(p->list_start, 0, sizeof(int32_t) * (nkeys + 1), s));
    p->nsub = 0;
    p->nlist = 0;
    p->mask_rows = 0;
    p->sub_of_set.assign(nsets + 1, 0);
    int rc0 = build_spread_items(p);  // empty tiles still write zeros
    if (rc0) return rc0;
    HN_HIP(hipStreamSynchronize(s));
    p->has_pts = 1;
    return 0;
  }

  // temporaries (original order)
  uint32_t *ka = nullptr, *kb = nullptr;
  int32_t *va = nullptr, *vb = nullptr;
  int32_t* ctmp[3] = {nullptr, nullptr, nullptr};
  float* otmp[3] = {nullptr, nullptr, nullptr};
  HN_HIP(tmp.alloc(&ka, M));
  HN_HIP(tmp.alloc(&kb, M));
  HN_HIP(tmp.alloc(&va, M));
  HN_HIP(tmp.alloc(&vb, M));
  for (int d = 0; d < D; ++d) {
    HN_HIP(tmp.alloc(&ctmp[d], M));
    HN_HIP(tmp.alloc(&otmp[d], M));
  }

  FoldArgs fa{};
  for (int d = 0; d < D; ++d) {
    const int64_t nf = p->nf[o + d];
    fa.x[d] = xd[d];
    fa.scale[d] = (double)nf / (2.0 * M_PI);
    fa.nf[d] = (double)nf;
    fa.nfi[d] = (int)nf;
    fa.inv_tile[d] = 1.0 / (double)p->tile[o + d];
    fa.nt[d] = p->nt[o + d];
    fa.cell[d] = ctmp[d];
    fa.off[d] = otmp[d];
  }
  fa.halfw = 0.5 * p->w;
  fa.M = M;
  fa.set_off = p->set_off;
  fa.nsets = nsets;
  fa.nbins = (uint32_t)p->nbins;
  fa.key = ka;
  fa.val = va;
  const unsigned nb = (unsigned)((M + kThreads - 1) / kThreads);
  if (D == 1) hipLaunchKernelGGL(k_fold_bin<1>, dim3(nb), dim3(kThreads), 0, s, fa);
  if (D == 2) hipLaunchKernelGGL(k_fold_bin<2>, dim3(nb), dim3(kThreads), 0, s, fa);
  if (D == 3) hipLaunchKernelGGL(k_fold_bin<3>, dim3(nb), dim3(kThreads), 0, s, fa);
  HN_LAUNCH_CHECK();

  const int bits = bits_for(nkeys);
  if ((rc = radix_sort(ka, va, kb, vb, M, bits, s))) return rc;

  {
    const unsigned g = (unsigned)((M + 1 + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(k_bin_start, dim3(g), dim3(kThreads), 0, s, ka, M, nkeys, p->bin_start);
    HN_LAUNCH_CHECK();
  }
  {
    PermArgs pa{};
    pa.perm = va;
    pa.M = M;
    for (int d = 0; d < D; ++d) {
      pa.cin[d] = ctmp[d];
      pa.oin[d] = otmp[d];
      pa.cout[d] = p->pt_cell[d];
      pa.oout[d] = p->pt_off[d];
    }
    pa.idx = p->pt_idx;
    if (D == 1) hipLaunchKernelGGL(k_permute<1>, dim3(nb), dim3(kThreads), 0, s, pa);
    if (D == 2) hipLaunchKernelGGL(k_permute<2>, dim3(nb), dim3(kThreads), 0, s, pa);
    if (D == 3) hipLaunchKernelGGL(k_permute<3>, dim3(nb), dim3(kThreads), 0, s, pa);
    HN_LAUNCH_CHECK();
  }

  if (injected_failure("setpts"))
    return set_error(HIPNUFFT_ERR_ALLOC, "injected failure in setpts (HIPNUFFT_INJECT_FAILURE)");

  // interp work items: the points of interp_span bins along the last axis,
  // split into chunks of max_sub points
  {
    int32_t* cnt = nullptr;
    HN_HIP(tmp.alloc(&cnt, nkeys));
    const unsigned g = (unsigned)((nkeys + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(k_sub_count, dim3(g), dim3(kThreads), 0, s, p->bin_start, nkeys,
                       (uint32_t)p->nbins, p->nt[2], p->interp_span, p->max_sub, cnt);
    HN_LAUNCH_CHECK();
    int64_t nsub = 0;
    if ((rc = scan_exclusive(cnt, cnt, nkeys, s, &nsub))) return rc;
    if ((rc = per_set_offsets(p, cnt, nsub, p->sub_of_set))) return rc;
    if (nsub > p->sub_cap || !p->sub_tile) {
      int64_t c1 = 0, c2 = 0;
      if ((rc = grow(&p->sub_tile, &c1, nsub))) return rc;
      if ((rc = grow(&p->sub_begin, &c2, nsub))) return rc;
      p->sub_cap = c1;
    }
    p->nsub = nsub;
    hipLaunchKernelGGL(k_sub_emit, dim3(g), dim3(kThreads), 0, s, p->bin_start, cnt, nkeys,
                       (uint32_t)p->nbins, p->nt[2], p->interp_span, p->max_sub, p->sub_tile,
                       p->sub_begin);
    HN_LAUNCH_CHECK();
  }
  if ((rc = build_interp_masks(p))) return rc;

  // spread lists: one entry per (point, touched tile), stably sorted by tile
  {
    ListArgs la{};
    for (int d = 0; d < D; ++d) {
      la.cell[d] = p->pt_cell[d];
      la.nf[d] = (int)p->nf[o + d];
      la.tile[d] = p->tile[o + d];
      la.nt[d] = p->nt[o + d];
    }
    la.w = p->w;
    la.M = M;
    la.set_off = p->set_off;
    la.nsets = nsets;
    la.nbins = (uint32_t)p->nbins;
    la.cnt = vb;  // reuse temporaries: vb = counts, then offsets
    if (D == 1) hipLaunchKernelGGL(k_list_count<1>, dim3(nb), dim3(kThreads), 0, s, la);
    if (D == 2) hipLaunchKernelGGL(k_list_count<2>, dim3(nb), dim3(kThreads), 0, s, la);
    if (D == 3) hipLaunchKernelGGL(k_list_count<3>, dim3(nb), dim3(kThreads), 0, s, la);
    HN_LAUNCH_CHECK();
    int64_t L = 0;
    if ((rc = scan_exclusive(vb, vb, M, s, &L))) return rc;
    if (L > INT32_MAX / 2) return set_error(HIPNUFFT_ERR_ARG, "spread lists too large");
    uint32_t *lk0 = nullptr, *lk1 = nullptr;
    int32_t* lv1 = nullptr;
    HN_HIP(tmp.alloc(&lk0, L));
    HN_HIP(tmp.alloc(&lk1, L));
    HN_HIP(tmp.alloc(&lv1, L));
    int64_t c = p->list_cap;
    if ((rc = grow(&p->list_pts, &c, L))) return rc;
    p->list_cap = c;
    p->nlist = L;
    la.eoff = vb;
    la.lkey = lk0;
    la.lval = p->list_pts;
    if (D == 1) hipLaunchKernelGGL(k_list_emit<1>, dim3(nb), dim3(kThreads), 0, s, la);
    if (D == 2) hipLaunchKernelGGL(k_list_emit<2>, dim3(nb), dim3(kThreads), 0, s, la);
    if (D == 3) hipLaunchKernelGGL(k_list_emit<3>, dim3(nb), dim3(kThreads), 0, s, la);
    HN_LAUNCH_CHECK();
    if ((rc = radix_sort(lk0, p->list_pts, lk1, lv1, L, bits, s))) return rc;
    const unsigned g = (unsigned)((L + 1 + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(k_bin_start, dim3(g), dim3(kThreads), 0, s, lk0, L, nkeys, p->list_start);
    HN_LAUNCH_CHECK();
  }

  if ((rc = build_spread_items(p))) return rc;

  HN_HIP(tmp.release());
  HN_HIP(hipStreamSynchronize(s));
  p->has_pts = 1;
  return 0;
}

}  // namespace hn
