// Spreading (type 1) and interpolation (type 2) with the exponential of
// semicircle kernel, tiled through LDS.
//
// Replaces the spread / interp stages of cufinufft's Plan.execute
// (operators/interfaces/cufinufft.py:127-133) and of finufft
// (operators/interfaces/finufft.py:64-76).
//
// Spreading is output-driven: one workgroup owns one tile of the fine grid
// (or one slice of a dense tile's list) for a group of transforms, sums the
// footprints of every point that touches the tile (the spread lists built at
// setpts, one entry per point/tile pair) in registers, each thread owning a
// few cells, and writes the finished tile to HBM exactly once with coalesced
// stores -- no atomics, no zero-fill pass over the grid, and a fixed
// summation order.
//
// Interpolation stages the tile plus a w-wide halo in LDS and gathers each
// sorted point's w^D footprint from LDS.  Dense bins are split into work
// items of at most max_sub points (the radial k-space centre).

#include <hip/hip_runtime.h>

#include "internal.h"

namespace hn {
namespace {

__device__ __forceinline__ float es_eval(float u, float inv_hw, float beta) {
  const float z = u * inv_hw;
  const float a = 1.0f - z * z;
  return a > 0.0f ? __expf(beta * (__fsqrt_rn(a) - 1.0f)) : 0.0f;
}

// derivative d/du of the kernel, clamped where it is singular (|z| -> 1,
// where the kernel itself is exp(-beta) ~ 1e-7)
__device__ __forceinline__ float es_deriv(float u, float inv_hw, float beta) {
  const float z = u * inv_hw;
  const float a = 1.0f - z * z;
  if (a <= 1e-6f) return 0.0f;
  const float s = __fsqrt_rn(a);
  return __expf(beta * (s - 1.0f)) * (-beta * z / s) * inv_hw;
}

struct SpreadArgs {
  PtsGeom g;
  const int32_t* list_pts;
  const int32_t* list_start;
  const int32_t* it_tile;   // work items (dense tiles split)
  const int32_t* it_begin;
  const int32_t* it_slot;
  const int32_t* dn_tile;   // dense tiles: partial slots to sum
  const int32_t* dn_slot0;
  const int32_t* dn_count;
  float2* part;             // [slot][ntrans][tile cells]
  int ntrans;
  int smax;
  int item0, dn0;           // first spread item / dense tile of the launch
  const float2* in;         // sorted order [t_local][M]
  int64_t in_stride;
  float2* grid;
  int64_t grid_stride;
  int k0, ts;               // SetRange of the launch
  int G;
};

struct InterpArgs {
  PtsGeom g;
  const int32_t* sub_tile;  // key (set * nbins + tile) per work item
  const int32_t* sub_begin;
  const int32_t* bin_start;
  int max_sub;
  int sub0;                 // first work item of the launch
  const float2* grid;
  int64_t grid_stride;
  float2* out;
  const float2* obs;
  const float* density;
  float* gtraj;
  float dscale[3];  // radians -> cells factor per axis (gradient)
  int k0, ts;       // SetRange of the launch
  int tg0, G;       // transforms [tg0, tg0 + G) of each set (gradient groups)
  float scale;
  int mode;
  // staging of the tile + halo region: magic multipliers ceil(2^32 / R) for
  // the region extent per axis (exact divisions by umulhi), whether one
  // conditional wrap per axis suffices, lanes per staged row (the last-axis
  // extent, at most the workgroup) and its magic multiplier, rows per pass
  uint32_t rmag[3];
  int wrap1;
  int stage_lanes, stage_rows;
  uint32_t lane_mag;
  int span;  // bins along the last axis per work item (region: span tiles + halo)
  int stage_chunks;  // 16-cell chunks a region row of the last axis can meet
  int mask_stage;    // stage only the chunks the points read (needs grid_bytes)
  int grid_bytes;    // bytes of one transform's grid, < 2^31
  const uint16_t* sub_mask;  // per work item and region row: the chunks read
  int lds_pad;               // k_interp_stream: LDS cells after each region row
  uint32_t lds_row_mag;      // and the magic multiplier of the padded row length
};

// dword 3 of a raw buffer resource on gfx9-family GPUs (gfx950 included):
// 32-bit data format, no swizzle, out-of-range offsets return zero
constexpr int kBufferWord3 = 0x00020000;

__device__ __forceinline__ int mdiv(int c, uint32_t magic) {
  return (int)__umulhi((uint32_t)c, magic);
}

// offset of (set k, transform t_local, original index i) in a k-space array
// laid out per set as [ts][M_k]
__device__ __forceinline__ int64_t ksp_index(const PtsGeom& g, int ts, int k, int t, int i) {
  const int64_t o = g.set_off[k];
  const int64_t mk = g.set_off[k + 1] - o;
  return ts * o + t * mk + (i - o);
}

template <int D>
__device__ __forceinline__ void tile_coords(int key, const int* nt, int* tc) {
#pragma unroll
  for (int d = D - 1; d >= 0; --d) {
    tc[d] = key % nt[d];
    key /= nt[d];
  }
}

// ---------------------------------------------------------------------------
// spreading: owner-computes gather, no atomics
// ---------------------------------------------------------------------------
// Cell ownership: the tile is a set of columns along the last (contiguous)
// axis; thread (c, s) owns kZ consecutive cells z0 = s * kZ ... of column c
// and keeps their sums for the G transforms in registers.  With 64 columns
// (8 x 8 x 16 tiles) a wave holds one segment s of every column, so a point
// whose footprint misses the segment is skipped by the whole wave at once,
// and the column weight is formed once per point for the kZ cells.
//
// The tile's list entries are staged through LDS kSpreadChunk at a time
// (footprint start relative to the tile, the w weights per axis, the G
// strengths); every thread then walks the staged points in list order, so
// each cell's sum is formed in the same order on every run: deterministic
// bit for bit, and thousands of radial spokes crossing the k-space centre
// cost no contention.
constexpr int kZ = 4;  // cells per thread along the last axis
constexpr int kSpreadChunk = kThreads;

template <int D, int G>
__global__ __launch_bounds__(kThreads) void k_spread(SpreadArgs a) {
  extern __shared__ float smem[];
  const int W = a.g.w;
  const int PW = D + D * W + 2 * G;  // floats per staged point
  const int item = a.item0 + blockIdx.x;
  const int key = a.it_tile[item];
  const int set = key / a.g.nbins;
  const int tile = key - set * a.g.nbins;
  const int slot = a.it_slot[item];
  const int t0 = blockIdx.y * G;         // first transform of the set handled here
  const int tb = (set - a.k0) * a.ts;    // chunk transform of the set's first
  const int Gv = min(G, a.ts - t0);
  const float inv_hw = 2.0f / W, beta = a.g.beta;

  int tc[3], org[3], ext[3], td[3], nf[3];
  tile_coords<D>(tile, a.g.nt, tc);
  int TC = 1;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    td[d] = a.g.tile[d];
    nf[d] = a.g.nf[d];
    org[d] = tc[d] * td[d];
    ext[d] = min(td[d], nf[d] - org[d]);
    TC *= td[d];
  }
  constexpr int DL = D - 1;  // last axis
  const int zl = td[DL];
  const int ncol = TC / zl;
  const int col = threadIdx.x % ncol;
  const int z0 = (threadIdx.x / ncol) * kZ;
  int cc[3] = {0, 0, 0};
  bool colown = true;
  {
    int r = col;
#pragma unroll
    for (int d = DL - 1; d >= 0; --d) {
      cc[d] = r % td[d];
      r /= td[d];
      colown = colown && cc[d] < ext[d];
    }
  }
  bool own[kZ];
#pragma unroll
  for (int j = 0; j < kZ; ++j) own[j] = colown && z0 + j < ext[DL];

  float2 acc[kZ][G];
#pragma unroll
  for (int j = 0; j < kZ; ++j)
#pragma unroll
    for (int g = 0; g < G; ++g) acc[j][g] = make_float2(0.f, 0.f);

  const int e0 = a.it_begin[item];
  const int e1 = min(e0 + a.smax, a.list_start[key + 1]);
  for (int eb = e0; eb < e1; eb += kSpreadChunk) {
    const int n = min(kSpreadChunk, e1 - eb);
    if ((int)threadIdx.x < n) {
      float* sp = smem + threadIdx.x * PW;
      const int p = a.list_pts[eb + threadIdx.x];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const float off = a.g.off[d][p];
        reinterpret_cast<int*>(sp)[d] = a.g.cell[d][p] - org[d];  // raw start, mod nf
        for (int j = 0; j < W; ++j) sp[D + d * W + j] = es_eval(off - (float)j, inv_hw, beta);
      }
      // strengths are in sorted-point order (density already applied), so
      // the entries of one tile read nearly contiguous runs
      float* sv = sp + D + D * W;
#pragma unroll
      for (int g = 0; g < G; ++g) {
        float2 v = make_float2(0.f, 0.f);
        if (g < Gv) v = a.in[(int64_t)(t0 + g) * a.in_stride + p];
        sv[2 * g] = v.x;
        sv[2 * g + 1] = v.y;
      }
    }
    __syncthreads();
    for (int q = 0; q < n; ++q) {
      const float* sp = smem + q * PW;
      const int* si = reinterpret_cast<const int*>(sp);
      // last axis: which of the kZ cells the footprint covers
      const int lsz = si[DL];
      int dz[kZ];
      bool hz[kZ];
      bool any = false;
#pragma unroll
      for (int j = 0; j < kZ; ++j) {
        int da = z0 + j - lsz;  // owned cells satisfy z0 + j < nf: one wrap each way
        if (da < 0) da += nf[DL];
        if (da >= nf[DL]) da -= nf[DL];
        hz[j] = own[j] && da < W;
        dz[j] = hz[j] ? da : 0;
        any = any || hz[j];
      }
      if (!any) continue;  // uniform across the wave for interior 3D tiles
      float wc = 1.0f;
      bool hc = true;
#pragma unroll
      for (int d = 0; d < DL; ++d) {
        int da = cc[d] - si[d];
        if (da < 0) da += nf[d];
        if (da >= nf[d]) da -= nf[d];
        hc = hc && da < W;
        wc *= sp[D + d * W + (hc ? da : 0)];  // index stays inside the record
      }
      if (!hc) wc = 0.0f;
      const float* wl = sp + D + DL * W;
      const float* sv = sp + D + D * W;
      float2 v[G];
#pragma unroll
      for (int g = 0; g < G; ++g) v[g] = make_float2(sv[2 * g], sv[2 * g + 1]);
#pragma unroll
      for (int j = 0; j < kZ; ++j) {
        if (!hz[j]) continue;
        const float f = wc * wl[dz[j]];
#pragma unroll
        for (int g = 0; g < G; ++g) {
          acc[j][g].x += f * v[g].x;
          acc[j][g].y += f * v[g].y;
        }
      }
    }
    __syncthreads();
  }

  // tile-local linear index of cell (cc, z0 + j) is col * zl + z0 + j
#pragma unroll
  for (int j = 0; j < kZ; ++j) {
    if (!own[j]) continue;
    if (slot >= 0) {  // partial tile of a dense tile: summed by k_spread_reduce
      const int r = col * zl + z0 + j;
#pragma unroll
      for (int g = 0; g < G; ++g)
        if (g < Gv) a.part[((int64_t)slot * a.ntrans + tb + t0 + g) * TC + r] = acc[j][g];
      continue;
    }
    int64_t gi = 0;
#pragma unroll
    for (int d = 0; d < DL; ++d) gi = gi * nf[d] + org[d] + cc[d];
    gi = gi * nf[DL] + org[DL] + z0 + j;
#pragma unroll
    for (int g = 0; g < G; ++g)
      if (g < Gv) a.grid[(int64_t)(tb + t0 + g) * a.grid_stride + gi] = acc[j][g];
  }
}

// Block-owner spreading for 3D tiles with edges of 4, 8 or 16 cells (4, 8
// or 16 blocks of 4 x 4 x 4 cells, at most 4 along an axis): lane l owns
// cell (l / 16, l / 4 % 4, l % 4) of each of the wave's blocks (wave v owns
// blocks v, v + 4, ...).  When an entry is staged, its thread records which
// blocks of the tile the footprint touches (a bit mask) and, per axis and
// block along it, where the lane weights of that block start in the entry's
// weight rows (padded with 3 zeros on each side).  A wave selects the
// entries that touch its blocks with one ballot per 64 entries and walks
// only those; a touched block costs three weight reads and the products per
// lane.  A cell outside the footprint adds f = 0 times
// the strength: its sum is unchanged, and every cell's sum is formed in
// list order with the products (w0 * w1) * w2 of the cell-owner kernel:
// the same bits.
constexpr int kBlkE = 4;            // block edge
constexpr int kBlkPerWave = 4;      // at most 16 blocks per tile
constexpr int kSpreadWaves = 4;     // k_spread_mfma waves per 8 transforms (4 or 8)
constexpr int kBlkChunk = 128;      // staged entries per round
constexpr int kBlkPad = 3;          // zeros before / after each weight row
constexpr int kBlkHead = 4;         // mask + one word of 4 byte offsets per axis

// floats per staged entry of the block kernel: header (mask, 3 words of
// byte offsets), G strengths, 3 padded weight rows; a multiple of 4 so that
// every record starts 16-byte aligned
__host__ __device__ constexpr int blk_record_floats(int W, int G) {
  return (kBlkHead + 2 * G + 3 * (W + 2 * kBlkPad) + 3) & ~3;
}

// Where one list entry of a tile sits, as loaded from HBM for the block and
// matrix kernels: its sorted point, the point's offsets and its footprint
// start relative to the tile origin (raw, mod nf).
struct BlkPos {
  float off[3];
  int st[3];
  int p;
};

template <int D = 3>
__device__ __forceinline__ void load_blk_pos(const SpreadArgs& a, int p, const int* org,
                                             BlkPos& r) {
  r.p = p;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    r.off[d] = 0.0f;
    r.st[d] = 0;
  }
#pragma unroll
  for (int d = 0; d < D; ++d) {
    r.off[d] = a.g.off[d][p];
    r.st[d] = a.g.cell[d][p] - org[d];
  }
}

// The staged record of an entry at sp: header (touched-block mask, per axis
// the byte offsets of each block's first lane weight), G strengths and the
// three padded weight rows.  The strengths are loaded first so that their
// round trip overlaps the weight evaluation.
template <int G>
__device__ __forceinline__ void write_blk_entry(float* sp, const BlkPos& r, const SpreadArgs& a,
                                                int t0, int Gv, int WP, const int* ext,
                                                const int* nf, const int* nb, float inv_hw,
                                                float beta) {
  const int W = a.g.w;
  float2 v[G];
#pragma unroll
  for (int g = 0; g < G; ++g)
    v[g] = g < Gv ? a.in[(int64_t)(t0 + g) * a.in_stride + r.p] : make_float2(0.f, 0.f);
  uint32_t hit[3] = {0u, 0u, 0u};  // per axis: blocks along it the footprint meets
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int st = r.st[d];
    float* wr = sp + kBlkHead + 2 * G + d * WP;
    for (int j = 0; j < kBlkPad; ++j) wr[j] = 0.0f;
    for (int j = 0; j < W; ++j) wr[kBlkPad + j] = es_eval(r.off[d] - (float)j, inv_hw, beta);
    for (int j = 0; j < kBlkPad; ++j) wr[kBlkPad + W + j] = 0.0f;
    // offset of the block's first cell from the footprint start, in
    // [-kBlkPad, W) when the block meets the footprint (nf >= W +
    // kBlkPad; a block inside the grid has org + its origin < nf, so
    // one wrap each way suffices); past the grid end: not met
    uint32_t packed = 0u;
    for (int bd = 0; bd < nb[d]; ++bd) {
      const int bo = bd * kBlkE;
      int x = bo - st;
      if (x < 0) x += nf[d];
      if (x >= nf[d]) x -= nf[d];
      if (x >= nf[d] - kBlkPad) x -= nf[d];
      if (bo < ext[d] && x < W) hit[d] |= 1u << bd;
      const int o = kBlkHead + 2 * G + d * WP + kBlkPad + min(x, W - 1);  // < 256
      packed |= (uint32_t)o << (8 * bd);
    }
    reinterpret_cast<uint32_t*>(sp)[1 + d] = packed;
  }
  uint32_t mask = 0u;
  for (int b0 = 0; b0 < nb[0]; ++b0)
    for (int b1 = 0; b1 < nb[1]; ++b1)
      for (int b2 = 0; b2 < nb[2]; ++b2)
        if ((hit[0] >> b0) & (hit[1] >> b1) & (hit[2] >> b2) & 1u)
          mask |= 1u << ((b0 * nb[1] + b1) * nb[2] + b2);
  reinterpret_cast<uint32_t*>(sp)[0] = mask;
  float* sv = sp + kBlkHead;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    sv[2 * g] = v[g].x;
    sv[2 * g + 1] = v[g].y;
  }
}

// Kernel weights of one axis of a matrix-kernel record: wr[x] for the td
// tile cells x (zero outside the footprint and past the grid end); returns
// the bits of the blocks of `edge` cells along the axis that the footprint
// meets.  When nf >= td + W the footprint's copies [st + k nf, st + k nf + W)
// lie nf apart, so one of them meets the tile: it starts at tile cell s, the
// representative of st mod nf in (-W, td), and only its W cells are
// evaluated.  Smaller grids take the cell-by-cell form.  A4: wr is 16-byte
// aligned (the zero fill is written four cells at a time).
template <bool A4>
__device__ __forceinline__ uint32_t axis_weights(float* wr, int st, float off, int nf, int td,
                                                 int ext, int W, int edge, float inv_hw,
                                                 float beta) {
  uint32_t hit = 0u;
  if (nf >= td + W) {
    int s = st < 0 ? st + nf : st;
    if (s >= td) s -= nf;
    if constexpr (A4) {
      for (int x = 0; x < td; x += 4)
        *reinterpret_cast<float4*>(wr + x) = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
      for (int x = 0; x < td; ++x) wr[x] = 0.0f;
    }
    const int lo = max(s, 0), hi = min(s + W, ext);  // tile cells [lo, hi)
    for (int x = lo; x < hi; ++x) wr[x] = es_eval(off - (float)(x - s), inv_hw, beta);
    if (lo < hi) hit = (2u << ((hi - 1) / edge)) - (1u << (lo / edge));
    return hit;
  }
  // cell x of the tile is footprint cell (x - st) mod nf, st in (-nf, nf)
  for (int x = 0; x < td; ++x) {
    int j = x - st;
    if (j < 0) j += nf;
    if (j >= nf) j -= nf;
    float w = 0.0f;
    if (x < ext && j < W) {
      w = es_eval(off - (float)j, inv_hw, beta);
      hit |= 1u << (x / edge);
    }
    wr[x] = w;
  }
  return hit;
}

// Record of an entry for the matrix-core kernel at sp: touched-block mask,
// G strengths and, per axis, the weight of every tile cell along it (zero
// outside the footprint and past the grid end).  A lane's weights for a
// block then sit at fixed offsets from the record: no header lookup before
// the reads.  Same weights as the block record, so the same products.
__host__ __device__ constexpr int mfma_record_floats(int G, int tiles_sum) {
  return (kBlkHead + 2 * G + tiles_sum + 3) & ~3;
}

template <int D, int G>
__device__ __forceinline__ void write_mfma_entry(float* sp, const BlkPos& r, const SpreadArgs& a,
                                                 int t0, int Gv, const int* td, const int* ext,
                                                 const int* nf, const int* nb, float inv_hw,
                                                 float beta) {
  const int W = a.g.w;
  float2 v[G];
#pragma unroll
  for (int g = 0; g < G; ++g)
    v[g] = g < Gv ? a.in[(int64_t)(t0 + g) * a.in_stride + r.p] : make_float2(0.f, 0.f);
  uint32_t hit[3] = {1u, 1u, 1u};
  float* wr = sp + kBlkHead + 2 * G;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    // blocks of 4 cells along each axis, 16 along the last in 2D
    hit[d] = axis_weights<(kBlkHead + 2 * G) % 4 == 0>(
        wr, r.st[d], r.off[d], nf[d], td[d], ext[d], W,
        D == 2 && d == 1 ? 4 * kBlkE : kBlkE, inv_hw, beta);
    wr += td[d];
  }
  uint32_t mask = 0u;
  for (int b0 = 0; b0 < nb[0]; ++b0)
    for (int b1 = 0; b1 < nb[1]; ++b1)
      for (int b2 = 0; b2 < nb[2]; ++b2)
        if ((hit[0] >> b0) & (hit[1] >> b1) & (hit[2] >> b2) & 1u)
          mask |= 1u << ((b0 * nb[1] + b1) * nb[2] + b2);
  reinterpret_cast<uint32_t*>(sp)[0] = mask;
  float* sv = sp + kBlkHead;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    sv[2 * g] = v[g].x;
    sv[2 * g + 1] = v[g].y;
  }
}

// Stage list entries [eb, eb + n) of a tile for the block and matrix
// kernels: thread q < n writes entry eb + q's record at smem + q * PW.
template <int G>
__device__ __forceinline__ void stage_blk_entries(const SpreadArgs& a, float* smem, int eb, int n,
                                                  int PW, int WP, const int* org, const int* ext,
                                                  const int* nf, const int* nb, int t0, int Gv,
                                                  float inv_hw, float beta) {
  if ((int)threadIdx.x < n) {
    BlkPos r;
    load_blk_pos(a, a.list_pts[eb + threadIdx.x], org, r);
    write_blk_entry<G>(smem + threadIdx.x * PW, r, a, t0, Gv, WP, ext, nf, nb, inv_hw, beta);
  }
}

template <int G>
__global__ __launch_bounds__(kThreads) void k_spread_blocks(SpreadArgs a) {
  extern __shared__ float smem[];
  const int W = a.g.w;
  const int WP = W + 2 * kBlkPad;
  const int PW = blk_record_floats(W, G);
  const int item = a.item0 + blockIdx.x;
  const int key = a.it_tile[item];
  const int set = key / a.g.nbins;
  const int tile = key - set * a.g.nbins;
  const int slot = a.it_slot[item];
  const int t0 = blockIdx.y * G;
  const int tb = (set - a.k0) * a.ts;
  const int Gv = min(G, a.ts - t0);
  const float inv_hw = 2.0f / W, beta = a.g.beta;

  int tc[3], org[3], ext[3], td[3], nf[3], nb[3];
  tile_coords<3>(tile, a.g.nt, tc);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    td[d] = a.g.tile[d];
    nf[d] = a.g.nf[d];
    org[d] = tc[d] * td[d];
    ext[d] = min(td[d], nf[d] - org[d]);
    nb[d] = td[d] / kBlkE;
  }
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / 64);
  const int lane = (int)threadIdx.x & 63;
  const int lc[3] = {lane >> 4, (lane >> 2) & 3, lane & 3};
  const int nbw = nb[0] * nb[1] * nb[2] / 4;  // blocks per wave (uniform)
  // the wave's blocks: index along each axis, mask bit, ownership of the
  // lane's cell; a block wholly past the grid end is never touched
  int bidx[kBlkPerWave][3];
  uint32_t bbit[kBlkPerWave];
  bool own[kBlkPerWave];
#pragma unroll
  for (int bi = 0; bi < kBlkPerWave; ++bi) {
    const int b = wave + 4 * bi;
    bidx[bi][2] = b % nb[2];
    bidx[bi][1] = (b / nb[2]) % nb[1];
    bidx[bi][0] = b / (nb[2] * nb[1]);
    bool live = bi < nbw;
#pragma unroll
    for (int d = 0; d < 3; ++d) live = live && bidx[bi][d] * kBlkE < ext[d];
    bbit[bi] = live ? 1u << b : 0u;
    own[bi] = live;
#pragma unroll
    for (int d = 0; d < 3; ++d) own[bi] = own[bi] && bidx[bi][d] * kBlkE + lc[d] < ext[d];
  }

  float2 acc[kBlkPerWave][G];
#pragma unroll
  for (int bi = 0; bi < kBlkPerWave; ++bi)
#pragma unroll
    for (int g = 0; g < G; ++g) acc[bi][g] = make_float2(0.f, 0.f);

  const int e0 = a.it_begin[item];
  const int e1 = min(e0 + a.smax, a.list_start[key + 1]);
  for (int eb = e0; eb < e1; eb += kBlkChunk) {
    const int n = min(kBlkChunk, e1 - eb);
    stage_blk_entries<G>(a, smem, eb, n, PW, WP, org, ext, nf, nb, t0, Gv, inv_hw, beta);
    __syncthreads();
    // per 64 staged entries: lane i reads the header of entry base + i, a
    // ballot marks the entries that touch one of the wave's blocks, and the
    // wave walks only those, lowest first (list order kept)
    uint32_t wbits = 0u;
#pragma unroll
    for (int bi = 0; bi < kBlkPerWave; ++bi) wbits |= bbit[bi];
    for (int base = 0; base < n; base += 64) {
      const int qi = base + lane;
      uint4 h = make_uint4(0u, 0u, 0u, 0u);
      if (qi < n) h = *reinterpret_cast<const uint4*>(smem + qi * PW);
      uint64_t todo = __ballot((h.x & wbits) != 0u);
      while (todo) {  // uniform
        const int i = __builtin_ctzll(todo);
        todo &= todo - 1;
        const float* sp = smem + (base + i) * PW;
        const uint32_t mask = (uint32_t)__builtin_amdgcn_readlane((int)h.x, i);
        const uint32_t pk[3] = {(uint32_t)__builtin_amdgcn_readlane((int)h.y, i),
                                (uint32_t)__builtin_amdgcn_readlane((int)h.z, i),
                                (uint32_t)__builtin_amdgcn_readlane((int)h.w, i)};
        float2 v[G];
#pragma unroll
        for (int g = 0; g < G; ++g) v[g] = reinterpret_cast<const float2*>(sp + kBlkHead)[g];
#pragma unroll
        for (int bi = 0; bi < kBlkPerWave; ++bi) {
          if (!(mask & bbit[bi])) continue;  // uniform
          const int ox = (pk[0] >> (8 * bidx[bi][0])) & 255;
          const int oy = (pk[1] >> (8 * bidx[bi][1])) & 255;
          const int oz = (pk[2] >> (8 * bidx[bi][2])) & 255;
          const float wx = sp[ox + lc[0]];
          const float wy = sp[oy + lc[1]];
          const float wz = sp[oz + lc[2]];
          const float f = (wx * wy) * wz;
#pragma unroll
          for (int g = 0; g < G; ++g) {
            acc[bi][g].x += f * v[g].x;
            acc[bi][g].y += f * v[g].y;
          }
        }
      }
    }
    __syncthreads();
  }

  const int TC = td[0] * td[1] * td[2];
#pragma unroll
  for (int bi = 0; bi < kBlkPerWave; ++bi) {
    if (!own[bi]) continue;
    const int c0 = bidx[bi][0] * kBlkE + lc[0];
    const int c1 = bidx[bi][1] * kBlkE + lc[1];
    const int c2 = bidx[bi][2] * kBlkE + lc[2];
    if (slot >= 0) {  // partial tile of a dense tile: summed by k_spread_reduce
      const int r = (c0 * td[1] + c1) * td[2] + c2;
#pragma unroll
      for (int g = 0; g < G; ++g)
        if (g < Gv) a.part[((int64_t)slot * a.ntrans + tb + t0 + g) * TC + r] = acc[bi][g];
      continue;
    }
    const int64_t gi = ((int64_t)(org[0] + c0) * nf[1] + org[1] + c1) * nf[2] + org[2] + c2;
#pragma unroll
    for (int g = 0; g < G; ++g)
      if (g < Gv) a.grid[(int64_t)(tb + t0 + g) * a.grid_stride + gi] = acc[bi][g];
  }
}

// Matrix-core spreading for the tiles of the block kernel, 8 * GH transforms
// per workgroup of 4 * GH waves.  For one 4 x 4 x 4 block and four entries
// touching it, the sums of the block's cells are a 64 x 4 by 4 x 16 product
// per 8 transforms: cells by entries (the weight products w0 w1 w2) times
// entries by values (the strengths, 8 transforms x re / im).  Each slice of
// 16 cells (fixed first offset) is one v_mfma_f32_16x16x4f32 per 8
// transforms: lane l supplies A[l % 16][l / 16] (cell l % 16 of the slice,
// entry l / 16), B[l / 16][l % 16] (entry l / 16, value l % 16) and holds
// D[4 (l / 16) + r][l % 16] in register r.  With GH = 2 the A operand (and
// the staging and the entry selection behind it) serves both halves of 16
// transforms.  A wave takes the entries touching each of its blocks four at
// a time in list order (after each staging round a wave lists, per block,
// the staged entries touching it: a ballot and a lane prefix count per 64
// entries), so every cell's sum has the same order on every run and for
// both GH: deterministic, equal to the other kernels to rounding.
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int D, int GH, int NWAVE>
__global__ __launch_bounds__(64 * NWAVE, 4 / GH) void k_spread_mfma(SpreadArgs a) {
  constexpr int G = 8 * GH;              // transforms per workgroup
  constexpr int BPW = 16 / NWAVE;        // blocks per wave (16 blocks at most)
  // blocks: 4 x 4 x 4 cells in 3D, 4 x 16 in 2D (the last axis holds the
  // 16 cells of an MFMA row set either way)
  constexpr int BL = D == 3 ? kBlkE : 4 * kBlkE;  // block edge along the last axis
  extern __shared__ float smem[];
  const int PW = mfma_record_floats(G, a.g.tile[0] + a.g.tile[1] + (D == 3 ? a.g.tile[2] : 0));
  const int item = a.item0 + blockIdx.x;
  const int key = a.it_tile[item];
  const int set = key / a.g.nbins;
  const int tile = key - set * a.g.nbins;
  const int slot = a.it_slot[item];
  const int t0 = blockIdx.y * G;
  const int tb = (set - a.k0) * a.ts;
  const int Gv = min(G, a.ts - t0);
  const float inv_hw = 2.0f / a.g.w, beta = a.g.beta;

  int tc[3] = {0, 0, 0}, org[3] = {0, 0, 0}, ext[3] = {1, 1, 1}, td[3] = {1, 1, 1},
      nf[3] = {1, 1, 1}, nb[3] = {1, 1, 1};
  tile_coords<D>(tile, a.g.nt, tc);
#pragma unroll
  for (int d = 0; d < D; ++d) {
    td[d] = a.g.tile[d];
    nf[d] = a.g.nf[d];
    org[d] = tc[d] * td[d];
    ext[d] = min(td[d], nf[d] - org[d]);
    nb[d] = td[d] / (d == D - 1 ? BL : kBlkE);
  }
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / 64);
  const int lane = (int)threadIdx.x & 63;
  const int kq = lane >> 4;               // entry of the group (A column, B row)
  const int cy = (lane >> 2) & 3, cz = lane & 3;  // A row: cell (slice, cy, cz)
  const int nblk = nb[0] * nb[1] * nb[2];
  int bidx[BPW][3];
  uint32_t bbit[BPW];
#pragma unroll
  for (int bi = 0; bi < BPW; ++bi) {
    const int b = wave + NWAVE * bi;
    bidx[bi][2] = b % nb[2];
    bidx[bi][1] = (b / nb[2]) % nb[1];
    bidx[bi][0] = b / (nb[2] * nb[1]);
    bool live = b < nblk;
#pragma unroll
    for (int d = 0; d < D; ++d) live = live && bidx[bi][d] * (d == D - 1 ? BL : kBlkE) < ext[d];
    bbit[bi] = live ? 1u << b : 0u;
  }

  f32x4 acc[BPW][kBlkE][GH];
#pragma unroll
  for (int bi = 0; bi < BPW; ++bi)
#pragma unroll
    for (int s = 0; s < kBlkE; ++s)
#pragma unroll
      for (int h = 0; h < GH; ++h) acc[bi][s][h] = f32x4{0.f, 0.f, 0.f, 0.f};

  constexpr int CHUNK = kBlkChunk * NWAVE / 4;  // staged entries per round (a multiple of 64)
  const int e0 = a.it_begin[item];
  const int e1 = min(e0 + a.smax, a.list_start[key + 1]);
  const int q = (int)threadIdx.x;
  // per block of the wave, the record offsets (floats from smem) of the
  // staged entries that touch it, in list order; a zero record after the
  // staged ones pads the last group of four
  const int zrec = CHUNK * PW;
  int16_t* lists = reinterpret_cast<int16_t*>(smem + zrec + PW) + wave * BPW * (CHUNK + 4);
  for (int i = (int)threadIdx.x; i < PW; i += 64 * NWAVE) smem[zrec + i] = 0.0f;
  // a lane's operands sit at fixed offsets in a record (block weights, value)
  int ow4[BPW], owy[BPW], owz[BPW];
#pragma unroll
  for (int bi = 0; bi < BPW; ++bi) {
    ow4[bi] = kBlkHead + 2 * G + kBlkE * bidx[bi][0];
    if constexpr (D == 3) {
      owy[bi] = kBlkHead + 2 * G + td[0] + kBlkE * bidx[bi][1] + cy;
      owz[bi] = kBlkHead + 2 * G + td[0] + td[1] + kBlkE * bidx[bi][2] + cz;
    } else {  // the row's cell along the last axis: 16 bidx + 4 cy + cz
      owy[bi] = kBlkHead + 2 * G + td[0] + BL * bidx[bi][1] + 4 * cy + cz;
      owz[bi] = 0;
    }
  }
  const int ov = kBlkHead + (lane & 15);
  // the point of the entry this thread stages next, read one round ahead
  int pnext = q < min(CHUNK, e1 - e0) ? a.list_pts[e0 + q] : 0;
  for (int eb = e0; eb < e1; eb += CHUNK) {
    const int n = min(CHUNK, e1 - eb);
    const int pcur = pnext;
    if (q < n) {
      BlkPos r;
      load_blk_pos<D>(a, pcur, org, r);
      write_mfma_entry<D, G>(smem + q * PW, r, a, t0, Gv, td, ext, nf, nb, inv_hw, beta);
    }
    // in flight during this round's lists and products
    if (q < min(CHUNK, e1 - eb - CHUNK)) pnext = a.list_pts[eb + CHUNK + q];
    __syncthreads();
    int ngroups[BPW];
#pragma unroll
    for (int bi = 0; bi < BPW; ++bi) {
      int16_t* lst = lists + bi * (CHUNK + 4);
      int cnt = 0;
      if (bbit[bi]) {  // uniform
        for (int base = 0; base < n; base += 64) {
          const int qi = base + lane;
          const uint32_t hm =
              qi < n ? reinterpret_cast<const uint32_t*>(smem + qi * PW)[0] : 0u;
          const bool hit = (hm & bbit[bi]) != 0u;
          const uint64_t bal = __ballot(hit);
          if (hit) lst[cnt + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u))] =
              (int16_t)(qi * PW);
          cnt += __popcll(bal);
        }
        if (lane < 4) lst[cnt + lane] = (int16_t)zrec;
      }
      ngroups[bi] = (cnt + 3) >> 2;
    }
#pragma unroll
    for (int bi = 0; bi < BPW; ++bi) {
      const int16_t* lst = lists + bi * (CHUNK + 4);
      // four entries per step (uniform count); the next step's record offset
      // is read before this step's products
      int o = ngroups[bi] > 0 ? lst[kq] : 0;
      for (int gq = 0; gq < ngroups[bi]; ++gq) {
        const int on = lst[4 * gq + 4 + kq];  // <= CHUNK + 3: inside the list
        const float* sp = smem + o;
        const float4 w4 = *reinterpret_cast<const float4*>(sp + ow4[bi]);
        const float wyz = D == 3 ? sp[owy[bi]] * sp[owz[bi]] : sp[owy[bi]];
        float av[kBlkE];
        av[0] = w4.x * wyz;
        av[1] = w4.y * wyz;
        av[2] = w4.z * wyz;
        av[3] = w4.w * wyz;
        float bv[GH];
#pragma unroll
        for (int h = 0; h < GH; ++h) bv[h] = sp[ov + 16 * h];
#pragma unroll
        for (int sx = 0; sx < kBlkE; ++sx)
#pragma unroll
          for (int h = 0; h < GH; ++h)
            acc[bi][sx][h] =
                __builtin_amdgcn_mfma_f32_16x16x4f32(av[sx], bv[h], acc[bi][sx][h], 0, 0, 0);
        o = on;
      }
    }
    __syncthreads();
  }

  // register r of lane l holds value l % 16 (transform 8 h + (l % 16) / 2,
  // re / im by parity) of cell (slice, l / 16, r); lanes 2g and 2g + 1 swap
  // halves so that each writes two whole complex cells
  const int TC = td[0] * td[1] * td[2];
  const bool odd = lane & 1;
  const int ry = lane >> 4;
#pragma unroll
  for (int bi = 0; bi < BPW; ++bi) {
    if (!bbit[bi]) continue;
#pragma unroll
    for (int sx = 0; sx < kBlkE; ++sx) {
#pragma unroll
      for (int hh = 0; hh < GH; ++hh) {
        float2 out[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          // even lane keeps cells r = 0, 1, odd lane cells r = 2, 3
          const float mine = odd ? acc[bi][sx][hh][2 + h] : acc[bi][sx][hh][h];
          const float give = odd ? acc[bi][sx][hh][h] : acc[bi][sx][hh][2 + h];
          const float other = __shfl_xor(give, 1);
          out[h] = odd ? make_float2(other, mine) : make_float2(mine, other);
        }
        const int g = 8 * hh + ((lane & 15) >> 1);
        if (g >= Gv) continue;
        const int c0 = bidx[bi][0] * kBlkE + sx;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          if constexpr (D == 3) {
            const int c1 = bidx[bi][1] * kBlkE + ry;
            const int c2 = bidx[bi][2] * kBlkE + (odd ? 2 : 0) + h;
            if (c0 >= ext[0] || c1 >= ext[1] || c2 >= ext[2]) continue;
            if (slot >= 0) {
              const int r = (c0 * td[1] + c1) * td[2] + c2;
              a.part[((int64_t)slot * a.ntrans + tb + t0 + g) * TC + r] = out[h];
            } else {
              const int64_t gi =
                  ((int64_t)(org[0] + c0) * nf[1] + org[1] + c1) * nf[2] + org[2] + c2;
              a.grid[(int64_t)(tb + t0 + g) * a.grid_stride + gi] = out[h];
            }
          } else {  // row cell 4 ry + r of the block's 16 along the last axis
            const int c1 = bidx[bi][1] * BL + 4 * ry + (odd ? 2 : 0) + h;
            if (c0 >= ext[0] || c1 >= ext[1]) continue;
            if (slot >= 0) {
              const int r = c0 * td[1] + c1;
              a.part[((int64_t)slot * a.ntrans + tb + t0 + g) * TC + r] = out[h];
            } else {
              const int64_t gi = (int64_t)(org[0] + c0) * nf[1] + org[1] + c1;
              a.grid[(int64_t)(tb + t0 + g) * a.grid_stride + gi] = out[h];
            }
          }
        }
      }
    }
  }
}

// Record of an entry for the few-transform matrix kernel: as
// write_mfma_entry (GT strengths, each axis's weight at every tile cell),
// with the touched-block mask over blocks of 4 x 4 x ZB cells.
template <int GT, int ZB>
__device__ __forceinline__ void write_mfmaz_entry(float* sp, const BlkPos& r,
                                                  const SpreadArgs& a, int t0, int Gv,
                                                  const int* td, const int* ext, const int* nf,
                                                  const int* nb, float inv_hw, float beta) {
  const int W = a.g.w;
  float2 v[GT];
#pragma unroll
  for (int g = 0; g < GT; ++g)
    v[g] = g < Gv ? a.in[(int64_t)(t0 + g) * a.in_stride + r.p] : make_float2(0.f, 0.f);
  uint32_t hit[3];
  float* wr = sp + kBlkHead + 2 * GT;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    hit[d] = axis_weights<false>(wr, r.st[d], r.off[d], nf[d], td[d], ext[d], W,
                                 d == 2 ? ZB : kBlkE, inv_hw, beta);
    wr += td[d];
  }
  uint32_t mask = 0u;
  for (int b0 = 0; b0 < nb[0]; ++b0)
    for (int b1 = 0; b1 < nb[1]; ++b1)
      for (int b2 = 0; b2 < nb[2]; ++b2)
        if ((hit[0] >> b0) & (hit[1] >> b1) & (hit[2] >> b2) & 1u)
          mask |= 1u << ((b0 * nb[1] + b1) * nb[2] + b2);
  reinterpret_cast<uint32_t*>(sp)[0] = mask;
  float* sv = sp + kBlkHead;
#pragma unroll
  for (int g = 0; g < GT; ++g) {
    sv[2 * g] = v[g].x;
    sv[2 * g + 1] = v[g].y;
  }
}

// Matrix-core spreading for GT = 1, 2 or 4 transforms per workgroup.  The
// 8-transform kernel above puts transforms x re / im in the 16 columns of
// each product, so with fewer transforms most of its products are zero.
// Here the columns are (z cell, transform, re / im) instead: for a block of
// 4 x 4 x ZB cells (ZB = 8 / GT) and four entries touching it, the block's
// sums are one v_mfma_f32_16x16x4f32 of A (16 xy cells of the block x 4
// entries: the products wx wy) by B (4 entries x 16 columns: wz times the
// strength).  Lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16] and
// holds D[4 (l / 16) + r][l % 16]: cells (x = l / 16, y = r) of column
// l % 16.  A wave owns blocks wave + 4 i and takes the entries touching each
// four at a time in list order: deterministic, equal to the other kernels
// to rounding.
template <int GT>
__global__ __launch_bounds__(kThreads) void k_spread_mfmaz(SpreadArgs a) {
  constexpr int ZB = 8 / GT;   // block edge along the last axis
  constexpr int BPW = 8;       // blocks per wave (32 per tile at most)
  extern __shared__ float smem[];
  const int PW = mfma_record_floats(GT, a.g.tile[0] + a.g.tile[1] + a.g.tile[2]);
  const int item = a.item0 + blockIdx.x;
  const int key = a.it_tile[item];
  const int set = key / a.g.nbins;
  const int tile = key - set * a.g.nbins;
  const int slot = a.it_slot[item];
  const int t0 = blockIdx.y * GT;
  const int tb = (set - a.k0) * a.ts;
  const int Gv = min(GT, a.ts - t0);
  const float inv_hw = 2.0f / a.g.w, beta = a.g.beta;

  int tc[3], org[3], ext[3], td[3], nf[3], nb[3];
  tile_coords<3>(tile, a.g.nt, tc);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    td[d] = a.g.tile[d];
    nf[d] = a.g.nf[d];
    org[d] = tc[d] * td[d];
    ext[d] = min(td[d], nf[d] - org[d]);
    nb[d] = td[d] / (d == 2 ? ZB : kBlkE);
  }
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / 64);
  const int lane = (int)threadIdx.x & 63;
  const int kq = lane >> 4;                 // entry of the group (A column, B row)
  const int ai = lane & 15;
  const int ax = ai >> 2, ay = ai & 3;      // A row: cell (ax, ay) of the block's xy face
  const int bz = ai / (2 * GT);             // B column: z cell, transform, re / im
  const int bg = (ai >> 1) % GT;
  const int bri = ai & 1;
  const int nblk = nb[0] * nb[1] * nb[2];
  int bidx[BPW][3];
  uint32_t bbit[BPW];
#pragma unroll
  for (int bi = 0; bi < BPW; ++bi) {
    const int b = wave + 4 * bi;
    bidx[bi][2] = b % nb[2];
    bidx[bi][1] = (b / nb[2]) % nb[1];
    bidx[bi][0] = b / (nb[2] * nb[1]);
    bool live = b < nblk;
    live = live && bidx[bi][0] * kBlkE < ext[0] && bidx[bi][1] * kBlkE < ext[1] &&
           bidx[bi][2] * ZB < ext[2];
    bbit[bi] = live ? 1u << b : 0u;
  }

  f32x4 acc[BPW];
#pragma unroll
  for (int bi = 0; bi < BPW; ++bi) acc[bi] = f32x4{0.f, 0.f, 0.f, 0.f};

  const int e0 = a.it_begin[item];
  const int e1 = min(e0 + a.smax, a.list_start[key + 1]);
  const int q = (int)threadIdx.x;
  for (int eb = e0; eb < e1; eb += kThreads) {
    const int n = min(kThreads, e1 - eb);
    if (q < n) {
      BlkPos r;
      load_blk_pos(a, a.list_pts[eb + q], org, r);
      write_mfmaz_entry<GT, ZB>(smem + q * PW, r, a, t0, Gv, td, ext, nf, nb, inv_hw, beta);
    }
    __syncthreads();
    for (int base = 0; base < n; base += 64) {
      const int qi = base + lane;
      const uint32_t hm = qi < n ? reinterpret_cast<const uint32_t*>(smem + qi * PW)[0] : 0u;
#pragma unroll
      for (int bi = 0; bi < BPW; ++bi) {
        if (!bbit[bi]) continue;  // uniform
        uint64_t todo = __ballot((hm & bbit[bi]) != 0u);
        while (todo) {  // uniform: four entries per step, list order
          int ek[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            ek[k] = todo ? __builtin_ctzll(todo) : -1;
            todo &= todo - 1;
          }
          const int e = kq == 0 ? ek[0] : kq == 1 ? ek[1] : kq == 2 ? ek[2] : ek[3];
          float av = 0.f, bv = 0.f;
          if (e >= 0) {
            const float* sp = smem + (base + e) * PW;
            const float* w = sp + kBlkHead + 2 * GT;
            av = w[kBlkE * bidx[bi][0] + ax] * w[td[0] + kBlkE * bidx[bi][1] + ay];
            bv = w[td[0] + td[1] + ZB * bidx[bi][2] + bz] * sp[kBlkHead + 2 * bg + bri];
          }
          acc[bi] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[bi], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }

  // lane l holds component bri of transform bg at cells (x = l / 16, y = r,
  // z = bz); lanes 2k and 2k + 1 (re and im of one column) swap halves so
  // that each writes two whole complex cells
  const int TC = td[0] * td[1] * td[2];
  const bool odd = lane & 1;
  if (bg >= Gv) return;
#pragma unroll
  for (int bi = 0; bi < BPW; ++bi) {
    if (!bbit[bi]) continue;
    float2 out[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const float mine = odd ? acc[bi][2 + h] : acc[bi][h];
      const float give = odd ? acc[bi][h] : acc[bi][2 + h];
      const float other = __shfl_xor(give, 1);
      out[h] = odd ? make_float2(other, mine) : make_float2(mine, other);
    }
    const int c0 = bidx[bi][0] * kBlkE + kq, c2 = bidx[bi][2] * ZB + bz;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int c1 = bidx[bi][1] * kBlkE + (odd ? 2 : 0) + h;
      if (c0 >= ext[0] || c1 >= ext[1] || c2 >= ext[2]) continue;
      if (slot >= 0) {
        const int r = (c0 * td[1] + c1) * td[2] + c2;
        a.part[((int64_t)slot * a.ntrans + tb + t0 + bg) * TC + r] = out[h];
      } else {
        const int64_t gi = ((int64_t)(org[0] + c0) * nf[1] + org[1] + c1) * nf[2] + org[2] + c2;
        a.grid[(int64_t)(tb + t0 + bg) * a.grid_stride + gi] = out[h];
      }
    }
  }
}

// Column-owner spreading for 3D tiles of 8 x 8 x 16 cells and 16 transforms
// per workgroup: lane c of every wave owns tile column c (16 cells along
// the contiguous axis), wave v owns transforms [4v, 4v + 4).  Per staged
// point the column weight is formed once per lane, and the point's cells
// along the column are found from its (wave-uniform, scalar) footprint start,
// so a lane only does arithmetic for the 5..7 cells the footprint covers.
// Same list order, same summation order per cell as k_spread: deterministic.
constexpr int kColsZ = 16;       // cells per column
constexpr int kColsCW = 4;       // transforms per wave
constexpr int kColsChunk = 128;  // staged points per round

__global__ __launch_bounds__(kThreads) void k_spread_cols(SpreadArgs a) {
  extern __shared__ float smem[];
  constexpr int D = 3, GW = 4 * kColsCW;  // transforms per workgroup
  const int W = a.g.w;
  const int PW = D + D * W + 2 * GW;
  const int item = a.item0 + blockIdx.x;
  const int key = a.it_tile[item];
  const int set = key / a.g.nbins;
  const int tile = key - set * a.g.nbins;
  const int slot = a.it_slot[item];
  const int tg = blockIdx.y * GW;        // first transform of the set handled here
  const int tb = (set - a.k0) * a.ts;    // chunk transform of the set's first
  const int Gv = min(GW, a.ts - tg);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int cw = wave * kColsCW;         // this wave's first transform (of the GW)
  const float inv_hw = 2.0f / W, beta = a.g.beta;

  int tc[3], org[3], ext[3], nf[3];
  tile_coords<D>(tile, a.g.nt, tc);
#pragma unroll
  for (int d = 0; d < D; ++d) {
    nf[d] = a.g.nf[d];
    org[d] = tc[d] * a.g.tile[d];
    ext[d] = min(a.g.tile[d], nf[d] - org[d]);
  }
  const int c0 = lane >> 3, c1 = lane & 7;
  const bool colown = c0 < ext[0] && c1 < ext[1];
  const int zext = ext[2];

  float2 acc[kColsZ][kColsCW];
#pragma unroll
  for (int z = 0; z < kColsZ; ++z)
#pragma unroll
    for (int g = 0; g < kColsCW; ++g) acc[z][g] = make_float2(0.f, 0.f);

  const int e0 = a.it_begin[item];
  const int e1 = min(e0 + a.smax, a.list_start[key + 1]);
  for (int eb = e0; eb < e1; eb += kColsChunk) {
    const int n = min(kColsChunk, e1 - eb);
    if ((int)threadIdx.x < n) {
      float* sp = smem + threadIdx.x * PW;
      const int p = a.list_pts[eb + threadIdx.x];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const float off = a.g.off[d][p];
        reinterpret_cast<int*>(sp)[d] = a.g.cell[d][p] - org[d];  // raw start, mod nf
        for (int j = 0; j < W; ++j) sp[D + d * W + j] = es_eval(off - (float)j, inv_hw, beta);
      }
      float* sv = sp + D + D * W;
      for (int g = 0; g < GW; ++g) {
        float2 v = make_float2(0.f, 0.f);
        if (g < Gv) v = a.in[(int64_t)(tg + g) * a.in_stride + p];
        sv[2 * g] = v.x;
        sv[2 * g + 1] = v.y;
      }
    }
    __syncthreads();
    if (cw < Gv) {  // uniform per wave
      for (int q = 0; q < n; ++q) {
        const float* sp = smem + q * PW;
        const int* si = reinterpret_cast<const int*>(sp);
        // column weight (per lane)
        int d0 = c0 - si[0], d1 = c1 - si[1];
        if (d0 < 0) d0 += nf[0];
        if (d0 >= nf[0]) d0 -= nf[0];
        if (d1 < 0) d1 += nf[1];
        if (d1 >= nf[1]) d1 -= nf[1];
        const bool hc = colown && d0 < W && d1 < W;
        const float wc = hc ? sp[D + d0] * sp[D + W + d1] : 0.0f;
        // footprint start along the column: the same for every lane
        const int s2 = __builtin_amdgcn_readfirstlane(si[2]);
        const float* w2 = sp + D + 2 * W;
        const float* sv = sp + D + D * W + 2 * cw;
        float2 str[kColsCW];
#pragma unroll
        for (int g = 0; g < kColsCW; ++g) str[g] = make_float2(sv[2 * g], sv[2 * g + 1]);
#pragma unroll
        for (int z = 0; z < kColsZ; ++z) {
          int da = z - s2;  // scalar
          if (da < 0) da += nf[2];
          if (da >= nf[2]) da -= nf[2];
          if (z < zext && da < W) {
            const float f = wc * w2[da];
#pragma unroll
            for (int g = 0; g < kColsCW; ++g) {
              acc[z][g].x += f * str[g].x;
              acc[z][g].y += f * str[g].y;
            }
          }
        }
      }
    }
    __syncthreads();
  }

  if (!colown) return;
  const int TC = a.g.tile[0] * a.g.tile[1] * a.g.tile[2];
  const int64_t gcol = (((int64_t)(org[0] + c0) * nf[1]) + org[1] + c1) * nf[2] + org[2];
#pragma unroll
  for (int g = 0; g < kColsCW; ++g) {
    if (cw + g >= Gv) break;  // uniform
    const int t = tb + tg + cw + g;
    if (slot >= 0) {  // partial tile of a dense tile: summed by k_spread_reduce
      float2* dst = a.part + ((int64_t)slot * a.ntrans + t) * TC + lane * kColsZ;
#pragma unroll
      for (int z = 0; z < kColsZ; ++z)
        if (z < zext) dst[z] = acc[z][g];
    } else {
      float2* dst = a.grid + (int64_t)t * a.grid_stride + gcol;
#pragma unroll
      for (int z = 0; z < kColsZ; ++z)
        if (z < zext) dst[z] = acc[z][g];
    }
  }
}

// Sum the partial tiles of each dense tile in slot order (deterministic).
template <int D>
__global__ __launch_bounds__(kThreads) void k_spread_reduce(SpreadArgs a) {
  const int j = a.dn0 + blockIdx.x;
  const int key = a.dn_tile[j], slot0 = a.dn_slot0[j], cnt = a.dn_count[j];
  const int set = key / a.g.nbins;
  const int tile = key - set * a.g.nbins;
  const int t = (set - a.k0) * a.ts + blockIdx.y;  // chunk transform
  int tc[3], org[3], ext[3], td[3];
  tile_coords<D>(tile, a.g.nt, tc);
  int TC = 1;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    td[d] = a.g.tile[d];
    org[d] = tc[d] * td[d];
    ext[d] = min(td[d], a.g.nf[d] - org[d]);
    TC *= td[d];
  }
  for (int r = threadIdx.x; r < TC; r += kThreads) {
    int rr = r, l[3];
    bool ok = true;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
      l[d] = rr % td[d];
      rr /= td[d];
      ok = ok && l[d] < ext[d];
    }
    if (!ok) continue;
    // the slots' loads are issued eight at a time (a dense centre tile has
    // up to ~80 partials per cell); the sum keeps the slot order
    const float2* src = a.part + ((int64_t)slot0 * a.ntrans + t) * TC + r;
    const int64_t sstep = (int64_t)a.ntrans * TC;
    float2 sum = make_float2(0.f, 0.f);
    int q = 0;
    for (; q + 8 <= cnt; q += 8) {
      float2 v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = src[(int64_t)(q + u) * sstep];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        sum.x += v[u].x;
        sum.y += v[u].y;
      }
    }
    for (; q < cnt; ++q) {
      const float2 v = src[(int64_t)q * sstep];
      sum.x += v.x;
      sum.y += v.y;
    }
    int64_t gi = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) gi = gi * a.g.nf[d] + org[d] + l[d];
    a.grid[(int64_t)t * a.grid_stride + gi] = sum;
  }
}


// ---------------------------------------------------------------------------
// interpolation (and the kernel-derivative gradient)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wrap_coord(int x, int nfd, int wrap1) {
  if (wrap1) {  // x inside [-nf, 2nf)
    if (x < 0) x += nfd;
    if (x >= nfd) x -= nfd;
  } else {      // the region can exceed the grid several times over (tiny grids)
    x %= nfd;
    if (x < 0) x += nfd;
  }
  return x;
}

// Stage the tile + halo region of transforms [t0, t0 + G) into LDS: cell
// (gg, l_0, .., l_{D-1}) at gg * RC + row-major l.  In 2D and 3D the region
// is NR rows of RL cells along the last axis: a table after the staged
// cells holds each row's wrapped grid offset, each thread owns one position
// along the last axis (its wrapped grid coordinate computed once) and walks
// the rows, kStageBatch loads in flight before their LDS stores.  The table
// is built here and synchronised before the loads; the caller synchronises
// after.
template <int D, int W, int NT = kThreads>
__device__ __forceinline__ void stage_region(const InterpArgs& a, float2* lds, int tile, int t0,
                                             int G, int* rorg, int* R, int& RC) {
  constexpr int kStageBatch = 24;
  int tc[3];
  tile_coords<D>(tile, a.g.nt, tc);
  RC = 1;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    R[d] = (d == D - 1 ? a.span : 1) * a.g.tile[d] + W;
    rorg[d] = tc[d] * a.g.tile[d] - W / 2;
    RC *= R[d];
  }
  if constexpr (D == 1) {
    const int nfd = a.g.nf[0];
    for (int gg = 0; gg < G; ++gg) {
      const float2* src = a.grid + (int64_t)(t0 + gg) * a.grid_stride;
      float2* dst = lds + gg * RC;
      for (int c0 = threadIdx.x; c0 < RC; c0 += kStageBatch * NT) {
        float2 v[kStageBatch];
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) {
          const int c = c0 + u * NT;
          if (c < RC) v[u] = src[wrap_coord(rorg[0] + c, nfd, a.wrap1)];
        }
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) {
          const int c = c0 + u * NT;
          if (c < RC) dst[c] = v[u];
        }
      }
    }
  } else {
    constexpr int DL = D - 1;
    const int RL = R[DL];
    const int NR = D == 2 ? R[0] : R[0] * R[1];
    int* rowoff = reinterpret_cast<int*>(lds + G * RC);
    for (int r = threadIdx.x; r < NR; r += NT) {
      int off;
      if constexpr (D == 2) {
        off = wrap_coord(rorg[0] + r, a.g.nf[0], a.wrap1);
      } else {
        const int l0 = mdiv(r, a.rmag[1]);
        const int l1 = r - l0 * R[1];
        off = wrap_coord(rorg[0] + l0, a.g.nf[0], a.wrap1) * a.g.nf[1] +
              wrap_coord(rorg[1] + l1, a.g.nf[1], a.wrap1);
      }
      rowoff[r] = off * a.g.nf[DL];
    }
    __syncthreads();
    // L = min(RL, NT) lanes per row, P = NT / L rows per pass (set by the host)
    const int L = a.stage_lanes, P = a.stage_rows;
    const int rs = mdiv((int)threadIdx.x, a.lane_mag);
    const int lz = (int)threadIdx.x - rs * L;
    if (rs >= P) return;
    for (int gg = 0; gg < G; ++gg) {
      const float2* src = a.grid + (int64_t)(t0 + gg) * a.grid_stride;
      float2* dst = lds + gg * RC;
      for (int z = lz; z < RL; z += L) {
        const int xz = wrap_coord(rorg[DL] + z, a.g.nf[DL], a.wrap1);
        for (int r0 = rs; r0 < NR; r0 += kStageBatch * P) {
          // each load waits for its table read: issuing all of them back to
          // back measured slower (the loads of every wave arrive at once)
          float2 v[kStageBatch];
#pragma unroll
          for (int u = 0; u < kStageBatch; ++u) {
            const int r = r0 + u * P;
            if (r < NR) v[u] = src[(uint32_t)(rowoff[r] + xz)];
          }
#pragma unroll
          for (int u = 0; u < kStageBatch; ++u) {
            const int r = r0 + u * P;
            if (r < NR) dst[r * RL + z] = v[u];
          }
        }
      }
    }
  }
}

// bins whose points a work item led by tile covers: span along the last
// axis, fewer at the grid end
template <int D>
__device__ __forceinline__ int span_bins(const InterpArgs& a, int tile) {
  const int nl = a.g.nt[D - 1];
  return min(a.span, nl - tile % nl);
}

// Region geometry of an interpolation work item led by `tile`: origin and
// extent per axis (span tiles along the last axis plus the kernel halo).
template <int D, int W>
__device__ __forceinline__ int region_of(const InterpArgs& a, int tile, int* rorg, int* R) {
  int tc[3];
  tile_coords<D>(tile, a.g.nt, tc);
  int RC = 1;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    R[d] = (d == D - 1 ? a.span : 1) * a.g.tile[d] + W;
    rorg[d] = tc[d] * a.g.tile[d] - W / 2;
    RC *= R[d];
  }
  return RC;
}

// Chunk masks of the work items' region rows (2D / 3D), built once per
// setpts.  A point reads the region rows (l0 + j0 [, l1 + j1]), j < W, and
// along the last axis the cells [l_last, l_last + W).  The last axis is cut
// into 16-cell chunks aligned with the grid's 128-byte lines (region cell z
// lies in chunk (z + phase) / 16, phase = the region origin mod 16), so a
// row's mask is the OR, over the points whose footprint covers the row, of
// the chunks their cells touch.  Without per-row atomics: each point ORs
// its chunks into the entry of its footprint's first row, then separable
// passes dilate the entries by W rows along each of the other axes.  One
// workgroup per work item; out[sub * NR + r] for the NR region rows.
template <int D, int W>
__global__ __launch_bounds__(kThreads) void k_interp_masks(InterpArgs a, uint16_t* out) {
  extern __shared__ uint32_t msk[];
  constexpr int DL = D - 1;
  const int sub = blockIdx.x;
  const int key = a.sub_tile[sub];
  const int set = key / a.g.nbins;
  const int tile = key - set * a.g.nbins;
  const int pb = a.sub_begin[sub];
  const int pe = min(pb + a.max_sub, a.bin_start[key + span_bins<D>(a, tile)]);
  int rorg[3], R[3];
  region_of<D, W>(a, tile, rorg, R);
  const int RB = D == 2 ? 1 : R[1];
  const int NR = D == 2 ? R[0] : R[0] * R[1];
  const int phase = ((rorg[DL] % 16) + 16) % 16;
  uint32_t* first = msk;
  uint32_t* dil = msk + NR;
  for (int r = threadIdx.x; r < NR; r += kThreads) first[r] = 0u;
  __syncthreads();
  for (int p = pb + threadIdx.x; p < pe; p += kThreads) {
    int l[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {  // the footprint origin of point_setup
      int v = a.g.cell[d][p] - rorg[d];
      if (v < 0) v += a.g.nf[d];
      if (v >= a.g.nf[d]) v -= a.g.nf[d];
      l[d] = min(max(v, 0), (d == DL ? a.span : 1) * a.g.tile[d]);
    }
    const int k0 = (l[DL] + phase) >> 4;
    const int k1 = min((l[DL] + W - 1 + phase) >> 4, 15);
    const uint32_t bits = (2u << k1) - (1u << k0);
    atomicOr(&first[D == 2 ? l[0] : l[0] * R[1] + l[1]], bits);
  }
  __syncthreads();
  for (int r = threadIdx.x; r < NR; r += kThreads) {  // covered along axis 0
    const int ra = D == 2 ? r : mdiv(r, a.rmag[1]);
    const int rb = r - ra * RB;
    uint32_t m = 0u;
    for (int i = 0; i < W && i <= ra; ++i) m |= first[(ra - i) * RB + rb];
    dil[r] = m;
  }
  __syncthreads();
  for (int r = threadIdx.x; r < NR; r += kThreads) {  // and along axis 1
    uint32_t m = dil[r];
    if constexpr (D == 3) {
      const int ra = mdiv(r, a.rmag[1]);
      const int rb = r - ra * RB;
      m = 0u;
      for (int j = 0; j < W && j <= rb; ++j) m |= dil[ra * RB + rb - j];
    }
    out[(int64_t)sub * NR + r] = (uint16_t)m;
  }
}

// Staging of only the chunks the work item's points read (the masks of
// k_interp_masks).  From the item's row masks a table of NC words per row is
// built after the G regions: the row's wrapped grid offset for a needed
// chunk, -1 otherwise, so each staged load reads one table word, as the
// whole-region staging reads one row offset.  A cell of an unneeded chunk
// loads nothing (a buffer load past the grid's end returns zero) and stores
// that zero where no point reads.  The caller synchronises after.
template <int D, int W, int NT>
__device__ __forceinline__ void stage_region_masked(const InterpArgs& a, float2* lds, int tile,
                                                    int sub, int t0, int G, int* rorg, int* R,
                                                    int& RC) {
  static_assert(D >= 2, "masked staging is for rows of 2D / 3D regions");
  constexpr int kStageBatch = 20;  // 24 spills with the chunk table
  constexpr int DL = D - 1;
  RC = region_of<D, W>(a, tile, rorg, R);
  const int RL = R[DL];
  const int NR = D == 2 ? R[0] : R[0] * R[1];
  const int NC = a.stage_chunks;
  const int phase = ((rorg[DL] % 16) + 16) % 16;
  int* table = reinterpret_cast<int*>(lds + G * RC);
  const uint16_t* mask = a.sub_mask + (int64_t)sub * NR;
  for (int r = threadIdx.x; r < NR; r += NT) {
    int off;
    if constexpr (D == 2) {
      off = wrap_coord(rorg[0] + r, a.g.nf[0], a.wrap1);
    } else {
      const int l0 = mdiv(r, a.rmag[1]);
      const int l1 = r - l0 * R[1];
      off = wrap_coord(rorg[0] + l0, a.g.nf[0], a.wrap1) * a.g.nf[1] +
            wrap_coord(rorg[1] + l1, a.g.nf[1], a.wrap1);
    }
    off *= a.g.nf[DL];
    const uint32_t m = mask[r];
    for (int k = 0; k < NC; ++k) table[r * NC + k] = ((m >> k) & 1u) ? off : -1;
  }
  __syncthreads();
  const int L = a.stage_lanes, P = a.stage_rows;
  const int rs = mdiv((int)threadIdx.x, a.lane_mag);
  const int lz = (int)threadIdx.x - rs * L;
  if (rs >= P) return;
  for (int gg = 0; gg < G; ++gg) {
    // the transform's grid as a raw buffer (every cell offset < 2^28, set
    // by the host): an offset past its end returns zero without a memory
    // access, which is how a cell of an unneeded chunk skips its load while
    // every load of a batch stays one unconditional instruction
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(a.grid + (int64_t)(t0 + gg) * a.grid_stride), (short)0, a.grid_bytes,
        kBufferWord3);
    float2* dst = lds + gg * RC;
    for (int z = lz; z < RL; z += L) {
      const int xz = wrap_coord(rorg[DL] + z, a.g.nf[DL], a.wrap1);
      const int kz = (z + phase) >> 4;
      for (int r0 = rs; r0 < NR; r0 += kStageBatch * P) {
        float2 v[kStageBatch];
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) {
          const int r = r0 + u * P;
          if (r < NR) {
            const int t = table[r * NC + kz];
            const uint32_t boff = t >= 0 ? (uint32_t)(t + xz) * 8u : 0x80000000u;
            // two 32-bit buffer loads, merged into one 64-bit load by the
            // compiler (its 64-bit buffer-load builtin lowers to a single
            // dword here)
            const uint32_t re = __builtin_amdgcn_raw_buffer_load_b32(rsrc, boff, 0, 0);
            const uint32_t im = __builtin_amdgcn_raw_buffer_load_b32(rsrc, boff + 4u, 0, 0);
            v[u] = make_float2(__builtin_bit_cast(float, re), __builtin_bit_cast(float, im));
          }
        }
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) {
          const int r = r0 + u * P;
          if (r < NR) dst[r * RL + z] = v[u];
        }
      }
    }
  }
}


template <int D, int W>
__device__ __forceinline__ void point_setup(const InterpArgs& a, int p, const int* rorg,
                                            int* ls, float (*wt)[W]) {
  const float inv_hw = 2.0f / W;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    int l = a.g.cell[d][p] - rorg[d];
    if (l < 0) l += a.g.nf[d];
    if (l >= a.g.nf[d]) l -= a.g.nf[d];
    // memory safety; a sorted point always satisfies it
    l = min(max(l, 0), (d == D - 1 ? a.span : 1) * a.g.tile[d]);
    ls[d] = l;
    const float off = a.g.off[d][p];
#pragma unroll
    for (int j = 0; j < W; ++j) wt[d][j] = es_eval(off - (float)j, inv_hw, a.g.beta);
  }
}

// Work items of one transform are laid out along blockIdx.x in tile order
// (last axis fastest).  Consecutive ids are dispatched to different XCDs, so
// the halo rows that neighbouring items share would be fetched into two
// L2s; the XCD remap gives each XCD a contiguous range of items (a slab of
// the grid), where neighbours along the last two axes run close in time on
// the same L2.  All XCDs still work on the same transform at once.
// Writes one interpolated value (scale, and in DC mode the residual and
// density) for point p, transform t0 + gg of the launch.
__device__ __forceinline__ void interp_store(const InterpArgs& a, int set, int t, int p, int idx,
                                             float dens, float sx, float sy) {
  const int64_t ko = ksp_index(a.g, a.ts, set, t, idx);
  float2 r = make_float2(a.scale * sx, a.scale * sy);
  if (a.mode == 1) {
    // residual for the spreading that follows: written in sorted order
    const float2 y = a.obs[ko];
    r.x = (r.x - y.x) * dens;
    r.y = (r.y - y.y) * dens;
    a.out[(int64_t)t * a.g.M + p] = r;
  } else {
    a.out[ko] = r;
  }
}

// Interpolation of one work item: stage its region (whole, or with MASKED
// only the chunks its points read), then gather each point's W^D footprint.
// In 3D a work item holds about a hundred points for 512 threads, so the
// points are spread over lp = 1, 2 or 4 consecutive lanes each (as many as
// the item allows, uniform per workgroup): lane s of a point sums the kernel
// rows j0 = s, s + lp, ..., and the partial sums meet by lane exchanges in
// a fixed order (deterministic for a given trajectory).
template <int D, int W, int NT, bool MASKED>
__global__ __launch_bounds__(NT, 4) void k_interp(InterpArgs a) {
  extern __shared__ float2 lds[];
  const int sub = a.sub0 + (int)xcd_remap(blockIdx.x, gridDim.x);
  const int key = a.sub_tile[sub];
  const int set = key / a.g.nbins;
  const int tile = key - set * a.g.nbins;
  const int pb = a.sub_begin[sub];
  const int pe = min(pb + a.max_sub, a.bin_start[key + span_bins<D>(a, tile)]);
  const int t0 = blockIdx.y * a.G;       // first transform of the set handled here
  const int G = min(a.G, a.ts - t0);
  const int tb = (set - a.k0) * a.ts;    // chunk transform of the set's first
  int rorg[3], R[3], RC;
  if constexpr (MASKED && D >= 2) {
    stage_region_masked<D, W, NT>(a, lds, tile, sub, tb + t0, G, rorg, R, RC);
  } else {
    stage_region<D, W, NT>(a, lds, tile, tb + t0, G, rorg, R, RC);
  }
  __syncthreads();

  if constexpr (D == 3) {
    const int n = pe - pb;
    int lg = 0;  // lp = 1 << lg lanes per point
    while (lg < 2 && (n << (lg + 1)) <= NT) ++lg;
    const int lp = 1 << lg;
    const int trips = ((n << lg) + NT - 1) / NT;   // the same for every lane
    const float inv_hw = 2.0f / W;
    for (int it = 0; it < trips; ++it) {
      const int q = it * NT + (int)threadIdx.x;
      const int p = pb + (q >> lg);
      const int s = q & (lp - 1);
      const bool live = p < pe;
      int ls[3] = {0, 0, 0};
      float wt[D][W];
      int idx = 0;
      float dens = 1.0f, off0 = 0.0f;
      if (live) {
        point_setup<D, W>(a, p, rorg, ls, wt);
        idx = a.g.idx[p];
        dens = (a.mode == 1 && a.density) ? a.density[idx] : 1.0f;
        off0 = a.g.off[0][p];
      }
      for (int gg = 0; gg < G; ++gg) {
        const float2* base = lds + gg * RC;
        float sx = 0.f, sy = 0.f;
        if (live) {
#pragma unroll 1
          for (int j0 = s; j0 < W; j0 += lp) {
            const float w0 = es_eval(off0 - (float)j0, inv_hw, a.g.beta);
            float px = 0.f, py = 0.f;
#pragma unroll
            for (int j1 = 0; j1 < W; ++j1) {
              const float2* row = base + ((ls[0] + j0) * R[1] + ls[1] + j1) * R[2] + ls[2];
              float rx = 0.f, ry = 0.f;
#pragma unroll
              for (int j2 = 0; j2 < W; ++j2) {
                const float2 v = row[j2];
                rx += wt[2][j2] * v.x;
                ry += wt[2][j2] * v.y;
              }
              px += wt[1][j1] * rx;
              py += wt[1][j1] * ry;
            }
            sx += w0 * px;
            sy += w0 * py;
          }
        }
        for (int m = 1; m < lp; m <<= 1) {  // lanes of one point are consecutive
          sx += __shfl_xor(sx, m);
          sy += __shfl_xor(sy, m);
        }
        if (live && s == 0) interp_store(a, set, t0 + gg, p, idx, dens, sx, sy);
      }
    }
    return;
  }

  for (int p = pb + threadIdx.x; p < pe; p += NT) {
    int ls[3];
    float wt[D][W];
    point_setup<D, W>(a, p, rorg, ls, wt);
    const int idx = a.g.idx[p];
    const float dens = (a.mode == 1 && a.density) ? a.density[idx] : 1.0f;
    for (int gg = 0; gg < G; ++gg) {
      const float2* base = lds + gg * RC;
      float sx = 0.f, sy = 0.f;
      if constexpr (D == 1) {
#pragma unroll
        for (int j = 0; j < W; ++j) {
          const float2 v = base[ls[0] + j];
          sx += wt[0][j] * v.x;
          sy += wt[0][j] * v.y;
        }
      } else {
#pragma unroll
        for (int j0 = 0; j0 < W; ++j0) {
          const float2* row = base + (ls[0] + j0) * R[1] + ls[1];
          float rx = 0.f, ry = 0.f;
#pragma unroll
          for (int j1 = 0; j1 < W; ++j1) {
            const float2 v = row[j1];
            rx += wt[1][j1] * v.x;
            ry += wt[1][j1] * v.y;
          }
          sx += wt[0][j0] * rx;
          sy += wt[0][j0] * ry;
        }
      }
      interp_store(a, set, t0 + gg, p, idx, dens, sx, sy);
    }
  }
}

// k_interp_stream configuration: threads per workgroup and LDS buffers (512
// threads with one buffer, two workgroups per CU, ran the 16-coil 192^3
// interpolation in 5.5 ms; 1024 threads with two buffers in 6.6 ms; 512
// with two, one workgroup per CU, 7.1 ms), region cells a buffer holds
// (8 x 8 x 16 tiles, two along the last axis, w <= 8), region rows the row
// tables hold, the widest kernel it is built for (its footprint loops unroll
// fully) and the LDS cells after each region row (rows of 41 or 42 cells
// measured the same as 39, rows of 40 slower)
constexpr int kStreamNT = 512;
constexpr int kStreamNB = 1;
constexpr int kStreamCells = 9728;
constexpr int kStreamRows = 256;
constexpr int kStreamMaxW = 8;
constexpr int kStreamPad = 0;
// 2D regions (32 x 32 tiles plus the halo: 1521 cells) are small, so 2D
// workgroups of 256 threads hold 2048 cells and several share a CU
constexpr int kStream2NT = 256;
constexpr int kStream2Cells = 2048;
constexpr int kStream2Rows = 128;

// 2D / 3D interpolation of many transforms of a work item in one workgroup
// (in 2D a region row is one line along the contiguous axis, in 3D one
// (x, y) pair's line).
// What does not depend on the transform is done once per item: the grid
// byte offset of every staged 32-bit word of the region (each thread owns
// the words tid + u * NT, u < K; a word of a chunk no point reads gets an
// offset past the grid's end, which loads nothing), and each point's kernel
// weights.  Each transform's region goes HBM -> LDS by direct-to-LDS buffer
// loads (no registers hold it; a wave's load writes 64 consecutive words,
// which the flat word order of the region gives), then its points gather
// their footprints from LDS.  With one buffer (NB = 1) two workgroups share
// a CU and one's loads overlap the other's gathers; with two (NB = 2) the
// loads of transform gg + 1 are issued before transform gg is gathered.
// Points are spread over lp = 1, 2, 4 or 8 consecutive lanes (as many as
// the item allows; it holds at most NT >> LGMIN points), lane s summing
// the kernel rows j0 = s, s + lp, ...; the partial sums meet by lane
// exchanges in a fixed order, so the result does not depend on the run.
// The footprint gathers bound the kernel: random 16-byte LDS reads.
template <int D, int W, int NT, int NB, int CELLS, int ROWS>
__global__ __launch_bounds__(NT, 4) void k_interp_stream(InterpArgs a) {
  // static LDS (NB buffers of CELLS cells and the row tables): the compiler
  // then knows how many workgroups share a CU
  __shared__ float2 lds[NB * CELLS + ROWS];
  constexpr int K = 2 * CELLS / NT;  // staged words per thread
  constexpr int DL = D - 1;                 // the contiguous axis
  constexpr uint32_t kSkip = 0x80000000u;
  constexpr int LGMIN = NT >= 1024 ? 1 : 0;  // items hold at most NT >> LGMIN points
  constexpr int KJ = (W + (1 << LGMIN) - 1) >> LGMIN;  // rows per lane at the fewest lanes
  const int sub = a.sub0 + (int)xcd_remap(blockIdx.x, gridDim.x);
  const int key = a.sub_tile[sub];
  const int set = key / a.g.nbins;
  const int tile = key - set * a.g.nbins;
  const int pb = a.sub_begin[sub];
  const int pe = min(pb + a.max_sub, a.bin_start[key + span_bins<D>(a, tile)]);
  const int t0 = blockIdx.y * a.G;
  const int G = min(a.G, a.ts - t0);
  const int tb = (set - a.k0) * a.ts + t0;  // chunk transform of the first handled here
  int rorg[3], R[3];
  region_of<D, W>(a, tile, rorg, R);
  const int NR = D == 3 ? R[0] * R[1] : R[0];  // region rows along the contiguous axis
  const int RL = R[DL] + a.lds_pad;  // LDS row length: the padding moves rows across banks
  const int RC = NR * RL;
  const int RCP = (RC + 31) & ~31;  // a buffer: whole 64-word wave segments
  const int tid = (int)threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);

  // per region row: wrapped grid offset of its first cell, chunk mask
  int* rowoff = reinterpret_cast<int*>(lds + NB * RCP);
  uint32_t* rowmask = reinterpret_cast<uint32_t*>(rowoff + NR);
  for (int r = tid; r < NR; r += NT) {
    if constexpr (D == 3) {
      const int l0 = mdiv(r, a.rmag[1]);
      const int l1 = r - l0 * R[1];
      rowoff[r] = (wrap_coord(rorg[0] + l0, a.g.nf[0], a.wrap1) * a.g.nf[1] +
                   wrap_coord(rorg[1] + l1, a.g.nf[1], a.wrap1)) * a.g.nf[2];
    } else {
      rowoff[r] = wrap_coord(rorg[0] + r, a.g.nf[0], a.wrap1) * a.g.nf[1];
    }
    rowmask[r] = a.sub_mask ? (uint32_t)a.sub_mask[(int64_t)sub * NR + r] : 0xffffffffu;
  }
  __syncthreads();
  const int phase = ((rorg[DL] % 16) + 16) % 16;
  uint32_t off[K];
#pragma unroll
  for (int u = 0; u < K; ++u) {
    const int wd = tid + u * NT;  // word: cell wd / 2, real (0) or imaginary (1) part
    const int c = wd >> 1;
    off[u] = kSkip;
    if (c < RC) {
      const int r = mdiv(c, a.lds_row_mag);
      const int z = c - r * RL;
      if (z < R[DL] && ((rowmask[r] >> ((z + phase) >> 4)) & 1u))
        off[u] = (uint32_t)(rowoff[r] + wrap_coord(rorg[DL] + z, a.g.nf[DL], a.wrap1)) * 8u +
                 4u * (uint32_t)(wd & 1);
    }
  }

  // the points: one trip of lp lanes each
  const int n = pe - pb;
  int lg = LGMIN;
  while (lg < 3 && (n << (lg + 1)) <= NT) ++lg;
  const int lp = 1 << lg;
  const int p = pb + (tid >> lg);
  const int s = tid & (lp - 1);
  const bool live = p < pe;
  int lbase = 0;  // the lane's first footprint row j0 = s, as an LDS cell index
  float wt1[W], wt2[W], w0[KJ];
  int64_t kbase = 0, kstep = 0;  // k-space offset of transform t: kbase + t * kstep
  float dens = 1.0f;
#pragma unroll
  for (int j = 0; j < W; ++j) wt1[j] = wt2[j] = 0.0f;
#pragma unroll
  for (int k = 0; k < KJ; ++k) w0[k] = 0.0f;
  if (live) {
    int ls[3];
    float wt[D][W];
    point_setup<D, W>(a, p, rorg, ls, wt);
#pragma unroll
    for (int j = 0; j < W; ++j) {  // wt2: the contiguous axis; wt1: the middle one (3D)
      wt1[j] = wt[D == 3 ? 1 : 0][j];
      wt2[j] = wt[DL][j];
    }
    const float inv_hw = 2.0f / W;
    const float off0 = a.g.off[0][p];
#pragma unroll
    for (int k = 0; k < KJ; ++k)  // the weights of the lane's rows s + k lp
      w0[k] = s + k * lp < W ? es_eval(off0 - (float)(s + k * lp), inv_hw, a.g.beta) : 0.0f;
    lbase = D == 3 ? ((ls[0] + s) * R[1] + ls[1]) * RL + ls[2] : (ls[0] + s) * RL + ls[1];
    const int idx = a.g.idx[p];
    dens = (a.mode == 1 && a.density) ? a.density[idx] : 1.0f;
    kbase = ksp_index(a.g, a.ts, set, 0, idx);
    kstep = ksp_index(a.g, a.ts, set, 1, idx) - kbase;
  }
  const int rstep = lp * (D == 3 ? R[1] : 1) * RL;  // LDS cells between the lane's rows

  // the region of transform t into buffer b: load u of wave w fills words
  // [u NT + 64 w, u NT + 64 w + 64) of the buffer
  auto stage = [&](int t, int b) {
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(a.grid + (int64_t)(tb + t) * a.grid_stride), (short)0, a.grid_bytes,
        kBufferWord3);
    float* buf = reinterpret_cast<float*>(lds + b * RCP);
#pragma unroll
    for (int u = 0; u < K; ++u) {
      const int w0d = u * NT + wave * 64;  // uniform
      if (w0d < 2 * RCP)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            rs, (__attribute__((address_space(3))) void*)(buf + w0d), 4, off[u], 0, 0, 0);
    }
  };
  if (NB == 2) stage(0, 0);
  for (int gg = 0; gg < G; ++gg) {
    float2 y = make_float2(0.f, 0.f);
    if constexpr (NB == 2) {
      // this thread's loads of transform gg have landed; after the barrier
      // every thread's have, and every lane is done with transform gg - 1
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      // the residual's observation before the next loads, so that waiting
      // for it leaves them in flight
      if (a.mode == 1 && live && s == 0) y = a.obs[kbase + (t0 + gg) * kstep];
      if (gg + 1 < G) stage(gg + 1, (gg + 1) & 1);
    } else {
      if (gg > 0) __syncthreads();  // every lane is done with transform gg - 1
      if (a.mode == 1 && live && s == 0) y = a.obs[kbase + (t0 + gg) * kstep];
      stage(gg, 0);
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
    }
    const float2* cur = lds + (NB == 2 ? (gg & 1) * RCP : 0);
    float sx = 0.f, sy = 0.f;
#pragma unroll
    for (int k = 0; k < KJ; ++k) {
      if (live && s + k * lp < W) {
        const float2* rowk = cur + lbase + k * rstep;
        float px = 0.f, py = 0.f;
        if constexpr (D == 3) {
#pragma unroll
          for (int j1 = 0; j1 < W; ++j1) {
            // the row's reads first, so that they are in flight together
            // (read one by one, each waited for its predecessor's result:
            // 4.52 -> 4.41 ms for the 16-coil 192^3 interpolation)
            const float2* row = rowk + j1 * RL;
            float2 c[W];
#pragma unroll
            for (int j2 = 0; j2 < W; ++j2) c[j2] = row[j2];
            float rx = 0.f, ry = 0.f;
#pragma unroll
            for (int j2 = 0; j2 < W; ++j2) {
              rx += wt2[j2] * c[j2].x;
              ry += wt2[j2] * c[j2].y;
            }
            px += wt1[j1] * rx;
            py += wt1[j1] * ry;
          }
        } else {  // 2D: row s + k lp of the footprint is one line along the contiguous axis
#pragma unroll
          for (int j2 = 0; j2 < W; ++j2) {
            const float2 c = rowk[j2];
            px += wt2[j2] * c.x;
            py += wt2[j2] * c.y;
          }
        }
        sx += w0[k] * px;
        sy += w0[k] * py;
      }
    }
    for (int m = 1; m < lp; m <<= 1) {
      sx += __shfl_xor(sx, m);
      sy += __shfl_xor(sy, m);
    }
    if (live && s == 0) {
      float2 r = make_float2(a.scale * sx, a.scale * sy);
      if (a.mode == 1) {
        r.x = (r.x - y.x) * dens;
        r.y = (r.y - y.y) * dens;
        a.out[(int64_t)(t0 + gg) * a.g.M + p] = r;
      } else {
        a.out[kbase + (t0 + gg) * kstep] = r;
      }
    }
  }
}

// gtraj[idx*D + d] += Re sum_t conj(dy[t][idx]) * scale * d/dx_d interp_t
template <int D, int W>
__global__ __launch_bounds__(kThreads) void k_grad_traj(InterpArgs a) {
  extern __shared__ float2 lds[];
  const int sub = a.sub0 + blockIdx.x;
  const int key = a.sub_tile[sub];
  const int set = key / a.g.nbins;
  const int tile = key - set * a.g.nbins;
  const int pb = a.sub_begin[sub];
  const int pe = min(pb + a.max_sub, a.bin_start[key + span_bins<D>(a, tile)]);
  const int G = a.G;  // transforms [tg0, tg0 + G) of the set in one workgroup
  const int tb = (set - a.k0) * a.ts + a.tg0;
  int rorg[3], R[3], RC;
  stage_region<D, W>(a, lds, tile, tb, G, rorg, R, RC);
  __syncthreads();
  const float inv_hw = 2.0f / W;

  for (int p = pb + threadIdx.x; p < pe; p += kThreads) {
    int ls[3];
    float wt[D][W], dw[D][W];
    point_setup<D, W>(a, p, rorg, ls, wt);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const float off = a.g.off[d][p];
#pragma unroll
      for (int j = 0; j < W; ++j) dw[d][j] = es_deriv(off - (float)j, inv_hw, a.g.beta);
    }
    const int idx = a.g.idx[p];
    float acc[3] = {0.f, 0.f, 0.f};
    for (int gg = 0; gg < G; ++gg) {
      const float2* base = lds + gg * RC;
      float2 der[3];
      if constexpr (D == 1) {
        float dx = 0.f, dy = 0.f;
#pragma unroll
        for (int j = 0; j < W; ++j) {
          const float2 v = base[ls[0] + j];
          dx += dw[0][j] * v.x;
          dy += dw[0][j] * v.y;
        }
        der[0] = make_float2(dx, dy);
      } else if constexpr (D == 2) {
        float2 d0 = {0.f, 0.f}, d1 = {0.f, 0.f};
#pragma unroll
        for (int j0 = 0; j0 < W; ++j0) {
          const float2* row = base + (ls[0] + j0) * R[1] + ls[1];
          float2 s = {0.f, 0.f}, sd = {0.f, 0.f};
#pragma unroll
          for (int j1 = 0; j1 < W; ++j1) {
            const float2 v = row[j1];
            s.x += wt[1][j1] * v.x;
            s.y += wt[1][j1] * v.y;
            sd.x += dw[1][j1] * v.x;
            sd.y += dw[1][j1] * v.y;
          }
          d0.x += dw[0][j0] * s.x;
          d0.y += dw[0][j0] * s.y;
          d1.x += wt[0][j0] * sd.x;
          d1.y += wt[0][j0] * sd.y;
        }
        der[0] = d0;
        der[1] = d1;
      } else {
        float2 d0 = {0.f, 0.f}, d1 = {0.f, 0.f}, d2 = {0.f, 0.f};
        const float off0 = a.g.off[0][p];
#pragma unroll 1
        for (int j0 = 0; j0 < W; ++j0) {
          const float w0 = es_eval(off0 - (float)j0, inv_hw, a.g.beta);
          const float dw0 = es_deriv(off0 - (float)j0, inv_hw, a.g.beta);
          float2 s1 = {0.f, 0.f}, s1d1 = {0.f, 0.f}, s1d2 = {0.f, 0.f};
#pragma unroll
          for (int j1 = 0; j1 < W; ++j1) {
            const float2* row = base + ((ls[0] + j0) * R[1] + ls[1] + j1) * R[2] + ls[2];
            float2 s = {0.f, 0.f}, sd = {0.f, 0.f};
#pragma unroll
            for (int j2 = 0; j2 < W; ++j2) {
              const float2 v = row[j2];
              s.x += wt[2][j2] * v.x;
              s.y += wt[2][j2] * v.y;
              sd.x += dw[2][j2] * v.x;
              sd.y += dw[2][j2] * v.y;
            }
            s1.x += wt[1][j1] * s.x;
            s1.y += wt[1][j1] * s.y;
            s1d1.x += dw[1][j1] * s.x;
            s1d1.y += dw[1][j1] * s.y;
            s1d2.x += wt[1][j1] * sd.x;
            s1d2.y += wt[1][j1] * sd.y;
          }
          d0.x += dw0 * s1.x;
          d0.y += dw0 * s1.y;
          d1.x += w0 * s1d1.x;
          d1.y += w0 * s1d1.y;
          d2.x += w0 * s1d2.x;
          d2.y += w0 * s1d2.y;
        }
        der[0] = d0;
        der[1] = d1;
        der[2] = d2;
      }
      const float2 y = a.obs[ksp_index(a.g, a.ts, set, a.tg0 + gg, idx)];  // dy
#pragma unroll
      for (int d = 0; d < D; ++d) acc[d] += y.x * der[d].x + y.y * der[d].y;  // Re(conj(dy) der)
    }
#pragma unroll
    for (int d = 0; d < D; ++d) a.gtraj[(int64_t)idx * D + d] += a.scale * a.dscale[d] * acc[d];
  }
}

#define HN_WIDTH_SWITCH(W, BODY)                 \
  switch (W) {                                   \
    case 2: { constexpr int WW = 2; BODY; } break;  \
    case 3: { constexpr int WW = 3; BODY; } break;  \
    case 4: { constexpr int WW = 4; BODY; } break;  \
    case 5: { constexpr int WW = 5; BODY; } break;  \
    case 6: { constexpr int WW = 6; BODY; } break;  \
    case 7: { constexpr int WW = 7; BODY; } break;  \
    case 8: { constexpr int WW = 8; BODY; } break;  \
    case 9: { constexpr int WW = 9; BODY; } break;  \
    case 10: { constexpr int WW = 10; BODY; } break; \
    case 11: { constexpr int WW = 11; BODY; } break; \
    case 12: { constexpr int WW = 12; BODY; } break; \
    case 13: { constexpr int WW = 13; BODY; } break; \
    case 14: { constexpr int WW = 14; BODY; } break; \
    case 15: { constexpr int WW = 15; BODY; } break; \
    case 16: { constexpr int WW = 16; BODY; } break; \
    default: return set_error(HIPNUFFT_ERR_ARG, "unsupported kernel width"); \
  }

#define HN_DIM_SWITCH(D, BODY)                    \
  switch (D) {                                    \
    case 1: { constexpr int DD = 1; BODY; } break; \
    case 2: { constexpr int DD = 2; BODY; } break; \
    case 3: { constexpr int DD = 3; BODY; } break; \
    default: return set_error(HIPNUFFT_ERR_ARG, "unsupported ndim"); \
  }

constexpr size_t kMaxLds = 64 * 1024;

}  // namespace

PtsGeom pts_geom(const hipnufft_plan_s* p) {
  PtsGeom g;
  g.D = p->ndim;
  const int o = 3 - p->ndim;
  for (int d = 0; d < p->ndim; ++d) {
    g.nf[d] = (int)p->nf[o + d];
    g.tile[d] = p->tile[o + d];
    g.nt[d] = p->nt[o + d];
    g.cell[d] = p->pt_cell[d];
    g.off[d] = p->pt_off[d];
  }
  g.w = p->w;
  g.beta = (float)p->beta;
  g.M = p->M;
  g.nbins = (int)p->nbins;
  g.nsets = p->nsets;
  g.set_off = p->set_off;
  g.idx = p->pt_idx;
  return g;
}

static int tile_cells(const hipnufft_plan_s* p, int halo) {
  int c = 1;
  for (int d = 3 - p->ndim; d < 3; ++d) c *= p->tile[d] + halo;
  return c;
}

static uint32_t magic_of(int64_t d) {  // ceil(2^32 / d), d >= 2
  return (uint32_t)((((uint64_t)1 << 32) + (uint64_t)d - 1) / (uint64_t)d);
}

// extent of an interpolation region along axis d (3-padded index): the
// work item's tiles (interp_span of them along the last axis) plus the
// kernel width
static int region_extent(const hipnufft_plan_s* p, int d) {
  return (d == 2 ? p->interp_span : 1) * p->tile[d] + p->w;
}

// 16-cell chunks (128-byte lines) a region row along the last axis can meet
// whatever its alignment
static int region_chunks(const hipnufft_plan_s* p) {
  return (region_extent(p, 2) + 15 + 15) / 16;
}

static void set_region_magic(const hipnufft_plan_s* p, InterpArgs* a, int nt = kThreads) {
  const int o = 3 - p->ndim;
  a->wrap1 = 1;
  for (int d = 0; d < p->ndim; ++d) {
    const int R = region_extent(p, o + d);  // >= 3
    a->rmag[d] = magic_of(R);
    if (p->nf[o + d] < R) a->wrap1 = 0;
  }
  const int RL = region_extent(p, 2);
  a->stage_lanes = std::min(RL, nt);
  a->stage_rows = nt / a->stage_lanes;
  a->lane_mag = magic_of(a->stage_lanes);
  a->span = p->interp_span;
  a->stage_chunks = region_chunks(p);
  const int64_t cells = grid_size(p);
  a->mask_stage = p->mask_rows > 0 && p->sub_mask && cells < ((int64_t)1 << 28);
  a->grid_bytes = a->mask_stage ? (int)(cells * (int64_t)sizeof(float2)) : 0;
  a->sub_mask = a->mask_stage ? p->sub_mask : nullptr;
}

// LDS of an interpolation workgroup staging G transforms: the regions and,
// in 2D / 3D, the table of row offsets after them
static size_t region_lds_bytes(const hipnufft_plan_s* p, int G) {
  size_t cells = 1;
  for (int d = 3 - p->ndim; d < 3; ++d) cells *= (size_t)region_extent(p, d);
  const size_t rows = p->ndim == 1 ? 0 : cells / (size_t)region_extent(p, 2);
  // the regions, then per row its offset (unmasked staging) or its table of
  // chunk offsets (masked staging)
  return sizeof(float2) * cells * (size_t)G + sizeof(int) * rows * (size_t)region_chunks(p);
}

// one transform's region may use all of a CU's LDS; several transforms per
// workgroup only while they fit the 64 KiB that keeps workgroups co-resident
constexpr size_t kMaxLdsRegion = 160 * 1024;

template <typename K>
static int allow_lds(K kernel, size_t lds) {
  if (lds <= kMaxLds) return 0;
  HN_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  return 0;
}

static int check_range(const hipnufft_plan_s* p, SetRange r) {
  if (r.k0 < 0 || r.k1 > p->nsets || r.k0 >= r.k1 || r.ts < 1)
    return set_error(HIPNUFFT_ERR_ARG, "invalid point-set range");
  if ((int64_t)(r.k1 - r.k0) * r.ts > p->ntrans)
    return set_error(HIPNUFFT_ERR_ARG, "chunk larger than ntrans");
  return 0;
}

namespace {

// ksorted[t][q] = in[set k of q, t, idx[q]] * density[idx[q]] for the sorted
// positions q of the launch's sets: one pass that reads the caller's k-space
// once (runs of consecutive samples stay contiguous after binning) and
// writes the spreading input in sorted-point order.
__global__ __launch_bounds__(kThreads) void k_gather_sorted(PtsGeom g, int ts, int64_t q0,
                                                            int64_t q1, const float2* in,
                                                            const float* density, float2* out) {
  const int64_t q = q0 + (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (q >= q1) return;
  const int t = blockIdx.y;
  const int k = find_set(g.set_off, g.nsets, q);
  const int32_t o = g.idx[q];
  float2 v = in[ksp_index(g, ts, k, t, o)];
  if (density) {
    const float d = density[o];
    v.x *= d;
    v.y *= d;
  }
  out[(int64_t)t * g.M + q] = v;
}

}  // namespace

int launch_spread(const hipnufft_plan_s* p, const float2* in, const float* density, SetRange r,
                  float2* grid, int64_t grid_stride) {
  int rc = check_range(p, r);
  if (rc) return rc;
  if (!p->ksorted || p->ksorted_cap < (int64_t)r.ts * p->M)
    return set_error(HIPNUFFT_ERR_ARG, "sorted k-space workspace missing");
  const int64_t q0 = p->set_off_h[r.k0], q1 = p->set_off_h[r.k1];
  if (q1 > q0) {
    const dim3 g((unsigned)((q1 - q0 + kThreads - 1) / kThreads), (unsigned)r.ts);
    hipLaunchKernelGGL(k_gather_sorted, g, dim3(kThreads), 0, p->stream, pts_geom(p), r.ts, q0,
                       q1, in, density, p->ksorted);
    HN_LAUNCH_CHECK();
  }
  return launch_spread_sorted(p, p->ksorted, r, grid, grid_stride);
}

int launch_spread_sorted(const hipnufft_plan_s* p, const float2* in, SetRange r, float2* grid,
                         int64_t grid_stride) {
  int rc = check_range(p, r);
  if (rc) return rc;
  SpreadArgs a{};
  a.g = pts_geom(p);
  a.list_pts = p->list_pts;
  a.list_start = p->list_start;
  a.it_tile = p->it_tile;
  a.it_begin = p->it_begin;
  a.it_slot = p->it_slot;
  a.dn_tile = p->dn_tile;
  a.dn_slot0 = p->dn_slot0;
  a.dn_count = p->dn_count;
  a.part = p->part;
  a.ntrans = p->ntrans;
  a.smax = p->spread_max;
  a.in = in;
  a.in_stride = p->M;
  a.grid = grid;
  a.grid_stride = grid_stride;
  a.k0 = r.k0;
  a.ts = r.ts;
  a.item0 = (int)p->item_of_set[r.k0];
  a.dn0 = (int)p->dense_of_set[r.k0];
  const int64_t nitems = p->item_of_set[r.k1] - p->item_of_set[r.k0];
  const int64_t ndense = p->dense_of_set[r.k1] - p->dense_of_set[r.k0];
  {  // every (column, kZ-cell segment) of a tile needs its own thread
    const int zl = p->tile[2];
    const int ncol = tile_cells(p, 0) / zl;
    if (ncol * ((zl + kZ - 1) / kZ) > kThreads)
      return set_error(HIPNUFFT_ERR_ARG, "spread tile has more than 256 column segments of 4 cells");
  }
  // kernel choice: column owners for 3D 8 x 8 x 16 tiles when asked for
  // (spread_group >= 16 and at least 8 transforms per set), else the
  // matrix-core kernel for 3D tiles of 4^3-cell blocks (edges 4, 8, 16),
  // else the cell-owner kernel
  const bool cols_ok = p->ndim == 3 && p->tile[0] == 8 && p->tile[1] == 8 &&
                       p->tile[2] == kColsZ && r.ts >= 8;
  bool blocks_ok = p->ndim == 3;
  {
    int nblk = 1;
    for (int d = 0; d < 3 && blocks_ok; ++d) {
      blocks_ok = p->tile[d] % kBlkE == 0 && p->tile[d] <= 4 * kBlkE &&
                  p->nf[d] >= p->w + kBlkPad;
      nblk *= p->tile[d] / kBlkE;
    }
    blocks_ok = blocks_ok && nblk % 4 == 0 && nblk <= 4 * kBlkPerWave;
  }
  // 2D matrix kernel: blocks of 4 x 16 cells, at most 16 of them per tile
  const bool mat2d_ok = p->ndim == 2 && p->tile[1] % kBlkE == 0 && p->tile[2] % (4 * kBlkE) == 0 &&
                        (p->tile[1] / kBlkE) * (p->tile[2] / (4 * kBlkE)) <= 16;
  int method = p->spread_method;
  if (method == HIPNUFFT_SPREAD_AUTO)
    method = (cols_ok && p->spread_group >= 16) ? HIPNUFFT_SPREAD_COLUMNS
             : blocks_ok                        ? HIPNUFFT_SPREAD_MATRIX
             : mat2d_ok                         ? HIPNUFFT_SPREAD_MATRIX
                                                : HIPNUFFT_SPREAD_CELLS;
  if (p->ndim == 2 && method == HIPNUFFT_SPREAD_MATRIX) {
    if (!mat2d_ok)
      return set_error(HIPNUFFT_ERR_ARG,
                       "2D matrix spreading needs tiles of at most 16 blocks of 4 x 16 cells");
    // 8 transforms per workgroup, 4 waves of 4 blocks
    a.G = 8;
    const size_t chunk = (size_t)kBlkChunk;
    const size_t lds = sizeof(float) * (chunk + 1) * mfma_record_floats(8, p->tile[1] + p->tile[2]) +
                       sizeof(int16_t) * 16 * (chunk + 4);
    if (nitems > 0) {
      const dim3 grd((unsigned)nitems, (unsigned)((r.ts + 7) / 8));
      StageScope timer(p, kStageSpread);
      if ((rc = allow_lds(k_spread_mfma<2, 1, 4>, lds))) return rc;
      hipLaunchKernelGGL((k_spread_mfma<2, 1, 4>), grd, dim3(256), lds, p->stream, a);
      HN_LAUNCH_CHECK();
    }
    if (ndense > 0) {
      const dim3 rg((unsigned)ndense, (unsigned)r.ts);
      hipLaunchKernelGGL(k_spread_reduce<2>, rg, dim3(kThreads), 0, p->stream, a);
      HN_LAUNCH_CHECK();
    }
    return 0;
  }
  if (method == HIPNUFFT_SPREAD_COLUMNS && !cols_ok)
    return set_error(HIPNUFFT_ERR_ARG,
                     "column spreading needs 3D 8 x 8 x 16 tiles and >= 8 transforms per set");
  if ((method == HIPNUFFT_SPREAD_BLOCKS || method == HIPNUFFT_SPREAD_MATRIX) && !blocks_ok)
    return set_error(HIPNUFFT_ERR_ARG,
                     "block spreading needs 3D tiles of 4, 8 or 16 blocks of 4^3 cells, "
                     "at most 4 along an axis");
  // launches of at most 4 transforms: the matrix kernel with (z cell,
  // transform, re / im) columns, on blocks of 4 x 4 x (8 / GT) cells
  int GT = p->ndim != 3 ? 0 : r.ts == 1 ? 1 : r.ts == 2 ? 2 : r.ts <= 4 ? 4 : 0;
  if (GT) {
    const int zb = 8 / GT;
    const int nblk = (p->tile[0] / kBlkE) * (p->tile[1] / kBlkE) * (p->tile[2] / zb);
    if (p->tile[2] % zb != 0 || nblk > 32) GT = 0;
  }
  if (method == HIPNUFFT_SPREAD_MATRIX && GT) {
    const size_t lds =
        sizeof(float) * (size_t)kThreads * mfma_record_floats(GT, p->tile[0] + p->tile[1] + p->tile[2]);
    if (nitems > 0) {
      const dim3 grd((unsigned)nitems, (unsigned)((r.ts + GT - 1) / GT));
      StageScope timer(p, kStageSpread);
      if (GT == 1) hipLaunchKernelGGL(k_spread_mfmaz<1>, grd, dim3(kThreads), lds, p->stream, a);
      if (GT == 2) hipLaunchKernelGGL(k_spread_mfmaz<2>, grd, dim3(kThreads), lds, p->stream, a);
      if (GT == 4) hipLaunchKernelGGL(k_spread_mfmaz<4>, grd, dim3(kThreads), lds, p->stream, a);
      HN_LAUNCH_CHECK();
    }
    if (ndense > 0) {
      const dim3 rg((unsigned)ndense, (unsigned)r.ts);
      hipLaunchKernelGGL(k_spread_reduce<3>, rg, dim3(kThreads), 0, p->stream, a);
      HN_LAUNCH_CHECK();
    }
    return 0;
  }
  if (method == HIPNUFFT_SPREAD_MATRIX) {
    // 16 transforms per workgroup (8 waves) when asked for (spread_group >=
    // 16) and a launch has more than 8; measured no faster than 8 at the 3D
    // headline, so 8 by default
    const int GH = (r.ts > 8 && p->spread_group >= 16) ? 2 : 1;
    const int NW = GH == 2 ? 8 : kSpreadWaves;
    a.G = 8 * GH;
    const size_t chunk = (size_t)kBlkChunk * NW / 4;
    const size_t lds =
        sizeof(float) * (chunk + 1) *  // the records and a zero record
            mfma_record_floats(8 * GH, p->tile[0] + p->tile[1] + p->tile[2]) +
        sizeof(int16_t) * 16 * (chunk + 4);  // the per-block entry lists
    if (nitems > 0) {
      const dim3 grd((unsigned)nitems, (unsigned)((r.ts + 8 * GH - 1) / (8 * GH)));
      StageScope timer(p, kStageSpread);
      if (GH == 2) {
        if ((rc = allow_lds(k_spread_mfma<3, 2, 8>, lds))) return rc;
        hipLaunchKernelGGL((k_spread_mfma<3, 2, 8>), grd, dim3(2 * kThreads), lds, p->stream, a);
      } else {
        if ((rc = allow_lds(k_spread_mfma<3, 1, kSpreadWaves>, lds))) return rc;
        hipLaunchKernelGGL((k_spread_mfma<3, 1, kSpreadWaves>), grd, dim3(64 * kSpreadWaves), lds,
                           p->stream, a);
      }
      HN_LAUNCH_CHECK();
    }
    if (ndense > 0) {
      const dim3 rg((unsigned)ndense, (unsigned)r.ts);
      hipLaunchKernelGGL(k_spread_reduce<3>, rg, dim3(kThreads), 0, p->stream, a);
      HN_LAUNCH_CHECK();
    }
    return 0;
  }
  if (method == HIPNUFFT_SPREAD_BLOCKS) {
    int G = p->spread_group > 0 ? p->spread_group : 8;
    G = G >= 8 ? 8 : G >= 4 ? 4 : G >= 2 ? 2 : 1;
    while (G > 1 && G / 2 >= r.ts) G /= 2;
    a.G = G;
    const size_t lds =
        sizeof(float) * (size_t)kBlkChunk * blk_record_floats(p->w, G);
    if (nitems > 0) {
      const dim3 grd((unsigned)nitems, (unsigned)((r.ts + G - 1) / G));
      StageScope timer(p, kStageSpread);
      switch (G) {
        case 8: hipLaunchKernelGGL(k_spread_blocks<8>, grd, dim3(kThreads), lds, p->stream, a); break;
        case 4: hipLaunchKernelGGL(k_spread_blocks<4>, grd, dim3(kThreads), lds, p->stream, a); break;
        case 2: hipLaunchKernelGGL(k_spread_blocks<2>, grd, dim3(kThreads), lds, p->stream, a); break;
        default: hipLaunchKernelGGL(k_spread_blocks<1>, grd, dim3(kThreads), lds, p->stream, a); break;
      }
      HN_LAUNCH_CHECK();
    }
    if (ndense > 0) {
      const dim3 rg((unsigned)ndense, (unsigned)r.ts);
      hipLaunchKernelGGL(k_spread_reduce<3>, rg, dim3(kThreads), 0, p->stream, a);
      HN_LAUNCH_CHECK();
    }
    return 0;
  }
  if (method == HIPNUFFT_SPREAD_COLUMNS) {
    const size_t lds = sizeof(float) * (size_t)kColsChunk * (3 + 3 * p->w + 2 * 4 * kColsCW);
    if (nitems > 0) {
      const dim3 grd((unsigned)nitems, (unsigned)((r.ts + 4 * kColsCW - 1) / (4 * kColsCW)));
      StageScope timer(p, kStageSpread);
      hipLaunchKernelGGL(k_spread_cols, grd, dim3(kThreads), lds, p->stream, a);
      HN_LAUNCH_CHECK();
    }
    if (ndense > 0) {
      const dim3 rg((unsigned)ndense, (unsigned)r.ts);
      hipLaunchKernelGGL(k_spread_reduce<3>, rg, dim3(kThreads), 0, p->stream, a);
      HN_LAUNCH_CHECK();
    }
    return 0;
  }
  int G = p->spread_group > 0 ? p->spread_group : 8;
  G = G >= 8 ? 8 : G >= 4 ? 4 : G >= 2 ? 2 : 1;
  while (G > 1 && G / 2 >= r.ts) G /= 2;
  auto lds_of = [&](int g) {
    return sizeof(float) * (size_t)kSpreadChunk * (p->ndim * p->w + p->ndim + 2 * g);
  };
  while (G > 1 && lds_of(G) > kMaxLds) G /= 2;
  a.G = G;
  const size_t lds = lds_of(G);
  if (nitems > 0) {
    const dim3 grd((unsigned)nitems, (unsigned)((r.ts + G - 1) / G));
    StageScope timer(p, kStageSpread);
#define HN_SPREAD(DD)                                                                        \
  switch (G) {                                                                               \
    case 8: hipLaunchKernelGGL((k_spread<DD, 8>), grd, dim3(kThreads), lds, p->stream, a); break; \
    case 4: hipLaunchKernelGGL((k_spread<DD, 4>), grd, dim3(kThreads), lds, p->stream, a); break; \
    case 2: hipLaunchKernelGGL((k_spread<DD, 2>), grd, dim3(kThreads), lds, p->stream, a); break; \
    default: hipLaunchKernelGGL((k_spread<DD, 1>), grd, dim3(kThreads), lds, p->stream, a); break; \
  }
    if (p->ndim == 1) { HN_SPREAD(1) }
    else if (p->ndim == 2) { HN_SPREAD(2) }
    else { HN_SPREAD(3) }
#undef HN_SPREAD
    HN_LAUNCH_CHECK();
  }
  if (ndense > 0) {
    const dim3 rg((unsigned)ndense, (unsigned)r.ts);
    if (p->ndim == 1) hipLaunchKernelGGL(k_spread_reduce<1>, rg, dim3(kThreads), 0, p->stream, a);
    if (p->ndim == 2) hipLaunchKernelGGL(k_spread_reduce<2>, rg, dim3(kThreads), 0, p->stream, a);
    if (p->ndim == 3) hipLaunchKernelGGL(k_spread_reduce<3>, rg, dim3(kThreads), 0, p->stream, a);
    HN_LAUNCH_CHECK();
  }
  return 0;
}

int build_interp_masks(hipnufft_plan_s* p) {
  p->mask_rows = 0;
  if (p->ndim < 2 || p->nsub <= 0 || region_chunks(p) > 16) return 0;
  int rows = 1;
  for (int d = 3 - p->ndim; d < 2; ++d) rows *= region_extent(p, d);
  const int64_t need = p->nsub * (int64_t)rows;
  if (need > p->sub_mask_cap || !p->sub_mask) {
    if (p->sub_mask) HN_HIP(hipFree(p->sub_mask));
    p->sub_mask = nullptr;
    p->sub_mask_cap = 0;
    HN_HIP(hipMalloc((void**)&p->sub_mask, sizeof(uint16_t) * need));
    p->sub_mask_cap = need;
  }
  InterpArgs a{};
  a.g = pts_geom(p);
  a.sub_tile = p->sub_tile;
  a.sub_begin = p->sub_begin;
  a.bin_start = p->bin_start;
  a.max_sub = p->max_sub;
  set_region_magic(p, &a);
  const size_t lds = 2 * sizeof(uint32_t) * (size_t)rows;
  HN_DIM_SWITCH(p->ndim, HN_WIDTH_SWITCH(p->w, {
    if constexpr (DD >= 2)
      hipLaunchKernelGGL((k_interp_masks<DD, WW>), dim3((unsigned)p->nsub), dim3(kThreads), lds,
                         p->stream, a, p->sub_mask);
  }));
  HN_LAUNCH_CHECK();
  p->mask_rows = rows;
  return 0;
}

int launch_interp(const hipnufft_plan_s* p, const float2* grid, int64_t grid_stride, SetRange r,
                  float2* out, const float2* obs, const float* density, float scale, int mode) {
  int rc = check_range(p, r);
  if (rc) return rc;
  const int64_t nsub = p->sub_of_set[r.k1] - p->sub_of_set[r.k0];
  if (nsub <= 0) return 0;
  InterpArgs a{};
  a.g = pts_geom(p);
  a.sub_tile = p->sub_tile;
  a.sub_begin = p->sub_begin;
  a.bin_start = p->bin_start;
  a.max_sub = p->max_sub;
  a.sub0 = (int)p->sub_of_set[r.k0];
  a.grid = grid;
  a.grid_stride = grid_stride;
  a.out = out;
  a.obs = obs;
  a.density = density;
  a.k0 = r.k0;
  a.ts = r.ts;
  a.scale = scale;
  a.mode = mode;
  if (region_lds_bytes(p, 1) > kMaxLdsRegion)
    return set_error(HIPNUFFT_ERR_ARG, "interp tile too large for LDS");
  if (p->ndim >= 2 && p->interp_group <= 0) {
    // every transform of an item in one workgroup (k_interp_stream) when
    // its region, grid and items fit the kernel
    constexpr int NB = kStreamNB;
    const bool d3 = p->ndim == 3;
    const int NT = d3 ? kStreamNT : kStream2NT;
    int64_t cells = 1;
    for (int d = 3 - p->ndim; d < 3; ++d) cells *= region_extent(p, d);
    const int64_t rows = cells / region_extent(p, 2);
    const int64_t padded = (rows * (region_extent(p, 2) + kStreamPad) + 31) / 32 * 32;
    if (padded <= (d3 ? kStreamCells : kStream2Cells) && rows <= (d3 ? kStreamRows : kStream2Rows) &&
        grid_size(p) < ((int64_t)1 << 28) && p->max_sub <= (NT >= 1024 ? NT / 2 : NT) &&
        p->w <= kStreamMaxW) {
      set_region_magic(p, &a, NT);
      a.lds_pad = kStreamPad;
      a.lds_row_mag = magic_of(region_extent(p, 2) + kStreamPad);
      a.grid_bytes = (int)(grid_size(p) * (int64_t)sizeof(float2));
      a.sub_mask = p->mask_rows > 0 ? p->sub_mask : nullptr;
      // transforms per workgroup: all of the set's, fewer when the items
      // alone would not fill the GPU several times over.  3D: at least 1024
      // workgroups (two per CU).  2D (eight per CU): at least 16384, dense
      // items then leave no tail (stacked 128^3 slices, 2048 transforms:
      // 0.37 ms less per type 2 than with 4096), but at least 16 transforms
      // each, over which the point set-up is shared (radial 256^2, 64 coils:
      // 0.228 ms at 4 transforms, 0.195 at 16)
      int Gs = r.ts;
      if (d3) {
        while (Gs > 1 && nsub * (int64_t)((r.ts + Gs - 1) / Gs) < 1024) Gs = (Gs + 1) / 2;
      } else {
        while (Gs > 16 && nsub * (int64_t)((r.ts + Gs - 1) / Gs) < 16384) Gs = (Gs + 1) / 2;
      }
      a.G = Gs;
      const dim3 grd((unsigned)nsub, (unsigned)((r.ts + Gs - 1) / Gs));
      StageScope timer(p, kStageInterp);
      HN_WIDTH_SWITCH(p->w, {
        if constexpr (WW <= kStreamMaxW) {
          if (d3)
            hipLaunchKernelGGL((k_interp_stream<3, WW, kStreamNT, NB, kStreamCells, kStreamRows>),
                               grd, dim3(NT), 0, p->stream, a);
          else
            hipLaunchKernelGGL((k_interp_stream<2, WW, kStream2NT, NB, kStream2Cells, kStream2Rows>),
                               grd, dim3(NT), 0, p->stream, a);
        }
      });
      HN_LAUNCH_CHECK();
      return 0;
    }
  }
  // regions that do not fit the streaming kernel: one transform per
  // workgroup in 3D, four in 2D
  const int g0 = p->interp_group > 0 ? p->interp_group : p->ndim == 2 ? 4 : 1;
  int G = std::max(1, std::min(g0, r.ts));
  while (G > 1 && region_lds_bytes(p, G) > kMaxLds) --G;
  a.G = G;
  const size_t lds = region_lds_bytes(p, G);
  // regions past 64 KiB leave room for two workgroups per CU: 512 threads
  // each keep as many waves resident as four 256-thread workgroups
  const int nt = lds > kMaxLds ? 2 * kThreads : kThreads;
  set_region_magic(p, &a, nt);
  const dim3 grd((unsigned)nsub, (unsigned)((r.ts + G - 1) / G));
  StageScope timer(p, kStageInterp);
  HN_DIM_SWITCH(p->ndim, HN_WIDTH_SWITCH(p->w, {
    if (nt == kThreads) {
      if (a.mask_stage)
        hipLaunchKernelGGL((k_interp<DD, WW, kThreads, true>), grd, dim3(kThreads), lds,
                           p->stream, a);
      else
        hipLaunchKernelGGL((k_interp<DD, WW, kThreads, false>), grd, dim3(kThreads), lds,
                           p->stream, a);
    } else if (a.mask_stage) {
      if ((rc = allow_lds(k_interp<DD, WW, 2 * kThreads, true>, lds))) return rc;
      hipLaunchKernelGGL((k_interp<DD, WW, 2 * kThreads, true>), grd, dim3(2 * kThreads), lds,
                         p->stream, a);
    } else {
      if ((rc = allow_lds(k_interp<DD, WW, 2 * kThreads, false>, lds))) return rc;
      hipLaunchKernelGGL((k_interp<DD, WW, 2 * kThreads, false>), grd, dim3(2 * kThreads), lds,
                         p->stream, a);
    }
  }));
  HN_LAUNCH_CHECK();
  return 0;
}

int launch_grad_traj(const hipnufft_plan_s* p, const float2* grid, int64_t grid_stride,
                     SetRange r, const float2* dy, float scale, float* gtraj) {
  int rc = check_range(p, r);
  if (rc) return rc;
  const int64_t nsub = p->sub_of_set[r.k1] - p->sub_of_set[r.k0];
  if (nsub <= 0) return 0;
  InterpArgs a{};
  a.g = pts_geom(p);
  a.sub_tile = p->sub_tile;
  a.sub_begin = p->sub_begin;
  a.bin_start = p->bin_start;
  a.max_sub = p->max_sub;
  a.sub0 = (int)p->sub_of_set[r.k0];
  a.grid = grid;
  a.grid_stride = grid_stride;
  a.obs = dy;
  a.gtraj = gtraj;
  a.scale = scale;
  a.k0 = r.k0;
  a.ts = r.ts;
  set_region_magic(p, &a);
  const int o = 3 - p->ndim;
  for (int d = 0; d < p->ndim; ++d) a.dscale[d] = (float)((double)p->nf[o + d] / (2.0 * M_PI));
  // each workgroup handles the transforms of its set that fit LDS; the
  // launches accumulate into gtraj in stream order
  if (region_lds_bytes(p, 1) > kMaxLdsRegion)
    return set_error(HIPNUFFT_ERR_ARG, "interp tile too large for LDS");
  int G = 1;
  while (region_lds_bytes(p, G + 1) <= kMaxLds) ++G;
  for (int tg0 = 0; tg0 < r.ts; tg0 += G) {
    const int Tg = std::min(G, r.ts - tg0);
    a.tg0 = tg0;
    a.G = Tg;
    const size_t lds = region_lds_bytes(p, Tg);
    StageScope timer(p, kStageGradTraj);
    HN_DIM_SWITCH(p->ndim, HN_WIDTH_SWITCH(p->w, {
      if ((rc = allow_lds(k_grad_traj<DD, WW>, lds))) return rc;
      hipLaunchKernelGGL((k_grad_traj<DD, WW>), dim3((unsigned)nsub), dim3(kThreads), lds,
                         p->stream, a);
    }));
    HN_LAUNCH_CHECK();
  }
  return 0;
}

}  // namespace hn
