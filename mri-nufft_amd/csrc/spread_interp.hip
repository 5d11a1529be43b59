// Spreading (type 1) and interpolation (type 2) with the exponential of
// semicircle kernel, tiled through LDS.
//
// Replaces the spread / interp stages of cufinufft's Plan.execute
// (operators/interfaces/cufinufft.py:127-133) and of finufft
// (operators/interfaces/finufft.py:64-76).
//
// Spreading is output-driven: one workgroup owns one tile of the fine grid
// for a group of transforms, accumulates the footprints of every point that
// touches the tile (the spread lists built at setpts, one entry per
// point/tile pair) into LDS with ds_add_f32, and writes the finished tile to
// HBM exactly once with coalesced stores -- no global atomics and no
// zero-fill pass over the grid.
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
  const float2* in;
  int64_t in_stride;
  const float* density;
  float2* grid;
  int64_t grid_stride;
  int T, G;
};

struct InterpArgs {
  PtsGeom g;
  const int32_t* sub_tile;
  const int32_t* sub_begin;
  const int32_t* bin_start;
  int max_sub;
  const float2* grid;
  int64_t grid_stride;
  float2* out;
  int64_t out_stride;
  const float2* obs;
  const float* density;
  float* gtraj;
  float dscale[3];  // radians -> cells factor per axis (gradient)
  int T, G;
  float scale;
  int mode;
};

template <int D>
__device__ __forceinline__ void tile_coords(int key, const int* nt, int* tc) {
#pragma unroll
  for (int d = D - 1; d >= 0; --d) {
    tc[d] = key % nt[d];
    key /= nt[d];
  }
}

// ---------------------------------------------------------------------------
// spreading
// ---------------------------------------------------------------------------
template <int D, int W>
__global__ __launch_bounds__(kThreads) void k_spread(SpreadArgs a) {
  extern __shared__ float2 lds[];
  float* ldsf = reinterpret_cast<float*>(lds);
  const int tile = blockIdx.x;
  const int t0 = blockIdx.y * a.G;
  const int G = min(a.G, a.T - t0);
  const float inv_hw = 2.0f / W, beta = a.g.beta;

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
  for (int c = threadIdx.x; c < G * TC; c += kThreads) lds[c] = make_float2(0.f, 0.f);
  __syncthreads();

  const int e0 = a.list_start[tile], e1 = a.list_start[tile + 1];
  for (int e = e0 + threadIdx.x; e < e1; e += kThreads) {
    const int p = a.list_pts[e];
    const int idx = a.g.idx[p];
    int li[D][W];
    float wt[D][W];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int c0 = a.g.cell[d][p];
      const float off = a.g.off[d][p];
#pragma unroll
      for (int j = 0; j < W; ++j) {
        int c = c0 + j;
        if (c >= a.g.nf[d]) c -= a.g.nf[d];
        const int l = c - org[d];
        li[d][j] = ((unsigned)l < (unsigned)ext[d]) ? l : -1;
        wt[d][j] = es_eval(off - (float)j, inv_hw, beta);
      }
    }
    const float dens = a.density ? a.density[idx] : 1.0f;
    for (int gg = 0; gg < G; ++gg) {
      float2 v = a.in[(int64_t)(t0 + gg) * a.in_stride + idx];
      v.x *= dens;
      v.y *= dens;
      float* base = ldsf + 2 * gg * TC;
      if constexpr (D == 1) {
#pragma unroll
        for (int j = 0; j < W; ++j)
          if (li[0][j] >= 0) {
            atomicAdd(base + 2 * li[0][j], wt[0][j] * v.x);
            atomicAdd(base + 2 * li[0][j] + 1, wt[0][j] * v.y);
          }
      } else if constexpr (D == 2) {
#pragma unroll
        for (int j0 = 0; j0 < W; ++j0) {
          if (li[0][j0] < 0) continue;
          const float vx = wt[0][j0] * v.x, vy = wt[0][j0] * v.y;
          float* row = base + 2 * li[0][j0] * td[1];
#pragma unroll
          for (int j1 = 0; j1 < W; ++j1)
            if (li[1][j1] >= 0) {
              atomicAdd(row + 2 * li[1][j1], wt[1][j1] * vx);
              atomicAdd(row + 2 * li[1][j1] + 1, wt[1][j1] * vy);
            }
        }
      } else {
        // axis 0 is looped (not unrolled): its weight is recomputed per tap
        // so no register array is indexed with a runtime index
        const int c00 = a.g.cell[0][p];
        const float off0 = a.g.off[0][p];
#pragma unroll 1
        for (int j0 = 0; j0 < W; ++j0) {
          int c = c00 + j0;
          if (c >= a.g.nf[0]) c -= a.g.nf[0];
          const int l0 = c - org[0];
          if ((unsigned)l0 >= (unsigned)ext[0]) continue;
          const float w0 = es_eval(off0 - (float)j0, inv_hw, beta);
          const float ux = w0 * v.x, uy = w0 * v.y;
#pragma unroll
          for (int j1 = 0; j1 < W; ++j1) {
            if (li[1][j1] < 0) continue;
            const float vx = wt[1][j1] * ux, vy = wt[1][j1] * uy;
            float* row = base + 2 * (l0 * td[1] + li[1][j1]) * td[2];
#pragma unroll
            for (int j2 = 0; j2 < W; ++j2)
              if (li[2][j2] >= 0) {
                atomicAdd(row + 2 * li[2][j2], wt[2][j2] * vx);
                atomicAdd(row + 2 * li[2][j2] + 1, wt[2][j2] * vy);
              }
          }
        }
      }
    }
  }
  __syncthreads();

  for (int c = threadIdx.x; c < G * TC; c += kThreads) {
    const int gg = c / TC;
    int r = c - gg * TC;
    int l[3];
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
      l[d] = r % td[d];
      r /= td[d];
    }
    bool ok = true;
    int64_t gi = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      ok &= l[d] < ext[d];
      gi = gi * a.g.nf[d] + org[d] + l[d];
    }
    if (ok) a.grid[(int64_t)(t0 + gg) * a.grid_stride + gi] = lds[c];
  }
}

// ---------------------------------------------------------------------------
// interpolation (and the kernel-derivative gradient)
// ---------------------------------------------------------------------------
template <int D, int W>
__device__ __forceinline__ void stage_region(const InterpArgs& a, float2* lds, int tile, int t0,
                                             int G, int* rorg, int* R, int& RC) {
  int tc[3];
  tile_coords<D>(tile, a.g.nt, tc);
  RC = 1;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    R[d] = a.g.tile[d] + W;
    rorg[d] = tc[d] * a.g.tile[d] - W / 2;
    RC *= R[d];
  }
  for (int c = threadIdx.x; c < G * RC; c += kThreads) {
    const int gg = c / RC;
    int r = c - gg * RC;
    int l[3];
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
      l[d] = r % R[d];
      r /= R[d];
    }
    int64_t gi = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      int x = (rorg[d] + l[d]) % a.g.nf[d];
      if (x < 0) x += a.g.nf[d];
      gi = gi * a.g.nf[d] + x;
    }
    lds[c] = a.grid[(int64_t)(t0 + gg) * a.grid_stride + gi];
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
    l = min(max(l, 0), a.g.tile[d]);  // memory safety; a sorted point always satisfies it
    ls[d] = l;
    const float off = a.g.off[d][p];
#pragma unroll
    for (int j = 0; j < W; ++j) wt[d][j] = es_eval(off - (float)j, inv_hw, a.g.beta);
  }
}

template <int D, int W>
__global__ __launch_bounds__(kThreads) void k_interp(InterpArgs a) {
  extern __shared__ float2 lds[];
  const int sub = blockIdx.x;
  const int tile = a.sub_tile[sub];
  const int pb = a.sub_begin[sub];
  const int pe = min(pb + a.max_sub, a.bin_start[tile + 1]);
  const int t0 = blockIdx.y * a.G;
  const int G = min(a.G, a.T - t0);
  int rorg[3], R[3], RC;
  stage_region<D, W>(a, lds, tile, t0, G, rorg, R, RC);
  __syncthreads();

  for (int p = pb + threadIdx.x; p < pe; p += kThreads) {
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
      } else if constexpr (D == 2) {
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
      } else {
        const float off0 = a.g.off[0][p];
        const float inv_hw = 2.0f / W;
#pragma unroll 1
        for (int j0 = 0; j0 < W; ++j0) {
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
      const int64_t o = (int64_t)(t0 + gg) * a.out_stride + idx;
      float2 r = make_float2(a.scale * sx, a.scale * sy);
      if (a.mode == 1) {
        const float2 y = a.obs[o];
        r.x = (r.x - y.x) * dens;
        r.y = (r.y - y.y) * dens;
      }
      a.out[o] = r;
    }
  }
}

// gtraj[idx*D + d] += Re sum_t conj(dy[t][idx]) * scale * d/dx_d interp_t
template <int D, int W>
__global__ __launch_bounds__(kThreads) void k_grad_traj(InterpArgs a) {
  extern __shared__ float2 lds[];
  const int sub = blockIdx.x;
  const int tile = a.sub_tile[sub];
  const int pb = a.sub_begin[sub];
  const int pe = min(pb + a.max_sub, a.bin_start[tile + 1]);
  const int G = a.T;  // the whole chunk in one workgroup
  int rorg[3], R[3], RC;
  stage_region<D, W>(a, lds, tile, 0, G, rorg, R, RC);
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
      const float2 y = a.obs[(int64_t)gg * a.out_stride + idx];  // dy
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
  g.idx = p->pt_idx;
  return g;
}

static int tile_cells(const hipnufft_plan_s* p, int halo) {
  int c = 1;
  for (int d = 3 - p->ndim; d < 3; ++d) c *= p->tile[d] + halo;
  return c;
}

int launch_spread(const hipnufft_plan_s* p, const float2* in, const float* density, int T,
                  float2* grid, int64_t grid_stride) {
  if (T <= 0) return 0;
  SpreadArgs a{};
  a.g = pts_geom(p);
  a.list_pts = p->list_pts;
  a.list_start = p->list_start;
  a.in = in;
  a.in_stride = p->M;
  a.density = density;
  a.grid = grid;
  a.grid_stride = grid_stride;
  a.T = T;
  const size_t cell_bytes = sizeof(float2) * (size_t)tile_cells(p, 0);
  int G = p->spread_group;
  while (G > 1 && cell_bytes * G > kMaxLds) --G;
  a.G = G;
  const size_t lds = cell_bytes * G;
  const dim3 grd((unsigned)p->nbins, (unsigned)((T + G - 1) / G));
  HN_DIM_SWITCH(p->ndim, HN_WIDTH_SWITCH(p->w, {
    hipLaunchKernelGGL((k_spread<DD, WW>), grd, dim3(kThreads), lds, p->stream, a);
  }));
  HN_LAUNCH_CHECK();
  return 0;
}

int launch_interp(const hipnufft_plan_s* p, const float2* grid, int64_t grid_stride, int T,
                  float2* out, const float2* obs, const float* density, float scale, int mode) {
  if (T <= 0 || p->nsub == 0) return 0;
  InterpArgs a{};
  a.g = pts_geom(p);
  a.sub_tile = p->sub_tile;
  a.sub_begin = p->sub_begin;
  a.bin_start = p->bin_start;
  a.max_sub = p->max_sub;
  a.grid = grid;
  a.grid_stride = grid_stride;
  a.out = out;
  a.out_stride = p->M;
  a.obs = obs;
  a.density = density;
  a.T = T;
  a.scale = scale;
  a.mode = mode;
  const size_t cell_bytes = sizeof(float2) * (size_t)tile_cells(p, p->w);
  if (cell_bytes > kMaxLds) return set_error(HIPNUFFT_ERR_ARG, "interp tile too large for LDS");
  int G = p->interp_group;
  while (G > 1 && cell_bytes * G > kMaxLds) --G;
  a.G = G;
  const size_t lds = cell_bytes * G;
  const dim3 grd((unsigned)p->nsub, (unsigned)((T + G - 1) / G));
  HN_DIM_SWITCH(p->ndim, HN_WIDTH_SWITCH(p->w, {
    hipLaunchKernelGGL((k_interp<DD, WW>), grd, dim3(kThreads), lds, p->stream, a);
  }));
  HN_LAUNCH_CHECK();
  return 0;
}

int launch_grad_traj(const hipnufft_plan_s* p, const float2* grid, int64_t grid_stride, int T,
                     const float2* dy, float scale, float* gtraj) {
  if (T <= 0 || p->nsub == 0) return 0;
  InterpArgs a{};
  a.g = pts_geom(p);
  a.sub_tile = p->sub_tile;
  a.sub_begin = p->sub_begin;
  a.bin_start = p->bin_start;
  a.max_sub = p->max_sub;
  a.grid = grid;
  a.grid_stride = grid_stride;
  a.out_stride = p->M;
  a.obs = dy;
  a.gtraj = gtraj;
  a.scale = scale;
  const int o = 3 - p->ndim;
  for (int d = 0; d < p->ndim; ++d) a.dscale[d] = (float)((double)p->nf[o + d] / (2.0 * M_PI));
  const size_t cell_bytes = sizeof(float2) * (size_t)tile_cells(p, p->w);
  // process the chunk in groups that fit LDS; each launch accumulates
  int G = (int)(kMaxLds / cell_bytes);
  if (G < 1) return set_error(HIPNUFFT_ERR_ARG, "interp tile too large for LDS");
  for (int t0 = 0; t0 < T; t0 += G) {
    const int Tg = min(G, T - t0);
    a.grid = grid + (int64_t)t0 * grid_stride;
    a.obs = dy + (int64_t)t0 * p->M;
    a.T = Tg;
    const size_t lds = cell_bytes * Tg;
    HN_DIM_SWITCH(p->ndim, HN_WIDTH_SWITCH(p->w, {
      hipLaunchKernelGGL((k_grad_traj<DD, WW>), dim3((unsigned)p->nsub), dim3(kThreads), lds,
                         p->stream, a);
    }));
    HN_LAUNCH_CHECK();
  }
  return 0;
}

}  // namespace hn
