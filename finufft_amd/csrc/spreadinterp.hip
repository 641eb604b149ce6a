// Spreading (type-1 direction) and interpolation (type-2 direction) kernels.
//
// Method 1 (point driven, the reference's GM / GM-sort,
// src/cuda/spread_nupts_driven_inst.cu and interp_nupts_driven_inst.cu): one
// lane per nonuniform point, visiting points in the sorted order so that
// neighbouring lanes touch neighbouring grid cells; spreading adds each of the
// ns^d contributions to the fine grid with a hardware float atomic.
//
// Method 2 (LDS subproblems, the reference's SM, spread_subprob_inst.cu and
// interp_subprob_inst.cu): one workgroup per subproblem (<= maxsub points of
// one bin). Spreading accumulates into a padded bin tile held in LDS and adds
// the tile to the grid once; interpolation stages the padded tile from HBM
// into LDS and gathers from it.
//
// The kernels are templates on the precision T, dimension DIM and the kernel
// width NS; this file is compiled once per (T, DIM) with FA_REAL / FA_DIM.
#include "spreadinterp.hpp"

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>

#include "common.hpp"

#ifndef FA_DIM
#error "FA_DIM must be 1, 2 or 3"
#endif
#ifndef FA_REAL
#error "FA_REAL must be float or double"
#endif

namespace fa {
namespace {

constexpr int SPREAD_THREADS = 256;

template <typename T>
__device__ __forceinline__ void lds_add(cplx_t<T> *p, T re, T im) {
  atomicAdd(&p->x, re);
  atomicAdd(&p->y, im);
}
template <typename T>
__device__ __forceinline__ void hbm_add(cplx_t<T> *p, T re, T im) {
  unsafeAtomicAdd(&p->x, re);
  unsafeAtomicAdd(&p->y, im);
}

__device__ __forceinline__ int wrap_once(int i, int n) { return i >= n ? i - n : i; }

// ---------------------------------------------------------------- method 1

// Coordinate d of the i-th point in sorted order.
template <typename T>
__device__ __forceinline__ T coord(const SpreadCtx<T> &ctx, int d, int i, int j) {
  return ctx.sxyz[d] ? ctx.sxyz[d][i] : ctx.xyz[d][j];
}


template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(SPREAD_THREADS)
    spread_gm_kernel(SpreadCtx<T> ctx, const cplx_t<T> *__restrict__ c,
                     cplx_t<T> *__restrict__ fw, int64_t nftot) {
  const int i = blockIdx.x * SPREAD_THREADS + threadIdx.x;
  if (i >= ctx.M) return;
  const int t = blockIdx.y;
  const int j = ctx.idx[i];
  T ker[DIM][NS];
  int st[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    st[d] = start_and_kernel<T, NS>(coord(ctx, d, i, j), ctx.nf[d], ker[d], ctx.kp);
    if (st[d] < 0) st[d] += ctx.nf[d];
  }
  const cplx_t<T> cv = c[int64_t(t) * ctx.M + j];
  cplx_t<T> *g = fw + int64_t(t) * nftot;
  if constexpr (DIM == 1) {
#pragma unroll
    for (int a = 0; a < NS; ++a) {
      const int ix = wrap_once(st[0] + a, ctx.nf[0]);
      hbm_add<T>(g + ix, cv.x * ker[0][a], cv.y * ker[0][a]);
    }
  } else if constexpr (DIM == 2) {
    for (int b = 0; b < NS; ++b) {
      const int iy = wrap_once(st[1] + b, ctx.nf[1]);
      const T wr = cv.x * ker[1][b], wi = cv.y * ker[1][b];
      cplx_t<T> *row = g + int64_t(iy) * ctx.nf[0];
#pragma unroll
      for (int a = 0; a < NS; ++a) {
        const int ix = wrap_once(st[0] + a, ctx.nf[0]);
        hbm_add<T>(row + ix, wr * ker[0][a], wi * ker[0][a]);
      }
    }
  } else {
    for (int e = 0; e < NS; ++e) {
      const int iz = wrap_once(st[2] + e, ctx.nf[2]);
      const T zr = cv.x * ker[2][e], zi = cv.y * ker[2][e];
      for (int b = 0; b < NS; ++b) {
        const int iy = wrap_once(st[1] + b, ctx.nf[1]);
        const T wr = zr * ker[1][b], wi = zi * ker[1][b];
        cplx_t<T> *row = g + (int64_t(iz) * ctx.nf[1] + iy) * ctx.nf[0];
#pragma unroll
        for (int a = 0; a < NS; ++a) {
          const int ix = wrap_once(st[0] + a, ctx.nf[0]);
          hbm_add<T>(row + ix, wr * ker[0][a], wi * ker[0][a]);
        }
      }
    }
  }
}

template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(SPREAD_THREADS)
    interp_gm_kernel(SpreadCtx<T> ctx, cplx_t<T> *__restrict__ c,
                     const cplx_t<T> *__restrict__ fw, int64_t nftot, int blk) {
  const int i = blockIdx.x * SPREAD_THREADS + threadIdx.x;
  if (i >= ctx.M) return;
  const int j = ctx.idx[i];
  T ker[DIM][NS];
  int st[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    st[d] = start_and_kernel<T, NS>(coord(ctx, d, i, j), ctx.nf[d], ker[d], ctx.kp);
    if (st[d] < 0) st[d] += ctx.nf[d];
  }
  int xi[NS];
#pragma unroll
  for (int a = 0; a < NS; ++a) xi[a] = wrap_once(st[0] + a, ctx.nf[0]);
  for (int t = 0; t < blk; ++t) {
    const cplx_t<T> *g = fw + int64_t(t) * nftot;
    T ar = 0, ai = 0;
    if constexpr (DIM == 1) {
#pragma unroll
      for (int a = 0; a < NS; ++a) {
        const cplx_t<T> v = g[xi[a]];
        ar = std::fma(v.x, ker[0][a], ar);
        ai = std::fma(v.y, ker[0][a], ai);
      }
    } else if constexpr (DIM == 2) {
      for (int b = 0; b < NS; ++b) {
        const int iy = wrap_once(st[1] + b, ctx.nf[1]);
        const cplx_t<T> *row = g + int64_t(iy) * ctx.nf[0];
        T rr = 0, ri = 0;
#pragma unroll
        for (int a = 0; a < NS; ++a) {
          const cplx_t<T> v = row[xi[a]];
          rr = std::fma(v.x, ker[0][a], rr);
          ri = std::fma(v.y, ker[0][a], ri);
        }
        ar = std::fma(rr, ker[1][b], ar);
        ai = std::fma(ri, ker[1][b], ai);
      }
    } else {
      for (int e = 0; e < NS; ++e) {
        const int iz = wrap_once(st[2] + e, ctx.nf[2]);
        T pr = 0, pi = 0;
        for (int b = 0; b < NS; ++b) {
          const int iy = wrap_once(st[1] + b, ctx.nf[1]);
          const cplx_t<T> *row = g + (int64_t(iz) * ctx.nf[1] + iy) * ctx.nf[0];
          T rr = 0, ri = 0;
#pragma unroll
          for (int a = 0; a < NS; ++a) {
            const cplx_t<T> v = row[xi[a]];
            rr = std::fma(v.x, ker[0][a], rr);
            ri = std::fma(v.y, ker[0][a], ri);
          }
          pr = std::fma(rr, ker[1][b], pr);
          pi = std::fma(ri, ker[1][b], pi);
        }
        ar = std::fma(pr, ker[2][e], ar);
        ai = std::fma(pi, ker[2][e], ai);
      }
    }
    c[int64_t(t) * ctx.M + j] = cplx_t<T>{ar, ai};
  }
}

// ---------------------------------------------------------------- method 2

constexpr int BIN_THREADS = 512;

// Grid coordinate of a tile coordinate g (tile offset + position). A tile
// extends at most binsize + ns - 2 cells past the start of its bin, bins start
// below nf, and binsize <= nf, ns < nf: g < 3 nf, so two conditional
// subtractions bring it into [0, nf).
__device__ __forceinline__ int wrap_grid(int g, int nf) {
  g = g >= nf ? g - nf : g;
  return g >= nf ? g - nf : g;
}

template <typename T>
using vec2_t = T __attribute__((ext_vector_type(2)));

// One LDS work unit of method 2: a bin (all of its points) or a subproblem
// (<= maxsub points of one bin). Bins are bins of the first support node
// (the sort's start_bins mode), so a bin of b nodes owns exactly the padded
// tile [off, off + b + ns - 1) in each dimension.
template <int DIM, int NS>
struct StartTile {
  int first, npts;
  int bin;
  int off[3];
  int P[3];
  int ntile;

  template <typename T>
  __device__ void set_geometry(const SpreadCtx<T> &ctx, int b) {
    bin = b;
    int rem = b;
    ntile = 1;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int bd = rem % ctx.nbins[d];
      rem /= ctx.nbins[d];
      off[d] = bd * ctx.binsize[d];
      P[d] = d < DIM ? ctx.binsize[d] + NS - 1 : 1;
      ntile *= P[d];
    }
  }

  // Work unit = subproblem s (false when s is past the device-side count).
  template <typename T>
  __device__ bool init_sub(const SpreadCtx<T> &ctx, int s) {
    const int nb_tot = ctx.nbins[0] * ctx.nbins[1] * ctx.nbins[2];
    if (s >= ctx.substart[nb_tot]) return false;
    const int b = ctx.sub2bin[s];
    first = ctx.binstart[b] + (s - ctx.substart[b]) * ctx.maxsub;
    npts = min(ctx.maxsub, ctx.binstart[b + 1] - first);
    set_geometry(ctx, b);
    return true;
  }

  // Work unit = whole bin b.
  template <typename T>
  __device__ void init_bin(const SpreadCtx<T> &ctx, int b) {
    first = ctx.binstart[b];
    npts = ctx.binstart[b + 1] - first;
    set_geometry(ctx, b);
  }

  // Tile offset of a point's first node s; in [0, binsize) by construction of
  // the sort key (the clamp only guards memory).
  __device__ __forceinline__ int local_start(int s, int d, int nf) const {
    const int l = (s < 0 ? s + nf : s) - off[d];
    return min(max(l, 0), P[d] - NS);
  }

  // Fine-grid index of tile cell n.
  template <typename T>
  __device__ int64_t grid_index(const SpreadCtx<T> &ctx, int n) const {
    int64_t gi = 0, stride = 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const int nd = n % P[d];
      n /= P[d];
      const int g = wrap_grid(nd + off[d], ctx.nf[d]);
      gi += g * stride;
      stride *= ctx.nf[d];
    }
    return gi;
  }
};

// Walks a tile in steps of NT cells: each thread's first cell and the step in
// tile coordinates are found once per unit, so that the walk needs no
// division per cell.
template <int DIM, int NT>
struct TileWalk {
  int x0, y0, z0, sx, sy, sz;
  __device__ void init(const int *P) {
    int r = threadIdx.x;
    x0 = r % P[0];
    r /= P[0];
    y0 = DIM > 1 ? r % P[1] : 0;
    z0 = DIM > 1 ? r / P[1] : 0;
    r = NT;
    sx = r % P[0];
    r /= P[0];
    sy = DIM > 1 ? r % P[1] : 0;
    sz = DIM > 1 ? r / P[1] : 0;
  }
  // Grid index of tile cell (x, y, z) of a tile at off (< 2^31: the plan
  // caps a vector's grid there), then the next cell of the walk.
  template <typename T>
  __device__ __forceinline__ int index_and_step(const SpreadCtx<T> &ctx, const int *P,
                                                const int *off, int &x, int &y, int &z) const {
    int gi = wrap_grid(x + off[0], ctx.nf[0]);
    if (DIM > 1) {
      int row = wrap_grid(y + off[1], ctx.nf[1]);
      if (DIM > 2) row += wrap_grid(z + off[2], ctx.nf[2]) * ctx.nf[1];
      gi += row * ctx.nf[0];
    }
    x += sx;
    if (x >= P[0]) {
      x -= P[0];
      ++y;
    }
    if (DIM > 1) {
      y += sy;
      if (y >= P[1]) {
        y -= P[1];
        ++z;
      }
      z += sz;
    }
    return gi;
  }
};

// Cells per thread whose global accesses a tile transfer keeps in flight:
// the loads of a chunk are all issued before any of them is used (a loop of
// single read-modify-writes would wait a full memory latency per cell,
// since a load may not pass an earlier store to the same array).
constexpr int TILE_CHUNK = 16;

// Adds an fp64 LDS tile (re, im planes) to grid g. plain: read-modify-write
// (the unit owns these cells for the launch: colored bins); otherwise
// hardware atomics (zero cells skipped).
template <int DIM, int NT, bool PLAIN, typename T>
__device__ __forceinline__ void tile_to_grid(const SpreadCtx<T> &ctx, const int *P,
                                             const int *off, int ntile, const double *re,
                                             const double *im, cplx_t<T> *__restrict__ g) {
  TileWalk<DIM, NT> w;
  w.init(P);
  int x = w.x0, y = w.y0, z = w.z0;
  for (int base = threadIdx.x; base < ntile; base += TILE_CHUNK * NT) {
    int gi[TILE_CHUNK];
    cplx_t<T> v[TILE_CHUNK];
#pragma unroll
    for (int u = 0; u < TILE_CHUNK; ++u) {
      gi[u] = w.index_and_step(ctx, P, off, x, y, z);
      if (PLAIN && base + u * NT < ntile) v[u] = g[gi[u]];
    }
#pragma unroll
    for (int u = 0; u < TILE_CHUNK; ++u) {
      const int n = base + u * NT;
      if (n < ntile) {
        const T vr = T(re[n]), vi = T(im[n]);
        if (PLAIN) {
          v[u].x += vr;
          v[u].y += vi;
          g[gi[u]] = v[u];
        } else if (vr != T(0) || vi != T(0)) {
          hbm_add<T>(g + gi[u], vr, vi);
        }
      }
    }
  }
}

// Stages the tile of grid g into LDS (interleaved complex).
template <int DIM, int NT, typename T>
__device__ __forceinline__ void grid_to_tile(const SpreadCtx<T> &ctx, const int *P,
                                             const int *off, int ntile,
                                             const cplx_t<T> *__restrict__ g, cplx_t<T> *tile) {
  TileWalk<DIM, NT> w;
  w.init(P);
  int x = w.x0, y = w.y0, z = w.z0;
  for (int base = threadIdx.x; base < ntile; base += TILE_CHUNK * NT) {
    cplx_t<T> v[TILE_CHUNK];
#pragma unroll
    for (int u = 0; u < TILE_CHUNK; ++u) {
      const int gi = w.index_and_step(ctx, P, off, x, y, z);
      if (base + u * NT < ntile) v[u] = g[gi];
    }
#pragma unroll
    for (int u = 0; u < TILE_CHUNK; ++u)
      if (base + u * NT < ntile) tile[base + u * NT] = v[u];
  }
}

// Spreads the unit's points into the LDS tile held as two fp64 planes.
// Accumulating in fp64 is what makes LDS spreading fast on this GPU: its fp64
// LDS atomic add runs about 20x faster than the fp32 one, and the sums are
// exact to fp32 rounding of each contribution.
template <typename T, int DIM, int NS>
__device__ __forceinline__ void spread_into_tile(const SpreadCtx<T> &ctx,
                                                 const StartTile<DIM, NS> &st,
                                                 const cplx_t<T> *__restrict__ c, double *re,
                                                 double *im) {
  for (int i = threadIdx.x; i < st.npts; i += BIN_THREADS) {
    const int q = st.first + i;
    const int j = ctx.idx[q];
    T ker[DIM][NS];
    int l[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      l[d] = st.local_start(start_and_kernel<T, NS>(coord(ctx, d, q, j), ctx.nf[d], ker[d], ctx.kp),
                            d, ctx.nf[d]);
    const cplx_t<T> cv = c[j];
    if constexpr (DIM == 1) {
#pragma unroll
      for (int a = 0; a < NS; ++a) {
        atomicAdd(re + l[0] + a, double(cv.x * ker[0][a]));
        atomicAdd(im + l[0] + a, double(cv.y * ker[0][a]));
      }
    } else if constexpr (DIM == 2) {
      for (int b = 0; b < NS; ++b) {
        const T wr = cv.x * ker[1][b], wi = cv.y * ker[1][b];
        const int row = (l[1] + b) * st.P[0] + l[0];
#pragma unroll
        for (int a = 0; a < NS; ++a) {
          atomicAdd(re + row + a, double(wr * ker[0][a]));
          atomicAdd(im + row + a, double(wi * ker[0][a]));
        }
      }
    } else {
      for (int e = 0; e < NS; ++e) {
        const T zr = cv.x * ker[2][e], zi = cv.y * ker[2][e];
        for (int b = 0; b < NS; ++b) {
          const T wr = zr * ker[1][b], wi = zi * ker[1][b];
          const int row = ((l[2] + e) * st.P[1] + l[1] + b) * st.P[0] + l[0];
#pragma unroll
          for (int a = 0; a < NS; ++a) {
            atomicAdd(re + row + a, double(wr * ker[0][a]));
            atomicAdd(im + row + a, double(wi * ker[0][a]));
          }
        }
      }
    }
  }
}

// Bin of the blk-th workgroup of the launch of parity color (col0, col1, col2).
template <int DIM, typename T>
__device__ __forceinline__ int colored_bin(const SpreadCtx<T> &ctx, int blk, int col0, int col1,
                                           int col2) {
  const int col[3] = {col0, col1, col2};
  int rem = blk, b = 0, stride = 1;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int hb = d < DIM ? (ctx.nbins[d] - col[d] + 1) / 2 : 1;
    const int bd = d < DIM ? 2 * (rem % hb) + col[d] : 0;
    rem /= hb;
    b += bd * stride;
    stride *= ctx.nbins[d];
  }
  return b;
}

// Colored whole-bin spreading: the launch covers the bins whose index parity
// in every dimension equals col[]; their tiles are disjoint (the plan checks an
// even number of equal bins of >= ns - 1 nodes per dimension), so each tile is
// added to the grid with plain loads and stores.
template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(BIN_THREADS)
    spread_bin_colored_kernel(SpreadCtx<T> ctx, const cplx_t<T> *__restrict__ c,
                              cplx_t<T> *__restrict__ fw, int64_t nftot, int col0, int col1,
                              int col2) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double *re = reinterpret_cast<double *>(lds_raw);
  StartTile<DIM, NS> st;
  st.init_bin(ctx, colored_bin<DIM>(ctx, blockIdx.x, col0, col1, col2));
  if (st.npts == 0) return;
  double *im = re + st.ntile;
  const int t = blockIdx.y;
  for (int n = threadIdx.x; n < st.ntile; n += BIN_THREADS) {
    re[n] = 0.0;
    im[n] = 0.0;
  }
  __syncthreads();
  spread_into_tile<T, DIM, NS>(ctx, st, c + int64_t(t) * ctx.M, re, im);
  __syncthreads();
  tile_to_grid<DIM, BIN_THREADS, true>(ctx, st.P, st.off, st.ntile, re, im,
                                       fw + int64_t(t) * nftot);
}

// Subproblem spreading with an atomic flush (any geometry, any density).
template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(BIN_THREADS)
    spread_sub_kernel(SpreadCtx<T> ctx, const cplx_t<T> *__restrict__ c,
                      cplx_t<T> *__restrict__ fw, int64_t nftot) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double *re = reinterpret_cast<double *>(lds_raw);
  StartTile<DIM, NS> st;
  if (!st.init_sub(ctx, blockIdx.x)) return;
  double *im = re + st.ntile;
  const int t = blockIdx.y;
  for (int n = threadIdx.x; n < st.ntile; n += BIN_THREADS) {
    re[n] = 0.0;
    im[n] = 0.0;
  }
  __syncthreads();
  spread_into_tile<T, DIM, NS>(ctx, st, c + int64_t(t) * ctx.M, re, im);
  __syncthreads();
  tile_to_grid<DIM, BIN_THREADS, false>(ctx, st.P, st.off, st.ntile, re, im,
                                        fw + int64_t(t) * nftot);
}

// ------------------------------------------------- output-driven spreading
//
// gpu_method 3, the reference's output-driven spreader
// (src/cuda/spread_output_driven_inst.cu:17-103): a subproblem's points are
// staged np at a time (each point's kernel values, first tile cell and
// strength in LDS); then the workgroup takes the staged points one after
// another, its threads splitting the point's ns^d support cells, so that no
// two threads update one cell at once and the fp64 tile takes plain
// read-modify-writes; a barrier separates consecutive points. The tile is
// added to the grid with atomics, as for subproblems.
//
// LDS: the fp64 tile (re, im planes of ntile cells), then np strengths (re
// and im arrays of T), np * DIM * NS kernel values (T) and np * DIM first
// cells (int); od_stage_bytes gives the staging part.
template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(BIN_THREADS)
    spread_od_kernel(SpreadCtx<T> ctx, const cplx_t<T> *__restrict__ c,
                     cplx_t<T> *__restrict__ fw, int64_t nftot, int np) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double *re = reinterpret_cast<double *>(lds_raw);
  StartTile<DIM, NS> st;
  if (!st.init_sub(ctx, blockIdx.x) || st.npts == 0) return;
  double *im = re + st.ntile;
  T *s_re = reinterpret_cast<T *>(im + st.ntile);
  T *s_im = s_re + np;
  T *s_ker = s_im + np;
  int *s_l = reinterpret_cast<int *>(s_ker + np * DIM * NS);
  const cplx_t<T> *cv = c + int64_t(blockIdx.y) * ctx.M;
  for (int n = threadIdx.x; n < st.ntile; n += BIN_THREADS) {
    re[n] = 0.0;
    im[n] = 0.0;
  }
  constexpr int NSD = DIM == 1 ? NS : DIM == 2 ? NS * NS : NS * NS * NS;
  for (int b0 = 0; b0 < st.npts; b0 += np) {
    const int nb = min(np, st.npts - b0);
    // the previous batch's last point was followed by a barrier (and the
    // tile clearing by the one below), so the staging arrays are free
    for (int i = threadIdx.x; i < nb; i += BIN_THREADS) {
      const int q = st.first + b0 + i;
      const int j = ctx.idx[q];
      T ker[DIM][NS];
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        s_l[i * DIM + d] = st.local_start(
            start_and_kernel<T, NS>(coord(ctx, d, q, j), ctx.nf[d], ker[d], ctx.kp), d,
            ctx.nf[d]);
#pragma unroll
        for (int a = 0; a < NS; ++a) s_ker[(i * DIM + d) * NS + a] = ker[d][a];
      }
      const cplx_t<T> v = cv[j];
      s_re[i] = v.x;
      s_im[i] = v.y;
    }
    __syncthreads();
    for (int i = 0; i < nb; ++i) {
      const T cr = s_re[i], ci = s_im[i];
      const T *k = s_ker + i * DIM * NS;
      const int *l = s_l + i * DIM;
      for (int s = threadIdx.x; s < NSD; s += BIN_THREADS) {
        const int a = s % NS;
        T w = k[a];
        int cell = l[0] + a;
        if constexpr (DIM > 1) {
          const int b = (s / NS) % NS;
          w *= k[NS + b];
          cell += (l[1] + b) * st.P[0];
        }
        if constexpr (DIM > 2) {
          const int e = s / (NS * NS);
          w *= k[2 * NS + e];
          cell += (l[2] + e) * st.P[0] * st.P[1];
        }
        re[cell] += double(cr * w);
        im[cell] += double(ci * w);
      }
      __syncthreads();
    }
  }
  __syncthreads();
  tile_to_grid<DIM, BIN_THREADS, false>(ctx, st.P, st.off, st.ntile, re, im,
                                        fw + int64_t(blockIdx.y) * nftot);
}


// ------------------------------------------------------- column spreading (3-D)
//
// The per-point LDS scatter above costs ns^3 LDS atomics per point. Here a
// wave owns the FOOT x FOOT (x, y) grid columns of one sub-column (the points
// whose first node lies in an S x S square of the bin, S = col_side(ns),
// FOOT = S + ns - 1 <= 8), one column per lane, and holds the column's whole
// depth in the tile (bin depth + ns - 1 planes, at most COL_ZC + ns - 1) in
// registers. A point with first z plane lz adds c wx wy wz[m] to planes
// lz + m, m < ns, of every lane's column: 2 ns FMAs per lane. The sort orders
// a sub-column's points by lz, so the walk visits the planes in a pass
// unrolled at compile time and meets each run of equal lz with its register
// indices fixed (no branch on lz, every accumulator of a run's loop updated
// in place). When the walk passes into another sub-column the touched planes
// of the columns are added to the LDS tile (one fp64 atomic add per lane,
// plane and part) and cleared. The tile is fp64 because this GPU's fp64 LDS
// atomic add runs about 20 times faster than its fp32 one.
//
// A workgroup's points (a bin, or a subproblem of a dense bin) are dealt to
// its waves as equal contiguous ranges of the sorted order, in which the
// points of a sub-column are contiguous; a sub-column split between two waves
// is added to the tile by both. A wave stages COL_BATCH points at a time: lane k
// evaluates the kernel of point p + k and writes a record of COL_REC floats:
// wz (8 floats, zero past ns: two broadcast 128-bit reads), then c wx as
// (re, im) pairs shifted to the lanes' columns (16 floats, one 64-bit read
// per lane) and wy shifted likewise (8 floats, one 32-bit read), all zero
// outside the support. The point's first z plane and sub-column stay in
// that lane's registers and are read with readlane.

// One point's staged values as one lane sees them.
struct ColRec {
  float4 z0, z1;  // wz[0..7]
  float2 xc;      // c wx[li]
  float y;        // wy[lj]
  __device__ __forceinline__ void load(const float *stage, int k, int xoff, int yoff) {
    const float *rk = stage + k * COL_REC;
    z0 = reinterpret_cast<const float4 *>(rk)[0];
    z1 = reinterpret_cast<const float4 *>(rk)[1];
    xc = *reinterpret_cast<const float2 *>(rk + xoff);
    y = rk[yoff];
  }
};

// Adds the run of points k, k + 1, ... of sub-column position `base`
// (lz + 64 * sub-column) whose first plane is Z, then the runs of planes
// Z + 1 .. COL_ZC - 1.
template <int NS, int ZD, int Z>
__device__ __forceinline__ void col_walk(int &k, int n, int kb, int base, int posv,
                                         const float *stage, int xoff, int yoff, float (&ar)[ZD],
                                         float (&ai)[ZD], int &zlo, int &zhi) {
  if constexpr (Z < COL_ZC) {
    if (k < n && __builtin_amdgcn_readlane(posv, kb + k) == base + Z) {
      zlo = min(zlo, Z);
      zhi = max(zhi, Z + NS);
      do {
        ColRec rec;
        rec.load(stage, k, xoff, yoff);
        const float wr = rec.xc.x * rec.y, wi = rec.xc.y * rec.y;
        const float wz[8] = {rec.z0.x, rec.z0.y, rec.z0.z, rec.z0.w,
                             rec.z1.x, rec.z1.y, rec.z1.z, rec.z1.w};
#pragma unroll
        for (int m = 0; m < NS; ++m) {
          ar[Z + m] = std::fma(wr, wz[m], ar[Z + m]);
          ai[Z + m] = std::fma(wi, wz[m], ai[Z + m]);
        }
        ++k;
      } while (k < n && __builtin_amdgcn_readlane(posv, kb + k) == base + Z);
    }
    col_walk<NS, ZD, Z + 1>(k, n, kb, base, posv, stage, xoff, yoff, ar, ai, zlo, zhi);
  }
}

// Adds planes [z0, z1) of the register columns to the tile at column cell
// `colcell` (lanes with `on`), then clears all planes (the untouched ones
// are zero already; clearing them too keeps every register write
// unconditional, which the compiler keeps in place). z0, z1 are
// wave-uniform; the plane stride PLANE is a compile-time constant, so each
// add uses the LDS instruction's address offset.
template <int ZD, int PLANE>
__device__ __forceinline__ void col_flush(float (&ar)[ZD], float (&ai)[ZD], double *re,
                                          double *im, int colcell, int z0, int z1, bool on) {
  double *cr = re + colcell, *ci = im + colcell;
#pragma unroll
  for (int p = 0; p < ZD; ++p) {
    if (p >= z0 && p < z1 && on) {
      atomicAdd(cr + p * PLANE, double(ar[p]));
      atomicAdd(ci + p * PLANE, double(ai[p]));
    }
  }
#pragma unroll
  for (int p = 0; p < ZD; ++p) ar[p] = ai[p] = 0.f;
}

// Accumulates points [u0, u1) of the unit into the fp64 LDS tile (re, im
// planes of P x P x P2 cells, x fastest, P = 8 + ns - 1).
template <int NS, typename Start>
__device__ __forceinline__ void col_spread_range(const SpreadCtx<float> &ctx,
                                                 const StartTile<3, NS> &st,
                                                 const float2 *__restrict__ c, double *re,
                                                 double *im, float *stage, int u0, int u1,
                                                 Start &&start) {
  constexpr int S = col_side(NS);
  constexpr int FOOT = S + NS - 1;
  constexpr int ZD = COL_ZC + NS - 1;  // register planes per column
  constexpr int P = 8 + NS - 1;        // tile side in x and y (bins of 8)
  constexpr int NSX = 8 / S;           // sub-columns per bin side
  const int lane = threadIdx.x & 63;
  const int li = lane & 7, lj = lane >> 3;
  const bool on = li < FOOT && lj < FOOT;
  const int xoff = 8 + 2 * li, yoff = 24 + lj;

  // Sorted index and coordinates of one lane's point of the next batch,
  // loaded while the current batch is processed.
  int fj = 0;
  float fx = 0.f, fy = 0.f, fz = 0.f;
  auto fetch = [&](int q) {
    if (q < u1) {
      fj = ctx.idx[q];
      fx = coord(ctx, 0, q, fj);
      fy = coord(ctx, 1, q, fj);
      fz = coord(ctx, 2, q, fj);
    }
  };
  // Every lane evaluates one point of a pass of 64 and keeps its record in
  // registers; the records go to the wave's stage (COL_BATCH = 32 records)
  // one half-pass at a time, each half walked before the next is written.
  // The first pass's loads are issued before the workgroup's start-up work
  // (clearing the tile), and each next pass's loads while the current pass
  // is walked, so that memory latency overlaps work.
  static_assert(COL_BATCH == 32, "two half-passes of 32 records per pass of 64 points");
  fetch(u0 + lane);
  float2 fc = make_float2(0.f, 0.f);
  if (u0 + lane < u1) fc = c[fj];
  start();

  float ar[ZD], ai[ZD];
#pragma unroll
  for (int k = 0; k < ZD; ++k) ar[k] = ai[k] = 0.f;
  int cur = -1;           // sub-column held in the registers
  int zlo = ZD, zhi = 0;  // planes [zlo, zhi) touched since the last flush
  int colcell = 0;

  for (int p = u0; p < u1; p += 64) {
    const int n = min(64, u1 - p);
    const float x = fx, y = fy, z = fz;
    const float2 cv = fc;
    fetch(p + 64 + lane);
    int posv = 0;  // lz + 64 * sub-column of this lane's point
    float4 rec[8];  // this lane's point's record (COL_REC floats)
    if (lane < n) {
      float kx[NS], ky[NS], kz[NS];
      const int s0 = start_and_kernel<float, NS>(x, ctx.nf[0], kx, ctx.kp);
      const int s1 = start_and_kernel<float, NS>(y, ctx.nf[1], ky, ctx.kp);
      const int s2 = start_and_kernel<float, NS>(z, ctx.nf[2], kz, ctx.kp);
      const int lxl = st.local_start(s0, 0, ctx.nf[0]);
      const int lyl = st.local_start(s1, 1, ctx.nf[1]);
      // < bin depth <= COL_ZC (the plan's bins); the clamp keeps every
      // point inside the planes the walk visits, so that it always advances
      const int lz = min(st.local_start(s2, 2, ctx.nf[2]), COL_ZC - 1);
      const int sx = lxl / S, sy = lyl / S;
      posv = lz + 64 * (sx + NSX * sy);
      const int lx = lxl - sx * S, ly = lyl - sy * S;
      // x and y weights shifted to the lanes' columns: column t holds
      // k[t - lx] (lx < S), zero outside the support.
      float wx8[8], wy8[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        float vx = 0.f, vy = 0.f;
#pragma unroll
        for (int o = 0; o < S; ++o) {
          if (t - o >= 0 && t - o < NS) {
            vx = lx == o ? kx[t - o] : vx;
            vy = ly == o ? ky[t - o] : vy;
          }
        }
        wx8[t] = vx;
        wy8[t] = vy;
      }
      float wz8[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) wz8[m] = m < NS ? kz[m] : 0.f;
      rec[0] = make_float4(wz8[0], wz8[1], wz8[2], wz8[3]);
      rec[1] = make_float4(wz8[4], wz8[5], wz8[6], wz8[7]);
#pragma unroll
      for (int t = 0; t < 8; t += 2)
        rec[2 + t / 2] = make_float4(cv.x * wx8[t], cv.y * wx8[t], cv.x * wx8[t + 1],
                                     cv.y * wx8[t + 1]);
      rec[6] = make_float4(wy8[0], wy8[1], wy8[2], wy8[3]);
      rec[7] = make_float4(wy8[4], wy8[5], wy8[6], wy8[7]);
    }
    if (p + 64 + lane < u1) fc = c[fj];  // the next pass's strengths
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int kb = h * COL_BATCH;
      const int nh = min(COL_BATCH, n - kb);
      if (nh <= 0) break;
      if (lane >= kb && lane < kb + nh) {
        float4 *r = reinterpret_cast<float4 *>(stage + (lane - kb) * COL_REC);
#pragma unroll
        for (int q = 0; q < 8; ++q) r[q] = rec[q];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");

      int k = 0;
      while (k < nh) {
        const int sc = __builtin_amdgcn_readlane(posv, kb + k) >> 6;
        if (sc != cur) {  // another sub-column: hand the columns to the tile
          if (cur >= 0) col_flush<ZD, P * P>(ar, ai, re, im, colcell, zlo, zhi, on);
          cur = sc;
          zlo = ZD;
          zhi = 0;
          colcell = ((sc / NSX) * S + lj) * P + (sc % NSX) * S + li;
        }
        col_walk<NS, ZD, 0>(k, nh, kb, 64 * sc, posv, stage, xoff, yoff, ar, ai, zlo, zhi);
      }
      // all lanes are done reading these records before restaging
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
  }
  if (cur >= 0) col_flush<ZD, P * P>(ar, ai, re, im, colcell, zlo, zhi, on);
}

// The unit's tile (zeroed) and its points dealt to the waves.
template <int NS>
__device__ __forceinline__ void col_spread_unit(const SpreadCtx<float> &ctx,
                                                const StartTile<3, NS> &st,
                                                const float2 *__restrict__ c, double *re,
                                                double *im, float *stage_all) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int first = __builtin_amdgcn_readfirstlane(st.first);
  const int npts = __builtin_amdgcn_readfirstlane(st.npts);
  const int u0 = first + int(int64_t(npts) * wave / COL_WAVES);
  const int u1 = first + int(int64_t(npts) * (wave + 1) / COL_WAVES);
  // every wave runs the range (an empty one too): the tile is cleared and
  // the workgroup synchronized inside, after the first loads are issued
  col_spread_range<NS>(ctx, st, c, re, im, stage_all + wave * (COL_BATCH * COL_REC), u0, u1, [&] {
    for (int n = threadIdx.x; n < st.ntile; n += 64 * COL_WAVES) {
      re[n] = 0.0;
      im[n] = 0.0;
    }
    __syncthreads();
  });
  __syncthreads();
}

// Whole-bin column spreading in parity colors with a plain tile flush; bins
// with more than dense_min points are left to spread_col_sub_kernel.
template <int NS>
__global__ void __launch_bounds__(64 * COL_WAVES, 4)
    spread_col_colored_kernel(SpreadCtx<float> ctx, const float2 *__restrict__ c,
                              float2 *__restrict__ fw, int64_t nftot, int col0, int col1,
                              int col2) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  StartTile<3, NS> st;
  st.init_bin(ctx, colored_bin<3>(ctx, blockIdx.x, col0, col1, col2));
  if (st.npts == 0 || st.npts > ctx.dense_min) return;
  double *re = reinterpret_cast<double *>(lds_raw);
  double *im = re + st.ntile;
  const int t = blockIdx.y;
  col_spread_unit<NS>(ctx, st, c + int64_t(t) * ctx.M, re, im,
                      reinterpret_cast<float *>(im + st.ntile));
  tile_to_grid<3, 64 * COL_WAVES, true>(ctx, st.P, st.off, st.ntile, re, im,
                                        fw + int64_t(t) * nftot);
}

// Column spreading of the dense bins' subproblems with an atomic tile flush.
template <int NS>
__global__ void __launch_bounds__(64 * COL_WAVES, 4)
    spread_col_sub_kernel(SpreadCtx<float> ctx, const float2 *__restrict__ c,
                          float2 *__restrict__ fw, int64_t nftot) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  StartTile<3, NS> st;
  if (!st.init_sub(ctx, blockIdx.x)) return;
  double *re = reinterpret_cast<double *>(lds_raw);
  double *im = re + st.ntile;
  const int t = blockIdx.y;
  col_spread_unit<NS>(ctx, st, c + int64_t(t) * ctx.M, re, im,
                      reinterpret_cast<float *>(im + st.ntile));
  tile_to_grid<3, 64 * COL_WAVES, false>(ctx, st.P, st.off, st.ntile, re, im,
                                         fw + int64_t(t) * nftot);
}

// Workgroup size of the multi-vector interpolation, points per thread per
// chunk, and tile cells per thread held in registers while the previous
// vector is interpolated (tiles larger than MV_PF * MV_THREADS cells are
// staged without the overlap). Chosen by a sweep on the batched 2-D workload
// (256 / 512 threads, 1 / 2 points, 1 / 6 / 12 cells per thread).
constexpr int MV_THREADS = 512;
constexpr int MV_PT = 2;
constexpr int MV_PF = 6;

// Interpolation of one subproblem over nvec vectors (the fused launch over
// many problems). The points are taken in chunks of up to MV_PT per thread;
// each point's kernel weights and tile offsets stay in registers while the
// tile of every vector in turn is staged and applied, so a point's kernel is
// evaluated once per round rather than once per vector. The next vector's
// tile is loaded into registers while the current one is interpolated.
// Global arrays are accessed through global-typed pointers and the Horner
// table through a constant-typed one (scalar loads).
template <typename T, int DIM, int NS>
__device__ __forceinline__ void interp_unit_mv(const SpreadCtx<T> &ctx, const KerParams<T> &kp,
                                               const StartTile<DIM, NS> &st, cplx_t<T> *tile,
                                               gptr<const vec2_t<T>> g0, int64_t nftot,
                                               gptr<vec2_t<T>> c0, int nvec) {
  constexpr int CH = MV_PT * MV_THREADS;
  const gptr<const int> idx = to_global(ctx.idx);
  const kptr<T> coef = to_const(kp.coef);
  gptr<const T> sx[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) sx[d] = to_global(ctx.sxyz[d]);
  TileWalk<DIM, MV_THREADS> walk;
  walk.init(st.P);
  const bool pipelined = st.ntile <= MV_PF * MV_THREADS;
  for (int c0i = 0; c0i < st.npts; c0i += CH) {
    const int n = min(CH, st.npts - c0i);
    // Strided dealing (see interp_unit): item i = 32 m + r takes point r G + m.
    const int G = (n + 31) >> 5;
    T ker[MV_PT][DIM][NS];
    int l[MV_PT][DIM];
    int j[MV_PT];
    bool on[MV_PT];
#pragma unroll
    for (int k = 0; k < MV_PT; ++k) {
      const int i = threadIdx.x + k * MV_THREADS;
      const int pos = (i & 31) * G + (i >> 5);
      on[k] = i < 32 * G && pos < n;
      j[k] = 0;
      if (on[k]) {
        const int q = st.first + c0i + pos;
        j[k] = idx[q];
#pragma unroll
        for (int d = 0; d < DIM; ++d)
          l[k][d] = st.local_start(
              start_and_kernel_c<T, NS>(sx[d][q], ctx.nf[d], ker[k][d], kp, coef), d, ctx.nf[d]);
      }
    }
    vec2_t<T> pre[MV_PF];
    if (pipelined) {  // vector 0's tile into registers
      int x = walk.x0, y = walk.y0, z = walk.z0;
#pragma unroll
      for (int k = 0; k < MV_PF; ++k) {
        const int m = threadIdx.x + k * MV_THREADS;
        if (m < st.ntile) pre[k] = g0[walk.index_and_step(ctx, st.P, st.off, x, y, z)];
      }
    }
    for (int v = 0; v < nvec; ++v) {
      __syncthreads();  // the previous tile is no longer read
      if (pipelined) {
#pragma unroll
        for (int k = 0; k < MV_PF; ++k) {
          const int m = threadIdx.x + k * MV_THREADS;
          if (m < st.ntile) tile[m] = cplx_t<T>{pre[k].x, pre[k].y};
        }
      } else {
        const gptr<const vec2_t<T>> g = g0 + int64_t(v) * nftot;
        int x = walk.x0, y = walk.y0, z = walk.z0;
        for (int m = threadIdx.x; m < st.ntile; m += MV_THREADS) {
          const vec2_t<T> w = g[walk.index_and_step(ctx, st.P, st.off, x, y, z)];
          tile[m] = cplx_t<T>{w.x, w.y};
        }
      }
      __syncthreads();
      if (pipelined && v + 1 < nvec) {  // next vector's tile, in flight during the sums
        const gptr<const vec2_t<T>> g = g0 + int64_t(v + 1) * nftot;
        int x = walk.x0, y = walk.y0, z = walk.z0;
#pragma unroll
        for (int k = 0; k < MV_PF; ++k) {
          const int m = threadIdx.x + k * MV_THREADS;
          if (m < st.ntile) pre[k] = g[walk.index_and_step(ctx, st.P, st.off, x, y, z)];
        }
      }
      const gptr<vec2_t<T>> cout = c0 + int64_t(v) * ctx.M;
#pragma unroll
      for (int k = 0; k < MV_PT; ++k) {
        if (!on[k]) continue;
        T ar = 0, ai = 0;
        if constexpr (DIM == 1) {
#pragma unroll
          for (int a = 0; a < NS; ++a) {
            const cplx_t<T> w = tile[l[k][0] + a];
            ar = std::fma(w.x, ker[k][0][a], ar);
            ai = std::fma(w.y, ker[k][0][a], ai);
          }
        } else if constexpr (DIM == 2) {
          for (int b = 0; b < NS; ++b) {
            const cplx_t<T> *row = tile + (l[k][1] + b) * st.P[0] + l[k][0];
            T rr = 0, ri = 0;
#pragma unroll
            for (int a = 0; a < NS; ++a) {
              const cplx_t<T> w = row[a];
              rr = std::fma(w.x, ker[k][0][a], rr);
              ri = std::fma(w.y, ker[k][0][a], ri);
            }
            ar = std::fma(rr, ker[k][1][b], ar);
            ai = std::fma(ri, ker[k][1][b], ai);
          }
        } else {
          for (int e = 0; e < NS; ++e) {
            T pr = 0, pi = 0;
            for (int b = 0; b < NS; ++b) {
              const cplx_t<T> *row =
                  tile + ((l[k][2] + e) * st.P[1] + l[k][1] + b) * st.P[0] + l[k][0];
              T rr = 0, ri = 0;
#pragma unroll
              for (int a = 0; a < NS; ++a) {
                const cplx_t<T> w = row[a];
                rr = std::fma(w.x, ker[k][0][a], rr);
                ri = std::fma(w.y, ker[k][0][a], ri);
              }
              pr = std::fma(rr, ker[k][1][b], pr);
              pi = std::fma(ri, ker[k][1][b], pi);
            }
            ar = std::fma(pr, ker[k][2][e], ar);
            ai = std::fma(pi, ker[k][2][e], ai);
          }
        }
        vec2_t<T> o;
        o.x = ar;
        o.y = ai;
        cout[j[k]] = o;
      }
    }
  }
}

// Bin interpolation: stage the padded tile, then gather per point. The unit is
// a whole bin, or (interp_sub) a subproblem of at most maxsub points of one
// bin, so that a dense bin is shared by several workgroups.
// Interpolates one unit (bin or subproblem) of vector grid g into cout.
template <typename T, int DIM, int NS>
__device__ __forceinline__ void interp_unit(const SpreadCtx<T> &ctx, const StartTile<DIM, NS> &st,
                                            cplx_t<T> *tile, const cplx_t<T> *__restrict__ g,
                                            cplx_t<T> *__restrict__ cout) {
  // Bank-ordered units: the unit's points are sorted by the bank class r of
  // their first tile cell (cell mod 32), and cls[r] .. cls[r + 1] are the
  // positions of class r in the unit. Rounds j < R (R the smallest class
  // count) deal point cls[r] + j to work item 32 j + r: the 32 lanes of a
  // half-wave start in 32 different classes, so every gather of the half-wave
  // hits distinct banks. The points beyond the full rounds (left[r] of class
  // r, lp[r] their prefix) are dealt strided (below).
  int *cls = nullptr, *lp = nullptr;
  int R = 0, nleft = st.npts;
  bool contiguous = false;
  if (ctx.bank) {
    cls = reinterpret_cast<int *>(tile + st.ntile);
    lp = cls + 33;
    if (threadIdx.x <= 32) {
      const int v = ctx.colstart[int64_t(st.bin) * 32 + threadIdx.x];
      cls[threadIdx.x] = min(max(v, st.first), st.first + st.npts);
    }
  }
  grid_to_tile<DIM, BIN_THREADS>(ctx, st.P, st.off, st.ntile, g, tile);
  __syncthreads();
  if (ctx.bank) {
    if (threadIdx.x == 0) {
      int r0 = st.npts, occupied = 0;
      for (int r = 0; r < 32; ++r) {
        r0 = min(r0, cls[r + 1] - cls[r]);
        occupied += cls[r + 1] > cls[r] ? 1 : 0;
      }
      int acc = 0;
      for (int r = 0; r < 32; ++r) {
        lp[r] = acc;
        acc += cls[r + 1] - cls[r] - r0;
      }
      lp[32] = acc;
      lp[33] = r0;
      lp[34] = r0 == 0 && occupied < 16 ? 1 : 0;
    }
    __syncthreads();
    R = lp[33];
    nleft = lp[32];
    contiguous = lp[34] != 0;
  }
  // Strided dealing of the remaining points: work item i = 32 g + m takes
  // remaining point m G + g (G = ceil(nleft / 32)). In a bank-ordered unit
  // these are in class order too, so the 32 lanes of a half-wave mostly
  // start in different classes. A unit whose points fall in fewer than 16
  // classes (a tight cluster, such as the centre of radial spokes) is dealt
  // contiguously instead: strided lanes would take points of one class in
  // different cells (the same bank at different addresses), while
  // consecutive points of the sorted order (stable within a class) mostly
  // share their cells, whose reads the LDS broadcasts.
  const int nfull = 32 * R;
  const int G = (nleft + 31) >> 5;
  for (int i = threadIdx.x; i < nfull + 32 * G; i += BIN_THREADS) {
    int pos;
    if (i < nfull) {
      pos = cls[i & 31] + (i >> 5) - st.first;
    } else {
      const int t = i - nfull;
      const int u = contiguous ? t : (t & 31) * G + (t >> 5);
      if (u >= nleft) continue;
      if (R > 0) {
        // the class holding remaining point u: the last r with lp[r] <= u
        int lo = 0, hi = 31;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (lp[mid] <= u)
            lo = mid;
          else
            hi = mid - 1;
        }
        pos = cls[lo] + R + (u - lp[lo]) - st.first;
      } else {
        pos = u;
      }
    }
    const int q = st.first + pos;
    const int j = ctx.idx[q];
    T ker[DIM][NS];
    int l[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      l[d] = st.local_start(start_and_kernel<T, NS>(coord(ctx, d, q, j), ctx.nf[d], ker[d], ctx.kp),
                            d, ctx.nf[d]);
    T ar = 0, ai = 0;
    if constexpr (DIM == 1) {
#pragma unroll
      for (int a = 0; a < NS; ++a) {
        const cplx_t<T> v = tile[l[0] + a];
        ar = std::fma(v.x, ker[0][a], ar);
        ai = std::fma(v.y, ker[0][a], ai);
      }
    } else if constexpr (DIM == 2) {
      for (int b = 0; b < NS; ++b) {
        const cplx_t<T> *row = tile + (l[1] + b) * st.P[0] + l[0];
        T rr = 0, ri = 0;
#pragma unroll
        for (int a = 0; a < NS; ++a) {
          const cplx_t<T> v = row[a];
          rr = std::fma(v.x, ker[0][a], rr);
          ri = std::fma(v.y, ker[0][a], ri);
        }
        ar = std::fma(rr, ker[1][b], ar);
        ai = std::fma(ri, ker[1][b], ai);
      }
    } else {
      for (int e = 0; e < NS; ++e) {
        T pr = 0, pi = 0;
        for (int b = 0; b < NS; ++b) {
          const cplx_t<T> *row = tile + ((l[2] + e) * st.P[1] + l[1] + b) * st.P[0] + l[0];
          T rr = 0, ri = 0;
#pragma unroll
          for (int a = 0; a < NS; ++a) {
            const cplx_t<T> v = row[a];
            rr = std::fma(v.x, ker[0][a], rr);
            ri = std::fma(v.y, ker[0][a], ri);
          }
          pr = std::fma(rr, ker[1][b], pr);
          pi = std::fma(ri, ker[1][b], pi);
        }
        ar = std::fma(pr, ker[2][e], ar);
        ai = std::fma(pi, ker[2][e], ai);
      }
    }
    cout[j] = cplx_t<T>{ar, ai};
  }
}

// Bin interpolation: stage the padded tile, then gather per point. The unit is
// a whole bin, or (interp_sub) a subproblem of at most maxsub points of one
// bin, so that a dense bin is shared by several workgroups. The tile holds
// interleaved complex values: one 8-byte (fp32) or 16-byte (fp64) LDS read
// per cell.
template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(BIN_THREADS)
    interp_bin_kernel(SpreadCtx<T> ctx, cplx_t<T> *__restrict__ c,
                      const cplx_t<T> *__restrict__ fw, int64_t nftot) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  cplx_t<T> *tile = reinterpret_cast<cplx_t<T> *>(lds_raw);
  StartTile<DIM, NS> st;
  if (ctx.interp_sub) {
    if (!st.init_sub(ctx, blockIdx.x)) return;
  } else {
    st.init_bin(ctx, blockIdx.x);
  }
  if (st.npts == 0) return;
  const int t = blockIdx.y;
  interp_unit<T, DIM, NS>(ctx, st, tile, fw + int64_t(t) * nftot, c + int64_t(t) * ctx.M);
}

// Interpolation of many problems (see interp_batch): workgroup b takes work
// unit b (a subproblem of one problem, over all vectors of the round).
template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(MV_THREADS)
    interp_batch_kernel(const InterpItem<T> *__restrict__ items, int nprob, KerParams<T> kp,
                        const cplx_t<T> *const *__restrict__ fw_base,
                        cplx_t<T> *const *__restrict__ c_base, int64_t c_vec_offset, int nvec,
                        int64_t total_units) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  cplx_t<T> *tile = reinterpret_cast<cplx_t<T> *>(lds_raw);
  const int64_t u = blockIdx.x;
  if (u >= total_units) return;
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].unit_begin <= u)
      lo = mid;
    else
      hi = mid - 1;
  }
  // lo is wave-uniform (readfirstlane: scalar loads of the descriptor);
  // interp_unit_mv accesses the problem's arrays through global-typed
  // pointers, as a kernel with the context as argument does (generic pointers
  // cost flat loads and ~100 vector registers, halving the occupancy).
  lo = __builtin_amdgcn_readfirstlane(lo);
  const InterpItem<T> it = items[lo];
  StartTile<DIM, NS> st;
  if (!st.init_sub(it.ctx, int(u - it.unit_begin)) || st.npts == 0) return;
  interp_unit_mv<T, DIM, NS>(
      it.ctx, kp, st, tile, to_global(reinterpret_cast<const vec2_t<T> *>(fw_base[lo])), it.nftot,
      to_global(reinterpret_cast<vec2_t<T> *>(c_base[lo])) + c_vec_offset * int64_t(it.ctx.M),
      nvec);
}

// Spreading of many problems (see spread_batch): workgroup b takes work unit
// b (a subproblem of one problem) and adds it to the problem's grid for
// every vector of the round in turn (fp64 tile, atomic flush, as
// spread_sub_kernel).
template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(BIN_THREADS)
    spread_batch_kernel(const InterpItem<T> *__restrict__ items, int nprob,
                        const cplx_t<T> *const *__restrict__ c_base, int64_t c_vec_offset,
                        cplx_t<T> *const *__restrict__ fw_base, int nvec, int64_t total_units) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double *re = reinterpret_cast<double *>(lds_raw);
  const int64_t u = blockIdx.x;
  if (u >= total_units) return;
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].unit_begin <= u)
      lo = mid;
    else
      hi = mid - 1;
  }
  lo = __builtin_amdgcn_readfirstlane(lo);
  const InterpItem<T> it = items[lo];
  StartTile<DIM, NS> st;
  if (!st.init_sub(it.ctx, int(u - it.unit_begin)) || st.npts == 0) return;
  double *im = re + st.ntile;
  const cplx_t<T> *c = c_base[lo] + c_vec_offset * int64_t(it.ctx.M);
  cplx_t<T> *fw = fw_base[lo];
  for (int t = 0; t < nvec; ++t) {
    for (int n = threadIdx.x; n < st.ntile; n += BIN_THREADS) {
      re[n] = 0.0;
      im[n] = 0.0;
    }
    __syncthreads();
    spread_into_tile<T, DIM, NS>(it.ctx, st, c + int64_t(t) * it.ctx.M, re, im);
    __syncthreads();
    tile_to_grid<DIM, BIN_THREADS, false>(it.ctx, st.P, st.off, st.ntile, re, im,
                                          fw + int64_t(t) * it.nftot);
    __syncthreads();  // the tile is read out before the next vector clears it
  }
}

// Interpolant and coordinate derivatives, one lane per point (sorted order),
// reading the fine grid from HBM: c = sum g kx ky kz and
// dc/dx_d = sum g (d k_d / dx_d) prod_{e != d} k_e, where d k / dx =
// (d k / d xl) (-nf / 2 pi). Used by the backward pass; the extra FMAs share
// the grid reads of the plain interpolation.
template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(SPREAD_THREADS)
    interp_grad_kernel(SpreadCtx<T> ctx, cplx_t<T> *__restrict__ c, cplx_t<T> *__restrict__ dc,
                       const cplx_t<T> *__restrict__ fw, int64_t nftot, int blk) {
  const int i = blockIdx.x * SPREAD_THREADS + threadIdx.x;
  if (i >= ctx.M) return;
  const int j = ctx.idx[i];
  T ker[3][NS], dker[3][NS];
  int st[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    st[d] = start_and_kernel_d<T, NS>(coord(ctx, d, i, j), ctx.nf[d], ker[d], dker[d], ctx.kp);
    if (st[d] < 0) st[d] += ctx.nf[d];
    const T scale = T(-double(ctx.nf[d]) / (2.0 * 3.14159265358979323846));
#pragma unroll
    for (int a = 0; a < NS; ++a) dker[d][a] *= scale;
  }
  int xi[NS];
#pragma unroll
  for (int a = 0; a < NS; ++a) xi[a] = wrap_once(st[0] + a, ctx.nf[0]);
  const int64_t mb = int64_t(ctx.M) * blk;
  for (int t = 0; t < blk; ++t) {
    const cplx_t<T> *g = fw + int64_t(t) * nftot;
    // v: value, d0..d2: derivatives (re, im)
    T v[2] = {0, 0}, d0[2] = {0, 0}, d1[2] = {0, 0}, d2[2] = {0, 0};
    const int ne = DIM > 2 ? NS : 1, nb = DIM > 1 ? NS : 1;
    for (int e = 0; e < ne; ++e) {
      const int iz = DIM > 2 ? wrap_once(st[2] + e, ctx.nf[2]) : 0;
      T pv[2] = {0, 0}, p0[2] = {0, 0}, p1[2] = {0, 0};
      for (int b = 0; b < nb; ++b) {
        const int iy = DIM > 1 ? wrap_once(st[1] + b, ctx.nf[1]) : 0;
        const cplx_t<T> *row = g + (int64_t(iz) * ctx.nf[1] + iy) * ctx.nf[0];
        T rv[2] = {0, 0}, r0[2] = {0, 0};
#pragma unroll
        for (int a = 0; a < NS; ++a) {
          const cplx_t<T> w = row[xi[a]];
          rv[0] = std::fma(w.x, ker[0][a], rv[0]);
          rv[1] = std::fma(w.y, ker[0][a], rv[1]);
          r0[0] = std::fma(w.x, dker[0][a], r0[0]);
          r0[1] = std::fma(w.y, dker[0][a], r0[1]);
        }
        if constexpr (DIM > 1) {
          const T ky = ker[1][b], dky = dker[1][b];
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            pv[u] = std::fma(rv[u], ky, pv[u]);
            p0[u] = std::fma(r0[u], ky, p0[u]);
            p1[u] = std::fma(rv[u], dky, p1[u]);
          }
        } else {
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            pv[u] = rv[u];
            p0[u] = r0[u];
          }
        }
      }
      if constexpr (DIM > 2) {
        const T kz = ker[2][e], dkz = dker[2][e];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          v[u] = std::fma(pv[u], kz, v[u]);
          d0[u] = std::fma(p0[u], kz, d0[u]);
          d1[u] = std::fma(p1[u], kz, d1[u]);
          d2[u] = std::fma(pv[u], dkz, d2[u]);
        }
      } else {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          v[u] = pv[u];
          d0[u] = p0[u];
          d1[u] = p1[u];
        }
      }
    }
    const int64_t o = int64_t(t) * ctx.M + j;
    if (c) c[o] = cplx_t<T>{v[0], v[1]};
    dc[o] = cplx_t<T>{d0[0], d0[1]};
    if constexpr (DIM > 1) dc[mb + o] = cplx_t<T>{d1[0], d1[1]};
    if constexpr (DIM > 2) dc[2 * mb + o] = cplx_t<T>{d2[0], d2[1]};
  }
}

// ---------------------------------------------------------------- launchers

template <typename T>
int64_t grid_total(const SpreadCtx<T> &ctx) {
  return int64_t(ctx.nf[0]) * ctx.nf[1] * ctx.nf[2];
}

// Raises a kernel's dynamic LDS limit once per (kernel, size); later launches
// (including those recorded into a graph) make no runtime call for it.
template <typename K>
void allow_lds(K kernel, size_t bytes) {
  if (bytes <= 48 * 1024) return;
  static std::mutex mu;
  static std::map<const void *, size_t> done;
  const void *key = reinterpret_cast<const void *>(kernel);
  std::lock_guard<std::mutex> lock(mu);
  auto it = done.find(key);
  if (it != done.end() && it->second >= bytes) return;
  hip_check(hipFuncSetAttribute(key, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)),
            "hipFuncSetAttribute(LDS)");
  done[key] = bytes;
}

template <typename T, int DIM, int NS>
void spread_ns(const SpreadCtx<T> &ctx, const cplx_t<T> *c, cplx_t<T> *fw, int blk,
               hipStream_t stream) {
  if (ctx.M == 0) return;
  const int64_t nftot = grid_total(ctx);
  if (ctx.method == 1 || ctx.spread_gm) {
    dim3 grid((ctx.M + SPREAD_THREADS - 1) / SPREAD_THREADS, blk);
    spread_gm_kernel<T, DIM, NS><<<grid, SPREAD_THREADS, 0, stream>>>(ctx, c, fw, nftot);
    launch_check("spread_gm_kernel");
  } else if (ctx.method == 3) {
    if (ctx.od_np < 1) fail(FINUFFT_ERR_INVALID_ARGUMENT, "output-driven spreading: np < 1");
    auto k = spread_od_kernel<T, DIM, NS>;
    allow_lds(k, ctx.lds_od);
    k<<<dim3(ctx.sub_capacity, blk), BIN_THREADS, ctx.lds_od, stream>>>(ctx, c, fw, nftot,
                                                                        ctx.od_np);
    launch_check("spread_od_kernel");
  } else if (ctx.colS) {
    if constexpr (DIM == 3 && std::is_same<T, float>::value && NS <= COL_MAX_NS) {
      if (ctx.colored) {
        auto k = spread_col_colored_kernel<NS>;
        allow_lds(k, ctx.lds_spread);
        for (int color = 0; color < 8; ++color) {
          int col[3] = {color & 1, (color >> 1) & 1, (color >> 2) & 1};
          int64_t nblk = 1;
          for (int d = 0; d < 3; ++d) nblk *= (ctx.nbins[d] - col[d] + 1) / 2;
          if (nblk == 0) continue;
          k<<<dim3(unsigned(nblk), blk), 64 * COL_WAVES, ctx.lds_spread, stream>>>(
              ctx, c, fw, nftot, col[0], col[1], col[2]);
          launch_check("spread_col_colored_kernel");
        }
      }
      if (ctx.n_dense_sub > 0) {
        auto k = spread_col_sub_kernel<NS>;
        allow_lds(k, ctx.lds_spread);
        k<<<dim3(unsigned(ctx.n_dense_sub), blk), 64 * COL_WAVES, ctx.lds_spread, stream>>>(
            ctx, c, fw, nftot);
        launch_check("spread_col_sub_kernel");
      }
    } else {
      fail(FINUFFT_ERR_INVALID_ARGUMENT, "column spreading needs 3-D fp32 with ns <= 8");
    }
  } else if (ctx.colored) {
    auto k = spread_bin_colored_kernel<T, DIM, NS>;
    allow_lds(k, ctx.lds_spread);
    for (int color = 0; color < (1 << DIM); ++color) {
      int col[3] = {0, 0, 0};
      int64_t nblk = 1;
      for (int d = 0; d < DIM; ++d) {
        col[d] = (color >> d) & 1;
        nblk *= (ctx.nbins[d] - col[d] + 1) / 2;
      }
      if (nblk == 0) continue;
      k<<<dim3(unsigned(nblk), blk), BIN_THREADS, ctx.lds_spread, stream>>>(
          ctx, c, fw, nftot, col[0], col[1], col[2]);
      launch_check("spread_bin_colored_kernel");
    }
  } else {
    auto k = spread_sub_kernel<T, DIM, NS>;
    allow_lds(k, ctx.lds_spread);
    dim3 grid(ctx.sub_capacity, blk);
    k<<<grid, BIN_THREADS, ctx.lds_spread, stream>>>(ctx, c, fw, nftot);
    launch_check("spread_sub_kernel");
  }
}

template <typename T, int DIM, int NS>
void interp_ns(const SpreadCtx<T> &ctx, cplx_t<T> *c, const cplx_t<T> *fw, int blk,
               hipStream_t stream) {
  if (ctx.M == 0) return;
  const int64_t nftot = grid_total(ctx);
  if (ctx.method == 1) {
    dim3 grid((ctx.M + SPREAD_THREADS - 1) / SPREAD_THREADS);
    interp_gm_kernel<T, DIM, NS><<<grid, SPREAD_THREADS, 0, stream>>>(ctx, c, fw, nftot, blk);
    launch_check("interp_gm_kernel");
  } else {
    auto k = interp_bin_kernel<T, DIM, NS>;
    const size_t lds = ctx.lds_bytes + (ctx.bank ? INTERP_BANK_LDS : 0);
    allow_lds(k, lds);
    dim3 grid(ctx.interp_sub ? ctx.sub_capacity : ctx.nbins[0] * ctx.nbins[1] * ctx.nbins[2],
              blk);
    k<<<grid, BIN_THREADS, lds, stream>>>(ctx, c, fw, nftot);
    launch_check("interp_bin_kernel");
  }
}

template <typename T, int DIM, int NS>
void interp_batch_ns(const InterpItem<T> *items, int nprob, const KerParams<T> &kp,
                     const cplx_t<T> *const *fw_base, cplx_t<T> *const *c_base,
                     int64_t c_vec_offset, int nvec, int64_t total_units, size_t lds_bytes,
                     hipStream_t stream) {
  auto k = interp_batch_kernel<T, DIM, NS>;
  allow_lds(k, lds_bytes);
  if (total_units <= 0) return;
  if (total_units > INT32_MAX) fail(FINUFFT_ERR_MAXNALLOC, "interp_batch: too many work units");
  k<<<unsigned(total_units), MV_THREADS, lds_bytes, stream>>>(items, nprob, kp, fw_base, c_base,
                                                              c_vec_offset, nvec, total_units);
  launch_check("interp_batch_kernel");
}

template <typename T, int DIM, int NS>
void spread_batch_ns(const InterpItem<T> *items, int nprob, const cplx_t<T> *const *c_base,
                     int64_t c_vec_offset, cplx_t<T> *const *fw_base, int nvec,
                     int64_t total_units, size_t lds_bytes, hipStream_t stream) {
  auto k = spread_batch_kernel<T, DIM, NS>;
  allow_lds(k, lds_bytes);
  if (total_units <= 0) return;
  if (total_units > INT32_MAX) fail(FINUFFT_ERR_MAXNALLOC, "spread_batch: too many work units");
  k<<<unsigned(total_units), BIN_THREADS, lds_bytes, stream>>>(items, nprob, c_base, c_vec_offset,
                                                                fw_base, nvec, total_units);
  launch_check("spread_batch_kernel");
}

template <typename T, int DIM, int NS>
void interp_grad_ns(const SpreadCtx<T> &ctx, cplx_t<T> *c, cplx_t<T> *dc, const cplx_t<T> *fw,
                    int blk, hipStream_t stream) {
  if (ctx.M == 0) return;
  dim3 grid((ctx.M + SPREAD_THREADS - 1) / SPREAD_THREADS);
  interp_grad_kernel<T, DIM, NS><<<grid, SPREAD_THREADS, 0, stream>>>(ctx, c, dc, fw,
                                                                      grid_total(ctx), blk);
  launch_check("interp_grad_kernel");
}

// Compile-time width dispatch over [2, max width of the precision].
template <typename T, template <typename, int, int> class Op, int NS = 2>
struct ns_dispatch {
  template <typename... A>
  static void run(int ns, A &&...a) {
    constexpr int MAXW = std::is_same<T, float>::value ? MAX_NS_F32 : MAX_NS;
    if constexpr (NS <= MAXW) {
      if (ns == NS) return Op<T, FA_DIM, NS>::go(std::forward<A>(a)...);
      return ns_dispatch<T, Op, NS + 1>::run(ns, std::forward<A>(a)...);
    } else {
      fail(FINUFFT_ERR_INVALID_ARGUMENT, "kernel width out of range: " + std::to_string(ns));
    }
  }
};

template <typename T, int DIM, int NS>
struct SpreadOp {
  static void go(const SpreadCtx<T> &ctx, const cplx_t<T> *c, cplx_t<T> *fw, int blk,
                 hipStream_t s) {
    spread_ns<T, DIM, NS>(ctx, c, fw, blk, s);
  }
};
template <typename T, int DIM, int NS>
struct InterpOp {
  static void go(const SpreadCtx<T> &ctx, cplx_t<T> *c, const cplx_t<T> *fw, int blk,
                 hipStream_t s) {
    interp_ns<T, DIM, NS>(ctx, c, fw, blk, s);
  }
};

template <typename T, int DIM, int NS>
struct InterpBatchOp {
  static void go(const InterpItem<T> *items, int nprob, const KerParams<T> &kp,
                 const cplx_t<T> *const *fw_base, cplx_t<T> *const *c_base,
                 int64_t c_vec_offset, int nvec, int64_t total_units, size_t lds_bytes,
                 hipStream_t s) {
    interp_batch_ns<T, DIM, NS>(items, nprob, kp, fw_base, c_base, c_vec_offset, nvec,
                                total_units, lds_bytes, s);
  }
};

template <typename T, int DIM, int NS>
struct SpreadBatchOp {
  static void go(const InterpItem<T> *items, int nprob, const cplx_t<T> *const *c_base,
                 int64_t c_vec_offset, cplx_t<T> *const *fw_base, int nvec, int64_t total_units,
                 size_t lds_bytes, hipStream_t s) {
    spread_batch_ns<T, DIM, NS>(items, nprob, c_base, c_vec_offset, fw_base, nvec, total_units,
                                lds_bytes, s);
  }
};

template <typename T, int DIM, int NS>
struct InterpGradOp {
  static void go(const SpreadCtx<T> &ctx, cplx_t<T> *c, cplx_t<T> *dc, const cplx_t<T> *fw,
                 int blk, hipStream_t s) {
    interp_grad_ns<T, DIM, NS>(ctx, c, dc, fw, blk, s);
  }
};

}  // namespace

template <typename T, int DIM>
void interp_grad_dim(const SpreadCtx<T> &ctx, cplx_t<T> *c, cplx_t<T> *dc, const cplx_t<T> *fw,
                     int blk, hipStream_t stream) {
  ns_dispatch<T, InterpGradOp>::run(ctx.ns, ctx, c, dc, fw, blk, stream);
}

template <typename T, int DIM>
void interp_batch_dim(const InterpItem<T> *items, int nprob, int ns, const KerParams<T> &kp,
                      const cplx_t<T> *const *fw_base, cplx_t<T> *const *c_base,
                      int64_t c_vec_offset, int nvec, int64_t total_units, size_t lds_bytes,
                      hipStream_t stream) {
  ns_dispatch<T, InterpBatchOp>::run(ns, items, nprob, kp, fw_base, c_base, c_vec_offset, nvec,
                                     total_units, lds_bytes, stream);
}

template <typename T, int DIM>
void spread_batch_dim(const InterpItem<T> *items, int nprob, int ns,
                      const cplx_t<T> *const *c_base, int64_t c_vec_offset,
                      cplx_t<T> *const *fw_base, int nvec, int64_t total_units, size_t lds_bytes,
                      hipStream_t stream) {
  ns_dispatch<T, SpreadBatchOp>::run(ns, items, nprob, c_base, c_vec_offset, fw_base, nvec,
                                     total_units, lds_bytes, stream);
}

template <typename T, int DIM>
void spread_dim(const SpreadCtx<T> &ctx, const cplx_t<T> *c, cplx_t<T> *fw, int blk,
                hipStream_t stream) {
  ns_dispatch<T, SpreadOp>::run(ctx.ns, ctx, c, fw, blk, stream);
}

template <typename T, int DIM>
void interp_dim(const SpreadCtx<T> &ctx, cplx_t<T> *c, const cplx_t<T> *fw, int blk,
                hipStream_t stream) {
  ns_dispatch<T, InterpOp>::run(ctx.ns, ctx, c, fw, blk, stream);
}

template void spread_dim<FA_REAL, FA_DIM>(const SpreadCtx<FA_REAL> &, const cplx_t<FA_REAL> *,
                                          cplx_t<FA_REAL> *, int, hipStream_t);
template void interp_dim<FA_REAL, FA_DIM>(const SpreadCtx<FA_REAL> &, cplx_t<FA_REAL> *,
                                          const cplx_t<FA_REAL> *, int, hipStream_t);
template void spread_batch_dim<FA_REAL, FA_DIM>(const InterpItem<FA_REAL> *, int, int,
                                                const cplx_t<FA_REAL> *const *, int64_t,
                                                cplx_t<FA_REAL> *const *, int, int64_t, size_t,
                                                hipStream_t);
template void interp_batch_dim<FA_REAL, FA_DIM>(const InterpItem<FA_REAL> *, int, int,
                                                const KerParams<FA_REAL> &,
                                                const cplx_t<FA_REAL> *const *,
                                                cplx_t<FA_REAL> *const *, int64_t, int,
                                                int64_t, size_t, hipStream_t);
template void interp_grad_dim<FA_REAL, FA_DIM>(const SpreadCtx<FA_REAL> &, cplx_t<FA_REAL> *,
                                               cplx_t<FA_REAL> *, const cplx_t<FA_REAL> *, int,
                                               hipStream_t);

}  // namespace fa
