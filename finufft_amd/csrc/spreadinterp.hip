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
#include <string>

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

// One LDS work unit of method 2: a bin (all of its points) or a subproblem
// (<= maxsub points of one bin). Bins are bins of the first support node
// (the sort's start_bins mode), so a bin of b nodes owns exactly the padded
// tile [off, off + b + ns - 1) in each dimension.
template <int DIM, int NS>
struct StartTile {
  int first, npts;
  int off[3];
  int P[3];
  int ntile;

  template <typename T>
  __device__ void set_geometry(const SpreadCtx<T> &ctx, int b) {
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
      int g = nd + off[d];
      g = g >= ctx.nf[d] ? g - ctx.nf[d] : g;
      gi += g * stride;
      stride *= ctx.nf[d];
    }
    return gi;
  }
};

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
  const int col[3] = {col0, col1, col2};
  int rem = blockIdx.x, b = 0, stride = 1;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int hb = d < DIM ? (ctx.nbins[d] - col[d] + 1) / 2 : 1;
    const int bd = d < DIM ? 2 * (rem % hb) + col[d] : 0;
    rem /= hb;
    b += bd * stride;
    stride *= ctx.nbins[d];
  }
  StartTile<DIM, NS> st;
  st.init_bin(ctx, b);
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
  cplx_t<T> *g = fw + int64_t(t) * nftot;
  for (int n = threadIdx.x; n < st.ntile; n += BIN_THREADS) {
    const int64_t gi = st.grid_index(ctx, n);
    cplx_t<T> v = g[gi];
    v.x += T(re[n]);
    v.y += T(im[n]);
    g[gi] = v;
  }
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
  cplx_t<T> *g = fw + int64_t(t) * nftot;
  for (int n = threadIdx.x; n < st.ntile; n += BIN_THREADS) {
    const T vr = T(re[n]), vi = T(im[n]);
    if (vr == T(0) && vi == T(0)) continue;
    hbm_add<T>(g + st.grid_index(ctx, n), vr, vi);
  }
}

// Whole-bin interpolation: stage the padded tile, then gather per point.
template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(BIN_THREADS)
    interp_bin_kernel(SpreadCtx<T> ctx, cplx_t<T> *__restrict__ c,
                      const cplx_t<T> *__restrict__ fw, int64_t nftot) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *re = reinterpret_cast<T *>(lds_raw);
  StartTile<DIM, NS> st;
  st.init_bin(ctx, blockIdx.x);
  if (st.npts == 0) return;
  T *im = re + st.ntile;
  const int t = blockIdx.y;
  const cplx_t<T> *g = fw + int64_t(t) * nftot;
  for (int n = threadIdx.x; n < st.ntile; n += BIN_THREADS) {
    const cplx_t<T> v = g[st.grid_index(ctx, n)];
    re[n] = v.x;
    im[n] = v.y;
  }
  __syncthreads();
  cplx_t<T> *cout = c + int64_t(t) * ctx.M;
  for (int i = threadIdx.x; i < st.npts; i += BIN_THREADS) {
    const int q = st.first + i;
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
        ar = std::fma(re[l[0] + a], ker[0][a], ar);
        ai = std::fma(im[l[0] + a], ker[0][a], ai);
      }
    } else if constexpr (DIM == 2) {
      for (int b = 0; b < NS; ++b) {
        const int row = (l[1] + b) * st.P[0] + l[0];
        T rr = 0, ri = 0;
#pragma unroll
        for (int a = 0; a < NS; ++a) {
          rr = std::fma(re[row + a], ker[0][a], rr);
          ri = std::fma(im[row + a], ker[0][a], ri);
        }
        ar = std::fma(rr, ker[1][b], ar);
        ai = std::fma(ri, ker[1][b], ai);
      }
    } else {
      for (int e = 0; e < NS; ++e) {
        T pr = 0, pi = 0;
        for (int b = 0; b < NS; ++b) {
          const int row = ((l[2] + e) * st.P[1] + l[1] + b) * st.P[0] + l[0];
          T rr = 0, ri = 0;
#pragma unroll
          for (int a = 0; a < NS; ++a) {
            rr = std::fma(re[row + a], ker[0][a], rr);
            ri = std::fma(im[row + a], ker[0][a], ri);
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

// ---------------------------------------------------------------- launchers

template <typename T>
int64_t grid_total(const SpreadCtx<T> &ctx) {
  return int64_t(ctx.nf[0]) * ctx.nf[1] * ctx.nf[2];
}

template <typename K>
void allow_lds(K kernel, size_t bytes) {
  if (bytes > 48 * 1024)
    hip_check(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)),
              "hipFuncSetAttribute(LDS)");
}

template <typename T, int DIM, int NS>
void spread_ns(const SpreadCtx<T> &ctx, const cplx_t<T> *c, cplx_t<T> *fw, int blk,
               hipStream_t stream) {
  if (ctx.M == 0) return;
  const int64_t nftot = grid_total(ctx);
  if (ctx.method == 1) {
    dim3 grid((ctx.M + SPREAD_THREADS - 1) / SPREAD_THREADS, blk);
    spread_gm_kernel<T, DIM, NS><<<grid, SPREAD_THREADS, 0, stream>>>(ctx, c, fw, nftot);
    launch_check("spread_gm_kernel");
  } else if (ctx.colored) {
    auto k = spread_bin_colored_kernel<T, DIM, NS>;
    allow_lds(k, ctx.lds_bytes);
    for (int color = 0; color < (1 << DIM); ++color) {
      int col[3] = {0, 0, 0};
      int64_t nblk = 1;
      for (int d = 0; d < DIM; ++d) {
        col[d] = (color >> d) & 1;
        nblk *= (ctx.nbins[d] - col[d] + 1) / 2;
      }
      if (nblk == 0) continue;
      k<<<dim3(unsigned(nblk), blk), BIN_THREADS, ctx.lds_bytes, stream>>>(
          ctx, c, fw, nftot, col[0], col[1], col[2]);
      launch_check("spread_bin_colored_kernel");
    }
  } else {
    auto k = spread_sub_kernel<T, DIM, NS>;
    allow_lds(k, ctx.lds_bytes);
    dim3 grid(ctx.sub_capacity, blk);
    k<<<grid, BIN_THREADS, ctx.lds_bytes, stream>>>(ctx, c, fw, nftot);
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
    allow_lds(k, ctx.lds_bytes);
    dim3 grid(ctx.nbins[0] * ctx.nbins[1] * ctx.nbins[2], blk);
    k<<<grid, BIN_THREADS, ctx.lds_bytes, stream>>>(ctx, c, fw, nftot);
    launch_check("interp_bin_kernel");
  }
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

}  // namespace

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

}  // namespace fa
