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
    st[d] = start_and_kernel<T, NS>(ctx.xyz[d][j], ctx.nf[d], ker[d], ctx.kp);
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
    st[d] = start_and_kernel<T, NS>(ctx.xyz[d][j], ctx.nf[d], ker[d], ctx.kp);
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

// Geometry of one subproblem: its points and its padded LDS tile. The tile
// spans bin + ceil(ns/2) cells on the low side and ceil(ns/2) + 1 on the high
// side; the extra high cell absorbs a point that floor(r * (1/binsize))
// rounded into the bin below its true one.
template <int DIM, int NS>
struct SubTile {
  static constexpr int H = (NS + 1) / 2;
  int first, npts;
  int off[3];
  int P[3];
  int ntile;

  template <typename T>
  __device__ bool init(const SpreadCtx<T> &ctx, int s) {
    const int nb_tot = ctx.nbins[0] * ctx.nbins[1] * ctx.nbins[2];
    if (s >= ctx.substart[nb_tot]) return false;
    const int b = ctx.sub2bin[s];
    first = ctx.binstart[b] + (s - ctx.substart[b]) * ctx.maxsub;
    npts = min(ctx.maxsub, ctx.binstart[b + 1] - first);
    int rem = b;
    ntile = 1;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int bd = rem % ctx.nbins[d];
      rem /= ctx.nbins[d];
      off[d] = bd * ctx.binsize[d];
      P[d] = d < DIM ? ctx.binsize[d] + 2 * H + 1 : 1;
      ntile *= P[d];
    }
    return true;
  }

  // Fine-grid index of tile cell n, or -1 when no point can reach it.
  template <typename T>
  __device__ int64_t grid_index(const SpreadCtx<T> &ctx, int n) const {
    int64_t gi = 0, stride = 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const int nd = n % P[d];
      n /= P[d];
      int g = nd + off[d] - H;
      if (g >= ctx.nf[d] + H) return -1;
      g = g < 0 ? g + ctx.nf[d] : (g >= ctx.nf[d] ? g - ctx.nf[d] : g);
      gi += g * stride;
      stride *= ctx.nf[d];
    }
    return gi;
  }
};

template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(SPREAD_THREADS)
    spread_sm_kernel(SpreadCtx<T> ctx, const cplx_t<T> *__restrict__ c,
                     cplx_t<T> *__restrict__ fw, int64_t nftot) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  cplx_t<T> *tile = reinterpret_cast<cplx_t<T> *>(lds_raw);
  SubTile<DIM, NS> st;
  if (!st.init(ctx, blockIdx.x)) return;
  constexpr int H = SubTile<DIM, NS>::H;
  const int t = blockIdx.y;
  for (int n = threadIdx.x; n < st.ntile; n += SPREAD_THREADS) tile[n] = cplx_t<T>{0, 0};
  __syncthreads();
  for (int i = threadIdx.x; i < st.npts; i += SPREAD_THREADS) {
    const int j = ctx.idx[st.first + i];
    T ker[DIM][NS];
    int l[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      l[d] = start_and_kernel<T, NS>(ctx.xyz[d][j], ctx.nf[d], ker[d], ctx.kp) - st.off[d] + H;
    const cplx_t<T> cv = c[int64_t(t) * ctx.M + j];
    if constexpr (DIM == 1) {
#pragma unroll
      for (int a = 0; a < NS; ++a) lds_add<T>(tile + l[0] + a, cv.x * ker[0][a], cv.y * ker[0][a]);
    } else if constexpr (DIM == 2) {
      for (int b = 0; b < NS; ++b) {
        const T wr = cv.x * ker[1][b], wi = cv.y * ker[1][b];
        cplx_t<T> *row = tile + (l[1] + b) * st.P[0] + l[0];
#pragma unroll
        for (int a = 0; a < NS; ++a) lds_add<T>(row + a, wr * ker[0][a], wi * ker[0][a]);
      }
    } else {
      for (int e = 0; e < NS; ++e) {
        const T zr = cv.x * ker[2][e], zi = cv.y * ker[2][e];
        for (int b = 0; b < NS; ++b) {
          const T wr = zr * ker[1][b], wi = zi * ker[1][b];
          cplx_t<T> *row = tile + ((l[2] + e) * st.P[1] + l[1] + b) * st.P[0] + l[0];
#pragma unroll
          for (int a = 0; a < NS; ++a) lds_add<T>(row + a, wr * ker[0][a], wi * ker[0][a]);
        }
      }
    }
  }
  __syncthreads();
  cplx_t<T> *g = fw + int64_t(t) * nftot;
  for (int n = threadIdx.x; n < st.ntile; n += SPREAD_THREADS) {
    const cplx_t<T> v = tile[n];
    if (v.x == T(0) && v.y == T(0)) continue;
    const int64_t gi = st.grid_index(ctx, n);
    if (gi >= 0) hbm_add<T>(g + gi, v.x, v.y);
  }
}

template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(SPREAD_THREADS)
    interp_sm_kernel(SpreadCtx<T> ctx, cplx_t<T> *__restrict__ c,
                     const cplx_t<T> *__restrict__ fw, int64_t nftot) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  cplx_t<T> *tile = reinterpret_cast<cplx_t<T> *>(lds_raw);
  SubTile<DIM, NS> st;
  if (!st.init(ctx, blockIdx.x)) return;
  constexpr int H = SubTile<DIM, NS>::H;
  const int t = blockIdx.y;
  const cplx_t<T> *g = fw + int64_t(t) * nftot;
  for (int n = threadIdx.x; n < st.ntile; n += SPREAD_THREADS) {
    const int64_t gi = st.grid_index(ctx, n);
    tile[n] = gi >= 0 ? g[gi] : cplx_t<T>{0, 0};
  }
  __syncthreads();
  for (int i = threadIdx.x; i < st.npts; i += SPREAD_THREADS) {
    const int j = ctx.idx[st.first + i];
    T ker[DIM][NS];
    int l[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      l[d] = start_and_kernel<T, NS>(ctx.xyz[d][j], ctx.nf[d], ker[d], ctx.kp) - st.off[d] + H;
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
          rr = std::fma(row[a].x, ker[0][a], rr);
          ri = std::fma(row[a].y, ker[0][a], ri);
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
            rr = std::fma(row[a].x, ker[0][a], rr);
            ri = std::fma(row[a].y, ker[0][a], ri);
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
  } else {
    auto k = spread_sm_kernel<T, DIM, NS>;
    allow_lds(k, ctx.lds_bytes);
    dim3 grid(ctx.sub_capacity, blk);
    k<<<grid, SPREAD_THREADS, ctx.lds_bytes, stream>>>(ctx, c, fw, nftot);
    launch_check("spread_sm_kernel");
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
    auto k = interp_sm_kernel<T, DIM, NS>;
    allow_lds(k, ctx.lds_bytes);
    dim3 grid(ctx.sub_capacity, blk);
    k<<<grid, SPREAD_THREADS, ctx.lds_bytes, stream>>>(ctx, c, fw, nftot);
    launch_check("interp_sm_kernel");
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
