// C ABI: the cuFINUFFT guru and simple interfaces (include/cufinufft.h) and
// the extensions of include/finufft_amd.h, over fa::Plan<T>.
//
// Argument checks and their order follow the reference C layer
// (src/cuda/c_interface.cpp:14-217): dim, then the mode array, then the plan
// constructor's own checks; every exception becomes an integer code.
#include <cufinufft.h>
#include <finufft_amd.h>

#include <algorithm>
#include <array>
#include <vector>
#include <cstdint>
#include <limits>
#include <memory>
#include <mutex>
#include <unordered_set>

#include "batch.hpp"
#include "kernel_math.hpp"
#include "density.hpp"
#include "mlp.hpp"
#include "plan.hpp"

namespace fa {
void lds_probe(int mode, int blocks, int reps, double *ms, double *wave_ops);
}

using fa::Plan;
using fa::safe_call;

namespace {

constexpr int64_t I32MAX = std::numeric_limits<int32_t>::max();

// Mode sizes must be in (0, INT32_MAX] with product <= INT32_MAX; type 3
// ignores them (c_interface.cpp:14-31).
bool invalid_modes(int type, int dim, const int64_t *m64, int m32[3]) {
  if (type == 3) {
    m32[0] = m32[1] = m32[2] = 1;
    return false;
  }
  int64_t tot = 1;
  for (int i = 0; i < dim; ++i) {
    if (m64[i] > I32MAX || m64[i] <= 0) return true;
    m32[i] = int(m64[i]);
    tot *= m64[i];
  }
  for (int i = dim; i < 3; ++i) m32[i] = 1;
  return tot > I32MAX;
}

// Live handles of one kind (fp32 plans, fp64 plans, batches). Every entry
// point looks its handle up before using it, so that a destroyed handle, a
// plan of the other precision or a stray pointer is refused with
// FINUFFT_ERR_PLAN_NOTVALID instead of being dereferenced.
template <typename K>
struct Live {
  std::mutex mu;
  std::unordered_set<const void *> set;

  static Live &get() {
    static Live l;
    return l;
  }
  void add(const void *p) {
    std::lock_guard<std::mutex> lk(mu);
    set.insert(p);
  }
  bool has(const void *p) {
    std::lock_guard<std::mutex> lk(mu);
    return set.count(p) != 0;
  }
  bool remove(const void *p) {
    std::lock_guard<std::mutex> lk(mu);
    return set.erase(p) != 0;
  }
};

template <typename T>
Plan<T> *as_plan(void *p) {
  if (!p) fa::fail(FINUFFT_ERR_PLAN_NOTVALID, "null plan");
  if (!Live<Plan<T>>::get().has(p))
    fa::fail(FINUFFT_ERR_PLAN_NOTVALID, "not a live plan of this precision");
  return static_cast<Plan<T> *>(p);
}

template <typename T>
void destroy_plan(void *p) {
  if (!p) fa::fail(FINUFFT_ERR_PLAN_NOTVALID, "null plan");
  Plan<T> *plan = as_plan<T>(p);
  // A batch holds raw pointers to its plans: a plan is destroyed only after
  // every batch made from it.
  if (plan->batch_refs() > 0)
    fa::fail(FINUFFT_ERR_INVALID_ARGUMENT, "plan is held by a live batch; destroy the batch first");
  if (!Live<Plan<T>>::get().remove(p))
    fa::fail(FINUFFT_ERR_PLAN_NOTVALID, "not a live plan of this precision");
  delete plan;
}

template <typename T, typename P>
int makeplan(int type, int dim, const int64_t *n_modes, int iflag, int ntr, T eps, P *out,
             const cufinufft_opts *opts) {
  return safe_call([&] {
    if (dim < 1 || dim > 3) fa::fail(FINUFFT_ERR_DIM_NOTVALID, "invalid dim");
    int m32[3];
    if (invalid_modes(type, dim, n_modes, m32))
      fa::fail(FINUFFT_ERR_NDATA_NOTVALID, "invalid mode sizes");
    cufinufft_opts o;
    if (opts)
      o = *opts;
    else
      cufinufft_default_opts(&o);
    if (!out) fa::fail(FINUFFT_ERR_INVALID_ARGUMENT, "null plan pointer");
    auto plan = std::make_unique<Plan<T>>(type, dim, m32, iflag, ntr, eps, o);
    Live<Plan<T>>::get().add(plan.get());
    *out = reinterpret_cast<P>(plan.release());
  });
}

template <typename T>
int setpts(void *plan, int64_t M, const T *x, const T *y, const T *z, int N, const T *s,
           const T *t, const T *u) {
  return safe_call([&] {
    if (M > I32MAX) fa::fail(FINUFFT_ERR_NDATA_NOTVALID, "M exceeds INT32_MAX");
    as_plan<T>(plan)->setpts(int(M), x, y, z, N, s, t, u);
  });
}

template <typename T>
int simple(int dim, int type, int ntr, int64_t nj, const std::array<const T *, 3> &xyz,
           fa::cplx_t<T> *cj, int iflag, T eps, const std::array<int64_t, 3> &modes, int64_t nk,
           const std::array<const T *, 3> &stu, fa::cplx_t<T> *fk, const cufinufft_opts *opts) {
  return safe_call([&] {
    if (dim < 1 || dim > 3) fa::fail(FINUFFT_ERR_DIM_NOTVALID, "invalid dim");
    if (nj > I32MAX || nk > I32MAX) fa::fail(FINUFFT_ERR_NDATA_NOTVALID, "size exceeds INT32_MAX");
    int m32[3];
    if (invalid_modes(type, dim, modes.data(), m32))
      fa::fail(FINUFFT_ERR_NDATA_NOTVALID, "invalid mode sizes");
    cufinufft_opts o;
    if (opts)
      o = *opts;
    else
      cufinufft_default_opts(&o);
    auto plan = std::make_unique<Plan<T>>(type, dim, m32, iflag, ntr, eps, o);
    plan->setpts(int(nj), xyz[0], xyz[1], xyz[2], int(nk), stu[0], stu[1], stu[2]);
    plan->execute(cj, fk);
  });
}

template <typename T>
fa::cplx_t<T> *nc(const fa::cplx_t<T> *p) {
  return const_cast<fa::cplx_t<T> *>(p);
}

}  // namespace

namespace {
// A batch handle remembers its precision.
struct BatchHandle {
  int is_double;
  std::unique_ptr<fa::BatchPlan<float>> f;
  std::unique_ptr<fa::BatchPlan<double>> d;
};
BatchHandle *as_batch(void *b) {
  if (!b) fa::fail(FINUFFT_ERR_PLAN_NOTVALID, "null batch");
  if (!Live<BatchHandle>::get().has(b)) fa::fail(FINUFFT_ERR_PLAN_NOTVALID, "not a live batch");
  return static_cast<BatchHandle *>(b);
}
template <typename T>
std::unique_ptr<fa::BatchPlan<T>> make_batch(int n, void *const *plans, size_t arena) {
  if (n < 1 || plans == nullptr) fa::fail(FINUFFT_ERR_INVALID_ARGUMENT, "batch: no plans");
  std::vector<Plan<T> *> v(n);
  for (int i = 0; i < n; ++i) v[i] = as_plan<T>(plans[i]);
  const size_t budget = arena ? arena : (size_t(192) << 20);
  const hipStream_t stream = v[0]->stream();
  return std::make_unique<fa::BatchPlan<T>>(std::move(v), budget, stream);
}
}  // namespace

extern "C" {

void cufinufft_default_opts(cufinufft_opts *opts) {
  if (!opts) return;
  // Values of src/cuda/c_interface.cpp:153-172.
  opts->modeord = 0;
  opts->gpu_device_id = 0;
  opts->gpu_spreadinterponly = 0;
  opts->gpu_method = 0;
  opts->gpu_sort = 1;
  opts->gpu_kerevalmeth = 1;
  opts->upsampfac = 0.0;
  opts->gpu_maxsubprobsize = 1024;
  opts->gpu_obinsizex = 0;
  opts->gpu_obinsizey = 0;
  opts->gpu_obinsizez = 0;
  opts->gpu_binsizex = 0;
  opts->gpu_binsizey = 0;
  opts->gpu_binsizez = 0;
  opts->gpu_maxbatchsize = 0;
  opts->gpu_np = 0;
  opts->debug = 0;
  opts->gpu_stream = nullptr;
}

int cufinufft_makeplan(int type, int dim, const int64_t *n_modes, int iflag, int ntr, double eps,
                       cufinufft_plan *d_plan_ptr, const cufinufft_opts *opts) {
  return makeplan<double>(type, dim, n_modes, iflag, ntr, eps, d_plan_ptr, opts);
}
int cufinufftf_makeplan(int type, int dim, const int64_t *n_modes, int iflag, int ntr, float eps,
                        cufinufftf_plan *d_plan_ptr, const cufinufft_opts *opts) {
  return makeplan<float>(type, dim, n_modes, iflag, ntr, eps, d_plan_ptr, opts);
}

int cufinufft_setpts(cufinufft_plan p, int64_t M, const double *x, const double *y,
                     const double *z, int N, const double *s, const double *t, const double *u) {
  return setpts<double>(p, M, x, y, z, N, s, t, u);
}
int cufinufftf_setpts(cufinufftf_plan p, int64_t M, const float *x, const float *y,
                      const float *z, int N, const float *s, const float *t, const float *u) {
  return setpts<float>(p, M, x, y, z, N, s, t, u);
}

int cufinufft_execute(cufinufft_plan p, hipDoubleComplex *c, hipDoubleComplex *fk) {
  return safe_call([&] { as_plan<double>(p)->execute(c, fk); });
}
int cufinufftf_execute(cufinufftf_plan p, hipFloatComplex *c, hipFloatComplex *fk) {
  return safe_call([&] { as_plan<float>(p)->execute(c, fk); });
}

int cufinufft_destroy(cufinufft_plan p) {
  return safe_call([&] { destroy_plan<double>(p); });
}
int cufinufftf_destroy(cufinufftf_plan p) {
  return safe_call([&] { destroy_plan<float>(p); });
}

int cufinufft_execute_adjoint(cufinufft_plan p, hipDoubleComplex *c, hipDoubleComplex *fk) {
  return safe_call([&] { as_plan<double>(p)->execute_adjoint(c, fk); });
}
int cufinufftf_execute_adjoint(cufinufftf_plan p, hipFloatComplex *c, hipFloatComplex *fk) {
  return safe_call([&] { as_plan<float>(p)->execute_adjoint(c, fk); });
}

int cufinufft_interp_grad(cufinufft_plan p, hipDoubleComplex *c, hipDoubleComplex *dc,
                          hipDoubleComplex *fk) {
  return safe_call([&] { as_plan<double>(p)->interp_grad(c, dc, fk); });
}
int cufinufftf_interp_grad(cufinufftf_plan p, hipFloatComplex *c, hipFloatComplex *dc,
                           hipFloatComplex *fk) {
  return safe_call([&] { as_plan<float>(p)->interp_grad(c, dc, fk); });
}


int finufft_amd_batch_makeplan(int is_double, int nplans, void *const *plans, size_t arena_bytes,
                               finufft_amd_batch *out) {
  return safe_call([&] {
    if (!out) fa::fail(FINUFFT_ERR_INVALID_ARGUMENT, "batch: null output handle");
    *out = nullptr;
    auto h = std::make_unique<BatchHandle>();
    h->is_double = is_double ? 1 : 0;
    if (is_double)
      h->d = make_batch<double>(nplans, plans, arena_bytes);
    else
      h->f = make_batch<float>(nplans, plans, arena_bytes);
    Live<BatchHandle>::get().add(h.get());
    *out = h.release();
  });
}

int finufft_amd_batch_execute(finufft_amd_batch batch, void *const *c, void *const *fk) {
  return safe_call([&] {
    BatchHandle *h = as_batch(batch);
    if (!c || !fk) fa::fail(FINUFFT_ERR_INVALID_ARGUMENT, "batch: null pointer array");
    if (h->is_double)
      h->d->execute(reinterpret_cast<double2 *const *>(c), reinterpret_cast<double2 *const *>(fk));
    else
      h->f->execute(reinterpret_cast<float2 *const *>(c), reinterpret_cast<float2 *const *>(fk));
  });
}

int finufft_amd_batch_destroy(finufft_amd_batch batch) {
  return safe_call([&] {
    if (!batch) fa::fail(FINUFFT_ERR_PLAN_NOTVALID, "null batch");
    if (!Live<BatchHandle>::get().remove(batch))
      fa::fail(FINUFFT_ERR_PLAN_NOTVALID, "not a live batch");
    delete static_cast<BatchHandle *>(batch);
  });
}

int finufft_amd_batch_info(finufft_amd_batch batch, int64_t *iout) {
  return safe_call([&] {
    BatchHandle *h = as_batch(batch);
    if (!iout) return;
    auto fill = [&](const auto &b) {
      iout[0] = b.nprob();
      iout[1] = b.vectors_per_round();
      iout[2] = b.groups();
      iout[3] = b.graph() ? 1 : 0;
      iout[4] = b.capture_stage();
      iout[5] = b.capture_error();
      iout[6] = b.chunks();
    };
    if (h->is_double)
      fill(*h->d);
    else
      fill(*h->f);
  });
}

int finufft_amd_batch_set_graph(finufft_amd_batch batch, int use_graph) {
  return safe_call([&] {
    BatchHandle *h = as_batch(batch);
    if (h->is_double)
      h->d->set_use_graph(use_graph != 0);
    else
      h->f->set_use_graph(use_graph != 0);
  });
}

int finufft_amd_mlp_forward(const float *params, int n, const float *x, float *y,
                            void *stream) {
  return safe_call([&] {
    if (n < 0 || (n > 0 && (!params || !x || !y)))
      fa::fail(FINUFFT_ERR_INVALID_ARGUMENT, "mlp_forward: invalid arguments");
    fa::mlp_forward(params, n, x, y, static_cast<hipStream_t>(stream));
  });
}

int finufft_amd_mlp_grad(const float *params, int n, const float *x, const float *target,
                         const float *mask, float *grad, float *loss, void *stream) {
  return safe_call([&] {
    if (n < 0 || (n > 0 && (!params || !x || !target || !mask || !grad || !loss)))
      fa::fail(FINUFFT_ERR_INVALID_ARGUMENT, "mlp_grad: invalid arguments");
    fa::mlp_grad(params, n, x, target, mask, grad, loss, static_cast<hipStream_t>(stream));
  });
}

int finufft_amd_mlp_adam(float *params, const float *grad, float *m1, float *m2, int step,
                         float lr, void *stream) {
  return safe_call([&] {
    if (!params || !grad || !m1 || !m2 || step < 1)
      fa::fail(FINUFFT_ERR_INVALID_ARGUMENT, "mlp_adam: invalid arguments");
    fa::mlp_adam(params, grad, m1, m2, step, lr, static_cast<hipStream_t>(stream));
  });
}

int finufft_amd_plan_info(const void *plan, int is_double, int64_t *iout, double *dout) {
  return safe_call([&] {
    auto fill = [&](const auto *p) {
      if (iout) {
        const int64_t v[16] = {p->type(),       p->dim(),       p->ns(),     p->nf(0),
                               p->nf(1),        p->nf(2),       p->method(), p->binsize(0),
                               p->binsize(1),   p->binsize(2),  p->batchsize(), p->M(),
                               p->nbins_total(), p->col_side(), p->colored(),
                               p->dense_subproblems()};
        for (int i = 0; i < 16; ++i) iout[i] = v[i];
      }
      if (dout) {
        dout[0] = p->beta();
        dout[1] = p->upsampfac();
        dout[2] = p->od_np();
        dout[3] = p->fft_path();
      }
    };
    void *q = const_cast<void *>(plan);
    if (is_double)
      fill(as_plan<double>(q));
    else
      fill(as_plan<float>(q));
  });
}

int finufft_amd_copy_sort(const void *plan, int is_double, int *perm, int *binstart) {
  return safe_call([&] {
    auto copy = [&](const auto *p) {
      const hipStream_t s = p->stream();
      if (perm && p->M() > 0)
        fa::hip_check(hipMemcpyAsync(perm, p->sort_perm(), size_t(p->M()) * sizeof(int),
                                     hipMemcpyDeviceToDevice, s),
                      "hipMemcpyAsync(perm)");
      if (binstart && p->bin_start())
        fa::hip_check(hipMemcpyAsync(binstart, p->bin_start(),
                                     (size_t(p->nbins_total()) + 1) * sizeof(int),
                                     hipMemcpyDeviceToDevice, s),
                      "hipMemcpyAsync(binstart)");
      fa::hip_check(hipStreamSynchronize(s), "hipStreamSynchronize(copy_sort)");
    };
    void *q = const_cast<void *>(plan);
    if (is_double)
      copy(as_plan<double>(q));
    else
      copy(as_plan<float>(q));
  });
}

int finufft_amd_set_stage_timing(void *plan, int is_double, int on) {
  return safe_call([&] {
    if (is_double)
      as_plan<double>(plan)->timer().enable(on != 0);
    else
      as_plan<float>(plan)->timer().enable(on != 0);
  });
}

int finufft_amd_set_fft_path(int path) {
  return safe_call([&] {
    if (path != 0 && path != 1) fa::fail(FINUFFT_ERR_INVALID_ARGUMENT, "fft path must be 0 or 1");
    fa::set_fft_path_setting(path);
  });
}

int finufft_amd_get_stage_times(void *plan, int is_double, double *ms) {
  return safe_call([&] {
    if (!ms) fa::fail(FINUFFT_ERR_INVALID_ARGUMENT, "null output");
    if (is_double)
      as_plan<double>(plan)->timer().collect(ms);
    else
      as_plan<float>(plan)->timer().collect(ms);
  });
}

int finufft_amd_density_stats(int dim, int64_t M, const void *x, const void *y, const void *z,
                              int is_double, int g, double *stats, void *stream) {
  return safe_call([&] {
    if (!stats) fa::fail(FINUFFT_ERR_INVALID_ARGUMENT, "density: null stats");
    auto s = static_cast<hipStream_t>(stream);
    if (is_double)
      fa::density_stats<double>(dim, M, static_cast<const double *>(x),
                                static_cast<const double *>(y), static_cast<const double *>(z),
                                g, stats, s);
    else
      fa::density_stats<float>(dim, M, static_cast<const float *>(x),
                               static_cast<const float *>(y), static_cast<const float *>(z), g,
                               stats, s);
  });
}

int finufft_amd_density_stats_noalloc(int dim, int64_t M, const void *x, const void *y,
                                      const void *z, int is_double, int g, double *d_stats,
                                      void *stream) {
  return safe_call([&] {
    auto s = static_cast<hipStream_t>(stream);
    if (is_double)
      fa::density_stats_noalloc<double>(dim, M, static_cast<const double *>(x),
                                        static_cast<const double *>(y),
                                        static_cast<const double *>(z), g, d_stats, s);
    else
      fa::density_stats_noalloc<float>(dim, M, static_cast<const float *>(x),
                                       static_cast<const float *>(y),
                                       static_cast<const float *>(z), g, d_stats, s);
  });
}

int finufft_amd_lds_probe(int mode, int blocks, int reps, double *ms, double *wave_ops) {
  return safe_call([&] {
    if (!ms || !wave_ops) fa::fail(FINUFFT_ERR_INVALID_ARGUMENT, "null output");
    fa::lds_probe(mode, blocks, reps, ms, wave_ops);
  });
}

int64_t finufft_amd_next235(int64_t n, int64_t f) { return fa::next235(n, f); }

void finufft_amd_gaussquad(int n, double *x, double *w) {
  if (n > 0 && x && w) fa::gaussquad(n, x, w);
}

int finufft_amd_kernel_shape(double tol, int is_double, double upsampfac, int kerevalmeth,
                             int dim, int type, int *ns, double *beta, int *formula) {
  return safe_call([&] {
    if (kerevalmeth < 0 || kerevalmeth > 2)
      fa::fail(FINUFFT_ERR_KERFORMULA_NOTVALID, "gpu_kerevalmeth must be 0, 1 or 2");
    if (upsampfac != 2.0 && upsampfac != 1.25) {
      if (kerevalmeth != 0) fa::fail(FINUFFT_ERR_HORNER_WRONG_BETA, "Horner needs sigma 2 or 1.25");
      if (upsampfac <= 1.0) fa::fail(FINUFFT_ERR_UPSAMPFAC_TOO_SMALL, "upsampfac must exceed 1");
    }
    fa::KernelShape ks;
    if (is_double) {
      const double e = std::max(tol, std::numeric_limits<double>::epsilon());
      ks = kerevalmeth == 1 ? fa::kernel_shape_pswf(e, upsampfac, dim, type, fa::MAX_NS, false)
                            : fa::kernel_shape(e, e / 10.0, upsampfac, fa::MAX_NS);
    } else {
      const float e = std::max(float(tol), std::numeric_limits<float>::epsilon());
      ks = kerevalmeth == 1
               ? fa::kernel_shape_pswf(double(e), upsampfac, dim, type, fa::MAX_NS_F32, true)
               : fa::kernel_shape(double(e), double(e / 10.0f), upsampfac, fa::MAX_NS_F32);
    }
    if (ns) *ns = ks.ns;
    if (beta) *beta = ks.beta;
    if (formula) *formula = ks.formula;
  });
}

int finufft_amd_horner_coeffs(int ns, double upsampfac, double *out, int nc_max) {
  if (ns < 2 || ns > fa::MAX_NS || !out) return 0;
  const int nc = std::min(nc_max, fa::horner_nc_sigma2(ns));
  const double bo = upsampfac == 2.0 ? (ns == 2 ? 2.20 : ns == 3 ? 2.26 : ns == 4 ? 2.38 : 2.30)
                                     : 0.97 * fa::PI * (1.0 - 1.0 / (2.0 * upsampfac));
  const std::vector<double> c = fa::horner_coeffs(ns, upsampfac, bo * ns, nc);
  std::copy(c.begin(), c.end(), out);
  return nc;
}

int finufft_amd_horner_coeffs_pswf(int ns, double beta, double *out, int nc_max) {
  if (ns < 2 || ns > fa::MAX_NS || !out || beta <= 0) return 0;
  const int nc = std::min(nc_max, fa::horner_nc_sigma2(ns));
  const std::vector<double> c = fa::horner_coeffs_pswf(ns, beta, nc);
  std::copy(c.begin(), c.end(), out);
  return nc;
}

void finufft_amd_pswf(double c, const double *z, int64_t n, double *out) {
  if (!z || !out || n <= 0 || !(c > 0)) return;
  const fa::Prolate0 psi(c);
  for (int64_t i = 0; i < n; ++i) out[i] = psi(z[i]);
}

void finufft_amd_fold_rescale_f(const float *x, int64_t m, int n, float *out) {
  for (int64_t i = 0; i < m; ++i) out[i] = fa::fold_rescale(x[i], n);
}
void finufft_amd_fold_rescale_d(const double *x, int64_t m, int n, double *out) {
  for (int64_t i = 0; i < m; ++i) out[i] = fa::fold_rescale(x[i], n);
}

int64_t finufft_amd_fine_grid_len(int64_t m, double upsampfac, int ns) {
  return fa::fine_grid_len(m, upsampfac, ns);
}

// ---- simple interfaces (c_interface.cpp:219-470)

#define FA_SIMPLE_1D(P, T, CT)                                                                 \
  int P##1d1many(int nt, int64_t nj, const T *x, const CT *c, int fl, T e, int64_t ms, CT *fk, \
                 const cufinufft_opts *o) {                                                    \
    return simple<T>(1, 1, nt, nj, {x, nullptr, nullptr}, nc<T>(c), fl, e, {ms, 1, 1}, 0, {},  \
                     fk, o);                                                                   \
  }                                                                                            \
  int P##1d1(int64_t nj, const T *x, const CT *c, int fl, T e, int64_t ms, CT *fk,             \
             const cufinufft_opts *o) {                                                        \
    return P##1d1many(1, nj, x, c, fl, e, ms, fk, o);                                          \
  }                                                                                            \
  int P##1d2many(int nt, int64_t nj, const T *x, CT *c, int fl, T e, int64_t ms, const CT *fk, \
                 const cufinufft_opts *o) {                                                    \
    return simple<T>(1, 2, nt, nj, {x, nullptr, nullptr}, c, fl, e, {ms, 1, 1}, 0, {},         \
                     nc<T>(fk), o);                                                            \
  }                                                                                            \
  int P##1d2(int64_t nj, const T *x, CT *c, int fl, T e, int64_t ms, const CT *fk,             \
             const cufinufft_opts *o) {                                                        \
    return P##1d2many(1, nj, x, c, fl, e, ms, fk, o);                                          \
  }                                                                                            \
  int P##1d3many(int nt, int64_t nj, const T *x, const CT *c, int fl, T e, int64_t nk,         \
                 const T *s, CT *fk, const cufinufft_opts *o) {                                \
    return simple<T>(1, 3, nt, nj, {x, nullptr, nullptr}, nc<T>(c), fl, e, {0, 0, 0}, nk,      \
                     {s, nullptr, nullptr}, fk, o);                                            \
  }                                                                                            \
  int P##1d3(int64_t nj, const T *x, const CT *c, int fl, T e, int64_t nk, const T *s, CT *fk, \
             const cufinufft_opts *o) {                                                        \
    return P##1d3many(1, nj, x, c, fl, e, nk, s, fk, o);                                       \
  }

#define FA_SIMPLE_2D(P, T, CT)                                                                 \
  int P##2d1many(int nt, int64_t nj, const T *x, const T *y, const CT *c, int fl, T e,         \
                 int64_t ms, int64_t mt, CT *fk, const cufinufft_opts *o) {                    \
    return simple<T>(2, 1, nt, nj, {x, y, nullptr}, nc<T>(c), fl, e, {ms, mt, 1}, 0, {}, fk,   \
                     o);                                                                       \
  }                                                                                            \
  int P##2d1(int64_t nj, const T *x, const T *y, const CT *c, int fl, T e, int64_t ms,         \
             int64_t mt, CT *fk, const cufinufft_opts *o) {                                    \
    return P##2d1many(1, nj, x, y, c, fl, e, ms, mt, fk, o);                                   \
  }                                                                                            \
  int P##2d2many(int nt, int64_t nj, const T *x, const T *y, CT *c, int fl, T e, int64_t ms,   \
                 int64_t mt, const CT *fk, const cufinufft_opts *o) {                          \
    return simple<T>(2, 2, nt, nj, {x, y, nullptr}, c, fl, e, {ms, mt, 1}, 0, {}, nc<T>(fk),   \
                     o);                                                                       \
  }                                                                                            \
  int P##2d2(int64_t nj, const T *x, const T *y, CT *c, int fl, T e, int64_t ms, int64_t mt,   \
             const CT *fk, const cufinufft_opts *o) {                                          \
    return P##2d2many(1, nj, x, y, c, fl, e, ms, mt, fk, o);                                   \
  }                                                                                            \
  int P##2d3many(int nt, int64_t nj, const T *x, const T *y, const CT *c, int fl, T e,         \
                 int64_t nk, const T *s, const T *t, CT *fk, const cufinufft_opts *o) {        \
    return simple<T>(2, 3, nt, nj, {x, y, nullptr}, nc<T>(c), fl, e, {0, 0, 0}, nk,            \
                     {s, t, nullptr}, fk, o);                                                  \
  }                                                                                            \
  int P##2d3(int64_t nj, const T *x, const T *y, const CT *c, int fl, T e, int64_t nk,         \
             const T *s, const T *t, CT *fk, const cufinufft_opts *o) {                        \
    return P##2d3many(1, nj, x, y, c, fl, e, nk, s, t, fk, o);                                 \
  }

#define FA_SIMPLE_3D(P, T, CT)                                                                 \
  int P##3d1many(int nt, int64_t nj, const T *x, const T *y, const T *z, const CT *c, int fl,  \
                 T e, int64_t ms, int64_t mt, int64_t mu, CT *fk, const cufinufft_opts *o) {   \
    return simple<T>(3, 1, nt, nj, {x, y, z}, nc<T>(c), fl, e, {ms, mt, mu}, 0, {}, fk, o);    \
  }                                                                                            \
  int P##3d1(int64_t nj, const T *x, const T *y, const T *z, const CT *c, int fl, T e,         \
             int64_t ms, int64_t mt, int64_t mu, CT *fk, const cufinufft_opts *o) {            \
    return P##3d1many(1, nj, x, y, z, c, fl, e, ms, mt, mu, fk, o);                            \
  }                                                                                            \
  int P##3d2many(int nt, int64_t nj, const T *x, const T *y, const T *z, CT *c, int fl, T e,   \
                 int64_t ms, int64_t mt, int64_t mu, const CT *fk, const cufinufft_opts *o) {  \
    return simple<T>(3, 2, nt, nj, {x, y, z}, c, fl, e, {ms, mt, mu}, 0, {}, nc<T>(fk), o);    \
  }                                                                                            \
  int P##3d2(int64_t nj, const T *x, const T *y, const T *z, CT *c, int fl, T e, int64_t ms,   \
             int64_t mt, int64_t mu, const CT *fk, const cufinufft_opts *o) {                  \
    return P##3d2many(1, nj, x, y, z, c, fl, e, ms, mt, mu, fk, o);                            \
  }                                                                                            \
  int P##3d3many(int nt, int64_t nj, const T *x, const T *y, const T *z, const CT *c, int fl,  \
                 T e, int64_t nk, const T *s, const T *t, const T *u, CT *fk,                  \
                 const cufinufft_opts *o) {                                                    \
    return simple<T>(3, 3, nt, nj, {x, y, z}, nc<T>(c), fl, e, {0, 0, 0}, nk, {s, t, u}, fk,   \
                     o);                                                                       \
  }                                                                                            \
  int P##3d3(int64_t nj, const T *x, const T *y, const T *z, const CT *c, int fl, T e,         \
             int64_t nk, const T *s, const T *t, const T *u, CT *fk,                           \
             const cufinufft_opts *o) {                                                        \
    return P##3d3many(1, nj, x, y, z, c, fl, e, nk, s, t, u, fk, o);                           \
  }

FA_SIMPLE_1D(cufinufft, double, hipDoubleComplex)
FA_SIMPLE_1D(cufinufftf, float, hipFloatComplex)
FA_SIMPLE_2D(cufinufft, double, hipDoubleComplex)
FA_SIMPLE_2D(cufinufftf, float, hipFloatComplex)
FA_SIMPLE_3D(cufinufft, double, hipDoubleComplex)
FA_SIMPLE_3D(cufinufftf, float, hipFloatComplex)

}  // extern "C"
