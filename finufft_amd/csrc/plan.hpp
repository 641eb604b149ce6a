// The NUFFT plan: kernel and grid choice at construction, point sort at
// setpts, spread/interp -> FFT -> deconvolve at execute.
//
// Mirrors the life cycle of the reference's cufinufft_plan_t
// (include/cufinufft/cufinufft_plan_t.hpp:205-499; src/cuda/makeplan.cu:223-398,
// setpts.cu:77-314, execute.cu:105-271) with these differences of form:
//   * the bin sort is deterministic (stable radix sort, sort.hpp);
//   * setpts for types 1/2 synchronizes with the host only for 3-D fp32
//     spreading plans (to size the dense-bin launch exactly); otherwise the
//     subproblem table is sized by an upper bound and its true length stays
//     on the GPU;
//   * the fine-grid work buffer is allocated once at plan time, not at every
//     execute, and grid clearing uses a kernel rather than a memset node, so
//     execute is capturable in a hipGraph;
//   * all ntransf vectors of a batch go through one launch per stage.
#pragma once

#include <memory>

#include <cufinufft_opts.h>

#include "common.hpp"
#include "fft.hpp"
#include "fft3.hpp"
#include "host_math.hpp"
#include "sort.hpp"
#include "spreadinterp.hpp"

namespace fa {

// Grid <-> modes transforms of plans made from now on: 0 (the default) the
// pruned line passes where the grid allows them, 1 rocFFT with a separate
// amplification / deconvolution pass.
int fft_path_setting();
void set_fft_path_setting(int path);

template <typename T>
class BatchPlan;

template <typename T>
class Plan {
  friend class BatchPlan<T>;

 public:
  using C = cplx_t<T>;

  Plan(int type, int dim, const int *nmodes, int iflag, int ntransf, T tol,
       const cufinufft_opts &opts);

  void setpts(int M, const T *x, const T *y, const T *z, int N, const T *s, const T *t,
              const T *u);
  // The plan's transform (type 1, 2 or 3).
  void execute(C *c, C *fk);
  // Adjoint of the plan's transform on the same points: type 1 -> type 2 with
  // the opposite sign and vice versa (the reference CPU library's
  // execute_adjoint, include/finufft/execute.hpp:318-357). Types 1 and 2 only.
  void execute_adjoint(C *c, C *fk);
  // Type-2 direction of the plan with the derivatives of the result with
  // respect to every point coordinate: c (may be null) and dc[dim][ntransf][M].
  void interp_grad(C *c, C *dc, C *fk);

  // introspection (tests, bench)
  int type() const { return type_; }
  int dim() const { return dim_; }
  int ns() const { return ns_; }
  double beta() const { return beta_; }
  int kernel_formula() const { return formula_; }
  double upsampfac() const { return opts_.upsampfac; }
  int method() const { return opts_.gpu_method; }
  int od_np() const { return od_np_; }
  int nf(int d) const { return nf_[d]; }
  int binsize(int d) const { return geom_.binsize[d]; }
  int batchsize() const { return batchsize_; }
  int M() const { return M_; }
  const int *sort_perm() const { return idx_.get(); }
  const int *bin_start() const { return binstart_.get(); }
  int nbins_total() const { return geom_.total(); }
  int col_side() const { return geom_.colS; }
  int colored() const { return colored_ ? 1 : 0; }
  int dense_subproblems() const { return n_dense_sub_; }
  // 1 when the grid <-> modes transforms run as pruned line passes (3-D
  // power-of-two grids), 0 when rocFFT with separate (de)convolution.
  int fft_path() const { return fft3_ ? 1 : 0; }
  hipStream_t stream() const { return stream_; }
  StageTimer &timer() { return timer_; }
  // Number of setpts calls so far (a batch records it and refuses to run
  // when the points have changed since), and the number of live batches
  // holding the plan (destroy is refused while it is not zero).
  uint64_t setpts_generation() const { return setpts_gen_; }
  int batch_refs() const { return batch_refs_; }
  void add_batch_ref(int d) { batch_refs_ += d; }

 private:
  void setup_kernel();
  void upload_coefficients();
  void setup_bins();
  void sort_points(const T *x, const T *y, const T *z);
  SpreadCtx<T> spread_ctx() const;
  void check_data(const C *c, const C *fk) const;
  void exec_type1(C *c, C *fk, int fft_sign, int ntr);
  void exec_type2(C *c, C *fk, int fft_sign, int ntr);
  void exec_type3(C *c, C *fk);
  void prepare_grid();
  // fw (blk vectors) -> fk with the deconvolution (type-1 direction; fw is
  // overwritten), and fk -> fw with the amplification (type-2 direction).
  void grid_to_modes(C *fw, C *fk, int fft_sign, int blk);
  void modes_to_grid(C *fk, C *fw, int fft_sign, int blk);

  // parameters
  int type_, dim_, ntransf_, iflag_;
  T tol_;
  cufinufft_opts opts_;
  int batchsize_ = 1;
  hipStream_t stream_ = nullptr;
  bool spread_only_ = false;
  bool auto_method_ = false;
  int user_bin_[3] = {0, 0, 0};  // the caller's gpu_binsize{x,y,z}

  // kernel
  int ns_ = 2;
  double beta_ = 0;
  int formula_ = 0;  // KernelFormula of the spreading kernel
  dbuf<T> coef_;

  // sizes
  int ms_[3] = {1, 1, 1};
  int nf_[3] = {1, 1, 1};
  int64_t nftot_ = 1;
  int64_t nmodes_ = 1;
  int M_ = 0;
  int Nk_ = 0;  // type-3 targets
  const T *xyz_[3] = {nullptr, nullptr, nullptr};

  // kernel Fourier series (types 1/2)
  dbuf<T> fwker_[3];

  // sort / subproblems
  BinGeom geom_;
  dbuf<int> idx_, binstart_, substart_, sub2bin_, colstart_;
  dbuf<T> sxyz_[3];  // coordinates in sorted order
  bool colored_ = false;
  int sub_capacity_ = 0;
  int sub_size_ = 1024;  // points per subproblem
  int dense_min_ = 0;    // column spreading: bins above this count use subproblems
  int n_dense_sub_ = 0;  // column spreading: number of those subproblems
  SortScratch scratch_;
  size_t lds_bytes_ = 0;   // interpolation tile
  size_t lds_spread_ = 0;  // spreading tile (fp64) plus staging
  bool spread_gm_ = false;  // spreading tile too large: point-driven spreading
  int user_np_ = 0;         // the caller's gpu_np (method 3)
  int od_np_ = 0;           // output-driven spreading: points staged per batch
  size_t lds_od_ = 0;       // its LDS bytes (fp64 tile + staging)

  // fine grid work buffer and FFT
  dbuf<C> fw_;
  FftPlan fft_;
  std::unique_ptr<Fft3<T>> fft3_;  // pruned line transforms, when supported
  dbuf<C> grad_scratch_;  // interp_grad when ntransf exceeds the batch size

  // type 3
  dbuf<T> xyzp_[3], stup_[3];
  dbuf<C> prephase_, deconv_, cpre_;
  double t3C_[3] = {0, 0, 0}, t3D_[3] = {0, 0, 0};
  std::unique_ptr<Plan<T>> inner_;
  bool t3_set_ = false;  // type 3: setpts has completed

  StageTimer timer_;
  uint64_t setpts_gen_ = 0;
  int batch_refs_ = 0;
};

}  // namespace fa
