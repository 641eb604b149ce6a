// Fused execution of many independent NUFFT problems.
//
// The reference runs one plan per problem, with one launch per stage and
// vector batch (src/cuda/execute.cu:105-204); for thousands of small, uneven
// problems the GPU then spends much of its time between launches. A BatchPlan
// takes a set of type-2 plans whose points are already set, groups problems of
// equal fine-grid shape, and cuts each group into chunks whose grids (for a
// round of V vectors) fit in a cache-sized budget. Every chunk runs
//   * one amplification launch over its problems (fk -> fine grid),
//   * one batched rocFFT call (its problems share the shape),
//   * one interpolation launch over its problems' subproblems,
// so that its grids stay in the GPU's last-level cache from stage to stage.
// The whole schedule is recorded once into a hipGraph and replayed on every
// execute (on a stream of the batch's own, ordered with the caller's stream by
// events). The per-problem input/output pointers live in a small device table
// that is refreshed before each replay, so the graph never changes.
#pragma once

#include <memory>
#include <vector>

#include "fft.hpp"
#include "fused_fft.hpp"
#include "plan.hpp"
#include "plan_kernels.hpp"

namespace fa {

template <typename T>
class BatchPlan {
 public:
  using C = cplx_t<T>;

  // plans: type-2 plans with points set, of equal dimension, ntransf, iflag
  // and kernel. chunk_bytes bounds the fine-grid memory of one chunk (the
  // vectors of a problem are processed in rounds of V, V dividing ntransf,
  // and a chunk never has fewer than one problem).
  BatchPlan(std::vector<Plan<T> *> plans, size_t chunk_bytes, hipStream_t stream);
  ~BatchPlan();
  BatchPlan(const BatchPlan &) = delete;
  BatchPlan &operator=(const BatchPlan &) = delete;

  // c[p]: values of problem p (ntransf x M_p), fk[p]: its modes
  // (ntransf x N_p); device pointers, host arrays of length nprob.
  void execute(C *const *c, C *const *fk);

  int nprob() const { return int(plans_.size()); }
  int vectors_per_round() const { return V_; }
  int groups() const { return int(groups_.size()); }
  int chunks() const { return int(chunks_.size()); }
  // 1 when the grids are made by the fused column/row transforms (2-D grids
  // whose sides fit one LDS buffer), 0 when by amplification + rocFFT.
  bool fused_fft() const { return fused_; }
  bool graph() const { return gexec_ != nullptr; }
  // Why the schedule could not be recorded (0: it was, or not tried yet):
  // capture stage (1 begin, 2 a call inside, 3 end, 4 instantiate) and the
  // HIP error code of that stage.
  int capture_stage() const { return capture_stage_; }
  int capture_error() const { return capture_error_; }
  void set_use_graph(bool on) { use_graph_ = on; }

 private:
  struct Group {
    int nf[3];
    std::vector<int> probs;
    int per_chunk = 1;  // problems per full chunk
    FftPlan fft_full, fft_rest;  // unused on the fused 2-D path
    dbuf<T> tw[2], rker[2];      // twiddles and reciprocal kernel series (fused 2-D path)
    int cw = 1, lcw = 0, rw = 1, lrw = 0;  // fused 2-D path: columns / rows per workgroup
    size_t lds_col = 0, lds_row = 0;
  };
  struct Chunk {
    int first, count;  // slots [first, first + count)
    int group;
    bool full;
    int64_t amp_blocks, units, col_blocks, row_blocks;
    int64_t cells;  // grid cells of the chunk's round (type 1 clears them)
  };
  void enqueue(hipStream_t s);

  std::vector<Plan<T> *> plans_;
  std::vector<uint64_t> gens_;  // each plan's setpts generation when the batch was made
  hipStream_t stream_ = nullptr;  // the plans' (caller's) stream
  // The schedule runs on a stream of its own (the legacy null stream cannot
  // be captured), ordered after and before the caller's stream by events.
  hipStream_t own_ = nullptr;
  hipEvent_t ev_in_ = nullptr, ev_out_ = nullptr;
  int type_ = 2, dim_ = 0, ntr_ = 1, iflag_ = 1, ns_ = 2, V_ = 1;
  size_t lds_bytes_ = 0;
  KerParams<T> kp_{};
  dbuf<C> arena_;
  std::vector<std::unique_ptr<Group>> groups_;
  std::vector<Chunk> chunks_;
  std::vector<int> slot_prob_;  // problem index of each slot (group-major order)
  bool fused_ = false;
  dbuf<AmpItem<T>> amp_items_;
  dbuf<FusedFftItem<T>> fft_items_;
  dbuf<InterpItem<T>> interp_items_;
  dbuf<C *> fw_base_;
  dbuf<const C *> fk_base_;
  dbuf<C *> c_base_;
  std::vector<const C *> fk_host_;
  std::vector<C *> c_host_;
  bool use_graph_ = true;
  bool graph_failed_ = false;
  int capture_stage_ = 0, capture_error_ = 0;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t gexec_ = nullptr;
};

}  // namespace fa
