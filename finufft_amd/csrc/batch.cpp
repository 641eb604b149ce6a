// Fused execution of many independent NUFFT problems (see batch.hpp).
#include "batch.hpp"

#include <algorithm>
#include <array>
#include <map>
#include <string>

namespace fa {

template <typename T>
BatchPlan<T>::BatchPlan(std::vector<Plan<T> *> plans, size_t chunk_bytes, hipStream_t stream)
    : plans_(std::move(plans)), stream_(stream) {
  if (plans_.empty()) fail(FINUFFT_ERR_INVALID_ARGUMENT, "batch: no plans");
  for (const Plan<T> *p : plans_)
    if (p == nullptr) fail(FINUFFT_ERR_PLAN_NOTVALID, "batch: null plan");
  const Plan<T> &p0 = *plans_[0];
  type_ = p0.type_;
  dim_ = p0.dim_;
  ntr_ = p0.ntransf_;
  iflag_ = p0.iflag_;
  ns_ = p0.ns_;
  for (const Plan<T> *p : plans_) {
    if ((p->type_ != 1 && p->type_ != 2) || p->type_ != type_ || p->spread_only_ ||
        p->opts_.gpu_method == 1)
      fail(FINUFFT_ERR_INVALID_ARGUMENT,
           "batch: the plans must be full transforms of one type (1 or 2) with gpu_method 2");
    // type 1: the subproblem table of the plan must cover every bin (the
    // column-window spreader of 3-D fp32 plans keeps only its dense bins)
    if (p->type_ == 1 && (p->geom_.colS != 0 || p->opts_.gpu_method == 3))
      fail(FINUFFT_ERR_INVALID_ARGUMENT,
           "batch: type-1 plans of the column-window or output-driven spreaders are not "
           "batched");
    if (p->dim_ != dim_ || p->ntransf_ != ntr_ || p->iflag_ != iflag_ || p->ns_ != ns_ ||
        p->opts_.upsampfac != p0.opts_.upsampfac ||
        p->opts_.gpu_kerevalmeth != p0.opts_.gpu_kerevalmeth)
      fail(FINUFFT_ERR_INVALID_ARGUMENT,
           "batch: plans must share dimension, ntransf, isign and kernel (width, upsampfac, "
           "evaluation method)");
    if (p->opts_.gpu_device_id != p0.opts_.gpu_device_id)
      fail(FINUFFT_ERR_INVALID_ARGUMENT, "batch: plans must share the device");
    // setpts builds the subproblem table that the batch reads below
    if (p->substart_.get() == nullptr)
      fail(FINUFFT_ERR_INVALID_ARGUMENT, "batch: every plan must have its points set");
  }
  device_guard dg(p0.opts_.gpu_device_id);
  const int n = nprob();
  kp_ = p0.spread_ctx().kp;  // identical for all plans (same width and upsampfac)

  // Subproblem counts stay on the device after setpts; read them once here.
  std::vector<int> nsub(n, 0);
  for (int i = 0; i < n; ++i) {
    const Plan<T> &p = *plans_[i];
    hip_check(hipMemcpyAsync(&nsub[i], p.substart_.get() + p.geom_.total(), sizeof(int),
                             hipMemcpyDeviceToHost, p.stream_),
              "hipMemcpyAsync(batch subproblems)");
    hip_check(hipStreamSynchronize(p.stream_), "hipStreamSynchronize(batch subproblems)");
  }

  // Problems of equal fine-grid shape form a group (one FFT shape).
  std::map<std::array<int, 3>, int> gid;
  for (int i = 0; i < n; ++i) {
    const Plan<T> &p = *plans_[i];
    const std::array<int, 3> key{p.nf_[0], p.nf_[1], p.nf_[2]};
    auto it = gid.find(key);
    if (it == gid.end()) {
      it = gid.emplace(key, int(groups_.size())).first;
      groups_.push_back(std::make_unique<Group>());
      for (int d = 0; d < 3; ++d) groups_.back()->nf[d] = key[d];
    }
    groups_[it->second]->probs.push_back(i);
  }

  // Vectors per round: the largest divisor of ntransf for which the largest
  // problem's grids fit the chunk budget.
  int64_t max_cells = 0;
  for (const Plan<T> *p : plans_) max_cells = std::max<int64_t>(max_cells, p->nftot_);
  V_ = 1;
  for (int v = ntr_; v >= 1; --v) {
    if (ntr_ % v) continue;
    if (size_t(v) * size_t(max_cells) * sizeof(C) <= chunk_bytes || v == 1) {
      V_ = v;
      break;
    }
  }

  // 2-D grids whose sides fit one LDS buffer are made by the fused column /
  // row transforms; otherwise by amplification and rocFFT.
  fused_ = dim_ == 2 && type_ == 2;
  for (size_t gi = 0; gi < groups_.size() && fused_; ++gi) {
    Group &g = *groups_[gi];
    fused_ = fused_fft_shape<T>(g.nf, g.cw, g.lcw, g.rw, g.lrw, g.lds_col, g.lds_row);
  }

  // Chunks: consecutive problems of one group, as many as fit the budget.
  const bool fp64 = std::is_same<T, double>::value;
  int64_t arena_cells = 0;
  for (int gi = 0; gi < int(groups_.size()); ++gi) {
    Group &g = *groups_[gi];
    const int64_t nft = int64_t(g.nf[0]) * g.nf[1] * g.nf[2];
    const int64_t per = int64_t(V_) * nft * int64_t(sizeof(C));
    g.per_chunk = int(std::max<int64_t>(1, std::min<int64_t>(int64_t(g.probs.size()),
                                                              int64_t(chunk_bytes) / per)));
    const int K = int(g.probs.size());
    const int rest = K % g.per_chunk;
    if (fused_) {
      const Plan<T> &p = *plans_[g.probs[0]];  // the group shares nf and kernel
      for (int d = 0; d < 2; ++d) {
        std::vector<T> tw;
        fft_twiddles<T>(g.nf[d], tw);
        g.tw[d].resize(tw.size());
        hip_check(hipMemcpy(g.tw[d].get(), tw.data(), tw.size() * sizeof(T),
                            hipMemcpyHostToDevice),
                  "hipMemcpy(batch twiddles)");
        std::vector<T> k(size_t(g.nf[d] / 2 + 1));
        hip_check(hipMemcpy(k.data(), p.fwker_[d].get(), k.size() * sizeof(T),
                            hipMemcpyDeviceToHost),
                  "hipMemcpy(batch kernel series)");
        for (T &v : k) v = T(1.0 / double(v));
        g.rker[d].resize(k.size());
        hip_check(hipMemcpy(g.rker[d].get(), k.data(), k.size() * sizeof(T),
                            hipMemcpyHostToDevice),
                  "hipMemcpy(batch kernel series)");
      }
    } else {
      g.fft_full.setup(dim_, g.nf, V_ * g.per_chunk, fp64);
      g.fft_full.prepare(iflag_);
      if (rest) {
        g.fft_rest.setup(dim_, g.nf, V_ * rest, fp64);
        g.fft_rest.prepare(iflag_);
      }
    }
    for (int k = 0; k < K; k += g.per_chunk) {
      Chunk c;
      c.first = int(slot_prob_.size()) + k;
      c.count = std::min(g.per_chunk, K - k);
      c.group = gi;
      c.full = c.count == g.per_chunk;
      c.amp_blocks = c.units = c.col_blocks = c.row_blocks = 0;
      c.cells = int64_t(c.count) * V_ * nft;
      chunks_.push_back(c);
      arena_cells = std::max<int64_t>(arena_cells, int64_t(c.count) * V_ * nft);
    }
    for (int i : g.probs) slot_prob_.push_back(i);
  }
  arena_.resize(size_t(arena_cells));

  // Work tables in slot order; prefix counts restart at every chunk.
  std::vector<AmpItem<T>> amp(n);
  std::vector<InterpItem<T>> itp(n);
  std::vector<FusedFftItem<T>> ffi(fused_ ? n : 0);
  std::vector<C *> fwb(n);
  for (Chunk &c : chunks_) {
    int64_t blocks = 0, units = 0, cblocks = 0, rblocks = 0;
    for (int s = c.first; s < c.first + c.count; ++s) {
      const int i = slot_prob_[s];
      const Plan<T> &p = *plans_[i];
      AmpItem<T> &a = amp[s];
      a.dim = dim_;
      a.modeord = p.opts_.modeord;
      for (int d = 0; d < 3; ++d) {
        a.ms[d] = p.ms_[d];
        a.nf[d] = p.nf_[d];
        a.rnf[d] = 1.0 / double(p.nf_[d]);
        a.ker[d] = p.fwker_[d < dim_ ? d : 0].get();
      }
      a.nmodes = p.nmodes_;
      a.nftot = p.nftot_;
      a.chunks = int((p.nftot_ + AMP_CHUNK - 1) / AMP_CHUNK);
      a.blk_begin = blocks;
      blocks += int64_t(a.chunks) * V_;

      InterpItem<T> &it = itp[s];
      it.ctx = p.spread_ctx();
      it.ctx.interp_sub = 1;
      it.nftot = p.nftot_;
      it.nsub = nsub[i];
      it.unit_begin = units;
      units += nsub[i];
      lds_bytes_ = std::max(lds_bytes_, type_ == 1 ? p.lds_spread_ : p.lds_bytes_);

      if (fused_) {
        FusedFftItem<T> &f = ffi[s];
        const Group &g = *groups_[c.group];
        f.modeord = p.opts_.modeord;
        for (int d = 0; d < 2; ++d) {
          f.ms[d] = p.ms_[d];
          f.nf[d] = p.nf_[d];
          f.rker[d] = g.rker[d].get();
          f.tw[d] = g.tw[d].get();
          radix_plan(p.nf_[d], f.rp[d]);
        }
        f.cw = g.cw;
        f.lcw = g.lcw;
        f.rw = g.rw;
        f.lrw = g.lrw;
        f.ncol = (f.ms[0] + f.cw - 1) / f.cw;
        f.nrow = (f.nf[1] + f.rw - 1) / f.rw;
        f.nmodes = p.nmodes_;
        f.nftot = p.nftot_;
        f.col_begin = cblocks;
        f.row_begin = rblocks;
        cblocks += int64_t(f.ncol) * V_;
        rblocks += int64_t(f.nrow) * V_;
      }

      fwb[s] = arena_.get() + int64_t(s - c.first) * V_ * p.nftot_;
    }
    c.amp_blocks = blocks;
    c.units = units;
    c.col_blocks = cblocks;
    c.row_blocks = rblocks;
  }
  if (fused_) {
    fused_fft_prepare<T>();
    fft_items_.resize(n);
    hip_check(hipMemcpy(fft_items_.get(), ffi.data(), n * sizeof(FusedFftItem<T>),
                        hipMemcpyHostToDevice),
              "hipMemcpy(batch transform items)");
  }
  amp_items_.resize(n);
  interp_items_.resize(n);
  fw_base_.resize(n);
  fk_base_.resize(n);
  c_base_.resize(n);
  hip_check(hipMemcpy(amp_items_.get(), amp.data(), n * sizeof(AmpItem<T>), hipMemcpyHostToDevice),
            "hipMemcpy(batch amp items)");
  hip_check(hipMemcpy(interp_items_.get(), itp.data(), n * sizeof(InterpItem<T>),
                      hipMemcpyHostToDevice),
            "hipMemcpy(batch interp items)");
  hip_check(hipMemcpy(fw_base_.get(), fwb.data(), n * sizeof(C *), hipMemcpyHostToDevice),
            "hipMemcpy(batch grids)");
  fk_host_.resize(n);
  c_host_.resize(n);
  // Raise the spreading / interpolation kernel's LDS limit now, outside any
  // capture.
  if (type_ == 1)
    spread_batch<T>(interp_items_.get(), n, dim_, ns_, c_base_.get(), 0, fw_base_.get(), V_, 0,
                    lds_bytes_, stream_);
  else
    interp_batch<T>(interp_items_.get(), n, dim_, ns_, kp_, fw_base_.get(), c_base_.get(), 0, V_,
                    0, lds_bytes_, stream_);
  hip_check(hipStreamCreateWithFlags(&own_, hipStreamNonBlocking), "hipStreamCreate(batch)");
  hip_check(hipEventCreateWithFlags(&ev_in_, hipEventDisableTiming), "hipEventCreate(batch)");
  hip_check(hipEventCreateWithFlags(&ev_out_, hipEventDisableTiming), "hipEventCreate(batch)");
  // The batch copied every plan's sort tables and subproblem counts: it is
  // valid for the points set so far, and the plans must outlive it.
  for (Plan<T> *p : plans_) {
    gens_.push_back(p->setpts_generation());
    p->add_batch_ref(1);
  }
}

template <typename T>
BatchPlan<T>::~BatchPlan() {
  for (size_t i = 0; i < gens_.size(); ++i) plans_[i]->add_batch_ref(-1);
  if (own_) (void)hipStreamSynchronize(own_);
  if (gexec_) (void)hipGraphExecDestroy(gexec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  if (ev_in_) (void)hipEventDestroy(ev_in_);
  if (ev_out_) (void)hipEventDestroy(ev_out_);
  if (own_) (void)hipStreamDestroy(own_);
}

template <typename T>
void BatchPlan<T>::enqueue(hipStream_t s) {
  for (int v0 = 0; v0 < ntr_; v0 += V_) {
    for (const Chunk &c : chunks_) {
      Group &g = *groups_[c.group];
      if (type_ == 1) {
        // clear the round's grids, spread every problem's points into them,
        // transform, and read the modes out with the deconvolution
        zero<T>(arena_.get(), c.cells, s);
        spread_batch<T>(interp_items_.get() + c.first, c.count, dim_, ns_,
                        c_base_.get() + c.first, v0, fw_base_.get() + c.first, V_, c.units,
                        lds_bytes_, s);
        (c.full ? g.fft_full : g.fft_rest).exec(arena_.get(), iflag_, s);
        // type 1 writes the modes: the table of fk pointers is an output here
        deconv_batch<T>(amp_items_.get() + c.first, fw_base_.get() + c.first,
                        const_cast<C *const *>(fk_base_.get()) + c.first, c.count, dim_, v0,
                        c.amp_blocks, s);
        continue;
      }
      if (fused_) {
        fused_fft_cols<T>(fft_items_.get() + c.first, fk_base_.get() + c.first,
                          fw_base_.get() + c.first, c.count, v0, iflag_, c.col_blocks, g.lds_col,
                          s);
        fused_fft_rows<T>(fft_items_.get() + c.first, fw_base_.get() + c.first, c.count, iflag_,
                          c.row_blocks, g.lds_row, s);
      } else {
        amplify_batch<T>(amp_items_.get() + c.first, fk_base_.get() + c.first,
                         fw_base_.get() + c.first, c.count, dim_, v0, c.amp_blocks, s);
        (c.full ? g.fft_full : g.fft_rest).exec(arena_.get(), iflag_, s);
      }
      interp_batch<T>(interp_items_.get() + c.first, c.count, dim_, ns_, kp_,
                      fw_base_.get() + c.first, c_base_.get() + c.first, v0, V_, c.units,
                      lds_bytes_, s);
    }
  }
}

template <typename T>
void BatchPlan<T>::execute(C *const *c, C *const *fk) {
  device_guard dg(plans_[0]->opts_.gpu_device_id);
  const int n = nprob();
  for (int i = 0; i < n; ++i)
    if (plans_[i]->setpts_generation() != gens_[i])
      fail(FINUFFT_ERR_INVALID_ARGUMENT,
           "batch: the points of plan " + std::to_string(i) +
               " were set again after the batch was made; make the batch again");
  for (int i = 0; i < n; ++i)
    if (c[i] == nullptr || fk[i] == nullptr)
      fail(FINUFFT_ERR_INVALID_ARGUMENT, "batch: null data pointer");
  for (int s = 0; s < n; ++s) {
    c_host_[s] = c[slot_prob_[s]];
    fk_host_[s] = fk[slot_prob_[s]];
  }
  // The previous execute may still read the pointer tables: wait for it, then
  // copy them synchronously (the host vectors may be rewritten at once).
  hip_check(hipEventSynchronize(ev_out_), "hipEventSynchronize(batch)");
  hip_check(hipMemcpy(fk_base_.get(), fk_host_.data(), n * sizeof(C *), hipMemcpyHostToDevice),
            "hipMemcpy(batch inputs)");
  hip_check(hipMemcpy(c_base_.get(), c_host_.data(), n * sizeof(C *), hipMemcpyHostToDevice),
            "hipMemcpy(batch outputs)");
  if (use_graph_ && !graph_failed_ && gexec_ == nullptr) {
    // Record the whole schedule once; a failed capture falls back to direct
    // launches for the life of the batch (the reason is kept for info()).
    auto failed = [&](int stage, hipError_t e) {
      if (capture_stage_ == 0) {
        capture_stage_ = stage;
        capture_error_ = int(e);
      }
    };
    hipError_t e = hipStreamBeginCapture(own_, hipStreamCaptureModeThreadLocal);
    bool ok = e == hipSuccess;
    if (!ok) failed(1, e);
    if (ok) {
      try {
        enqueue(own_);
      } catch (const nufft_error &) {
        ok = false;
        failed(2, hipGetLastError());
      }
      hipGraph_t g = nullptr;
      e = hipStreamEndCapture(own_, &g);
      if (e != hipSuccess || g == nullptr) {
        if (ok) failed(3, e);
        ok = false;
      }
      if (ok) {
        e = hipGraphInstantiate(&gexec_, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
          gexec_ = nullptr;
          ok = false;
          failed(4, e);
        }
      }
      if (ok) {
        graph_ = g;
      } else if (g) {
        (void)hipGraphDestroy(g);
      }
    }
    if (!ok) {
      (void)hipGetLastError();
      graph_failed_ = true;
    }
  }
  hip_check(hipEventRecord(ev_in_, stream_), "hipEventRecord(batch in)");
  hip_check(hipStreamWaitEvent(own_, ev_in_, 0), "hipStreamWaitEvent(batch in)");
  if (gexec_ && use_graph_) {
    hip_check(hipGraphLaunch(gexec_, own_), "hipGraphLaunch(batch)");
  } else {
    enqueue(own_);
  }
  hip_check(hipEventRecord(ev_out_, own_), "hipEventRecord(batch out)");
  hip_check(hipStreamWaitEvent(stream_, ev_out_, 0), "hipStreamWaitEvent(batch out)");
}

template class BatchPlan<float>;
template class BatchPlan<double>;

}  // namespace fa
