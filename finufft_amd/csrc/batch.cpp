// Fused execution of many independent NUFFT problems (see batch.hpp).
#include "batch.hpp"

#include <algorithm>
#include <map>
#include <string>

namespace fa {

template <typename T>
BatchPlan<T>::BatchPlan(std::vector<Plan<T> *> plans, size_t arena_bytes, hipStream_t stream)
    : plans_(std::move(plans)), stream_(stream) {
  if (plans_.empty()) fail(FINUFFT_ERR_INVALID_ARGUMENT, "batch: no plans");
  for (const Plan<T> *p : plans_)
    if (p == nullptr) fail(FINUFFT_ERR_PLAN_NOTVALID, "batch: null plan");
  const Plan<T> &p0 = *plans_[0];
  dim_ = p0.dim_;
  ntr_ = p0.ntransf_;
  iflag_ = p0.iflag_;
  ns_ = p0.ns_;
  for (const Plan<T> *p : plans_) {
    if (p->type_ != 2 || p->spread_only_ || p->opts_.gpu_method == 1)
      fail(FINUFFT_ERR_INVALID_ARGUMENT,
           "batch: every plan must be a full type-2 transform with gpu_method 2");
    if (p->dim_ != dim_ || p->ntransf_ != ntr_ || p->iflag_ != iflag_ || p->ns_ != ns_)
      fail(FINUFFT_ERR_INVALID_ARGUMENT,
           "batch: plans must share dimension, ntransf, isign and kernel width");
    if (p->opts_.gpu_device_id != p0.opts_.gpu_device_id)
      fail(FINUFFT_ERR_INVALID_ARGUMENT, "batch: plans must share the device");
  }
  device_guard dg(p0.opts_.gpu_device_id);
  const int n = nprob();

  // Subproblem counts stay on the device after setpts; read them once here.
  std::vector<int> nsub(n, 0);
  for (int i = 0; i < n; ++i) {
    const Plan<T> &p = *plans_[i];
    hip_check(hipMemcpyAsync(&nsub[i], p.substart_.get() + p.geom_.total(), sizeof(int),
                             hipMemcpyDeviceToHost, p.stream_),
              "hipMemcpyAsync(batch subproblems)");
    hip_check(hipStreamSynchronize(p.stream_), "hipStreamSynchronize(batch subproblems)");
  }

  // Problems of equal fine-grid shape share one batched FFT.
  std::map<std::array<int, 3>, int> gid;
  std::vector<int> group_of(n), pos_in_group(n);
  for (int i = 0; i < n; ++i) {
    const Plan<T> &p = *plans_[i];
    const std::array<int, 3> key{p.nf_[0], p.nf_[1], p.nf_[2]};
    auto it = gid.find(key);
    if (it == gid.end()) {
      it = gid.emplace(key, int(groups_.size())).first;
      groups_.push_back(std::make_unique<Group>());
      for (int d = 0; d < 3; ++d) groups_.back()->nf[d] = key[d];
    }
    group_of[i] = it->second;
    pos_in_group[i] = int(groups_[it->second]->probs.size());
    groups_[it->second]->probs.push_back(i);
  }

  // Vectors per round: the largest divisor of ntransf whose grids fit.
  int64_t cells = 0;
  for (const Plan<T> *p : plans_) cells += p->nftot_;
  V_ = 1;
  for (int v = ntr_; v >= 1; --v) {
    if (ntr_ % v) continue;
    if (size_t(v) * size_t(cells) * sizeof(C) <= arena_bytes || v == 1) {
      V_ = v;
      break;
    }
  }
  arena_.resize(size_t(V_) * size_t(cells));
  int64_t off = 0;
  for (auto &g : groups_) {
    g->off = off;
    const int64_t nft = int64_t(g->nf[0]) * g->nf[1] * g->nf[2];
    off += nft * V_ * int64_t(g->probs.size());
    g->fft.setup(dim_, g->nf, V_ * int(g->probs.size()), std::is_same<T, double>::value);
    g->fft.prepare(iflag_);
  }

  // Work tables.
  std::vector<AmpItem<T>> amp(n);
  std::vector<InterpItem<T>> itp(n);
  std::vector<C *> fwb(n);
  int64_t blocks = 0, units = 0;
  for (int i = 0; i < n; ++i) {
    const Plan<T> &p = *plans_[i];
    AmpItem<T> &a = amp[i];
    a.dim = dim_;
    a.modeord = p.opts_.modeord;
    for (int d = 0; d < 3; ++d) {
      a.ms[d] = p.ms_[d];
      a.nf[d] = p.nf_[d];
      a.ker[d] = p.fwker_[d < dim_ ? d : 0].get();
    }
    a.nmodes = p.nmodes_;
    a.nftot = p.nftot_;
    a.chunks = int((p.nftot_ + AMP_CHUNK - 1) / AMP_CHUNK);
    a.blk_begin = blocks;
    blocks += int64_t(a.chunks) * V_;

    InterpItem<T> &it = itp[i];
    it.ctx = p.spread_ctx();
    it.ctx.interp_sub = 1;
    it.nftot = p.nftot_;
    it.nsub = nsub[i];
    it.unit_begin = units;
    units += int64_t(nsub[i]) * V_;
    lds_bytes_ = std::max(lds_bytes_, p.lds_bytes_);

    const Group &g = *groups_[group_of[i]];
    fwb[i] = arena_.get() + g.off + int64_t(pos_in_group[i]) * V_ * p.nftot_;
  }
  amp_blocks_ = blocks;
  interp_units_ = units;
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
  // Raise the interpolation kernel's LDS limit now, outside any capture.
  interp_batch<T>(interp_items_.get(), n, dim_, ns_, fw_base_.get(), c_base_.get(), 0, 0,
                  lds_bytes_, stream_);
  hip_check(hipStreamCreateWithFlags(&own_, hipStreamNonBlocking), "hipStreamCreate(batch)");
  hip_check(hipEventCreateWithFlags(&ev_in_, hipEventDisableTiming), "hipEventCreate(batch)");
  hip_check(hipEventCreateWithFlags(&ev_out_, hipEventDisableTiming), "hipEventCreate(batch)");
}

template <typename T>
BatchPlan<T>::~BatchPlan() {
  if (own_) (void)hipStreamSynchronize(own_);
  if (gexec_) (void)hipGraphExecDestroy(gexec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  if (ev_in_) (void)hipEventDestroy(ev_in_);
  if (ev_out_) (void)hipEventDestroy(ev_out_);
  if (own_) (void)hipStreamDestroy(own_);
}

template <typename T>
void BatchPlan<T>::enqueue(hipStream_t s) {
  const int n = nprob();
  for (int v0 = 0; v0 < ntr_; v0 += V_) {
    amplify_batch<T>(amp_items_.get(), fk_base_.get(), fw_base_.get(), n, V_, v0, amp_blocks_,
                     s);
    for (auto &g : groups_) g->fft.exec(arena_.get() + g->off, iflag_, s);
    interp_batch<T>(interp_items_.get(), n, dim_, ns_, fw_base_.get(), c_base_.get(), v0,
                    interp_units_, lds_bytes_, s);
  }
}

template <typename T>
void BatchPlan<T>::execute(C *const *c, C *const *fk) {
  device_guard dg(plans_[0]->opts_.gpu_device_id);
  const int n = nprob();
  for (int i = 0; i < n; ++i) {
    if (c[i] == nullptr || fk[i] == nullptr)
      fail(FINUFFT_ERR_INVALID_ARGUMENT, "batch: null data pointer");
    c_host_[i] = c[i];
    fk_host_[i] = fk[i];
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
