"""Learned cost model: which plan options run a given problem fastest.

A small network (finufft_amd/csrc/mlp.hip: 8 inputs, two ReLU layers of 32,
one output per option) predicts the log execute time of each option from

    dim, log2 N (modes), log2 M, log2 (M / N), log2 ntransf, -log10 tol,
    precision (0 fp32, 1 fp64), transform type (1 or 2)

for the options (ARMS): LDS subproblems (gpu_method 2, the default), global
memory (gpu_method 1), all vectors of an execute in one pass
(gpu_maxbatchsize 32) and upsampling factor 1.25. Training data
come from sweeps of the library's own options on the GPU (`sweep`): every
option is timed and its error against the FP64 direct sum measured at sampled
outputs; an option whose error exceeds max(tol, 1.5 x the default option's
error) is rejected for that problem class and never chosen there. The network
is trained on the GPU with the library's own forward, gradient and Adam
kernels and runs inference with the forward kernel. Without weights, or for
features outside the training range, the library's own heuristic is used
(no option override).

The reference has no learned component: it picks methods from fixed tables
(src/cuda/heuristics.cu:270-280, src/cuda/makeplan.cu:289-329).

Command line (on a GPU):
    python -m finufft_amd.costmodel train --out finufft_amd/costmodel_weights.json
"""

from __future__ import annotations

import argparse
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

from . import _lib

ARMS = [
    {"upsampfac": 2.0, "gpu_method": 2},
    {"upsampfac": 2.0, "gpu_method": 1},
    {"upsampfac": 2.0, "gpu_method": 2, "gpu_maxbatchsize": 32},
    {"upsampfac": 1.25, "gpu_method": 2},
]
NIN, NHID, NOUT = 8, 32, 4
NPARAM = NHID * NIN + NHID + NHID * NHID + NHID + NOUT * NHID + NOUT
WEIGHTS = Path(__file__).resolve().parent / "costmodel_weights.json"


def features(ttype, n_modes, M, ntransf, tol, dtype):
    n_tot = float(np.prod(n_modes))
    return np.array([
        len(n_modes),
        math.log2(max(n_tot, 1.0)),
        math.log2(max(M, 1)),
        math.log2(max(M, 1) / max(n_tot, 1.0)),
        math.log2(max(ntransf, 1)),
        -math.log10(tol),
        1.0 if dtype in ("complex128", "double") else 0.0,
        float(ttype),
    ], np.float32)


def problem_class(ttype, dim, tol, dtype):
    """Key of the rejection table: (type, dim, precision, tol decade)."""
    return f"{ttype}:{dim}:{dtype}:{int(round(-math.log10(tol)))}"


# ------------------------------------------------------------ HIP network


def _torch():
    import torch

    return torch


def _ptr(t, n=None, name="array"):
    """Device pointer of a contiguous float32 GPU tensor of n elements (the
    kernels index it without knowing its size)."""
    torch = _torch()
    if not (isinstance(t, torch.Tensor) and t.dtype == torch.float32 and t.is_cuda
            and t.is_contiguous()):
        raise TypeError(f"{name} must be a contiguous float32 tensor on the GPU")
    if n is not None and t.numel() != n:
        raise ValueError(f"{name} has {t.numel()} elements, {n} expected")
    return t.data_ptr()


def gpu_forward(params, x):
    """y[n][NOUT] from x[n][NIN] with the library's forward kernel (torch
    float32 tensors on the GPU)."""
    torch = _torch()
    n = int(x.shape[0])
    y = torch.empty((n, NOUT), dtype=torch.float32, device=x.device)
    ier = _lib.lib().finufft_amd_mlp_forward(_ptr(params, NPARAM, "params"), n,
                                             _ptr(x, n * NIN, "x"), _ptr(y), None)
    if ier != 0:
        raise RuntimeError(f"finufft_amd_mlp_forward failed ({ier})")
    return y


def gpu_grad(params, x, t, m):
    """(loss, grad) of the masked squared error with the library's kernel."""
    torch = _torch()
    n = int(x.shape[0])
    g = torch.zeros(NPARAM, dtype=torch.float32, device=x.device)
    loss = torch.zeros(1, dtype=torch.float32, device=x.device)
    ier = _lib.lib().finufft_amd_mlp_grad(_ptr(params, NPARAM, "params"), n,
                                          _ptr(x, n * NIN, "x"), _ptr(t, n * NOUT, "target"),
                                          _ptr(m, n * NOUT, "mask"), _ptr(g), _ptr(loss), None)
    if ier != 0:
        raise RuntimeError(f"finufft_amd_mlp_grad failed ({ier})")
    return loss, g


def init_params(seed=0):
    r = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((NIN, NHID), (NHID, NHID), (NHID, NOUT)):
        parts.append(r.normal(0, math.sqrt(2.0 / fan_in), (fan_out, fan_in)).reshape(-1))
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts).astype(np.float32)


def train(x, t, m, epochs=3000, lr=3e-3, seed=0, log=None):
    """Fits the network on the GPU (full-batch Adam, the library's kernels).
    x, t, m: numpy arrays (already normalized inputs). Returns (params, loss)."""
    torch = _torch()
    dev = torch.device("cuda", 0)
    X = torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)
    T = torch.from_numpy(np.ascontiguousarray(t, np.float32)).to(dev)
    Mk = torch.from_numpy(np.ascontiguousarray(m, np.float32)).to(dev)
    P = torch.from_numpy(init_params(seed)).to(dev)
    m1 = torch.zeros_like(P)
    m2 = torch.zeros_like(P)
    lib = _lib.lib()
    loss = None
    for step in range(1, epochs + 1):
        loss, g = gpu_grad(P, X, T, Mk)
        ier = lib.finufft_amd_mlp_adam(_ptr(P), _ptr(g), _ptr(m1), _ptr(m2), step, float(lr), None)
        if ier != 0:
            raise RuntimeError(f"finufft_amd_mlp_adam failed ({ier})")
        if log and step % 500 == 0:
            log(f"epoch {step}: loss {float(loss):.4g}")
    return P.cpu().numpy(), float(loss)


# ------------------------------------------------------------ the model


class CostModel:
    def __init__(self, params, mean, std, lo, hi, rejected=None, meta=None, arms=None):
        self.arms = [dict(a) for a in (arms or ARMS)]  # the options the network was trained on
        self.params = np.asarray(params, np.float32)
        self.mean = np.asarray(mean, np.float32)
        self.std = np.asarray(std, np.float32)
        self.lo = np.asarray(lo, np.float32)
        self.hi = np.asarray(hi, np.float32)
        self.rejected = rejected or {}
        self.meta = meta or {}
        self._dev_params = None

    @classmethod
    def load(cls, path=WEIGHTS):
        d = json.loads(Path(path).read_text())
        return cls(d["params"], d["mean"], d["std"], d["lo"], d["hi"], d.get("rejected"),
                   d.get("meta"), d.get("arms"))

    def save(self, path):
        Path(path).write_text(json.dumps({
            "params": [float(v) for v in self.params],
            "mean": self.mean.tolist(), "std": self.std.tolist(),
            "lo": self.lo.tolist(), "hi": self.hi.tolist(),
            "rejected": self.rejected, "meta": self.meta,
            "arms": self.arms,
        }, indent=1))

    def predict(self, feats):
        """Predicted log execute times (n x NOUT) for raw feature rows."""
        torch = _torch()
        dev = torch.device("cuda", 0)
        if self._dev_params is None:
            self._dev_params = torch.from_numpy(self.params).to(dev)
        z = (np.atleast_2d(feats) - self.mean) / self.std
        return gpu_forward(self._dev_params, torch.from_numpy(z.astype(np.float32)).to(dev)).cpu(
        ).numpy()

    def in_range(self, f):
        span = self.hi - self.lo
        return bool(np.all(f >= self.lo - 0.1 * span) and np.all(f <= self.hi + 0.1 * span))

    def choose(self, ttype, n_modes, M, ntransf, tol, dtype, methods=(1, 2)):
        """Plan options (a dict for Plan(**opts)) of the predicted fastest
        allowed option among those with gpu_method in `methods`, or {} (the
        library's heuristic) when the features are outside the training
        range."""
        return self.choose_many([(ttype, n_modes, M, ntransf, tol, dtype)], methods)[0]

    def choose_many(self, problems, methods=(1, 2)):
        """choose() for many problems with one inference launch."""
        feats = np.stack([features(*p) for p in problems])
        pred = self.predict(feats)
        out = []
        for f, y, (ttype, n_modes, M, ntransf, tol, dtype) in zip(feats, pred, problems):
            if not self.in_range(f):
                out.append({})
                continue
            banned = set(self.rejected.get(problem_class(ttype, len(n_modes), tol, dtype), []))
            order = [int(i) for i in np.argsort(y)
                     if int(i) not in banned and self.arms[int(i)]["gpu_method"] in methods]
            out.append(dict(self.arms[order[0]]) if order else {})
        return out


_default = None


def default_model():
    """The committed model, or None when no weights are present."""
    global _default
    if _default is None and WEIGHTS.exists():
        _default = CostModel.load(WEIGHTS)
    return _default


def choose_opts(ttype, n_modes, M, ntransf, tol, dtype):
    m = default_model()
    return m.choose(ttype, n_modes, M, ntransf, tol, dtype) if m else {}


# ------------------------------------------------------------ sweeps


def time_arm(ttype, n_modes, M, ntransf, tol, dtype, arm, reps=5, seed=0, err_samples=48):
    """(ms per execute, relative error at sampled outputs) of one option."""
    torch = _torch()

    from . import accuracy as acc
    from ._plan import Plan

    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    real = torch.float64 if dtype == "complex128" else torch.float32
    dim = len(n_modes)
    x = [torch.rand(M, generator=g, device=dev, dtype=real) * (2 * math.pi) - math.pi
         for _ in range(dim)]
    plan = Plan(ttype, tuple(n_modes), n_trans=ntransf, eps=tol, dtype=dtype, **arm)
    plan.setpts(*x)
    ctype = torch.complex128 if dtype == "complex128" else torch.complex64
    if ttype == 2:
        data = torch.randn((ntransf, *n_modes), generator=g, device=dev, dtype=ctype)
    else:
        data = torch.randn((ntransf, M), generator=g, device=dev, dtype=ctype)
    if ntransf == 1:
        data = data[0]
    out = plan.execute(data)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        plan.execute(data, out=out)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    d0 = data if ntransf == 1 else data[0]
    o0 = out if ntransf == 1 else out[0]
    r = torch.Generator(device="cpu")
    r.manual_seed(seed + 1)
    if ttype == 2:
        idx = torch.randint(0, M, (err_samples,), generator=r)
        ref = acc.type2_at(x, d0, plan._isign, idx)
        err = acc.relerr(o0[idx.to(dev)], ref)
    else:
        idx = torch.randint(0, int(np.prod(n_modes)), (err_samples,), generator=r)
        ref = acc.type1_at(x, d0, plan._isign, tuple(n_modes), idx)
        err = acc.relerr(o0.reshape(-1)[idx.to(dev)], ref)
    del plan
    return ms, err


def sweep_configs(seed=0, n=160):
    """Problem configurations of the sweep: 2-D and 3-D, types 1 and 2, fp32
    (three in four) and fp64, sizes around the uneven small-problem and the
    3-D workloads."""
    r = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        dim = int(r.choice([2, 2, 3]))
        ttype = int(r.choice([1, 2]))
        if dim == 2:
            side = int(np.exp(r.uniform(np.log(24), np.log(640))))
            M = int(np.exp(r.uniform(np.log(5e2), np.log(3e5))))
        else:
            side = int(np.exp(r.uniform(np.log(12), np.log(96))))
            M = int(np.exp(r.uniform(np.log(1e3), np.log(2e6))))
        ntr = int(r.choice([1, 4, 32]))
        if r.uniform() < 0.75:
            dtype, tol = "complex64", float(r.choice([1e-3, 1e-5, 1e-6]))
        else:
            dtype, tol = "complex128", float(r.choice([1e-6, 1e-9, 1e-12]))
        out.append((ttype, (side,) * dim, M, ntr, tol, dtype))
    return out


def build_model(configs, log=print, epochs=3000):
    """Sweeps every option of every configuration, trains the network and
    returns the model and the raw sweep records."""
    recs = []
    rejected = {}
    for i, (ttype, nm, M, ntr, tol, dtype) in enumerate(configs):
        times, errs = [], []
        for arm in ARMS:
            try:
                ms, err = time_arm(ttype, nm, M, ntr, tol, dtype, arm, seed=i)
            except Exception as e:  # noqa: BLE001 - an option that cannot plan is skipped
                ms, err = float("nan"), float("inf")
                log(f"config {i} arm {arm}: {e}")
            times.append(ms)
            errs.append(err)
        base = errs[0]
        ok = [math.isfinite(t) and e <= max(tol, 1.5 * base) for t, e in zip(times, errs)]
        key = problem_class(ttype, len(nm), tol, dtype)
        for a in range(NOUT):
            if not ok[a] and math.isfinite(times[a]):
                rejected.setdefault(key, [])
                if a not in rejected[key]:
                    rejected[key].append(a)
        recs.append({"ttype": ttype, "n_modes": list(nm), "M": M, "ntransf": ntr, "tol": tol,
                     "dtype": dtype, "ms": times, "err": errs, "ok": ok})
        if (i + 1) % 20 == 0:
            log(f"swept {i + 1}/{len(configs)}")
    X = np.stack([features(r["ttype"], r["n_modes"], r["M"], r["ntransf"], r["tol"], r["dtype"])
                  for r in recs])
    T = np.array([[math.log(t) if (o and math.isfinite(t)) else 0.0 for t, o in
                   zip(r["ms"], r["ok"])] for r in recs], np.float32)
    Mk = np.array([[1.0 if o else 0.0 for o in r["ok"]] for r in recs], np.float32)
    mean, std = X.mean(0), X.std(0) + 1e-3
    params, loss = train((X - mean) / std, T, Mk, epochs=epochs, log=log)
    model = CostModel(params, mean, std, X.min(0), X.max(0), rejected,
                      {"configs": len(recs), "final_loss": loss})
    return model, recs


def evaluate(model, recs):
    """Measured time of the model's choice vs the default option over sweep
    records: (default_ms, chosen_ms, best_ms) summed over the records, and the
    geometric means over records of default / chosen and default / best."""
    d = c = b = 0.0
    lc = lb = 0.0
    n = 0
    for r in recs:
        ms = r["ms"]
        if not (r["ok"][0] and math.isfinite(ms[0])):
            continue
        opts = model.choose(r["ttype"], r["n_modes"], r["M"], r["ntransf"], r["tol"], r["dtype"])
        a = model.arms.index(opts) if opts in model.arms else 0
        chosen = ms[a] if (r["ok"][a] and math.isfinite(ms[a])) else ms[0]
        best = min(t for t, o in zip(ms, r["ok"]) if o and math.isfinite(t))
        d, c, b = d + ms[0], c + chosen, b + best
        lc += math.log(ms[0] / chosen)
        lb += math.log(ms[0] / best)
        n += 1
    n = max(n, 1)
    return d, c, b, math.exp(lc / n), math.exp(lb / n)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("cmd", choices=["train"])
    ap.add_argument("--out", default=str(WEIGHTS))
    ap.add_argument("--configs", type=int, default=240)
    ap.add_argument("--holdout", type=int, default=60)
    ap.add_argument("--records", default="")
    args = ap.parse_args(argv)

    def log(msg):
        print(f"[costmodel] {msg}", file=sys.stderr, flush=True)

    t0 = time.time()
    train_cfg = sweep_configs(0, args.configs)
    model, recs = build_model(train_cfg, log=log)
    hold = sweep_configs(1, args.holdout)
    _, hrecs = build_model(hold, log=log, epochs=1)  # sweep only; its network is unused
    d, c, b, gc, gb = evaluate(model, hrecs)
    model.meta.update({"holdout_default_ms": d, "holdout_chosen_ms": c, "holdout_best_ms": b,
                       "holdout_geomean_speedup": gc, "holdout_geomean_best": gb,
                       "sweep_s": time.time() - t0})
    model.save(args.out)
    if args.records:
        Path(args.records).write_text(json.dumps({"train": recs, "holdout": hrecs}))
    log(f"held-out total time: default {d:.1f} ms, model {c:.1f} ms, best {b:.1f} ms; "
        f"per-problem geometric-mean speedup: model {gc:.3f}, best {gb:.3f}")
    print(json.dumps(model.meta))
    return 0


if __name__ == "__main__":
    sys.exit(main())
