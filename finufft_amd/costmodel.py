"""Learned cost model: which plan options run a given problem fastest.

A small network (finufft_amd/csrc/mlp.hip: 11 inputs, two ReLU layers of 32,
one output per option) predicts the log execute time of each option from

    dim, log2 N (modes), log2 M, log2 (M / N), log2 ntransf, -log10 tol,
    precision (0 fp32, 1 fp64), transform type (1 or 2),
    point density: log2 (max / mean) and log2 (99th percentile / mean) of
    the point counts per cell of a coarse grid, and the occupied fraction
    of its cells (finufft_amd_density_stats, a HIP kernel)

for the options (ARMS): the library's defaults (LDS methods, upsampling 2,
the kernel width for tol), the global-memory method (gpu_method 1), all
vectors of an execute in one pass (gpu_maxbatchsize 32), upsampling factor
1.25, a kernel one width class wider (the plan made for tol / 10), two
alternative bin shapes per (dimension, type) (BIN_SHAPES), and output-driven
spreading (gpu_method 3; types 1 and 3 only).

Training data come from sweeps of the library's own options on the GPU
(`sweep`): every option is timed and its error against the FP64 direct sum
measured at sampled outputs. An option whose error exceeds tol is masked out
of that sample, and rejected for the problem's class (type, dimension,
precision, tol decade) if it exceeds tol anywhere in the sweep: it is never
chosen there. A class where the default options themselves miss tol is
recorded (`default_misses`): there the model leaves the library's heuristic
in place and reports the miss. The network is trained on the GPU with the
library's own forward, gradient and Adam kernels and runs inference with the
forward kernel. Without weights, or for features outside the training range,
the library's own heuristic is used (no option override).

The reference has no learned component: it picks methods and bin sizes from
fixed tables (src/cuda/heuristics.cu:270-281 and :325-445,
src/cuda/makeplan.cu:289-329).

Command line (on a GPU):
    python -m finufft_amd.costmodel train --out finufft_amd/costmodel_weights.json
"""

from __future__ import annotations

import argparse
import ctypes
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

from . import _lib

ARMS = [
    {},
    {"gpu_method": 1},
    {"gpu_maxbatchsize": 32},
    {"upsampfac": 1.25},
    {"eps_scale": 0.1},
    {"bins": 0},
    {"bins": 1},
    {"gpu_method": 3},
]
# Alternatives to the heuristic's bin shape, (x, y[, z]) with x fastest, per
# (dimension, type): shallower column bins and wide LDS tiles for spreading,
# narrower and deeper tiles for interpolation.
BIN_SHAPES = {
    (2, 1): ((32, 32), (64, 32)),
    (2, 2): ((32, 16), (128, 32)),
    (3, 1): ((8, 8, 4), (16, 16, 4)),
    (3, 2): ((16, 8, 8), (32, 8, 4)),
}
NIN, NHID, NOUT = 11, 32, len(ARMS)
NPARAM = NHID * NIN + NHID + NHID * NHID + NHID + NOUT * NHID + NOUT
WEIGHTS = Path(__file__).resolve().parent / "costmodel_weights.json"
DEFAULT_MARGIN = 0.15  # log units (a predicted gain of 16 %)
# coarse grid side of the density statistics per dimension (g^dim cells)
DENSITY_GRID = {1: 4096, 2: 128, 3: 32}


def features(ttype, n_modes, M, ntransf, tol, dtype, density):
    n_tot = float(np.prod(n_modes))
    mx, p99, occ = density
    return np.array([
        len(n_modes),
        math.log2(max(n_tot, 1.0)),
        math.log2(max(M, 1)),
        math.log2(max(M, 1) / max(n_tot, 1.0)),
        math.log2(max(ntransf, 1)),
        -math.log10(tol),
        1.0 if dtype in ("complex128", "double") else 0.0,
        float(ttype),
        math.log2(max(mx, 1e-3)),
        math.log2(max(p99, 1e-3)),
        float(occ),
    ], np.float32)


def problem_class(ttype, dim, tol, dtype):
    """Key of the rejection table: (type, dim, precision, tol decade)."""
    return f"{ttype}:{dim}:{dtype}:{int(round(-math.log10(tol)))}"


def arm_valid(arm, ttype, dim):
    """Whether option `arm` applies: bin shapes where BIN_SHAPES has them,
    output-driven spreading for the types that spread."""
    if arm.get("gpu_method") == 3 and ttype == 2:
        return False
    return "bins" not in arm or (dim, ttype) in BIN_SHAPES


def plan_args(arm, ttype, n_modes, tol):
    """(eps, Plan keyword options) of option `arm` for a problem."""
    opts = {k: v for k, v in arm.items() if k not in ("eps_scale", "bins")}
    eps = tol * arm.get("eps_scale", 1.0)
    if "bins" in arm:
        shape = BIN_SHAPES[(len(n_modes), ttype)][arm["bins"]]
        for name, b in zip(("gpu_binsizex", "gpu_binsizey", "gpu_binsizez"), shape):
            opts[name] = int(b)
    return eps, opts


def density_of(pts, stream=None):
    """(max / mean, 99th percentile / mean, occupied fraction) of the point
    counts on the coarse grid, from device coordinate tensors (one per
    dimension, array order as for Plan.setpts), with the library's kernel."""
    dim = len(pts)
    M = int(pts[0].numel())
    is_double = 1 if str(pts[0].dtype).endswith("float64") else 0
    # the C interface takes coordinates x fastest, as setpts does
    ptrs = [int(p.data_ptr()) if M > 0 else None for p in reversed(pts)] + [None] * (3 - dim)
    out = (ctypes.c_double * 3)()
    ier = _lib.lib().finufft_amd_density_stats(dim, M, ptrs[0], ptrs[1], ptrs[2], is_double,
                                               DENSITY_GRID[dim], out, stream)
    if ier != 0:
        raise RuntimeError(f"finufft_amd_density_stats failed ({ier})")
    return tuple(float(v) for v in out)


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
        if log and step % 1000 == 0:
            log(f"epoch {step}: loss {float(loss):.4g}")
    return P.cpu().numpy(), float(loss)


# ------------------------------------------------------------ the model


class CostModel:
    def __init__(self, params, mean, std, lo, hi, rejected=None, meta=None, arms=None,
                 default_misses=None):
        self.arms = [dict(a) for a in (arms or ARMS)]  # the options the network was trained on
        self.params = np.asarray(params, np.float32)
        self.mean = np.asarray(mean, np.float32)
        self.std = np.asarray(std, np.float32)
        self.lo = np.asarray(lo, np.float32)
        self.hi = np.asarray(hi, np.float32)
        self.rejected = rejected or {}
        self.default_misses = default_misses or []
        self.meta = meta or {}
        # An option other than the defaults is chosen only when its predicted
        # log time is below the defaults' by more than this margin: smaller
        # predicted gains are within the network's error on held-out sweeps.
        self.margin = float(self.meta.get("margin", DEFAULT_MARGIN))
        self._dev_params = None

    @classmethod
    def load(cls, path=WEIGHTS):
        d = json.loads(Path(path).read_text())
        if len(d["params"]) != NPARAM or len(d["mean"]) != NIN:
            raise ValueError(f"{path}: weights of another network layout")
        return cls(d["params"], d["mean"], d["std"], d["lo"], d["hi"], d.get("rejected"),
                   d.get("meta"), d.get("arms"), d.get("default_misses"))

    def save(self, path):
        Path(path).write_text(json.dumps({
            "params": [float(v) for v in self.params],
            "mean": self.mean.tolist(), "std": self.std.tolist(),
            "lo": self.lo.tolist(), "hi": self.hi.tolist(),
            "rejected": self.rejected, "default_misses": self.default_misses,
            "meta": self.meta, "arms": self.arms,
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

    def misses_tol(self, ttype, dim, tol, dtype):
        """True where the sweep saw the default options miss tol in this
        problem class (the model then keeps the heuristic and says so)."""
        return problem_class(ttype, dim, tol, dtype) in self.default_misses

    def choose(self, ttype, n_modes, M, ntransf, tol, dtype, density, arms=None):
        """Index into self.arms of the predicted fastest allowed option (the
        defaults, 0, unless another is predicted faster by more than
        self.margin in log time), or None (the library's heuristic):
        features outside the training range, a class never swept, or one
        where the defaults miss tol. `arms` limits the candidates
        (indices)."""
        return self.choose_many([(ttype, n_modes, M, ntransf, tol, dtype, density)], arms)[0]

    def choose_many(self, problems, arms=None):
        feats = np.stack([features(*p) for p in problems])
        pred = self.predict(feats)
        out = []
        for f, y, (ttype, n_modes, M, ntransf, tol, dtype, _) in zip(feats, pred, problems):
            key = problem_class(ttype, len(n_modes), tol, dtype)
            if not self.in_range(f) or key not in self.rejected or key in self.default_misses:
                out.append(None)
                continue
            banned = set(self.rejected[key])
            order = [int(i) for i in np.argsort(y)
                     if int(i) not in banned and (arms is None or int(i) in arms)
                     and arm_valid(self.arms[int(i)], ttype, len(n_modes))]
            pick = order[0] if order else None
            if pick not in (None, 0) and 0 in order and y[pick] > y[0] - self.margin:
                pick = 0  # the defaults, unless the predicted gain exceeds the margin
            out.append(pick)
        return out

    def plan_args(self, choice, ttype, n_modes, tol):
        """(eps, Plan options) of choose()'s result (None: the defaults)."""
        if choice is None:
            return tol, {}
        return plan_args(self.arms[choice], ttype, n_modes, tol)


_default = None


def default_model():
    """The committed model, or None when no (current-layout) weights are
    present."""
    global _default
    if _default is None and WEIGHTS.exists():
        try:
            _default = CostModel.load(WEIGHTS)
        except (ValueError, KeyError):
            return None
    return _default


# ------------------------------------------------------------ sweeps


def make_points(kind, dim, M, seed, real="float32", device="cuda"):
    """Seeded point sets on the GPU (array order): uniform on [-pi, pi)^dim,
    a mixture of 16 Gaussians (sigma 0.05 pi, means uniform), or radial
    spokes (lines through the origin in uniformly random directions, points
    evenly spaced in radius from 0 to pi)."""
    torch = _torch()
    dtype = getattr(torch, real)
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    if kind == "uniform":
        return [(torch.rand(M, generator=g, device=device, dtype=torch.float64) * 2 - 1)
                .mul_(math.pi).to(dtype) for _ in range(dim)]
    if kind == "clustered":
        mu = (torch.rand(16, dim, generator=g, device=device, dtype=torch.float64) * 2 - 1) * math.pi
        which = torch.randint(0, 16, (M,), generator=g, device=device)
        p = mu[which] + torch.randn(M, dim, generator=g, device=device,
                                    dtype=torch.float64) * (0.05 * math.pi)
        return [p[:, d].to(dtype).contiguous() for d in range(dim)]
    if kind == "spokes":
        nsp = max(1, int(round(math.sqrt(M) * (0.4 if dim == 3 else 0.2))))
        per = max(1, M // nsp)
        u = torch.randn(nsp, dim, generator=g, device=device, dtype=torch.float64)
        u = u / u.norm(dim=1, keepdim=True)
        rad = torch.linspace(0, math.pi, per, device=device, dtype=torch.float64)
        p = (u[:, None, :] * rad[None, :, None]).reshape(-1, dim)
        return [p[:, d].to(dtype).contiguous() for d in range(dim)]
    raise ValueError(kind)


def time_arm(ttype, n_modes, pts, ntransf, tol, dtype, arm, reps=5, seed=0, err_samples=48):
    """(ms per execute, relative error at sampled outputs) of one option on
    the given device points."""
    torch = _torch()

    from . import accuracy as acc
    from ._plan import Plan

    dev = pts[0].device
    M = int(pts[0].numel())
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    eps, opts = plan_args(arm, ttype, n_modes, tol)
    plan = Plan(ttype, tuple(n_modes), n_trans=ntransf, eps=eps, dtype=dtype, **opts)
    plan.setpts(*pts)
    ctype = torch.complex128 if dtype == "complex128" else torch.complex64
    if ttype == 2:
        data = torch.randn((ntransf, *n_modes), generator=g, device=dev, dtype=ctype)
    else:
        data = torch.randn((ntransf, M), generator=g, device=dev, dtype=ctype)
    if ntransf == 1:
        data = data[0]
    out = plan.execute(data)
    plan.execute(data, out=out)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    # at least `reps` executes and about 3 ms of them: single small problems
    # run for tens of microseconds, where a few repetitions time mostly noise
    t0 = time.perf_counter()
    plan.execute(data, out=out)
    torch.cuda.synchronize()
    one = max(time.perf_counter() - t0, 1e-6)
    reps = int(min(200, max(reps, math.ceil(3e-3 / one))))
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
        ref = acc.type2_at(pts, d0, plan._isign, idx)
        err = acc.relerr(o0[idx.to(dev)], ref)
    else:
        idx = torch.randint(0, int(np.prod(n_modes)), (err_samples,), generator=r)
        ref = acc.type1_at(pts, d0, plan._isign, tuple(n_modes), idx)
        err = acc.relerr(o0.reshape(-1)[idx.to(dev)], ref)
    del plan
    return ms, err


def sweep_configs(seed=0, n=200):
    """Problem configurations of the sweep: 2-D and 3-D, types 1 and 2,
    uniform, clustered and radial-spoke points, fp32 (three in four) and
    fp64, sizes around the uneven small-problem, the 2-D and the 3-D
    workloads; problems drawn as the uneven small-problem workload's; plus,
    for the first sweep, the 3-D headline shape (256^3 modes, 1e8 points,
    uniform and spokes) and the 2-D one (1000^2, 1e7)."""
    r = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        dim = int(r.choice([2, 2, 3]))
        ttype = int(r.choice([1, 2]))
        kind = str(r.choice(["uniform", "uniform", "clustered", "spokes"]))
        if dim == 2:
            side = int(np.exp(r.uniform(np.log(24), np.log(1024))))
            M = int(np.exp(r.uniform(np.log(5e2), np.log(4e6))))
        else:
            side = int(np.exp(r.uniform(np.log(12), np.log(160))))
            M = int(np.exp(r.uniform(np.log(1e3), np.log(1.6e7))))
        ntr = int(r.choice([1, 4, 32])) if M * side ** dim < 2e9 else 1
        if r.uniform() < 0.75:
            dtype, tol = "complex64", float(r.choice([1e-3, 1e-5, 1e-6]))
        else:
            dtype, tol = "complex128", float(r.choice([1e-6, 1e-9, 1e-12]))
        out.append((ttype, (side,) * dim, M, ntr, tol, dtype, kind))
    # problems of the uneven small-problem workload (2-D type 2, 32 vectors)
    for _ in range(24 if seed == 0 else 8):
        side = int(np.exp(r.uniform(np.log(32), np.log(512))))
        M = int(np.exp(r.uniform(np.log(1e3), np.log(1e5))))
        out.append((2, (side, side), M, 32, 1e-5, "complex64", "uniform"))
    if seed == 0:
        for ttype in (1, 2):
            for kind in ("uniform", "spokes"):
                out.append((ttype, (256, 256, 256), 100_000_000, 1, 1e-6, "complex64", kind))
            for kind in ("uniform", "clustered"):
                out.append((ttype, (1000, 1000), 10_000_000, 1, 1e-5, "complex64", kind))
    return out


def sweep(configs, log=print):
    """Times every option of every configuration; returns the records."""
    torch = _torch()
    recs = []
    t0 = time.time()
    for i, (ttype, nm, M, ntr, tol, dtype, kind) in enumerate(configs):
        real = "float64" if dtype == "complex128" else "float32"
        pts = make_points(kind, len(nm), M, seed=1000 + i, real=real)
        M = int(pts[0].numel())
        dens = density_of(pts)
        times, errs = [], []
        for arm in ARMS:
            if not arm_valid(arm, ttype, len(nm)):
                times.append(float("nan"))
                errs.append(float("inf"))
                continue
            try:
                ms, err = time_arm(ttype, nm, pts, ntr, tol, dtype, arm, seed=i,
                                   reps=3 if M * ntr > 1e7 else 5)
            except Exception as e:  # noqa: BLE001 - an option that cannot plan is skipped
                ms, err = float("nan"), float("inf")
                log(f"config {i} arm {arm}: {e}")
            times.append(ms)
            errs.append(err)
        recs.append({"ttype": ttype, "n_modes": list(nm), "M": M, "ntransf": ntr, "tol": tol,
                     "dtype": dtype, "kind": kind, "density": list(dens), "ms": times,
                     "err": errs,
                     "ok": [math.isfinite(t) and e <= tol for t, e in zip(times, errs)]})
        del pts
        torch.cuda.empty_cache()
        if (i + 1) % 10 == 0 or M > 1e7:
            log(f"swept {i + 1}/{len(configs)} ({time.time() - t0:.0f} s)")
    return recs


def rejection_tables(recs):
    """(rejected, default_misses): per class the options that missed tol in
    some record, and the classes where the default options missed tol."""
    rejected, misses = {}, []
    for r in recs:
        key = problem_class(r["ttype"], len(r["n_modes"]), r["tol"], r["dtype"])
        rejected.setdefault(key, [])
        for a, (t, o) in enumerate(zip(r["ms"], r["ok"])):
            if math.isfinite(t) and not o and a not in rejected[key]:
                rejected[key].append(a)
        if math.isfinite(r["ms"][0]) and not r["ok"][0] and key not in misses:
            misses.append(key)
    return rejected, misses


def fit(recs, log=print, epochs=4000):
    """Trains the network on sweep records; returns the model."""
    X = np.stack([features(r["ttype"], r["n_modes"], r["M"], r["ntransf"], r["tol"], r["dtype"],
                           r["density"]) for r in recs])
    T = np.array([[math.log(t) if (o and math.isfinite(t)) else 0.0 for t, o in
                   zip(r["ms"], r["ok"])] for r in recs], np.float32)
    Mk = np.array([[1.0 if o else 0.0 for o in r["ok"]] for r in recs], np.float32)
    mean, std = X.mean(0), X.std(0) + 1e-3
    params, loss = train((X - mean) / std, T, Mk, epochs=epochs, log=log)
    rejected, misses = rejection_tables(recs)
    return CostModel(params, mean, std, X.min(0), X.max(0), rejected,
                     {"configs": len(recs), "final_loss": loss}, ARMS, misses)


def evaluate(model, recs):
    """Measured time of the model's choice vs the default option over sweep
    records whose default meets tol: (default_ms, chosen_ms, best_ms) summed,
    the geometric means of default / chosen and default / best, and the
    number of records where the choice missed tol (0 by construction of the
    rejection table when the records come from the training classes)."""
    d = c = b = 0.0
    lc = lb = 0.0
    n = 0
    missed = 0
    for r in recs:
        ms = r["ms"]
        if not (r["ok"][0] and math.isfinite(ms[0])):
            continue
        a = model.choose(r["ttype"], r["n_modes"], r["M"], r["ntransf"], r["tol"], r["dtype"],
                         r["density"])
        a = 0 if a is None else a
        if not r["ok"][a]:
            missed += 1
        chosen = ms[a] if (r["ok"][a] and math.isfinite(ms[a])) else ms[0]
        best = min(t for t, o in zip(ms, r["ok"]) if o and math.isfinite(t))
        d, c, b = d + ms[0], c + chosen, b + best
        lc += math.log(ms[0] / chosen)
        lb += math.log(ms[0] / best)
        n += 1
    n = max(n, 1)
    return {"default_ms": d, "chosen_ms": c, "best_ms": b, "geomean_speedup": math.exp(lc / n),
            "geomean_best": math.exp(lb / n), "records": n, "choice_missed_tol": missed}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("cmd", choices=["train"])
    ap.add_argument("--out", default=str(WEIGHTS))
    ap.add_argument("--configs", type=int, default=200)
    ap.add_argument("--holdout", type=int, default=50)
    ap.add_argument("--records", default="")
    args = ap.parse_args(argv)

    def log(msg):
        print(f"[costmodel] {msg}", file=sys.stderr, flush=True)

    t0 = time.time()
    recs = sweep(sweep_configs(0, args.configs), log=log)
    model = fit(recs, log=log)
    hrecs = sweep(sweep_configs(1, args.holdout), log=log)
    ev = evaluate(model, hrecs)
    model.meta.update({"holdout_" + k: v for k, v in ev.items()})
    model.meta["sweep_s"] = time.time() - t0
    model.save(args.out)
    if args.records:
        Path(args.records).write_text(json.dumps({"train": recs, "holdout": hrecs}))
    log(f"held-out: {ev}")
    print(json.dumps(model.meta))
    return 0


if __name__ == "__main__":
    sys.exit(main())
