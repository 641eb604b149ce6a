#!/usr/bin/env python3
"""Benchmark of the NUFFT hot path on one MI355X.

Workload (the headline configuration): 3-D type 1 and type 2, fp32
(complex64 data), tol 1e-6, M = 1e8 nonuniform points i.i.d. uniform on
[-pi, pi)^3, N = 256^3 modes, complex N(0,1) strengths / coefficients,
isign +1 (type 1) and -1 (type 2). Inputs are generated on the GPU (seeded)
and resident in HBM before timing. Plans use the library's default options
(LDS methods for both types: the column-window spreader for type 1, bank-
ordered subproblem interpolation for type 2).

One step = one type-1 execute + one type-2 execute over the same points.
value = 2 M / (time per step) NU pts/s. setpts (the sort) is timed separately
and reported, not counted.

The second part of the metric, the uneven small-problem comparison
(`batched`), runs 2000 independent 2-D type-2 problems (N log-uniform in
[32, 512], M log-uniform in [1e3, 1e5], 32 vectors each, fp32 tol 1e-5) as a
loop of per-plan executes and as one fused batch replayed from a HIP graph,
and reports the speedup at the errors measured for both.

Extra fields:
  roofline      the dominant stage (spread or interp, the longer), its
                algorithmic bytes per launch over its average launch time
                (the stage time, timed with HIP events on the plan's stream
                over the timed region, over its launches: one per parity
                color for colored spreading); traffic, the HBM bytes per
                launch the kernel moves, and its limiter come from the
                committed counter summary of this workload (used when the
                summary's kernel time matches this run's within 25%);
  other_configs the 2-D type 1 + 2 workloads (uniform and clustered points)
                and the 3-D differentiation workload, each with its time,
                throughput, errors and spread / interp roofline figures;
  stages        average ms per execute of every stage of both plans;
  error         relative l2 error against an FP64 direct sum on the GPU at
                sampled modes (type 1) and sampled points (type 2);
  cpu_baseline  the oracle's CPU restatement (C + OpenMP spread/interp,
                scipy.fft FFT) on a bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--M 1e8] [--N 256]
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

METRIC = "NU pts/s 3D type1/2 fp32 @tol1e-6; batched+learned vs plain speedup, same err"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the headline kernels, from rocprofv3 FETCH_SIZE /
# WRITE_SIZE runs of this workload (calibrated against a plain copy), with
# the command that made them and what the SQ counters say bounds each kernel
TRAFFIC = ROOT / "profiles" / "hbm_traffic_3d_fp32_M1e8_N256.json"
STAGE_KERNEL = {"spread (type 1)": "spread_col_colored_kernel<7>",
                "interp (type 2)": "interp_bin_kernel<float, 3, 7>"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo)."""
    try:
        for ln in Path("/proc/cpuinfo").read_text().splitlines():
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def finufft_probe():
    """Whether a preinstalled reference CPU finufft is importable or on the
    loader path (probed only; nothing is installed)."""
    import importlib.util
    import subprocess

    found = []
    if importlib.util.find_spec("finufft") is not None:
        found.append("python module finufft")
    try:
        res = subprocess.run(["ldconfig", "-p"], capture_output=True, text=True, timeout=10)
        if "finufft" in res.stdout:
            found.append("libfinufft on the loader path")
    except (OSError, subprocess.SubprocessError):
        pass
    return ", ".join(found) if found else "absent (no finufft module, no libfinufft)"


def measured_traffic(stage, ms_per_launch):
    """(HBM bytes per launch, source, limiter) of the stage's kernel from the
    committed counter summary, or (None, None, None) when it has no entry.
    The summary's kernel time is given too: a reading taken at a kernel time
    far from this run's is stale and not used."""
    try:
        rec = json.loads(TRAFFIC.read_text())["kernels"][STAGE_KERNEL[stage]]
    except (OSError, KeyError, ValueError):
        return None, None, None
    if not 0.8 <= rec["ms_per_launch"] / ms_per_launch <= 1.25:
        return None, None, None
    return int(rec["hbm_bytes_per_launch"]), TRAFFIC.name, rec.get("limiter")


def cpu_baseline(N, tol, M_full, M_sample, threads):
    """Times the oracle's CPU restatement on a sample of the workload and
    scales the spread/interp part to the full point count."""
    import numpy as np

    from oracle import cpu

    r = np.random.default_rng(0)
    pts = [r.uniform(-np.pi, np.pi, M_sample).astype(np.float32) for _ in range(3)]
    c = (r.standard_normal(M_sample) + 1j * r.standard_normal(M_sample)).astype(np.complex64)
    f = (r.standard_normal((N, N, N)) + 1j * r.standard_normal((N, N, N))).astype(np.complex64)
    out = {}
    for ttype, data in ((1, c), (2, f)):
        p = cpu.CpuPlan(ttype, (N, N, N), eps=tol, nthreads=threads)
        p.setpts(*pts)
        p.execute(data)  # warm (page-in, FFT plan)
        t0 = time.perf_counter()
        p.execute(data)
        total = time.perf_counter() - t0
        # fixed cost (grid FFT + deconvolution) from a run with no points
        pz = cpu.CpuPlan(ttype, (N, N, N), eps=tol, nthreads=threads)
        pz.setpts(*[a[:1] for a in pts])
        t0 = time.perf_counter()
        pz.execute(data[:1] if ttype == 1 else data)
        fixed = time.perf_counter() - t0
        per_pt = max(total - fixed, 0.0) / M_sample
        out[ttype] = fixed + per_pt * M_full
    value = 2 * M_full / (out[1] + out[2])
    return {
        "value": value,
        "unit": "NU pts/s",
        "cores": threads,
        "cpu_model": cpu_model(),
        "reference_finufft": finufft_probe(),
        "kind": "port",
        "sample": (f"type1+type2 at N={N}^3, tol {tol}, on M={M_sample:.0e} of the {M_full:.0e} "
                   f"points; spread/interp time scaled linearly to M={M_full:.0e}, grid FFT "
                   f"(scipy.fft) and deconvolution timed at full size; projected s/step "
                   f"t1={out[1]:.2f} t2={out[2]:.2f}"),
    }


def batched_type1_comparison(nprob, ntr, tol, reps):
    """The same uneven workload in the type-1 direction (strengths -> modes):
    a loop of per-plan executes against the fused type-1 batch (per chunk one
    spreading launch over all its problems, a batched FFT and one
    deconvolution launch, replayed as a HIP graph); errors at sampled modes of
    sampled problems against the FP64 direct sum."""
    import numpy as np
    import torch

    from finufft_amd import Plan
    from finufft_amd import accuracy as acc
    from finufft_amd.batch import BatchPlan

    dev = torch.device("cuda", 0)
    r = np.random.default_rng(6)
    Ns = np.rint(np.exp(r.uniform(np.log(32), np.log(512), nprob))).astype(int)
    Ms = np.rint(np.exp(r.uniform(np.log(1e3), np.log(1e5), nprob))).astype(int)
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    plans, pts, cs, outs = [], [], [], []
    for n, m in zip(Ns, Ms):
        x = [torch.rand(int(m), generator=g, device=dev) * (2 * math.pi) - math.pi
             for _ in range(2)]
        c = torch.complex(torch.randn((ntr, int(m)), generator=g, device=dev),
                          torch.randn((ntr, int(m)), generator=g, device=dev))
        p = Plan(1, (int(n), int(n)), n_trans=ntr, eps=tol, dtype="complex64", isign=1)
        p.setpts(*x)
        plans.append(p)
        pts.append(x)
        cs.append(c)
        outs.append(torch.empty((ntr, int(n), int(n)), dtype=torch.complex64, device=dev))
    torch.cuda.synchronize()

    def plain():
        for p, c, o in zip(plans, cs, outs):
            p.execute(c, out=o)

    samp = r.choice(nprob, min(8, nprob), replace=False)
    midx = {int(i): torch.randint(0, int(Ns[i]) ** 2, (32,), device=dev) for i in samp}

    def errors():
        e = 0.0
        for i in samp:
            n = int(Ns[i])
            ref = acc.type1_at(pts[i], cs[i][0], 1, (n, n), midx[int(i)])
            e = max(e, acc.relerr(outs[i][0].reshape(-1)[midx[int(i)]], ref))
        return e

    plain()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        plain()
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t0) / reps
    e_plain = errors()
    b = BatchPlan(plans)
    b.execute(cs, outs)  # records the graph
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.execute(cs, outs)
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / reps
    e_batch = errors()
    evals = float(np.sum(Ms)) * ntr
    info = b.info()
    log(f"batched type 1: per-plan loop {t_plain * 1e3:.1f} ms, batch {t_batch * 1e3:.1f} ms, "
        f"errors {e_plain:.2e} / {e_batch:.2e}")
    return {
        "workload": f"{nprob} 2D type-1 problems, N log-uniform [32,512], M log-uniform "
                    f"[1e3,1e5], ntransf {ntr}, fp32 tol {tol:g}",
        "speedup": t_plain / t_batch, "plain_ms": round(t_plain * 1e3, 3),
        "batched_ms": round(t_batch * 1e3, 3), "plain_nu_evals_per_s": evals / t_plain,
        "batched_nu_evals_per_s": evals / t_batch, "error_plain": e_plain,
        "error_batched": e_batch, "vectors_per_round": info["vectors_per_round"],
        "graph": info["graph"],
    }


def batched_comparison(nprob, ntr, tol, reps, log_every=500):
    """The uneven small-problem workload: nprob independent 2-D type-2
    problems, side N log-uniform in [32, 512], M log-uniform in [1e3, 1e5]
    points uniform on [-pi, pi)^2, ntr complex normal coefficient vectors each,
    fp32 at tol. Times a loop of per-plan executes against the fused batch
    (per cache-sized chunk of equal-shape problems: a column transform with
    the amplification folded in, a row transform and one interpolation
    launch, replayed as a HIP graph). Returns the
    speedup, both throughputs (NU point evaluations per second, M ntr per
    problem) and the errors at sampled points of sampled problems."""
    import numpy as np
    import torch

    from finufft_amd import Plan
    from finufft_amd import accuracy as acc
    from finufft_amd.batch import BatchPlan

    dev = torch.device("cuda", 0)
    r = np.random.default_rng(3)
    Ns = np.rint(np.exp(r.uniform(np.log(32), np.log(512), nprob))).astype(int)
    Ms = np.rint(np.exp(r.uniform(np.log(1e3), np.log(1e5), nprob))).astype(int)
    g = torch.Generator(device=dev)
    g.manual_seed(4)
    plans, pts, fks, outs = [], [], [], []
    t0 = time.perf_counter()
    for i, (n, m) in enumerate(zip(Ns, Ms)):
        x = [torch.rand(int(m), generator=g, device=dev) * (2 * math.pi) - math.pi
             for _ in range(2)]
        fk = torch.complex(torch.randn((ntr, n, n), generator=g, device=dev),
                           torch.randn((ntr, n, n), generator=g, device=dev))
        p = Plan(2, (int(n), int(n)), n_trans=ntr, eps=tol, dtype="complex64", isign=-1)
        p.setpts(*x)
        plans.append(p)
        pts.append(x)
        fks.append(fk)
        outs.append(torch.empty((ntr, int(m)), dtype=torch.complex64, device=dev))
        if log_every and (i + 1) % log_every == 0:
            log(f"batched workload: {i + 1}/{nprob} problems planned")
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0

    def plain():
        for p, fk, o in zip(plans, fks, outs):
            p.execute(fk, out=o)

    plain()  # FFT plans, page-in
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        plain()
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t0) / reps
    samp = r.choice(nprob, min(8, nprob), replace=False)
    pidx = {int(i): torch.randint(0, int(Ms[i]), (32,), device=dev) for i in samp}

    def errors():
        e = 0.0
        for i in samp:
            ref = acc.type2_at(pts[i], fks[i][0], -1, pidx[int(i)])
            e = max(e, acc.relerr(outs[i][0][pidx[int(i)]], ref))
        return e

    e_plain = errors()
    log(f"batched workload: setup {t_setup:.1f} s, per-plan loop {t_plain * 1e3:.1f} ms")

    t0 = time.perf_counter()
    b = BatchPlan(plans)
    for o in outs:
        o.zero_()
    b.execute(fks, outs)  # records the graph
    torch.cuda.synchronize()
    t_bsetup = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(reps):
        b.execute(fks, outs)
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / reps
    e_batch = errors()
    info = b.info()
    evals = float(np.sum(Ms)) * ntr
    log(f"batched workload: fused {t_batch * 1e3:.1f} ms ({info})")
    del b

    # Learned options: the cost model picks, per problem, among the options
    # the fused executor runs (defaults, upsampling 1.25, the wider kernel);
    # problems of equal choice form one batch each.
    learned = None
    from finufft_amd import costmodel

    model = costmodel.default_model()
    if model is not None:
        fused_arms = [i for i, a in enumerate(model.arms)
                      if set(a) <= {"upsampfac", "eps_scale"}]
        probs = [(2, (int(n), int(n)), int(m), ntr, tol, "complex64", costmodel.density_of(pts[i]))
                 for i, (n, m) in enumerate(zip(Ns, Ms))]
        choices = model.choose_many(probs, arms=fused_arms)
        del plans
        torch.cuda.empty_cache()
        lplans = []
        for i, (n, m) in enumerate(zip(Ns, Ms)):
            eps, opts = model.plan_args(choices[i], 2, (int(n), int(n)), tol)
            p = Plan(2, (int(n), int(n)), n_trans=ntr, eps=eps, dtype="complex64", isign=-1,
                     **opts)
            p.setpts(*pts[i])
            lplans.append(p)
        groups = {}
        for i, a in enumerate(choices):
            groups.setdefault(0 if a is None else a, []).append(i)
        batches = [(BatchPlan([lplans[i] for i in idx]), idx) for idx in groups.values()]
        for o in outs:
            o.zero_()

        def run_learned():
            for bb, idx in batches:
                bb.execute([fks[i] for i in idx], [outs[i] for i in idx])

        run_learned()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            run_learned()
        torch.cuda.synchronize()
        t_learned = (time.perf_counter() - t0) / reps
        e_learned = errors()
        learned = {
            "speedup": t_plain / t_learned,
            "ms": round(t_learned * 1e3, 3),
            "nu_evals_per_s": evals / t_learned,
            "error": e_learned,
            "choices": {json.dumps(model.arms[k]): len(v) for k, v in groups.items()},
        }
        log(f"batched workload: fused + learned {t_learned * 1e3:.1f} ms ({learned['choices']})")
        del batches, lplans
    return {
        "workload": (f"{nprob} 2D type-2 problems, N log-uniform [32,512], M log-uniform "
                     f"[1e3,1e5], ntransf {ntr}, fp32 tol {tol:g}"),
        "speedup": t_plain / t_batch,
        "plain_ms": round(t_plain * 1e3, 3),
        "batched_ms": round(t_batch * 1e3, 3),
        "plain_nu_evals_per_s": evals / t_plain,
        "batched_nu_evals_per_s": evals / t_batch,
        "error_plain": e_plain,
        "error_batched": e_batch,
        "setup_s": round(t_setup, 2),
        "batch_setup_s": round(t_bsetup, 2),
        "vectors_per_round": info["vectors_per_round"],
        "grid_shapes": info["grid_shapes"],
        "graph": info["graph"],
        "learned": learned,
    }


def spokes_3d(nspokes, per, seed, dev):
    """Points on 3-D radial spokes: nspokes directions uniform on the sphere,
    per samples each evenly spaced in radius from 0 to pi."""
    import torch

    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    u = torch.randn(nspokes, 3, generator=g, device=dev, dtype=torch.float64)
    u = u / u.norm(dim=1, keepdim=True)
    rad = torch.linspace(0, math.pi, per, device=dev, dtype=torch.float64)
    p = (u[:, None, :] * rad[None, :, None]).reshape(-1, 3)
    return [p[:, d].to(torch.float32).contiguous() for d in range(3)]


def c3_comparison(pts, c, f, shape, tol, steps, err_samples):
    """The headline problem with the library's default options ("plain")
    and with the options the cost model selects from the problem and its
    points' density ("selected"): ms per step (type 1 + type 2), NU pts/s and
    the errors against the FP64 direct sum, for one point set."""
    import torch

    from finufft_amd import Plan
    from finufft_amd import accuracy as acc
    from finufft_amd import costmodel

    dev = c.device
    M = int(pts[0].numel())
    model = costmodel.default_model()
    dens = costmodel.density_of(pts)
    out = {"density": [round(v, 3) for v in dens]}
    variants = {"plain": (None, None)}
    if model is not None:
        a1 = model.choose(1, shape, M, 1, tol, "complex64", dens)
        a2 = model.choose(2, shape, M, 1, tol, "complex64", dens)
        variants["selected"] = (a1, a2)
    r = torch.Generator(device="cpu")
    r.manual_seed(5)
    midx = torch.randint(0, math.prod(shape), (err_samples,), generator=r)
    pidx = torch.randint(0, M, (err_samples,), generator=r)
    ref1 = acc.type1_at(pts, c, 1, shape, midx)
    ref2 = acc.type2_at(pts, f, -1, pidx)
    fk_out = torch.empty(shape, dtype=torch.complex64, device=dev)
    c_out = torch.empty(M, dtype=torch.complex64, device=dev)
    for name, (a1, a2) in variants.items():
        eps1, o1 = (tol, {}) if model is None else model.plan_args(a1, 1, shape, tol)
        eps2, o2 = (tol, {}) if model is None else model.plan_args(a2, 2, shape, tol)
        p1 = Plan(1, shape, eps=eps1, dtype="complex64", isign=1, **o1)
        p2 = Plan(2, shape, eps=eps2, dtype="complex64", isign=-1, **o2)
        p1.setpts(*pts)
        p2.setpts(*pts)
        p1.execute(c, out=fk_out)
        p2.execute(f, out=c_out)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            p1.execute(c, out=fk_out)
            p2.execute(f, out=c_out)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / steps * 1e3
        # stage split from one more (untimed) step with HIP events
        p1.set_stage_timing(True)
        p2.set_stage_timing(True)
        p1.stage_times()
        p2.stage_times()
        p1.execute(c, out=fk_out)
        p2.execute(f, out=c_out)
        st1 = {k: round(v, 3) for k, v in p1.stage_times().items() if v > 0}
        st2 = {k: round(v, 3) for k, v in p2.stage_times().items() if v > 0}
        p1.set_stage_timing(False)
        p2.set_stage_timing(False)
        e1 = acc.relerr(fk_out.reshape(-1)[midx.to(dev)], ref1)
        e2 = acc.relerr(c_out[pidx.to(dev)], ref2)
        i1, i2 = p1.info(), p2.info()
        out[name] = {
            "ms_per_step": round(ms, 3), "nu_pts_per_s": 2 * M / (ms * 1e-3),
            "error": {"type1": e1, "type2": e2},
            "stages_ms": {"type1": st1, "type2": st2},
            "type1": {"options": model.arms[a1] if (model and a1 is not None) else {},
                      "ns": i1["ns"], "gpu_method": i1["gpu_method"],
                      "binsize": i1["binsize"], "upsampfac": i1.get("upsampfac")},
            "type2": {"options": model.arms[a2] if (model and a2 is not None) else {},
                      "ns": i2["ns"], "gpu_method": i2["gpu_method"],
                      "binsize": i2["binsize"], "upsampfac": i2.get("upsampfac")},
        }
        log(f"c3 {name}: {ms:.2f} ms/step, errors {e1:.2e} / {e2:.2e}, "
            f"type1 {out[name]['type1']}, type2 {out[name]['type2']}")
        del p1, p2
        torch.cuda.empty_cache()
    if "selected" in out:
        out["speedup"] = out["plain"]["ms_per_step"] / out["selected"]["ms_per_step"]
    return out


def gaussian_mixture_2d(M, seed, dev):
    """M points from a mixture of 16 Gaussians (sigma 0.05 pi, means uniform
    on [-pi, pi)^2, equal weights); the library folds points outside the
    period."""
    import torch

    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    means = torch.rand(16, 2, generator=g, device=dev) * (2 * math.pi) - math.pi
    comp = torch.randint(0, 16, (M,), generator=g, device=dev)
    p = means[comp] + 0.05 * math.pi * torch.randn(M, 2, generator=g, device=dev)
    return [p[:, d].contiguous() for d in range(2)]


def _stage_roofline(M, dim, nf, ms, which):
    """Algorithmic bytes of a spread or interp stage (one vector) and the
    fraction of the HBM peak at the stage's measured time."""
    nbytes = (M * (4 + 4 * dim + 8) + 8 * nf) if which == "spread" else (
        8 * nf + M * (4 + 4 * dim) + 8 * M)
    gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"stage": which, "ms": round(ms, 4), "alg_bytes": nbytes, "achieved_gbs": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4)}


def type12_config(pts, shape, tol, steps, err_samples, seed):
    """Type 1 + type 2 (fp32, default options) on the given points: ms per
    step, NU pts/s, errors at sampled modes / points, and the spread and
    interp stages against the HBM roofline."""
    import torch

    from finufft_amd import Plan
    from finufft_amd import accuracy as acc

    dev = pts[0].device
    M, dim = int(pts[0].numel()), len(shape)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    c = torch.complex(torch.randn(M, generator=g, device=dev),
                      torch.randn(M, generator=g, device=dev))
    f = torch.complex(torch.randn(shape, generator=g, device=dev),
                      torch.randn(shape, generator=g, device=dev))
    p1 = Plan(1, shape, eps=tol, dtype="complex64", isign=1)
    p2 = Plan(2, shape, eps=tol, dtype="complex64", isign=-1)
    p1.setpts(*pts)
    p2.setpts(*pts)
    fk = torch.empty(shape, dtype=torch.complex64, device=dev)
    co = torch.empty(M, dtype=torch.complex64, device=dev)
    p1.execute(c, out=fk)
    p2.execute(f, out=co)
    p1.set_stage_timing(True)
    p2.set_stage_timing(True)
    p1.stage_times()
    p2.stage_times()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        p1.execute(c, out=fk)
        p2.execute(f, out=co)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    s1 = {k: v / steps for k, v in p1.stage_times().items()}
    s2 = {k: v / steps for k, v in p2.stage_times().items()}
    r = torch.Generator(device="cpu")
    r.manual_seed(seed + 1)
    midx = torch.randint(0, math.prod(shape), (err_samples,), generator=r)
    pidx = torch.randint(0, M, (err_samples,), generator=r)
    e1 = acc.relerr(fk.reshape(-1)[midx.to(dev)], acc.type1_at(pts, c, 1, shape, midx))
    e2 = acc.relerr(co[pidx.to(dev)], acc.type2_at(pts, f, -1, pidx))
    i1, i2 = p1.info(), p2.info()
    out = {
        "ms_per_step": round(ms, 3), "nu_pts_per_s": 2 * M / (ms * 1e-3),
        "error": {"type1": e1, "type2": e2},
        "spread": _stage_roofline(M, dim, math.prod(i1["nf"]), s1["spread"], "spread"),
        "interp": _stage_roofline(M, dim, math.prod(i2["nf"]), s2["interp"], "interp"),
        "type1": {"ns": i1["ns"], "binsize": i1["binsize"], "nf": i1["nf"]},
        "type2": {"ns": i2["ns"], "binsize": i2["binsize"], "nf": i2["nf"]},
    }
    del p1, p2
    torch.cuda.empty_cache()
    return out


def differentiation_config(steps, err_samples):
    """3-D type 2 with gradients (fp32, tol 1e-6, N = 128^3): 1000 spokes x
    2000 radii on [0, pi] plus uniform jitter in [-1e-3, 1e-3], loss
    0.5 ||T2(f, x) - y||^2, forward + backward (gradients in f and in the
    three coordinate arrays) per step; the forward's error against the FP64
    direct sum at sampled points. (Finite-difference and FP64 gradient checks
    are in the GPU tests.)"""
    import torch

    from finufft_amd import accuracy as acc
    from finufft_amd.autograd import NufftType2

    dev = torch.device("cuda", 0)
    N, tol = 128, 1e-6
    sp = spokes_3d(1000, 2000, 7, dev)
    g = torch.Generator(device=dev)
    g.manual_seed(8)
    pts = [(p + (torch.rand(p.numel(), generator=g, device=dev) * 2e-3 - 1e-3)).contiguous()
           for p in sp]
    M = int(pts[0].numel())
    f = torch.complex(torch.randn((N, N, N), generator=g, device=dev),
                      torch.randn((N, N, N), generator=g, device=dev))
    y = torch.complex(torch.randn(M, generator=g, device=dev),
                      torch.randn(M, generator=g, device=dev))
    mod = NufftType2((N, N, N), eps=tol)

    def step():
        fr = f.clone().requires_grad_(True)
        xr = [p.clone().requires_grad_(True) for p in pts]
        cv = mod(fr, *xr)
        loss = 0.5 * (cv - y).abs().pow(2).sum()
        loss.backward()
        return cv.detach()

    cv = step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    r = torch.Generator(device="cpu")
    r.manual_seed(9)
    pidx = torch.randint(0, M, (err_samples,), generator=r)
    err = acc.relerr(cv[pidx.to(dev)], acc.type2_at(pts, f, -1, pidx))
    return {"workload": "3D type-2 fp32 tol 1e-6, N=128^3, M=2e6 (1000 spokes x 2000 radii, "
                        "jitter 1e-3), forward + backward (gradients in f and x)",
            "ms_per_step": round(ms, 3), "nu_pts_per_s": M / (ms * 1e-3), "error_forward": err}


def other_configs(steps, err_samples):
    """The other single-GPU workloads of the metric's configuration list:
    2-D type 1 + 2 at M = 1e7, N = 1000^2, fp32 tol 1e-5 on uniform and
    Gaussian-mixture points, and the 3-D differentiation problem."""
    import torch

    dev = torch.device("cuda", 0)
    out = {}
    g = torch.Generator(device=dev)
    g.manual_seed(10)
    M = 10_000_000
    uni = [torch.rand(M, generator=g, device=dev) * (2 * math.pi) - math.pi for _ in range(2)]
    out["2d_uniform"] = type12_config(uni, (1000, 1000), 1e-5, steps, err_samples, 11)
    out["2d_uniform"]["workload"] = "2D type1+type2 fp32 tol 1e-5, M=1e7 uniform, N=1000^2"
    del uni
    mix = gaussian_mixture_2d(M, 12, dev)
    out["2d_clustered"] = type12_config(mix, (1000, 1000), 1e-5, steps, err_samples, 13)
    out["2d_clustered"]["workload"] = ("2D type1+type2 fp32 tol 1e-5, M=1e7 from 16 Gaussians "
                                       "(sigma 0.05 pi), N=1000^2")
    del mix
    torch.cuda.empty_cache()
    out["3d_differentiation"] = differentiation_config(steps, err_samples)
    for k, v in out.items():
        log(f"{k}: {v['ms_per_step']} ms/step, {v['nu_pts_per_s']:.3e} NU pts/s")
    torch.cuda.empty_cache()
    return out


PROBE_MODES = ("ds_add_f32", "ds_add_f64", "ds_add_u32", "rmw_f32", "ds_read_b32",
               "ds_read_b32_uniform", "ds_read_b128", "ds_read_b128_uniform", "v_fma_f32",
               "ds_add_f32_rotating", "ds_add_f64_rotating", "ds_read_b64_stride1",
               "ds_read_b64_stride2", "ds_read_b64_stride16", "ds_read_b64_stride32",
               "ds_read_b64_stride64", "ds_add_f32_denorm_flush")


def lds_probe():
    """Wave-level operations per clock per CU of the LDS operations the
    spreading kernels choose between (and of v_fma_f32 for scale)."""
    import ctypes

    import torch

    from finufft_amd import _lib

    assert torch.cuda.is_available(), "the probe needs a GPU"
    props = torch.cuda.get_device_properties(0)
    ncu = props.multi_processor_count
    clk_khz = ctypes.c_int(0)
    hip = ctypes.CDLL("libamdhip64.so")
    # hipDeviceAttributeClockRate (kHz)
    hip.hipDeviceGetAttribute(ctypes.byref(clk_khz), 5, 0)
    clk = clk_khz.value * 1e3 if clk_khz.value > 0 else 2.4e9
    lib = _lib.lib()
    out = {"cus": ncu, "clock_hz": clk, "rates": {}}
    for mode, name in enumerate(PROBE_MODES):
        ms, ops = ctypes.c_double(0), ctypes.c_double(0)
        ier = lib.finufft_amd_lds_probe(mode, 8 * ncu, 2001, ctypes.byref(ms), ctypes.byref(ops))
        if ier != 0:
            raise RuntimeError(f"lds probe {name}: error {ier}")
        rate = ops.value / (ms.value * 1e-3) / clk / ncu
        out["rates"][name] = round(rate, 4)
        log(f"lds probe {name}: {ms.value:.3f} ms, {rate:.3f} wave ops/clk/CU")
    print(json.dumps(out), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--M", type=float, default=1e8)
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--cpu-sample", type=float, default=2e6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--err-samples", type=int, default=256)
    ap.add_argument("--method1", type=int, default=0, help="gpu_method of the type-1 plan (0 auto)")
    ap.add_argument("--method2", type=int, default=0, help="gpu_method of the type-2 plan (0 auto)")
    ap.add_argument("--binsize1", type=str, default="",
                    help="gpu_binsize x,y,z of the type-1 plan (default: the plan's choice)")
    ap.add_argument("--binsize2", type=str, default="",
                    help="gpu_binsize x,y,z of the type-2 plan (default: the plan's choice)")
    ap.add_argument("--batched", type=int, default=2000,
                    help="problems of the uneven small-problem comparison (0: skip)")
    ap.add_argument("--batched-ntr", type=int, default=32)
    ap.add_argument("--c3-compare", type=int, default=1,
                    help="time the headline problem with the default and the cost-model-"
                         "selected options on uniform and radial-spoke points (0: skip)")
    ap.add_argument("--batched-reps", type=int, default=3)
    ap.add_argument("--batched1", type=int, default=500,
                    help="problems of the type-1 batched comparison (0: skip)")
    ap.add_argument("--others", type=int, default=1,
                    help="also time the 2-D (uniform, clustered) and 3-D differentiation "
                         "workloads (0: skip)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the run, copy a tensor of this many MiB (a plain copy kernel of "
                         "known size, the reference point for HBM counter readings)")
    ap.add_argument("--lds-probe", action="store_true",
                    help="only measure LDS / VALU operation rates (wave instructions per clock "
                         "per CU) with the library's probe kernels and print them as JSON")
    args = ap.parse_args(argv)

    if args.lds_probe:
        return lds_probe()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(the library runs one plan on one GPU; it shards nothing across GPUs)",
              file=sys.stderr)
        return 2

    import torch

    from finufft_amd import Plan
    from finufft_amd import accuracy as acc

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda", 0)
    M, N, tol = int(args.M), args.N, args.tol
    shape = (N, N, N)

    g = torch.Generator(device=dev)
    g.manual_seed(0)
    pts = [(torch.rand(M, generator=g, device=dev, dtype=torch.float32) * (2 * math.pi) - math.pi)
           for _ in range(3)]
    g.manual_seed(1)
    c = torch.complex(torch.randn(M, generator=g, device=dev),
                      torch.randn(M, generator=g, device=dev))
    f = torch.complex(torch.randn(shape, generator=g, device=dev),
                      torch.randn(shape, generator=g, device=dev))
    fk_out = torch.empty(shape, dtype=torch.complex64, device=dev)
    c_out = torch.empty(M, dtype=torch.complex64, device=dev)
    torch.cuda.synchronize()
    log(f"inputs ready: M={M:.3g} N={N}^3")

    def bins(spec):
        if not spec:
            return {}
        b = [int(v) for v in spec.split(",")]
        return {"gpu_binsizex": b[0], "gpu_binsizey": b[1], "gpu_binsizez": b[2]}

    p1 = Plan(1, shape, eps=tol, dtype="complex64", isign=1, gpu_method=args.method1,
              **bins(args.binsize1))
    p2 = Plan(2, shape, eps=tol, dtype="complex64", isign=-1, gpu_method=args.method2,
              **bins(args.binsize2))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    p1.setpts(*pts)
    torch.cuda.synchronize()
    t_set1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    p2.setpts(*pts)
    torch.cuda.synchronize()
    t_set2 = time.perf_counter() - t0
    info1, info2 = p1.info(), p2.info()
    log(f"setpts: {t_set1*1e3:.1f} ms / {t_set2*1e3:.1f} ms; plan1 {info1}; plan2 {info2}")

    for _ in range(args.warmup):
        p1.execute(c, out=fk_out)
        p2.execute(f, out=c_out)
    torch.cuda.synchronize()
    log("warmup done")

    p1.set_stage_timing(True)
    p2.set_stage_timing(True)
    p1.stage_times()
    p2.stage_times()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        p1.execute(c, out=fk_out)
        p2.execute(f, out=c_out)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    s1 = {k: v / args.steps for k, v in p1.stage_times().items()}
    s2 = {k: v / args.steps for k, v in p2.stage_times().items()}
    p1.set_stage_timing(False)
    p2.set_stage_timing(False)
    ms_step = elapsed / args.steps * 1e3
    value = 2 * M / (elapsed / args.steps)
    log(f"timed {args.steps} steps: {ms_step:.2f} ms/step; type1 stages {s1}; type2 stages {s2}")

    # roofline of the dominant kernel
    nf1 = math.prod(info1["nf"])
    nf2 = math.prod(info2["nf"])
    spread_bytes = M * (4 + 4 * 3 + 8) + 8 * nf1
    interp_bytes = 8 * nf2 + M * (4 + 4 * 3) + 8 * M
    # launches per stage: a colored spread runs one launch per parity color
    # (plus one for dense-bin subproblems), interpolation one launch
    sp_launches = (2 ** 3 if info1["colored"] else 1) + (1 if info1["dense_subproblems"] else 0)
    cand = [("spread (type 1)", s1["spread"], spread_bytes, sp_launches),
            ("interp (type 2)", s2["interp"], interp_bytes, 1)]
    name, ms, nbytes, nl = max(cand, key=lambda t: t[1])
    achieved = nbytes / (ms * 1e-3) / 1e9
    traffic, source, limiter = measured_traffic(name, ms / nl)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": name, "launches": nl, "ms_per_launch": round(ms / nl, 4),
                "alg_bytes": nbytes // nl, "traffic_source": source, "limiter": limiter}

    # accuracy against the FP64 direct sum at sampled modes / points
    r = torch.Generator(device="cpu")
    r.manual_seed(2)
    S = args.err_samples
    midx = torch.randint(0, N ** 3, (S,), generator=r)
    pidx = torch.randint(0, M, (S,), generator=r)
    ref1 = acc.type1_at(pts, c, 1, shape, midx)
    e1 = acc.relerr(fk_out.reshape(-1)[midx.to(dev)], ref1)
    ref2 = acc.type2_at(pts, f, -1, pidx)
    e2 = acc.relerr(c_out[pidx.to(dev)], ref2)
    log(f"relerr type1 {e1:.3e} type2 {e2:.3e}")

    if args.copy_probe_mib:
        n = args.copy_probe_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {args.copy_probe_mib} MiB read + {args.copy_probe_mib} MiB written")
        del src, dst

    c3 = None
    if args.c3_compare:
        del p1, p2
        torch.cuda.empty_cache()
        steps = max(1, min(args.steps, 5))
        c3 = {"uniform": c3_comparison(pts, c, f, shape, tol, steps, args.err_samples)}
        del pts
        torch.cuda.empty_cache()
        nsp, per = (4096, 24414) if M == 100_000_000 else (4096, max(1, M // 4096))
        spk = spokes_3d(nsp, per, 6, dev)
        c3["radial"] = c3_comparison(spk, c[: spk[0].numel()], f, shape, tol, steps,
                                     args.err_samples)
        c3["radial"]["points"] = f"{nsp} spokes x {per} radii"
        del spk
        torch.cuda.empty_cache()
    else:
        del p1, p2, pts

    others = None
    if args.others:
        others = other_configs(max(1, min(args.steps, 5)), min(args.err_samples, 64))

    batched = None
    if args.batched > 0:
        del c, f, fk_out, c_out
        torch.cuda.empty_cache()
        batched = batched_comparison(args.batched, args.batched_ntr, 1e-5, args.batched_reps)
        torch.cuda.empty_cache()
    batched1 = None
    if args.batched1 > 0:
        batched1 = batched_type1_comparison(args.batched1, args.batched_ntr, 1e-5,
                                            args.batched_reps)
        torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu:
        threads = min(16, len(os.sched_getaffinity(0)))
        log(f"cpu baseline on {threads} threads ...")
        cpu = cpu_baseline(N, tol, M, int(args.cpu_sample), threads)
        log(f"cpu baseline {cpu['value']:.3e} NU pts/s")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "NU pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded uniform points, complex normal data, generated on the GPU)",
        "config": {
            "workload": f"3D type1+type2 fp32 tol {tol:g}, M={M:.0e} uniform on [-pi,pi)^3, "
                        f"N={N}^3 modes",
            "M": M, "N": [N, N, N], "tol": tol,
            "type1": {"gpu_method": info1["gpu_method"], "binsize": info1["binsize"],
                      "ns": info1["ns"], "nf": info1["nf"]},
            "type2": {"gpu_method": info2["gpu_method"], "binsize": info2["binsize"],
                      "ns": info2["ns"], "nf": info2["nf"]},
            "parallelism": "single GPU",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stages_ms": {"type1": {k: round(v, 4) for k, v in s1.items()},
                      "type2": {k: round(v, 4) for k, v in s2.items()}},
        "setpts_ms": {"type1": round(t_set1 * 1e3, 2), "type2": round(t_set2 * 1e3, 2)},
        "error": {"type1": e1, "type2": e2, "samples": S},
        "c3_plain_vs_selected": c3,
        "other_configs": others,
        "batched": batched,
        "batched_type1": batched1,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
