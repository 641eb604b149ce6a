"""The learned cost model: its HIP network kernels against the FP64 numpy
oracle (forward, gradient, training), option choice and heuristic fallback,
and what the committed model gains over the default options."""

import math

import numpy as np
import pytest

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


def _data(n, seed):
    r = np.random.default_rng(seed)
    p = r.normal(0, 0.3, omlp.NPARAM).astype(np.float32)
    x = r.normal(0, 1, (n, omlp.IN)).astype(np.float32)
    t = r.normal(0, 1, (n, omlp.OUT)).astype(np.float32)
    m = (r.uniform(size=(n, omlp.OUT)) > 0.25).astype(np.float32)
    return p, x, t, m


def test_layout_matches_library():
    from finufft_amd import costmodel

    assert costmodel.NPARAM == omlp.NPARAM


def test_forward_and_grad_vs_oracle(torch_gpu):
    from finufft_amd import costmodel

    torch = torch_gpu
    p, x, t, m = _data(300, 1)
    dev = torch.device("cuda", 0)
    P, X, T, Mk = (torch.from_numpy(a).to(dev) for a in (p, x, t, m))
    y = costmodel.gpu_forward(P, X).cpu().numpy()
    y_ref, _ = omlp.forward(p.astype(np.float64), x.astype(np.float64))
    assert np.max(np.abs(y - y_ref)) < 1e-4 * max(1.0, np.max(np.abs(y_ref)))
    loss, g = costmodel.gpu_grad(P, X, T, Mk)
    g_ref = omlp.grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64),
                      m.astype(np.float64))
    l_ref = omlp.loss(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64),
                      m.astype(np.float64))
    assert abs(float(loss) - l_ref) < 1e-4 * max(1.0, l_ref)
    g = g.cpu().numpy()
    assert np.linalg.norm(g - g_ref) < 1e-4 * np.linalg.norm(g_ref)


def test_training_fits(torch_gpu):
    """Full-batch Adam with the library's kernels fits a smooth target."""
    from finufft_amd import costmodel

    r = np.random.default_rng(2)
    x = r.normal(0, 1, (400, omlp.IN)).astype(np.float32)
    t = np.stack([np.sin(x[:, 0]) + 0.5 * x[:, 1], x[:, 2] ** 2 * 0.3, x[:, 3] - x[:, 4],
                  np.cos(x[:, 5])], axis=1).astype(np.float32)
    m = np.ones_like(t)
    l0 = omlp.loss(costmodel.init_params(0).astype(np.float64), x.astype(np.float64),
                   t.astype(np.float64), m)
    params, loss = costmodel.train(x, t, m, epochs=1500, lr=3e-3)
    l_ref = omlp.loss(params.astype(np.float64), x.astype(np.float64), t.astype(np.float64), m)
    print(l0, loss, l_ref)
    assert l_ref < 0.1 * l0
    assert abs(loss - l_ref) < 0.05 * l0  # the last reported loss is that of the last step


def test_choose_and_fallback(torch_gpu):
    from finufft_amd import costmodel

    p = costmodel.init_params(3)
    lo = np.zeros(costmodel.NIN, np.float32)
    hi = np.ones(costmodel.NIN, np.float32)
    model = costmodel.CostModel(p, np.zeros(costmodel.NIN), np.ones(costmodel.NIN), lo, hi)
    # far outside the (tiny) training range: the library's heuristic
    assert model.choose(2, (256, 256), 10 ** 5, 32, 1e-5, "complex64") == {}
    f = costmodel.features(2, (256, 256), 10 ** 5, 32, 1e-5, "complex64")
    model.lo, model.hi = f - 1, f + 1
    choice = model.choose(2, (256, 256), 10 ** 5, 32, 1e-5, "complex64")
    assert choice in costmodel.ARMS
    # a rejected option is never chosen
    best = costmodel.ARMS.index(choice)
    model.rejected = {costmodel.problem_class(2, 2, 1e-5, "complex64"): [best]}
    assert model.choose(2, (256, 256), 10 ** 5, 32, 1e-5, "complex64") != choice


def test_committed_model_gain(torch_gpu):
    """On problems the model was not trained on, the model's options are not
    slower in total than the default ones and keep the error."""
    from finufft_amd import costmodel

    model = costmodel.default_model()
    if model is None:
        pytest.skip("no committed cost-model weights")
    cfgs = [(2, (256, 256), 50_000, 32, 1e-5, "complex64"),
            (2, (480, 480), 3_000, 32, 1e-5, "complex64"),
            (2, (48, 48), 40_000, 32, 1e-5, "complex64"),
            (1, (64, 64, 64), 1_000_000, 1, 1e-6, "complex64"),
            (1, (40, 40, 40), 20_000, 32, 1e-3, "complex64"),
            (2, (100, 100), 8_000, 32, 1e-6, "complex64"),
            (1, (200, 200), 150_000, 4, 1e-9, "complex128")]
    t_def = t_mod = 0.0
    log_gain = 0.0
    for i, c in enumerate(cfgs):
        opts = model.choose(*c) or costmodel.ARMS[0]
        ms0, e0 = costmodel.time_arm(*c, costmodel.ARMS[0], seed=100 + i)
        ms1, e1 = costmodel.time_arm(*c, opts, seed=100 + i)
        print(c, opts, ms0, ms1, e0, e1)
        assert e1 <= max(c[4], 1.5 * e0) * 1.5
        t_def += ms0
        t_mod += ms1
        log_gain += math.log(ms0 / ms1)
    gmean = math.exp(log_gain / len(cfgs))
    print("default", t_def, "model", t_mod, "total gain", t_def / t_mod, "per-problem gain", gmean)
    assert t_mod <= 1.1 * t_def
    # the gain: many-vector problems of these sizes run faster in one pass
    assert gmean > 1.05
