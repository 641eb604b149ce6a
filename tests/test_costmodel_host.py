"""Host logic of the learned cost model (no GPU): features, the rejection
table key and tables, the training-range fallback, the choice among the
allowed options, the options' plan arguments and the weights file. The
network's forward pass is the FP64 restatement (oracle/mlp.py) in place of
the HIP kernel, which tests/test_emu_mlp.py and tests/test_gpu_costmodel.py
check against it."""

import math

import numpy as np
import pytest

from finufft_amd import costmodel
from oracle import mlp as omlp

UNIFORM = (1.5, 1.3, 1.0)  # density statistics of a uniform point set


class HostModel(costmodel.CostModel):
    """The cost model with the oracle's forward pass."""

    def predict(self, feats):
        z = (np.atleast_2d(feats) - self.mean) / self.std
        y, _ = omlp.forward(self.params.astype(np.float64), z.astype(np.float64))
        return y


def _model(seed=0, lo=None, hi=None, classes=()):
    p = costmodel.init_params(seed)
    n = costmodel.NIN
    lo = np.full(n, -1e9, np.float32) if lo is None else lo
    hi = np.full(n, 1e9, np.float32) if hi is None else hi
    return HostModel(p, np.zeros(n), np.ones(n), lo, hi, {k: [] for k in classes})


def test_layout():
    assert costmodel.NIN == omlp.IN and costmodel.NOUT == omlp.OUT == len(costmodel.ARMS)
    assert costmodel.NPARAM == omlp.NPARAM


def test_features():
    f = costmodel.features(2, (256, 128), 2 ** 20, 32, 1e-5, "complex64", (4.0, 2.0, 0.5))
    assert f.dtype == np.float32 and f.shape == (costmodel.NIN,)
    assert list(f) == pytest.approx([2, 15, 20, 5, 5, 5, 0, 2, 2, 1, 0.5])
    g = costmodel.features(1, (64, 64, 64), 1, 1, 1e-12, "complex128", (0, 0, 0))
    assert g[0] == 3 and g[5] == pytest.approx(12) and g[6] == 1 and g[7] == 1
    assert np.all(np.isfinite(g))


def test_problem_class():
    assert costmodel.problem_class(2, 3, 1e-6, "complex64") == "2:3:complex64:6"
    assert costmodel.problem_class(1, 2, 2e-5, "complex128") == "1:2:complex128:5"


def test_plan_args():
    assert costmodel.plan_args({}, 1, (64, 64, 64), 1e-6) == (1e-6, {})
    eps, o = costmodel.plan_args({"eps_scale": 0.1}, 1, (64, 64, 64), 1e-6)
    assert eps == pytest.approx(1e-7) and o == {}
    eps, o = costmodel.plan_args({"bins": 0}, 1, (64, 64, 64), 1e-6)
    assert (o["gpu_binsizex"], o["gpu_binsizey"], o["gpu_binsizez"]) == (8, 8, 4)
    eps, o = costmodel.plan_args({"bins": 1}, 2, (100, 200), 1e-5)
    assert (o["gpu_binsizex"], o["gpu_binsizey"]) == costmodel.BIN_SHAPES[(2, 2)][1]
    assert "gpu_binsizez" not in o
    assert costmodel.plan_args({"gpu_method": 1, "gpu_maxbatchsize": 32}, 2, (8, 8), 1e-3) == (
        1e-3, {"gpu_method": 1, "gpu_maxbatchsize": 32})
    assert not costmodel.arm_valid({"bins": 0}, 3, 3) and costmodel.arm_valid({}, 3, 1)
    # output-driven spreading only for the types that spread
    assert not costmodel.arm_valid({"gpu_method": 3}, 2, 3)
    assert costmodel.arm_valid({"gpu_method": 3}, 1, 3) and costmodel.arm_valid({"gpu_method": 3}, 3, 2)


def test_choice_is_the_predicted_fastest_allowed_option():
    key = costmodel.problem_class(2, 2, 1e-5, "complex64")
    m = _model(1, classes=[key, costmodel.problem_class(2, 2, 1e-6, "complex64")])
    m.margin = 0.0
    prob = (2, (256, 256), 10 ** 5, 32, 1e-5, "complex64", UNIFORM)
    # the options that apply to 2-D type 2, in order of predicted time
    valid = [i for i in range(costmodel.NOUT) if costmodel.arm_valid(costmodel.ARMS[i], 2, 2)]
    y = m.predict(costmodel.features(*prob))[0]
    ranked = [int(i) for i in np.argsort(y) if int(i) in valid]
    best = ranked[0]
    assert m.choose(*prob) == best
    # a rejected option is skipped for its class only
    m.rejected[key] = [best]
    second = ranked[1]
    assert m.choose(*prob) == second
    other = (2, (256, 256), 10 ** 5, 32, 1e-6, "complex64", UNIFORM)
    y2 = m.predict(costmodel.features(*other))[0]
    assert m.choose(*other) == min(valid, key=lambda i: y2[i])
    # the candidates filter
    allowed = [i for i in range(costmodel.NOUT) if i not in (best, second)]
    assert m.choose(*prob, arms=allowed) in allowed
    eps, opts = m.plan_args(m.choose(*prob), 2, (256, 256), 1e-5)
    assert eps <= 1e-5


def test_margin_keeps_the_defaults_for_small_predicted_gains():
    """Another option replaces the defaults only when predicted faster by more
    than the margin (log time)."""
    key = costmodel.problem_class(2, 2, 1e-5, "complex64")
    m = _model(2, classes=[key])
    prob = (2, (256, 256), 10 ** 5, 32, 1e-5, "complex64", UNIFORM)
    y = m.predict(costmodel.features(*prob))[0]
    allowed = [i for i in range(costmodel.NOUT)
               if costmodel.arm_valid(costmodel.ARMS[i], 2, 2)]
    best = min(allowed, key=lambda i: y[i])
    gain = y[0] - y[best]
    assert best != 0 and gain > 0  # this seed's network prefers another option
    m.margin = gain + 1e-3
    assert m.choose(*prob) == 0
    m.margin = gain - 1e-3
    assert m.choose(*prob) == best
    assert costmodel.CostModel(m.params, m.mean, m.std, m.lo, m.hi,
                               meta={"margin": 0.3}).margin == 0.3


def test_unswept_class_or_default_miss_is_the_heuristic():
    key = costmodel.problem_class(1, 3, 1e-6, "complex64")
    m = _model(5, classes=[key])
    prob = (1, (64, 64, 64), 10 ** 6, 1, 1e-6, "complex64", UNIFORM)
    assert m.choose(*prob) is not None
    assert m.choose(1, (64, 64, 64), 10 ** 6, 1, 1e-9, "complex64", UNIFORM) is None
    m.default_misses = [key]
    assert m.choose(*prob) is None and m.misses_tol(1, 3, 1e-6, "complex64")
    assert m.plan_args(None, 1, (64, 64, 64), 1e-6) == (1e-6, {})


def test_bin_arms_only_where_defined():
    """The bin-shape options exist for 2-D and 3-D types 1 and 2 only."""
    key = costmodel.problem_class(1, 1, 1e-6, "complex64")
    m = _model(6, classes=[key])
    bins = [i for i, a in enumerate(costmodel.ARMS) if "bins" in a]
    for seed in range(20):
        m.params = costmodel.init_params(seed)
        got = m.choose(1, (4096,), 10 ** 5, 1, 1e-6, "complex64", UNIFORM)
        assert got not in bins


def test_outside_the_training_range_is_the_heuristic():
    f = costmodel.features(2, (256, 256), 10 ** 5, 32, 1e-5, "complex64", UNIFORM)
    m = _model(2, lo=f - 0.5, hi=f + 0.5, classes=[costmodel.problem_class(2, 2, 1e-5,
                                                                           "complex64")])
    assert m.choose(2, (256, 256), 10 ** 5, 32, 1e-5, "complex64", UNIFORM) is not None
    # 2^6 times more points: log2 M is 6 above the range, beyond its 10 % margin
    assert m.choose(2, (256, 256), 64 * 10 ** 5, 32, 1e-5, "complex64", UNIFORM) is None
    # a much denser point set: the density features leave the range
    assert m.choose(2, (256, 256), 10 ** 5, 32, 1e-5, "complex64", (200.0, 50.0, 0.1)) is None


def test_every_option_rejected_is_the_heuristic():
    key = costmodel.problem_class(1, 3, 1e-6, "complex64")
    m = _model(3, classes=[key])
    m.rejected[key] = list(range(costmodel.NOUT))
    assert m.choose(1, (64, 64, 64), 10 ** 6, 1, 1e-6, "complex64", UNIFORM) is None


def test_rejection_tables_use_tol_itself():
    """An option is rejected for a class when its error exceeds tol in any
    record (not a multiple of the default's error); a class where the default
    misses tol is listed."""
    inf = float("inf")
    base = {"ttype": 1, "n_modes": [32, 32, 32], "tol": 1e-6, "dtype": "complex64"}
    recs = [
        dict(base, ms=[1.0] * 8, err=[9e-7, 8e-7, 1.1e-6, 2e-6, 1e-7, 9e-7, 9e-7, inf]),
        dict(base, ms=[1.0] * 8, err=[9e-7] * 8),
        dict(base, tol=1e-5, ms=[1.0] * 8, err=[2e-5] + [1e-6] * 7),
    ]
    for r in recs:
        r["ok"] = [e <= r["tol"] for e in r["err"]]
    rej, misses = costmodel.rejection_tables(recs)
    assert sorted(rej["1:3:complex64:6"]) == [2, 3, 7]
    assert rej["1:3:complex64:5"] == [0] and misses == ["1:3:complex64:5"]


def test_weights_roundtrip(tmp_path):
    m = _model(4)
    m.rejected = {"2:2:complex64:5": [3]}
    m.default_misses = ["1:3:complex64:6"]
    m.meta = {"note": "x"}
    path = tmp_path / "w.json"
    m.save(path)
    back = costmodel.CostModel.load(path)
    assert np.array_equal(back.params, m.params) and back.rejected == m.rejected
    assert back.arms == m.arms and back.meta == m.meta
    assert back.default_misses == m.default_misses


def test_committed_weights_are_consistent():
    """The committed model loads, has the layout of the network, bans only
    options it has, and picks a valid option on a problem inside its range."""
    if not costmodel.WEIGHTS.exists():
        pytest.skip("no committed cost-model weights")
    m = costmodel.CostModel.load(costmodel.WEIGHTS)
    assert m.params.shape == (costmodel.NPARAM,) and m.arms == costmodel.ARMS
    assert all(np.isfinite(m.params)) and all(m.std > 0)
    for key, banned in m.rejected.items():
        ttype, dim, dtype, dec = key.split(":")
        assert int(ttype) in (1, 2, 3) and int(dim) in (1, 2, 3)
        assert all(0 <= i < len(m.arms) for i in banned)
    host = HostModel(m.params, m.mean, m.std, m.lo, m.hi, m.rejected, m.meta, m.arms,
                     m.default_misses)
    y = host.predict((m.lo + m.hi) / 2)
    assert y.shape == (1, costmodel.NOUT) and np.all(np.isfinite(y))
    got = host.choose(2, (256, 256), 50_000, 32, 1e-5, "complex64", UNIFORM)
    assert got is None or 0 <= got < len(m.arms)
    assert math.isfinite(float(m.meta.get("holdout_geomean_speedup", 1.0)))
