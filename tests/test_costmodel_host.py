"""Host logic of the learned cost model (no GPU): features, the rejection
table key, the training-range fallback, the choice among the allowed options
and the weights file. The network's forward pass is the FP64 restatement
(oracle/mlp.py) in place of the HIP kernel, which tests/test_emu_mlp.py and
tests/test_gpu_costmodel.py check against it."""

import math

import numpy as np
import pytest

from finufft_amd import costmodel
from oracle import mlp as omlp


class HostModel(costmodel.CostModel):
    """The cost model with the oracle's forward pass."""

    def predict(self, feats):
        z = (np.atleast_2d(feats) - self.mean) / self.std
        y, _ = omlp.forward(self.params.astype(np.float64), z.astype(np.float64))
        return y


def _model(seed=0, lo=None, hi=None):
    p = costmodel.init_params(seed)
    n = costmodel.NIN
    lo = np.full(n, -1e9, np.float32) if lo is None else lo
    hi = np.full(n, 1e9, np.float32) if hi is None else hi
    return HostModel(p, np.zeros(n), np.ones(n), lo, hi)


def test_features():
    f = costmodel.features(2, (256, 128), 2 ** 20, 32, 1e-5, "complex64")
    assert f.dtype == np.float32 and f.shape == (costmodel.NIN,)
    assert list(f) == pytest.approx([2, 15, 20, 5, 5, 5, 0, 2])
    g = costmodel.features(1, (64, 64, 64), 1, 1, 1e-12, "complex128")
    assert g[0] == 3 and g[5] == pytest.approx(12) and g[6] == 1 and g[7] == 1


def test_problem_class():
    assert costmodel.problem_class(2, 3, 1e-6, "complex64") == "2:3:complex64:6"
    assert costmodel.problem_class(1, 2, 2e-5, "complex128") == "1:2:complex128:5"


def test_choice_is_the_predicted_fastest_allowed_option():
    m = _model(1)
    prob = (2, (256, 256), 10 ** 5, 32, 1e-5, "complex64")
    y = m.predict(costmodel.features(*prob))[0]
    best = int(np.argmin(y))
    assert m.choose(*prob, methods=(1, 2)) == costmodel.ARMS[best]
    # a rejected option is skipped for its class only
    m.rejected = {costmodel.problem_class(2, 2, 1e-5, "complex64"): [best]}
    second = [int(i) for i in np.argsort(y) if int(i) != best][0]
    assert m.choose(*prob) == costmodel.ARMS[second]
    assert m.choose(2, (256, 256), 10 ** 5, 32, 1e-6, "complex64") == costmodel.ARMS[best]
    # the methods filter
    allowed = [i for i, a in enumerate(costmodel.ARMS) if a["gpu_method"] == 2
               and i != best]
    got = m.choose(*prob, methods=(2,))
    assert got["gpu_method"] == 2 and costmodel.ARMS.index(got) in allowed


def test_outside_the_training_range_is_the_heuristic():
    f = costmodel.features(2, (256, 256), 10 ** 5, 32, 1e-5, "complex64")
    m = _model(2, lo=f - 0.5, hi=f + 0.5)
    assert m.choose(2, (256, 256), 10 ** 5, 32, 1e-5, "complex64") != {}
    # 2^6 times more points: log2 M is 6 above the range, beyond its 10 % margin
    assert m.choose(2, (256, 256), 64 * 10 ** 5, 32, 1e-5, "complex64") == {}


def test_every_option_rejected_is_the_heuristic():
    m = _model(3)
    m.rejected = {costmodel.problem_class(1, 3, 1e-6, "complex64"):
                  list(range(len(costmodel.ARMS)))}
    assert m.choose(1, (64, 64, 64), 10 ** 6, 1, 1e-6, "complex64") == {}


def test_weights_roundtrip(tmp_path):
    m = _model(4)
    m.rejected = {"2:2:complex64:5": [3]}
    m.meta = {"note": "x"}
    path = tmp_path / "w.json"
    m.save(path)
    back = costmodel.CostModel.load(path)
    assert np.array_equal(back.params, m.params) and back.rejected == m.rejected
    assert back.arms == m.arms and back.meta == m.meta


def test_committed_weights_are_consistent():
    """The committed model loads, has the layout of the network, bans only
    options it has, and picks a valid option on a problem inside its range."""
    if not costmodel.WEIGHTS.exists():
        pytest.skip("no committed cost-model weights")
    m = costmodel.CostModel.load(costmodel.WEIGHTS)
    assert m.params.shape == (costmodel.NPARAM,)
    assert all(np.isfinite(m.params)) and all(m.std > 0)
    for key, banned in m.rejected.items():
        ttype, dim, dtype, dec = key.split(":")
        assert int(ttype) in (1, 2, 3) and int(dim) in (1, 2, 3)
        assert all(0 <= i < len(m.arms) for i in banned)
    host = HostModel(m.params, m.mean, m.std, m.lo, m.hi, m.rejected, m.meta, m.arms)
    f_mid = (m.lo + m.hi) / 2
    y = host.predict(f_mid)
    assert y.shape == (1, costmodel.NOUT) and np.all(np.isfinite(y))
    got = host.choose(2, (256, 256), 50_000, 32, 1e-5, "complex64")
    assert got == {} or got in m.arms
    assert math.isfinite(float(m.meta.get("holdout_geomean_speedup", 1.0)))
