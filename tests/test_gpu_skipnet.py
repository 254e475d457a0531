"""Pass-limit network on the GPU (BF16 matrix cores) against its CPU oracle
(oracle/skipnet.py): forward outputs, loss gradients, training progress,
and the learned pass skip producing the plain encoder's codestream."""
import sys
from pathlib import Path

import numpy as np
import pytest

from bigwhoop_amd import bwc, skipnet

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import skipnet as ref  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if bwc.lib().bwc_amd_device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need the MI355X")


def _problem(seed, n):
    rng = np.random.default_rng(seed)
    X = rng.normal(0.2, 1.0, (n, 16)).astype(np.float32)
    w = ref.init_weights(seed, mu=rng.normal(0, 0.1, 16), sc=rng.uniform(0.7, 1.5, 16)).astype(np.float32)
    w[ref.B1:ref.W2] = rng.normal(0, 0.1, 64)
    w[ref.B2:ref.W3] = rng.normal(0, 0.1, 64)
    w[ref.B3] = 0.3
    t = rng.normal(1.0, 0.6, n).astype(np.float32)
    return X, w, t


@pytest.mark.parametrize("n", [1, 16, 37, 1000])
def test_forward_matches_oracle(n):
    X, w, _ = _problem(1, n)
    y = skipnet.gpu_predict(X, w)
    yr, _ = ref.forward(w, X, bf=True)
    # same bf16 roundings; fp32 sums in another order can move an
    # activation across a bf16 rounding boundary, hence the tolerance
    assert np.max(np.abs(y - yr)) <= 2e-2 * (1 + np.max(np.abs(yr))), np.max(np.abs(y - yr))
    assert np.median(np.abs(y - yr)) <= 2e-3 * (1 + np.max(np.abs(yr)))


@pytest.mark.parametrize("n", [16, 100, 256])
def test_gradient_matches_oracle(n):
    X, w, t = _problem(2, n)
    loss, g = skipnet.gpu_grad(X, t, w)
    lr, gr = ref.loss_and_grad(w, X, t, bf=True)
    assert abs(loss - lr) <= 1e-2 * lr
    for name, a, b in [("W1", ref.W1, ref.B1), ("b1", ref.B1, ref.W2), ("W2", ref.W2, ref.B2),
                       ("b2", ref.B2, ref.W3), ("w3", ref.W3, ref.B3), ("b3", ref.B3, ref.B3 + 1)]:
        d = np.linalg.norm(g[a:b] - gr[a:b])
        assert d <= 2e-2 * np.linalg.norm(gr[a:b]) + 1e-6, (name, d, np.linalg.norm(gr[a:b]))


def test_training_fits_a_known_function():
    rng = np.random.default_rng(5)
    X = rng.normal(0, 1, (4096, 16)).astype(np.float32)
    t = (1.5 + 0.8 * X[:, 0] - 0.5 * np.abs(X[:, 3]) + 0.3 * X[:, 7] * X[:, 8]).astype(np.float32)
    w0 = skipnet.init_weights(3, np.zeros(16, np.float32), np.ones(16, np.float32))
    y0 = skipnet.gpu_predict(X, w0)
    w, loss = skipnet.gpu_train(X, t, w0, epochs=30, lr=3e-3, seed=1)
    y = skipnet.gpu_predict(X, w)
    assert loss[-1] < 0.1 * loss[0], loss
    assert np.mean((y - t) ** 2) < 0.1 * np.mean((y0 - t) ** 2)
