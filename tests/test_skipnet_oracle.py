"""The pass-limit network's CPU oracle, pinned by mathematics: its float64
gradient against central finite differences of its own loss, and its bf16
restatement of the GPU arithmetic close to the float64 one."""
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import skipnet as sn  # noqa: E402


def _problem(seed=0, n=48):
    rng = np.random.default_rng(seed)
    X = rng.normal(0.3, 1.0, (n, 16))
    w = sn.init_weights(seed, mu=rng.normal(0, 0.2, 16), sc=rng.uniform(0.5, 2, 16)).astype(np.float64)
    w[sn.B1:sn.W2] = rng.normal(0, 0.1, 64)
    w[sn.B2:sn.W3] = rng.normal(0, 0.1, 64)
    t = rng.normal(1.0, 0.5, n)
    return w, X, t


def test_gradient_matches_finite_differences():
    w, X, t = _problem()
    _, g = sn.loss_and_grad(w, X, t, bf=False)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(np.arange(sn.W1, sn.B1), 12), rng.choice(np.arange(sn.B1, sn.W2), 6),
                          rng.choice(np.arange(sn.W2, sn.B2), 12), rng.choice(np.arange(sn.B2, sn.W3), 6),
                          rng.choice(np.arange(sn.W3, sn.B3), 6), [sn.B3]])
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (sn.loss_and_grad(wp, X, t, bf=False)[0] - sn.loss_and_grad(wm, X, t, bf=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_asymmetric_weight_applies_below_target():
    # a network whose output is the constant b3: gradient is 2 c (b3 - t)
    w = np.zeros(sn.NPARAM)
    w[sn.SC:] = 1
    X = np.zeros((2, 16))
    w[sn.B3] = 1.0
    _, g = sn.loss_and_grad(w, X, np.array([2.0, 2.0]), bf=False)
    assert np.isclose(g[sn.B3], 2 * 4 * (1.0 - 2.0))
    _, g = sn.loss_and_grad(w, X, np.array([0.5, 0.5]), bf=False)
    assert np.isclose(g[sn.B3], 2 * 1 * (1.0 - 0.5))


def test_bf16_restatement_close_to_float64():
    w, X, t = _problem(seed=3, n=64)
    y64, _ = sn.forward(w, X, bf=False)
    y16, _ = sn.forward(w.astype(np.float32), X.astype(np.float32), bf=True)
    assert np.max(np.abs(y16 - y64)) < 0.05 * (1 + np.max(np.abs(y64)))
    l64, g64 = sn.loss_and_grad(w, X, t, bf=False)
    l16, g16 = sn.loss_and_grad(w.astype(np.float32), X.astype(np.float32), t.astype(np.float32), bf=True)
    assert abs(l16 - l64) < 0.05 * l64
    assert np.linalg.norm(g16 - g64) < 0.05 * np.linalg.norm(g64)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 0.0], np.float32)
    r = sn.bf16(x)
    # 1 + 2^-8 is exactly halfway between bf16 neighbours 1 and 1 + 2^-7: ties to even (1.0)
    assert r.tolist() == [1.0, 1.0, 1.0, 1.0 + 2 ** -7, -2.5, 0.0]


def test_limits_clamp_to_complete_coding():
    y = np.array([-1.0, 0.1, 1.0, 100.0])
    K = np.array([5, 5, 1, 3])
    assert sn.limits(y, K, 2).tolist() == [2, 3, 1, 7]


def test_committed_networks_and_choice():
    """Both committed weight sets load (5345 float32 values, finite), and the
    selector keeps the headline population's network for 3-D fields up to
    2 bpv and takes the wide one otherwise."""
    from bigwhoop_amd import skipnet

    a, b = skipnet.default_weights(), skipnet.wide_weights()
    assert a.size == b.size == skipnet.NPARAM
    assert np.isfinite(a).all() and np.isfinite(b).all()
    assert not np.array_equal(a, b)
    w, m = skipnet.choose(3, 2.0)
    assert np.array_equal(w, a) and m == 6
    for nd, r in ((3, 4.0), (1, 1.0), (2, 0.5), (4, 2.0)):
        w, m = skipnet.choose(nd, r)
        assert np.array_equal(w, b) and m == 8
