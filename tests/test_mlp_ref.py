"""The CPU restatement of the proposal network (tests/mlp_ref.py) that the GPU network is checked against: its
backward equals central finite differences of its forward in FP64 (no bf16 rounding), its bf16 rounding is
round-to-nearest-even, and its encoding has the documented structure."""
import numpy as np

import mlp_ref as M


def _params(seed=0):
    rng = np.random.default_rng(seed)
    off, count = M.layout()
    p = rng.normal(0, 0.2, count)
    p[off["w1"][0]:off["w1"][0] + off["w1"][1]].reshape(64, 128)[63] = 0.0
    return p


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 0.0], np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6  # ties to even
    assert r[3] == 1.0 and r[4] == -2.5 and r[5] == 0.0
    y = np.random.default_rng(0).normal(size=1000).astype(np.float32)
    assert np.abs(M.bf16(y) - y).max() <= np.abs(y).max() * 2 ** -8


def test_encoding_structure():
    o = np.array([[0.0, 0.0, -4.0], [3.0, 1.0, 2.0]])
    d = np.array([[0.0, 0.0, 2.0], [-1.0, 0.0, 0.0]])
    x = M.encode(o, d)
    assert x.shape == (2, 64) and np.all(x[:, 63] == 0)
    np.testing.assert_allclose(x[0, 0:3], [0, 0, -1])
    np.testing.assert_allclose(x[0, 3:6], [0, 0, 0], atol=1e-12)          # closest point to the origin
    np.testing.assert_allclose(x[1, 3:6], [0, 0.25, 0.5])
    np.testing.assert_allclose(x[0, 36:39], [0, 0, 1])                     # unit direction
    assert np.all(np.abs(x[:, 6:36]) <= 1) and np.all(np.abs(x[:, 39:63]) <= 1)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    x = M.encode(rng.normal(size=(48, 3)) * 3, rng.normal(size=(48, 3)))
    target = np.exp(rng.normal(1.0, 0.3, 48))
    target[::7] = -1.0  # ignored rows
    p = _params(2)
    _, g = M.loss_and_grad(p, x, target, 0.7, rnd=False)
    off, count = M.layout()
    idx = np.concatenate([o + rng.choice(s, 8, replace=False) for o, s in off.values() if s > 1] + [[count - 1]])
    idx = idx[np.abs(g[idx]) > 0]
    assert len(idx) > 30
    h = 1e-6
    for i in idx:
        e = np.zeros(count)
        e[i] = h
        fd = (M.loss_and_grad(p + e, x, target, 0.7, rnd=False)[0] -
              M.loss_and_grad(p - e, x, target, 0.7, rnd=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(g[i]), (i, fd, g[i])
