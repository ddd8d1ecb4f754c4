"""The FP64 network oracle: back-propagation pinned by central finite differences."""
import numpy as np

from oracle import mlp_oracle as M


def test_gradients_match_finite_differences():
    dims = (5, 7, 3, 4)
    rng = np.random.default_rng(1)
    p = M.init_params(dims, seed=2) + rng.normal(0, 0.05, size=M.init_params(dims).size)
    x = rng.normal(size=(11, 5))
    t = rng.normal(size=(11, 4))
    norm = np.concatenate([rng.normal(size=5), rng.uniform(0.5, 2, 5), rng.normal(size=4), rng.uniform(0.5, 2, 4)])
    loss, g, _ = M.loss_and_grad(p, x, t, norm, dims)
    eps = 1e-6
    idx = rng.choice(p.size, 60, replace=False)
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (M.loss_and_grad(pp, x, t, norm, dims)[0] - M.loss_and_grad(pm, x, t, norm, dims)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd))


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.001953125, 3.14159265, -2.5e-3, 65504.0])
    r = M.bf16_round(v)
    assert r[0] == 1.0
    assert r[1] == 1.0  # 1 + 2^-8 is half-way between 1 and 1 + 2^-7: ties go to the even mantissa
    assert r[2] == 1.0
    assert M.bf16_round(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6  # tie, odd neighbour -> up
    assert abs(r[3] - np.pi) < np.pi * 2 ** -8
    assert np.all(np.abs(r - v) <= np.abs(v) * 2 ** -8)
