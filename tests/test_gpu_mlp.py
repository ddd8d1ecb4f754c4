"""BF16-MFMA surrogate network against its FP64 oracle (needs an MI355X)."""
import numpy as np
import pytest

from oracle import mlp_oracle as M
from portal_amd.api import Portal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net():
    pt = Portal(0)
    dims = (20, 256, 4, 19)
    pt.mlp_create(*dims, seed=7)
    yield pt, dims
    pt.close()


def _data(dims, B, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(B, dims[0])) * 2 + 1
    t = np.tanh(x[:, : dims[3]] * 0.7) + 0.1 * x[:, 1:2]
    norm = np.concatenate([np.full(dims[0], 1.0), np.full(dims[0], 2.0), np.zeros(dims[3]), np.ones(dims[3])])
    return x, t, norm


def test_forward_matches_oracle(net):
    pt, dims = net
    x, t, norm = _data(dims, 1000, 1)
    pt.mlp_set_params(None, norm)
    p, nm = pt.mlp_get_params()
    y = pt.mlp_forward(x)
    yo, _, _ = M.forward(p.astype(np.float64), x, nm.astype(np.float64), dims, bf16=True)
    rel = np.linalg.norm(y - yo) / np.linalg.norm(yo)
    print("forward rel err", rel)
    assert rel < 2e-3


def test_gradients_match_oracle(net):
    pt, dims = net
    x, t, norm = _data(dims, 4096, 2)
    pt.mlp_set_params(None, norm)
    p, nm = pt.mlp_get_params()
    loss, g = pt.mlp_train_step(x, t, 0.0, want_grads=True)
    lo, go, _ = M.loss_and_grad(p.astype(np.float64), x, t, nm.astype(np.float64), dims, bf16=True)
    assert loss == pytest.approx(lo, rel=2e-3)
    o = 0
    for a, b in M.layer_dims(*dims):
        for n in (a * b, b):
            gg, oo = g[o:o + n], go[o:o + n]
            rel = np.linalg.norm(gg - oo) / max(np.linalg.norm(oo), 1e-30)
            print(f"block of {n}: grad rel err {rel:.3e}")
            assert rel < 5e-3
            o += n
    # the parameters did not move with lr = 0
    p2, _ = pt.mlp_get_params()
    np.testing.assert_array_equal(p, p2)


def test_large_batch_forward_and_gradients(net):
    """65,536 rows: the forward products and the back-propagated dZ take the 128 x 128 tile
    (the small batches above take the 64 x 64 one), with 16-byte, row-vector and element
    operand loads all in use; errors to the same bounds."""
    pt, dims = net
    x, t, norm = _data(dims, 65536, 4)
    pt.mlp_set_params(M.init_params(dims, seed=9).astype(np.float32), norm)
    p, nm = pt.mlp_get_params()
    y = pt.mlp_forward(x)
    yo, _, _ = M.forward(p.astype(np.float64), x, nm.astype(np.float64), dims, bf16=True)
    rel = np.linalg.norm(y - yo) / np.linalg.norm(yo)
    print("65,536-row forward rel err", rel)
    assert rel < 2e-3
    loss, g = pt.mlp_train_step(x, t, 0.0, want_grads=True)
    lo, go, _ = M.loss_and_grad(p.astype(np.float64), x, t, nm.astype(np.float64), dims, bf16=True)
    assert loss == pytest.approx(lo, rel=2e-3)
    rel = np.linalg.norm(g - go) / np.linalg.norm(go)
    print("65,536-row gradient rel err", rel)
    assert rel < 5e-3


def test_training_reduces_loss(net):
    pt, dims = net
    x, t, norm = _data(dims, 16384, 3)
    pt.mlp_set_params(M.init_params(dims, seed=5).astype(np.float32), norm)
    hist, ms = pt.mlp_train(x, t, epochs=15, batch=1024, lr=2e-3, seed=1)
    print("loss history", hist, "ms", ms)
    assert hist[-1] < 0.2 * hist[0]
