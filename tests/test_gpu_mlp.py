"""The learned cost model's HIP kernels (BF16 matrix cores, FP32
accumulation) against an FP32 CPU reference that rounds the same matmul
operands to bfloat16. Tolerances: 2e-3 relative on the forward, and on each
gradient block 2e-2 of that block's largest magnitude (only summation order
differs; ReLU masks at exactly-zero pre-activations are measure-zero)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_forward_matches_reference(gat, gpu):
    from graphautotune_amd.tuning import mlp
    for in_dim, hidden, batch in ((96, 256, 37), (40, 64, 130), (96, 256, 1)):
        m = mlp.MLPCostModel(in_dim, hidden, seed=3)
        rng = np.random.default_rng(1)
        X = rng.normal(size=(batch, in_dim)).astype(np.float32)
        y = m.forward(X)
        ref, _ = mlp.reference_forward(m.params(), X.astype(np.float64), in_dim, hidden)
        np.testing.assert_allclose(y, ref, rtol=2e-3, atol=2e-3 * np.abs(ref).max())


def test_gradients_match_reference(gat, gpu):
    from graphautotune_amd.tuning import mlp
    in_dim, hidden, batch = 96, 256, 50
    m = mlp.MLPCostModel(in_dim, hidden, seed=5)
    rng = np.random.default_rng(2)
    X = rng.normal(size=(batch, in_dim)).astype(np.float32)
    t = rng.normal(size=batch).astype(np.float32)
    g, loss = m.gradients(X, t)
    rg, rloss = mlp.reference_gradients(m.params(), X.astype(np.float64), t.astype(np.float64), in_dim, hidden)
    assert abs(loss - rloss) <= 2e-3 * max(1.0, abs(rloss))
    sizes = [in_dim * hidden, hidden, hidden * hidden, hidden, hidden, 1]
    o = 0
    for name, n in zip(["W1", "b1", "W2", "b2", "W3", "b3"], sizes):
        a, b = g[o:o + n], rg[o:o + n]
        scale = max(np.abs(b).max(), 1e-12)
        assert np.abs(a - b).max() <= 2e-2 * scale, (name, np.abs(a - b).max(), scale)
        o += n


def test_training_fits_and_ranks(gat, gpu):
    """A smooth synthetic runtime surface: after GPU training the model's
    ranking of held-out points correlates with the truth."""
    from graphautotune_amd.tuning import mlp
    rng = np.random.default_rng(0)
    X = rng.uniform(size=(400, 96)).astype(np.float32)
    secs = np.exp(X[:, 0] - X[:, 1] + 0.5 * X[:, 2] * X[:, 3] + 0.3 * np.sin(3 * X[:, 4]))
    m = mlp.MLPCostModel(96, 256, seed=1)
    l0 = m.fit(X[:300], secs[:300], steps=1, lr=1e-3)
    l1 = m.fit(X[:300], secs[:300], steps=600, lr=1e-3)
    assert l1 < 0.5 * l0
    pred = m.predict(X[300:])
    from graphautotune_amd.tuning.study import spearman
    assert spearman(pred, secs[300:]) > 0.7
