"""Cost-model MLP: the float64 reference is pinned by finite differences
(CPU); the GPU forward/backward kernels are checked against it (gpu)."""
import numpy as np
import pytest

from nqueens import costmodel
from oracle import mlp_ref


def _data(m=96, seed=3):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, (m, 24)).astype(np.float32)
    t = rng.normal(0, 1, m).astype(np.float32)
    return x, t


def test_reference_gradient_finite_differences():
    x, t = _data(32)
    p = costmodel.init_params(5).astype(np.float64)
    p[1536:1600] = 0.05          # non-zero biases so that every unit is active somewhere
    _, g = mlp_ref.loss_grad(p, x, t)
    rng = np.random.default_rng(0)
    for k in rng.choice(p.size, 60, replace=False).tolist() + [p.size - 1, 0, 1536, 5696]:
        h = 1e-6
        a, b = p.copy(), p.copy()
        a[k] += h
        b[k] -= h
        fd = (mlp_ref.loss(a, x, t) - mlp_ref.loss(b, x, t)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), k


def test_init_and_spearman():
    p = costmodel.init_params(0)
    assert p.dtype == np.float32 and p.size == mlp_ref.NPARAM
    assert costmodel.spearman(np.arange(10.0), np.arange(10.0) ** 3) == pytest.approx(1.0)
    assert costmodel.spearman(np.arange(10.0), -np.arange(10.0)) == pytest.approx(-1.0)


@pytest.mark.gpu
def test_gpu_forward_matches_reference(ctx):
    x, _ = _data(300)
    p = costmodel.init_params(1)
    y = ctx.mlp_forward(p, x)
    ref, _, _ = mlp_ref.forward(p, x)
    np.testing.assert_allclose(y, ref, rtol=1e-4, atol=1e-5)


@pytest.mark.gpu
def test_gpu_gradient_matches_reference(ctx):
    x, t = _data(257)
    p = costmodel.init_params(2)
    p[1536:1600] = 0.05
    loss, g = ctx.mlp_loss_grad(p, x, t)
    rloss, rg = mlp_ref.loss_grad(p, x, t)
    assert loss == pytest.approx(rloss, rel=1e-4)
    # FP32 accumulation over 257 rows: relative 1e-4 of the gradient scale
    np.testing.assert_allclose(g, rg, rtol=1e-3, atol=1e-4 * np.abs(rg).max())


@pytest.mark.gpu
def test_gpu_features_shape_and_range(ctx):
    from nqueens import gen_constellations
    rec = gen_constellations(12, 5)
    ctx.load(12, rec)
    f = ctx.features()
    assert f.shape == (rec.shape[0], 24) and np.isfinite(f).all()
    assert (f >= 0).all() and (f <= 2).all()


@pytest.mark.gpu
def test_gpu_training_learns_node_counts(ctx):
    """Trained on N=13..15 node counts, the model ranks N=16's constellations
    by work far better than chance and better than any single feature."""
    X, y, ns = costmodel.collect(ctx, Ns=(13, 14, 15, 16))
    tr, te = ns < 16, ns == 16
    model, rep = costmodel.train(ctx, X[tr], y[tr], epochs=25, batch=256, lr=2e-3, holdout=0.1)
    assert rep["train_loss_last"] < rep["train_loss_first"]
    pred = model.predict_features(ctx, np.ascontiguousarray(X[te]))
    rho = costmodel.spearman(pred, y[te])
    best_single = max(abs(costmodel.spearman(X[te][:, k], y[te])) for k in range(24)
                      if np.std(X[te][:, k]) > 0)
    print(f"spearman={rho:.3f} best single feature={best_single:.3f} mae={np.mean(np.abs(pred - y[te])):.3f}")
    assert rho > 0.7
    assert rho > best_single
