"""GPU checks of the surrogate: gradients and forward against the FP64
oracle, BF16 inference against FP32, features against numpy, training, and
the hybrid's exact fallback.

Tolerances: FP32 gradients within 2e-4 relative (of the largest |g|) of the
FP64 oracle; FP32 forward within 1e-5; BF16 inference within 2e-2 of FP32
(8-bit mantissa activations); features within 1e-4 relative of FP64 numpy."""
import numpy as np
import pytest

from gmx_correlation_amd import _lib, ksg, surrogate, synth
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu(hiplib):
    if not _lib.probe():
        pytest.fail("no gfx950 device visible to the HIP library")
    return hiplib


def test_gradients_match_fp64_oracle(gpu):
    rng = np.random.default_rng(0)
    p = M.init_params(3).astype(np.float32)
    X = rng.standard_normal((300, 16)).astype(np.float32)
    y = rng.random(300).astype(np.float32)
    mlp = surrogate.Mlp(512)
    mlp.set_params(p)
    g, loss = mlp.grad(X, y)
    g64 = M.grad(p.astype(np.float64), X.astype(np.float64), y.astype(np.float64))
    assert np.max(np.abs(g - g64)) <= 2e-4 * np.max(np.abs(g64))
    assert abs(loss - M.loss(p.astype(np.float64), X.astype(np.float64), y)) <= 1e-5
    out, _ = mlp.forward(X, bf16=False)
    o64, _ = M.forward(p.astype(np.float64), X.astype(np.float64))
    assert np.max(np.abs(out - o64)) <= 1e-5
    ob, _ = mlp.forward(X, bf16=True)
    assert np.max(np.abs(ob - o64)) <= 2e-2 * max(1.0, np.max(np.abs(o64)))


def test_bf16_inference_ragged_sizes(gpu):
    rng = np.random.default_rng(1)
    mlp = surrogate.Mlp(256)
    mlp.set_params(M.init_params(5).astype(np.float32))
    for n in (1, 31, 33, 200, 1000):
        X = rng.standard_normal((n, 16)).astype(np.float32)
        ob, _ = mlp.forward(X, bf16=True)
        of, _ = mlp.forward(X, bf16=False)
        assert ob.shape == (n, 2)
        assert np.max(np.abs(ob - of)) <= 2e-2 * max(1.0, np.max(np.abs(of)))


def test_training_decreases_loss(gpu):
    rng = np.random.default_rng(2)
    X = rng.standard_normal((8192, 16)).astype(np.float32)
    X[:, 14:] = 0
    y = (0.5 + 0.3 * np.tanh(X[:, 0] - 0.5 * X[:, 3])).astype(np.float32)
    s = surrogate.Surrogate(4096, seed=0).fit(X, y, epochs=40, batch=1024, lr=3e-3)
    assert s.history[-1] < s.history[0] - 1.0
    mu, sigma, _ = s.predict(X)
    assert np.sqrt(np.mean((mu - y) ** 2)) < 0.02


@pytest.fixture(scope="module")
def small_plan(gpu):
    c = synth.generate(synth.TrajSpec(natoms=40, nframes=1000, nmodes=6, seed=21, n_switch=3,
                                      n_rotor=2))
    rows, psi = ksg.kraskov_prepare(ksg.center(c))
    plan = ksg.KsgPlan(40, 1000, 20)
    plan.load(rows, psi)
    yield plan, rows
    plan.close()


def test_features_match_numpy(small_plan, oracle_lib):
    plan, rows = small_plan
    a1, a2 = np.tril_indices(40, -1)
    feat, ms = surrogate.pair_features(plan, a1, a2, nsample=16)
    Xc = rows - rows.mean(axis=1, keepdims=True)
    C = Xc @ Xc.T / rows.shape[1]
    gmi = oracle_lib.gauss_matrix_from_cov(C, 40)[a1, a2]
    assert np.max(np.abs(feat[:, 0] - gmi)) <= 1e-4 * max(1.0, np.max(np.abs(gmi)))
    for q in (0, 17, 500):
        idx = np.r_[3 * a1[q]:3 * a1[q] + 3, 3 * a2[q]:3 * a2[q] + 3]
        R = np.corrcoef(rows[idx])
        assert np.allclose(np.sort(np.linalg.eigvalsh(R)), feat[q, 1:7], atol=1e-4)
    import scipy.stats
    kur = scipy.stats.kurtosis(rows, axis=1).reshape(40, 3).mean(axis=1)
    assert np.allclose(np.minimum(kur[a1], kur[a2]), feat[:, 7], atol=1e-4)
    assert np.all(feat[:, 9] <= feat[:, 11] + 1e-6) and np.all(feat[:, 9] > 0)
    assert np.all(feat[:, 14:] == 0)


def test_hybrid_fallback_is_exact_and_gains(small_plan):
    """Train on half of the pairs' exact r, run the hybrid on the rest:
    every fallback pair carries the exact KSG value, the audit measures the
    kept pairs, and the trained network beats the untrained feature map
    r(Gaussian MI) in RMSE against exact KSG (the gain of the learned part)."""
    plan, _ = small_plan
    a1, a2 = np.tril_indices(40, -1)
    r_exact = surrogate.mi_to_r(plan.run_pairs(a1, a2))
    feat, _ = surrogate.pair_features(plan, a1, a2, nsample=16)
    rng = np.random.default_rng(0)
    tr = rng.random(len(a1)) < 0.5
    model = surrogate.Surrogate(1024).fit(feat[tr], r_exact[tr], epochs=150, batch=128, lr=3e-3)
    res = surrogate.hybrid(plan, a1[~tr], a2[~tr], model, band=3e-2, audit=0.2, alpha=0.05,
                           seed=1)
    fb = res["fallback"]
    assert np.array_equal(res["r"][fb], r_exact[~tr][fb])
    kept = res["kept"]
    if kept.any():   # conformal coverage at level 0.95, on a few hundred pairs
        assert np.mean(np.abs(res["r"][kept] - r_exact[~tr][kept]) <= 3e-2) >= 0.85
    mu = res["mu"]
    rmse_net = np.sqrt(np.mean((mu - r_exact[~tr]) ** 2))
    rmse_feat = np.sqrt(np.mean((feat[~tr, 13] - r_exact[~tr]) ** 2))
    print(f"hybrid test: fallback {res['fallback_fraction']:.3f}, rmse net {rmse_net:.4f} vs "
          f"r(Gaussian MI) {rmse_feat:.4f}, audit max|dr| {res['audit_max_abs_dr']:.4f}")
    assert rmse_net < rmse_feat
