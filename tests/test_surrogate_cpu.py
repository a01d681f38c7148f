"""Pins the FP64 surrogate oracle by finite differences (no GPU)."""
import numpy as np

from oracle import mlp_oracle as M


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init_params(1)
    X = rng.standard_normal((37, 16))
    y = rng.random(37)
    g = M.grad(p, X, y)
    idx = rng.choice(M.NPARAMS, 60, replace=False)
    h = 1e-6
    for t in idx:
        e = np.zeros(M.NPARAMS); e[t] = h
        fd = (M.loss(p + e, X, y) - M.loss(p - e, X, y)) / (2 * h)
        assert abs(fd - g[t]) <= 1e-7 + 1e-5 * abs(fd), (t, fd, g[t])


def test_oracle_loss_closed_form():
    """Zero weights: mu = b3[0], s = b3[1]; loss = 0.5(s + mean (y-mu)^2 e^-s)."""
    p = np.zeros(M.NPARAMS)
    p[-2], p[-1] = 0.3, -1.0
    y = np.array([0.1, 0.5, 0.9])
    X = np.zeros((3, 16))
    want = 0.5 * (-1.0 + np.mean((y - 0.3) ** 2) * np.exp(1.0))
    assert abs(M.loss(p, X, y) - want) < 1e-15


def test_all_pairs_order_check():
    """The hybrid hands a fallback of every pair, in run_all's order, to
    run_all; any other list goes through run_pairs."""
    from gmx_correlation_amd.surrogate import _all_pairs_in_order
    a1, a2 = np.tril_indices(6, -1)
    assert _all_pairs_in_order(a1.astype(np.int32), a2.astype(np.int32), 6)
    assert not _all_pairs_in_order(a1[::-1], a2[::-1], 6)
    assert not _all_pairs_in_order(a1[:-1], a2[:-1], 6)
    assert not _all_pairs_in_order(a2, a1, 6)
