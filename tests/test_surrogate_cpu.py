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


def test_calibrated_threshold_bounds_the_kept_pairs():
    """Errors grow with sigma faster than a fixed multiple of it (the network
    is overconfident at large sigma): the threshold is the one where the
    kept pairs' 1 - alpha error quantile reaches the band, which for
    |err| = sigma^2 * |z| is tau with tau^2 * z_(1-alpha/2) = band; fresh
    pairs kept by it meet the band in at least 1 - alpha of cases."""
    from scipy.stats import norm
    from gmx_correlation_amd import surrogate
    rng = np.random.default_rng(0)
    n, alpha, band = 40000, 0.05, 1e-2
    sig = rng.uniform(0.0, 0.3, n)
    res = sig ** 2 * rng.standard_normal(n)
    tau = surrogate.calibrate_threshold(res, sig, band, alpha)
    # subset sigma <= tau: sigma uniform on [0, tau], |err| = sigma^2 |z|;
    # its 95% quantile is band; solve by Monte Carlo on the same law
    sub = np.sort(np.abs(np.linspace(0, tau, 200001)[:, None] ** 2 *
                         norm.ppf(np.linspace(0.0005, 0.9995, 1000))[None, :]).ravel())
    assert abs(sub[int(0.95 * len(sub))] - band) / band < 0.03
    sig2 = rng.uniform(0.0, 0.3, n)
    res2 = sig2 ** 2 * rng.standard_normal(n)
    keep = sig2 <= tau
    assert np.mean(np.abs(res2[keep]) <= band) >= 1 - alpha - 0.005
    # no qualifying threshold: nothing kept
    assert surrogate.calibrate_threshold(np.ones(500), np.ones(500), 1e-3, 0.01) == float("-inf")
    # too few calibration pairs for the level
    assert surrogate.calibrate_threshold(np.zeros(50), np.ones(50), 1e-3, 0.01) == float("-inf")


class _FakePlan:
    """Exact r = known values; records which pairs were run exactly."""
    def __init__(self, mi):
        self.mi, self.natoms, self.ran = mi, 10**6, []

    def run_pairs(self, a1, a2):
        self.ran.append(np.asarray(a1).copy())
        return self.mi[np.asarray(a1)]

    def run_all(self):
        raise AssertionError("not an all-pairs list")

    def stats(self):
        return {"ms_pairs": 0.0, "ms_fixup": 0.0}


class _FakeModel:
    def __init__(self, mu, sigma):
        self.mu, self.sigma = mu, sigma

    def predict(self, feat, bf16=True):
        return self.mu, self.sigma, 0.0


def _run_hybrid(monkeypatch, r_exact, mu, sigma, **kw):
    from gmx_correlation_amd import surrogate
    n = len(r_exact)
    monkeypatch.setattr(surrogate, "pair_features", lambda plan, a1, a2, ns: (np.zeros((n, 16)), 0.0))
    mi = -1.5 * np.log1p(-r_exact ** 2)          # inverse of r = sqrt(1 - exp(-2/3 MI))
    plan = _FakePlan(mi)
    a1 = np.arange(n, dtype=np.int32)             # pair p is "atom" p: the fake looks it up
    res = surrogate.hybrid(plan, a1, np.zeros(n, np.int32), _FakeModel(mu, sigma), **kw)
    return res, plan


def test_hybrid_gate_keeps_only_calibrated_pairs(monkeypatch):
    """Half of the pairs are predicted to 1e-4 with sigma 1e-4, half are off
    by 0.05 with sigma 0.05: at band 1e-3 the calibrated gate keeps only the
    accurate half (minus the audit), and every other pair is exact."""
    rng = np.random.default_rng(1)
    n = 20000
    r_exact = rng.uniform(0.2, 0.8, n)
    good = rng.random(n) < 0.5
    sigma = np.where(good, 1e-4, 0.05)
    mu = r_exact + sigma * rng.standard_normal(n) * 0.5
    res, plan = _run_hybrid(monkeypatch, r_exact, mu, sigma, band=1e-3, audit=0.05, alpha=0.01)
    kept = res["kept"]
    assert not res["failed_closed"]
    assert kept.any() and np.all(good[kept])
    assert np.all(~kept[~good])
    assert np.allclose(res["r"][~kept], r_exact[~kept], rtol=0, atol=1e-12)
    assert np.max(np.abs(res["r"][kept] - r_exact[kept])) <= 1e-3


def test_hybrid_fails_closed_when_the_check_half_breaks_the_band(monkeypatch):
    """Pairs whose sigma is small but whose mu is wrong, present only in the
    audit's check half: the gate would keep them, the check sees the
    violation, and every pair is run as exact KSG."""
    n = 4000
    rng = np.random.default_rng(2)
    r_exact = rng.uniform(0.2, 0.8, n)
    sigma = np.full(n, 1e-4)
    mu = r_exact.copy()
    res0, _ = _run_hybrid(monkeypatch, r_exact, mu, sigma, band=1e-3, audit=0.1, seed=5)
    aud = np.flatnonzero(res0["audited"])
    mu[aud[1::2][:8]] += 0.2          # wrong and confident, in the check half only
    res, _ = _run_hybrid(monkeypatch, r_exact, mu, sigma, band=1e-3, audit=0.1, seed=5)
    assert res["failed_closed"] and res["check_max_abs_dr"] > 1e-3
    assert not res["kept"].any()
    assert np.allclose(res["r"], r_exact, rtol=0, atol=1e-12)
