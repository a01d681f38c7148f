"""Learned surrogate of the KSG r(MI) with exact-KSG fallback.

The surrogate has no counterpart in the reference.  It predicts r(MI) of a
pair (the quantity pearsify() writes to correl.dat, lib/correlation_core.cpp:
461-469) from cheap per-pair features, with a predicted uncertainty; pairs
whose predicted sigma is too large, plus a random audit fraction, are sent to
the exact GPU KSG (KsgPlan.run_pairs), which also measures the surrogate's
error on the audited pairs.  Network, training (forward, backward, Adam) and
BF16 inference all run in the HIP library.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _lib
from .ksg import KsgPlan, pearsify

NFEAT = 14          # informative inputs; two zero pads make 16


def mi_to_r(mi: np.ndarray) -> np.ndarray:
    return pearsify(np.asarray(mi, dtype=np.float64))


def pair_features(plan: KsgPlan, a1, a2, nsample: int = 32) -> tuple[np.ndarray, float]:
    """[npairs][16] float32 features of the pairs on a loaded plan, and the
    device time in ms."""
    a1 = np.ascontiguousarray(a1, dtype=np.int32)
    a2 = np.ascontiguousarray(a2, dtype=np.int32)
    feat = np.empty((len(a1), 16), np.float32)
    ms = ctypes.c_double(0.0)
    _lib.check(_lib.load().gmxcorr_ksg_plan_features(
        plan._h, len(a1), _lib.ptr(a1, _lib._ip), _lib.ptr(a2, _lib._ip), nsample,
        _lib.ptr(feat, _lib._fp), ctypes.byref(ms)))
    return feat, ms.value


class Mlp:
    """ctypes handle of the device MLP (16 -> 64 -> 64 -> 2)."""

    def __init__(self, max_batch: int = 8192):
        L = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(L.gmxcorr_mlp_create(max_batch, ctypes.byref(h)))
        self._h = h
        self.max_batch = max_batch
        self.nparams = L.gmxcorr_mlp_num_params()

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (self.nparams,)
        _lib.check(_lib.load().gmxcorr_mlp_set_params(self._h, _lib.ptr(p, _lib._fp)))

    def get_params(self) -> np.ndarray:
        p = np.empty(self.nparams, np.float32)
        _lib.check(_lib.load().gmxcorr_mlp_get_params(self._h, _lib.ptr(p, _lib._fp)))
        return p

    def grad(self, x, y) -> tuple[np.ndarray, float]:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.empty(self.nparams, np.float32)
        loss = ctypes.c_float(0.0)
        _lib.check(_lib.load().gmxcorr_mlp_grad(self._h, _lib.ptr(x, _lib._fp),
                                                _lib.ptr(y, _lib._fp), len(y),
                                                _lib.ptr(g, _lib._fp), ctypes.byref(loss)))
        return g, loss.value

    def train_epoch(self, x, y, batch: int, lr: float) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = ctypes.c_float(0.0)
        _lib.check(_lib.load().gmxcorr_mlp_train_epoch(self._h, _lib.ptr(x, _lib._fp),
                                                       _lib.ptr(y, _lib._fp), len(y), batch,
                                                       lr, ctypes.byref(loss)))
        return loss.value

    def forward(self, x, bf16: bool = True) -> tuple[np.ndarray, float]:
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty((len(x), 2), np.float32)
        ms = ctypes.c_double(0.0)
        _lib.check(_lib.load().gmxcorr_mlp_forward(self._h, _lib.ptr(x, _lib._fp), len(x),
                                                   1 if bf16 else 0, _lib.ptr(out, _lib._fp),
                                                   ctypes.byref(ms)))
        return out, ms.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.load().gmxcorr_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def init_params(seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    parts = [rng.standard_normal((64, 16)) * np.sqrt(1 / 16), np.zeros(64),
             rng.standard_normal((64, 64)) * np.sqrt(1 / 64), np.zeros(64),
             rng.standard_normal((2, 64)) * np.sqrt(1 / 64) * 0.1, np.zeros(2)]
    return np.concatenate([q.ravel() for q in parts]).astype(np.float32)


class Surrogate:
    """Feature standardisation + the device MLP."""

    def __init__(self, max_batch: int = 8192, seed: int = 0):
        self.mlp = Mlp(max_batch)
        self.seed = seed
        self.mean = np.zeros(16, np.float32)
        self.std = np.ones(16, np.float32)
        self.history: list[float] = []

    def _norm(self, feat):
        x = (np.asarray(feat, np.float32) - self.mean) / self.std
        x[:, NFEAT:] = 0.0
        return np.ascontiguousarray(x)

    def fit(self, feat, r, epochs: int = 150, batch: int = 4096, lr: float = 2e-3) -> "Surrogate":
        feat = np.asarray(feat, np.float32)
        r = np.asarray(r, np.float32)
        self.mean = feat.mean(axis=0)
        self.std = feat.std(axis=0) + 1e-6
        self.mean[NFEAT:] = 0.0
        self.std[NFEAT:] = 1.0
        x = self._norm(feat)
        p = init_params(self.seed)
        # output bias: mean target, log of the target variance
        p[-2] = float(r.mean())
        p[-1] = float(np.log(r.var() + 1e-12))
        self.mlp.set_params(p)
        rng = np.random.default_rng(self.seed)
        batch = min(batch, len(r))
        for ep in range(epochs):
            perm = rng.permutation(len(r))
            step_lr = lr * (0.5 * (1 + np.cos(np.pi * ep / epochs)))   # cosine decay
            self.history.append(self.mlp.train_epoch(x[perm], r[perm], batch, step_lr))
        return self

    def predict(self, feat, bf16: bool = True) -> tuple[np.ndarray, np.ndarray, float]:
        out, ms = self.mlp.forward(self._norm(feat), bf16=bf16)
        mu = out[:, 0].astype(np.float64)
        sigma = np.exp(0.5 * out[:, 1].astype(np.float64))
        return mu, sigma, ms


def _all_pairs_in_order(a1: np.ndarray, a2: np.ndarray, natoms: int) -> bool:
    """True when (a1, a2) lists every pair a2 < a1 of natoms atoms in the
    order of KsgPlan.run_all (a1 = 1, 2, ..., a2 = 0 .. a1 - 1)."""
    if len(a1) != natoms * (natoms - 1) // 2:
        return False
    t1, t2 = np.tril_indices(natoms, -1)
    return bool(np.array_equal(a1, t1) and np.array_equal(a2, t2))


def calibrate_threshold(resid: np.ndarray, sigma: np.ndarray, band: float, alpha: float) -> float:
    """Largest sigma threshold tau whose kept calibration pairs meet the band:
    among the m calibration pairs with sigma <= tau, the
    ceil((m + 1)(1 - alpha))-th smallest |residual| (the split-conformal
    quantile of that subset) is at most band.  Calibrating on the subset the
    gate keeps, not on all pairs, is what makes the bound hold for the kept
    pairs: a network's sigma is least reliable exactly where it is small.
    Returns -inf when no threshold qualifies (nothing is kept)."""
    err = np.abs(np.asarray(resid, np.float64))
    sig = np.asarray(sigma, np.float64)
    order = np.argsort(sig, kind="stable")
    s_sorted, e_sorted = sig[order], err[order]
    n = len(order)
    best = 0
    m_min = int(np.ceil((1.0 - alpha) / alpha))          # smallest m with a finite quantile
    # the gate keeps every pair with sigma <= tau, so a kept prefix ends at
    # the last of the pairs sharing its sigma; beyond 8192 such ends, steps
    # of 0.2% in m (the kept share is resolved to that)
    ends = np.flatnonzero(np.r_[s_sorted[1:] != s_sorted[:-1], True]) + 1
    if len(ends) > 8192:
        grid = np.geomspace(max(m_min, 1), max(n, 1), 4096)
        ends = np.unique(ends[np.maximum(np.searchsorted(ends, grid, side="right") - 1, 0)])
    for m in ends:
        if m < m_min:
            continue
        rank = int(np.ceil((m + 1) * (1.0 - alpha)))
        if rank > m:
            continue
        q = np.partition(e_sorted[:m], rank - 1)[rank - 1]
        if q <= band:
            best = max(best, int(m))
    return float(s_sorted[best - 1]) if best else float("-inf")


def hybrid(plan: KsgPlan, a1, a2, model: Surrogate, band: float = 1e-3, audit: float = 0.01,
           alpha: float = 0.01, seed: int = 0, nsample: int = 32,
           sigma_max: float | None = None) -> dict:
    """r(MI) of the pairs: the surrogate's mu where it is trusted, exact GPU
    KSG everywhere else.

    A random audit fraction of the pairs is always run as exact KSG first.
    Half of the audit calibrates the gate (calibrate_threshold: the largest
    sigma threshold whose kept calibration pairs have their (1 - alpha)
    split-conformal error quantile within band).  The other half checks it:
    if more than an alpha share (plus one pair) of its pairs that the gate
    keeps is off by more than band, the hybrid fails closed and every pair
    runs as exact KSG.  The gate bounds the (1 - alpha) quantile of the kept
    pairs' error, not its maximum.  With sigma_max set, the raw gate sigma <=
    sigma_max is used instead (no calibration)."""
    a1 = np.asarray(a1, np.int32)
    a2 = np.asarray(a2, np.int32)
    t0 = time.perf_counter()
    feat, ms_feat = pair_features(plan, a1, a2, nsample)
    mu, sigma, ms_inf = model.predict(feat, bf16=True)
    r = np.clip(mu, 0.0, 1.0)
    rng = np.random.default_rng(seed)
    audited = rng.random(len(a1)) < audit
    aud_idx = np.flatnonzero(audited)
    ms_ksg = 0.0
    r_aud = np.zeros(0)
    if len(aud_idx):
        r_aud = mi_to_r(plan.run_pairs(a1[aud_idx], a2[aud_idx]))
        ms_ksg += plan.stats()["ms_pairs"] + plan.stats()["ms_fixup"]
    res_aud = r[aud_idx] - r_aud
    cal = np.zeros(len(aud_idx), bool)
    cal[::2] = True
    if sigma_max is not None:
        thresh = sigma_max
    else:
        thresh = calibrate_threshold(res_aud[cal], sigma[aud_idx][cal], band, alpha)
    gate = sigma <= thresh
    chk = ~cal & gate[aud_idx]
    check_err = np.abs(res_aud[chk])
    check_max = float(check_err.max()) if chk.any() else 0.0
    check_miss = int(np.sum(check_err > band))
    failed_closed = sigma_max is None and check_miss > alpha * len(check_err) + 1
    keep = gate & ~audited & (not failed_closed)
    fallback = ~keep
    r[aud_idx] = r_aud
    rest = fallback & ~audited
    if rest.any():
        if fallback.all() and _all_pairs_in_order(a1, a2, plan.natoms):
            # every pair exact: the all-pairs launch (2x2 atom tiles keep more
            # rows in L2) gives the same values
            r = mi_to_r(plan.run_all())
        else:
            r[rest] = mi_to_r(plan.run_pairs(a1[rest], a2[rest]))
        ms_ksg += plan.stats()["ms_pairs"] + plan.stats()["ms_fixup"]
    wall = time.perf_counter() - t0
    kept_aud = audited & gate
    return {
        "r": r, "mu": mu, "sigma": sigma, "fallback": fallback, "audited": audited,
        "kept": keep, "fallback_fraction": float(fallback.mean()),
        "band": band, "alpha": alpha, "sigma_threshold": thresh,
        "failed_closed": bool(failed_closed), "check_max_abs_dr": check_max,
        "check_pairs": int(len(check_err)), "check_beyond_band": check_miss,
        "audit_max_abs_dr": float(np.max(np.abs(res_aud[gate[aud_idx]])))
        if kept_aud.any() else 0.0,
        "audit_kept_pairs": int(kept_aud.sum()),
        "ms_features": ms_feat, "ms_inference": ms_inf, "ms_exact_ksg": ms_ksg,
        "wall_s": wall,
    }
