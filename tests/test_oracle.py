"""Pins the CPU oracle: against the reference's own outputs (golden fixtures
written from the compiled reference sources) and against what the
mathematics fixes.  No GPU."""
import json
import os

import numpy as np
import pytest
import scipy.special
import scipy.stats

from gmx_correlation_amd import synth
from tests.conftest import GOLDEN, golden


@pytest.fixture(scope="module")
def small(oracle_lib):
    c = synth.generate(synth.SMALL_CPU)
    traj = oracle_lib.center(c)
    kx, psi = oracle_lib.prepare(traj)
    return c, traj, kx, psi


def test_generator_is_frozen():
    meta = json.load(open(os.path.join(GOLDEN, "meta.json")))
    assert synth.coords_sha256(synth.generate(synth.SMALL_CPU)) == meta["small_cpu_sha256"]
    assert synth.coords_sha256(synth.generate(synth.README_247)) == meta["readme247_sha256"]


def test_prepare_matches_reference(small):
    _, _, kx, psi = small
    assert np.array_equal(kx[:6], golden("small_cpu_prepare_rows_a0a1.npy"))
    assert np.array_equal(psi, golden("small_cpu_psi.npy"))


def test_psi_table_closed_form(small):
    """psi[m] = -0.57721566490153 + H_{m-1}; the truncated Euler constant
    cancels in MI, the rest is scipy's digamma."""
    _, _, _, psi = small
    m = np.arange(1, len(psi))
    shift = -0.57721566490153 + np.euler_gamma
    assert np.allclose(psi[1:] + shift, scipy.special.digamma(m), rtol=0, atol=1e-12)
    assert psi[0] == 0.0


def test_normalised_rows_are_shifted_zscores(small):
    _, traj, kx, _ = small
    z = (traj[0, :, 0] - traj[0, :, 0].mean()) / traj[0, :, 0].std(ddof=1)
    assert kx[0].min() == 0.0
    assert np.allclose(kx[0], z - z.min(), atol=1e-12)


def test_ksg_matrix_bitexact_vs_reference(small, oracle_lib):
    _, _, kx, psi = small
    ref = golden("small_cpu_k10_mi.npy")
    grid = oracle_lib.ksg_matrix(kx, psi, 64, 10, nthreads=4)
    assert np.array_equal(grid, ref)


@pytest.mark.parametrize("k", [1, 3, 25, 499])
def test_ksg_other_k_bitexact_vs_reference(small, oracle_lib, k):
    _, _, kx, psi = small
    ref = golden(f"small_cpu_k{k}_pairs.npy")
    got = [oracle_lib.ksg_pair(kx, psi, a1, a2, k) for a1, a2 in ((1, 0), (17, 5), (63, 62))]
    assert np.array_equal(np.array(got), ref)


def test_brute_force_recount_agrees(small, oracle_lib):
    """Without ties the brute-force FP64 recount and the grid search are the
    same numbers, frame by frame."""
    _, _, kx, psi = small
    for a1, a2 in ((1, 0), (40, 7)):
        vg, dg = oracle_lib.ksg_pair(kx, psi, a1, a2, 10, detail=True)
        vb, db = oracle_lib.ksg_pair(kx, psi, a1, a2, 10, method="brute", detail=True)
        assert not db["tie"].any()
        assert np.array_equal(dg["nx"], db["nx"]) and np.array_equal(dg["ny"], db["ny"])
        assert np.array_equal(dg["epsx"], db["epsx"]) and np.array_equal(dg["epsk"], db["epsk"])
        assert vg == vb


def test_brute_force_by_hand_tiny(oracle_lib):
    """5 frames of 1-D-like data worked out independently in numpy."""
    rng = np.random.default_rng(5)
    kx = rng.random((6, 9))
    psi = np.concatenate([[0.0], -0.57721566490153 + np.concatenate([[0.0], np.cumsum(1.0 / np.arange(1, 9))])])
    k = 2
    acc = 0.0
    for i in range(9):
        dA = np.abs(kx[:3, :] - kx[:3, i:i + 1]).max(axis=0)
        dB = np.abs(kx[3:6, :] - kx[3:6, i:i + 1]).max(axis=0)
        dJ = np.maximum(dA, dB)
        others = [j for j in range(9) if j != i]
        nn = sorted(others, key=lambda j: dJ[j])[:k]
        ex, ey = dA[nn].max(), dB[nn].max()
        nx = sum(dA[j] <= ex for j in others)
        ny = sum(dB[j] <= ey for j in others)
        acc += psi[nx] + psi[ny]
    expect = psi[9] + (psi[k] - 1.0 / k) - acc / 9
    # atoms 0 and 1 of a 2-atom workspace
    got = oracle_lib.ksg_pair(np.ascontiguousarray(kx), psi, 0, 1, k)
    assert abs(got - expect) < 1e-13


def test_k_equals_n_minus_one_gives_zero(small):
    """All counts are N-1, so MI = psi[N] - psi[N-1] - 1/(N-1) = 0 up to
    rounding of the recurrence."""
    ref = golden("small_cpu_k499_pairs.npy")
    assert np.all(np.abs(ref) < 1e-12)


def test_symmetry_and_permutation(small, oracle_lib):
    _, _, kx, psi = small
    v = oracle_lib.ksg_pair(kx, psi, 9, 4, 10)
    assert v == oracle_lib.ksg_pair(kx, psi, 4, 9, 10) or abs(v - oracle_lib.ksg_pair(kx, psi, 4, 9, 10)) < 1e-12
    perm = np.random.default_rng(1).permutation(kx.shape[1])
    assert abs(oracle_lib.ksg_pair(kx[:, perm].copy(), psi, 9, 4, 10) - v) < 1e-12


def test_gaussian_closed_form_statistical(oracle_lib):
    """Three independent unit-variance components correlated with r gives
    MI = -1.5 log(1 - r^2) (tests/test_kraskov.py:58-62 of the reference)."""
    rng = np.random.default_rng(7)
    N, r = 2000, 0.5
    x = rng.standard_normal((N, 3))
    y = r * x + np.sqrt(1 - r * r) * rng.standard_normal((N, 3))
    traj = np.stack([x, y]).astype(np.float64)
    kx, psi = oracle_lib.prepare(traj - traj.mean(axis=1, keepdims=True))
    mi = oracle_lib.ksg_pair(kx, psi, 1, 0, 3)
    # KSG-2 is biased low in 6-D; the reference test allows 0.08-0.30 nats
    assert abs(mi - (-1.5 * np.log(1 - r * r))) < 0.08


def test_ties_fixture(oracle_lib):
    """Quantised coordinates: the grid restatement reproduces the reference's
    choice among equidistant neighbours bit-exactly."""
    from tests.golden.make_golden import tie_spec
    c = tie_spec()
    kx, psi = oracle_lib.prepare(oracle_lib.center(c))
    got = oracle_lib.ksg_matrix(kx, psi, 12, 4)
    assert np.array_equal(got, golden("ties12_k4_mi.npy"))
    # and the brute recount sees the ties
    _, d = oracle_lib.ksg_pair(kx, psi, 3, 1, 4, method="brute", detail=True)
    assert d["tie"].any()


def test_pearsify_closed_form(oracle_lib):
    m = np.array([[2000.0, 0.3], [-0.1, 0.0]])
    r = oracle_lib.pearsify(m)
    assert r[0, 0] == 1.0 and r[1, 0] == 0.0 and r[1, 1] == 0.0
    assert abs(r[0, 1] - np.sqrt(1 - np.exp(-0.2))) < 1e-15


# ---------------------------------------------------------------- dense paths

def test_lu_det_vs_numpy(oracle_lib):
    rng = np.random.default_rng(3)
    for n in (3, 6):
        a = rng.standard_normal((n, n))
        assert abs(oracle_lib.lu_det(a) - np.linalg.det(a)) < 1e-10 * max(1, abs(np.linalg.det(a)))
    assert oracle_lib.lu_det(np.zeros((3, 3))) == 0.0


def test_covariance_vs_numpy(small, oracle_lib):
    _, traj, _, _ = small
    t = traj[:10]
    X = t.transpose(1, 0, 2).reshape(t.shape[1], -1)
    assert np.allclose(oracle_lib.covariance(t), X.T @ X / X.shape[0], atol=1e-13)


def test_gauss_matrix_matches_reference(small, oracle_lib):
    _, traj, _, _ = small
    assert np.allclose(oracle_lib.gauss_matrix(traj), golden("small_cpu_gauss_mi.npy"),
                       rtol=1e-12, atol=1e-12)


def test_gauss_closed_form(oracle_lib):
    """MI of a 6-D Gaussian with known covariance is
    0.5*(log det C11 + log det C22 - log det C)."""
    rng = np.random.default_rng(2)
    A = rng.standard_normal((6, 6))
    C = A @ A.T + 6 * np.eye(6)
    expect = 0.5 * (np.linalg.slogdet(C[:3, :3])[1] + np.linalg.slogdet(C[3:, 3:])[1]
                    - np.linalg.slogdet(C)[1])
    assert abs(oracle_lib.gauss_matrix_from_cov(C, 2)[0, 1] - expect) < 1e-12


def test_erfinv_vs_scipy(oracle_lib):
    xs = np.linspace(-0.9999, 0.9999, 4001)
    got = np.array([oracle_lib.erfinv(x) for x in xs])
    assert np.max(np.abs(got - scipy.special.erfinv(xs))) < 1e-9


def test_gcmi_matrix_matches_reference(small, oracle_lib):
    _, traj, _, _ = small
    assert np.allclose(oracle_lib.gcmi_matrix(traj), golden("small_cpu_gcmi_mi.npy"),
                       rtol=1e-12, atol=1e-12)


def test_copula_midranks(oracle_lib):
    """Tied values share the mean normal score of their rank range
    (tests/test_gcmi.py:111-116 of the reference)."""
    traj = np.zeros((1, 6, 3))
    traj[0, :, 0] = [3.0, 1.0, 1.0, 2.0, 5.0, 4.0]
    z = oracle_lib.copula(traj)
    ranks = scipy.stats.rankdata(traj[0, :, 0])
    scores = np.sqrt(2) * scipy.special.erfinv(2 * (np.arange(6) + 0.5) / 6 - 1)
    expect = [scores[int(r) - 1] if r == int(r) else 0.5 * (scores[0] + scores[1]) for r in ranks]
    assert np.allclose(z[0], expect, atol=1e-9)


@pytest.mark.parametrize("lag", [1, 5, 10])
def test_te_matches_reference(small, oracle_lib, lag):
    c, _, _, _ = small
    traj = oracle_lib.center(np.ascontiguousarray(c[:300, :16]))
    assert np.allclose(oracle_lib.te_matrix(traj, 5, lag), golden(f"te16_k5_lag{lag}.npy"),
                       rtol=1e-12, atol=1e-12)


def test_te_k_above_128_matches_reference(small, oracle_lib):
    """k = 200 at lag 3, and k = 150 on the same slice quantised to 0.02 nm
    (exact ties at the k-th distance), 6 atoms x 400 frames."""
    c, _, _, _ = small
    c6 = np.ascontiguousarray(c[:400, 16:22])
    assert np.allclose(oracle_lib.te_matrix(oracle_lib.center(c6), 200, 3),
                       golden("te6_k200_lag3.npy"), rtol=1e-12, atol=1e-12)
    cq = np.round(c6 / 0.02).astype(np.float32) * np.float32(0.02)
    assert np.allclose(oracle_lib.te_matrix(oracle_lib.center(cq), 150, 1),
                       golden("te6q_k150_lag1.npy"), rtol=1e-12, atol=1e-12)


def test_te_short_series_is_zero(oracle_lib):
    traj = np.random.default_rng(0).standard_normal((3, 8, 3))
    assert not oracle_lib.te_matrix(traj, 5, 4).any()
