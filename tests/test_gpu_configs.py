"""The three larger workloads at their own sizes, on the GPU, against the
oracle.

- Dense path, 1000 atoms x 10^4 frames (Student-t noise, 50 modes): the
  split-FP16 MFMA second-moment GEMM against a host FP64 GEMM, Gaussian MI
  against the oracle's FP64 log-dets of that FP64 matrix, GCMI copula scores
  and sampled pair MI against the oracle (lib/correlation_core.cpp:140-166,
  280-311; lib/gcmi.cpp:95-146, 190-231), the GEMM also in four frame slabs;
  transfer entropy at lags 1, 5, 10 on ordered pairs against the oracle's
  FP64 restatement (lib/transfer_entropy.cpp:63-221).
- Surrogate: trained on exact KSG of 500 atoms x 4000 frames, hybrid on
  2000 atoms x 5000 frames (k = 100): every pair sent to exact KSG equals the
  plain run, and plain pairs equal the oracle's grid restatement
  (lib/kraskov.cpp:1228-1331).
- Batched: the 32 uneven trajectories in persistent launches, 20 pairs of
  every trajectory against the oracle.

Tolerances (stated against FP64): GEMM entries 2e-6 relative to
sqrt(C_ii C_jj); dense MI 1e-4 nats; copula scores 1e-12; TE 1e-10 nats; KSG
MI 1e-9 nats except at an exact FP64 tie at the k-th distance, where the GPU
takes every tied frame (the reference CUDA kernel's rule) and the oracle
keeps the reference CPU's grid-order choice: there |dr| <= 1e-3.
"""
import concurrent.futures as cf

import numpy as np
import pytest

from gmx_correlation_amd import _lib, batch, dense, ksg, surrogate, synth

pytestmark = pytest.mark.gpu

R_TOL = 1e-3
MI_TOL = 1e-9


@pytest.fixture(scope="module")
def gpu(hiplib):
    if not _lib.probe():
        pytest.fail("no gfx950 device visible to the HIP library")
    return hiplib


def _pmap(fn, args, workers=16):
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(lambda a: fn(*a), args))


@pytest.fixture(scope="module")
def dense_traj():
    print(f"configs: generating {synth.DENSE_1000}", flush=True)
    return ksg.center(synth.generate(synth.DENSE_1000))


def _off_pairs_hold_ties(oracle_lib, rows, psi, k, a1, a2, got, want):
    """Pairs whose GPU MI is more than 1e-9 nats off the grid restatement must
    hold an exact FP64 tie at the k-th distance and equal the FP64
    brute-force recount under the all-ties rule.  Returns how many there
    were."""
    off = np.flatnonzero(np.abs(got - want) > MI_TOL)
    for p in off:
        v, d = oracle_lib.ksg_pair(rows, psi, int(a1[p]), int(a2[p]), k, "brute", True)
        assert int(d["tie"].sum()) > 0, (a1[p], a2[p])
        assert abs(got[p] - v) <= 1e-12, (a1[p], a2[p], got[p], v)
    return len(off)


def _rel_cov_err(got, want):
    d = np.sqrt(np.outer(np.diag(want), np.diag(want)))
    return float(np.max(np.abs(got - want) / d))


def test_dense_1000x10000_gauss(gpu, dense_traj, oracle_lib):
    t = dense_traj
    n, N = t.shape[0], t.shape[1]
    print("configs: gauss 1000 x 10000", flush=True)
    mat, cov, st = dense.gauss_corrmatrix(t, return_cov=True, stats=True)
    X = np.ascontiguousarray(t.transpose(1, 0, 2).reshape(N, -1))
    c64 = X.T @ X / N
    err = _rel_cov_err(cov, c64)
    mi64 = oracle_lib.gauss_matrix_from_cov(c64, n)
    iu = np.triu_indices(n, 1)
    dmi = float(np.max(np.abs(mat[iu] - mi64[iu])))
    print(f"gauss 1000x10000: cov rel err {err:.2e}, max|dMI| vs FP64 {dmi:.2e}, "
          f"gemm {st['ms_gemm']:.3f} ms", flush=True)
    assert err <= 2e-6
    assert dmi <= 1e-4
    assert np.all(np.diag(mat) == 2000.0)
    # the same GEMM in four frame slabs (partial matrices summed)
    _lib.check(gpu.gmxcorr_dense_set_slab_frames(2560))
    try:
        mat_s, cov_s = dense.gauss_corrmatrix(t, return_cov=True)
    finally:
        _lib.check(gpu.gmxcorr_dense_set_slab_frames(0))
    assert _rel_cov_err(cov_s, c64) <= 2e-6
    assert float(np.max(np.abs(mat_s[iu] - mi64[iu]))) <= 1e-4


def test_dense_1000x10000_gcmi(gpu, dense_traj, oracle_lib):
    t = dense_traj
    n, N = t.shape[0], t.shape[1]
    print("configs: gcmi 1000 x 10000", flush=True)
    mat, cov, z = dense.gcmi_corrmatrix(t, return_cov=True, return_z=True)
    zo = oracle_lib.copula(t)
    assert float(np.max(np.abs(z - zo))) <= 1e-12
    c64 = zo @ zo.T / N
    assert _rel_cov_err(cov, c64) <= 2e-6
    rng = np.random.default_rng(22)
    a1 = rng.integers(0, n, 400)
    a2 = (a1 + rng.integers(1, n, 400)) % n
    want = np.array(_pmap(lambda x, y: oracle_lib.gcmi_pair(zo, int(x), int(y)),
                          list(zip(a1, a2))))
    dmi = float(np.max(np.abs(mat[a1, a2] - want)))
    print(f"gcmi 1000x10000: 400 sampled pairs max|dMI| vs oracle {dmi:.2e}", flush=True)
    assert dmi <= 1e-4


@pytest.mark.parametrize("lag", [1, 5, 10])
def test_dense_1000x10000_te(gpu, dense_traj, oracle_lib, lag):
    t = dense_traj
    n = t.shape[0]
    rng = np.random.default_rng(lag)
    a1 = rng.integers(0, n, 6)
    a2 = (a1 + rng.integers(1, n, 6)) % n
    print(f"configs: TE 1000 x 10000 lag {lag}, 6 ordered pairs", flush=True)
    te, ms = dense.te_pairs(t, a1, a2, k=100, lag=lag)
    want = np.array(_pmap(lambda x, y: oracle_lib.te_pair(t, int(x), int(y), 100, lag),
                          list(zip(a1, a2))))
    d = float(np.max(np.abs(te - want)))
    print(f"TE lag {lag}: max|dTE| {d:.2e}, kernel {ms:.1f} ms", flush=True)
    assert d <= 1e-10
    # the same pairs through the per-target prefix stores (the automatic
    # path of the full matrix): bit-identical
    dense.te_set_path(3)
    try:
        te3, ms3 = dense.te_pairs(t, a1, a2, k=100, lag=lag)
        st = dense.te_stats()
    finally:
        dense.te_set_path(0)
    print(f"TE lag {lag} prefix stores: kernel {ms3:.1f} ms, {st}", flush=True)
    assert st["prefix_path"] and st["prefix_resolved"] > 0
    assert np.array_equal(te3, te)


def test_surrogate_train_500x4000_hybrid_2000x5000(gpu, oracle_lib):
    k = 100
    spec = synth.SURR_TRAIN
    print(f"configs: surrogate training set {spec}", flush=True)
    rows, psi = ksg.kraskov_prepare(ksg.center(synth.generate(spec)))
    a1, a2 = np.tril_indices(spec.natoms, -1)
    with ksg.KsgPlan(spec.natoms, spec.nframes, k) as plan:
        plan.load(rows, psi)
        r_train = surrogate.mi_to_r(plan.run_all())
        feat, _ = surrogate.pair_features(plan, a1, a2)
    model = surrogate.Surrogate(8192, seed=0).fit(feat, r_train, epochs=60, batch=4096, lr=2e-3)
    assert model.history[-1] < model.history[0]
    del rows, feat
    spec = synth.SURR_INFER
    print(f"configs: surrogate inference set {spec}", flush=True)
    rows, psi = ksg.kraskov_prepare(ksg.center(synth.generate(spec)))
    n = spec.natoms
    a1, a2 = np.tril_indices(n, -1)
    with ksg.KsgPlan(n, spec.nframes, k) as plan:
        plan.load(rows, psi)
        mi_plain = plan.run_all()
        r_plain = surrogate.mi_to_r(mi_plain)
        res = surrogate.hybrid(plan, a1, a2, model, band=3e-2, audit=0.01, alpha=0.01, seed=1)
    fb = res["fallback"]
    kept = res["kept"]
    err_kept = np.abs(res["r"][kept] - r_plain[kept])
    print(f"hybrid 2000x5000 at band 3e-2: fallback {res['fallback_fraction']:.3f}, sigma "
          f"threshold {res['sigma_threshold']:.3e}, kept pairs within band "
          f"{np.mean(err_kept <= 3e-2) if kept.any() else 1.0:.4f}, max|dr| "
          f"{err_kept.max() if kept.any() else 0.0:.3e}", flush=True)
    assert 0.0 < res["fallback_fraction"] < 1.0 and not res["failed_closed"]
    # split-conformal coverage at level 0.99 over about 10^4 calibration pairs
    assert np.mean(err_kept <= 3e-2) >= 0.98
    # every pair sent to exact KSG equals the plain run; a pair list splits
    # rows into other ranges than the all-pairs launch, so allow the last bits
    assert float(np.max(np.abs(res["r"][fb] - r_plain[fb]))) <= 1e-12
    # plain pairs against the oracle's grid restatement
    pick = np.linspace(0, len(a1) - 1, 12).astype(np.int64)
    want = oracle_lib.ksg_pairs(rows, psi, k, a1[pick], a2[pick], 0)
    assert float(np.max(np.abs(ksg.pearsify(mi_plain[pick]) - ksg.pearsify(want)))) <= R_TOL
    _off_pairs_hold_ties(oracle_lib, rows, psi, k, a1[pick], a2[pick], mi_plain[pick], want)


def test_batched_32_uneven(gpu, oracle_lib):
    specs = synth.batched_specs()
    b = batch.KsgBatch()
    keep = []
    for t, (spec, k) in enumerate(specs):
        rows, psi = ksg.kraskov_prepare(ksg.center(synth.generate(spec)))
        b.add(rows, psi, k)
        keep.append((rows, psi, k))
    print(f"configs: batched, {len(specs)} trajectories", flush=True)
    st = b.run(max_row_bytes=200 * 10**9, nblocks=2048)
    print(f"batched: {st['npairs']} pairs in {st['ms_total'] / 1e3:.1f} s, {st['nwaves']} waves",
          flush=True)
    jobs, where = [], []
    for t, (spec, k) in enumerate(specs):
        rows, psi, _ = keep[t]
        mi, _, _ = b.result(t)
        a1, a2 = np.tril_indices(spec.natoms, -1)
        pick = np.linspace(0, len(a1) - 1, 20).astype(np.int64)
        jobs.append((rows, psi, k, a1[pick], a2[pick]))
        where.append(mi[pick])
    got = _pmap(lambda r, p, k, x, y: oracle_lib.ksg_pairs(r, p, k, x, y, 1), jobs)
    worst_r, off = 0.0, 0
    for (rows, psi, k, x, y), mi, want in zip(jobs, where, got):
        worst_r = max(worst_r, float(np.max(np.abs(ksg.pearsify(mi) - ksg.pearsify(want)))))
        off += _off_pairs_hold_ties(oracle_lib, rows, psi, k, x, y, mi, want)
    print(f"batched: 640 sampled pairs, max|dr| {worst_r:.2e}, {off} with exact ties",
          flush=True)
    assert worst_r <= R_TOL
