"""GPU checks of the persistent batched launch and the autotuner: every
trajectory of an uneven batch (several memory waves) gives the same MI as
its own single-trajectory plan (same kernels, so within 1e-12: only the
order of FP64 recount atomics may differ) and as the FP64 brute recount."""
import numpy as np
import pytest

from gmx_correlation_amd import _lib, batch, ksg, synth

pytestmark = pytest.mark.gpu

SHAPES = [(7, 300, 5), (9, 900, 10), (6, 2500, 20), (5, 5000, 30), (8, 1500, 7)]


@pytest.fixture(scope="module")
def gpu(hiplib):
    if not _lib.probe():
        pytest.fail("no gfx950 device visible to the HIP library")
    return hiplib


@pytest.fixture(scope="module")
def trajs():
    out = []
    for t, (n, N, k) in enumerate(SHAPES):
        c = synth.generate(synth.TrajSpec(natoms=n, nframes=N, nmodes=4, seed=700 + t))
        rows, psi = ksg.kraskov_prepare(ksg.center(c))
        out.append((rows, psi, k))
    return out


def test_batch_matches_single_plans(gpu, trajs, oracle_lib):
    b = batch.KsgBatch()
    for rows, psi, k in trajs:
        b.add(rows, psi, k)
    # a small row budget forces several memory waves
    budget = max(b.traj_bytes(t) for t in range(len(trajs)))
    st = b.run(max_row_bytes=budget, nblocks=64)
    assert st["nwaves"] >= 2
    for t, (rows, psi, k) in enumerate(trajs):
        n, N = rows.shape[0] // 3, rows.shape[1]
        mi, ms_done, wave = b.result(t)
        with ksg.KsgPlan(n, N, k) as plan:
            plan.load(rows, psi)
            ref = plan.run_all()
        assert np.max(np.abs(mi - ref)) <= 1e-12, t
        assert ms_done > 0 and 0 <= wave < st["nwaves"]
    # one pair per trajectory against the FP64 brute-force recount
    for t, (rows, psi, k) in enumerate(trajs[:3]):
        mi, _, _ = b.result(t)
        v = oracle_lib.ksg_pair(rows, psi, 1, 0, k, method="brute")
        assert abs(mi[0] - v) <= 1e-12


@pytest.mark.parametrize("t", [1, 2])
def test_batch_explicit_paths(gpu, trajs, t):
    """Every compiled batch path that covers a trajectory gives the same MI
    (N = 900: kinds 0 and 2; N = 2500: kinds 1 and 3)."""
    rows, psi, k = trajs[t]
    ref = None
    for kind, s4 in batch.fitting_paths(rows.shape[1], k):
        b = batch.KsgBatch()
        b.add(rows, psi, k, kind, s4)
        b.run(nblocks=32)
        mi, _, _ = b.result(0)
        if ref is None:
            ref = mi
        assert np.max(np.abs(mi - ref)) <= 1e-12, (kind, s4)


@pytest.mark.parametrize("N,k,path", [(2000, 200, (2, 1)), (6000, 100, (3, 2)), (6000, 200, (3, 3))])
def test_batch_long_rows(gpu, oracle_lib, N, k, path):
    """Independent atoms, whose (k+1)-th joint neighbour lies beyond the
    prefix read (512 frames, or 768 at k = 200) in most rows: the batch kernel queues those rows and
    the cooperative streaming pass resolves them; counts stay exact, so MI
    equals the FP64 brute recount."""
    rng = np.random.default_rng(N + k)
    coords = rng.standard_normal((N, 4, 3)).astype(np.float32)
    rows, psi = ksg.kraskov_prepare(ksg.center(coords))
    b = batch.KsgBatch()
    b.add(rows, psi, k, *path)
    st = b.run(nblocks=64)
    assert st["long_rows"] > 1000
    mi, _, _ = b.result(0)
    a1, a2 = np.tril_indices(4, -1)
    for p in range(len(a1)):
        v = oracle_lib.ksg_pair(rows, psi, int(a1[p]), int(a2[p]), k, method="brute")
        assert abs(mi[p] - v) <= 1e-12, p


def test_tuner_returns_a_fitting_path(gpu, trajs):
    rows, psi, k = trajs[3]                      # N = 5000
    (kind, s4), timings = batch.tune_variant(rows, psi, k, natoms_probe=4, reps=1)
    assert (kind, s4) in batch.fitting_paths(rows.shape[1], k)
    assert all(v > 0 for v in timings.values())


def test_batch_rejects_bad_input(gpu):
    b = batch.KsgBatch()
    with pytest.raises(_lib.GmxcorrError):
        b.add(np.zeros((6, 10)), np.zeros(11), 20)
    with pytest.raises(_lib.GmxcorrError):
        b.add(np.zeros((6, 100)), np.zeros(101), 5, kind=0, s4=1)
