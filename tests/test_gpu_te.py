"""GPU transfer entropy against the reference's own output (golden, 16 atoms
x 300 frames, k = 5, lags 1/5/10; 6 atoms x 400 frames at k = 200 and, with
ties, k = 150) and the FP64 oracle.  Distances, the k-th
neighbour distance and the strict counts are FP64 exactly as the reference
computes them, so only the summation order differs: |dTE| <= 1e-12."""
import numpy as np
import pytest

from gmx_correlation_amd import _lib, dense, ksg, synth
from tests.conftest import golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu(hiplib):
    if not _lib.probe():
        pytest.fail("no gfx950 device visible to the HIP library")
    return hiplib


@pytest.mark.parametrize("lag", [1, 5, 10])
def test_te_matches_reference(gpu, lag):
    c = synth.generate(synth.SMALL_CPU)
    traj = ksg.center(np.ascontiguousarray(c[:300, :16]))
    mat = dense.te_matrix(traj, k=5, lag=lag)
    assert np.max(np.abs(mat - golden(f"te16_k5_lag{lag}.npy"))) <= 1e-12


def test_te_k100_and_ties_vs_oracle(gpu, oracle_lib):
    """k = 100 (the reference default) uses the 128-slot heap; quantised
    coordinates give exact ties at the k-th distance."""
    c = synth.generate(synth.TrajSpec(natoms=4, nframes=400, nmodes=3, seed=5))
    c = np.round(c / 0.02).astype(np.float32) * np.float32(0.02)
    traj = ksg.center(c)
    for k, lag in ((100, 1), (7, 3)):
        mat = dense.te_matrix(traj, k=k, lag=lag)
        ref = oracle_lib.te_matrix(traj, k, lag)
        assert np.max(np.abs(mat - ref)) <= 1e-12, (k, lag)


def test_te_pairs_and_short_series(gpu, oracle_lib):
    traj = ksg.center(synth.generate(synth.TrajSpec(natoms=5, nframes=250, nmodes=3, seed=8)))
    te, ms = dense.te_pairs(traj, [0, 4], [3, 1], k=10, lag=2)
    ref = [oracle_lib.te_pair(traj, 0, 3, 10, 2), oracle_lib.te_pair(traj, 4, 1, 10, 2)]
    assert np.max(np.abs(te - ref)) <= 1e-12 and ms > 0
    assert not dense.te_matrix(traj[:, :12], k=10, lag=4).any()


def test_te_duplicate_frames_vs_oracle(gpu, oracle_lib):
    """Three atoms move on two levels of one axis, so many 9-D points
    coincide exactly: at k = 5 the k-th distance is 0 for most queries (the
    query's own frame then has the root key too), at k = 40 and 100 the
    root is positive with many frames tied at it; the fourth atom is
    continuous."""
    rng = np.random.default_rng(21)
    c = np.zeros((300, 4, 3), np.float32)
    c[:, :, 0] = rng.integers(0, 2, (300, 4))
    c[:, 3, :] += rng.standard_normal((300, 3)).astype(np.float32)
    traj = ksg.center(c)
    for k, lag in ((5, 1), (40, 2), (100, 1)):
        mat = dense.te_matrix(traj, k=k, lag=lag)
        ref = oracle_lib.te_matrix(traj, k, lag)
        assert np.max(np.abs(mat - ref)) <= 1e-12, (k, lag)


def test_te_matrix_over_two_launch_chunks(gpu, oracle_lib):
    """2100 atoms x 100 frames: 4,407,900 ordered pairs, more than one
    launch pair of the heap and count kernels takes (2^29 queries, 4,194,304
    pairs at two query blocks per pair), so the matrix is computed in two
    chunks; pairs on both sides of the chunk boundary (target atom >= 1998
    lies in the second) equal the oracle."""
    rng = np.random.default_rng(41)
    c = rng.standard_normal((100, 2100, 3)).astype(np.float32)
    traj = ksg.center(c)
    mat = dense.te_matrix(traj, k=5, lag=1)
    for a1, a2 in ((0, 1), (1000, 5), (1997, 2098), (1998, 300), (2050, 0), (2099, 2098)):
        ref = oracle_lib.te_pair(traj, a1, a2, 5, 1)
        assert abs(mat[a1, a2] - ref) <= 1e-12, (a1, a2)


def test_te_k_above_128_matches_reference(gpu):
    """k above the heap kernel's 128 goes through the radix-select pass:
    the reference's own output at k = 200 (lag 3) and, on coordinates
    quantised to 0.02 nm (exact ties at the k-th distance), k = 150."""
    c = synth.generate(synth.SMALL_CPU)
    c6 = np.ascontiguousarray(c[:400, 16:22])
    mat = dense.te_matrix(ksg.center(c6), k=200, lag=3)
    assert np.max(np.abs(mat - golden("te6_k200_lag3.npy"))) <= 1e-12
    cq = np.round(c6 / 0.02).astype(np.float32) * np.float32(0.02)
    mat = dense.te_matrix(ksg.center(cq), k=150, lag=1)
    assert np.max(np.abs(mat - golden("te6q_k150_lag1.npy"))) <= 1e-12


@pytest.mark.parametrize("path", [1, 2])
def test_te_largest_k_vs_oracle(gpu, oracle_lib, path):
    """k = N - lag - 1, the largest the estimator takes (the k-th neighbour
    is the farthest frame), through each pass-1 kernel; and k = 399 of 401
    valid frames through the radix select."""
    traj = ksg.center(synth.generate(synth.TrajSpec(natoms=3, nframes=130, nmodes=2, seed=9)))
    dense.te_set_path(path)
    try:
        mat = dense.te_matrix(traj, k=127, lag=2)
        assert np.max(np.abs(mat - oracle_lib.te_matrix(traj, 127, 2))) <= 1e-12
        mat = dense.te_matrix(traj, k=128, lag=1)
        assert np.max(np.abs(mat - oracle_lib.te_matrix(traj, 128, 1))) <= 1e-12
    finally:
        dense.te_set_path(0)
    traj = ksg.center(synth.generate(synth.TrajSpec(natoms=3, nframes=402, nmodes=2, seed=10)))
    mat = dense.te_matrix(traj, k=400, lag=1)
    assert np.max(np.abs(mat - oracle_lib.te_matrix(traj, 400, 1))) <= 1e-12


@pytest.mark.parametrize("case", ["continuous", "quantised", "duplicates"])
def test_te_select_pass_equals_heap_pass(gpu, case):
    """Both pass-1 kernels find the same k-th neighbour key, so the TE
    values are identical, bit for bit, at k <= 128 where both run."""
    rng = np.random.default_rng(31)
    if case == "duplicates":
        c = np.zeros((300, 5, 3), np.float32)
        c[:, :, 0] = rng.integers(0, 2, (300, 5))
        c[:, 4, :] += rng.standard_normal((300, 3)).astype(np.float32)
    else:
        c = synth.generate(synth.TrajSpec(natoms=5, nframes=700, nmodes=3, seed=12))
        if case == "quantised":
            c = np.round(c / 0.02).astype(np.float32) * np.float32(0.02)
    traj = ksg.center(c)
    try:
        for k, lag in ((1, 1), (5, 2), (100, 1), (128, 4)):
            dense.te_set_path(1)
            heap = dense.te_matrix(traj, k=k, lag=lag)
            dense.te_set_path(2)
            sel = dense.te_matrix(traj, k=k, lag=lag)
            assert np.array_equal(heap, sel), (case, k, lag)
    finally:
        dense.te_set_path(0)


def _cap_short_queries(traj, a1, a2, k, lag, stride=32, rank=14):
    """Queries of TE(a2 -> a1) whose heap cap (the rank-th smallest FP32 key
    of every stride-th frame) has fewer than k + 1 keys below it."""
    X = np.asarray(traj[a1], np.float64)   # centred trajectories are [atom][frame][3]
    Y = np.asarray(traj[a2], np.float64)
    Nv = len(X) - lag
    P = np.concatenate([X[lag:lag + Nv], X[:Nv], Y[:Nv]], axis=1)
    short = 0
    for q in range(Nv):
        keys = np.abs(P - P[q]).max(axis=1).astype(np.float32)
        s = np.sort(keys[::stride])
        short += len(s) >= rank and int((keys < s[rank - 1]).sum()) < k + 1
    return short


def test_te_heap_cap_and_its_fallback(gpu, oracle_lib):
    """From 448 frames on, the heap pass offers only keys below the 14th
    smallest key of every 32nd frame.  A random trajectory takes that path
    throughout.  In the second, every 32nd 9-D point of TE(1 -> 0) is the
    same point (atom 0 at frames j and j + 1 and atom 1 at frame j, j a
    multiple of 32), the others spread over a cube around it: the cap of
    most queries holds fewer than k + 1 keys (checked here in numpy), and
    their waves repeat the frame loop without it.  Both equal the FP64
    oracle."""
    c = synth.generate(synth.TrajSpec(natoms=3, nframes=1000, nmodes=3, seed=17))
    traj = ksg.center(c)
    for k, lag in ((100, 1), (10, 5)):
        assert _cap_short_queries(traj, 0, 1, k, lag) == 0
        mat = dense.te_matrix(traj, k=k, lag=lag)
        ref = oracle_lib.te_matrix(traj, k, lag)
        assert np.max(np.abs(mat - ref)) <= 1e-12, (k, lag)
    rng = np.random.default_rng(77)
    c = rng.uniform(-10.0, 10.0, (1001, 3, 3)).astype(np.float32)
    for j in range(0, 1000, 32):
        c[j, 0] = c[j + 1, 0] = c[j, 1] = 0.0
    traj = ksg.center(c)
    short = _cap_short_queries(traj, 0, 1, 100, 1)
    assert short > 100, short
    mat = dense.te_matrix(traj, k=100, lag=1)
    ref = oracle_lib.te_matrix(traj, 100, 1)
    assert np.max(np.abs(mat - ref)) <= 1e-12


def test_te_grouped_count_pass_equals_single(gpu, oracle_lib):
    """The count kernel takes four consecutive ordered pairs with one target
    atom together (the matrix is launched target-major, a pair list in
    stable target-major order) and any other pair alone.  A scrambled list
    with two or six pairs per target equals the matrix entry for entry, bit
    for bit, and the oracle on a few pairs."""
    traj = ksg.center(synth.generate(synth.TrajSpec(natoms=7, nframes=600, nmodes=3, seed=23)))
    k, lag = 20, 2
    mat = dense.te_matrix(traj, k=k, lag=lag)
    rng = np.random.default_rng(5)
    pairs = [(t, s) for t in range(7) for s in range(7) if s != t]
    keep = [p for p in pairs if not (p[0] == 2 and p[1] > 2) and not (p[0] == 5 and p[1] > 1)]
    pick = rng.permutation(len(keep))
    a1 = np.array([keep[i][0] for i in pick], np.int32)
    a2 = np.array([keep[i][1] for i in pick], np.int32)
    te, _ = dense.te_pairs(traj, a1, a2, k=k, lag=lag)
    assert np.array_equal(te, mat[a1, a2])
    for p in range(3):
        assert abs(te[p] - oracle_lib.te_pair(traj, int(a1[p]), int(a2[p]), k, lag)) <= 1e-12


def _te_both(traj, k, lag, a1=None, a2=None):
    """TE through the heap pass (path 1) and the prefix stores (path 3), and
    the prefix run's counts."""
    try:
        dense.te_set_path(1)
        heap = (dense.te_matrix(traj, k=k, lag=lag) if a1 is None
                else dense.te_pairs(traj, a1, a2, k=k, lag=lag)[0])
        dense.te_set_path(3)
        pre = (dense.te_matrix(traj, k=k, lag=lag) if a1 is None
               else dense.te_pairs(traj, a1, a2, k=k, lag=lag)[0])
        st = dense.te_stats()
    finally:
        dense.te_set_path(0)
    return heap, pre, st


@pytest.mark.parametrize("case", ["continuous", "quantised", "duplicates"])
def test_te_prefix_path_equals_heap_path_whole_rows(gpu, case):
    """Up to 1024 valid frames the stores hold every frame of a query, so the
    select resolves the root from them alone (or lists the query when the
    root's bin is crowded); TE equals the heap pass's bit for bit."""
    rng = np.random.default_rng(31)
    if case == "duplicates":
        c = np.zeros((300, 5, 3), np.float32)
        c[:, :, 0] = rng.integers(0, 2, (300, 5))
        c[:, 4, :] += rng.standard_normal((300, 3)).astype(np.float32)
    else:
        c = synth.generate(synth.TrajSpec(natoms=5, nframes=700, nmodes=3, seed=12))
        if case == "quantised":
            c = np.round(c / 0.02).astype(np.float32) * np.float32(0.02)
    traj = ksg.center(c)
    for k, lag in ((1, 1), (5, 2), (100, 1), (128, 4)):
        heap, pre, st = _te_both(traj, k, lag)
        assert st["prefix_path"] and st["queries"] == 20 * (traj.shape[1] - lag)
        assert np.array_equal(heap, pre), (case, k, lag)


def test_te_prefix_path_beyond_the_stores(gpu, oracle_lib):
    """3000 frames: the stores hold the 1024 frames nearest in the target's
    own distances.  On a coupled trajectory most queries resolve there and
    the rest go to the heap kernel; with a target atom that barely moves
    (its distances tiny beside the source's) the neighbour lies anywhere in
    that order and most queries are listed.  Both equal the heap pass bit
    for bit, and the FP64 oracle on sampled pairs."""
    traj = ksg.center(synth.generate(synth.TrajSpec(natoms=4, nframes=3000, nmodes=3, seed=41)))
    heap, pre, st = _te_both(traj, 100, 1)
    assert np.array_equal(heap, pre)
    assert st["prefix_resolved"] > st["heap_listed"] > 0, st
    assert abs(pre[1, 2] - oracle_lib.te_pair(traj, 1, 2, 100, 1)) <= 1e-12
    rng = np.random.default_rng(43)
    c = rng.standard_normal((3000, 3, 3)).astype(np.float32)
    c[:, 0, :] *= np.float32(1e-3)
    traj = ksg.center(c)
    a1 = np.array([0, 0, 1], np.int32)
    a2 = np.array([1, 2, 0], np.int32)
    heap, pre, st = _te_both(traj, 20, 3, a1, a2)
    assert np.array_equal(heap, pre)
    assert st["heap_listed"] > st["queries"] // 2, st
    for p in range(3):
        ref = oracle_lib.te_pair(traj, int(a1[p]), int(a2[p]), 20, 3)
        assert abs(pre[p] - ref) <= 1e-12, p


def test_te_prefix_path_chosen_automatically(gpu, oracle_lib):
    """The automatic path takes the stores for more than 256 valid frames
    when the pairs average at least 8 per target atom (10 atoms: 9), not
    for a list of one pair per target, nor at 256 valid frames."""
    traj = ksg.center(synth.generate(synth.TrajSpec(natoms=10, nframes=2100, nmodes=4, seed=45)))
    mat = dense.te_matrix(traj, k=50, lag=2)
    assert dense.te_stats()["prefix_path"]
    try:
        dense.te_set_path(1)
        assert np.array_equal(mat, dense.te_matrix(traj, k=50, lag=2))
    finally:
        dense.te_set_path(0)
    dense.te_pairs(traj, [0, 1], [1, 0], k=50, lag=2)
    assert not dense.te_stats()["prefix_path"]
    dense.te_matrix(traj[:, :258], k=50, lag=2)
    assert not dense.te_stats()["prefix_path"]
    dense.te_matrix(traj[:, :259], k=50, lag=2)
    assert dense.te_stats()["prefix_path"]
    assert abs(mat[3, 7] - oracle_lib.te_pair(traj, 3, 7, 50, 2)) <= 1e-12


def test_te_prefix_path_over_two_launch_chunks(gpu):
    """2100 atoms x 100 frames: 4.4 M ordered pairs in two launch chunks, the
    boundary inside one target's pairs (its stores built in both chunks);
    the prefix path equals the heap pass on pairs on both sides of it."""
    traj = ksg.center(synth.generate(synth.TrajSpec(natoms=2100, nframes=100, nmodes=4, seed=47)))
    rng = np.random.default_rng(49)
    a1 = rng.choice(2100, 300, replace=False).astype(np.int32)
    # target 1998's pairs straddle the chunk boundary of the full matrix
    # (pair 4,194,304 of 2100 x 2099, launched target-major)
    a1 = np.concatenate([a1, np.full(2099, 1998, np.int32)])
    a2 = np.array([(t + 1 + rng.integers(0, 2098)) % 2100 for t in a1[:300]] +
                  [s for s in range(2100) if s != 1998], np.int32)
    heap, pre, st = _te_both(traj, 10, 1, a1, a2)
    assert np.array_equal(heap, pre)
    full = None
    try:
        dense.te_set_path(3)
        full = dense.te_matrix(traj, k=10, lag=1)
    finally:
        dense.te_set_path(0)
    assert np.array_equal(full[a1, a2], heap)
