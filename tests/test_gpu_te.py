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
