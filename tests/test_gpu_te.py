"""GPU transfer entropy against the reference's own output (golden, 16 atoms
x 300 frames, k = 5, lags 1/5/10) and the FP64 oracle.  Distances, the k-th
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
