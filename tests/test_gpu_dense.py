"""GPU parity of the Gaussian and GCMI paths (split-precision FP16 MFMA
second-moment GEMM with FP32 accumulation, FP64 log-determinants).

Tolerances, stated against FP64: the GEMM's entries within 2e-6 of the FP64
product relative to sqrt(C_ii C_jj); MI within 1e-4 nats of the reference's
float64 output (the GEMM error propagated through 3x3/6x6 log-dets); copula
scores within 1e-12 of the FP64 oracle."""
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


@pytest.fixture(scope="module")
def traj():
    return ksg.center(synth.generate(synth.SMALL_CPU))


def _rel_cov_err(got, want):
    d = np.sqrt(np.outer(np.diag(want), np.diag(want)))
    return np.max(np.abs(got - want) / d)


def test_gauss_matches_reference(gpu, traj, oracle_lib):
    mat, cov, st = dense.gauss_corrmatrix(traj, return_cov=True, stats=True)
    X = traj.transpose(1, 0, 2).reshape(traj.shape[1], -1)
    want_cov = X.T @ X / X.shape[0]
    err = _rel_cov_err(cov, want_cov)
    ref = golden("small_cpu_gauss_mi.npy")
    dmi = np.max(np.abs(mat - ref))
    print(f"gauss 64x500: cov rel err {err:.2e}, max|dMI| {dmi:.2e}, stats {st}")
    assert err <= 2e-6
    assert dmi <= 1e-4
    assert np.all(np.diag(mat) == 2000.0)
    # the FP64 log-det stage alone is exact: oracle on the GPU's own covariance
    assert np.max(np.abs(oracle_lib.gauss_matrix_from_cov(cov, 64) - mat)) <= 1e-10


def test_gcmi_matches_reference(gpu, traj, oracle_lib):
    mat, cov, z = dense.gcmi_corrmatrix(traj, return_cov=True, return_z=True)
    zo = oracle_lib.copula(traj)
    assert np.max(np.abs(z - zo)) <= 1e-12
    want = zo @ zo.T / zo.shape[1]
    assert _rel_cov_err(cov, want) <= 2e-6
    ref = golden("small_cpu_gcmi_mi.npy")
    dmi = np.max(np.abs(mat - ref))
    print(f"gcmi 64x500: max|dMI| {dmi:.2e}")
    assert dmi <= 1e-4


def test_gcmi_copula_ties(gpu, oracle_lib):
    """Quantised coordinates: tied values share the mean score of their run."""
    from tests.golden.make_golden import tie_spec
    t = ksg.center(tie_spec())
    _, z = dense.gcmi_corrmatrix(t, return_z=True)
    assert np.max(np.abs(z - oracle_lib.copula(t))) <= 1e-12


def test_gauss_closed_form(gpu):
    """Large-N Gaussian sample: MI approaches -1.5 log(1 - r^2)."""
    rng = np.random.default_rng(4)
    N, r = 20000, 0.6
    x = rng.standard_normal((N, 3))
    y = r * x + np.sqrt(1 - r * r) * rng.standard_normal((N, 3))
    t = np.stack([x, y])
    t = t - t.mean(axis=1, keepdims=True)
    mi = dense.gauss_corrmatrix(t)[0, 1]
    assert abs(mi - (-1.5 * np.log(1 - r * r))) < 0.02


def test_gcmi_rejects_short(gpu):
    with pytest.raises(ValueError):
        dense.gcmi_corrmatrix(np.zeros((3, 5, 3)))
    with pytest.raises(_lib.GmxcorrError):
        dense.gauss_corrmatrix(np.zeros((1, 50, 3)))
