"""The reference's unchanged dispatcher linked against libgmxcorr_hip.so.

oracle/Makefile (target `dispatch`) compiles lib/correlation_core.cpp and
lib/kraskov.cpp where they lie, with GMX_CORRELATION_USE_CUDA defined, plus
oracle/dispatch_link.cpp, and links the HIP library in place of the CUDA
backend (CMakeLists.txt:127-135).  The link resolving at all shows that the
library exports the two C++ symbols the dispatcher calls
(lib/kraskov_gpu.h:20,32).  Without a visible gfx950 device the probe is
false and the dispatcher's own CPU loop must run
(lib/correlation_core.cpp:337-348): its matrix equals the reference's golden
output.  The executable is built from reference sources, so it exists only
where the reference checkout does, and it is never taken to a GPU machine.
"""
import os
import subprocess

import numpy as np
import pytest

from gmx_correlation_amd import _lib, synth
from tests.conftest import ROOT, golden

REF = "/root/reference/lib"
EXE = os.path.join(ROOT, "oracle", "_ref", "dispatch_link")

pytestmark = pytest.mark.skipif(not os.path.isdir(REF),
                                reason="needs the reference checkout (CPU container only)")


@pytest.fixture(scope="module")
def dispatch_exe(hiplib):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "dispatch"], check=True)
    assert os.path.exists(EXE)
    return EXE


def _run(exe, coords, k, use_gpu, tmp_path):
    T, n, _ = coords.shape
    cin = tmp_path / "coords.f32"
    mout = tmp_path / "mat.f64"
    np.ascontiguousarray(coords, dtype=np.float32).tofile(cin)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1") if not _lib.probe() else dict(os.environ)
    p = subprocess.run([exe, str(cin), str(T), str(n), str(k), str(int(use_gpu)), "2", str(mout)],
                       capture_output=True, text=True, env=env, timeout=300)
    assert p.returncode == 0, p.stderr
    return p, np.fromfile(mout, dtype=np.float64).reshape(n, n)


def test_dispatcher_links_both_gpu_symbols(dispatch_exe):
    out = subprocess.run(["nm", "-u", "-C", dispatch_exe], capture_output=True, text=True,
                         check=True).stdout
    assert "kraskov_gpu_probe()" in out
    assert "kraskov_corrmatrix_gpu(t_kraskov const*, int, double*, int)" in out
    ldd = subprocess.run(["ldd", dispatch_exe], capture_output=True, text=True).stdout
    assert "libgmxcorr_hip.so" in ldd


@pytest.mark.skipif(_lib.probe() if os.path.exists(_lib.LIB_PATH) else False,
                    reason="a gfx950 device is visible: the GPU branch would run")
def test_dispatcher_falls_back_to_its_cpu_loop_without_a_device(dispatch_exe, tmp_path):
    coords = synth.generate(synth.SMALL_CPU)
    p, mat = _run(dispatch_exe, coords, 10, True, tmp_path)
    assert "probe=0" in p.stdout
    assert "falling back to CPU" in p.stderr
    want = golden("small_cpu_k10_mi.npy")
    assert np.array_equal(mat, want)


def test_dispatcher_k_above_128_takes_the_cpu_loop(dispatch_exe, tmp_path):
    """k > 128 never reaches the library (lib/correlation_core.cpp:339-340)."""
    spec = synth.TrajSpec(natoms=6, nframes=300, nmodes=3, seed=5)
    coords = synth.generate(spec)
    p, mat = _run(dispatch_exe, coords, 150, True, tmp_path)
    assert "k ≤ 128" in p.stderr or "k ≤ 128" in p.stderr
    assert np.all(np.diag(mat) == 2000.0)
    assert np.allclose(mat, mat.T)
