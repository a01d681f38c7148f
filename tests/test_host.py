"""Host-side mirror and C-ABI library checks that need no GPU."""
import os
import re
import subprocess

import numpy as np
import pytest

from gmx_correlation_amd import _lib, ksg, synth
from tests.conftest import GOLDEN, ROOT, golden


def _declared_symbols():
    """Function names declared in include/*.h (C ABI) plus the two C++
    boundary symbols of include/kraskov_gpu.h."""
    names = set()
    src = open(os.path.join(ROOT, "include", "gmxcorr.h")).read()
    for m in re.finditer(r"^\s*(?:const\s+)?(?:[\w\*]+\s+)+\**(gmxcorr_\w+)\s*\(", src, re.M):
        names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(hiplib):
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    declared = _declared_symbols()
    assert declared, "no declarations parsed"
    assert declared <= exported, declared - exported
    assert set(_lib.CXX_SYMBOLS) <= exported
    assert set(_lib.C_SYMBOLS) == declared


def test_cxx_boundary_names_demangle(hiplib):
    out = subprocess.run(["c++filt"] + _lib.CXX_SYMBOLS, capture_output=True, text=True).stdout
    assert "kraskov_gpu_probe()" in out
    assert "kraskov_corrmatrix_gpu(t_kraskov const*, int, double*, int)" in out


def test_library_loads_and_probe_is_boolean(hiplib):
    assert _lib.probe() in (True, False)
    assert hiplib.gmxcorr_ksg_max_frames() >= 2000


def test_plan_rejects_bad_shapes_without_gpu(hiplib):
    import ctypes
    h = ctypes.c_void_p()
    assert hiplib.gmxcorr_ksg_plan_create(1, 100, 5, ctypes.byref(h)) == -1
    assert hiplib.gmxcorr_ksg_plan_create(4, 10, 10, ctypes.byref(h)) == -1
    assert hiplib.gmxcorr_ksg_plan_create(4, 10, 0, ctypes.byref(h)) == -1
    assert b"k" in hiplib.gmxcorr_last_error()


def test_variant_rules_without_gpu(hiplib):
    """The code-path table of the pair kernels (host logic, no GPU): kind 2
    up to 2048 frames, kind 3 above with a 512-, 768- or 1024-frame prefix, both
    for k + 1 <= 256; the streaming kinds by register tile; and the batch
    tuner's Python table agrees with the library on every batch path."""
    from gmx_correlation_amd import batch
    f = hiplib.gmxcorr_ksg_variant_fits
    assert f(2, 1, 2000, 100) == 1 and f(2, 1, 2049, 100) == 0
    assert f(2, 1, 2000, 255) == 1 and f(2, 1, 2000, 256) == 0
    assert f(3, 2, 5000, 100) == 1 and f(3, 4, 5000, 100) == 1
    assert f(3, 3, 5000, 100) == 1
    assert f(3, 1, 5000, 100) == 0 and f(3, 5, 5000, 100) == 0
    assert f(3, 4, 2048, 10) == 0 and f(3, 4, 8192, 10) == 1 and f(3, 4, 8193, 10) == 0
    assert f(0, 8, 2048, 10) == 1 and f(0, 8, 2049, 10) == 0 and f(0, 7, 1000, 10) == 0
    assert f(1, 2, 2048, 10) == 1 and f(1, 8, 8192, 10) == 1 and f(1, 9, 9000, 10) == 0
    for N in (300, 900, 2048, 2049, 5000, 8192):
        for k in (10, 100, 254, 255):
            fit = set(batch.fitting_paths(N, k))
            for kind, s4 in batch.BATCH_PATHS:
                assert ((kind, s4) in fit) == bool(f(kind, s4, N, k)), (N, k, kind, s4)


@pytest.mark.parametrize("spec", [synth.SMALL_CPU, synth.README_247])
def test_center_and_prepare_bitexact_vs_oracle(oracle_lib, spec):
    c = synth.generate(spec)
    t = ksg.center(c)
    assert np.array_equal(t, oracle_lib.center(c))
    rows, psi = ksg.kraskov_prepare(t)
    r2, p2 = oracle_lib.prepare(t)
    assert np.array_equal(rows, r2) and np.array_equal(psi, p2)


def test_prepare_matches_reference_fixture():
    rows, psi = ksg.kraskov_prepare(ksg.center(synth.generate(synth.SMALL_CPU)))
    assert np.array_equal(rows[:6], golden("small_cpu_prepare_rows_a0a1.npy"))
    assert np.array_equal(psi, golden("small_cpu_psi.npy"))


def test_write_matrix_byte_exact_vs_reference(tmp_path, oracle_lib):
    """correl.dat written by the reference's write_matrix (fixture) and by the
    mirror from the same r(MI) values are identical files."""
    c = synth.generate(synth.SMALL_CPU)[:, :8]
    rows, psi = ksg.kraskov_prepare(ksg.center(c))
    mi = oracle_lib.ksg_matrix(rows, psi, 8, 10)
    path = tmp_path / "correl.dat"
    ksg.write_matrix(ksg.pearsify(mi), str(path))
    assert path.read_bytes() == open(os.path.join(GOLDEN, "small8_correl.dat"), "rb").read()
    back = ksg.read_matrix(str(path))
    assert np.allclose(back, ksg.pearsify(mi), rtol=1e-5)


def test_pearsify_mirror(oracle_lib):
    """numpy's exp may differ from libm's by an ulp; correl.dat keeps 6 digits."""
    m = golden("small_cpu_k10_mi.npy")
    assert np.max(np.abs(ksg.pearsify(m) - oracle_lib.pearsify(m))) <= 1e-15


def test_kraskov_corrmatrix_input_checks():
    t = np.zeros((1, 50, 3))
    with pytest.raises(ValueError):
        ksg.kraskov_corrmatrix(t, 5)
    with pytest.raises(ValueError):
        ksg.kraskov_corrmatrix(np.zeros((3, 5, 3)), 5)
    with pytest.raises(RuntimeError):
        ksg.kraskov_corrmatrix(np.random.default_rng(0).random((3, 300, 3)), 129)
    with pytest.raises(RuntimeError):
        ksg.kraskov_corrmatrix(np.random.default_rng(0).random((3, 300, 3)), 5, use_gpu=False)


def test_batched_specs_deterministic():
    a = synth.batched_specs()
    b = synth.batched_specs()
    assert a == b and len(a) == 32
    assert all(50 <= s.natoms <= 600 and 1000 <= s.nframes <= 8000 and k in (10, 50, 100)
               for s, k in a)


@pytest.mark.parametrize("case", ["small8", "flat5"])
def test_write_xpm_byte_exact_vs_reference(tmp_path, oracle_lib, case):
    """XPM written by the reference's write_xpm_matrix (fixtures) and by the
    mirror from the same matrix are identical files; flat5 has equal
    off-diagonal values (the writer's 0..1 range fallback)."""
    if case == "small8":
        c = synth.generate(synth.SMALL_CPU)[:, :8]
        rows, psi = ksg.kraskov_prepare(ksg.center(c))
        m = oracle_lib.pearsify(oracle_lib.ksg_matrix(rows, psi, 8, 10))
        title, legend, fixture = "C-alpha correlation matrix", "r(MI)", "small8_correl.xpm"
    else:
        m = np.full((5, 5), 0.25)
        np.fill_diagonal(m, 1.0)
        title, legend, fixture = "flat correlation matrix", "MI (bits)", "flat5.xpm"
    path = tmp_path / "out.xpm"
    ksg.write_xpm_matrix(m, str(path), title, legend)
    assert path.read_bytes() == open(os.path.join(GOLDEN, fixture), "rb").read()


def test_dump_format_and_round_trip(tmp_path):
    """-dump text (src/gmx_correlation.cpp:225-244): header, then frame-major
    lines of an atom's three fluctuations in %.10g; read_dump inverts it."""
    t = np.array([[[0.1, -2.0, 1.0 / 3.0], [1e-12, 5.0, -0.25]],
                  [[7.0, 8.0, 9.0], [-1.5, 0.0, 123456789.123]]])   # [natoms=2][nframes=2][3]
    path = tmp_path / "traj.dump"
    ksg.write_dump(t, str(path))
    assert path.read_text().splitlines() == [
        "2 2",
        "0.1 -2 0.3333333333",
        "7 8 9",
        "1e-12 5 -0.25",
        "-1.5 0 123456789.1",
    ]
    back = ksg.read_dump(str(path))
    assert back.shape == (2, 2, 3)
    assert np.array_equal(back, np.vectorize(lambda v: float("%.10g" % v))(t))


def test_dump_feeds_prepare(tmp_path):
    """A dumped trajectory read back gives the same normalised rows to the
    precision of the text (10 significant digits)."""
    traj = ksg.center(synth.generate(synth.SMALL_CPU)[:, :6])
    path = tmp_path / "small.dump"
    ksg.write_dump(traj, str(path))
    rows_a, psi_a = ksg.kraskov_prepare(traj)
    rows_b, psi_b = ksg.kraskov_prepare(ksg.read_dump(str(path)))
    assert np.array_equal(psi_a, psi_b)
    assert np.max(np.abs(rows_a - rows_b)) < 1e-7


def test_read_dump_rejects_malformed(tmp_path):
    p = tmp_path / "bad.dump"
    p.write_text("2 3\n1 2 3\n")
    with pytest.raises(ValueError):
        ksg.read_dump(str(p))
    p.write_text("oops\n")
    with pytest.raises(ValueError):
        ksg.read_dump(str(p))


def test_plan_layout_memory_sizing_without_gpu(hiplib):
    """Host arithmetic of the plan layout: all rows resident when they fit
    the budget, memory waves of two atom blocks when they do not, the
    coordinate path above 8192 frames or below two atoms' rows."""
    P = lambda n: (n + 255) // 256 * 256
    per_sorted = lambda n: n * P(n) * 4 + n * ((64 + (max(n, 512) + 127) // 64 * 64) * 4
                                               + 64 * 4 + 512 * 2)
    per_store = lambda n: n * P(n) * 4 + n * ((64 + 1024) * 4 + 1024 * 2 + 16 * 4)
    # 247 x 2000 (8.98 GB of rows) with 10 GB: resident
    assert ksg.plan_layout(247, 2000, 100, 10 * 10**9) == {"slots": 247, "wave_atoms": 0,
                                                           "stores": 0}
    assert 247 * per_sorted(2000) == 8978944000
    # with 1 GB: two blocks of B atoms per wave
    B = 10**9 // (2 * per_sorted(2000))
    assert ksg.plan_layout(247, 2000, 100, 10**9) == {"slots": 2 * B, "wave_atoms": B,
                                                      "stores": 0}
    # long trajectories keep the prefix stores while a block holds two atoms
    lay = ksg.plan_layout(2000, 5000, 100, 200 * 10**9)
    assert lay["stores"] == 1 and lay["wave_atoms"] == 200 * 10**9 // (2 * per_store(5000))
    assert ksg.plan_layout(50, 5000, 300, 10**12)["stores"] == 0   # k + 1 > 256: no stores
    # above 8192 frames, or not even two atoms' rows: coordinates only
    assert ksg.plan_layout(247, 8193, 100, 10**12) == {"slots": 0, "wave_atoms": 0, "stores": 0}
    assert ksg.plan_layout(10, 600, 10, 10**6) == {"slots": 0, "wave_atoms": 0, "stores": 0}
    with pytest.raises(_lib.GmxcorrError):
        ksg.plan_layout(1, 600, 10, 10**9)


def test_te_argument_errors_before_the_device(hiplib):
    """The TE entry points reject a bad kernel choice, k above 128 when the
    heap or the prefix stores are forced, more than 16384 valid frames for
    the prefix stores, and non-finite coordinates before they look for a
    device (so these hold without one)."""
    from gmx_correlation_amd import dense
    traj = ksg.center(synth.generate(synth.TrajSpec(natoms=2, nframes=300, nmodes=2, seed=1)))
    with pytest.raises(_lib.GmxcorrError):
        dense.te_set_path(4)
    try:
        for path in (1, 3):
            dense.te_set_path(path)
            with pytest.raises(_lib.GmxcorrError, match="128"):
                dense.te_matrix(traj, k=129, lag=1)
        long = np.zeros((2, 16386, 3))
        with pytest.raises(_lib.GmxcorrError, match="16384"):
            dense.te_matrix(long, k=5, lag=1)
    finally:
        dense.te_set_path(0)
    bad = traj.copy()
    bad[1, 7, 2] = np.nan
    with pytest.raises(_lib.GmxcorrError, match="non-finite"):
        dense.te_matrix(bad, k=5, lag=1)
