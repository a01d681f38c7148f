"""GPU parity of the HIP KSG path against the oracle and the reference's
golden outputs, through the C ABI.

Tolerances: neighbour counts are integers and must match the FP64 brute-force
recount exactly; MI differs from the reference only by the order of the FP64
digamma sum (|dMI| <= 1e-9 asserted); the stated product tolerance is
|dr(MI)| <= 1e-3 (the reference's GPU-vs-CPU band, README.md:208-211)."""
import ctypes
import os

import numpy as np
import pytest

from gmx_correlation_amd import _lib, ksg, synth
from tests.conftest import GOLDEN, golden

pytestmark = pytest.mark.gpu

R_TOL = 1e-3
MI_TOL = 1e-9


@pytest.fixture(scope="module")
def gpu(hiplib):
    if not _lib.probe():
        pytest.fail("no gfx950 device visible to the HIP library")
    return hiplib


@pytest.fixture(scope="module")
def small_ws():
    rows, psi = ksg.kraskov_prepare(ksg.center(synth.generate(synth.SMALL_CPU)))
    return rows, psi


def test_wave_primitives(gpu):
    """DPP scan / reductions used by the kernels, against numpy."""
    out = np.zeros(67, np.int32)
    _lib.check(gpu.gmxcorr_selftest_wave(_lib.ptr(out, _lib._ip)))
    lanes = np.arange(64)
    v = (lanes * 5 + 3) % 7 + 1
    assert np.array_equal(out[:64], np.cumsum(v))
    assert out[64] == v.sum()
    assert out[65] == ((lanes * 37) % 64).max()
    assert out[66] == np.float32(((lanes * 11) % 64).max() * 0.5).view(np.int32)


def test_small_matrix_matches_reference(gpu, small_ws):
    rows, psi = small_ws
    with ksg.KsgPlan(64, 500, 10) as plan:
        plan.load(rows, psi)
        plan.run_all()
        mat = plan.matrix()
        st = plan.stats()
    ref = golden("small_cpu_k10_mi.npy")
    assert np.max(np.abs(mat - ref)) <= MI_TOL
    assert np.max(np.abs(ksg.pearsify(mat) - ksg.pearsify(ref))) <= R_TOL
    assert np.all(np.diag(mat) == 2000.0)
    assert st["npairs"] == 2016


@pytest.mark.parametrize("k", [1, 3, 25, 499])
def test_other_k_match_reference(gpu, small_ws, k):
    rows, psi = small_ws
    pairs = ((1, 0), (17, 5), (63, 62))
    with ksg.KsgPlan(64, 500, k) as plan:
        plan.load(rows, psi)
        got = plan.run_pairs([p[0] for p in pairs], [p[1] for p in pairs])
    assert np.max(np.abs(got - golden(f"small_cpu_k{k}_pairs.npy"))) <= MI_TOL


def test_counts_bitexact_vs_fp64_recount(gpu, small_ws, oracle_lib):
    rows, psi = small_ws
    a1 = np.array([1, 40, 63, 20]); a2 = np.array([0, 7, 2, 19])
    with ksg.KsgPlan(64, 500, 10) as plan:
        plan.load(rows, psi)
        mi, d = plan.run_pairs(a1, a2, detail=True)
    for p in range(len(a1)):
        v, ref = oracle_lib.ksg_pair(rows, psi, int(a1[p]), int(a2[p]), 10, method="brute",
                                     detail=True)
        assert np.array_equal(d["nx"][p], ref["nx"])
        assert np.array_equal(d["ny"][p], ref["ny"])
        assert abs(mi[p] - v) <= 1e-12


def test_symmetry_exact(gpu, small_ws):
    """Swapping the atoms swaps nx and ny row by row: MI is bit-identical."""
    rows, psi = small_ws
    rng = np.random.default_rng(3)
    a1 = rng.integers(0, 64, 50); a2 = (a1 + rng.integers(1, 64, 50)) % 64
    with ksg.KsgPlan(64, 500, 10) as plan:
        plan.load(rows, psi)
        x = plan.run_pairs(a1, a2)
        y = plan.run_pairs(a2, a1)
    assert np.array_equal(x, y)


def test_ties_recounted_in_fp64(gpu, oracle_lib):
    """Quantised coordinates: exact ties at the k-th distance are detected,
    recounted in FP64 and counted; counts equal the FP64 recount."""
    from tests.golden.make_golden import tie_spec
    rows, psi = ksg.kraskov_prepare(ksg.center(tie_spec()))
    a1, a2 = np.tril_indices(12, -1)
    with ksg.KsgPlan(12, 80, 4) as plan:
        plan.load(rows, psi)
        mi, d = plan.run_pairs(a1, a2, detail=True)
        st = plan.stats()
    assert st["fixup_rows"] > 0 and st["ties"] > 0
    for p in range(len(a1)):
        v, ref = oracle_lib.ksg_pair(rows, psi, int(a1[p]), int(a2[p]), 4, method="brute",
                                     detail=True)
        assert np.array_equal(d["nx"][p], ref["nx"]), p
        assert np.array_equal(d["ny"][p], ref["ny"]), p
        assert np.array_equal(d["flags"][p] == 2, ref["tie"].astype(bool)), p
        assert abs(mi[p] - v) <= 1e-12
    # against the reference's own grid-order tie choice the band still holds
    ref = golden("ties12_k4_mi.npy")[a1, a2]
    print("max |dr| vs reference with ties:",
          np.max(np.abs(ksg.pearsify(mi) - ksg.pearsify(ref))))


def test_readme_workload_matches_reference(gpu, oracle_lib):
    """247 atoms x 2000 frames, k = 100: the whole upper triangle against the
    reference's float64 CPU output, plus per-frame counts of sampled pairs
    against the FP64 recount."""
    path = os.path.join(GOLDEN, "readme247_k100_mi_triu.npy")
    if not os.path.exists(path):
        pytest.skip("247-atom golden fixture not generated")
    rows, psi = ksg.kraskov_prepare(ksg.center(synth.generate(synth.README_247)))
    with ksg.KsgPlan(247, 2000, 100) as plan:
        plan.load(rows, psi)
        plan.run_all()
        mat = plan.matrix()
        st = plan.stats()
        a1 = np.array([1, 200, 246]); a2 = np.array([0, 13, 245])
        _, d = plan.run_pairs(a1, a2, detail=True)
    iu = np.triu_indices(247, 1)
    got = mat[iu]
    ref = np.load(path, allow_pickle=False)
    dmi = np.abs(got - ref)
    dr = np.max(np.abs(ksg.pearsify(got) - ksg.pearsify(ref)))
    off = np.nonzero(dmi > MI_TOL)[0]
    print(f"247x2000 k=100: max|dMI|={dmi.max():.3e} max|dr|={dr:.3e} "
          f"pairs off the reference beyond {MI_TOL}: {len(off)} stats={st}")
    assert dr <= R_TOL
    # float32 inputs repeat coordinate values, so some frames have exact FP64
    # ties at the k-th distance (st["ties"]); only pairs holding such a row
    # may leave the reference (grid order vs all-ties rule), and those must
    # equal the tie-aware FP64 brute-force recount
    assert st["ties"] > 0 and len(off) <= st["ties"]
    for q in off[:6]:
        b1, b2 = int(iu[1][q]), int(iu[0][q])     # stored as (a1 > a2)
        v, ref_d = oracle_lib.ksg_pair(rows, psi, b1, b2, 100, method="brute", detail=True)
        assert ref_d["tie"].any()
        assert abs(got[q] - v) <= 1e-12
    for p in range(len(a1)):
        _, ref_d = oracle_lib.ksg_pair(rows, psi, int(a1[p]), int(a2[p]), 100, method="brute",
                                       detail=True)
        assert np.array_equal(d["nx"][p], ref_d["nx"])
        assert np.array_equal(d["ny"][p], ref_d["ny"])


@pytest.mark.parametrize("kind,s4", [(0, 8), (2, 1)])
def test_single_wave_variants_bitexact_2000(gpu, oracle_lib, kind, s4):
    """Both one-wave-per-row paths at 2000 frames, k = 100 (kind 0 streams
    both distance rows, kind 2 reads the sorted prefix of the sparser atom):
    per-frame counts equal the FP64 brute-force recount."""
    c = synth.generate(synth.TrajSpec(natoms=6, nframes=2000, nmodes=5, seed=7))
    rows, psi = ksg.kraskov_prepare(ksg.center(c))
    a1 = np.array([1, 5, 3]); a2 = np.array([0, 2, 4])
    with ksg.KsgPlan(6, 2000, 100) as plan:
        plan.load(rows, psi)
        plan.set_variant(kind, s4)
        mi, d = plan.run_pairs(a1, a2, detail=True)
    for p in range(len(a1)):
        v, ref = oracle_lib.ksg_pair(rows, psi, int(a1[p]), int(a2[p]), 100, method="brute",
                                     detail=True)
        assert np.array_equal(d["nx"][p], ref["nx"]), p
        assert np.array_equal(d["ny"][p], ref["ny"]), p
        assert abs(mi[p] - v) <= 1e-12


def test_prefix_overflow_rows_take_streaming_path(gpu, oracle_lib):
    """Independent atoms at k = 200: the (k+1)-th joint neighbour lies beyond
    the 512 nearest frames of either atom in most rows (about sqrt(201*2000)
    = 634 frames are needed), so the sorted-prefix path hands those rows to
    the long-row kernel (streaming path); counts still equal the FP64
    recount."""
    rng = np.random.default_rng(11)
    coords = rng.standard_normal((2000, 4, 3)).astype(np.float32)
    rows, psi = ksg.kraskov_prepare(ksg.center(coords))
    a1 = np.array([1, 3]); a2 = np.array([0, 2])
    with ksg.KsgPlan(4, 2000, 200) as plan:
        plan.load(rows, psi)
        mi, d = plan.run_pairs(a1, a2, detail=True)
        st = plan.stats()
    assert st["long_rows"] > 1000
    for p in range(len(a1)):
        v, ref = oracle_lib.ksg_pair(rows, psi, int(a1[p]), int(a2[p]), 200, method="brute",
                                     detail=True)
        assert np.array_equal(d["nx"][p], ref["nx"]), p
        assert np.array_equal(d["ny"][p], ref["ny"]), p
        assert abs(mi[p] - v) <= 1e-12


@pytest.mark.parametrize("variant", [None, (3, 2), (3, 3), (3, 4), (1, 0)])
@pytest.mark.parametrize("N,k,npairs", [(3000, 50, 2), (5000, 100, 1), (8000, 10, 1)])
def test_long_trajectories(gpu, oracle_lib, N, k, npairs, variant):
    """N > 2048: the default path, the prefix-store path (kind 3) over 512,
    768 and 1024 prefix frames, and the 4-waves-per-row streaming kernel
    (kind 1); counts and MI against the FP64 brute-force recount."""
    c = synth.generate(synth.TrajSpec(natoms=5, nframes=N, nmodes=4, seed=N))
    rows, psi = ksg.kraskov_prepare(ksg.center(c))
    a1 = np.array([1, 4])[:npairs]; a2 = np.array([0, 2])[:npairs]
    with ksg.KsgPlan(5, N, k) as plan:
        plan.load(rows, psi)
        if variant is not None:
            kind, s4 = variant
            plan.set_variant(kind, s4 if kind == 3 else (N + 1023) // 1024)
        mi, d = plan.run_pairs(a1, a2, detail=True)
    for p in range(npairs):
        v, ref = oracle_lib.ksg_pair(rows, psi, int(a1[p]), int(a2[p]), k, method="brute",
                                     detail=True)
        assert np.array_equal(d["nx"][p], ref["nx"])
        assert np.array_equal(d["ny"][p], ref["ny"])
        assert abs(mi[p] - v) <= 1e-12


def test_cxx_boundary_entry_point(gpu, small_ws):
    """Call the exported kraskov_corrmatrix_gpu(const t_kraskov*, int,
    double*, int) with a t_kraskov laid out as lib/correlation_core.h:45-52."""
    rows, psi = small_ws

    class TKraskov(ctypes.Structure):
        _fields_ = [("x", ctypes.POINTER(ctypes.POINTER(ctypes.c_double))),
                    ("psi", ctypes.POINTER(ctypes.c_double)),
                    ("scal", ctypes.POINTER(ctypes.c_double)),
                    ("N", ctypes.c_int), ("dim", ctypes.c_int)]

    row_arrays = [np.ascontiguousarray(r) for r in rows]
    ptrs = (ctypes.POINTER(ctypes.c_double) * len(row_arrays))(
        *[r.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for r in row_arrays])
    scal = np.ones(len(row_arrays))
    kr = TKraskov(ptrs, psi.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                  scal.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 500, 192)
    mat = np.zeros((64, 64))
    fn = getattr(gpu, "_Z22kraskov_corrmatrix_gpuPK9t_kraskoviPdi")
    fn.argtypes = [ctypes.POINTER(TKraskov), ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                   ctypes.c_int]
    fn(ctypes.byref(kr), 64, mat.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 10)
    assert np.max(np.abs(mat - golden("small_cpu_k10_mi.npy"))) <= MI_TOL
    probe = getattr(gpu, "_Z17kraskov_gpu_probev")
    probe.restype = ctypes.c_bool
    assert probe()


def test_one_shot_matches_plan(gpu, small_ws):
    traj = ksg.center(synth.generate(synth.SMALL_CPU))
    mat = ksg.kraskov_corrmatrix(traj, 10)
    assert np.max(np.abs(mat - golden("small_cpu_k10_mi.npy"))) <= MI_TOL


def test_rejects_constant_component(gpu):
    traj = np.random.default_rng(0).standard_normal((3, 200, 3))
    traj[1, :, 2] = 0.0
    with pytest.raises(_lib.GmxcorrError) as e:
        ksg.kraskov_corrmatrix(traj, 5)
    assert e.value.code == -1


def test_rejects_rows_that_are_not_zscores(gpu, small_ws):
    """The pair kernels assume every distance is below 2048 (z-scored rows)."""
    rows, psi = small_ws
    wide = rows.copy()
    wide[5, 7] += 5000.0
    with ksg.KsgPlan(64, 500, 10) as plan:
        with pytest.raises(_lib.GmxcorrError) as e:
            plan.load(wide, psi)
        assert "spread" in str(e.value)
        plan.load(rows, psi)            # the plan stays usable


def test_rejects_bad_pair_indices(gpu, small_ws):
    rows, psi = small_ws
    with ksg.KsgPlan(64, 500, 10) as plan:
        plan.load(rows, psi)
        with pytest.raises(_lib.GmxcorrError):
            plan.run_pairs([3], [3])
        with pytest.raises(_lib.GmxcorrError):
            plan.run_pairs([64], [0])


@pytest.mark.parametrize("k,s4", [(250, 4), (180, 3), (100, 2)])
def test_long_prefix_overflow_rows_take_cooperative_path(gpu, oracle_lib, k, s4):
    """Independent atoms at 6000 frames: the (k+1)-th joint neighbour lies
    about sqrt((k+1)*6000) frames deep in either atom's sorted row (1227 at
    k = 250, 1042 at k = 180, 778 at k = 100), beyond the prefix read (1024,
    768 or 512 frames) in most rows, so those rows run the cooperative
    streaming row; counts equal the FP64 recount."""
    rng = np.random.default_rng(12)
    coords = rng.standard_normal((6000, 4, 3)).astype(np.float32)
    rows, psi = ksg.kraskov_prepare(ksg.center(coords))
    a1 = np.array([1, 3]); a2 = np.array([0, 2])
    with ksg.KsgPlan(4, 6000, k) as plan:
        plan.load(rows, psi)
        plan.set_variant(3, s4)
        mi, d = plan.run_pairs(a1, a2, detail=True)
        st = plan.stats()
    assert st["long_rows"] > 1000
    for p in range(len(a1)):
        v, ref = oracle_lib.ksg_pair(rows, psi, int(a1[p]), int(a2[p]), k, method="brute",
                                     detail=True)
        assert np.array_equal(d["nx"][p], ref["nx"]), p
        assert np.array_equal(d["ny"][p], ref["ny"]), p
        assert abs(mi[p] - v) <= 1e-12


@pytest.mark.parametrize("s4", [2, 4])
def test_long_trajectories_with_ties(gpu, oracle_lib, s4):
    """Coordinates on a quarter-unit grid at 3000 frames: distance rows hold
    long runs of equal values, at the prefix boundary too (the prefix keeps
    the first tied frames in frame order), and the joint and marginal radii
    tie; counts equal the FP64 recount."""
    rng = np.random.default_rng(13)
    coords = np.round(rng.standard_normal((3000, 4, 3)) * 4.0) / 4.0
    rows, psi = ksg.kraskov_prepare(ksg.center(coords.astype(np.float32)))
    a1 = np.array([1, 3, 2]); a2 = np.array([0, 2, 0])
    with ksg.KsgPlan(4, 3000, 20) as plan:
        plan.load(rows, psi)
        plan.set_variant(3, s4)
        mi, d = plan.run_pairs(a1, a2, detail=True)
        st = plan.stats()
    assert st["ties"] > 0
    for p in range(len(a1)):
        v, ref = oracle_lib.ksg_pair(rows, psi, int(a1[p]), int(a2[p]), 20, method="brute",
                                     detail=True)
        assert np.array_equal(d["nx"][p], ref["nx"]), p
        assert np.array_equal(d["ny"][p], ref["ny"]), p
        assert abs(mi[p] - v) <= 1e-12


@pytest.mark.parametrize("natoms,N,k", [
    (5, 2048, 255),   # sorted-prefix path at its largest N and k + 1 = 256
    (5, 2049, 100),   # prefix-store path at its smallest N
    (5, 8192, 5),     # the largest N
    (7, 300, 10),     # rows shorter than the 512-frame prefix; odd atom count
    (3, 513, 256),    # k + 1 = 257: beyond the prefix paths, streaming kernel
    (2, 102, 100),    # N = k + 2, a single pair
])
def test_boundary_shapes_all_pairs(gpu, oracle_lib, natoms, N, k):
    """All pairs through the tiled launch (2x2 atom tiles plus the ragged
    remainder) at the edges of each path's range; every MI equals the
    oracle's grid restatement to 1e-12 nats, and the first pair's per-frame
    counts equal the FP64 brute-force recount."""
    c = synth.generate(synth.TrajSpec(natoms=natoms, nframes=N, nmodes=3, seed=N + k))
    rows, psi = ksg.kraskov_prepare(ksg.center(c))
    a1, a2 = np.tril_indices(natoms, -1)
    order = np.lexsort((a2, a1))   # the reference's order: a1 = 1.., a2 < a1
    a1, a2 = a1[order], a2[order]
    with ksg.KsgPlan(natoms, N, k) as plan:
        plan.load(rows, psi)
        mi = plan.run_all()
        _, d = plan.run_pairs(a1[:1], a2[:1], detail=True)
    ref = oracle_lib.ksg_pairs(rows, psi, k, a1, a2)
    assert np.max(np.abs(mi - ref)) <= 1e-12
    _, rd = oracle_lib.ksg_pair(rows, psi, int(a1[0]), int(a2[0]), k, method="brute", detail=True)
    assert np.array_equal(d["nx"][0], rd["nx"])
    assert np.array_equal(d["ny"][0], rd["ny"])


@pytest.mark.parametrize("N,k", [(2000, 100), (3000, 50)])
def test_xtc_precision_coordinates(gpu, oracle_lib, N, k):
    """Coordinates rounded to 0.001 nm, the precision XTC files store: the
    distance rows hold many equal values and the k-th neighbour ties often.
    Sorted-prefix path at 2000 frames, prefix-store path at 3000; per-frame
    counts equal the FP64 brute-force recount."""
    c = synth.generate(synth.TrajSpec(natoms=4, nframes=N, nmodes=3, seed=31))
    c = (np.round(c.astype(np.float64) * 1000.0) / 1000.0).astype(np.float32)
    rows, psi = ksg.kraskov_prepare(ksg.center(c))
    a1 = np.array([1, 3, 2]); a2 = np.array([0, 2, 1])
    with ksg.KsgPlan(4, N, k) as plan:
        plan.load(rows, psi)
        mi, d = plan.run_pairs(a1, a2, detail=True)
        st = plan.stats()
    assert st["fixup_rows"] > 0
    for p in range(len(a1)):
        v, ref = oracle_lib.ksg_pair(rows, psi, int(a1[p]), int(a2[p]), k, method="brute",
                                     detail=True)
        assert np.array_equal(d["nx"][p], ref["nx"]), p
        assert np.array_equal(d["ny"][p], ref["ny"]), p
        assert abs(mi[p] - v) <= 1e-12


def test_all_pairs_over_two_launches(gpu, oracle_lib):
    """1500 atoms x 64 frames: 1,124,250 pairs in 281,063 2x2 tiles times 128
    row ranges, 36.0 M blocks, more than one dispatch carries (2^22 blocks of
    256 work-items), so the sorted-prefix kernel goes out in nine launches;
    3000 pairs spread over the matrix equal the oracle."""
    c = synth.generate(synth.TrajSpec(natoms=1500, nframes=64, nmodes=6, seed=64))
    rows, psi = ksg.kraskov_prepare(ksg.center(c))
    with ksg.KsgPlan(1500, 64, 5) as plan:
        plan.load(rows, psi)
        mi = plan.run_all()
    a1, a2 = np.tril_indices(1500, -1)
    order = np.lexsort((a2, a1))   # the reference's order: a1 = 1.., a2 < a1
    a1, a2 = a1[order], a2[order]
    pick = np.linspace(0, len(a1) - 1, 3000).astype(np.int64)
    ref = oracle_lib.ksg_pairs(rows, psi, 5, a1[pick], a2[pick])
    assert np.max(np.abs(mi[pick] - ref)) <= 1e-12


def test_long_trajectory_all_pairs_over_two_launches(gpu, oracle_lib):
    """720 atoms x 2100 frames (prefix-store path, 512-frame prefix at k =
    10): 64,710 2x2 tiles times 132 chunks of 16 rows, 8.54 M blocks, so the
    kernel goes out in three launches (every tile has chunks in more than
    one).  600 pairs spread over the matrix equal the oracle's grid
    restatement, except pairs with an exact FP64 tie at the k-th distance,
    where the GPU counts every tied frame: those equal the FP64 brute-force
    recount, which follows the same rule."""
    c = synth.generate(synth.TrajSpec(natoms=720, nframes=2100, nmodes=6, seed=21))
    rows, psi = ksg.kraskov_prepare(ksg.center(c))
    with ksg.KsgPlan(720, 2100, 10) as plan:
        plan.load(rows, psi)
        mi = plan.run_all()
    a1, a2 = np.tril_indices(720, -1)
    order = np.lexsort((a2, a1))
    a1, a2 = a1[order], a2[order]
    pick = np.linspace(0, len(a1) - 1, 600).astype(np.int64)
    ref = oracle_lib.ksg_pairs(rows, psi, 10, a1[pick], a2[pick])
    off = np.nonzero(np.abs(mi[pick] - ref) > 1e-12)[0]
    assert len(off) <= len(pick) // 10
    for i in off:
        p = int(pick[i])
        v = oracle_lib.ksg_pair(rows, psi, int(a1[p]), int(a2[p]), 10, method="brute")
        assert abs(mi[p] - v) <= 1e-12, p
