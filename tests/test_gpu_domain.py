"""The boundary's whole input domain on the GPU, through the C ABI.

The reference kernel accepts any frame count and needs O(natoms * N) device
memory (lib/kraskov_gpu.cu:129,147).  Here:
- frame counts above the distance-row kernels' 8192 run on the coordinate
  path (kind 4): every row in FP64 from the coordinates;
- plans whose distance rows do not fit the row budget run in memory waves of
  two atom blocks;
- the coordinate path on any shape equals the other paths row by row;
- the device-pointer entry point equals the plan's coordinate path.
Counts are integers and must equal the FP64 brute-force recount exactly; MI
within 1e-12 nats of the recount's MI (same per-row terms, another FP64
summation order)."""
import concurrent.futures as cf
import ctypes

import numpy as np
import pytest

from gmx_correlation_amd import _lib, ksg, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu(hiplib):
    if not _lib.probe():
        pytest.fail("no gfx950 device visible to the HIP library")
    return hiplib


def _brute_all(oracle_lib, rows, psi, a1, a2, k):
    """FP64 brute-force recount of several pairs in parallel threads (the
    oracle releases the GIL inside ctypes)."""
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        futs = [ex.submit(oracle_lib.ksg_pair, rows, psi, int(x), int(y), k, "brute", True)
                for x, y in zip(a1, a2)]
        return [f.result() for f in futs]


def _check_counts(mi, d, ref, tol=1e-12):
    for p, (v, r) in enumerate(ref):
        assert np.array_equal(d["nx"][p], r["nx"]), p
        assert np.array_equal(d["ny"][p], r["ny"]), p
        assert np.array_equal(d["flags"][p] == 2, r["tie"].astype(bool)), p
        assert abs(mi[p] - v) <= tol, (p, mi[p], v)


@pytest.mark.parametrize("nframes,k", [(8193, 10), (10000, 100)])
def test_frames_above_8192_on_the_coordinate_path(gpu, oracle_lib, nframes, k):
    """N = 8193 and 10^4 (the reference kernel loops over any N): per-row
    counts bit-exact against the FP64 brute-force recount."""
    spec = synth.TrajSpec(natoms=4, nframes=nframes, nmodes=3, seed=nframes, n_switch=1)
    print(f"domain: coordinate path {spec} k={k}", flush=True)
    rows, psi = ksg.kraskov_prepare(ksg.center(synth.generate(spec)))
    a1 = np.array([1, 3, 3]); a2 = np.array([0, 0, 2])
    with ksg.KsgPlan(4, nframes, k) as plan:
        plan.load(rows, psi)
        mi_all = plan.run_all()
        st = plan.stats()
        assert st["kind"] == 4 and st["slots"] == 0 and st["row_bytes"] == 0
        mi, d = plan.run_pairs(a1, a2, detail=True)
    _check_counts(mi, d, _brute_all(oracle_lib, rows, psi, a1, a2, k))
    # all-pairs order (a1 = 1.., a2 < a1): (1,0) is 0, (3,0) is 3, (3,2) is 5
    assert np.array_equal(mi_all[[0, 3, 5]], mi)


def test_cxx_boundary_above_8192_frames(gpu, oracle_lib):
    """kraskov_corrmatrix_gpu() with N = 8200: computed, not refused."""
    spec = synth.TrajSpec(natoms=3, nframes=8200, nmodes=2, seed=5)
    rows, psi = ksg.kraskov_prepare(ksg.center(synth.generate(spec)))

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
                  scal.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 8200, 9)
    mat = np.zeros((3, 3))
    fn = getattr(gpu, "_Z22kraskov_corrmatrix_gpuPK9t_kraskoviPdi")
    fn.argtypes = [ctypes.POINTER(TKraskov), ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                   ctypes.c_int]
    fn(ctypes.byref(kr), 3, mat.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 20)
    assert np.all(np.diag(mat) == 2000.0)
    ref = _brute_all(oracle_lib, rows, psi, [1, 2, 2], [0, 0, 1], 20)
    for (x, y), (v, _) in zip(((1, 0), (2, 0), (2, 1)), ref):
        assert abs(mat[x, y] - v) <= 1e-12 and mat[x, y] == mat[y, x]


@pytest.mark.parametrize("kind", [2, 0])
def test_coordinate_path_equals_row_paths(gpu, kind):
    """Kind 4 against the distance-row paths on 64 atoms x 500 frames: the
    same counts in every row, MI to 1e-12."""
    rows, psi = ksg.kraskov_prepare(ksg.center(synth.generate(synth.SMALL_CPU)))
    rng = np.random.default_rng(8)
    a1 = rng.integers(0, 64, 40); a2 = (a1 + rng.integers(1, 64, 40)) % 64
    with ksg.KsgPlan(64, 500, 10) as plan:
        plan.load(rows, psi)
        plan.set_variant(kind, 2 if kind == 0 else 0)
        m0, d0 = plan.run_pairs(a1, a2, detail=True)
        plan.set_variant(4)
        m4, d4 = plan.run_pairs(a1, a2, detail=True)
        assert plan.stats()["kind"] == 4
    assert np.array_equal(d0["nx"], d4["nx"]) and np.array_equal(d0["ny"], d4["ny"])
    assert np.array_equal(d0["flags"] == 2, d4["flags"] == 2)
    assert np.max(np.abs(m0 - m4)) <= 1e-12


def _per_atom_bytes(N, stores):
    P = (N + 255) // 256 * 256
    b = N * P * 4
    if N <= 2048:
        sp = 64 + (max(N, 512) + 64 + 63) // 64 * 64
        b += N * (sp * 4 + 64 * 4 + 512 * 2)
    if stores:
        b += N * ((64 + 1024) * 4 + 1024 * 2 + 16 * 4)
    return b


@pytest.mark.parametrize("natoms,nframes,k,B", [(13, 600, 20, 3), (9, 3000, 50, 2)])
def test_memory_waves_equal_resident_rows(gpu, oracle_lib, natoms, nframes, k, B):
    """A row budget of 2B atoms' rows: the pairs run in memory waves over
    blocks of B atoms (sorted-row path at 600 frames, prefix-store path at
    3000) and give the resident plan's MI; per-row counts of sampled pairs
    equal the FP64 recount."""
    spec = synth.TrajSpec(natoms=natoms, nframes=nframes, nmodes=4, seed=natoms * 7,
                          n_switch=2, n_rotor=1)
    print(f"domain: memory waves {spec} k={k} B={B}", flush=True)
    rows, psi = ksg.kraskov_prepare(ksg.center(synth.generate(spec)))
    stores = nframes > 2048
    budget = 2 * B * _per_atom_bytes(nframes, stores) + 4096
    lay = ksg.plan_layout(natoms, nframes, k, budget)
    assert lay == {"slots": 2 * B, "wave_atoms": B, "stores": int(stores)}
    with ksg.KsgPlan(natoms, nframes, k) as res:
        res.load(rows, psi)
        mi_res = res.run_all()
        assert res.stats()["slots"] == natoms
    nb = (natoms + B - 1) // B
    with ksg.KsgPlan(natoms, nframes, k, max_row_bytes=budget) as wv:
        wv.load(rows, psi)
        mi_w = wv.run_all()
        st = wv.stats()
        # every pair of blocks, and every block of two or more atoms with itself
        sizes = [min(natoms, (b + 1) * B) - b * B for b in range(nb)]
        assert st["wave_atoms"] == B
        assert st["waves"] == nb * (nb - 1) // 2 + sum(s >= 2 for s in sizes)
        mat = wv.matrix()
        a1 = np.array([natoms - 1, B, 1, natoms - 1]); a2 = np.array([0, B - 1, 0, natoms - 2])
        mi, d = wv.run_pairs(a1, a2, detail=True)
    assert np.max(np.abs(mi_w - mi_res)) <= 1e-12
    il = np.tril_indices(natoms, -1)
    order = np.lexsort((il[1], il[0]))
    assert np.array_equal(mat[il[0][order], il[1][order]], mi_w)
    _check_counts(mi, d, _brute_all(oracle_lib, rows, psi, a1, a2, k))


def test_coordinate_path_when_no_rows_fit(gpu, oracle_lib):
    """A budget below two atoms' rows: the plan takes the coordinate path."""
    spec = synth.TrajSpec(natoms=5, nframes=700, nmodes=3, seed=71)
    rows, psi = ksg.kraskov_prepare(ksg.center(synth.generate(spec)))
    with ksg.KsgPlan(5, 700, 15, max_row_bytes=1 << 20) as plan:
        plan.load(rows, psi)
        mi = plan.run_all()
        st = plan.stats()
    assert st["slots"] == 0 and st["kind"] == 4
    with ksg.KsgPlan(5, 700, 15) as res:
        res.load(rows, psi)
        assert np.max(np.abs(res.run_all() - mi)) <= 1e-12


def test_device_pointer_entry_equals_plan(gpu):
    """gmxcorr_ksg_pairs_device on device buffers = the plan's kind 4, bit
    for bit (same kernels, same row ranges)."""
    from tests.hipmem import DevArray
    rows, psi = ksg.kraskov_prepare(ksg.center(synth.generate(synth.SMALL_CPU)))
    a1, a2 = np.tril_indices(64, -1)
    with ksg.KsgPlan(64, 500, 10) as plan:
        plan.load(rows, psi)
        plan.set_variant(4)
        want = plan.run_pairs(a1, a2)
    dx, dpsi = DevArray(rows), DevArray(psi)
    d1, d2 = DevArray(a1.astype(np.int32)), DevArray(a2.astype(np.int32))
    dmi = DevArray(shape=(len(a1),), dtype=np.float64)
    ksg.ksg_pairs_device(dx.ptr, 64, 500, dpsi.ptr, 10, len(a1), d1.ptr, d2.ptr, dmi.ptr)
    assert np.array_equal(dmi.download(), want)


@pytest.mark.parametrize("nframes,k", [(1500, 50), (3000, 20)])
def test_xtc_file_frames_bitexact_vs_fp64_recount(gpu, oracle_lib, tmp_path, nframes, k):
    """Frames written to an XTC file at 0.001 nm and read back (real MD
    trajectories arrive quantised like this): many frames share a distance
    exactly, and the counts still equal the FP64 brute-force recount row by
    row, with the exact ties flagged."""
    spec = synth.TrajSpec(natoms=10, nframes=nframes, nmodes=4, seed=nframes + 1, n_switch=1)
    path = str(tmp_path / "q.xtc")
    ksg.write_xtc(path, synth.generate(spec), precision=1000.0)
    x, *_ = ksg.read_xtc(path)
    print(f"domain: XTC frames {spec} k={k}", flush=True)
    rows, psi = ksg.kraskov_prepare(ksg.center(x))
    a1 = np.array([1, 5, 9, 9]); a2 = np.array([0, 2, 8, 0])
    with ksg.KsgPlan(10, nframes, k) as plan:
        plan.load(rows, psi)
        mi, d = plan.run_pairs(a1, a2, detail=True)
        st = plan.stats()
    print(f"domain: XTC frames: {st['fixup_rows']} rows recounted, {st['ties']} exact ties",
          flush=True)
    _check_counts(mi, d, _brute_all(oracle_lib, rows, psi, a1, a2, k))


@pytest.mark.parametrize("nframes,k,kind,quantum", [(2000, 10, 2, 0.001), (3000, 20, 3, 0.001),
                                                    (3000, 20, 3, 0.0)])
def test_repeated_runs_are_bit_identical(gpu, nframes, k, kind, quantum):
    """Sums that blocks and recount waves add in any order (the long kind's
    row chunks, FP64-recounted rows) are kept in 64-bit fixed point, so
    every run of a plan gives the same bits; coordinates on the 0.001 nm
    grid make many recounted rows.  The same holds for the batched launch,
    whose persistent blocks take work in a different order every run."""
    c = synth.generate(synth.TrajSpec(natoms=24, nframes=nframes, nmodes=4, seed=nframes + k))
    if quantum:
        c = np.round(c / quantum).astype(np.float32) * np.float32(quantum)
    rows, psi = ksg.kraskov_prepare(ksg.center(c))
    with ksg.KsgPlan(24, nframes, k) as plan:
        plan.load(rows, psi)
        plan.set_variant(kind, 2 if kind == 3 else 0)
        runs = [plan.run_all() for _ in range(4)]
        st = plan.stats()
    assert st["fixup_rows"] > 0 or kind == 3
    for r in runs[1:]:
        assert np.array_equal(r, runs[0])
    from gmx_correlation_amd import batch
    outs = []
    for _ in range(3):
        b = batch.KsgBatch()
        b.add(rows, psi, k)
        b.add(rows[:30], psi, k)
        b.run(max_row_bytes=20 * 10**9, nblocks=512)
        outs.append((b.result(0)[0].copy(), b.result(1)[0].copy()))
        b.close()
    for o in outs[1:]:
        assert np.array_equal(o[0], outs[0][0]) and np.array_equal(o[1], outs[0][1])
