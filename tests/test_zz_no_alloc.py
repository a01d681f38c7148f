"""gmxcorr_ksg_pairs_device_noalloc (static device scratch, no allocation)
against gmxcorr_ksg_pairs_device (allocated scratch) through the C ABI: two
calls in a row, the second on a smaller input, each bit-identical to the
allocating entry point on the same input (same kernels and row ranges; the
coordinate path's sums have a fixed order); a pair list above the scratch
maximum is refused."""
import ctypes

import numpy as np
import pytest

from gmx_correlation_amd import _lib, ksg, synth
from tests.hipmem import DevArray, hip

pytestmark = pytest.mark.gpu

MAX_PAIRS = ((32 << 20) - 256) // 8


@pytest.fixture(scope="module")
def gpu(hiplib):
    if not _lib.probe():
        pytest.fail("no gfx950 device visible to the HIP library")
    return hiplib


def _inputs(spec, k, pairs=None):
    rows, psi = ksg.kraskov_prepare(ksg.center(synth.generate(spec)))
    n = spec.natoms
    a1, a2 = np.tril_indices(n, -1) if pairs is None else pairs
    return dict(rows=rows, psi=psi, k=k, n=n, N=spec.nframes,
                a1=np.asarray(a1, np.int32), a2=np.asarray(a2, np.int32))


def _call(L, fn, inp):
    dx, dpsi = DevArray(inp["rows"]), DevArray(inp["psi"])
    d1, d2 = DevArray(inp["a1"]), DevArray(inp["a2"])
    dmi = DevArray(shape=(len(inp["a1"]),), dtype=np.float64)
    rc = getattr(L, fn)(dx.ptr, inp["n"], inp["N"], dpsi.ptr, inp["k"], len(inp["a1"]), d1.ptr,
                        d2.ptr, dmi.ptr, None)
    _lib.check(rc)
    assert hip().hipDeviceSynchronize() == 0
    return dmi.download()


def test_noalloc_equals_allocating_entry_twice(gpu):
    first = _inputs(synth.SMALL_CPU, 10)                                  # 2016 pairs x 500
    second = _inputs(synth.TrajSpec(natoms=9, nframes=300, nmodes=3, seed=41, n_switch=1), 7)
    print("no-alloc: call 1 64 atoms x 500 frames k=10; call 2 9 atoms x 300 frames k=7",
          flush=True)
    got1 = _call(gpu, "gmxcorr_ksg_pairs_device_noalloc", first)
    got2 = _call(gpu, "gmxcorr_ksg_pairs_device_noalloc", second)
    want1 = _call(gpu, "gmxcorr_ksg_pairs_device", first)
    want2 = _call(gpu, "gmxcorr_ksg_pairs_device", second)
    assert np.array_equal(got1, want1)
    assert np.array_equal(got2, want2)
    # and against the plan's coordinate path, i.e. the tested KSG semantics
    with ksg.KsgPlan(9, 300, 7) as plan:
        plan.load(second["rows"], second["psi"])
        plan.set_variant(4)
        assert np.array_equal(plan.run_pairs(second["a1"], second["a2"]), got2)


def test_noalloc_refuses_lists_above_its_scratch(gpu):
    inp = _inputs(synth.TrajSpec(natoms=3, nframes=50, nmodes=2, seed=2), 3)
    dx, dpsi = DevArray(inp["rows"]), DevArray(inp["psi"])
    d1, d2 = DevArray(inp["a1"]), DevArray(inp["a2"])
    dmi = DevArray(shape=(3,), dtype=np.float64)
    rc = gpu.gmxcorr_ksg_pairs_device_noalloc(dx.ptr, 3, 50, dpsi.ptr, 3, MAX_PAIRS + 1, d1.ptr,
                                              d2.ptr, dmi.ptr, None)
    assert rc == -1
    assert b"scratch maximum" in gpu.gmxcorr_last_error()
