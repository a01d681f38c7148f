"""The shape autotuner against a full sweep.

batch.tune_shape times every candidate shape (row ranges per pair x block
tile shape for the sorted-prefix kind, frame rows per block x tile shape for
the long-trajectory kind) by the pair kernel over the first eighth of its
all-pairs work items (KsgPlan.probe: every atom, the launch's block order,
fewer frame rows).  Its pick,
timed on the whole trajectory, must be within 8% of the fastest shape of a
full sweep on the whole trajectory (run-to-run spread of a launch is about
1%, and shapes near the optimum differ by a few percent).  Cases: the
247-atom x 2000-frame workload (k = 100), and two trajectories of the
batched workload, one per kind (146 x 1933, k = 10; 206 x 3858, k = 50)."""
import numpy as np
import pytest

from gmx_correlation_amd import _lib, batch, ksg, synth

pytestmark = pytest.mark.gpu

MARGIN = 1.08


@pytest.fixture(scope="module")
def gpu(hiplib):
    if not _lib.probe():
        pytest.fail("no gfx950 device visible to the HIP library")
    return hiplib


def _case(name):
    if name == "readme247":
        return synth.README_247, 100
    t = {"batched15": 15, "batched10": 10}[name]
    return synth.batched_specs()[t]


@pytest.mark.parametrize("name", ["readme247", "batched15", "batched10"])
def test_shape_pick_within_margin_of_full_sweep(gpu, name):
    spec, k = _case(name)
    rows, psi = ksg.kraskov_prepare(ksg.center(synth.generate(spec)))
    kind = 2 if spec.nframes <= 2048 else 3
    depth2 = (k + 1) * spec.nframes   # the plan's default prefix length of kind 3
    s4 = 1 if kind == 2 else 2 if depth2 <= 2**18 else 3 if depth2 <= 2**19 else 4
    print(f"tuner: {name} {spec.natoms} x {spec.nframes} k={k} kind {kind} s4 {s4}", flush=True)
    pick, probe_ms = batch.tune_shape(rows, psi, k, kind, s4)
    full = batch.time_shapes(rows, psi, k, kind, s4, batch.shape_candidates(kind))
    best = min(full.values())
    got = full[batch.shape_key(pick)]
    print(f"tuner: {name}: pick {batch.shape_key(pick)} {got:.2f} ms, best of the full sweep "
          f"{min(full, key=full.get)} {best:.2f} ms; full {full}", flush=True)
    assert got <= MARGIN * best
