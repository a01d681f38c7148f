"""XTC trajectory reader and writer (gmxcorr_xtc_*, host code, no GPU).

No GROMACS-written XTC file is available offline, so the reader is pinned by
the format's own rules rather than by foreign files: frames written with the
format's encoder (xdr3dfcoord rules: integer coordinates round(x * precision)
offset by their minimum, mixed-radix large atoms, runs of small differences
with the first two atoms swapped, the smallidx schedule) decode to exactly
round(x * precision) * (1 / precision); the XDR header and the compressed
block's fields hold what the format defines; frames of nine atoms or fewer
are stored as plain floats.  Parity with GROMACS-written files is unpinned."""
import os
import struct

import numpy as np
import pytest

from gmx_correlation_amd import _lib, ksg, synth


def _quantised(c, precision):
    """What the format stores: round half away from zero of x * precision
    (in float), decoded as int * (1 / precision) in float."""
    lf = c.astype(np.float32) * np.float32(precision)
    q = np.where(lf >= 0, np.floor(lf + np.float32(0.5)), np.ceil(lf - np.float32(0.5)))
    return q.astype(np.float32) * np.float32(np.float32(1.0) / np.float32(precision)), q


def _water(nmol, frames, seed):
    """Molecules of three atoms 0.1 nm apart (water-like): consecutive atoms
    are close, so the writer uses runs of small steps."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0, 5, (frames, nmol, 3))
    off = np.array([[0, 0, 0], [0.0957, 0, 0], [-0.024, 0.0927, 0]])
    return (centres[:, :, None, :] + off[None, None]).reshape(frames, 3 * nmol, 3).astype(np.float32)


@pytest.mark.parametrize("natoms,precision", [(10, 1000.0), (64, 1000.0), (1000, 1000.0),
                                              (300, 100.0), (300, 10000.0)])
def test_round_trip_is_the_quantised_trajectory(hiplib, tmp_path, natoms, precision):
    c = synth.generate(synth.TrajSpec(natoms=natoms, nframes=20, nmodes=4, seed=natoms))
    c = c - 3.0    # negative coordinates too
    path = str(tmp_path / "t.xtc")
    steps, times = np.arange(20) * 500, np.arange(20, dtype=np.float32) * 2.5
    boxes = np.tile(np.diag([7.0, 8.0, 9.0]).astype(np.float32), (20, 1, 1))
    ksg.write_xtc(path, c, precision, steps, times, boxes)
    x, st, tm, bx, prec = ksg.read_xtc(path)
    want, _ = _quantised(c, precision)
    assert x.shape == c.shape and prec == precision
    assert np.array_equal(x, want)
    assert np.array_equal(st, steps) and np.array_equal(tm, times) and np.array_equal(bx, boxes)


def test_water_runs_and_swaps(hiplib, tmp_path):
    """Runs of small steps (the swap of the first two atoms, smallidx moving
    up and down) decode exactly, and compress well below random data."""
    w = _water(400, 10, 1)
    rnd = np.random.default_rng(2).uniform(0, 5, w.shape).astype(np.float32)
    pw, pr = str(tmp_path / "w.xtc"), str(tmp_path / "r.xtc")
    ksg.write_xtc(pw, w)
    ksg.write_xtc(pr, rnd)
    x, *_ = ksg.read_xtc(pw)
    assert np.array_equal(x, _quantised(w, 1000.0)[0])
    assert np.array_equal(ksg.read_xtc(pr)[0], _quantised(rnd, 1000.0)[0])
    assert os.path.getsize(pw) < 0.8 * os.path.getsize(pr)


def test_large_coordinate_range_uses_per_coordinate_bits(hiplib, tmp_path):
    """A range above 2^24 integer units (here 30,000 nm at 0.001 nm) is
    written coordinate by coordinate, not in mixed radix."""
    rng = np.random.default_rng(3)
    c = rng.uniform(-15000, 15000, (3, 50, 3)).astype(np.float32)
    p = str(tmp_path / "l.xtc")
    ksg.write_xtc(p, c)
    assert np.array_equal(ksg.read_xtc(p)[0], _quantised(c, 1000.0)[0])


def test_small_frames_are_plain_floats(hiplib, tmp_path):
    c = np.random.default_rng(4).standard_normal((5, 9, 3)).astype(np.float32)
    p = str(tmp_path / "s.xtc")
    ksg.write_xtc(p, c)
    x, _, _, _, prec = ksg.read_xtc(p)
    assert np.array_equal(x, c) and prec == 0.0
    raw = open(p, "rb").read()
    frame = 4 * (4 + 9 + 1) + 9 * 3 * 4          # header, box, natoms, floats
    assert len(raw) == 5 * frame
    assert struct.unpack(">27f", raw[14 * 4:14 * 4 + 108]) == tuple(c[0].ravel().tolist())


def test_header_and_block_fields(hiplib, tmp_path):
    """XDR words as the format defines them: magic 1995, natoms, step, time,
    box; then natoms, precision, minint, maxint (the integer coordinates'
    extremes), smallidx in [9, 72], the byte count and its padding."""
    c = synth.generate(synth.TrajSpec(natoms=30, nframes=1, nmodes=2, seed=5))
    p = str(tmp_path / "h.xtc")
    ksg.write_xtc(p, c, 1000.0, [7], [1.5], [np.eye(3, dtype=np.float32) * 4])
    raw = open(p, "rb").read()
    magic, n, step = struct.unpack(">3i", raw[:12])
    time, = struct.unpack(">f", raw[12:16])
    box = struct.unpack(">9f", raw[16:52])
    n2, = struct.unpack(">i", raw[52:56])
    prec, = struct.unpack(">f", raw[56:60])
    mins, maxs = struct.unpack(">3i", raw[60:72]), struct.unpack(">3i", raw[72:84])
    smallidx, nbytes = struct.unpack(">2i", raw[84:92])
    _, q = _quantised(c[0], 1000.0)
    assert (magic, n, step, time, n2, prec) == (1995, 30, 7, 1.5, 30, 1000.0)
    assert box == (4.0, 0, 0, 0, 4.0, 0, 0, 0, 4.0)
    assert mins == tuple(q.min(axis=0).astype(int)) and maxs == tuple(q.max(axis=0).astype(int))
    assert 9 <= smallidx <= 72
    assert len(raw) == 92 + (nbytes + 3) // 4 * 4


def test_atom_selection_and_errors(hiplib, tmp_path):
    c = synth.generate(synth.TrajSpec(natoms=40, nframes=6, nmodes=2, seed=6))
    p = str(tmp_path / "a.xtc")
    ksg.write_xtc(p, c)
    sel = [3, 17, 39]
    x, *_ = ksg.read_xtc(p, atoms=sel)
    assert np.array_equal(x, _quantised(c, 1000.0)[0][:, sel])
    assert ksg.read_xtc(p, max_frames=2)[0].shape == (2, 40, 3)
    bad = str(tmp_path / "bad.xtc")
    open(bad, "wb").write(b"\x00\x00\x07\xcb" + open(p, "rb").read()[4:60])   # cut short
    with pytest.raises(_lib.GmxcorrError):
        ksg.read_xtc(bad)
    open(bad, "wb").write(b"\x00\x00\x00\x01" + b"\x00" * 60)
    with pytest.raises(_lib.GmxcorrError):
        ksg.read_xtc(bad)


def test_xtc_frames_feed_the_reference_pipeline(hiplib, tmp_path, oracle_lib):
    """Frames read back from XTC go through center / kraskov_prepare like any
    coordinates: the oracle's KSG of the decoded frames equals its KSG of the
    quantised coordinates."""
    c = synth.generate(synth.TrajSpec(natoms=12, nframes=300, nmodes=3, seed=7))
    p = str(tmp_path / "k.xtc")
    ksg.write_xtc(p, c)
    x, *_ = ksg.read_xtc(p)
    rows, psi = ksg.kraskov_prepare(ksg.center(x))
    rq, pq = ksg.kraskov_prepare(ksg.center(_quantised(c, 1000.0)[0]))
    assert np.array_equal(rows, rq)
    assert oracle_lib.ksg_pair(rows, psi, 3, 1, 5) == oracle_lib.ksg_pair(rq, pq, 3, 1, 5)
