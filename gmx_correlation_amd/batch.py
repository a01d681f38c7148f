"""Batched KSG matrices of many independent trajectories (uneven atom and
frame counts, different k) with one persistent pair-kernel launch per memory
wave, plus the autotuner that picks each trajectory's kernel path.

No counterpart in the reference, which computes one trajectory per run
(lib/correlation_core.cpp:324-459); every trajectory's MI follows the same
KSG-2 definition (lib/kraskov.cpp:1228-1331).
"""
from __future__ import annotations

import ctypes
import json
import os
import time

import numpy as np

from . import _lib
from .ksg import KsgPlan

# Code paths compiled into the persistent batch kernel: (kind, s4) with
# kind 0 = one wave per frame row streaming both distance rows (N <= 256*s4),
# kind 1 = four waves per row (N <= 1024*s4), kind 2 = one wave per row over
# the 512 nearest frames of the sparser atom (N <= 2048, k + 1 <= 256), kind 3
# = the same over the first 256*s4 frames of 1024-frame prefix stores
# (2048 < N <= 8192, k + 1 <= 256; s4 = 2, 3 or 4).
BATCH_PATHS = [(2, 1), (3, 2), (3, 3), (3, 4), (0, 2), (0, 4), (0, 6), (0, 8), (1, 3), (1, 4),
               (1, 6), (1, 8)]
PREFIX_MAX_FRAMES = 2048
LONG_MAX_FRAMES = 8192
PREFIX_MAX_T = 256
# Largest share of a trajectory's rows that may fall beyond the prefix of
# kind 3 before a longer prefix is faster: a long row costs about 4x a
# prefix row and the 1024-frame prefix about 1.4x the 512-frame one
# (break-even near 6%).
LONG_SHARE_MAX = 0.04


def fitting_paths(nframes: int, k: int) -> list[tuple[int, int]]:
    def fits(kind: int, s4: int) -> bool:
        if kind == 2:
            return nframes <= PREFIX_MAX_FRAMES and k + 1 <= PREFIX_MAX_T
        if kind == 3:
            return PREFIX_MAX_FRAMES < nframes <= LONG_MAX_FRAMES and k + 1 <= PREFIX_MAX_T
        return nframes <= (256 if kind == 0 else 1024) * s4
    return [(kd, s) for kd, s in BATCH_PATHS if fits(kd, s)]


def tune_variant(rows: np.ndarray, psi: np.ndarray, k: int, natoms_probe: int = 10,
                 reps: int = 2, paths=None) -> tuple[tuple[int, int], dict]:
    """Time every batch code path that covers the trajectory's frame count on
    the pairs of `natoms_probe` atoms spread evenly over the trajectory
    (standalone launches, HIP events; pair kernel plus long-row pass and FP64
    recount) and return the fastest (kind, s4) with all timings in ms.
    Spread atoms stand for the all-pairs mix: most pairs of a chain are far
    apart and weakly coupled, and their (k+1)-th neighbour lies deepest in
    the sorted rows."""
    N = rows.shape[1]
    natoms = rows.shape[0] // 3
    m = min(natoms_probe, natoms)
    probe = np.unique(np.round(np.linspace(0, natoms - 1, m)).astype(np.int64))
    m = len(probe)
    sub = rows.reshape(natoms, 3, N)[probe].reshape(3 * m, N)
    cands = paths if paths is not None else fitting_paths(N, k)
    # keep only the two smallest of each streaming kind (larger register
    # tiles only add work); every prefix length of kind 3
    by_kind: dict[int, list[int]] = {}
    for kind, s4 in sorted(cands):
        by_kind.setdefault(kind, []).append(s4)
    cands = [(kd, s) for kd, ss in by_kind.items() for s in (ss if kd == 3 else ss[:2])]
    a1, a2 = np.tril_indices(m, -1)
    timings = {}
    long_share: dict[tuple[int, int], float] = {}
    with KsgPlan(m, N, k) as plan:
        plan.load(np.ascontiguousarray(sub), psi)
        for kind, s4 in cands:
            plan.set_variant(kind, s4)
            plan.run_pairs(a1, a2)                      # warm-up
            t = []
            for _ in range(reps):
                plan.run_pairs(a1, a2)
                st = plan.stats()
                t.append(st["ms_pairs"] + st["ms_fixup"])
            long_share[(kind, s4)] = st["long_rows"] / float(len(a1) * N)
            timings[f"{kind}:{s4}"] = float(np.min(t))
    best = min(timings, key=timings.get)
    kind, s4 = (int(v) for v in best.split(":"))
    if kind == 3:
        # a probe of a few pairs leaves the GPU idle and hides the cost of
        # the long rows, which a full launch pays at about 4x a prefix row:
        # the shortest prefix, no shorter than the fastest, whose share of
        # rows beyond it stays within LONG_SHARE_MAX (else the whole store)
        for c in range(s4, 5):
            if c == 4 or long_share.get((3, c), 1.0) <= LONG_SHARE_MAX:
                s4 = c
                break
    return (kind, s4), timings


# Shape candidates: row ranges per pair and the all-pairs block shape (2x2
# atom tile, or four consecutive pairs) for the one-wave kinds, frame rows
# per block for kind 3.  The kinds' defaults are 128 ranges, 16 rows, tiles.
SHAPE_SPLITS = (32, 64, 128, 256)
SHAPE_CHUNKS = (4, 8, 16, 32, 64)
SHAPE_TILES = (1, 2)


def shape_candidates(kind: int) -> list[dict]:
    if kind == 3:
        return [{"nsplit": 0, "chunk": c, "tile": t} for c in SHAPE_CHUNKS for t in SHAPE_TILES]
    if kind in (0, 2, 4):
        return [{"nsplit": n, "chunk": 0, "tile": t} for n in SHAPE_SPLITS for t in SHAPE_TILES]
    return [{"nsplit": 0, "chunk": 0, "tile": 0}]


def shape_key(shape: dict) -> str:
    return f"{shape['nsplit']}:{shape['chunk']}:{shape['tile']}"


def time_shapes(rows: np.ndarray, psi: np.ndarray, k: int, kind: int, s4: int,
                shapes: list[dict], reps: int = 3, fraction: float | None = None) -> dict:
    """ms of the trajectory's all-pairs run for each shape, the best of
    `reps` after one warm-up: the whole run (pair kernel, long rows,
    recount, finalise) or, with `fraction`, the pair kernel over that
    fraction of its work items (KsgPlan.probe)."""
    natoms, N = rows.shape[0] // 3, rows.shape[1]
    out = {}
    with KsgPlan(natoms, N, k) as plan:
        plan.load(np.ascontiguousarray(rows), psi)
        plan.set_variant(kind, s4)
        for shp in shapes:
            plan.set_shape(shp["nsplit"], shp["chunk"], shp["tile"])
            t = []
            for r in range(reps + 1):
                if fraction is not None:
                    ms = plan.probe(fraction)
                else:
                    plan.run_all_device()
                    st = plan.stats()
                    ms = st["ms_pairs"] + st["ms_fixup"]
                if r:          # the first run is the warm-up
                    t.append(ms)
            out[shape_key(shp)] = float(np.min(t))
    return out


def tune_shape(rows: np.ndarray, psi: np.ndarray, k: int, kind: int, s4: int,
               fraction: float = 0.125, reps: int = 3) -> tuple[dict, dict]:
    """Row-block and tile shape of path (kind, s4) for this trajectory: every
    candidate of shape_candidates(kind) timed by the pair kernel's probe over
    the first `fraction` of the all-pairs work items.  The shapes act
    through which distance rows the blocks in flight on one XCD share in its
    L2; the probe keeps every atom and the launch's block order (tiles,
    row-range-major or chunk-major), so the blocks in flight share rows as
    in the full run, over fewer frame rows.  (Timing a subset of the atoms
    does not: with fewer tiles per row range the order, and the best shape,
    change.)  Returns the fastest shape and all probe timings (ms)."""
    shapes = shape_candidates(kind)
    timings = time_shapes(rows, psi, k, kind, s4, shapes, reps, fraction)
    best = min(shapes, key=lambda sh: timings[shape_key(sh)])
    return best, timings


class Tuner:
    """Per-(frame count, k) cache of tuned paths, optionally kept in a JSON
    file between runs."""

    def __init__(self, path: str | None = None):
        self.path = path
        self.cache: dict[str, dict] = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.cache = json.load(f)

    def choose(self, rows, psi, k) -> tuple[int, int]:
        key = f"{rows.shape[1]}:{k}"
        if key not in self.cache:
            (kind, s4), timings = tune_variant(rows, psi, k)
            self.cache[key] = {"kind": kind, "s4": s4, "ms": timings}
            self._save()
        c = self.cache[key]
        return c["kind"], c["s4"]

    def choose_chunk(self, rows, psi, k, kind, s4, fraction: float = 1 / 32, reps: int = 1) -> int:
        """Frame rows per kind-3 work item of the batch (0 for other kinds):
        each candidate timed by the probe over `fraction` of the all-pairs
        work items, best of `reps` after a warm-up (a whole trajectory's
        plan is built for it, so the probe is kept short)."""
        if kind != 3:
            return 0
        key = f"{rows.shape[1]}:{k}:{kind}:{s4}:chunk"
        if key not in self.cache:
            shapes = [{"nsplit": 0, "chunk": c, "tile": 1} for c in SHAPE_CHUNKS]
            t = time_shapes(rows, psi, k, kind, s4, shapes, reps=reps, fraction=fraction)
            best = min(shapes, key=lambda sh: t[shape_key(sh)])
            self.cache[key] = {"chunk": best["chunk"], "ms": t}
            self._save()
        return self.cache[key]["chunk"]

    def _save(self):
        if self.path:
            os.makedirs(os.path.dirname(os.path.abspath(self.path)), exist_ok=True)
            with open(self.path, "w") as f:
                json.dump(self.cache, f, indent=1)


class KsgBatch:
    def __init__(self):
        L = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(L.gmxcorr_ksg_batch_create(ctypes.byref(h)))
        self._h = h
        self.shapes: list[tuple[int, int, int]] = []

    def add(self, rows: np.ndarray, psi: np.ndarray, k: int, kind: int = -1, s4: int = 0) -> int:
        rows = np.ascontiguousarray(rows, dtype=np.float64)
        psi = np.ascontiguousarray(psi, dtype=np.float64)
        natoms, N = rows.shape[0] // 3, rows.shape[1]
        rc = _lib.load().gmxcorr_ksg_batch_add(self._h, _lib.ptr(rows), _lib.ptr(psi), natoms, N,
                                               k, kind, s4)
        if rc < 0:
            _lib.check(rc)
        self.shapes.append((natoms, N, k))
        return rc

    def set_chunk(self, t: int, chunk: int) -> None:
        """Frame rows per kind-3 work item of trajectory t (0 = default 4)."""
        _lib.check(_lib.load().gmxcorr_ksg_batch_set_chunk(self._h, t, chunk))

    def traj_bytes(self, t: int) -> int:
        """Device bytes of trajectory t's rows in a memory wave."""
        v = _lib.load().gmxcorr_ksg_batch_traj_bytes(self._h, t)
        if v < 0:
            _lib.check(int(v))
        return int(v)

    def run(self, max_row_bytes: int = 200 * 10**9, nblocks: int = 2048) -> dict:
        t0 = time.perf_counter()
        _lib.check(_lib.load().gmxcorr_ksg_batch_run(self._h, int(max_row_bytes), nblocks))
        st = _lib.BatchStats()
        _lib.check(_lib.load().gmxcorr_ksg_batch_stats(self._h, ctypes.byref(st)))
        out = st.as_dict()
        out["wall_s"] = time.perf_counter() - t0
        return out

    def result(self, t: int) -> tuple[np.ndarray, float, int]:
        natoms = self.shapes[t][0]
        n = natoms * (natoms - 1) // 2
        mi = np.empty(n)
        ms = ctypes.c_double(0.0)
        wave = ctypes.c_int(0)
        _lib.check(_lib.load().gmxcorr_ksg_batch_result(self._h, t, _lib.ptr(mi), n,
                                                        ctypes.byref(ms), ctypes.byref(wave)))
        return mi, ms.value, wave.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.load().gmxcorr_ksg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
