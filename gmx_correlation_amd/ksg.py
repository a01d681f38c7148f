"""Host-side mirror of the reference's KSG interface, backed by the HIP library.

Names, argument meaning and error behaviour follow the reference:
  center()             src/gmx_correlation.cpp:184-223 (t_traj fluctuations)
  kraskov_prepare()    lib/kraskov.cpp:2168-2220
  kraskov_corrmatrix() lib/correlation_core.cpp:324-459 (GPU branch :337-348)
  pearsify()           lib/correlation_core.cpp:461-469
  write_matrix()       lib/correlation_core.cpp:202-225
  read_matrix()        bin/analyze_matrix.py:71-102 (the format's consumer)
  read_xtc()           the frames GROMACS hands to the analysis module
                       (src/gmx_correlation.cpp:158-182), from an XTC file

One deliberate difference: the reference's dispatcher silently falls back to
its CPU loop when no GPU is present or k > 128.  This package has no CPU
path; it raises instead (the CPU path stays in the reference's own code).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

EULER_GAMMA_TRUNC = 0.57721566490153   # the reference's psi[1] = -0.57721566490153
GPU_K_MAX = 128                         # lib/correlation_core.cpp:339


def center(coords: np.ndarray) -> np.ndarray:
    """float32 frames [nframes][natoms][3] -> mean-centred double
    fluctuations [natoms][nframes][3] (t_traj.x).  The mean is a sequential
    double sum over frames divided by nframes."""
    c = np.asarray(coords, dtype=np.float32).astype(np.float64)
    T = c.shape[0]
    mean = np.cumsum(c, axis=0)[-1] / T            # sequential sum, as the reference
    return np.ascontiguousarray((c - mean[None]).transpose(1, 0, 2))


def kraskov_prepare(traj: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """t_traj.x [natoms][N][3] -> (x rows [3*natoms][N], psi [N+1]).

    Per component: mean and sample SD (N-1) from sequential sums, z-score,
    subtract the minimum.  psi[0]=0, psi[1]=-0.57721566490153,
    psi[m+1]=psi[m]+1/m.  A zero SD gives non-finite rows, which the GPU
    library rejects."""
    t = np.asarray(traj, dtype=np.float64)
    n, N, _ = t.shape
    x = np.ascontiguousarray(t.transpose(0, 2, 1).reshape(3 * n, N))
    me = np.cumsum(x, axis=1)[:, -1] / N
    dev = x - me[:, None]
    s = np.sqrt(np.cumsum(dev * dev, axis=1)[:, -1] / (N - 1))
    with np.errstate(divide="ignore", invalid="ignore"):
        z = dev / s[:, None]
    lo = np.minimum(z.min(axis=1), 1000000.0)
    rows = np.ascontiguousarray(z - lo[:, None])
    inc = np.empty(N)
    inc[0] = -EULER_GAMMA_TRUNC
    inc[1:] = 1.0 / np.arange(1, N, dtype=np.float64)
    psi = np.empty(N + 1)
    psi[0] = 0.0
    psi[1:] = np.cumsum(inc)
    return rows, psi


def plan_layout(natoms: int, nframes: int, k: int, max_row_bytes: int) -> dict:
    """How a plan of this shape holds its distance rows for a row budget
    (host arithmetic, no device): slots = natoms (all rows resident),
    2 * wave_atoms (memory waves of two atom blocks) or 0 (coordinate path
    only); stores = 1 when the long-trajectory prefix stores fit too."""
    slots, wave, stores = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _lib.check(_lib.load().gmxcorr_ksg_plan_layout(natoms, nframes, k, int(max_row_bytes),
                                                   ctypes.byref(slots), ctypes.byref(wave),
                                                   ctypes.byref(stores)))
    return {"slots": slots.value, "wave_atoms": wave.value, "stores": stores.value}


class KsgPlan:
    """Device-resident KSG run: distance rows stay in HBM between runs.

    max_row_bytes: budget for distance rows (0: free device memory less the
    pair buffers).  When every atom's rows do not fit, pairs run in memory
    waves of two atom blocks; above 8192 frames (or when not even two atoms'
    rows fit) every row is computed from the coordinates (kind 4)."""

    def __init__(self, natoms: int, nframes: int, k: int, max_row_bytes: int = 0):
        L = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(L.gmxcorr_ksg_plan_create_budget(natoms, nframes, k, int(max_row_bytes),
                                                    ctypes.byref(h)))
        self._h = h
        self.natoms, self.nframes, self.k = natoms, nframes, k

    def load(self, rows: np.ndarray, psi: np.ndarray) -> "KsgPlan":
        rows = np.ascontiguousarray(rows, dtype=np.float64)
        psi = np.ascontiguousarray(psi, dtype=np.float64)
        if rows.shape != (3 * self.natoms, self.nframes) or psi.shape != (self.nframes + 1,):
            raise ValueError("rows must be [3*natoms][nframes] and psi [nframes+1]")
        _lib.check(_lib.load().gmxcorr_ksg_plan_load(self._h, _lib.ptr(rows), _lib.ptr(psi)))
        return self

    def run_all(self) -> np.ndarray:
        """MI of all pairs in the reference's order (a1 = 1.., a2 < a1)."""
        L = _lib.load()
        _lib.check(L.gmxcorr_ksg_plan_run_all(self._h))
        n = self.natoms * (self.natoms - 1) // 2
        out = np.empty(n)
        _lib.check(L.gmxcorr_ksg_plan_result(self._h, _lib.ptr(out), n))
        return out

    def run_all_device(self) -> None:
        """Run without copying results back (timing)."""
        _lib.check(_lib.load().gmxcorr_ksg_plan_run_all(self._h))

    def matrix(self) -> np.ndarray:
        mat = np.zeros((self.natoms, self.natoms))
        _lib.check(_lib.load().gmxcorr_ksg_plan_matrix(self._h, _lib.ptr(mat)))
        return mat

    def run_pairs(self, a1, a2, detail: bool = False):
        a1 = np.ascontiguousarray(a1, dtype=np.int32)
        a2 = np.ascontiguousarray(a2, dtype=np.int32)
        n = len(a1)
        L = _lib.load()
        nx = ny = fl = None
        if detail:
            nx = np.empty((n, self.nframes), np.int32)
            ny = np.empty((n, self.nframes), np.int32)
            fl = np.empty((n, self.nframes), np.uint8)
        _lib.check(L.gmxcorr_ksg_plan_run_pairs(
            self._h, n, _lib.ptr(a1, _lib._ip), _lib.ptr(a2, _lib._ip),
            _lib.ptr(nx, _lib._ip), _lib.ptr(ny, _lib._ip), _lib.ptr(fl, _lib._u8p)))
        out = np.empty(n)
        _lib.check(L.gmxcorr_ksg_plan_result(self._h, _lib.ptr(out), n))
        if detail:
            return out, dict(nx=nx, ny=ny, flags=fl)
        return out

    def set_variant(self, kind: int, s4: int = 0) -> None:
        """Pair-kernel variant: kind 0 one wave per row streaming both rows
        (256*s4 frames), kind 1 four waves per row (1024*s4 frames), kind 2
        one wave per row over the sorted prefix (N <= 2048, k + 1 <= 256),
        kind 3 the same over the first 256*s4 frames of the prefix stores
        (2048 < N <= 8192, s4 = 2, 3 or 4), kind 4 every row in FP64 from
        the coordinates (any N), kind -1 default."""
        _lib.check(_lib.load().gmxcorr_ksg_plan_set_variant(self._h, kind, s4))

    def set_shape(self, nsplit: int = 0, chunk: int = 0, tile: int = 0) -> None:
        """Row-block and tile shape: nsplit row ranges per pair (kinds 0, 2,
        4; 0 = default), chunk frame rows per block of kind 3 (0 = default
        4), tile 0/1 = 2x2 atom tiles in the all-pairs run, 2 = runs of four
        pairs sharing their first atom."""
        _lib.check(_lib.load().gmxcorr_ksg_plan_set_shape(self._h, nsplit, chunk, tile))

    def probe(self, fraction: float) -> float:
        """ms of the all-pairs pair kernel over the first `fraction` of its
        work items (the autotuner's timing probe; leaves no result)."""
        ms = ctypes.c_double(0.0)
        _lib.check(_lib.load().gmxcorr_ksg_plan_probe(self._h, float(fraction), ctypes.byref(ms)))
        return ms.value

    def stats(self) -> dict:
        st = _lib.KsgStats()
        _lib.check(_lib.load().gmxcorr_ksg_plan_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.load().gmxcorr_ksg_plan_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ksg_pairs_device(xrows_ptr: int, natoms: int, nframes: int, psi_ptr: int, k: int,
                     npairs: int, a1_ptr: int, a2_ptr: int, mi_ptr: int, stream: int = 0) -> None:
    """MI of a pair list on the coordinate path, all arrays already in device
    memory (raw device addresses): xrows [3*natoms][nframes] float64, psi
    [nframes+1] float64, a1/a2 [npairs] int32, mi [npairs] float64 (out).
    The caller guarantees 0 <= a1, a2 < natoms and a1 != a2."""
    _lib.check(_lib.load().gmxcorr_ksg_pairs_device(
        xrows_ptr, natoms, nframes, psi_ptr, k, npairs, a1_ptr, a2_ptr, mi_ptr, stream or None))


def kraskov_corrmatrix(traj: np.ndarray, k: int, use_gpu: bool = True) -> np.ndarray:
    """MI matrix [natoms][natoms] (nats, diagonal 2000) of a t_traj.

    Raises ValueError for natoms < 2 or nframes <= k (the reference throws
    std::runtime_error, lib/correlation_core.cpp:327-330) and RuntimeError
    when the GPU path cannot run (no device, k > 128)."""
    t = np.asarray(traj, dtype=np.float64)
    natoms, nframes = t.shape[0], t.shape[1]
    if natoms < 2:
        raise ValueError("kraskov_corrmatrix: need at least 2 atoms")
    if nframes <= k:
        raise ValueError("kraskov_corrmatrix: nframes must exceed k (number of nearest neighbours)")
    if not use_gpu:
        raise RuntimeError("this package implements only the GPU path (use_gpu=True)")
    if k > GPU_K_MAX:
        raise RuntimeError(f"GPU path supports k <= {GPU_K_MAX}; k={k} requested")
    rows, psi = kraskov_prepare(t)
    mat = np.zeros((natoms, natoms))
    _lib.check(_lib.load().gmxcorr_ksg_matrix(_lib.ptr(rows), natoms, nframes, _lib.ptr(psi), k,
                                              _lib.ptr(mat)))
    return mat


def pearsify(mat: np.ndarray, dim: int = 3) -> np.ndarray:
    """r = sqrt(1 - exp(-2/dim * MI)) for MI > 0 else 0 (diagonal 2000 -> 1)."""
    m = np.asarray(mat, dtype=np.float64)
    with np.errstate(over="ignore"):
        r = np.sqrt(1.0 - np.exp(-2.0 / dim * m))
    return np.where(m > 0.0, r, 0.0)


def write_matrix(mat: np.ndarray, path: str) -> None:
    """correl.dat: "n x n [", then mat[j*n+i] for i, j (transposed) as
    "%10.6g " with a newline after every 21 values, then "]"."""
    m = np.asarray(mat, dtype=np.float64)
    n = m.shape[0]
    parts = [f"{n} x {n} [\n"]
    ncol = 0
    for i in range(n):
        for j in range(n):
            parts.append("%10.6g " % m[j, i])
            ncol += 1
            if ncol > 20:
                ncol = 0
                parts.append("\n")
    parts.append("]\n")
    with open(path, "w") as f:
        f.write("".join(parts))


def read_matrix(path: str) -> np.ndarray:
    """Inverse of write_matrix: returns mat with mat[j, i] as written."""
    with open(path) as f:
        header = f.readline().split()
        n = int(header[0])
        vals = [float(v) for v in f.read().replace("]", " ").split()]
    if len(vals) != n * n:
        raise ValueError(f"expected {n * n} values, found {len(vals)}")
    return np.array(vals).reshape(n, n).T.copy()


def write_xpm_matrix(mat: np.ndarray, path: str, title: str, legend: str) -> None:
    """XPM image of the matrix (lib/correlation_core.cpp:227-278): 80 grey
    levels '!'..'p' spanning [min, max] of the off-diagonal values (0..1 when
    they are all equal), white at the minimum; row i of the image holds
    mat[j*n+i] for j, as correl.dat does."""
    m = np.asarray(mat, dtype=np.float64)
    n = m.shape[0]
    off = m.T[~np.eye(n, dtype=bool)] if n > 1 else np.array([])
    lo = float(off.min()) if off.size else 1e100
    hi = float(off.max()) if off.size else -1e100
    if lo >= hi:
        lo, hi = 0.0, 1.0
    nlev = 80
    out = ["/* XPM */\nstatic char *gmx_correlation_xpm[] = {\n", f'"{n} {n} {nlev} 1",\n']
    for lev in range(nlev):
        shade = 255 - int(255.0 * lev / (nlev - 1))
        out.append('"%c c #%02x%02x%02x",\n' % (33 + lev, shade, shade, shade))
    out.append(f"/* title: {title} */\n/* legend: {legend} */\n")
    for i in range(n):
        # truncation toward zero of the C cast, then the clamp
        lev = np.trunc((m[:, i] - lo) / (hi - lo) * (nlev - 1))
        lev = np.clip(lev, 0, nlev - 1).astype(int)
        out.append('"' + "".join(chr(33 + int(v)) for v in lev) + '"' + ("" if i + 1 == n else ",") + "\n")
    out.append("};\n")
    with open(path, "w") as f:
        f.write("".join(out))


def write_dump(traj: np.ndarray, path: str) -> None:
    """The -dump text of the mean-centred trajectory
    (src/gmx_correlation.cpp:225-244): "natoms nframes", then per frame, per
    atom, "%.10g %.10g %.10g"."""
    t = np.asarray(traj, dtype=np.float64)
    natoms, nframes = t.shape[0], t.shape[1]
    lines = [f"{natoms} {nframes}\n"]
    for fr in range(nframes):
        for a in range(natoms):
            x = t[a, fr]
            lines.append("%.10g %.10g %.10g\n" % (x[0], x[1], x[2]))
    with open(path, "w") as f:
        f.write("".join(lines))


def read_dump(path: str) -> np.ndarray:
    """Inverse of write_dump: t_traj.x as [natoms][nframes][3] float64, so a
    trajectory dumped by the reference tool feeds kraskov_prepare and the GPU
    path without GROMACS.  Raises ValueError on a malformed file."""
    with open(path) as f:
        head = f.readline().split()
        if len(head) != 2:
            raise ValueError(f"{path}: first line must be 'natoms nframes'")
        natoms, nframes = int(head[0]), int(head[1])
        if natoms < 1 or nframes < 1:
            raise ValueError(f"{path}: natoms and nframes must be positive")
        vals = np.array(f.read().split(), dtype=np.float64)
    if vals.size != natoms * nframes * 3:
        raise ValueError(f"{path}: expected {natoms * nframes * 3} values, found {vals.size}")
    return np.ascontiguousarray(vals.reshape(nframes, natoms, 3).transpose(1, 0, 2))


def read_xtc(path: str, atoms=None, max_frames: int | None = None):
    """Frames of an XTC trajectory: float32 coordinates [nframes][natoms][3]
    in nm (the layout center() takes), plus steps, times, boxes [nframes][3][3]
    and the file's precision.  atoms: optional atom indices to keep (a static
    selection, as the reference's -select).  Host-only: no device needed."""
    L = _lib.load()
    h = ctypes.c_void_p()
    _lib.check(L.gmxcorr_xtc_open(path.encode(), 0, ctypes.byref(h)))
    frames, steps, times, boxes = [], [], [], []
    prec = ctypes.c_float(0.0)
    buf = np.empty((0, 3), np.float32)
    try:
        while max_frames is None or len(frames) < max_frames:
            n, st, tm = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_float(0.0)
            box = np.empty(9, np.float32)
            args = (ctypes.byref(n), ctypes.byref(st), ctypes.byref(tm), _lib.ptr(box, _lib._fp))
            rc = L.gmxcorr_xtc_read(h, *args, _lib.ptr(buf, _lib._fp), len(buf),
                                    ctypes.byref(prec))
            if rc == _lib.GMXCORR_EOF:
                break
            if rc != _lib.GMXCORR_OK and n.value > len(buf):
                buf = np.empty((n.value, 3), np.float32)      # sized from the header
                rc = L.gmxcorr_xtc_read(h, *args, _lib.ptr(buf, _lib._fp), len(buf),
                                        ctypes.byref(prec))
            _lib.check(rc)
            x = buf[:n.value] if atoms is None else buf[np.asarray(atoms)]
            frames.append(x.copy())
            steps.append(st.value)
            times.append(tm.value)
            boxes.append(box.reshape(3, 3).copy())
    finally:
        L.gmxcorr_xtc_close(h)
    if not frames:
        raise ValueError(f"{path}: no XTC frames")
    return (np.ascontiguousarray(np.stack(frames)), np.array(steps), np.array(times, np.float32),
            np.stack(boxes), float(prec.value))


def write_xtc(path: str, coords, precision: float = 1000.0, steps=None, times=None,
              boxes=None) -> None:
    """Write float32 frames [nframes][natoms][3] (nm) as an XTC trajectory."""
    c = np.ascontiguousarray(coords, dtype=np.float32)
    T, n, _ = c.shape
    L = _lib.load()
    h = ctypes.c_void_p()
    _lib.check(L.gmxcorr_xtc_open(path.encode(), 1, ctypes.byref(h)))
    try:
        for f in range(T):
            box = (np.zeros(9, np.float32) if boxes is None
                   else np.ascontiguousarray(boxes[f], np.float32).reshape(9))
            _lib.check(L.gmxcorr_xtc_write(h, n, int(steps[f]) if steps is not None else f,
                                           float(times[f]) if times is not None else float(f),
                                           _lib.ptr(box, _lib._fp), _lib.ptr(c[f], _lib._fp),
                                           float(precision)))
    finally:
        L.gmxcorr_xtc_close(h)
