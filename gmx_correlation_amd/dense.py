"""Host-side mirror of the reference's Gaussian and GCMI matrices, backed by
the HIP library (no CPU path).

  gauss_corrmatrix()  lib/correlation_core.cpp:280-311  (-linear)
  gcmi_corrmatrix()   lib/gcmi.cpp:246-317              (-gcmi)

Both take t_traj.x [natoms][nframes][3] (mean-centred) and return the
[natoms][natoms] MI matrix in nats with the 2000 diagonal.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def _traj(traj) -> np.ndarray:
    t = np.ascontiguousarray(traj, dtype=np.float64)
    if t.ndim != 3 or t.shape[2] != 3:
        raise ValueError("traj must be [natoms][nframes][3]")
    return t


def gauss_corrmatrix(traj, return_cov: bool = False, stats: bool = False):
    t = _traj(traj)
    n, N, _ = t.shape
    mat = np.zeros((n, n))
    cov = np.zeros((3 * n, 3 * n)) if return_cov else None
    st = _lib.DenseStats()
    _lib.check(_lib.load().gmxcorr_gauss_matrix(_lib.ptr(t), n, N, _lib.ptr(mat), _lib.ptr(cov),
                                                ctypes.byref(st)))
    out = [mat]
    if return_cov:
        out.append(cov)
    if stats:
        out.append(st.as_dict())
    return out[0] if len(out) == 1 else tuple(out)


def gcmi_corrmatrix(traj, return_cov: bool = False, return_z: bool = False, stats: bool = False):
    t = _traj(traj)
    n, N, _ = t.shape
    if n < 2:
        raise ValueError("gcmi_corrmatrix: need at least 2 atoms")
    if N < 6:
        raise ValueError("gcmi_corrmatrix: need at least 6 frames for a non-singular 6x6 covariance")
    mat = np.zeros((n, n))
    cov = np.zeros((3 * n, 3 * n)) if return_cov else None
    z = np.zeros((3 * n, N)) if return_z else None
    st = _lib.DenseStats()
    _lib.check(_lib.load().gmxcorr_gcmi_matrix(_lib.ptr(t), n, N, _lib.ptr(mat), _lib.ptr(cov),
                                               _lib.ptr(z), ctypes.byref(st)))
    out = [mat]
    if return_cov:
        out.append(cov)
    if return_z:
        out.append(z)
    if stats:
        out.append(st.as_dict())
    return out[0] if len(out) == 1 else tuple(out)


def te_matrix(traj, k: int = 100, lag: int = 1, return_ms: bool = False):
    """Transfer-entropy matrix mat[a1][a2] = TE(a2 -> a1) of the raw
    mean-centred trajectory (lib/transfer_entropy.cpp:63-221)."""
    t = _traj(traj)
    n, N, _ = t.shape
    mat = np.zeros((n, n))
    ms = ctypes.c_double(0.0)
    _lib.check(_lib.load().gmxcorr_te_matrix(_lib.ptr(t), n, N, k, lag, _lib.ptr(mat),
                                             ctypes.byref(ms)))
    return (mat, ms.value) if return_ms else mat


def te_pairs(traj, a1, a2, k: int = 100, lag: int = 1):
    """TE(a2[p] -> a1[p]) for a list of ordered pairs, and the kernel ms."""
    t = _traj(traj)
    n, N, _ = t.shape
    a1 = np.ascontiguousarray(a1, dtype=np.int32)
    a2 = np.ascontiguousarray(a2, dtype=np.int32)
    out = np.empty(len(a1))
    ms = ctypes.c_double(0.0)
    _lib.check(_lib.load().gmxcorr_te_pairs(_lib.ptr(t), n, N, k, lag, len(a1),
                                            _lib.ptr(a1, _lib._ip), _lib.ptr(a2, _lib._ip),
                                            _lib.ptr(out), ctypes.byref(ms)))
    return out, ms.value


def te_set_path(path: int) -> None:
    """Pass-1 kernel of later TE calls in this process: 0 automatic (prefix
    stores for k <= 128 and 256 < N - lag <= 16384 when the pairs average 8
    or more per target, else the heap for k <= 128, the radix select above),
    1 heap (k <= 128), 2 radix select, 3 prefix stores (k <= 128,
    N - lag <= 16384)."""
    _lib.check(_lib.load().gmxcorr_te_set_path(int(path)))


def te_stats() -> dict:
    """Counts of the last TE call: queries, those resolved from the prefix
    stores, those left to the heap kernel, and whether the prefix path ran."""
    out = (ctypes.c_longlong * 4)()
    _lib.check(_lib.load().gmxcorr_te_stats(out))
    return {"queries": out[0], "prefix_resolved": out[1], "heap_listed": out[2],
            "prefix_path": bool(out[3])}
