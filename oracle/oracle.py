"""ctypes access to the CPU oracle (liboracle.so) and, when it has been built,
to the reference's own CPU code (_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
benchmark's CPU-baseline leg, as the checker.  The product package never
imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_u8p = ctypes.POINTER(ctypes.c_ubyte)
_i = ctypes.c_int


def build(with_ref: bool | None = None) -> None:
    """Compile liboracle.so, and _ref/libref.so if the reference is present."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if with_ref is None:
        with_ref = os.path.isdir("/root/reference/lib")
    if with_ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _ptr(a, typ):
    return a.ctypes.data_as(typ)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(with_ref=False)
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_center.argtypes = [_fp, _i, _i, _dp]
        L.oracle_prepare.argtypes = [_dp, _i, _i, _dp, _dp]
        L.oracle_ksg_pair_grid.argtypes = [_dp, _i, _i, _i, _i, _dp, _ip, _ip, _dp, _dp, _dp]
        L.oracle_ksg_pair_grid.restype = ctypes.c_double
        L.oracle_ksg_pair_brute.argtypes = [_dp, _i, _i, _i, _i, _dp, _ip, _ip, _dp, _dp, _dp,
                                            _u8p, _ip]
        L.oracle_ksg_pair_brute.restype = ctypes.c_double
        L.oracle_ksg_matrix.argtypes = [_dp, _i, _i, _i, _dp, _dp, _i, _i]
        L.oracle_ksg_pairs.argtypes = [_dp, _i, _i, _dp, _i, _ip, _ip, _dp, _i]
        L.oracle_ksg_pairs_baseline.argtypes = [_dp, _i, _i, _dp, _i, _ip, _ip, _dp, _i]
        L.oracle_max_threads.restype = _i
        L.oracle_pearsify.argtypes = [_dp, _i, _i]
        L.oracle_lu_det.argtypes = [_dp, _i]
        L.oracle_lu_det.restype = ctypes.c_double
        L.oracle_covariance.argtypes = [_dp, _i, _i, _dp]
        L.oracle_gauss_pair.argtypes = [_dp, _i, _i, _i]
        L.oracle_gauss_pair.restype = ctypes.c_double
        L.oracle_gauss_matrix.argtypes = [_dp, _i, _i, _dp]
        L.oracle_gauss_matrix_from_cov.argtypes = [_dp, _i, _dp]
        L.oracle_erfinv.argtypes = [ctypes.c_double]
        L.oracle_erfinv.restype = ctypes.c_double
        L.oracle_copula.argtypes = [_dp, _i, _i, _dp]
        L.oracle_gcmi_pair.argtypes = [_dp, _i, _i, _i]
        L.oracle_gcmi_pair.restype = ctypes.c_double
        L.oracle_gcmi_matrix.argtypes = [_dp, _i, _i, _dp, _i]
        L.oracle_gcmi_matrix.restype = _i
        L.oracle_te_pair.argtypes = [_dp, _i, _i, _i, _i, _i, _dp]
        L.oracle_te_pair.restype = ctypes.c_double
        L.oracle_te_psi.argtypes = [_i, _dp]
        L.oracle_te_matrix.argtypes = [_dp, _i, _i, _i, _i, _dp, _i]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        R = ctypes.CDLL(REF_PATH)
        R.ref_center.argtypes = [_fp, _i, _i, _dp]
        R.ref_prepare.argtypes = [_fp, _i, _i, _dp, _dp]
        R.ref_ksg_pair.argtypes = [_fp, _i, _i, _i, _i, _i]
        R.ref_ksg_pair.restype = ctypes.c_double
        R.ref_ksg_matrix.argtypes = [_fp, _i, _i, _i, _i, _dp]
        R.ref_ksg_matrix.restype = _i
        R.ref_ksg_pairs.argtypes = [_fp, _i, _i, _i, _i, _ip, _ip, _dp, _i]
        R.ref_gauss_matrix.argtypes = [_fp, _i, _i, _dp]
        R.ref_gcmi_matrix.argtypes = [_fp, _i, _i, _i, _dp]
        R.ref_gcmi_matrix.restype = _i
        R.ref_te_matrix.argtypes = [_fp, _i, _i, _i, _i, _i, _dp]
        R.ref_pearsify.argtypes = [_dp, _i]
        R.ref_write_matrix.argtypes = [_dp, _i, ctypes.c_char_p]
        R.ref_write_xpm.argtypes = [_dp, _i, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]
        _ref = R
    return _ref


# ---------------------------------------------------------------- oracle API

def center(coords: np.ndarray) -> np.ndarray:
    c = np.ascontiguousarray(coords, dtype=np.float32)
    T, n, _ = c.shape
    out = np.empty((n, T, 3), dtype=np.float64)
    lib().oracle_center(_ptr(c, _fp), T, n, _ptr(out, _dp))
    return out


def prepare(traj: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    t = np.ascontiguousarray(traj, dtype=np.float64)
    n, T, _ = t.shape
    kx = np.empty((3 * n, T), dtype=np.float64)
    psi = np.empty(T + 1, dtype=np.float64)
    lib().oracle_prepare(_ptr(t, _dp), n, T, _ptr(kx, _dp), _ptr(psi, _dp))
    return kx, psi


def ksg_pair(kx, psi, a1, a2, k, method="grid", detail=False):
    kx = np.ascontiguousarray(kx, dtype=np.float64)
    T = kx.shape[1]
    nx = np.empty(T, np.int32); ny = np.empty(T, np.int32)
    ex = np.empty(T); ey = np.empty(T); ek = np.empty(T)
    if method == "grid":
        v = lib().oracle_ksg_pair_grid(_ptr(kx, _dp), T, a1, a2, k, _ptr(psi, _dp),
                                       _ptr(nx, _ip), _ptr(ny, _ip), _ptr(ex, _dp),
                                       _ptr(ey, _dp), _ptr(ek, _dp))
        tie = None
    else:
        tie = np.empty(T, np.uint8)
        nt = ctypes.c_int(0)
        v = lib().oracle_ksg_pair_brute(_ptr(kx, _dp), T, a1, a2, k, _ptr(psi, _dp),
                                        _ptr(nx, _ip), _ptr(ny, _ip), _ptr(ex, _dp),
                                        _ptr(ey, _dp), _ptr(ek, _dp), _ptr(tie, _u8p),
                                        ctypes.byref(nt))
    if not detail:
        return v
    return v, dict(nx=nx, ny=ny, epsx=ex, epsy=ey, epsk=ek, tie=tie)


def ksg_matrix(kx, psi, natoms, k, nthreads=0, method="grid") -> np.ndarray:
    kx = np.ascontiguousarray(kx, dtype=np.float64)
    T = kx.shape[1]
    mat = np.empty((natoms, natoms), dtype=np.float64)
    lib().oracle_ksg_matrix(_ptr(kx, _dp), natoms, T, k, _ptr(psi, _dp), _ptr(mat, _dp),
                            nthreads, 0 if method == "grid" else 1)
    return mat


def ksg_pairs(kx, psi, k, a1, a2, nthreads=0) -> np.ndarray:
    kx = np.ascontiguousarray(kx, dtype=np.float64)
    a1 = np.ascontiguousarray(a1, dtype=np.int32)
    a2 = np.ascontiguousarray(a2, dtype=np.int32)
    out = np.empty(len(a1))
    lib().oracle_ksg_pairs(_ptr(kx, _dp), kx.shape[1], k, _ptr(psi, _dp), len(a1),
                           _ptr(a1, _ip), _ptr(a2, _ip), _ptr(out, _dp), nthreads)
    return out


def ksg_pairs_baseline(kx, psi, k, a1, a2, nthreads=0) -> np.ndarray:
    """The timed CPU baseline: same MI as ksg_pairs, with the reference's
    memory layout and allocation pattern (oracle_ksg_pair_baseline)."""
    kx = np.ascontiguousarray(kx, dtype=np.float64)
    a1 = np.ascontiguousarray(a1, dtype=np.int32)
    a2 = np.ascontiguousarray(a2, dtype=np.int32)
    out = np.empty(len(a1))
    lib().oracle_ksg_pairs_baseline(_ptr(kx, _dp), kx.shape[1], k, _ptr(psi, _dp), len(a1),
                                    _ptr(a1, _ip), _ptr(a2, _ip), _ptr(out, _dp), nthreads)
    return out


def max_threads() -> int:
    return int(lib().oracle_max_threads())


def pearsify(mat: np.ndarray, dim: int = 3) -> np.ndarray:
    m = np.array(mat, dtype=np.float64, order="C", copy=True)
    lib().oracle_pearsify(_ptr(m, _dp), m.shape[0], dim)
    return m


def lu_det(m: np.ndarray) -> float:
    m = np.ascontiguousarray(m, dtype=np.float64)
    return lib().oracle_lu_det(_ptr(m, _dp), m.shape[0])


def covariance(traj) -> np.ndarray:
    t = np.ascontiguousarray(traj, dtype=np.float64)
    n, T, _ = t.shape
    cov = np.empty((3 * n, 3 * n))
    lib().oracle_covariance(_ptr(t, _dp), n, T, _ptr(cov, _dp))
    return cov


def gauss_matrix(traj) -> np.ndarray:
    t = np.ascontiguousarray(traj, dtype=np.float64)
    n, T, _ = t.shape
    mat = np.empty((n, n))
    lib().oracle_gauss_matrix(_ptr(t, _dp), n, T, _ptr(mat, _dp))
    return mat


def gauss_matrix_from_cov(cov, natoms) -> np.ndarray:
    c = np.ascontiguousarray(cov, dtype=np.float64)
    mat = np.empty((natoms, natoms))
    lib().oracle_gauss_matrix_from_cov(_ptr(c, _dp), natoms, _ptr(mat, _dp))
    return mat


def erfinv(x: float) -> float:
    return lib().oracle_erfinv(float(x))


def copula(traj) -> np.ndarray:
    t = np.ascontiguousarray(traj, dtype=np.float64)
    n, T, _ = t.shape
    z = np.empty((3 * n, T))
    lib().oracle_copula(_ptr(t, _dp), n, T, _ptr(z, _dp))
    return z


def gcmi_pair(z, a1, a2) -> float:
    z = np.ascontiguousarray(z, dtype=np.float64)
    return lib().oracle_gcmi_pair(_ptr(z, _dp), z.shape[1], a1, a2)


def gcmi_matrix(traj, nthreads=0) -> np.ndarray:
    t = np.ascontiguousarray(traj, dtype=np.float64)
    n, T, _ = t.shape
    mat = np.empty((n, n))
    rc = lib().oracle_gcmi_matrix(_ptr(t, _dp), n, T, _ptr(mat, _dp), nthreads)
    if rc != 0:
        raise RuntimeError("gcmi_corrmatrix: need at least 2 atoms and 6 frames")
    return mat


def te_psi(n: int) -> np.ndarray:
    psi = np.empty(n + 1)
    lib().oracle_te_psi(n, _ptr(psi, _dp))
    return psi


def te_pair(traj, a1, a2, k, lag) -> float:
    t = np.ascontiguousarray(traj, dtype=np.float64)
    n, T, _ = t.shape
    psi = te_psi(T - lag)
    return lib().oracle_te_pair(_ptr(t, _dp), T, a1, a2, k, lag, _ptr(psi, _dp))


def te_matrix(traj, k, lag, nthreads=0) -> np.ndarray:
    t = np.ascontiguousarray(traj, dtype=np.float64)
    n, T, _ = t.shape
    mat = np.empty((n, n))
    lib().oracle_te_matrix(_ptr(t, _dp), n, T, k, lag, _ptr(mat, _dp), nthreads)
    return mat


# ------------------------------------------------------------- reference API

def ref_prepare(coords):
    c = np.ascontiguousarray(coords, dtype=np.float32)
    T, n, _ = c.shape
    kx = np.empty((3 * n, T)); psi = np.empty(T + 1)
    ref().ref_prepare(_ptr(c, _fp), T, n, _ptr(kx, _dp), _ptr(psi, _dp))
    return kx, psi


def ref_center(coords):
    c = np.ascontiguousarray(coords, dtype=np.float32)
    T, n, _ = c.shape
    out = np.empty((n, T, 3))
    ref().ref_center(_ptr(c, _fp), T, n, _ptr(out, _dp))
    return out


def ref_ksg_matrix(coords, k, nthreads=0):
    c = np.ascontiguousarray(coords, dtype=np.float32)
    T, n, _ = c.shape
    mat = np.empty((n, n))
    if ref().ref_ksg_matrix(_ptr(c, _fp), T, n, k, nthreads, _ptr(mat, _dp)) != 0:
        raise RuntimeError("kraskov_corrmatrix rejected the input")
    return mat


def ref_ksg_pairs(coords, k, a1, a2, nthreads=0):
    c = np.ascontiguousarray(coords, dtype=np.float32)
    T, n, _ = c.shape
    a1 = np.ascontiguousarray(a1, dtype=np.int32)
    a2 = np.ascontiguousarray(a2, dtype=np.int32)
    out = np.empty(len(a1))
    ref().ref_ksg_pairs(_ptr(c, _fp), T, n, k, len(a1), _ptr(a1, _ip), _ptr(a2, _ip),
                        _ptr(out, _dp), nthreads)
    return out


def ref_gauss_matrix(coords):
    c = np.ascontiguousarray(coords, dtype=np.float32)
    T, n, _ = c.shape
    mat = np.empty((n, n))
    ref().ref_gauss_matrix(_ptr(c, _fp), T, n, _ptr(mat, _dp))
    return mat


def ref_gcmi_matrix(coords, nthreads=0):
    c = np.ascontiguousarray(coords, dtype=np.float32)
    T, n, _ = c.shape
    mat = np.empty((n, n))
    if ref().ref_gcmi_matrix(_ptr(c, _fp), T, n, nthreads, _ptr(mat, _dp)) != 0:
        raise RuntimeError("gcmi_corrmatrix rejected the input")
    return mat


def ref_te_matrix(coords, k, lag, nthreads=0):
    c = np.ascontiguousarray(coords, dtype=np.float32)
    T, n, _ = c.shape
    mat = np.empty((n, n))
    ref().ref_te_matrix(_ptr(c, _fp), T, n, k, lag, nthreads, _ptr(mat, _dp))
    return mat


def ref_write_matrix(mat, path):
    m = np.ascontiguousarray(mat, dtype=np.float64)
    ref().ref_write_matrix(_ptr(m, _dp), m.shape[0], path.encode())


def ref_write_xpm(mat, path, title, legend):
    m = np.ascontiguousarray(mat, dtype=np.float64)
    ref().ref_write_xpm(_ptr(m, _dp), m.shape[0], path.encode(), title.encode(), legend.encode())
