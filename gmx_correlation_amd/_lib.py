"""ctypes binding of libgmxcorr_hip.so (include/gmxcorr.h).

There is no CPU fallback anywhere in this package: if the library is missing
or has no usable gfx950 device, the calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libgmxcorr_hip.so")

GMXCORR_OK = 0
STATUS_NAMES = {-1: "EINVAL", -2: "ENODEV", -3: "ENOMEM", -4: "EHIP", -5: "EUNSUP"}
GMXCORR_EOF = 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_u8p = ctypes.POINTER(ctypes.c_ubyte)
_fp = ctypes.POINTER(ctypes.c_float)
_i = ctypes.c_int
_ll = ctypes.c_longlong
_llp = ctypes.POINTER(ctypes.c_longlong)


class KsgStats(ctypes.Structure):
    _fields_ = [
        ("ms_load", ctypes.c_double),
        ("ms_build", ctypes.c_double),
        ("ms_pairs", ctypes.c_double),
        ("ms_fixup", ctypes.c_double),
        ("npairs", _ll),
        ("fixup_rows", _ll),
        ("long_rows", _ll),
        ("ties", _ll),
        ("row_bytes", _ll),
        ("nframes", _i),
        ("k", _i),
        ("natoms", _i),
        ("pitch", _i),
        ("kind", _i),
        ("s4", _i),
        ("nsplit", _i),
        ("chunk", _i),
        ("slots", _i),
        ("wave_atoms", _i),
        ("waves", _i),
        ("all_pairs", _i),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class DenseStats(ctypes.Structure):
    _fields_ = [
        ("ms_load", ctypes.c_double),
        ("ms_copula", ctypes.c_double),
        ("ms_split", ctypes.c_double),
        ("ms_gemm", ctypes.c_double),
        ("ms_pairs", ctypes.c_double),
        ("gemm_flops", ctypes.c_double),
        ("natoms", _i),
        ("nframes", _i),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("ms_total", ctypes.c_double),
        ("ms_build", ctypes.c_double),
        ("ms_pairs", ctypes.c_double),
        ("ms_fixup", ctypes.c_double),
        ("npairs", _ll),
        ("fixup_rows", _ll),
        ("ties", _ll),
        ("nwaves", _i),
        ("long_rows", _ll),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class GmxcorrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gmxcorr {STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


# Symbols declared in include/gmxcorr.h and include/kraskov_gpu.h.
C_SYMBOLS = [
    "gmxcorr_last_error", "gmxcorr_probe", "gmxcorr_selftest_wave", "gmxcorr_ksg_max_frames",
    "gmxcorr_ksg_matrix",
    "gmxcorr_ksg_plan_create", "gmxcorr_ksg_plan_load", "gmxcorr_ksg_plan_run_all",
    "gmxcorr_ksg_plan_run_pairs", "gmxcorr_ksg_plan_result", "gmxcorr_ksg_plan_matrix",
    "gmxcorr_ksg_plan_stats", "gmxcorr_ksg_plan_destroy", "gmxcorr_ksg_plan_set_variant",
    "gmxcorr_ksg_variant_fits", "gmxcorr_ksg_plan_probe", "gmxcorr_ksg_plan_layout", "gmxcorr_ksg_plan_create_budget",
    "gmxcorr_ksg_plan_set_shape", "gmxcorr_ksg_pairs_device",
    "gmxcorr_ksg_pairs_device_noalloc",
    "gmxcorr_gauss_matrix", "gmxcorr_gcmi_matrix", "gmxcorr_dense_set_slab_frames",
    "gmxcorr_te_matrix", "gmxcorr_te_pairs", "gmxcorr_te_set_path", "gmxcorr_te_stats",
    "gmxcorr_ksg_batch_create", "gmxcorr_ksg_batch_add", "gmxcorr_ksg_batch_run",
    "gmxcorr_ksg_batch_traj_bytes", "gmxcorr_ksg_batch_set_chunk",
    "gmxcorr_ksg_batch_result", "gmxcorr_ksg_batch_stats", "gmxcorr_ksg_batch_destroy",
    "gmxcorr_ksg_plan_features", "gmxcorr_mlp_num_params", "gmxcorr_mlp_create",
    "gmxcorr_mlp_set_params", "gmxcorr_mlp_get_params", "gmxcorr_mlp_grad",
    "gmxcorr_mlp_train_epoch", "gmxcorr_mlp_forward", "gmxcorr_mlp_destroy",
    "gmxcorr_xtc_open", "gmxcorr_xtc_read", "gmxcorr_xtc_write", "gmxcorr_xtc_close",
]
CXX_SYMBOLS = ["_Z17kraskov_gpu_probev", "_Z22kraskov_corrmatrix_gpuPK9t_kraskoviPdi"]

_lib = None


def load() -> ctypes.CDLL:
    """Load the in-tree library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run gmx_correlation_amd.build.build() first")
    L = ctypes.CDLL(LIB_PATH)
    L.gmxcorr_last_error.restype = ctypes.c_char_p
    L.gmxcorr_probe.restype = _i
    L.gmxcorr_ksg_max_frames.restype = _i
    L.gmxcorr_selftest_wave.argtypes = [_ip]
    L.gmxcorr_ksg_matrix.argtypes = [_dp, _i, _i, _dp, _i, _dp]
    L.gmxcorr_ksg_plan_create.argtypes = [_i, _i, _i, ctypes.POINTER(ctypes.c_void_p)]
    L.gmxcorr_ksg_plan_load.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.gmxcorr_ksg_plan_run_all.argtypes = [ctypes.c_void_p]
    L.gmxcorr_ksg_plan_run_pairs.argtypes = [ctypes.c_void_p, _ll, _ip, _ip, _ip, _ip, _u8p]
    L.gmxcorr_ksg_plan_result.argtypes = [ctypes.c_void_p, _dp, _ll]
    L.gmxcorr_ksg_plan_matrix.argtypes = [ctypes.c_void_p, _dp]
    L.gmxcorr_ksg_plan_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(KsgStats)]
    L.gmxcorr_ksg_plan_destroy.argtypes = [ctypes.c_void_p]
    L.gmxcorr_ksg_plan_destroy.restype = None
    L.gmxcorr_ksg_plan_set_variant.argtypes = [ctypes.c_void_p, _i, _i]
    L.gmxcorr_ksg_plan_layout.argtypes = [_i, _i, _i, _ll, _ip, _ip, _ip]
    L.gmxcorr_ksg_plan_create_budget.argtypes = [_i, _i, _i, _ll, ctypes.POINTER(ctypes.c_void_p)]
    L.gmxcorr_ksg_plan_set_shape.argtypes = [ctypes.c_void_p, _i, _i, _i]
    L.gmxcorr_ksg_pairs_device.argtypes = [ctypes.c_void_p, _i, _i, ctypes.c_void_p, _i, _ll,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p]
    L.gmxcorr_ksg_pairs_device_noalloc.argtypes = L.gmxcorr_ksg_pairs_device.argtypes
    L.gmxcorr_ksg_variant_fits.argtypes = [_i, _i, _i, _i]
    L.gmxcorr_ksg_plan_probe.argtypes = [ctypes.c_void_p, ctypes.c_double,
                                         ctypes.POINTER(ctypes.c_double)]
    L.gmxcorr_dense_set_slab_frames.argtypes = [_i]
    L.gmxcorr_gauss_matrix.argtypes = [_dp, _i, _i, _dp, _dp, ctypes.POINTER(DenseStats)]
    L.gmxcorr_gcmi_matrix.argtypes = [_dp, _i, _i, _dp, _dp, _dp, ctypes.POINTER(DenseStats)]
    _fp = ctypes.POINTER(ctypes.c_float)
    _vp = ctypes.c_void_p
    L.gmxcorr_ksg_plan_features.argtypes = [_vp, _ll, _ip, _ip, _i, _fp, _dp]
    L.gmxcorr_mlp_num_params.restype = _i
    L.gmxcorr_mlp_create.argtypes = [_i, ctypes.POINTER(ctypes.c_void_p)]
    L.gmxcorr_mlp_set_params.argtypes = [_vp, _fp]
    L.gmxcorr_mlp_get_params.argtypes = [_vp, _fp]
    L.gmxcorr_mlp_grad.argtypes = [_vp, _fp, _fp, _i, _fp, _fp]
    L.gmxcorr_mlp_train_epoch.argtypes = [_vp, _fp, _fp, _ll, _i, ctypes.c_float, _fp]
    L.gmxcorr_mlp_forward.argtypes = [_vp, _fp, _ll, _i, _fp, _dp]
    L.gmxcorr_mlp_destroy.argtypes = [_vp]
    L.gmxcorr_mlp_destroy.restype = None
    L.gmxcorr_te_matrix.argtypes = [_dp, _i, _i, _i, _i, _dp, _dp]
    L.gmxcorr_te_pairs.argtypes = [_dp, _i, _i, _i, _i, _ll, _ip, _ip, _dp, _dp]
    L.gmxcorr_te_set_path.argtypes = [_i]
    L.gmxcorr_te_stats.argtypes = [_llp]
    L.gmxcorr_ksg_batch_create.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    L.gmxcorr_ksg_batch_add.argtypes = [_vp, _dp, _dp, _i, _i, _i, _i, _i]
    L.gmxcorr_ksg_batch_traj_bytes.argtypes = [_vp, _i]
    L.gmxcorr_ksg_batch_traj_bytes.restype = _ll
    L.gmxcorr_ksg_batch_run.argtypes = [_vp, _ll, _i]
    L.gmxcorr_ksg_batch_set_chunk.argtypes = [_vp, _i, _i]
    L.gmxcorr_ksg_batch_result.argtypes = [_vp, _i, _dp, _ll, _dp, _ip]
    L.gmxcorr_ksg_batch_stats.argtypes = [_vp, ctypes.POINTER(BatchStats)]
    L.gmxcorr_ksg_batch_destroy.argtypes = [_vp]
    L.gmxcorr_ksg_batch_destroy.restype = None
    L.gmxcorr_xtc_open.argtypes = [ctypes.c_char_p, _i, ctypes.POINTER(ctypes.c_void_p)]
    L.gmxcorr_xtc_read.argtypes = [_vp, _ip, _ip, _fp, _fp, _fp, _i, _fp]
    L.gmxcorr_xtc_write.argtypes = [_vp, _i, _i, ctypes.c_float, _fp, _fp, ctypes.c_float]
    L.gmxcorr_xtc_close.argtypes = [_vp]
    L.gmxcorr_xtc_close.restype = None
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != GMXCORR_OK:
        raise GmxcorrError(rc, load().gmxcorr_last_error().decode(errors="replace"))


def ptr(a: np.ndarray, typ=_dp):
    return a.ctypes.data_as(typ) if a is not None else None


def probe() -> bool:
    return bool(load().gmxcorr_probe())


class TKraskov(ctypes.Structure):
    """t_kraskov as lib/correlation_core.h:45-52 lays it out."""
    _fields_ = [("x", ctypes.POINTER(ctypes.POINTER(ctypes.c_double))),
                ("psi", ctypes.POINTER(ctypes.c_double)),
                ("scal", ctypes.POINTER(ctypes.c_double)),
                ("N", ctypes.c_int), ("dim", ctypes.c_int)]


def kraskov_corrmatrix_gpu(rows: np.ndarray, psi: np.ndarray, natoms: int, k: int) -> np.ndarray:
    """Call the exported C++ kraskov_corrmatrix_gpu(const t_kraskov*, int,
    double*, int) (lib/kraskov_gpu.h:32) on prepared rows [3 natoms][N] and
    psi [N + 1], as the reference's dispatcher does after kraskov_prepare
    (lib/correlation_core.cpp:337-348).  Returns the MI matrix."""
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    psi = np.ascontiguousarray(psi, dtype=np.float64)
    nrows, N = rows.shape
    dp = ctypes.POINTER(ctypes.c_double)
    ptrs = (dp * nrows)(*[rows[c].ctypes.data_as(dp) for c in range(nrows)])
    scal = np.ones(nrows)
    kr = TKraskov(ptrs, psi.ctypes.data_as(dp), scal.ctypes.data_as(dp), N, nrows)
    mat = np.zeros((natoms, natoms))
    fn = getattr(load(), CXX_SYMBOLS[1])
    fn.argtypes = [ctypes.POINTER(TKraskov), ctypes.c_int, dp, ctypes.c_int]
    fn.restype = None
    fn(ctypes.byref(kr), natoms, mat.ctypes.data_as(dp), k)
    return mat
