"""ctypes bindings of the test oracle.

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py use it, and only as the checker.  Two libraries:

* liboracle.so - the plain-C restatement of the reference fusion path
  (pct_oracle.c; every function cites the reference lines it follows);
* _ref/libpctref.so - the reference CPU program itself
  (src/CPU/NewPatchedBrainPureC.cpp), compiled from the reference tree by
  oracle/Makefile where that tree exists.  It pins the restatement.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int, c_long

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libpctref.so")
FP = POINTER(c_float)
DP = POINTER(c_double)

_oracle = None
_ref = None


def _fp(a):
    return a.ctypes.data_as(FP)


def oracle_lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_PATH):
            raise ImportError(f"{ORACLE_PATH} missing: run make -C oracle")
        l = ctypes.CDLL(ORACLE_PATH)
        l.oracle_normalize_labels.argtypes = [FP, ctypes.c_size_t]
        l.oracle_normalize_labels.restype = None
        l.oracle_div.argtypes = [c_double, c_double]
        l.oracle_div.restype = c_double
        l.oracle_fuse.argtypes = [c_int] * 5 + [FP] * 4 + [c_double, c_int]
        l.oracle_fuse.restype = None
        l.oracle_voxel.argtypes = [c_int] * 5 + [FP] * 3 + [c_double, c_int, c_int, c_int, DP, DP]
        l.oracle_voxel.restype = c_double
        l.oracle_voxels.argtypes = [c_int] * 5 + [FP] * 3 + [c_double, POINTER(c_long), c_long,
                                                              DP, c_int]
        l.oracle_voxels.restype = None
        l.oracle_median.argtypes = [FP, c_int]
        l.oracle_median.restype = c_double
        l.oracle_regularize.argtypes = [c_int] * 5 + [FP]
        l.oracle_regularize.restype = c_int
        l.oracle_round.argtypes = [FP, ctypes.c_size_t]
        l.oracle_round.restype = None
        l.oracle_to_hu.argtypes = [FP, ctypes.c_size_t]
        l.oracle_to_hu.restype = None
        l.oracle_postprocess.argtypes = [c_int] * 4 + [FP]
        l.oracle_postprocess.restype = None
        _oracle = l
    return _oracle


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref_lib():
    global _ref
    if _ref is None:
        if not ref_available():
            raise ImportError(f"{REF_PATH} missing (needs the reference tree; make -C oracle)")
        l = ctypes.CDLL(REF_PATH)
        l.pctref_fuse.argtypes = [c_int] * 5 + [FP] * 4 + [c_double, c_int, c_int]
        l.pctref_fuse.restype = None
        l.pctref_patch_similarity.argtypes = [c_int] * 3 + [FP, FP] + [c_int] * 7 + [c_double]
        l.pctref_patch_similarity.restype = c_double
        l.pctref_median.argtypes = [FP, c_int]
        l.pctref_median.restype = c_double
        l.pctref_regularize.argtypes = [c_int] * 5 + [FP]
        l.pctref_regularize.restype = None
        l.pctref_round.argtypes = [FP, c_int]
        l.pctref_round.restype = None
        l.pctref_matrixio_read.argtypes = [ctypes.c_char_p, POINTER(FP), POINTER(c_int)]
        l.pctref_matrixio_read.restype = c_int
        l.pctref_matrixio_store.argtypes = [ctypes.c_char_p, FP, c_int, c_int, c_int, c_int]
        l.pctref_matrixio_store.restype = None
        l.pctref_read_config.argtypes = [ctypes.c_char_p, POINTER(c_int), POINTER(FP),
                                         POINTER(FP), POINTER(FP)]
        l.pctref_read_config.restype = c_int
        l.pctref_free.argtypes = [ctypes.c_void_p]
        l.pctref_free.restype = None
        _ref = l
    return _ref


def _dims(img, atlas):
    nz, nx, ny = img.shape
    return nx, ny, nz, atlas.shape[0]


def _check(img, atlas, label):
    for a in (img, atlas, label):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    assert atlas.shape == label.shape and atlas.shape[1:] == img.shape


# ---- restatement --------------------------------------------------------------

def div(sigma: float, beta: float) -> float:
    return oracle_lib().oracle_div(sigma, beta)


def normalize_labels(label: np.ndarray) -> np.ndarray:
    oracle_lib().oracle_normalize_labels(_fp(label), label.size)
    return label


def fuse(img, atlas, label_norm, nh: int, sigma: float, beta: float, threads: int = 0):
    """Reference-equivalent fused value (pre-zeroed output, interior written)."""
    _check(img, atlas, label_norm)
    nx, ny, nz, n = _dims(img, atlas)
    out = np.zeros(img.shape, np.float32)
    threads = threads or (os.cpu_count() or 1)
    oracle_lib().oracle_fuse(nx, ny, nz, n, nh, _fp(img), _fp(atlas), _fp(label_norm), _fp(out),
                             div(sigma, beta), threads)
    return out


def voxel(img, atlas, label_norm, nh, sigma, beta, z, x, y):
    """(P, S, P/S) in float64 for one voxel."""
    nx, ny, nz, n = _dims(img, atlas)
    P, S = c_double(), c_double()
    r = oracle_lib().oracle_voxel(nx, ny, nz, n, nh, _fp(img), _fp(atlas), _fp(label_norm),
                                  div(sigma, beta), z, x, y, ctypes.byref(P), ctypes.byref(S))
    return P.value, S.value, r


def voxels(img, atlas, label_norm, nh, sigma, beta, flat_idx, threads: int = 0):
    """float64 P/S (NaN rule applied) at the given flat indices."""
    _check(img, atlas, label_norm)
    nx, ny, nz, n = _dims(img, atlas)
    idx = np.ascontiguousarray(flat_idx, dtype=np.int64)
    out = np.empty(idx.size, np.float64)
    threads = threads or (os.cpu_count() or 1)
    oracle_lib().oracle_voxels(nx, ny, nz, n, nh, _fp(img), _fp(atlas), _fp(label_norm),
                               div(sigma, beta), idx.ctypes.data_as(POINTER(c_long)), idx.size,
                               out.ctypes.data_as(DP), threads)
    return out


def median(patch) -> float:
    p = np.ascontiguousarray(patch, np.float32)
    return oracle_lib().oracle_median(_fp(p), p.size)


def regularize(arr, nh: int) -> int:
    nz, nx, ny = arr.shape
    return oracle_lib().oracle_regularize((nh - 1) // 2, 1, nx, ny, nz, _fp(arr))


def round_(arr):
    oracle_lib().oracle_round(_fp(arr), arr.size)
    return arr


def to_hu(arr):
    oracle_lib().oracle_to_hu(_fp(arr), arr.size)
    return arr


def postprocess(arr, nh: int):
    nz, nx, ny = arr.shape
    oracle_lib().oracle_postprocess(nh, nx, ny, nz, _fp(arr))
    return arr


# ---- the compiled reference -----------------------------------------------------

def ref_fuse(img, atlas, label_norm, nh: int, sigma: float, beta: float, threads: int = 0,
             multicore: bool = True):
    """The reference's own MultiAtlasPatchBasedSegmentation[OpenMP]."""
    _check(img, atlas, label_norm)
    nx, ny, nz, n = _dims(img, atlas)
    out = np.zeros(img.shape, np.float32)
    threads = threads or (os.cpu_count() or 1)
    # div exactly as the reference's main() computes it (cpp:1601)
    d = 2 * 27 * beta * pow(sigma, 2)
    ref_lib().pctref_fuse(nx, ny, nz, n, nh, _fp(img), _fp(atlas), _fp(label_norm), _fp(out),
                          d, threads, int(multicore))
    return out


def ref_median(patch) -> float:
    p = np.ascontiguousarray(patch, np.float32)
    return ref_lib().pctref_median(_fp(p), p.size)


def ref_regularize(arr, nh: int) -> None:
    nz, nx, ny = arr.shape
    ref_lib().pctref_regularize((nh - 1) // 2, 1, nx, ny, nz, _fp(arr))


def ref_round(arr):
    ref_lib().pctref_round(_fp(arr), arr.size)
    return arr


def ref_matrixio_store(path: str, arr, num, rows, cols, depth):
    a = np.ascontiguousarray(arr, np.float32)
    ref_lib().pctref_matrixio_store(path.encode(), _fp(a), num, rows, cols, depth)


def ref_matrixio_read(path: str):
    data = FP()
    dims = (c_int * 4)()
    rc = ref_lib().pctref_matrixio_read(path.encode(), ctypes.byref(data), dims)
    if rc != 0:
        raise ValueError("reference reader returned no dimensions")
    rows, cols, depth, num = list(dims)
    n = rows * cols * depth * num
    arr = np.ctypeslib.as_array(data, shape=(n,)).copy()
    ref_lib().pctref_free(ctypes.cast(data, ctypes.c_void_p))
    return arr, [rows, cols, depth, num]


def ref_read_config(path: str):
    dims = (c_int * 4)()
    im, at, lb = FP(), FP(), FP()
    rc = ref_lib().pctref_read_config(path.encode(), dims, ctypes.byref(im), ctypes.byref(at),
                                      ctypes.byref(lb))
    if rc != 0:
        raise ValueError("reference readMatFile failed")
    nx, ny, nz, n = list(dims)
    N = nx * ny * nz
    out = {
        "img": np.ctypeslib.as_array(im, shape=(N,)).copy().reshape(nz, nx, ny),
        "atlas": np.ctypeslib.as_array(at, shape=(N * n,)).copy().reshape(n, nz, nx, ny),
        "label": np.ctypeslib.as_array(lb, shape=(N * n,)).copy().reshape(n, nz, nx, ny),
    }
    for p in (im, at, lb):
        ref_lib().pctref_free(ctypes.cast(p, ctypes.c_void_p))
    return out
