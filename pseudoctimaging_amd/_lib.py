"""ctypes binding of libpct.so, the C ABI declared in include/pct/pct.h.

The library is built in-tree (pseudoctimaging_amd/lib/libpct.so) by
``pseudoctimaging_amd.build.build()``.  Loading fails loudly when it is
missing: there is no CPU fallback for the fusion path.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, c_char_p, c_double, c_float, c_int,
                    c_longlong, c_size_t, c_uint64, c_void_p)

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libpct.so")

PCT_OK = 0
PCT_ERR_INVALID = -1
PCT_ERR_NO_DEVICE = -2
PCT_ERR_HIP = -3
PCT_ERR_NO_MEMORY = -4
PCT_ERR_IO = -5
PCT_ERR_FORMAT = -6
PCT_ERR_DIMS = -7
PCT_ERR_UNSUPPORTED = -8
PCT_ERR_TOO_LARGE = -9

CHOICE_GM = 0
CHOICE_GM2 = 1
CHOICE_SM = 2
CHOICE_DIRECT = 5   # the one-lane-per-voxel kernel (no reference counterpart)
CHOICE_GPU_TEST_TIME = 4


class Problem(Structure):
    """pct_problem"""
    _fields_ = [("nx", c_int), ("ny", c_int), ("nz", c_int), ("n_atlas", c_int),
                ("nh", c_int), ("patch", c_int), ("sigma", c_double), ("beta", c_double),
                ("div", c_double)]


class TileConfig(Structure):
    """pct_tile_config"""
    _fields_ = [("id", c_int), ("seg_width", c_int), ("bx", c_int), ("bz", c_int),
                ("waves", c_int), ("atlas_group", c_int), ("lds_bytes", c_int),
                ("items", c_longlong), ("est_cost", c_double), ("measured_ms", c_double),
                ("form", c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Subject(Structure):
    """pct_subject"""
    _fields_ = [("pb", Problem), ("d_img", c_void_p), ("d_atlas", c_void_p),
                ("d_atlas_label_norm", c_void_p), ("d_out", c_void_p)]


PCT_MAX_BATCH = 16
PCT_BATCH_CLASSES = 8


class BatchChoice(Structure):
    """pct_batch_choice: tile shape and atlas group per half window (NH-1)/2"""
    _fields_ = [("shape_id", c_int * PCT_BATCH_CLASSES), ("atlas_group", c_int * PCT_BATCH_CLASSES),
                ("serial", c_int)]


class VolumeSet(Structure):
    """pct_volume_set"""
    _fields_ = [("nx", c_int), ("ny", c_int), ("nz", c_int), ("n_atlas", c_int),
                ("img", POINTER(c_float)), ("atlas", POINTER(c_float)),
                ("label", POINTER(c_float))]


class PrunedOpts(Structure):
    """pct_pruned_opts"""
    _fields_ = [("topk", c_int), ("samples", c_int), ("margin", c_float), ("shape_id", c_int),
                ("max_slabs", c_int), ("embed_all", c_int), ("max_err", c_float)]


class TrainOpts(Structure):
    """pct_train_opts"""
    _fields_ = [("steps", c_int), ("anchors", c_int), ("lr", c_float), ("tau", c_float),
                ("teacher_t", c_float), ("seed", c_uint64)]


PCT_MLP_NPARAMS = 7314

FP = POINTER(c_float)

# name -> (restype, argtypes); every function declared in include/pct/pct.h
SIGNATURES = {
    "pct_strerror": (c_char_p, [c_int]),
    "pct_abi_version": (c_int, []),
    "pct_weight_divisor": (c_double, [c_double, c_double]),
    "pct_normalize_labels": (None, [FP, c_size_t]),
    "pct_normalize_labels_dev": (c_int, [c_void_p, c_size_t, c_void_p]),
    "pct_fuse_host": (c_int, [POINTER(Problem), FP, FP, FP, FP, c_int]),
    "pct_pseudo_ct_host": (c_int, [POINTER(Problem), FP, FP, FP, FP, c_int]),
    "pct_fuse_dev": (c_int, [POINTER(Problem), c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                             c_void_p]),
    "pct_fuse_workspace_size": (c_size_t, [POINTER(Problem), c_int]),
    "pct_fuse_dev_noalloc": (c_int, [POINTER(Problem), c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_int, c_void_p]),
    "pct_fuse_dev_ws": (c_int, [POINTER(Problem), c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                c_void_p, c_size_t, c_void_p]),
    "pct_tile_config_count": (c_int, []),
    "pct_tile_config_choose": (c_int, [POINTER(Problem), POINTER(TileConfig)]),
    "pct_autotune": (c_int, [POINTER(Problem), c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                             c_void_p, POINTER(TileConfig)]),
    "pct_tile_config_force": (c_int, [c_int, c_int]),
    "pct_fuse_batch_dev": (c_int, [POINTER(Subject), c_int, c_void_p]),
    "pct_fuse_batch_workspace_size": (c_size_t, [POINTER(Subject), c_int]),
    "pct_fuse_batch_dev_ws": (c_int, [POINTER(Subject), c_int, c_void_p, c_size_t, c_void_p]),
    "pct_autotune_batch": (c_int, [POINTER(Subject), c_int, c_int, c_void_p,
                                   POINTER(TileConfig)]),
    "pct_fuse_batch_config": (c_int, [POINTER(Subject), c_int, c_int, c_int]),
    "pct_fuse_batch_choice_get": (c_int, [POINTER(Subject), c_int, POINTER(BatchChoice)]),
    "pct_fuse_batch_choice_set": (c_int, [POINTER(Subject), c_int, POINTER(BatchChoice)]),
    "pct_mlp_init": (c_int, [FP, c_uint64, c_float, c_float]),
    "pct_embed_dev": (c_int, [FP, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p]),
    "pct_fuse_pruned_dev": (c_int, [POINTER(Problem), FP, POINTER(PrunedOpts), c_void_p, c_void_p,
                                    c_void_p, c_void_p, POINTER(c_double), c_void_p]),
    "pct_mlp_loss_grad_dev": (c_int, [POINTER(Problem), FP, POINTER(c_int), c_int, c_float,
                                      c_float, c_void_p, c_void_p, POINTER(c_double), FP,
                                      c_void_p]),
    "pct_mlp_train_dev": (c_int, [POINTER(Problem), FP, POINTER(TrainOpts), c_void_p, c_void_p,
                                  FP, c_void_p]),
    "pct_kernel_timing": (c_int, [c_int]),
    "pct_kernel_timing_read": (c_int, [POINTER(c_double), POINTER(c_int)]),
    "pct_regularize": (c_int, [c_int, c_int, c_int, c_int, FP]),
    "pct_round": (None, [FP, c_size_t]),
    "pct_to_hu": (None, [FP, c_size_t]),
    "pct_postprocess": (c_int, [c_int, c_int, c_int, c_int, FP]),
    "pct_regularize_dev": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_void_p]),
    "pct_postprocess_dev": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_void_p]),
    "pct_matrixio_read": (c_int, [c_char_p, POINTER(FP), POINTER(c_int)]),
    "pct_matrixio_write": (c_int, [c_char_p, FP, c_int, c_int, c_int, c_int]),
    "pct_matrixio_write_sidecar": (c_int, [c_char_p]),
    "pct_read_config": (c_int, [c_char_p, c_uint64, POINTER(VolumeSet)]),
    "pct_volume_set_free": (None, [POINTER(VolumeSet)]),
    "pct_mat_read": (c_int, [c_char_p, POINTER(VolumeSet)]),
    "pct_mat_write": (c_int, [c_char_p, c_char_p, FP, c_int, c_int, c_int]),
    "pct_free": (None, [c_void_p]),
    "pct_synth_uniform": (c_int, [c_int, c_int, c_int, c_int, c_uint64, FP, FP, FP]),
    "pct_synth_tissue": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, c_uint64, FP, FP,
                                 FP]),
    "pct_device_count": (c_int, []),
    "pct_set_device": (c_int, [c_int]),
    "pct_release": (None, []),
}

_lib = None


class PctError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        msg = strerror(status)
        super().__init__(f"{what}: {msg} ({status})" if what else f"{msg} ({status})")


def lib() -> ctypes.CDLL:
    """The loaded libpct.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with pseudoctimaging_amd.build.build() "
                "(the fusion path has no CPU fallback)")
        # torch ships its own libamdhip64.so.7.  Loaded after libpct.so, it
        # would bind to the runtime libpct.so brought in (same soname) and
        # fail to initialise; loaded first, both share torch's runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def strerror(status: int) -> str:
    try:
        return lib().pct_strerror(status).decode()
    except ImportError:
        return f"status {status}"


def check(status: int, what: str = "") -> None:
    if status != PCT_OK:
        raise PctError(status, what)


def fptr(a) -> ctypes.POINTER(c_float):
    """float* of a C-contiguous float32 numpy array."""
    return a.ctypes.data_as(FP)
