"""Python host interface over libpct.so, mirroring the reference program's
operator surface for the fusion path.

Arrays follow the reference layout: a volume is a C-contiguous float32 array
of shape (nz, nx, ny) (linear index z*nx*ny + x*ny + y,
src/CPU/NewPatchedBrainPureC.cpp:1300) and atlases stack on a leading axis,
shape (n, nz, nx, ny) (cpp:1288).

Reference entry point -> this module:
  MultiPatchBasedSegmentationGpu (src/GPU/PatchBasedPseudoCT.cu:1635)
      -> multi_patch_based_segmentation_gpu (host arrays) / fuse_device (HBM)
  MultiAtlasPatchBasedSegmentationOpenMP (cpp:1350)
      -> the same function; the fusion has no CPU implementation here
  label normalisation (cpp:1583-1593)   -> normalize_labels
  div = 2*27*beta*sigma^2 (cpp:1601)    -> weight_divisor
  RegularizationStep / RoundMatrix / HU rescale (cpp:1059-1106, 1221-1227,
  1725-1739)                            -> regularization_step, round_matrix,
                                           to_hu, postprocess
  MatrixIO<float>::Read / Store (cpp:252-320) -> matrixio_read / matrixio_write
  readMatFile (cpp:908-1018)            -> read_config
  main() (cpp:1502-1781, cu:1898-2239)  -> pseudo_ct (and the pct_cli program)
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_int

import numpy as np

from . import _lib
from ._lib import (CHOICE_DIRECT, CHOICE_GM, CHOICE_GM2, CHOICE_SM, PctError, Problem,  # noqa: F401
                   TileConfig, check, fptr, lib)


def _f32(a, name: str) -> np.ndarray:
    a = np.asarray(a)
    if a.dtype != np.float32:
        raise TypeError(f"{name} must be float32, got {a.dtype}")
    if not a.flags.c_contiguous:
        raise ValueError(f"{name} must be C-contiguous")
    return a


def problem(shape, n_atlas: int, nh: int, sigma: float = 1.0, beta: float = 1.0,
            div: float = 0.0) -> Problem:
    """pct_problem for a volume of shape (nz, nx, ny).  div > 0 gives the
    weight divisor itself (the `div` argument of the reference's wrappers,
    cu:1635 / cpp:1350), and sigma and beta are then not used."""
    nz, nx, ny = (int(s) for s in shape)
    return Problem(nx=nx, ny=ny, nz=nz, n_atlas=int(n_atlas), nh=int(nh), patch=3,
                   sigma=float(sigma), beta=float(beta), div=float(div))


def weight_divisor(sigma: float, beta: float) -> float:
    """div = 2*27*beta*sigma^2 (cpp:1601)."""
    return lib().pct_weight_divisor(float(sigma), float(beta))


def normalize_labels(label: np.ndarray) -> np.ndarray:
    """In place: L' = 255*((L+1024)/4095) in float32 (cpp:1583-1593)."""
    label = _f32(label, "label")
    lib().pct_normalize_labels(fptr(label), label.size)
    return label


def release() -> None:
    """pct_release: frees the library's per-device workspaces, cached host-call
    buffers (device and pinned), streams, batch graphs and model copies.  The
    next call recreates what it needs."""
    lib().pct_release()


def device_count() -> int:
    return lib().pct_device_count()


def multi_patch_based_segmentation_gpu(img, atlas, atlas_label, nh: int, sigma: float = 1.0,
                                       beta: float = 1.0, choice: int = CHOICE_SM,
                                       div: float = 0.0) -> np.ndarray:
    """Fused value of every voxel (interior: P/S or NaN, border: 0).

    `atlas_label` is the normalised CT (see normalize_labels).  Host arrays in
    and out, as the reference wrapper (cu:1687-1744); the library keeps its
    device buffers for the next call.  div > 0 is the reference's `div`
    argument (cu:1635) and replaces sigma and beta.
    """
    img = _f32(img, "img")
    atlas = _f32(atlas, "atlas")
    atlas_label = _f32(atlas_label, "atlas_label")
    if img.ndim != 3 or atlas.ndim != 4 or atlas.shape != atlas_label.shape \
            or atlas.shape[1:] != img.shape:
        raise ValueError("expected img (nz,nx,ny) and atlas/atlas_label (n,nz,nx,ny)")
    pb = problem(img.shape, atlas.shape[0], nh, sigma, beta, div)
    out = np.zeros(img.shape, np.float32)
    check(lib().pct_fuse_host(byref(pb), fptr(img), fptr(atlas), fptr(atlas_label), fptr(out),
                              int(choice)), "pct_fuse_host")
    return out


def pseudo_ct_hu(img, atlas, atlas_label, nh: int, sigma: float = 1.0, beta: float = 1.0,
                 choice: int = CHOICE_SM) -> np.ndarray:
    """Final HU volume of the reference GPU program's SM block (cu:2152-2179):
    fusion, NaN fill, rounding and HU rescale on the device, host arrays in
    and out.  `atlas_label` is the normalised CT."""
    img = _f32(img, "img")
    atlas = _f32(atlas, "atlas")
    atlas_label = _f32(atlas_label, "atlas_label")
    if img.ndim != 3 or atlas.ndim != 4 or atlas.shape != atlas_label.shape \
            or atlas.shape[1:] != img.shape:
        raise ValueError("expected img (nz,nx,ny) and atlas/atlas_label (n,nz,nx,ny)")
    pb = problem(img.shape, atlas.shape[0], nh, sigma, beta)
    out = np.zeros(img.shape, np.float32)
    check(lib().pct_pseudo_ct_host(byref(pb), fptr(img), fptr(atlas), fptr(atlas_label),
                                   fptr(out), int(choice)), "pct_pseudo_ct_host")
    return out


def _dev_ptr(t, name: str):
    import torch
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise TypeError(f"{name} must be a device tensor")
    if t.dtype != torch.float32 or not t.is_contiguous():
        raise TypeError(f"{name} must be contiguous float32")
    return ctypes.c_void_p(t.data_ptr())


def check_volume_set(img, atlas, atlas_label, out=None, stream=None) -> None:
    """Checks, before anything is launched, that device tensors form one
    problem: img (nz, nx, ny), atlas and atlas_label (n, nz, nx, ny), out
    shaped as img, every tensor on one device and the stream on it too.  The
    kernels take their sizes from img and atlas alone, so a smaller label or
    out tensor would otherwise be read or written past its end."""
    if getattr(img, "ndim", None) != 3:
        raise ValueError("img must be (nz, nx, ny)")
    if getattr(atlas, "ndim", None) != 4 or tuple(atlas.shape[1:]) != tuple(img.shape):
        raise ValueError("atlas must be (n, nz, nx, ny) with (nz, nx, ny) = img.shape")
    if atlas_label is not None and tuple(atlas_label.shape) != tuple(atlas.shape):
        raise ValueError("atlas_label must have the shape of atlas")
    if out is not None and tuple(out.shape) != tuple(img.shape):
        raise ValueError("out must have the shape of img")
    tensors = [t for t in (img, atlas, atlas_label, out) if t is not None]
    devices = {str(getattr(t, "device", "host")) for t in tensors}
    if len(devices) != 1:
        raise ValueError(f"tensors on different devices: {sorted(devices)}")
    if stream is not None and getattr(stream, "device", None) is not None \
            and str(stream.device) != str(img.device):
        raise ValueError(f"stream on {stream.device}, tensors on {img.device}")


def fuse_device(img, atlas, atlas_label, out, nh: int, sigma: float = 1.0, beta: float = 1.0,
                choice: int = CHOICE_SM, stream=None, div: float = 0.0) -> None:
    """Fusion on tensors already resident in HBM (torch only carries memory
    and the stream).  Enqueued on `stream` (default: torch's current one).
    div > 0 replaces sigma and beta, as in multi_patch_based_segmentation_gpu."""
    import torch
    check_volume_set(img, atlas, atlas_label, out, stream)
    if stream is None:
        stream = torch.cuda.current_stream(img.device)
    pb = problem(img.shape, atlas.shape[0], nh, sigma, beta, div)
    check(lib().pct_fuse_dev(byref(pb), _dev_ptr(img, "img"), _dev_ptr(atlas, "atlas"),
                             _dev_ptr(atlas_label, "atlas_label"), _dev_ptr(out, "out"),
                             int(choice), ctypes.c_void_p(stream.cuda_stream)), "pct_fuse_dev")


def normalize_labels_device(label, stream=None) -> None:
    import torch
    if stream is None:
        stream = torch.cuda.current_stream(label.device)
    check(lib().pct_normalize_labels_dev(_dev_ptr(label, "label"), label.numel(),
                                         ctypes.c_void_p(stream.cuda_stream)),
          "pct_normalize_labels_dev")


def tile_config(shape, n_atlas: int, nh: int) -> dict:
    """The tiled-kernel configuration the library would use (no GPU work)."""
    pb = problem(shape, n_atlas, nh)
    cfg = TileConfig()
    check(lib().pct_tile_config_choose(byref(pb), byref(cfg)), "pct_tile_config_choose")
    return cfg.as_dict()


def autotune(img, atlas, atlas_label, out, nh: int, sigma: float = 1.0, beta: float = 1.0,
             reps: int = 2, stream=None) -> dict:
    import torch
    check_volume_set(img, atlas, atlas_label, out, stream)
    if stream is None:
        stream = torch.cuda.current_stream(img.device)
    pb = problem(img.shape, atlas.shape[0], nh, sigma, beta)
    cfg = TileConfig()
    check(lib().pct_autotune(byref(pb), _dev_ptr(img, "img"), _dev_ptr(atlas, "atlas"),
                             _dev_ptr(atlas_label, "atlas_label"), _dev_ptr(out, "out"),
                             int(reps), ctypes.c_void_p(stream.cuda_stream), byref(cfg)),
          "pct_autotune")
    return cfg.as_dict()


def force_tile_config(config_id: int, atlas_group: int = 0) -> None:
    check(lib().pct_tile_config_force(int(config_id), int(atlas_group)), "pct_tile_config_force")


# ---- batches of subjects -------------------------------------------------------

def _subjects(subjects):
    """ctypes array of pct_subject from dicts with device tensors img, atlas,
    label (normalised CT), out and the window parameters nh, sigma, beta."""
    devices = set()
    for s in subjects:
        check_volume_set(s["img"], s["atlas"], s["label"], s["out"])
        devices.add(str(s["img"].device))
    if len(devices) > 1:
        raise ValueError(f"subjects on different devices: {sorted(devices)}")
    arr = (_lib.Subject * len(subjects))()
    for i, s in enumerate(subjects):
        arr[i].pb = problem(s["img"].shape, s["atlas"].shape[0], s["nh"], s.get("sigma", 1.0),
                            s.get("beta", 1.0))
        arr[i].d_img = _dev_ptr(s["img"], "img")
        arr[i].d_atlas = _dev_ptr(s["atlas"], "atlas")
        arr[i].d_atlas_label_norm = _dev_ptr(s["label"], "label")
        arr[i].d_out = _dev_ptr(s["out"], "out")
    return arr


def fuse_batch(subjects, stream=None) -> None:
    """Fuses every subject (list of dicts, see _subjects) in one graph-captured
    persistent launch; asynchronous on `stream`."""
    import torch
    arr = _subjects(subjects)
    if stream is None:
        stream = torch.cuda.current_stream(subjects[0]["img"].device)
    check_volume_set(subjects[0]["img"], subjects[0]["atlas"], None, None, stream)
    check(lib().pct_fuse_batch_dev(arr, len(subjects), ctypes.c_void_p(stream.cuda_stream)),
          "pct_fuse_batch_dev")


def autotune_batch(subjects, reps: int = 1, stream=None) -> dict:
    import torch
    arr = _subjects(subjects)
    if stream is None:
        stream = torch.cuda.current_stream(subjects[0]["img"].device)
    check_volume_set(subjects[0]["img"], subjects[0]["atlas"], None, None, stream)
    cfg = TileConfig()
    check(lib().pct_autotune_batch(arr, len(subjects), int(reps),
                                   ctypes.c_void_p(stream.cuda_stream), byref(cfg)),
          "pct_autotune_batch")
    return cfg.as_dict()


def set_batch_config(subjects, config_id: int, atlas_group: int) -> None:
    arr = _subjects(subjects)
    check(lib().pct_fuse_batch_config(arr, len(subjects), int(config_id), int(atlas_group)),
          "pct_fuse_batch_config")


def batch_choice(subjects) -> dict:
    """The configuration fuse_batch uses for these subjects:
    {"classes": {half_window: (shape_id, atlas_group)}, "serial": bool}."""
    arr = _subjects(subjects)
    ch = _lib.BatchChoice()
    check(lib().pct_fuse_batch_choice_get(arr, len(subjects), byref(ch)),
          "pct_fuse_batch_choice_get")
    classes = {hw: (ch.shape_id[hw], ch.atlas_group[hw])
               for hw in range(1, _lib.PCT_BATCH_CLASSES) if ch.shape_id[hw] >= 0}
    return {"classes": classes, "serial": bool(ch.serial)}


def set_batch_choice(subjects, classes: dict, serial: bool = False) -> None:
    """Sets the tile shape and atlas group per half window (NH-1)/2."""
    arr = _subjects(subjects)
    ch = _lib.BatchChoice()
    for hw in range(_lib.PCT_BATCH_CLASSES):
        ch.shape_id[hw], ch.atlas_group[hw] = classes.get(hw, (-1, 1))
    ch.serial = int(bool(serial))
    check(lib().pct_fuse_batch_choice_set(arr, len(subjects), byref(ch)),
          "pct_fuse_batch_choice_set")


# ---- post-processing --------------------------------------------------------

def regularization_step(arr: np.ndarray, nh: int) -> int:
    """In-place NaN median fill (cpp:1059-1106); returns the number of passes."""
    arr = _f32(arr, "arr")
    nz, nx, ny = arr.shape
    rc = lib().pct_regularize(int(nh), nx, ny, nz, fptr(arr))
    if rc < 0:
        check(rc, "pct_regularize")
    return rc


def round_matrix(arr: np.ndarray) -> np.ndarray:
    arr = _f32(arr, "arr")
    lib().pct_round(fptr(arr), arr.size)
    return arr


def to_hu(arr: np.ndarray) -> np.ndarray:
    arr = _f32(arr, "arr")
    lib().pct_to_hu(fptr(arr), arr.size)
    return arr


def postprocess(arr: np.ndarray, nh: int) -> np.ndarray:
    """The chain of cpp:1725-1739 in place: NaN fill, round, back to HU."""
    arr = _f32(arr, "arr")
    nz, nx, ny = arr.shape
    check(lib().pct_postprocess(int(nh), nx, ny, nz, fptr(arr)), "pct_postprocess")
    return arr


def regularization_step_device(arr, nh: int, stream=None) -> None:
    """pct_regularize_dev on a (nz, nx, ny) float32 device tensor."""
    import torch
    if stream is None:
        stream = torch.cuda.current_stream(arr.device)
    nz, nx, ny = arr.shape
    check(lib().pct_regularize_dev(int(nh), nx, ny, nz, _dev_ptr(arr, "arr"),
                                   ctypes.c_void_p(stream.cuda_stream)), "pct_regularize_dev")


def postprocess_device(arr, nh: int, stream=None) -> None:
    """pct_postprocess_dev (NaN fill, round, HU) on a device tensor, in place."""
    import torch
    if stream is None:
        stream = torch.cuda.current_stream(arr.device)
    nz, nx, ny = arr.shape
    check(lib().pct_postprocess_dev(int(nh), nx, ny, nz, _dev_ptr(arr, "arr"),
                                    ctypes.c_void_p(stream.cuda_stream)), "pct_postprocess_dev")


def pseudo_ct(img, atlas_mri, atlas_ct_hu, nh: int, sigma: float = 1.0, beta: float = 1.0,
              choice: int = CHOICE_SM):
    """main()'s pipeline (cpp:1583-1739): normalise the atlas CT, fuse,
    post-process.  Returns (fused value before rounding, final HU volume)."""
    label = np.array(atlas_ct_hu, dtype=np.float32, copy=True, order="C")
    normalize_labels(label)
    fused = multi_patch_based_segmentation_gpu(img, atlas_mri, label, nh, sigma, beta, choice)
    hu = postprocess(fused.copy(), nh)
    return fused, hu


def fused_to_hu_raw(fused: np.ndarray) -> np.ndarray:
    """HU of the fused value before any rounding: 4095*v/255 - 1024 (float64)."""
    return 4095.0 * (np.asarray(fused, np.float64) / 255.0) - 1024.0


# ---- files and inputs -------------------------------------------------------

def matrixio_read(path: str):
    """(array of shape (num, depth, rows, cols) squeezed to the reference's
    memory order, dims) where dims = [rows, cols, depth, num] (cpp:252-287)."""
    data = _lib.FP()
    dims = (c_int * 4)()
    check(lib().pct_matrixio_read(path.encode(), byref(data), dims), f"read {path}")
    rows, cols, depth, num = list(dims)
    n = rows * cols * depth * num
    arr = np.ctypeslib.as_array(data, shape=(n,)).copy()
    lib().pct_free(ctypes.cast(data, ctypes.c_void_p))
    return arr, [rows, cols, depth, num]


def matrixio_write(path: str, arr, num: int, rows: int, cols: int, depth: int) -> None:
    arr = _f32(np.ascontiguousarray(arr), "arr")
    if arr.size != num * rows * cols * depth:
        raise ValueError("size does not match dimensions")
    check(lib().pct_matrixio_write(path.encode(), fptr(arr), num, rows, cols, depth),
          f"write {path}")


def matrixio_write_sidecar(path: str) -> None:
    """Writes the binary sidecar <path>.pctbin that matrixio_read prefers
    while the text file is unchanged."""
    check(lib().pct_matrixio_write_sidecar(path.encode()), f"sidecar {path}")


def read_config(path: str, seed: int = 0) -> dict:
    """readMatFile (cpp:908-1018): returns img (nz,nx,ny), atlas and label
    (n,nz,nx,ny) with the label still in HU.  A path ending in .mat is read
    as a MAT-file (mat_read)."""
    vs = _lib.VolumeSet()
    check(lib().pct_read_config(path.encode(), int(seed), byref(vs)), f"read_config {path}")
    return _take_volume_set(vs)


def mat_read(path: str) -> dict:
    """ReadMatFile of the MATLAB build (cpp:584-704) without libmat: the
    structs atlas / atlas_label and the target image of a MAT-file, in the
    library layout (img (nz,nx,ny), atlas and label (n,nz,nx,ny))."""
    vs = _lib.VolumeSet()
    check(lib().pct_mat_read(path.encode(), byref(vs)), f"mat_read {path}")
    return _take_volume_set(vs)


def mat_write(path: str, name: str, arr) -> None:
    """WriteOutputMatFile (cpp:848-906): a (nz,nx,ny) float32 volume as the
    nx x ny x nz single-precision MATLAB array `name`."""
    arr = _f32(arr, "arr")
    nz, nx, ny = arr.shape
    check(lib().pct_mat_write(path.encode(), name.encode(), fptr(arr), nx, ny, nz),
          f"mat_write {path}")


def _take_volume_set(vs) -> dict:
    try:
        nz, nx, ny, n = vs.nz, vs.nx, vs.ny, vs.n_atlas
        N = nx * ny * nz
        img = np.ctypeslib.as_array(vs.img, shape=(N,)).copy().reshape(nz, nx, ny)
        atlas = np.ctypeslib.as_array(vs.atlas, shape=(N * n,)).copy().reshape(n, nz, nx, ny)
        label = np.ctypeslib.as_array(vs.label, shape=(N * n,)).copy().reshape(n, nz, nx, ny)
    finally:
        lib().pct_volume_set_free(byref(vs))
    return {"img": img, "atlas": atlas, "label": label}


def synth_tissue(nx: int, ny: int, nz: int, n_atlas: int, max_shift: int,
                 independent_anatomy: bool = False, seed: int = 0):
    """Seeded tissue-class phantom: (target MRI, atlas MRI, atlas CT in HU)."""
    img = np.empty((nz, nx, ny), np.float32)
    amri = np.empty((n_atlas, nz, nx, ny), np.float32)
    act = np.empty((n_atlas, nz, nx, ny), np.float32)
    check(lib().pct_synth_tissue(nx, ny, nz, n_atlas, max_shift, int(bool(independent_anatomy)),
                                 int(seed), fptr(img), fptr(amri), fptr(act)), "synth_tissue")
    return img, amri, act


def synth_uniform(nx: int, ny: int, nz: int, n_atlas: int, seed: int = 0):
    """Seeded data of the reference's random mode: every value 10*U[0,1)."""
    img = np.empty((nz, nx, ny), np.float32)
    atlas = np.empty((n_atlas, nz, nx, ny), np.float32)
    label = np.empty((n_atlas, nz, nx, ny), np.float32)
    check(lib().pct_synth_uniform(nx, ny, nz, n_atlas, int(seed), fptr(img), fptr(atlas),
                                  fptr(label)), "synth_uniform")
    return img, atlas, label
