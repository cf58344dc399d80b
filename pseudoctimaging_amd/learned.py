"""Learned candidate proposal: host interface over the embedding MLP, its
GPU training and the pruned fusion in libpct.so.

No counterpart in the reference program, whose fusion weighs every one of
the n*NH^3 candidates of a voxel (src/CPU/NewPatchedBrainPureC.cpp:1399-1432,
src/GPU/PatchBasedPseudoCT.cu:1635).  The pruned fusion keeps, per tile of
target voxels and per atlas, the `topk` window offsets whose embeddings are
closest to the target's, weighs them exactly in FP32, estimates the rest from
a lattice of exactly weighted samples, and sends tiles with an ambiguous
selection through the full search.

Arrays follow the layout of `api`: volumes (nz, nx, ny), atlases (n, nz, nx,
ny), float32, on the GPU as torch tensors (torch carries memory and streams).
"""
from __future__ import annotations

import ctypes
from ctypes import byref

import numpy as np

from ._lib import (PCT_MLP_NPARAMS, PrunedOpts, TrainOpts, check, fptr, lib)
from .api import _dev_ptr, _f32, check_volume_set, problem

# parameter vector layout (input-major weights)
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3, OFF_MU, OFF_SD = (
    0, 2048, 2112, 6208, 6272, 7296, 7312, 7313)
N_FEAT, N_HID, N_EMB = 30, 64, 16


def _stream(t, stream):
    import torch
    return stream if stream is not None else torch.cuda.current_stream(t.device)


def mlp_init(seed: int, mu: float, inv_sd: float) -> np.ndarray:
    """He-initialised parameter vector (float32, PCT_MLP_NPARAMS values)."""
    p = np.zeros(PCT_MLP_NPARAMS, np.float32)
    check(lib().pct_mlp_init(fptr(p), int(seed), float(mu), float(inv_sd)), "pct_mlp_init")
    return p


def intensity_stats(img) -> tuple[float, float]:
    """mu and inv_sd of a target volume (the features' normalisation)."""
    a = np.asarray(img, np.float64)
    sd = float(a.std())
    return float(a.mean()), (1.0 / sd if sd > 0 else 1.0)


def unpack(params: np.ndarray) -> dict:
    """Named views of a parameter vector."""
    p = np.asarray(params)
    return {"W1": p[OFF_W1:OFF_B1].reshape(32, N_HID), "b1": p[OFF_B1:OFF_W2],
            "W2": p[OFF_W2:OFF_B2].reshape(N_HID, N_HID), "b2": p[OFF_B2:OFF_W3],
            "W3": p[OFF_W3:OFF_B3].reshape(N_HID, N_EMB), "b3": p[OFF_B3:OFF_MU],
            "mu": p[OFF_MU], "inv_sd": p[OFF_SD]}


def embed_device(params: np.ndarray, vol, out=None, stream=None):
    """BF16 embeddings (raw 16-bit words in an int16 tensor of shape
    (nz, nx, ny, 16)) of a device volume."""
    import torch
    params = _f32(params, "params")
    nz, nx, ny = vol.shape
    if out is None:
        out = torch.empty((nz, nx, ny, N_EMB), dtype=torch.int16, device=vol.device)
    if out.dtype != torch.int16 or not out.is_contiguous() or out.numel() != vol.numel() * N_EMB:
        raise ValueError("out must be a contiguous int16 tensor of shape (nz, nx, ny, 16)")
    s = _stream(vol, stream)
    check(lib().pct_embed_dev(fptr(params), _dev_ptr(vol, "vol"), nx, ny, nz,
                              ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(s.cuda_stream)),
          "pct_embed_dev")
    return out


def bf16_words_to_float(words) -> np.ndarray:
    """float32 values of raw BF16 words (numpy int16/uint16 array)."""
    w = np.asarray(words).astype(np.uint16).astype(np.uint32) << 16
    return w.view(np.float32)


# Error bar of the pruned fusion's fallback, in label units (16 HU): at
# 222x222x112, n=10, NH=11 it flags no tile at sigma 2.5 (1.90 HU MAE against
# the exact fusion, 7.9x faster) and every tile at sigma 1 and 0.25, where the
# pruned sums would miss by 11.6 and 124 HU MAE.
DEFAULT_MAX_ERR = 1.0


def fuse_pruned_device(img, atlas, atlas_label, out, nh: int, sigma: float, beta: float,
                       params: np.ndarray, topk: int = 32, samples: int = 27,
                       margin: float = 0.0, shape_id: int = -1, max_slabs: int = 3,
                       embed_all: bool = False, stream=None, report_fallback: bool = True,
                       max_err: float = DEFAULT_MAX_ERR):
    """Learned-pruned fusion on device tensors.  Returns the fraction of
    (tile, atlas) items that took the full search (synchronises) when
    report_fallback, else None (asynchronous).  max_slabs bounds the dz
    planes the kept offsets come from (0: any).  The default 3 keeps them on
    the lattice samples' planes: at 222x222x112, n=10, NH=11 it measured
    faster than 4 (3.49 vs 3.80 ms) at the same error (1.90 HU MAE).
    max_err (label units) is the bar of the per-voxel error estimate above
    which a tile takes the full search; 0 switches the estimate off."""
    params = _f32(params, "params")
    check_volume_set(img, atlas, atlas_label, out, stream)
    s = _stream(img, stream)
    pb = problem(img.shape, atlas.shape[0], nh, sigma, beta)
    opts = PrunedOpts(topk=int(topk), samples=int(samples), margin=float(margin),
                      shape_id=int(shape_id), max_slabs=int(max_slabs),
                      embed_all=int(bool(embed_all)), max_err=float(max_err))
    frac = ctypes.c_double(-1.0)
    check(lib().pct_fuse_pruned_dev(byref(pb), fptr(params), byref(opts),
                                    _dev_ptr(img, "img"), _dev_ptr(atlas, "atlas"),
                                    _dev_ptr(atlas_label, "atlas_label"), _dev_ptr(out, "out"),
                                    byref(frac) if report_fallback else None,
                                    ctypes.c_void_p(s.cuda_stream)), "pct_fuse_pruned_dev")
    return frac.value if report_fallback else None


def mlp_loss_grad(img, atlas, nh: int, sigma: float, beta: float, params: np.ndarray,
                  anchors, tau: float = 1.0, teacher_t: float = 0.0, stream=None):
    """(loss, gradient) of the training objective at `params` for the given
    anchors ([n][4] = z, x, y, atlas), computed on the GPU.  teacher_t <= 0
    uses the fusion divisor as teacher temperature."""
    params = _f32(params, "params")
    anc = np.ascontiguousarray(np.asarray(anchors, np.int32).reshape(-1, 4))
    check_volume_set(img, atlas, None, None, stream)
    s = _stream(img, stream)
    pb = problem(img.shape, atlas.shape[0], nh, sigma, beta)
    loss = ctypes.c_double(0.0)
    grad = np.zeros(PCT_MLP_NPARAMS, np.float32)
    check(lib().pct_mlp_loss_grad_dev(byref(pb), fptr(params),
                                      anc.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                      int(anc.shape[0]), float(tau), float(teacher_t),
                                      _dev_ptr(img, "img"),
                                      _dev_ptr(atlas, "atlas"), byref(loss), fptr(grad),
                                      ctypes.c_void_p(s.cuda_stream)), "pct_mlp_loss_grad_dev")
    return loss.value, grad


def mlp_train(img, atlas, nh: int, sigma: float, beta: float, params: np.ndarray,
              steps: int = 200, anchors: int = 256, lr: float = 2e-3, tau: float = 1.0,
              teacher_t: float = 0.5, seed: int = 0, stream=None):
    """Trains a copy of `params` on the GPU; returns (params, loss per step).
    teacher_t is the temperature of the SSD ordering the embedding learns
    (<= 0: the fusion divisor)."""
    p = np.array(_f32(params, "params"), copy=True)
    check_volume_set(img, atlas, None, None, stream)
    s = _stream(img, stream)
    pb = problem(img.shape, atlas.shape[0], nh, sigma, beta)
    opts = TrainOpts(steps=int(steps), anchors=int(anchors), lr=float(lr), tau=float(tau),
                     teacher_t=float(teacher_t), seed=int(seed))
    trace = np.zeros(max(int(steps), 1), np.float32)
    check(lib().pct_mlp_train_dev(byref(pb), fptr(p), byref(opts), _dev_ptr(img, "img"),
                                  _dev_ptr(atlas, "atlas"), fptr(trace),
                                  ctypes.c_void_p(s.cuda_stream)), "pct_mlp_train_dev")
    return p, trace[:int(steps)]
