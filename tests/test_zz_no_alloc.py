"""pct_fuse_dev_noalloc, the entry point that allocates nothing: its output
against pct_fuse_dev's on two calls in a row, the second smaller (the scratch
array keeps the first call's contents), and its size limit."""
import ctypes
from ctypes import byref

import numpy as np
import pytest

import pseudoctimaging_amd as pct
from pseudoctimaging_amd import _lib
from parity import compare

pytestmark = pytest.mark.gpu


def _case(nx, ny, nz, n, seed):
    img, amri, act = pct.synth_tissue(nx, ny, nz, n, 2, False, seed)
    lab = act.copy()
    pct.normalize_labels(lab)
    return img, amri, lab


def _noalloc(d_img, d_amri, d_lab, d_out, nh, choice, stream):
    pb = pct.problem(d_img.shape, d_amri.shape[0], nh, 2.5, 1.0)
    return _lib.lib().pct_fuse_dev_noalloc(
        byref(pb), ctypes.c_void_p(d_img.data_ptr()), ctypes.c_void_p(d_amri.data_ptr()),
        ctypes.c_void_p(d_lab.data_ptr()), ctypes.c_void_p(d_out.data_ptr()), int(choice),
        ctypes.c_void_p(stream.cuda_stream))


def test_noalloc_matches_device_call_twice(gpu):
    import torch
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    # the second call is smaller: voxels the first call left in the scratch
    # array must not reach its output
    for shape, n, nh, seed in (((40, 64, 60), 3, 11, 31), ((22, 40, 36), 2, 7, 32)):
        nz, nx, ny = shape
        print(f"case noalloc {shape} n={n} NH={nh}", flush=True)
        img, amri, lab = _case(nx, ny, nz, n, seed)
        d_img, d_amri, d_lab = (torch.from_numpy(x).to(dev) for x in (img, amri, lab))
        ref = torch.full(img.shape, -7.0, device=dev)
        pct.fuse_device(d_img, d_amri, d_lab, ref, nh, 2.5, 1.0, stream=stream)
        out = torch.full(img.shape, -9.0, device=dev)
        assert _noalloc(d_img, d_amri, d_lab, out, nh, pct.CHOICE_SM, stream) == 0
        torch.cuda.synchronize()
        a, b = out.cpu().numpy(), ref.cpu().numpy()
        st = compare(a, b)
        r = (nh - 1) // 2 + 1
        border = np.ones(img.shape, bool)
        border[r:nz - r, r:nx - r, r:ny - r] = False
        assert not a[border].any()
        print(st)


def test_noalloc_direct_choice_and_size_limit(gpu):
    import torch
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    img, amri, lab = _case(24, 30, 20, 2, 33)
    d_img, d_amri, d_lab = (torch.from_numpy(x).to(dev) for x in (img, amri, lab))
    ref = torch.zeros(img.shape, device=dev)
    pct.fuse_device(d_img, d_amri, d_lab, ref, 5, 2.5, 1.0, pct.CHOICE_DIRECT, stream)
    out = torch.full(img.shape, -1.0, device=dev)
    assert _noalloc(d_img, d_amri, d_lab, out, 5, pct.CHOICE_DIRECT, stream) == 0
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().view(np.uint32), ref.cpu().numpy().view(np.uint32))
    # beyond the 64 MiB of partial sums (8 388 608 voxels): refused before
    # any launch, so the small buffers passed here are never touched
    pb = pct.problem((130, 256, 256), 1, 7, 2.5, 1.0)
    rc = _lib.lib().pct_fuse_dev_noalloc(
        byref(pb), ctypes.c_void_p(d_img.data_ptr()), ctypes.c_void_p(d_amri.data_ptr()),
        ctypes.c_void_p(d_lab.data_ptr()), ctypes.c_void_p(out.data_ptr()), pct.CHOICE_SM,
        ctypes.c_void_p(stream.cuda_stream))
    assert rc == _lib.PCT_ERR_TOO_LARGE
