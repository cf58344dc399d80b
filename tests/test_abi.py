"""The C ABI library loads on a machine without a GPU, exports every function
include/pct/pct.h declares, and validates its arguments before touching a
device."""
import ctypes
import os
import re

import pytest

import pseudoctimaging_amd as pct
from pseudoctimaging_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pct", "pct.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(pct_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declarations_are_all_exported_and_bound():
    names = declared_functions()
    assert len(names) >= 25
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), f"{n} declared in pct.h but not exported"
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_reference_and_oracle_libraries_load():
    import oracle
    oracle.oracle_lib()
    if oracle.ref_available():
        oracle.ref_lib()


def test_version_and_strerror():
    L = _lib.lib()
    assert L.pct_abi_version() == 2
    for code in range(0, -10, -1):
        assert L.pct_strerror(code)
    assert L.pct_strerror(-99) == b"unknown status"


def _pb(**kw):
    d = dict(nx=30, ny=30, nz=30, n_atlas=2, nh=7, patch=3, sigma=2.5, beta=1.0)
    d.update(kw)
    return _lib.Problem(**d)


@pytest.mark.parametrize("kw", [dict(nh=0), dict(patch=5), dict(sigma=0.0), dict(beta=-1.0),
                                dict(nx=0), dict(n_atlas=0), dict(sigma=float("nan"))])
def test_invalid_problems_rejected(kw):
    L = _lib.lib()
    pb = _pb(**kw)
    buf = (ctypes.c_float * 8)()
    rc = L.pct_fuse_host(ctypes.byref(pb), buf, buf, buf, buf, 2)
    assert rc == _lib.PCT_ERR_INVALID
    cfg = _lib.TileConfig()
    assert L.pct_tile_config_choose(ctypes.byref(pb), ctypes.byref(cfg)) == _lib.PCT_ERR_INVALID


def test_null_pointers_and_bad_choice_rejected():
    L = _lib.lib()
    pb = _pb()
    buf = (ctypes.c_float * 8)()
    assert L.pct_fuse_host(None, buf, buf, buf, buf, 2) == _lib.PCT_ERR_INVALID
    assert L.pct_fuse_host(ctypes.byref(pb), None, buf, buf, buf, 2) == _lib.PCT_ERR_INVALID
    assert L.pct_fuse_dev(ctypes.byref(pb), buf, buf, buf, buf, 3, None) == _lib.PCT_ERR_INVALID


def test_no_device_is_reported():
    if pct.device_count() > 0:
        pytest.skip("a GPU is present")
    L = _lib.lib()
    pb = _pb()
    buf = (ctypes.c_float * 8)()
    assert L.pct_fuse_host(ctypes.byref(pb), buf, buf, buf, buf, 2) == _lib.PCT_ERR_NO_DEVICE
    assert L.pct_set_device(0) == _lib.PCT_ERR_NO_DEVICE
    with pytest.raises(pct.PctError):
        pct.multi_patch_based_segmentation_gpu(*pct.synth_tissue(10, 10, 10, 1, 1, False, 0)[:2],
                                               pct.synth_tissue(10, 10, 10, 1, 1, False, 0)[2],
                                               3)


@pytest.mark.parametrize("nh", [3, 5, 7, 9, 11, 13, 15])
def test_tile_configuration_fits_the_chip(nh):
    cfg = pct.tile_config((112, 222, 222), 10, nh)
    assert 0 < cfg["lds_bytes"] <= 160 * 1024
    assert cfg["items"] > 0 and cfg["atlas_group"] >= 1
    L = _lib.lib()
    pb = _pb(nx=222, ny=222, nz=112, n_atlas=10, nh=nh)
    ws = L.pct_fuse_workspace_size(ctypes.byref(pb), 2)
    groups = (10 + cfg["atlas_group"] - 1) // cfg["atlas_group"]
    assert ws >= 2 * groups * 222 * 222 * 112 * 4 and ws % 256 == 0


def test_large_window_has_no_tiled_configuration():
    L = _lib.lib()
    pb = _pb(nh=17)
    cfg = _lib.TileConfig()
    assert L.pct_tile_config_choose(ctypes.byref(pb), ctypes.byref(cfg)) == \
        _lib.PCT_ERR_UNSUPPORTED
    # the direct kernel serves such windows: no tiled workspace needed
    assert L.pct_fuse_workspace_size(ctypes.byref(pb), 2) == 0


def test_forced_configuration_round_trip():
    try:
        pct.force_tile_config(3, 2)
        cfg = pct.tile_config((60, 70, 80), 4, 7)
        assert cfg["id"] == 3 and cfg["atlas_group"] == 2
    finally:
        pct.force_tile_config(-1, 0)
    with pytest.raises(pct.PctError):
        pct.force_tile_config(999, 1)


def test_device_wrappers_check_shapes_before_any_launch():
    # the kernels size every buffer from img and atlas: a label or output
    # tensor of another shape must be refused on the host, before a launch
    import torch

    from pseudoctimaging_amd import learned
    img = torch.zeros(6, 7, 8)
    atlas = torch.zeros(2, 6, 7, 8)
    small_label = torch.zeros(1, 6, 7, 8)
    small_out = torch.zeros(6, 7, 7)
    calls = [
        lambda: pct.fuse_device(img, atlas, small_label, torch.zeros(6, 7, 8), 3),
        lambda: pct.fuse_device(img, atlas, atlas.clone(), small_out, 3),
        lambda: pct.fuse_device(img, torch.zeros(2, 6, 7, 9), torch.zeros(2, 6, 7, 9),
                                torch.zeros(6, 7, 8), 3),
        lambda: pct.fuse_device(img[0], atlas, atlas.clone(), torch.zeros(6, 7, 8), 3),
        lambda: pct.autotune(img, atlas, small_label, torch.zeros(6, 7, 8), 3),
        lambda: pct.fuse_batch([{"img": img, "atlas": atlas, "label": small_label,
                                 "out": torch.zeros(6, 7, 8), "nh": 3}]),
        lambda: learned.fuse_pruned_device(img, atlas, atlas.clone(), small_out, 3, 2.5, 1.0,
                                           learned.mlp_init(0, 0.0, 1.0)),
        lambda: learned.mlp_train(img, torch.zeros(2, 6, 7), 3, 2.5, 1.0,
                                  learned.mlp_init(0, 0.0, 1.0), steps=1),
    ]
    for call in calls:
        with pytest.raises(ValueError):
            call()
    # matching shapes pass the check; host tensors are then refused as such
    pct.check_volume_set(img, atlas, atlas.clone(), torch.zeros(6, 7, 8))
    with pytest.raises(TypeError):
        pct.fuse_device(img, atlas, atlas.clone(), torch.zeros(6, 7, 8), 3,
                        stream=type("S", (), {"device": "cpu", "cuda_stream": 0})())
