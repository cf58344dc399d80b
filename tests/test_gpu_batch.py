"""Batched fusion (one graph-captured persistent launch over many subjects).

Each subject's output must equal its individual tiled run with the same tile
shape and atlas group bit for bit, and the oracle within the plain path's
tolerance.  The captured graph is replayed several times with eager work and
synchronisations in between, and every replay is checked.
"""
import numpy as np
import pytest

import oracle
import pseudoctimaging_amd as pct
from parity import compare

pytestmark = pytest.mark.gpu

# nx, ny, nz, n, nh: uneven sizes, atlas counts and windows, including a
# window without a tiled kernel (NH=17) and an empty interior
SUBJECTS = [
    (40, 44, 36, 1, 7),
    (48, 50, 40, 3, 9),
    (56, 62, 44, 2, 11),
    (30, 34, 30, 2, 3),
    (26, 27, 25, 1, 17),
    (10, 10, 10, 1, 11),
    (44, 40, 38, 4, 7),
]
BATCH_SHAPES = [7, 8, 9, 10, 11, 13, 16, 17, 19]


def make_subjects(specs, seed=0):
    import torch
    dev = torch.device("cuda:0")
    host, subs = [], []
    for k, (nx, ny, nz, n, nh) in enumerate(specs):
        img, amri, act = pct.synth_tissue(nx, ny, nz, n, 2, False, seed + k)
        lab = act.copy()
        pct.normalize_labels(lab)
        host.append((img, amri, lab, nh))
        subs.append({"img": torch.from_numpy(img).to(dev), "atlas": torch.from_numpy(amri).to(dev),
                     "label": torch.from_numpy(lab).to(dev),
                     "out": torch.full(img.shape, -7.0, device=dev), "nh": nh,
                     "sigma": 2.5, "beta": 1.0})
    return host, subs


@pytest.mark.parametrize("shape_id", BATCH_SHAPES)
@pytest.mark.parametrize("group", [1, 2])
def test_batch_equals_individual_runs(gpu, shape_id, group):
    import torch
    print(f"case batch shape {shape_id} group {group}", flush=True)
    host, subs = make_subjects(SUBJECTS)
    pct.set_batch_config(subs, shape_id, group)
    pct.fuse_batch(subs)
    torch.cuda.synchronize()
    try:
        pct.force_tile_config(shape_id, group)
        for (img, amri, lab, nh), s in zip(host, subs):
            single = pct.multi_patch_based_segmentation_gpu(img, amri, lab, nh, 2.5, 1.0)
            got = s["out"].cpu().numpy()
            assert np.array_equal(got.view(np.uint32), single.view(np.uint32)), nh
    finally:
        pct.force_tile_config(-1, 0)


def _check_against_individual(host, subs, classes):
    """Each subject equals its own tiled run with its class's shape and group."""
    try:
        for (img, amri, lab, nh), s in zip(host, subs):
            sid, grp = classes.get((nh - 1) // 2, (BATCH_SHAPES[0], 1))
            pct.force_tile_config(sid, grp)
            single = pct.multi_patch_based_segmentation_gpu(img, amri, lab, nh, 2.5, 1.0)
            got = s["out"].cpu().numpy()
            assert np.array_equal(got.view(np.uint32), single.view(np.uint32)), (nh, sid, grp)
    finally:
        pct.force_tile_config(-1, 0)


@pytest.mark.parametrize("serial", [False, True])
def test_batch_per_class_configuration(gpu, serial):
    """Window classes with different tile shapes and atlas groups, run
    concurrently or one after another, in one graph."""
    import torch
    print(f"case batch per-class serial {serial}", flush=True)
    host, subs = make_subjects(SUBJECTS, seed=50)
    classes = {1: (11, 1), 3: (7, 2), 4: (9, 1), 5: (13, 2)}
    pct.set_batch_choice(subs, classes, serial)
    got = pct.batch_choice(subs)
    assert got["serial"] == serial
    # the NH=17 subject has no tiled class; (10, 10, 10) has no interior
    assert got["classes"] == classes
    for replay in range(2):
        for s in subs:
            s["out"].fill_(-3.0)
        pct.fuse_batch(subs)
        torch.cuda.synchronize()
        _check_against_individual(host, subs, classes)


def test_batch_against_oracle_and_graph_replays(gpu):
    import torch
    host, subs = make_subjects(SUBJECTS, seed=100)
    refs = [oracle.fuse(img, amri, lab, nh, 2.5, 1.0) for (img, amri, lab, nh) in host]
    scratch = torch.randn(1 << 20, device="cuda:0")
    for replay in range(4):
        for s in subs:
            s["out"].fill_(-1.0)
        pct.fuse_batch(subs)
        # the program's own eager work between replays
        scratch.mul_(1.0001).add_(0.5)
        torch.cuda.synchronize()
        for ref, s in zip(refs, subs):
            compare(s["out"].cpu().numpy(), ref)
        print(f"replay {replay} ok", flush=True)


def test_batch_autotune_and_workspace_variant(gpu):
    import torch
    from pseudoctimaging_amd import _lib
    host, subs = make_subjects(SUBJECTS[:4], seed=7)
    cfg = pct.autotune_batch(subs, reps=1)
    assert cfg["id"] in BATCH_SHAPES and cfg["measured_ms"] > 0
    choice = pct.batch_choice(subs)
    assert set(choice["classes"]) == {3, 4, 5, 1}
    assert all(sid in BATCH_SHAPES for sid, _ in choice["classes"].values())
    pct.fuse_batch(subs)
    torch.cuda.synchronize()
    _check_against_individual(host, subs, choice["classes"])
    first = [s["out"].clone() for s in subs]
    arr = pct.api._subjects(subs)
    L = _lib.lib()
    need = L.pct_fuse_batch_workspace_size(arr, len(subs))
    assert need > 0
    ws = torch.empty(need // 4 + 64, device="cuda:0")
    for s in subs:
        s["out"].zero_()
    rc = L.pct_fuse_batch_dev_ws(arr, len(subs), _lib.c_void_p(ws.data_ptr()), need,
                                 _lib.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    for a, s in zip(first, subs):
        assert torch.equal(a, s["out"])
    too_many = pct.api._subjects(subs * 5)
    assert L.pct_fuse_batch_dev(too_many, 20, None) == _lib.PCT_ERR_TOO_LARGE
