"""pct_release frees every cached resource of the library (device
workspaces, the host-call buffers and pinned staging ring, the pruned call's
second stream, batch graphs, model copies); the next calls recreate them and
return the same results bit for bit."""
import numpy as np
import pytest

import pseudoctimaging_amd as pct
from pseudoctimaging_amd import learned

pytestmark = pytest.mark.gpu


def _run_all(img, amri, lab, dev_in, subs):
    import torch
    nh = 9
    fused = pct.multi_patch_based_segmentation_gpu(img, amri, lab, nh, 2.5, 1.0)
    hu = pct.pseudo_ct_hu(img, amri, lab, nh, 2.5, 1.0)
    p = learned.mlp_init(4, *learned.intensity_stats(img))
    pruned = torch.full(img.shape, -1.0, device="cuda:0")
    learned.fuse_pruned_device(*dev_in, pruned, nh, 2.5, 1.0, p, topk=16, samples=27)
    for s in subs:
        s["out"].fill_(-1.0)
    pct.fuse_batch(subs)
    torch.cuda.synchronize()
    return [fused, hu, pruned.cpu().numpy()] + [s["out"].cpu().numpy() for s in subs]


def test_release_then_reuse(gpu):
    import torch
    dev = torch.device("cuda:0")
    img, amri, act = pct.synth_tissue(40, 46, 32, 3, 2, False, 61)
    lab = act.copy()
    pct.normalize_labels(lab)
    dev_in = [torch.from_numpy(a).to(dev) for a in (img, amri, lab)]
    subs = []
    for k, (nx, ny, nz, n, nh) in enumerate([(30, 34, 28, 2, 7), (36, 30, 26, 1, 11)]):
        i2, a2, c2 = pct.synth_tissue(nx, ny, nz, n, 2, False, 70 + k)
        l2 = c2.copy()
        pct.normalize_labels(l2)
        subs.append({"img": torch.from_numpy(i2).to(dev), "atlas": torch.from_numpy(a2).to(dev),
                     "label": torch.from_numpy(l2).to(dev),
                     "out": torch.full(i2.shape, -7.0, device=dev), "nh": nh, "sigma": 2.5,
                     "beta": 1.0})
    first = _run_all(img, amri, lab, dev_in, subs)
    for rep in range(2):
        pct.release()
        again = _run_all(img, amri, lab, dev_in, subs)
        for k, (a, b) in enumerate(zip(first, again)):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (rep, k)
        print(f"after release {rep}: equal", flush=True)
