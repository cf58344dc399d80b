"""GPU parity of the plain FP32 fusion (tiled SM kernel and direct GM/GM2
kernel) against the oracle, through the C ABI."""
import numpy as np
import pytest

import oracle
import pseudoctimaging_amd as pct
from parity import MAX_DHU, compare

pytestmark = pytest.mark.gpu


def _case(nx, ny, nz, n, shift=2, seed=0, independent=False):
    img, amri, act = pct.synth_tissue(nx, ny, nz, n, shift, independent, seed)
    lab = act.copy()
    pct.normalize_labels(lab)
    return img, amri, lab


SMALL = [
    # nx, ny, nz, n, nh, sigma, beta
    (24, 26, 20, 2, 3, 2.5, 1.0),
    (30, 34, 22, 2, 7, 2.5, 1.0),
    (28, 40, 24, 3, 9, 2.5, 1.0),
    (34, 30, 26, 2, 11, 2.5, 1.0),
    (33, 37, 25, 1, 5, 1.0, 1.0),
    (40, 70, 21, 2, 7, 2.5, 0.5),
]


@pytest.mark.parametrize("nx,ny,nz,n,nh,sigma,beta", SMALL)
@pytest.mark.parametrize("choice", [pct.CHOICE_SM, pct.CHOICE_DIRECT])
def test_plain_vs_oracle(gpu, nx, ny, nz, n, nh, sigma, beta, choice):
    print(f"case {nx}x{ny}x{nz} n={n} nh={nh} choice={choice}", flush=True)
    img, amri, lab = _case(nx, ny, nz, n, seed=nx * 7 + nh)
    got = pct.multi_patch_based_segmentation_gpu(img, amri, lab, nh, sigma, beta, choice)
    ref = oracle.fuse(img, amri, lab, nh, sigma, beta)
    st = compare(got, ref)
    # border voxels are exactly 0
    r = (nh - 1) // 2 + 1
    mask = np.ones(img.shape, bool)
    mask[r:nz - r, r:nx - r, r:ny - r] = False
    assert np.all(got[mask] == 0)
    print(st)


def test_all_tile_shapes_agree(gpu):
    img, amri, lab = _case(44, 70, 30, 3, seed=3)
    ref = oracle.fuse(img, amri, lab, 7, 2.5, 1.0)
    outs = []
    try:
        for sid in range(pct._lib.lib().pct_tile_config_count()):
            for grp in (1, 2, 3):
                pct.force_tile_config(sid, grp)
                got = pct.multi_patch_based_segmentation_gpu(img, amri, lab, 7, 2.5, 1.0,
                                                             pct.CHOICE_SM)
                compare(got, ref)
                outs.append(got)
    finally:
        pct.force_tile_config(-1, 0)
    assert len(outs) >= 3


def test_deterministic(gpu):
    img, amri, lab = _case(36, 66, 28, 2, seed=5)
    a = pct.multi_patch_based_segmentation_gpu(img, amri, lab, 9, 2.5, 1.0)
    b = pct.multi_patch_based_segmentation_gpu(img, amri, lab, 9, 2.5, 1.0)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


# The cross-term shapes (tile configurations 15-19) weigh with
# SSD' = TT + AA - C only where every scaled MRI value is at most 0.25 in
# magnitude and every CT value at most 2^64.  An item with a larger target
# value keeps the squared differences throughout; a slab with a larger atlas
# value, or a CT value beyond 2^64 or not a number (which the pair pre-pass
# marks with +inf), is weighed with squared differences on the cross-term
# scale.  Each case below steers the kernels into one of those paths, and
# every path must meet the same bar; near a CT spike of 2e19 the HU bar has no
# meaning (a float ulp there is far above 0.5 HU), and the relative bar and
# the NaN pattern remain.
XT_SHAPES = [15, 16, 17, 18, 19]


def _xt_variants():
    img, amri, act = pct.synth_tissue(40, 44, 30, 2, 2, False, 11)
    lab = act.copy()
    pct.normalize_labels(lab)
    spiked_atlas = amri.copy()
    spiked_atlas[1, 10:12, 20, 20:22] = 30.0     # scaled beyond the bound: slab fallback
    spiked_img = img.copy()
    spiked_img[12, 15, 18] = 25.0                 # a target value beyond it: squared items
    uimg, uatlas, uct = pct.synth_uniform(40, 44, 30, 2, 12)   # values up to 10
    ulab = uct.copy()
    pct.normalize_labels(ulab)
    spiked_lab = lab.copy()
    spiked_lab[0, 14, 21, 9] = 2e19               # CT beyond 2^64: slab fallback
    nan_lab = lab.copy()
    nan_lab[1, 9, 30, 25] = np.nan                # CT not a number: NaN outputs around it
    return [("tissue", img, amri, lab, MAX_DHU), ("atlas-spike", img, spiked_atlas, lab, MAX_DHU),
            ("target-spike", spiked_img, amri, lab, MAX_DHU),
            ("both", spiked_img, spiked_atlas, lab, MAX_DHU),
            ("uniform", uimg, uatlas, ulab, MAX_DHU), ("ct-spike", img, amri, spiked_lab, np.inf),
            ("ct-nan", img, amri, nan_lab, MAX_DHU)]


@pytest.mark.parametrize("nh,sigma", [(7, 2.5), (9, 1.0), (5, 0.5)])
def test_cross_term_range_fallbacks(gpu, nh, sigma):
    try:
        for name, img, amri, lab, max_dhu in _xt_variants():
            ref = oracle.fuse(img, amri, lab, nh, sigma, 1.0)
            for sid in XT_SHAPES:
                print(f"case {name} nh={nh} sigma={sigma} shape={sid}", flush=True)
                pct.force_tile_config(sid, 1)
                got = pct.multi_patch_based_segmentation_gpu(img, amri, lab, nh, sigma, 1.0,
                                                             pct.CHOICE_SM)
                compare(got, ref, max_dhu=max_dhu)
    finally:
        pct.force_tile_config(-1, 0)
