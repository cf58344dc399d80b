"""Learned candidate proposal on the GPU: MFMA embedding against the FP64 /
BF16-emulating restatement, training gradients against the FP64 gradient,
training progress, the pruned fusion's degenerate cases (bit-exact with the
plain tiled kernel) and its accuracy against the exact fusion, with and
without a trained network."""
import numpy as np
import pytest

import mlp_ref
import pseudoctimaging_amd as pct
from pseudoctimaging_amd import learned

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _random_params(seed, mu, inv_sd, bias=0.1):
    p = learned.mlp_init(seed, mu, inv_sd)
    rng = np.random.default_rng(seed)
    for lo, hi in ((learned.OFF_B1, learned.OFF_W2), (learned.OFF_B2, learned.OFF_W3),
                   (learned.OFF_B3, learned.OFF_MU)):
        p[lo:hi] = rng.normal(0, bias, hi - lo).astype(np.float32)
    return p


def test_embedding_matches_reference(gpu):
    import torch
    print("case embed 11x13x21", flush=True)
    rng = np.random.default_rng(3)
    vol = rng.normal(100.0, 30.0, (11, 13, 21)).astype(np.float32)   # ny = 21: ragged chunk
    mu, inv_sd = learned.intensity_stats(vol)
    p = _random_params(7, mu, inv_sd)
    out = learned.embed_device(p, _dev(vol))
    torch.cuda.synchronize()
    got = learned.bf16_words_to_float(out.cpu().numpy()).astype(np.float64)
    ref16 = mlp_ref.embed_volume(p, vol, bf16=True)
    ref64 = mlp_ref.embed_volume(p, vol, bf16=False)
    scale = np.sqrt((ref64 ** 2).mean())
    # against the same BF16 arithmetic: accumulation order and the odd
    # rounding flip of a hidden activation only
    err16 = np.abs(got - ref16)
    assert err16.max() <= 0.03 * scale, err16.max() / scale
    assert err16.mean() <= 0.002 * scale
    # against FP64: the BF16 error budget
    err64 = np.abs(got - ref64)
    assert err64.mean() <= 0.02 * scale, err64.mean() / scale


def test_embedding_layout_is_not_transposed(gpu):
    # A single non-zero output column: only embedding dimension 5 may vary.
    import torch
    rng = np.random.default_rng(4)
    vol = rng.normal(0.0, 1.0, (6, 7, 16)).astype(np.float32)
    p = _random_params(1, 0.0, 1.0, bias=0.0)
    W3 = p[learned.OFF_W3:learned.OFF_B3].reshape(64, 16)
    W3[:, np.arange(16) != 5] = 0.0
    out = learned.embed_device(p, _dev(vol))
    torch.cuda.synchronize()
    got = learned.bf16_words_to_float(out.cpu().numpy())
    assert np.all(got[..., np.arange(16) != 5] == 0.0)
    assert np.abs(got[..., 5]).max() > 0.0


def _training_data(shape=(22, 24, 20), n=2, seed=11):
    img, amri, act = pct.synth_tissue(shape[1], shape[2], shape[0], n, 2, False, seed)
    return img, amri, act


def test_training_gradient_matches_fp64(gpu):
    img, amri, _ = _training_data()
    nh, sigma, beta, tau = 3, 2.5, 1.0, 0.8
    mu, inv_sd = learned.intensity_stats(img)
    p = _random_params(5, mu, inv_sd)
    r = (nh - 1) // 2 + 1
    rng = np.random.default_rng(9)
    anchors = np.stack([rng.integers(r, img.shape[0] - r, 12), rng.integers(r, img.shape[1] - r, 12),
                        rng.integers(r, img.shape[2] - r, 12), rng.integers(0, 2, 12)], axis=1)
    teacher_t = 0.4
    loss, g = learned.mlp_loss_grad(_dev(img), _dev(amri), nh, sigma, beta, p, anchors, tau,
                                    teacher_t)
    loss64, g64 = mlp_ref.loss_grad(p, img, amri, [tuple(a) for a in anchors], nh, teacher_t,
                                    tau)
    print(f"loss gpu {loss:.8g} fp64 {loss64:.8g}", flush=True)
    assert abs(loss - loss64) <= 1e-4 * abs(loss64) + 1e-7
    rel = np.linalg.norm(g - g64) / np.linalg.norm(g64)
    assert rel <= 1e-3, rel
    big = np.abs(g64) > 1e-3 * np.abs(g64).max()
    assert np.allclose(g[big], g64[big], rtol=2e-2, atol=1e-4 * np.abs(g64).max())
    assert np.all(g[30 * 64:2048] == 0) and np.all(g[7312:] == 0)
    # deterministic
    _, g2 = learned.mlp_loss_grad(_dev(img), _dev(amri), nh, sigma, beta, p, anchors, tau,
                                  teacher_t)
    assert np.array_equal(g, g2)
    # teacher_t <= 0 is the fusion divisor
    div = pct.weight_divisor(sigma, beta)
    loss_d, g_d = learned.mlp_loss_grad(_dev(img), _dev(amri), nh, sigma, beta, p, anchors, tau)
    loss64_d, _ = mlp_ref.loss_grad(p, img, amri, [tuple(a) for a in anchors], nh, div, tau)
    assert abs(loss_d - loss64_d) <= 1e-3 * abs(loss64_d) + 1e-7


def test_training_reduces_loss(gpu):
    img, amri, _ = _training_data(shape=(30, 32, 28), n=2)
    mu, inv_sd = learned.intensity_stats(img)
    p0 = learned.mlp_init(1, mu, inv_sd)
    p, trace = learned.mlp_train(_dev(img), _dev(amri), 7, 2.5, 1.0, p0, steps=80, anchors=64,
                                 lr=3e-3, tau=1.0, seed=2)
    print("loss trace", trace[:3], trace[-3:], flush=True)
    assert np.all(np.isfinite(trace))
    assert trace[-10:].mean() < 0.8 * trace[:10].mean()
    assert not np.array_equal(p, p0)
    assert np.array_equal(p[7312:], p0[7312:])


def _fusion_case(shape, n, nh, seed):
    img, amri, act = pct.synth_tissue(shape[1], shape[2], shape[0], n, 2, False, seed)
    lab = act.copy()
    pct.normalize_labels(lab)
    return img, amri, lab


@pytest.mark.parametrize("shape_id", [8, 13])
def test_pruned_all_candidates_is_bit_exact(gpu, shape_id):
    # NH = 3: topk = 27 keeps every offset, no samples -> the pruned kernel
    # evaluates every candidate in the plain kernel's order and the tail is 0.
    import torch
    img, amri, lab = _fusion_case((26, 40, 34), 3, 3, 21)
    p = learned.mlp_init(0, *learned.intensity_stats(img))
    dimg, dam, dlab = _dev(img), _dev(amri), _dev(lab)
    out = torch.full(img.shape, -3.0, device="cuda:0")
    frac = learned.fuse_pruned_device(dimg, dam, dlab, out, 3, 2.5, 1.0, p, topk=27, samples=0,
                                      margin=0.0, shape_id=shape_id)
    assert frac == 0.0
    ref = torch.full(img.shape, -5.0, device="cuda:0")
    try:
        pct.force_tile_config(shape_id, 1)
        pct.fuse_device(dimg, dam, dlab, ref, 3, 2.5, 1.0)
    finally:
        pct.force_tile_config(-1, 0)
    torch.cuda.synchronize()
    a, b = out.cpu().numpy(), ref.cpu().numpy()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("shape,nh", [((40, 37, 70), 11), ((30, 45, 131), 7), ((36, 33, 29), 15)])
def test_tail_term_constant_mri_is_window_mean(gpu, shape, nh):
    # With a constant MRI every candidate weight is exp(0) = 1, so the pruned
    # fusion, kept offsets and lattice samples weighed exactly and every other
    # candidate through the tail term, must return the mean of the atlas CT
    # over the NH^3 window (the closed form of the exact fusion).  This pins
    # the window sums of the CT on ragged tile edges.
    import torch
    nz, nx, ny = shape
    n = 2
    hw = (nh - 1) // 2
    r = hw + 1
    rng = np.random.default_rng(nh)
    img = np.full(shape, 0.5, np.float32)
    amri = np.full((n,) + shape, 0.5, np.float32)
    lab = (rng.random((n,) + shape) * 255.0).astype(np.float32)
    p = learned.mlp_init(2, 0.5, 1.0)
    out = torch.full(shape, -7.0, device="cuda:0")
    print(f"case tail window mean {shape} NH={nh}", flush=True)
    frac = learned.fuse_pruned_device(_dev(img), _dev(amri), _dev(lab), out, nh, 2.5, 1.0, p,
                                      topk=16, samples=27, margin=0.0, max_err=0.0)
    torch.cuda.synchronize()
    assert frac == 0.0
    got = out.cpu().numpy()
    # window sums by a 3-d summed-volume table in float64
    c = np.zeros((n, nz + 1, nx + 1, ny + 1))
    c[:, 1:, 1:, 1:] = lab.astype(np.float64).cumsum(1).cumsum(2).cumsum(3)
    k = 2 * hw + 1
    z0, z1 = r - hw, nz - r - hw   # window starts of the interior voxels
    x0, x1 = r - hw, nx - r - hw
    y0, y1 = r - hw, ny - r - hw
    box = np.zeros((n, nz - 2 * r, nx - 2 * r, ny - 2 * r))
    for dz, sz in ((k, 1), (0, -1)):
        for dx, sx in ((k, 1), (0, -1)):
            for dy, sy in ((k, 1), (0, -1)):
                box += sz * sx * sy * c[:, z0 + dz:z1 + dz, x0 + dx:x1 + dx, y0 + dy:y1 + dy]
    want = box.sum(0) / (n * k ** 3)
    inner = got[r:nz - r, r:nx - r, r:ny - r].astype(np.float64)
    assert np.isfinite(inner).all()
    err = np.abs(inner - want) / np.maximum(np.abs(want), 1.0)
    assert err.max() < 2e-5, (err.max(), np.unravel_index(err.argmax(), err.shape))
    # the border is zero
    border = got.copy()
    border[r:nz - r, r:nx - r, r:ny - r] = 0.0
    assert not border.any()


def test_pruned_full_fallback_is_bit_exact(gpu):
    # negative margin: every item takes the full search
    import torch
    img, amri, lab = _fusion_case((30, 36, 40), 2, 9, 22)
    p = learned.mlp_init(0, *learned.intensity_stats(img))
    dimg, dam, dlab = _dev(img), _dev(amri), _dev(lab)
    out = torch.zeros(img.shape, device="cuda:0")
    frac = learned.fuse_pruned_device(dimg, dam, dlab, out, 9, 2.5, 1.0, p, topk=32,
                                      margin=-1.0, shape_id=13)
    assert frac == 1.0
    ref = torch.zeros(img.shape, device="cuda:0")
    try:
        pct.force_tile_config(13, 1)
        pct.fuse_device(dimg, dam, dlab, ref, 9, 2.5, 1.0)
    finally:
        pct.force_tile_config(-1, 0)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().view(np.uint32), ref.cpu().numpy().view(np.uint32))


@pytest.mark.parametrize("shape", [(13, 20, 15), (14, 13, 40), (20, 14, 13)])
def test_pruned_thin_interior(gpu, shape):
    # NH = 11 (r = 6) on volumes whose interior is 1-3 voxels thick along one
    # axis: probe clusters, window planes and window sums sit against the
    # faces.  The full fallback equals the plain kernel bit for bit; the
    # pruned fusion stays finite inside, zero on the border, and close to
    # the exact fusion.
    import torch
    nh, r = 11, 6
    img, amri, lab = _fusion_case(shape, 1, nh, 25)
    p = learned.mlp_init(6, *learned.intensity_stats(img))
    dimg, dam, dlab = _dev(img), _dev(amri), _dev(lab)
    print(f"case thin interior {shape}", flush=True)
    ref = torch.zeros(img.shape, device="cuda:0")
    try:
        pct.force_tile_config(13, 1)
        pct.fuse_device(dimg, dam, dlab, ref, nh, 2.5, 1.0)
    finally:
        pct.force_tile_config(-1, 0)
    full = torch.full(img.shape, -1.0, device="cuda:0")
    frac = learned.fuse_pruned_device(dimg, dam, dlab, full, nh, 2.5, 1.0, p, topk=32,
                                      margin=-1.0, shape_id=13)
    assert frac == 1.0
    out = torch.full(img.shape, -1.0, device="cuda:0")
    frac0 = learned.fuse_pruned_device(dimg, dam, dlab, out, nh, 2.5, 1.0, p, topk=32,
                                       margin=0.0, max_err=0.0)
    torch.cuda.synchronize()
    assert frac0 == 0.0
    e, f, o = ref.cpu().numpy(), full.cpu().numpy(), out.cpu().numpy()
    assert np.array_equal(f.view(np.uint32), e.view(np.uint32))
    nz, nx, ny = shape
    inner = (slice(r, nz - r), slice(r, nx - r), slice(r, ny - r))
    assert np.isfinite(o[inner]).all()
    border = o.copy()
    border[inner] = 0.0
    assert not border.any()
    mae, _ = _hu_mae(o[inner], e[inner])
    assert mae < 10.0, mae


@pytest.mark.parametrize("case", [((30, 40, 44), 2, 11, 32, 4), ((26, 36, 30), 3, 9, 16, 0),
                                  ((24, 30, 28), 2, 7, 64, 2), ((20, 22, 26), 1, 5, 7, 4)])
def test_radix_selection_equals_sorted_selection(gpu, case):
    # margin 0 selects the kept offsets by radix select; a positive margin
    # needs the order statistics and sorts.  Both keep the topk smallest
    # (score, offset) pairs, so where the sorting path flags no item the
    # outputs are bit-identical.
    import torch
    shape, n, nh, topk, slabs = case
    print(f"case radix vs sort {case}", flush=True)
    img, amri, lab = _fusion_case(shape, n, nh, 24)
    p = _random_params(5, *learned.intensity_stats(img))
    dimg, dam, dlab = _dev(img), _dev(amri), _dev(lab)
    fast = torch.full(img.shape, -1.0, device="cuda:0")
    sort = torch.full(img.shape, -2.0, device="cuda:0")
    kw = dict(topk=topk, samples=27, shape_id=13, max_slabs=slabs, max_err=0.0)
    f0 = learned.fuse_pruned_device(dimg, dam, dlab, fast, nh, 2.5, 1.0, p, margin=0.0, **kw)
    f1 = learned.fuse_pruned_device(dimg, dam, dlab, sort, nh, 2.5, 1.0, p, margin=1e-37, **kw)
    torch.cuda.synchronize()
    assert f0 == 0.0 and f1 == 0.0
    a, b = fast.cpu().numpy(), sort.cpu().numpy()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("shape,nh", [((30, 40, 44), 11), ((16, 17, 50), 11), ((40, 70, 21), 11),
                                      ((24, 30, 28), 5), ((20, 22, 26), 7), ((34, 36, 40), 15)])
def test_window_embedding_equals_full(gpu, shape, nh):
    # The pruned path embeds only the atlas voxels the selection reads (the
    # window planes of the scored dz); with every voxel embedded instead, the
    # result is bit-identical.
    import torch
    img, amri, lab = _fusion_case(shape, 2, nh, 23)
    p = _random_params(9, *learned.intensity_stats(img))
    dimg, dam, dlab = _dev(img), _dev(amri), _dev(lab)
    outs = []
    for full in (False, True):
        out = torch.full(img.shape, -9.0, device="cuda:0")
        learned.fuse_pruned_device(dimg, dam, dlab, out, nh, 2.5, 1.0, p, topk=16, samples=27,
                                   max_err=0.0,
                                   margin=0.05, embed_all=full)
        torch.cuda.synchronize()
        outs.append(out.cpu().numpy())
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


def _hu_mae(a, b):
    ok = np.isfinite(a) & np.isfinite(b)
    return float(np.abs(4095.0 * (a[ok] - b[ok]) / 255.0).mean()), ok


def _pruned_maes(sigma, steps=150, seed=31, shape=(48, 64, 60), n=3, nh=11, max_err=0.0):
    """MAE (HU) of the pruned fusion against the exact fusion with an
    untrained and a trained network, plus the fallback fractions."""
    import torch
    img, amri, lab = _fusion_case(shape, n, nh, seed)
    dimg, dam, dlab = _dev(img), _dev(amri), _dev(lab)
    exact = torch.zeros(img.shape, device="cuda:0")
    pct.fuse_device(dimg, dam, dlab, exact, nh, sigma, 1.0)
    mu, inv_sd = learned.intensity_stats(img)
    p0 = learned.mlp_init(3, mu, inv_sd)
    p, _ = learned.mlp_train(dimg, dam, nh, sigma, 1.0, p0, steps=steps, anchors=128, lr=3e-3,
                             seed=4)
    res = {}
    e = exact.cpu().numpy()
    for name, params in (("untrained", p0), ("trained", p)):
        out = torch.zeros(img.shape, device="cuda:0")
        frac = learned.fuse_pruned_device(dimg, dam, dlab, out, nh, sigma, 1.0, params, topk=32,
                                          samples=27, margin=0.0, max_err=max_err)
        torch.cuda.synchronize()
        o = out.cpu().numpy()
        mae, _ = _hu_mae(o, e)
        res[name] = (mae, frac)
        print(f"sigma {sigma} max_err {max_err} {name}: MAE {mae:.3f} HU, fallback {frac:.3f}",
              flush=True)
        # the pruned sums carry the tail estimate of every skipped candidate,
        # so a voxel is NaN exactly where the exact one is
        assert np.array_equal(np.isnan(o), np.isnan(e))
    return res


def test_pruned_accuracy_flat_weights(gpu):
    # sigma = 2.5: every weight lies in [0.946, 1]; the tail term carries the
    # result and the bar is <= 5 HU MAE against the exact fusion, with and
    # without the error check (whose default flags nothing here)
    for max_err in (0.0, learned.DEFAULT_MAX_ERR):
        res = _pruned_maes(2.5, max_err=max_err)
        assert res["trained"][0] <= 5.0
        assert res["untrained"][0] <= 5.0


def test_network_gain_selective_weights(gpu):
    # sigma = 0.25: weights span orders of magnitude, and a per-tile top-k
    # with a sampled tail is far from the exact fusion (tens of HU).  The
    # error check flags those tiles for the full search: with the default
    # bar the result is within 5 HU MAE of the exact fusion.
    bare = _pruned_maes(0.25)
    checked = _pruned_maes(0.25, max_err=learned.DEFAULT_MAX_ERR)
    for name in ("untrained", "trained"):
        assert checked[name][0] <= 5.0, (name, checked[name])
        assert checked[name][1] > 0.0
        assert bare[name][0] > checked[name][0]


def test_pruned_error_fallback_mixed_is_bit_exact(gpu):
    # A bar that flags some tiles and not others.  With one atlas a voxel's
    # output comes from one (tile, atlas) item, so every flagged tile must
    # equal the plain kernel of the same tile shape bit for bit, and the
    # count of such tiles must be the reported fraction of the tiles.
    import math

    import torch
    nz, nx, ny = 36, 70, 66
    nh, r = 11, 6
    img, amri, lab = _fusion_case((nz, nx, ny), 1, nh, 51)
    p = learned.mlp_init(0, *learned.intensity_stats(img))
    dimg, dam, dlab = _dev(img), _dev(amri), _dev(lab)
    ref = torch.zeros(img.shape, device="cuda:0")
    try:
        pct.force_tile_config(13, 1)
        pct.fuse_device(dimg, dam, dlab, ref, nh, 2.5, 1.0)
    finally:
        pct.force_tile_config(-1, 0)
    e = ref.cpu().numpy()
    frac, o = None, None
    for max_err in (0.5, 0.4, 0.6, 0.3, 0.75, 0.2, 0.9):
        out = torch.full(img.shape, -1.0, device="cuda:0")
        frac = learned.fuse_pruned_device(dimg, dam, dlab, out, nh, 2.5, 1.0, p, topk=32,
                                          samples=27, shape_id=13, max_err=max_err)
        print(f"max_err {max_err}: fallback {frac:.3f}", flush=True)
        if 0.0 < frac < 1.0:
            o = out.cpu().numpy()
            break
    assert o is not None, "no bar gave a mixed fallback"
    # tiles of shape 13 {32, 2, 4, 8}: 4 z planes, 32 x rows, 30 y columns
    tz_n, tx_n, ty_n = (math.ceil((d - 2 * r) / t) for d, t in ((nz, 4), (nx, 32), (ny, 30)))
    exact = 0
    for tz in range(tz_n):
        for tx in range(tx_n):
            for ty in range(ty_n):
                sl = (slice(r + 4 * tz, min(r + 4 * tz + 4, nz - r)),
                      slice(r + 32 * tx, min(r + 32 * tx + 32, nx - r)),
                      slice(r + 30 * ty, min(r + 30 * ty + 30, ny - r)))
                exact += int(np.array_equal(o[sl].view(np.uint32), e[sl].view(np.uint32)))
    n_tiles = tz_n * tx_n * ty_n
    print(f"{exact} of {n_tiles} tiles exact, fallback {frac:.4f}", flush=True)
    assert exact == round(frac * n_tiles)


def _ranking_quality(params, img, amri, nh, anchors):
    """Mean Spearman correlation between embedding distance and exact patch
    SSD over the window offsets, and mean recall of the exact 32 best
    offsets among the 32 closest embeddings."""
    import torch
    from scipy.stats import spearmanr
    et = learned.bf16_words_to_float(learned.embed_device(params, _dev(img)).cpu().numpy())
    ea = [learned.bf16_words_to_float(learned.embed_device(params, _dev(a)).cpu().numpy())
          for a in amri]
    torch.cuda.synchronize()
    hw = (nh - 1) // 2
    r = np.arange(-hw, hw + 1)
    dz, dx, dy = (v.ravel() for v in np.meshgrid(r, r, r, indexing="ij"))
    p = np.arange(-1, 2)
    pz, px, py = (v.ravel() for v in np.meshgrid(p, p, p, indexing="ij"))
    rhos, recalls = [], []
    for (z, x, y, at) in anchors:
        tp = img[z + pz, x + px, y + py].astype(np.float64)
        A = amri[at]
        ssd = np.array([((tp - A[z + a + pz, x + b + px, y + c + py]) ** 2).sum()
                        for a, b, c in zip(dz, dx, dy)])
        dist = ((et[z, x, y] - ea[at][z + dz, x + dx, y + dy]) ** 2).sum(axis=1)
        rhos.append(spearmanr(ssd, dist).correlation)
        best = set(np.argsort(ssd, kind="stable")[:32])
        near = set(np.argsort(dist, kind="stable")[:32])
        recalls.append(len(best & near) / 32.0)
    return float(np.mean(rhos)), float(np.mean(recalls))


def test_network_gain_ranking(gpu):
    # What training buys: the embedding distance orders the window offsets
    # like the exact method's patch SSD.  Trained on one phantom, measured on
    # another (held out), against the same network untrained.
    nh = 11
    img, amri, _ = _training_data(shape=(40, 48, 44), n=2, seed=41)
    mu, inv_sd = learned.intensity_stats(img)
    p0 = learned.mlp_init(5, mu, inv_sd)
    p, _ = learned.mlp_train(_dev(img), _dev(amri), nh, 2.5, 1.0, p0, steps=200, anchors=128,
                             lr=3e-3, seed=6)
    test_img, test_amri, _ = _training_data(shape=(40, 48, 44), n=2, seed=42)
    rng = np.random.default_rng(12)
    rr = (nh - 1) // 2 + 1
    anchors = np.stack([rng.integers(rr, 40 - rr, 60), rng.integers(rr, 48 - rr, 60),
                        rng.integers(rr, 44 - rr, 60), rng.integers(0, 2, 60)], axis=1)
    rho0, rec0 = _ranking_quality(p0, test_img, test_amri, nh, anchors)
    rho1, rec1 = _ranking_quality(p, test_img, test_amri, nh, anchors)
    print(f"held-out ranking: untrained spearman {rho0:.3f} recall@32 {rec0:.3f}; "
          f"trained spearman {rho1:.3f} recall@32 {rec1:.3f}", flush=True)
    assert rho1 > rho0 + 0.05
    assert rec1 >= rec0
