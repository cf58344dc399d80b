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


def _hu_mae(a, b):
    ok = np.isfinite(a) & np.isfinite(b)
    return float(np.abs(4095.0 * (a[ok] - b[ok]) / 255.0).mean()), ok


def test_pruned_accuracy_and_network_gain(gpu):
    import torch
    shape, n, nh = (48, 64, 60), 3, 11
    img, amri, lab = _fusion_case(shape, n, nh, 31)
    dimg, dam, dlab = _dev(img), _dev(amri), _dev(lab)
    exact = torch.zeros(img.shape, device="cuda:0")
    pct.fuse_device(dimg, dam, dlab, exact, nh, 2.5, 1.0)
    mu, inv_sd = learned.intensity_stats(img)
    p0 = learned.mlp_init(3, mu, inv_sd)
    p, trace = learned.mlp_train(dimg, dam, nh, 2.5, 1.0, p0, steps=150, anchors=128, lr=3e-3,
                                 seed=4)
    res = {}
    for name, params in (("untrained", p0), ("trained", p)):
        out = torch.zeros(img.shape, device="cuda:0")
        frac = learned.fuse_pruned_device(dimg, dam, dlab, out, nh, 2.5, 1.0, params, topk=32,
                                          samples=27, margin=0.0)
        torch.cuda.synchronize()
        mae, ok = _hu_mae(out.cpu().numpy(), exact.cpu().numpy())
        res[name] = mae
        print(f"{name}: MAE {mae:.3f} HU, fallback {frac:.3f}", flush=True)
        # NaN pattern: the pruned sums include every exact weight's tail
        # estimate, so a voxel is NaN only where the exact one is
        e = exact.cpu().numpy()
        o = out.cpu().numpy()
        assert np.array_equal(np.isnan(o), np.isnan(e))
    assert res["trained"] <= 5.0
    assert res["trained"] < res["untrained"]
