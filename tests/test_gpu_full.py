"""GPU parity at the headline sizes and through the whole pipeline.

* 222x222x112 (n=5, NH=9 and n=10, NH=11): the reference's own value at
  6144 sampled voxels (golden fixtures computed by the compiled reference on
  crops holding each voxel's whole neighbourhood), plus invariants that hold
  at full size (zero border, determinism, atlas-order invariance).
* 64x64x32, n=2, NH=7: the full volume against the oracle and the golden
  sample of the reference's output.
* Final HU after the reference's post-processing chain: equal except where
  the fused value lies within 1e-3 of a rounding boundary (the output moves
  in 4095/255 HU steps), and then by exactly one step.
"""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
import pseudoctimaging_amd as pct
from pseudoctimaging_amd import _lib
from parity import MAX_REL, compare

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name)))


@pytest.mark.parametrize("name", ["sample_222x222x112_n5_nh9.npz", "sample_222x222x112_n10_nh11.npz"])
def test_full_size_against_reference_samples(gpu, name):
    g = load(name)
    nx, ny, nz, n, nh = (int(v) for v in g["params"][:5])
    sigma, beta, seed, shift = float(g["params"][5]), float(g["params"][6]), \
        int(g["params"][7]), int(g["params"][8])
    print(f"case {name}: {nx}x{ny}x{nz} n={n} nh={nh}", flush=True)
    img, amri, act = pct.synth_tissue(nx, ny, nz, n, shift, False, seed)
    assert digest(img, amri, act) == str(g["input_sha256"]), "generator changed"
    lab = act.copy()
    pct.normalize_labels(lab)
    out = pct.multi_patch_based_segmentation_gpu(img, amri, lab, nh, sigma, beta, pct.CHOICE_SM)
    got = out.reshape(-1)[g["index"]]
    st = compare(got, g["ref_fused"])
    err64 = np.abs(pct.fused_to_hu_raw(got) - pct.fused_to_hu_raw(g["fp64"]))
    print(st, "max |dHU| vs fp64", float(err64.max()))
    # size-independent properties of the whole output
    r = (nh - 1) // 2 + 1
    border = np.ones(out.shape, bool)
    border[r:nz - r, r:nx - r, r:ny - r] = False
    assert np.all(out[border] == 0)
    inner = out[r:nz - r, r:nx - r, r:ny - r]
    assert np.isfinite(inner).all()
    assert inner.min() >= lab.min() - 1e-3 and inner.max() <= lab.max() + 1e-3
    again = pct.multi_patch_based_segmentation_gpu(img, amri, lab, nh, sigma, beta,
                                                   pct.CHOICE_SM)
    assert np.array_equal(out.view(np.uint32), again.view(np.uint32))


def test_222x222x112_atlas_order_invariance(gpu):
    g = load("sample_222x222x112_n10_nh11.npz")
    nx, ny, nz, n, nh = (int(v) for v in g["params"][:5])
    img, amri, act = pct.synth_tissue(nx, ny, nz, n, 3, False, int(g["params"][7]))
    lab = act.copy()
    pct.normalize_labels(lab)
    a = pct.multi_patch_based_segmentation_gpu(img, amri, lab, nh, 2.5, 1.0)
    perm = np.roll(np.arange(n), 3)
    b = pct.multi_patch_based_segmentation_gpu(img, np.ascontiguousarray(amri[perm]),
                                               np.ascontiguousarray(lab[perm]), nh, 2.5, 1.0)
    r = (nh - 1) // 2 + 1
    inner = (slice(r, nz - r), slice(r, nx - r), slice(r, ny - r))
    np.testing.assert_allclose(b[inner], a[inner], rtol=2e-6)


def test_222x222x112_linear_in_ct(gpu):
    # The weights depend on the MRI only, so out(a L + b) = a out(L) + b
    # holds for the whole 222x222x112, n=10, NH=11 volume.
    g = load("sample_222x222x112_n10_nh11.npz")
    nx, ny, nz, n, nh = (int(v) for v in g["params"][:5])
    img, amri, act = pct.synth_tissue(nx, ny, nz, n, 3, False, int(g["params"][7]))
    lab = act.copy()
    pct.normalize_labels(lab)
    a = pct.multi_patch_based_segmentation_gpu(img, amri, lab, nh, 2.5, 1.0)
    alpha, beta = 0.5, 12.0
    lab2 = (alpha * lab + beta).astype(np.float32)
    b = pct.multi_patch_based_segmentation_gpu(img, amri, lab2, nh, 2.5, 1.0)
    r = (nh - 1) // 2 + 1
    inner = (slice(r, nz - r), slice(r, nx - r), slice(r, ny - r))
    want = alpha * a[inner].astype(np.float64) + beta
    # 1e-3 in label units is 0.016 HU
    assert np.abs(b[inner] - want).max() < 1e-3
    assert not b[0].any() and not b[:, :, ny - 1].any()


def test_222x222x112_constant_mri_is_window_mean(gpu):
    # A constant MRI makes every weight exp(0) = 1: the fusion is the mean of
    # the atlas CT over the NH^3 window, at the full 222x222x112, n=10 size.
    g = load("sample_222x222x112_n10_nh11.npz")
    nx, ny, nz, n, nh = (int(v) for v in g["params"][:5])
    _, _, act = pct.synth_tissue(nx, ny, nz, n, 3, False, int(g["params"][7]))
    lab = act.copy()
    pct.normalize_labels(lab)
    img = np.full((nz, nx, ny), 0.5, np.float32)
    amri = np.full((n, nz, nx, ny), 0.5, np.float32)
    got = pct.multi_patch_based_segmentation_gpu(img, amri, lab, nh, 2.5, 1.0)
    hw = (nh - 1) // 2
    r = hw + 1
    k = 2 * hw + 1
    c = np.zeros((nz + 1, nx + 1, ny + 1))
    c[1:, 1:, 1:] = lab.astype(np.float64).sum(0).cumsum(0).cumsum(1).cumsum(2)
    s0 = [slice(r - hw, d - r - hw) for d in (nz, nx, ny)]
    box = np.zeros((nz - 2 * r, nx - 2 * r, ny - 2 * r))
    for mz in (0, 1):
        for mx in (0, 1):
            for my in (0, 1):
                sl = tuple(slice(s.start + k * m, s.stop + k * m) for s, m in zip(s0, (mz, mx, my)))
                box += (-1) ** (3 - mz - mx - my) * c[sl]
    want = box / (n * k ** 3)
    inner = got[r:nz - r, r:nx - r, r:ny - r].astype(np.float64)
    # each work item sums K = 1331 float terms: the bar is the plain path's
    # relative tolerance (at a window mean of about 128 label units that is
    # 0.013 label units, 0.2 HU).  Every weight here is the same value, so
    # the rounding of a float running sum is the same at every add; its bias
    # is at most sum_k ulp(k c) / 2 / (K c) = 2^-25 K = 4e-5 relative, the
    # bound on the mean error
    rel = np.abs(inner - want) / np.abs(want)
    assert rel.max() < MAX_REL, rel.max()
    assert rel.mean() < 2.0 ** -25 * 1331, rel.mean()


def test_64x64x32_full_volume(gpu):
    with open(os.path.join(GOLDEN, "full_64x64x32_n2_nh7.json")) as f:
        g = json.load(f)
    nz, nx, ny = g["shape_zxy"]
    img, amri, act = pct.synth_tissue(nx, ny, nz, g["n_atlas"], g["max_shift"], False, g["seed"])
    assert digest(img, amri, act) == g["input_sha256"]
    lab = act.copy()
    pct.normalize_labels(lab)
    ref = oracle.fuse(img, amri, lab, g["nh"], g["sigma"], g["beta"])
    for choice in (pct.CHOICE_SM, pct.CHOICE_DIRECT):
        out = pct.multi_patch_based_segmentation_gpu(img, amri, lab, g["nh"], g["sigma"],
                                                     g["beta"], choice)
        compare(out, ref)
        idx = np.array(g["sample_index"])
        compare(out.reshape(-1)[idx], np.array(g["sample_fused"], np.float32))


def _final_hu_agrees(fused_gpu, fused_ref, hu_ref, nh):
    hu = pct.postprocess(fused_gpu.copy(), nh)
    diff = hu != hu_ref
    if not diff.any():
        return 0
    # a rounding step may flip only where the fused value sits next to a .5
    # boundary (or next to such a voxel, through the median fill)
    v = np.where(np.isnan(fused_ref), 0, fused_ref).astype(np.float64)
    near = np.abs(v - np.floor(v) - 0.5) < 1e-3
    from scipy.ndimage import binary_dilation
    near = binary_dilation(near, iterations=1)
    assert np.all(near[diff]), "final HU differs away from a rounding boundary"
    step = np.abs(hu[diff] - hu_ref[diff])
    assert np.all((step == 16) | (step == 17)), step
    return int(diff.sum())


@pytest.mark.parametrize("name", ["small_tissue.npz", "small_uniform.npz"])
@pytest.mark.parametrize("choice", [pct.CHOICE_SM, pct.CHOICE_GM, pct.CHOICE_GM2,
                                    pct.CHOICE_DIRECT])
def test_golden_small_through_pipeline(gpu, name, choice):
    g = load(name)
    nh = int(g["params"][4])
    sigma, beta = float(g["params"][5]), float(g["params"][6])
    print(f"case {name} choice {choice}", flush=True)
    fused, hu = pct.pseudo_ct(g["img"], g["atlas"], g["ct_hu"], nh, sigma, beta, choice)
    st = compare(fused, g["fused"])
    flips = _final_hu_agrees(fused, g["fused"], g["hu"], nh)
    print(st, "final HU flips:", flips)


EDGE = [
    # nx, ny, nz, n, nh, sigma, beta: window and volume edge cases
    (10, 10, 10, 2, 11, 2.5, 1.0),   # empty interior: all zeros
    (13, 15, 12, 2, 1, 2.5, 1.0),    # NH=1: one candidate (direct kernel)
    (20, 22, 19, 2, 8, 2.5, 1.0),    # even NH behaves as NH-1
    (30, 33, 29, 1, 13, 2.5, 1.0),   # widest tiled windows
    (34, 35, 33, 1, 15, 2.5, 1.0),
    (24, 25, 23, 1, 17, 2.5, 1.0),   # beyond the tiled kernels: direct kernel
    (15, 64, 14, 3, 3, 2.5, 1.0),    # y wider than one 30-wide segment
    (16, 16, 16, 2, 13, 2.5, 1.0),   # interior of 2x2x2 voxels
]


@pytest.mark.parametrize("nx,ny,nz,n,nh,sigma,beta", EDGE)
def test_edge_cases(gpu, nx, ny, nz, n, nh, sigma, beta):
    print(f"case edge {nx}x{ny}x{nz} n={n} nh={nh}", flush=True)
    img, amri, act = pct.synth_tissue(nx, ny, nz, n, 2, False, nh + nx)
    lab = act.copy()
    pct.normalize_labels(lab)
    ref = oracle.fuse(img, amri, lab, nh, sigma, beta)
    for choice in (pct.CHOICE_SM, pct.CHOICE_DIRECT):
        out = pct.multi_patch_based_segmentation_gpu(img, amri, lab, nh, sigma, beta, choice)
        compare(out, ref)
    if 2 * ((nh - 1) // 2 + 1) >= min(nx, ny, nz):
        assert np.all(out == 0)


def test_uniform_random_mode_with_nan(gpu):
    # the reference's random mode (10*U[0,1) data) at a small sigma: the
    # weights span hundreds of orders of magnitude and many voxels are NaN
    img, atlas, label = pct.synth_uniform(20, 22, 18, 2, 99)
    for sigma in (1.0, 0.5, 0.33, 0.25):
        ref = oracle.fuse(img, atlas, label, 3, sigma, 1.0)
        for choice in (pct.CHOICE_SM, pct.CHOICE_DIRECT):
            out = pct.multi_patch_based_segmentation_gpu(img, atlas, label, 3, sigma, 1.0, choice)
            st = compare(out, ref)
            print(sigma, choice, st)


@pytest.mark.parametrize("case", [
    # nx, ny, nz, n, nh, forced (shape, group) or None, choice
    (40, 66, 30, 10, 9, None, pct.CHOICE_SM),         # 4 chunks of uneven group counts
    (36, 40, 28, 3, 11, (7, 1), pct.CHOICE_SM),       # 3 chunks
    (36, 40, 28, 10, 7, (11, 3), pct.CHOICE_SM),      # groups of 3 atlases, the last short
    (30, 34, 26, 5, 11, (13, 5), pct.CHOICE_SM),      # one group: one chunk
    (30, 34, 26, 2, 17, None, pct.CHOICE_SM),         # no tiled kernel: direct path
    (30, 34, 26, 2, 9, None, pct.CHOICE_DIRECT),      # the direct kernel by choice
    (30, 34, 26, 2, 9, None, pct.CHOICE_GM2),         # GM2 runs the tiled kernel
])
def test_host_call_overlapped_copies_equal_device_call(gpu, case):
    """pct_fuse_host copies the atlases in chunks that overlap the fusion; its
    output equals pct_fuse_dev's bit for bit, also across calls that reuse
    the cached device buffers with smaller and larger problems."""
    import torch
    nx, ny, nz, n, nh, forced, choice = case
    print(f"case host overlap {case}", flush=True)
    dev = torch.device("cuda:0")
    try:
        if forced:
            pct.force_tile_config(*forced)
        for k, scale in enumerate((1.0, 0.5, 1.0)):
            sx, sy = max(12, int(nx * scale)), max(12, int(ny * scale))
            img, amri, act = pct.synth_tissue(sx, sy, nz, n, 2, False, 40 + k)
            lab = act.copy()
            pct.normalize_labels(lab)
            host = pct.multi_patch_based_segmentation_gpu(img, amri, lab, nh, 2.5, 1.0, choice)
            out = torch.empty(img.shape, device=dev)
            pct.fuse_device(torch.from_numpy(img).to(dev), torch.from_numpy(amri).to(dev),
                            torch.from_numpy(lab).to(dev), out, nh, 2.5, 1.0, choice)
            torch.cuda.synchronize()
            assert np.array_equal(out.cpu().numpy().view(np.uint32), host.view(np.uint32)), k
    finally:
        pct.force_tile_config(-1, 0)


@pytest.mark.parametrize("choice", [pct.CHOICE_SM, pct.CHOICE_DIRECT])
def test_host_call_staging_of_large_buffers(gpu, choice):
    """Host buffers larger than the pinned staging ring (four 16 MiB pieces):
    75 MB volumes go over in five pieces each, in and out, and the output
    equals the device-buffer call bit for bit."""
    import torch
    nx, ny, nz, n, nh = 384, 384, 128, 2, 3
    print(f"case host staging {nx}x{ny}x{nz} n={n} NH={nh} choice={choice}", flush=True)
    rng = np.random.default_rng(17)
    img = rng.random((nz, nx, ny), dtype=np.float32)
    amri = rng.random((n, nz, nx, ny), dtype=np.float32)
    lab = (rng.random((n, nz, nx, ny), dtype=np.float32) * 255.0).astype(np.float32)
    host = pct.multi_patch_based_segmentation_gpu(img, amri, lab, nh, 2.5, 1.0, choice)
    dev = torch.device("cuda:0")
    out = torch.empty(img.shape, device=dev)
    pct.fuse_device(torch.from_numpy(img).to(dev), torch.from_numpy(amri).to(dev),
                    torch.from_numpy(lab).to(dev), out, nh, 2.5, 1.0, choice)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().view(np.uint32), host.view(np.uint32))
    # the post chain on the device, through the same staging
    hu = pct.pseudo_ct_hu(img, amri, lab, nh, 2.5, 1.0, choice)
    ref = pct.postprocess(host.copy(), nh)
    assert np.array_equal(hu.view(np.uint32), ref.view(np.uint32))


def test_device_entry_points(gpu):
    import torch
    img, amri, act = pct.synth_tissue(40, 66, 30, 3, 2, False, 21)
    lab = act.copy()
    pct.normalize_labels(lab)
    host = pct.multi_patch_based_segmentation_gpu(img, amri, lab, 9, 2.5, 1.0)
    dev = torch.device("cuda:0")
    ti, ta = torch.from_numpy(img).to(dev), torch.from_numpy(amri).to(dev)
    tl = torch.from_numpy(act.copy()).to(dev)
    pct.normalize_labels_device(tl)
    torch.cuda.synchronize()
    assert np.array_equal(tl.cpu().numpy().view(np.uint32), lab.view(np.uint32))
    out = torch.empty(img.shape, device=dev)
    pct.fuse_device(ti, ta, tl, out, 9, 2.5, 1.0)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().view(np.uint32), host.view(np.uint32))
    # caller-provided workspace
    L = _lib.lib()
    pb = pct.problem(img.shape, 3, 9, 2.5, 1.0)
    need = L.pct_fuse_workspace_size(_lib.ctypes.byref(pb), 2)
    ws = torch.empty(need // 4 + 64, dtype=torch.float32, device=dev)
    out2 = torch.zeros_like(out)
    s = torch.cuda.current_stream()
    rc = L.pct_fuse_dev_ws(_lib.ctypes.byref(pb), _lib.c_void_p(ti.data_ptr()),
                           _lib.c_void_p(ta.data_ptr()), _lib.c_void_p(tl.data_ptr()),
                           _lib.c_void_p(out2.data_ptr()), 2, _lib.c_void_p(ws.data_ptr()), need,
                           _lib.c_void_p(s.cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    assert torch.equal(out, out2)
    # too small a workspace is rejected before anything runs
    rc = L.pct_fuse_dev_ws(_lib.ctypes.byref(pb), _lib.c_void_p(ti.data_ptr()),
                           _lib.c_void_p(ta.data_ptr()), _lib.c_void_p(tl.data_ptr()),
                           _lib.c_void_p(out2.data_ptr()), 2, _lib.c_void_p(ws.data_ptr()),
                           need - 256, _lib.c_void_p(s.cuda_stream))
    assert rc == _lib.PCT_ERR_INVALID


def test_autotune_keeps_results(gpu):
    import torch
    img, amri, act = pct.synth_tissue(48, 70, 40, 4, 2, False, 5)
    lab = act.copy()
    pct.normalize_labels(lab)
    ref = oracle.fuse(img, amri, lab, 7, 2.5, 1.0)
    dev = torch.device("cuda:0")
    ti, ta, tl = (torch.from_numpy(a).to(dev) for a in (img, amri, lab))
    out = torch.empty(img.shape, device=dev)
    cfg = pct.autotune(ti, ta, tl, out, 7, 2.5, 1.0, reps=1)
    assert cfg["measured_ms"] > 0
    assert pct.tile_config(img.shape, 4, 7)["id"] == cfg["id"]
    pct.fuse_device(ti, ta, tl, out, 7, 2.5, 1.0)
    torch.cuda.synchronize()
    compare(out.cpu().numpy(), ref)


@pytest.mark.parametrize("choice", [pct.CHOICE_SM, pct.CHOICE_GM, pct.CHOICE_DIRECT])
def test_cli_end_to_end(gpu, tmp_path, choice):
    # EXEC_CHOICE 2 (SM) and 0 (GM) run the tiled kernel, 5 the direct one
    nx, ny, nz, n, nh = 26, 34, 24, 2, 5
    img, amri, act = pct.synth_tissue(nx, ny, nz, n, 2, False, 33)
    paths = [str(tmp_path / f) for f in ("atlas.txt", "label.txt", "im.txt")]
    pct.matrixio_write(paths[0], amri, n, nx, ny, nz)
    pct.matrixio_write(paths[1], act, n, nx, ny, nz)
    pct.matrixio_write(paths[2], img, 1, nx, ny, nz)
    cfg = str(tmp_path / "data.cfg")
    with open(cfg, "w") as f:
        f.write("0\n" + "\n".join(paths) + "\n")
    cli = os.path.join(os.path.dirname(_lib.LIB_PATH), "pct_cli")
    res = str(tmp_path / "result")
    r = subprocess.run([cli, cfg, res, str(choice), str(nh), "2.5", "1", "0"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    hu, dims = pct.matrixio_read(res + ".mat")
    assert dims == [nx, ny, nz, 1]
    # the reference pipeline on the same (text-rounded) inputs
    d = pct.read_config(cfg)
    lab = d["label"].copy()
    oracle.normalize_labels(lab)
    fused = oracle.fuse(d["img"], d["atlas"], lab, nh, 2.5, 1.0)
    ref_hu = oracle.postprocess(fused.copy(), nh)
    got = hu.reshape(nz, nx, ny)
    gpu_fused = pct.multi_patch_based_segmentation_gpu(d["img"], d["atlas"], lab, nh, 2.5, 1.0,
                                                       choice)
    compare(gpu_fused, fused)
    _final_hu_agrees(gpu_fused, fused, ref_hu, nh)
    assert np.array_equal(got, pct.postprocess(gpu_fused.copy(), nh))


def test_cli_gpu_test_time_sweep(gpu, tmp_path):
    # EXEC_CHOICE 4 over a window list: every (NH, atlas count) point is timed,
    # with the reference's bandwidth line, and the averaged table is written
    cfg = str(tmp_path / "rand.cfg")
    with open(cfg, "w") as f:
        f.write("1\n3\n20\n22\n18\n")          # random mode: num nx ny nz
    cli = os.path.join(os.path.dirname(_lib.LIB_PATH), "pct_cli")
    table = str(tmp_path / "final.txt")
    r = subprocess.run([cli, cfg, str(tmp_path / "r"), "4", "7", "2.5", "1", "0", "--sweep-nh",
                        "3,5", "--many", "2", "--results", table, "--seed", "4"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-3000:], r.stderr[-1000:])
    assert r.returncode == 0
    assert r.stdout.count("Effective BandWidth") == 2 * 3
    assert r.stdout.count("Elapsed Time") == 2 * 3 * 2 + 2 * 3
    lines = open(table).read().splitlines()
    assert lines[0] == "Num of executions to calculate a result 2 sampled 2 times"
    points = [(int(l.split()[1]), int(l.split()[3])) for l in lines[1:]]
    assert points == [(3, 3), (3, 2), (3, 1), (5, 3), (5, 2), (5, 1)]


def test_cli_mat_file_round_trip(gpu, tmp_path):
    # the MATLAB build's flow: inputs from a MAT-file, the HU result written
    # as a MAT-file variable named after the result
    import scipy.io
    nx, ny, nz, n, nh = 24, 30, 22, 2, 5
    img, amri, act = pct.synth_tissue(nx, ny, nz, n, 2, False, 51)

    def matlab(v):                                   # library (z, x, y) -> MATLAB (x, y, z)
        return np.ascontiguousarray(np.transpose(v, (1, 2, 0)))
    data = str(tmp_path / "subject.mat")
    scipy.io.savemat(data, {"img": matlab(img),
                            "atlas": {f"MRI{k + 1}": matlab(amri[k]) for k in range(n)},
                            "atlas_label": {f"MRI{k + 1}": matlab(act[k]) for k in range(n)}},
                     do_compression=True)
    cli = os.path.join(os.path.dirname(_lib.LIB_PATH), "pct_cli")
    res = str(tmp_path / "pseudo_ct")
    r = subprocess.run([cli, data, res, "2", str(nh), "2.5", "1", "0"], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    got = scipy.io.loadmat(res + ".mat")["pseudo_ct"]
    assert got.dtype == np.float32 and got.shape == (nx, ny, nz)
    lab = act.copy()
    pct.normalize_labels(lab)
    fused = pct.multi_patch_based_segmentation_gpu(img, amri, lab, nh, 2.5, 1.0)
    want = pct.postprocess(fused.copy(), nh)
    assert np.array_equal(np.transpose(got, (2, 0, 1)), want)


def test_cli_cpu_surface_equals_gpu_surface(gpu, tmp_path):
    # the CPU program's command line (data result EXEC_CHOICE NH TH sigma
    # beta) and the GPU program's (data result EXEC_CHOICE NH sigma beta)
    # give the same HU volume
    nx, ny, nz, n, nh = 30, 28, 24, 2, 7
    img, amri, act = pct.synth_tissue(nx, ny, nz, n, 2, False, 44)
    paths = [str(tmp_path / f) for f in ("a.txt", "l.txt", "i.txt")]
    pct.matrixio_write(paths[0], amri, n, nx, ny, nz)
    pct.matrixio_write(paths[1], act, n, nx, ny, nz)
    pct.matrixio_write(paths[2], img, 1, nx, ny, nz)
    cfg = str(tmp_path / "data.cfg")
    with open(cfg, "w") as f:
        f.write("0\n" + "\n".join(paths) + "\n")
    cli = os.path.join(os.path.dirname(_lib.LIB_PATH), "pct_cli")
    outs = []
    for k, argv in enumerate(([cli, "--cpu-surface", cfg, str(tmp_path / "cpu"), "1", str(nh),
                               "4", "2.5", "1"],
                              [cli, "--cpu-surface", cfg, str(tmp_path / "cpu2"), "2", str(nh),
                               "2", "2.5", "1"],
                              [cli, cfg, str(tmp_path / "gpu"), "2", str(nh), "2.5", "1"])):
        r = subprocess.run(argv, capture_output=True, text=True, timeout=300)
        print(r.stdout[-1500:], r.stderr[-1500:])
        assert r.returncode == 0
        if k == 1:   # SingleCoreTestTime: 10 samples of 10 calls
            assert r.stdout.count("Iter ") == 10
        hu, dims = pct.matrixio_read(argv[3 if k < 2 else 2] + ".mat")
        assert dims == [nx, ny, nz, 1]
        outs.append(hu)
    assert np.array_equal(outs[0], outs[2]) and np.array_equal(outs[1], outs[2])


def test_divisor_argument(gpu):
    # div given directly (the reference wrapper's argument, cu:1635) equals
    # the sigma / beta call bit for bit, and another div matches the oracle
    img, amri, act = pct.synth_tissue(30, 34, 24, 2, 2, False, 45)
    lab = act.copy()
    pct.normalize_labels(lab)
    a = pct.multi_patch_based_segmentation_gpu(img, amri, lab, 7, 2.5, 1.0)
    b = pct.multi_patch_based_segmentation_gpu(img, amri, lab, 7, 0.0, 0.0,
                                               div=pct.weight_divisor(2.5, 1.0))
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    div = 40.0
    got = pct.multi_patch_based_segmentation_gpu(img, amri, lab, 7, div=div)
    # sigma with 2 * 27 * sigma^2 = div, beta 1
    ref = oracle.fuse(img, amri, lab, 7, float(np.sqrt(div / 54.0)), 1.0)
    compare(got, ref)
