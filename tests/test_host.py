"""Host-side code of libpct (no GPU): MatrixIO, the configuration file,
generators, pre/post-processing, the CLI surface.  The reference's own
functions (oracle/_ref) are the yardstick where they exist."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import pseudoctimaging_amd as pct
from pseudoctimaging_amd import _lib

CLI = os.path.join(os.path.dirname(_lib.LIB_PATH), "pct_cli")


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


# ---- MatrixIO ---------------------------------------------------------------

def test_matrixio_write_matches_reference_bytes(ref, tmp_path):
    rng = np.random.default_rng(0)
    a = (rng.standard_normal(2 * 3 * 4 * 5) * 1000).astype(np.float32)
    a[:5] = [0.0, -0.0, 1e-7, 123456.78, -3.5]
    p1, p2 = str(tmp_path / "ours.txt"), str(tmp_path / "ref.txt")
    pct.matrixio_write(p1, a, 2, 3, 4, 5)
    oracle.ref_matrixio_store(p2, a, 2, 3, 4, 5)
    assert open(p1, "rb").read() == open(p2, "rb").read()


def test_matrixio_read_matches_reference(ref, tmp_path):
    rng = np.random.default_rng(1)
    a = (rng.random(3 * 4 * 5 * 2) * 255).astype(np.float32)
    p = str(tmp_path / "m.txt")
    oracle.ref_matrixio_store(p, a, 2, 3, 4, 5)
    ours, dims = pct.matrixio_read(p)
    theirs, rdims = oracle.ref_matrixio_read(p)
    assert dims == rdims == [3, 4, 5, 2]
    assert np.array_equal(bits(ours), bits(theirs))
    # 6 significant digits survive the text round trip
    np.testing.assert_allclose(ours, a, rtol=1e-5)


def test_matrixio_read_at_scale_matches_reference(ref, tmp_path):
    """Two million values over many lines, mixed blanks, CRLF and '+' signs:
    the parallel parser cuts this text into many chunks."""
    rng = np.random.default_rng(2)
    a = (rng.standard_normal(2_000_000) * 300).astype(np.float32)
    toks = [f"{v:.9g}" for v in a[:1000]] + [f"{v:g}" for v in a[1000:]]
    pos = int(np.nonzero(a[:1000] > 0)[0][0])
    toks[pos] = "+" + toks[pos]
    seps = rng.choice([" ", "  ", "\t", "\n", "\r\n", " \n"], size=len(toks))
    p = str(tmp_path / "big.txt")
    with open(p, "w", newline="") as f:
        f.write("2\n100\n100\n100\n")
        f.write("".join(t + s for t, s in zip(toks, seps)))
    ours, dims = pct.matrixio_read(p)
    theirs, rdims = oracle.ref_matrixio_read(p)
    assert dims == rdims == [100, 100, 100, 2]
    assert np.array_equal(bits(ours), bits(theirs))


def test_matrixio_sidecar(tmp_path):
    rng = np.random.default_rng(3)
    a = (rng.random(4 * 5 * 6 * 2) * 255).astype(np.float32)
    p = str(tmp_path / "m.txt")
    pct.matrixio_write(p, a, 2, 4, 5, 6)
    text, dims = pct.matrixio_read(p)
    pct.matrixio_write_sidecar(p)
    side = p + ".pctbin"
    assert os.path.exists(side) and not os.path.exists(side + ".tmp")
    got, gdims = pct.matrixio_read(p)
    assert gdims == dims and np.array_equal(bits(got), bits(text))
    # the sidecar's values are what is read while the text is unchanged
    raw = bytearray(open(side, "rb").read())
    payload = np.frombuffer(bytes(raw[-4 * a.size:]), np.float32).copy()
    payload[7] = 12345.0
    raw[-4 * a.size:] = payload.tobytes()
    open(side, "wb").write(bytes(raw))
    got, _ = pct.matrixio_read(p)
    assert got[7] == 12345.0 and np.array_equal(np.delete(got, 7), np.delete(text, 7))
    # a changed text file makes the sidecar stale
    b = a * 0.5
    pct.matrixio_write(p, b, 2, 4, 5, 6)
    st = os.stat(p)
    os.utime(p, ns=(st.st_atime_ns, st.st_mtime_ns + 1_000_000_000))
    got, _ = pct.matrixio_read(p)
    np.testing.assert_allclose(got, b, rtol=1e-5)
    # a truncated sidecar is ignored
    pct.matrixio_write_sidecar(p)
    open(side, "r+b").truncate(100)
    got2, _ = pct.matrixio_read(p)
    assert np.array_equal(bits(got2), bits(got))
    with pytest.raises(pct.PctError) as e:
        pct.matrixio_write_sidecar(str(tmp_path / "missing.txt"))
    assert e.value.status == _lib.PCT_ERR_IO


def test_matrixio_errors(tmp_path):
    p = str(tmp_path / "bad.txt")
    with open(p, "w") as f:
        f.write("1\n2\n2\n2\n1 2 3\n")          # 3 values for a 2x2x2 volume
    with pytest.raises(pct.PctError) as e:
        pct.matrixio_read(p)
    assert e.value.status == _lib.PCT_ERR_FORMAT
    with pytest.raises(pct.PctError) as e:
        pct.matrixio_read(str(tmp_path / "missing.txt"))
    assert e.value.status == _lib.PCT_ERR_IO
    with open(p, "w") as f:
        f.write("1\n1\n1\n2\n1 x\n")
    with pytest.raises(pct.PctError):
        pct.matrixio_read(p)


def test_matrixio_header_products_that_wrap_are_refused(tmp_path):
    # 65536^4 = 2^64 wraps a 64-bit product to 0: the reader must refuse the
    # header before it allocates, not parse zero values into a zero-byte
    # buffer and report 65536^3 dims
    p = str(tmp_path / "wrap.txt")
    with open(p, "w") as f:
        f.write("65536\n65536\n65536\n65536\n")
    with pytest.raises(pct.PctError) as e:
        pct.matrixio_read(p)
    assert e.value.status == _lib.PCT_ERR_TOO_LARGE
    # one volume of 2^31 elements is refused too, however few atlases
    with open(p, "w") as f:
        f.write("1\n2048\n1024\n1024\n")
    with pytest.raises(pct.PctError) as e:
        pct.matrixio_read(p)
    assert e.value.status == _lib.PCT_ERR_TOO_LARGE
    # a sidecar with a wrapping header is ignored; the text is read instead
    a = np.arange(8, dtype=np.float32)
    pct.matrixio_write(p, a, 1, 2, 2, 2)
    pct.matrixio_write_sidecar(p)
    side = p + ".pctbin"
    raw = bytearray(open(side, "rb").read())
    hdr = np.frombuffer(bytes(raw[8:40]), np.int64).copy()
    hdr[:] = 65536
    raw[8:40] = hdr.tobytes()
    open(side, "wb").write(bytes(raw))
    got, dims = pct.matrixio_read(p)
    assert dims == [2, 2, 2, 1] and np.array_equal(got, a)


# ---- configuration file (cpp:908-1018) --------------------------------------

def _write_set(tmp_path, img, atlas, label):
    n, nz, nx, ny = atlas.shape
    pa, pl, pi = (str(tmp_path / f) for f in ("atlas.txt", "label.txt", "im.txt"))
    # MatrixIO dims are num, rows (nx), cols (ny), depth (nz), values in memory order
    pct.matrixio_write(pa, atlas, n, nx, ny, nz)
    pct.matrixio_write(pl, label, n, nx, ny, nz)
    pct.matrixio_write(pi, img, 1, nx, ny, nz)
    cfg = str(tmp_path / "data.cfg")
    with open(cfg, "w") as f:
        f.write(f"0\n{pa}\n{pl}\n{pi}\n")
    return cfg


def test_read_config_text_mode_matches_reference(ref, tmp_path):
    img, amri, act = pct.synth_tissue(7, 8, 6, 2, 1, False, 3)
    cfg = _write_set(tmp_path, img, amri, act)
    ours = pct.read_config(cfg)
    theirs = oracle.ref_read_config(cfg)
    for k in ("img", "atlas", "label"):
        assert ours[k].shape == theirs[k].shape
        assert np.array_equal(bits(ours[k]), bits(theirs[k]))


def test_read_config_dimension_mismatch(tmp_path):
    img, amri, act = pct.synth_tissue(7, 8, 6, 2, 1, False, 3)
    cfg = _write_set(tmp_path, img, amri, act)
    pct.matrixio_write(str(tmp_path / "im.txt"), np.zeros(6 * 7 * 9, np.float32), 1, 7, 9, 6)
    with pytest.raises(pct.PctError) as e:
        pct.read_config(cfg)
    assert e.value.status == _lib.PCT_ERR_DIMS


def test_read_config_random_mode(tmp_path):
    cfg = str(tmp_path / "rand.cfg")
    with open(cfg, "w") as f:
        f.write("1\n3\n9\n10\n11\n")   # random flag, num, nx, ny, nz
    a = pct.read_config(cfg, seed=5)
    b = pct.read_config(cfg, seed=5)
    c = pct.read_config(cfg, seed=6)
    assert a["img"].shape == (11, 9, 10) and a["atlas"].shape == (3, 11, 9, 10)
    for k in ("img", "atlas", "label"):
        assert np.array_equal(a[k], b[k])
        assert not np.array_equal(a[k], c[k])
        assert a[k].min() >= 0 and a[k].max() < 10
    assert abs(a["atlas"].mean() - 5) < 0.2


# ---- generators ---------------------------------------------------------------

def test_tissue_generator_properties():
    img, amri, act = pct.synth_tissue(40, 44, 36, 3, 2, False, 11)
    img2, amri2, act2 = pct.synth_tissue(40, 44, 36, 3, 2, False, 11)
    assert np.array_equal(bits(img), bits(img2)) and np.array_equal(bits(act), bits(act2))
    # four classes at the quartiles: CT class means -1000/-100/40/1000, sd 20
    cls = np.digitize(act[0], [-550, -30, 520])
    frac = np.bincount(cls.ravel(), minlength=4) / cls.size
    assert np.all(np.abs(frac - 0.25) < 0.06), frac
    for c, m in enumerate([-1000, -100, 40, 1000]):
        v = act[0][cls == c]
        assert abs(v.mean() - m) < 2 and abs(v.std() - 20) < 2
    # MRI class means 0/0.8/0.5/0.2 with sd 0.03, same anatomy as the CT
    for c, m in enumerate([0.0, 0.8, 0.5, 0.2]):
        v = amri[0][cls == c]
        assert abs(v.mean() - m) < 0.01 and abs(v.std() - 0.03) < 0.005


def test_shared_anatomy_is_a_shift_of_the_target():
    img, amri, act = pct.synth_tissue(30, 32, 28, 2, 2, False, 4)
    tcls = np.digitize(img, [0.1, 0.35, 0.65])      # MRI: 0 air, 0.2 bone, 0.5 soft, 0.8 fat
    acls = np.digitize(amri[0], [0.1, 0.35, 0.65])
    best = max(((dz, dx, dy) for dz in range(-2, 3) for dx in range(-2, 3) for dy in range(-2, 3)),
               key=lambda s: np.mean(acls[4:-4, 4:-4, 4:-4] ==
                                     tcls[4 + s[0]:24 + s[0], 4 + s[1]:26 + s[1],
                                          4 + s[2]:28 + s[2]]))
    dz, dx, dy = best
    agree = np.mean(acls[4:-4, 4:-4, 4:-4] == tcls[4 + dz:24 + dz, 4 + dx:26 + dx, 4 + dy:28 + dy])
    assert agree > 0.99
    ind = pct.synth_tissue(30, 32, 28, 2, 2, True, 4)[1]
    iagree = np.mean(np.digitize(ind[0], [0.1, 0.35, 0.65]) == tcls)
    assert iagree < 0.6


def test_uniform_generator():
    img, atlas, label = pct.synth_uniform(8, 9, 10, 2, 7)
    assert img.shape == (10, 8, 9) and atlas.shape == (2, 10, 8, 9)
    assert img.min() >= 0 and img.max() < 10


# ---- pre/post-processing -----------------------------------------------------

def test_normalize_and_divisor_match_oracle():
    rng = np.random.default_rng(2)
    v = (rng.random(10000) * 4095 - 1024).astype(np.float32)
    a, b = v.copy(), v.copy()
    pct.normalize_labels(a)
    oracle.normalize_labels(b)
    assert np.array_equal(bits(a), bits(b))
    for s, be in [(2.5, 1.0), (1.0, 1.0), (0.3, 2.0)]:
        assert pct.weight_divisor(s, be) == oracle.div(s, be)


@pytest.mark.parametrize("name", ["small_tissue.npz", "small_uniform.npz"])
def test_postprocess_matches_golden(name):
    g = dict(np.load(os.path.join(os.path.dirname(__file__), "golden", name)))
    nh = int(g["params"][4])
    hu = pct.postprocess(g["fused"].copy(), nh)
    assert np.array_equal(bits(hu), bits(g["hu"]))


def test_regularize_stops_without_progress():
    f = np.full((8, 8, 8), np.nan, np.float32)
    passes = pct.regularization_step(f, 3)   # nothing can ever be filled
    assert passes == 1 and np.isnan(f[2:-2, 2:-2, 2:-2]).all()


# ---- CLI surface ----------------------------------------------------------------

def test_cli_usage_without_arguments():
    r = subprocess.run([CLI], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "EXEC_CHOICE" in r.stdout


def test_cli_fails_cleanly_without_device(tmp_path):
    if pct.device_count() > 0:
        pytest.skip("a GPU is present")
    cfg = str(tmp_path / "rand.cfg")
    with open(cfg, "w") as f:
        f.write("1\n1\n12\n12\n12\n")
    r = subprocess.run([CLI, cfg, str(tmp_path / "out"), "2", "3"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode != 0 and "gfx950" in r.stderr


# ---- the CPU program's argument surface (cpp:1484-1565) ----------------------

def _cli():
    return os.path.join(os.path.dirname(_lib.LIB_PATH), "pct_cli")


def test_cli_cpu_surface_parses_threads_sigma_beta(tmp_path):
    # data result EXEC_CHOICE NH [TH] [sigma beta]: sigma and beta only with
    # TH (cpp:1552-1558); the reference GPU surface would read TH as sigma
    cfg = str(tmp_path / "rand.cfg")
    with open(cfg, "w") as f:
        f.write("1\n2\n12\n11\n10\n")
    cases = [(["1", "11", "12", "2.5", "1"], ("MultiCore", "11", "12", "2.5", "1")),
             (["0", "7", "4"], ("SingleCore", "7", "4", "1", "1")),
             (["3", "9"], ("MultiCoreTestTime", "9", "1", "1", "1"))]
    for argv, (choice, nh, th, sigma, beta) in cases:
        r = subprocess.run([_cli(), "--cpu-surface", cfg, str(tmp_path / "res")] + argv,
                           capture_output=True, text=True, timeout=60, cwd=str(tmp_path))
        out = r.stdout
        assert f"[3] Choice {choice}\n" in out, out
        assert f"[4] NH {nh}\n[5] TH {th}\n[6] sigma {sigma}\n[7] beta {beta}\n" in out, out
        # no GPU here: the run stops at the device, after reading the data
        if pct.device_count() == 0:
            assert r.returncode != 0 and "device 0" in r.stderr


def test_cli_cpu_surface_dumps_random_inputs(tmp_path):
    # random mode writes the generated inputs before the CT normalisation
    # (cpp:974-978): atlas.txt and label.txt with num volumes, im.txt with one
    cfg = str(tmp_path / "rand.cfg")
    with open(cfg, "w") as f:
        f.write("1\n2\n12\n11\n10\n")
    subprocess.run([_cli(), "--cpu-surface", cfg, str(tmp_path / "res"), "1", "5", "2",
                    "--seed", "5"], capture_output=True, text=True, timeout=60,
                   cwd=str(tmp_path))
    img, atlas, label = pct.synth_uniform(12, 11, 10, 2, 5)
    for name, want, num in (("atlas.txt", atlas, 2), ("label.txt", label, 2), ("im.txt", img, 1)):
        got, dims = pct.matrixio_read(str(tmp_path / name))
        assert dims == [12, 11, 10, num]
        # the text keeps 6 significant digits, as the reference's Store
        np.testing.assert_allclose(got, want.reshape(-1), rtol=5e-6, atol=1e-6)
    # the GPU surface does not write them (cu:1446-1449 has the writes commented out)
    other = tmp_path / "gpu"
    other.mkdir()
    subprocess.run([_cli(), cfg, str(other / "res"), "2", "5"], capture_output=True, text=True,
                   timeout=60, cwd=str(other))
    assert not (other / "atlas.txt").exists()


def test_problem_divisor_validation():
    from ctypes import byref
    L = _lib.lib()
    ok = pct.problem((20, 20, 20), 1, 5, 0.0, 0.0, div=337.5)
    assert L.pct_fuse_workspace_size(byref(ok), pct.CHOICE_SM) > 0
    for bad in (pct.problem((20, 20, 20), 1, 5, 0.0, 0.0, div=0.0),
                pct.problem((20, 20, 20), 1, 5, 2.5, 1.0, div=-1.0),
                pct.problem((20, 20, 20), 1, 5, 2.5, 1.0, div=float("nan")),
                pct.problem((20, 20, 20), 1, 5, 2.5, 1.0, div=float("inf"))):
        assert L.pct_fuse_workspace_size(byref(bad), pct.CHOICE_SM) == 0
