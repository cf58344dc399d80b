"""MAT-file (level 5) input and output without libmat, checked against
scipy.io, an independent reader and writer of the format.

The reference's MATLAB build reads `atlas` / `atlas_label` structs and the
target image from a MAT-file (src/CPU/NewPatchedBrainPureC.cpp:584-801) and
converts MATLAB's column-major arrays so that element (x, y, z) of an
nx x ny x nz array lands at (z*nx + x)*ny + y (cpp:377-474)."""
import struct

import numpy as np
import pytest
import scipy.io

import pseudoctimaging_amd as pct
from pseudoctimaging_amd import _lib


def _volumes(nx=7, ny=5, nz=4, n=3, seed=0, dtype=np.float32):
    rng = np.random.default_rng(seed)
    img = rng.normal(0.5, 0.2, (nx, ny, nz)).astype(dtype)          # MATLAB (x, y, z)
    atlas = [rng.normal(0.5, 0.2, (nx, ny, nz)).astype(dtype) for _ in range(n)]
    label = [rng.normal(40, 300, (nx, ny, nz)).astype(dtype) for _ in range(n)]
    return img, atlas, label


def _save(path, img, atlas, label, compress, img_name="img"):
    mdict = {img_name: img,
             "atlas": {f"MRI{i + 1}": a for i, a in enumerate(atlas)},
             "atlas_label": {f"MRI{i + 1}": a for i, a in enumerate(label)}}
    scipy.io.savemat(path, mdict, do_compression=compress, format="5")


def _layout(vol):
    """MATLAB (x, y, z) array -> library (z, x, y) layout."""
    return np.ascontiguousarray(np.transpose(vol, (2, 0, 1))).astype(np.float32)


@pytest.mark.parametrize("compress", [False, True])
def test_read_scipy_written_file(tmp_path, compress):
    img, atlas, label = _volumes()
    p = str(tmp_path / "in.mat")
    _save(p, img, atlas, label, compress)
    d = pct.mat_read(p)
    assert d["img"].shape == (4, 7, 5) and d["atlas"].shape == (3, 4, 7, 5)
    assert np.array_equal(d["img"], _layout(img))
    for k in range(3):
        assert np.array_equal(d["atlas"][k], _layout(atlas[k]))
        assert np.array_equal(d["label"][k], _layout(label[k]))
    # index formula of the reference's conversion, element by element
    x, y, z = 5, 3, 2
    assert d["img"][z, x, y] == img[x, y, z]
    # the configuration entry point takes a .mat path as the MATLAB build does
    c = pct.read_config(p)
    assert np.array_equal(c["img"], d["img"]) and np.array_equal(c["label"], d["label"])


def test_numeric_classes_are_converted(tmp_path):
    img, atlas, label = _volumes(dtype=np.float64)
    atlas = [np.round(a * 100).astype(np.int16) for a in atlas]
    p = str(tmp_path / "cls.mat")
    _save(p, img, atlas, label, True, img_name="target")
    d = pct.mat_read(p)
    assert np.array_equal(d["img"], _layout(img.astype(np.float32)))
    assert np.array_equal(d["atlas"][1], _layout(atlas[1].astype(np.float32)))


def test_write_is_read_by_scipy(tmp_path):
    rng = np.random.default_rng(4)
    vol = rng.normal(0, 500, (6, 9, 8)).astype(np.float32)     # library (nz, nx, ny)
    p = str(tmp_path / "out.mat")
    pct.mat_write(p, "result", vol)
    m = scipy.io.loadmat(p)
    a = m["result"]
    assert a.dtype == np.float32 and a.shape == (9, 8, 6)        # MATLAB nx x ny x nz
    assert np.array_equal(_layout(a), vol)


def test_round_trip_through_own_reader(tmp_path):
    img, atlas, label = _volumes(nx=6, ny=6, nz=3, n=1, seed=3)
    p = str(tmp_path / "rt.mat")
    _save(p, img, atlas, label, False)
    d = pct.mat_read(p)
    q = str(tmp_path / "img_again.mat")
    pct.mat_write(q, "img", d["img"])
    assert np.array_equal(scipy.io.loadmat(q)["img"], img)


def test_format_and_shape_errors(tmp_path):
    img, atlas, label = _volumes(n=2)
    bad = str(tmp_path / "shape.mat")
    _save(bad, img, atlas, [label[0], label[1][:, :, :3]], False)
    with pytest.raises(pct.PctError) as e:
        pct.mat_read(bad)
    assert e.value.status == _lib.PCT_ERR_DIMS
    count = str(tmp_path / "count.mat")
    _save(count, img, atlas, label[:1], False)
    with pytest.raises(pct.PctError) as e:
        pct.mat_read(count)
    assert e.value.status == _lib.PCT_ERR_DIMS
    missing = str(tmp_path / "missing.mat")
    scipy.io.savemat(missing, {"img": img})
    with pytest.raises(pct.PctError) as e:
        pct.mat_read(missing)
    assert e.value.status == _lib.PCT_ERR_FORMAT
    junk = str(tmp_path / "junk.mat")
    with open(junk, "wb") as f:
        f.write(b"not a mat file" * 20)
    with pytest.raises(pct.PctError) as e:
        pct.mat_read(junk)
    assert e.value.status == _lib.PCT_ERR_FORMAT
    truncated = str(tmp_path / "trunc.mat")
    _save(truncated, img, atlas, label, True)
    raw = open(truncated, "rb").read()
    with open(truncated, "wb") as f:
        f.write(raw[:len(raw) // 2])
    with pytest.raises(pct.PctError) as e:
        pct.mat_read(truncated)
    assert e.value.status == _lib.PCT_ERR_FORMAT
    with pytest.raises(pct.PctError) as e:
        pct.mat_read(str(tmp_path / "absent.mat"))
    assert e.value.status == _lib.PCT_ERR_IO


def _element(mi_type, payload):
    pad = (-len(payload)) % 8
    return struct.pack("<II", mi_type, len(payload)) + payload + b"\0" * pad


def _matrix(name, cls, dims, payload=b"", fields=None):
    """One miMATRIX element written by hand, so that its dims can be any
    int32 values whatever its payload holds."""
    body = _element(6, struct.pack("<II", cls, 0))                  # array flags
    body += _element(5, struct.pack("<%di" % len(dims), *dims))      # dimensions
    body += _element(1, name.encode())                               # name
    if fields is not None:
        flen = 32
        body += _element(5, struct.pack("<i", flen))
        body += _element(1, b"".join(f.encode().ljust(flen, b"\0") for f, _ in fields))
        for _, sub in fields:
            body += sub
    else:
        body += _element(7, payload)                                 # miSINGLE values
    return _element(14, body)


def _mat_file(path, elements):
    head = b"MATLAB 5.0 MAT-file, crafted".ljust(116, b" ") + b"\0" * 8
    head += struct.pack("<H", 0x0100) + b"IM"
    with open(path, "wb") as f:
        f.write(head + b"".join(elements))


def test_dims_whose_product_wraps_are_refused(tmp_path):
    # 2^22 * 2^21 * 2^21 = 2^64: a size_t product wraps to 0, so an unchecked
    # reader would accept an empty payload and then fill 2^64 elements
    dims = [1 << 22, 1 << 21, 1 << 21]
    vol = _matrix("", 7, dims)
    path = str(tmp_path / "wrap.mat")
    _mat_file(path, [_matrix("img", 7, dims),
                     _matrix("atlas", 2, [1, 1], fields=[("MRI1", vol)]),
                     _matrix("atlas_label", 2, [1, 1], fields=[("CT1", vol)])])
    with pytest.raises(pct.PctError) as e:
        pct.mat_read(path)
    assert e.value.status in (_lib.PCT_ERR_FORMAT, _lib.PCT_ERR_TOO_LARGE)
    # a product that fits 64 bits but exceeds one volume's limit
    dims = [2048, 1024, 1024]
    vol = _matrix("", 7, dims)
    _mat_file(path, [_matrix("img", 7, dims),
                     _matrix("atlas", 2, [1, 1], fields=[("MRI1", vol)]),
                     _matrix("atlas_label", 2, [1, 1], fields=[("CT1", vol)])])
    with pytest.raises(pct.PctError) as e:
        pct.mat_read(path)
    assert e.value.status in (_lib.PCT_ERR_FORMAT, _lib.PCT_ERR_TOO_LARGE)


def test_crafted_writer_matches_scipy_reader(tmp_path):
    # the hand-written elements above are valid MAT-v5: a small crafted file
    # reads back through the library with the values it holds
    img = np.arange(24, dtype=np.float32).reshape(2, 3, 4)
    col = img.transpose(2, 1, 0).astype("<f4").tobytes()   # MATLAB order of a 2x3x4 array
    dims = [2, 3, 4]
    path = str(tmp_path / "ok.mat")
    _mat_file(path, [_matrix("img", 7, dims, col),
                     _matrix("atlas", 2, [1, 1], fields=[("MRI1", _matrix("", 7, dims, col))]),
                     _matrix("atlas_label", 2, [1, 1],
                             fields=[("CT1", _matrix("", 7, dims, col))])])
    assert np.array_equal(scipy.io.loadmat(path)["img"], img)
    d = pct.mat_read(path)
    assert np.array_equal(d["img"], _layout(img))
