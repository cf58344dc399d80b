"""The plain path at a size whose atlas arrays exceed 2^31 elements, so every
atlas offset, workspace offset and output index needs 64-bit arithmetic.

1024 x 1024 x 256 voxels, n = 10, NH = 5: about 44 GB of device memory.  A
fused voxel depends only on the voxels within NH/2 + 1 of it, so each
sampled voxel is checked against the FP64 oracle run on a crop around it,
where it is interior.  Samples sit at the ends of the volume, where the
offsets are largest.
"""
import numpy as np
import pytest

import oracle
import pseudoctimaging_amd as pct
from parity import compare

pytestmark = pytest.mark.gpu


def test_atlas_arrays_beyond_2_31_elements(gpu):
    import torch
    nz, nx, ny, n, nh = 256, 1024, 1024, 10, 5
    hw = (nh - 1) // 2
    r = hw + 1
    assert nz * nx * ny * n > 2 ** 31
    dev = "cuda:0"
    print(f"case large {nx}x{ny}x{nz} n={n} NH={nh}: generating on the device", flush=True)
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    img = torch.rand((nz, nx, ny), generator=g, device=dev)
    amri = torch.rand((n, nz, nx, ny), generator=g, device=dev)
    lab = torch.rand((n, nz, nx, ny), generator=g, device=dev) * 255.0
    out = torch.full((nz, nx, ny), -1.0, device=dev)
    print("fusing", flush=True)
    pct.fuse_device(img, amri, lab, out, nh, 2.5, 1.0)
    torch.cuda.synchronize()
    print("checking samples", flush=True)
    rng = np.random.default_rng(3)
    samples = [(r, r, r), (nz - r - 1, nx - r - 1, ny - r - 1), (nz - r - 1, r, ny - r - 1),
               (r, nx - r - 1, r)]
    samples += [(int(rng.integers(r, nz - r)), int(rng.integers(r, nx - r)),
                 int(rng.integers(r, ny - r))) for _ in range(12)]
    got, ref = [], []
    R = r   # crop radius: the sample is interior in the crop
    for z, x, y in samples:
        zs, xs, ys = slice(z - R, z + R + 1), slice(x - R, x + R + 1), slice(y - R, y + R + 1)
        ci = np.ascontiguousarray(img[zs, xs, ys].cpu().numpy())
        ca = np.ascontiguousarray(amri[:, zs, xs, ys].cpu().numpy())
        cl = np.ascontiguousarray(lab[:, zs, xs, ys].cpu().numpy())
        c = 2 * R + 1
        ref.append(float(oracle.voxels(ci, ca, cl, nh, 2.5, 1.0, [(R * c + R) * c + R])[0]))
        got.append(float(out[z, x, y].item()))
    st = compare(np.array(got, np.float32), np.array(ref, np.float64))
    print(st, flush=True)
    # the border is zero on every face, the far ones included
    for face in (out[0], out[nz - 1], out[:, 0], out[:, nx - 1], out[:, :, 0], out[:, :, ny - 1]):
        assert float(face.abs().max().item()) == 0.0
    assert float(out[r:nz - r, r:nx - r, r:ny - r].min().item()) > 0.0
