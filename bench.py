#!/usr/bin/env python3
"""Benchmark of the multi-atlas patch fusion on one MI355X.

Workload (the headline configuration): target 222x222x112 float32, n=10
atlas MRI/CT pairs, patch 3x3x3, NH=11, sigma=2.5, beta=1, synthetic
tissue-class phantom (seeded).  One step = one fusion of the whole volume on
inputs resident in HBM (tiled SM kernel + finalize pass).  `value` is
full-volume voxels (222*222*112) per second.

Extra fields:
  roofline      the tiled fusion kernel against MI355X's FP32 vector peak:
                achieved = FP32 FLOP the kernel executes per launch (rocprof
                SQ_INSTS_VALU_* counters x 64 lanes, FMA counted twice, from
                profiles/, used only when the profile's source stamp matches
                the kernel sources built here) / kernel time; frac =
                achieved / peak; issue_frac = VALU instructions issued /
                the non-packed issue ceiling.  effective_tflops counts the
                reference's direct formulation instead (86 FLOP per
                candidate-voxel pair: 27 sub + 27 FMA + scale + exp + 3 for
                the accumulation), which the kernel's box-sum reuse beats.
                Kernel time from HIP events around each launch.
  cpu_baseline  the reference CPU path (oracle/_ref, compiled from the
                reference source) or the oracle port, OpenMP over z with a
                static schedule, timed on a crop of the same inputs that is
                one round of planes for the thread count, and scaled by the
                rounds and rows of the full volume.
  parity_sample float64 oracle vs the GPU output on sampled voxels.
  end_to_end    host buffers in and out around the fusion and the host post
                chain, the region the reference's GPU program times.
  learned_pruned, batched  the learned candidate path and the uneven batch.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
FLOPS_PER_CVP = 86
METRIC = ("pseudo-CT voxels/s @222\u00d7222\u00d7112 n=10 NH=11: plain vs learned-pruned "
          "@\u22645HU MAE")


# the files whose contents define the tiled fusion kernels: a profile is
# attached to a run only when it was taken of a build of the same contents
KERNEL_SOURCES = ("pseudoctimaging_amd/csrc/pct_tiled.h", "pseudoctimaging_amd/csrc/pct_fuse.hip",
                  "pseudoctimaging_amd/csrc/pct_learned.hip", "pseudoctimaging_amd/csrc/Makefile")
PROFILE = os.path.join("profiles", "fuse_tiled_222x222x112_n10_nh11_traffic.json")
VALU_ISSUE_PER_S = 1024 * 2.4e9 / 4   # wave64 FP32 instructions, non-packed


def kernel_source_stamp() -> str:
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(rel.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def profiled_counters(kernel: str):
    """The committed counter profile of `kernel` at 222x222x112, n=10, NH=11,
    and a note: (entry, None) when the profile was taken of these kernel
    sources and holds that kernel, else (None, why not)."""
    tfile = os.path.join(ROOT, PROFILE)
    if not os.path.exists(tfile):
        return None, f"no profile file {PROFILE}"
    with open(tfile) as f:
        profiles = json.load(f)
    stamp = kernel_source_stamp()
    kernels = profiles.get("kernels", {})
    if profiles.get("source_stamp") != stamp:
        return None, (f"profile taken of kernel sources {profiles.get('source_stamp')}, built "
                      f"here {stamp}")
    if kernel not in kernels:
        return None, f"the timed kernel {kernel} is not among the profiled kernels {sorted(kernels)}"
    return kernels[kernel], None


def counter_roofline(entry, kernel_ms):
    """achieved / frac / issue_frac of one launch of `kernel_ms` from a
    profile entry's VALU counters."""
    v = entry["valu"]
    executed = v["executed_fp32_flop_per_launch"] / (kernel_ms * 1e-3) / 1e12
    return {"achieved": round(executed, 3), "frac": round(executed / PEAK_FP32_TFLOPS, 4),
            "issue_frac": round((v["SQ_INSTS_VALU"] - v["SQ_INSTS_VALU_TRANS_F32"])
                                / (kernel_ms * 1e-3) / VALU_ISSUE_PER_S, 4)}


# tile shape of the pruned kernel by shape id: (W, BX, BZ, waves)
PRUNED_SHAPES = {13: (32, 2, 4, 8), 8: (32, 4, 2, 4)}


def host_cpu_info() -> dict:
    """CPUs this process may run on and the host's topology (lscpu)."""
    info = {"affinity_cpus": len(os.sched_getaffinity(0)), "logical_cpus": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = dict(line.split(":", 1) for line in out.splitlines() if ":" in line)
        kv = {k.strip(): v.strip() for k, v in kv.items()}
        info["model"] = kv.get("Model name")
        sockets = int(kv.get("Socket(s)", "0") or 0)
        cores = int(kv.get("Core(s) per socket", "0") or 0)
        info["physical_cores"] = sockets * cores if sockets and cores else None
        info["threads_per_core"] = int(kv.get("Thread(s) per core", "0") or 0) or None
    except (OSError, ValueError, subprocess.SubprocessError):
        info["model"] = None
    return info


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def interior(n: int, r: int) -> int:
    return max(0, n - 2 * r)


def cpu_baseline(img, amri, lab, nh, sigma, beta, planes, rows, threads):
    """Times the reference fusion on a crop: `planes` interior z planes and
    `rows` interior x rows of the full-size inputs (all y)."""
    import numpy as np

    import oracle
    r = (nh - 1) // 2 + 1
    zc, xc = planes + 2 * r, rows + 2 * r
    cimg = np.ascontiguousarray(img[:zc, :xc, :])
    camri = np.ascontiguousarray(amri[:, :zc, :xc, :])
    clab = np.ascontiguousarray(lab[:, :zc, :xc, :])
    kind = "reference" if oracle.ref_available() else "port"
    # the reference program prints progress to stdout; keep stdout for the
    # one JSON line by pointing file descriptor 1 at stderr meanwhile
    import ctypes
    libc = ctypes.CDLL(None)
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        t0 = time.perf_counter()
        if kind == "reference":
            oracle.ref_fuse(cimg, camri, clab, nh, sigma, beta, threads=threads)
        else:
            oracle.fuse(cimg, camri, clab, nh, sigma, beta, threads=threads)
        dt = time.perf_counter() - t0
    finally:
        libc.fflush(None)
        os.dup2(saved, 1)
        os.close(saved)
    vox_sample = planes * rows * interior(img.shape[2], r)
    return kind, dt, vox_sample


# The uneven batch: (nx, ny, nz, atlases, NH) per subject
BATCH = [(160, 160, 96, 1, 7), (192, 192, 96, 3, 9), (222, 222, 112, 5, 11),
         (222, 222, 112, 7, 11), (256, 256, 128, 10, 9), (128, 128, 64, 10, 7)]


def batched_leg(steps: int, warmup: int, seed: int):
    """Six subjects of uneven size, atlas count and window in one
    graph-captured persistent launch; returns a result dict."""
    import numpy as np
    import torch

    import pseudoctimaging_amd as pct
    dev = torch.device("cuda:0")
    subs, vox, cvp = [], 0, 0
    for k, (nx, ny, nz, n, nh) in enumerate(BATCH):
        img, amri, act = pct.synth_tissue(nx, ny, nz, n, 3, False, seed + 100 + k)
        lab = act.copy()
        pct.normalize_labels(lab)
        subs.append({"img": torch.from_numpy(img).to(dev), "atlas": torch.from_numpy(amri).to(dev),
                     "label": torch.from_numpy(lab).to(dev),
                     "out": torch.empty(img.shape, device=dev), "nh": nh, "sigma": 2.5,
                     "beta": 1.0})
        r = (nh - 1) // 2 + 1
        vox += nx * ny * nz
        cvp += interior(nx, r) * interior(ny, r) * interior(nz, r) * n * nh ** 3
        del img, amri, act, lab
    tuned = pct.autotune_batch(subs, reps=2)
    choice = pct.batch_choice(subs)
    for _ in range(warmup):
        pct.fuse_batch(subs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        pct.fuse_batch(subs)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    finite = all(bool(torch.isfinite(s["out"]).all()) for s in subs)
    # the same subjects one call each, every call with its own tuned configuration
    seq = 0.0
    for s in subs:
        def one(s=s):
            pct.fuse_device(s["img"], s["atlas"], s["label"], s["out"], s["nh"], 2.5, 1.0)
        pct.autotune(s["img"], s["atlas"], s["label"], s["out"], s["nh"], 2.5, 1.0, reps=1)
        one()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            one()
        torch.cuda.synchronize()
        seq += (time.perf_counter() - t1) / steps
    return {"subjects": [list(b) for b in BATCH], "ms_per_batch": dt * 1e3,
            "voxels_per_s": vox / dt, "cvp_per_s": cvp / dt,
            "achieved_tflops_86_per_cvp": FLOPS_PER_CVP * cvp / dt / 1e12,
            "config": tuned,
            "per_class": {f"NH{2 * hw + 1}": {"shape": sid, "atlas_group": grp}
                          for hw, (sid, grp) in choice["classes"].items()},
            "classes_serial": choice["serial"],
            "sequential_tuned_ms": seq * 1e3,
            "outputs_finite": finite, "launch": "one hipGraph per batch"}


def learned_leg(d_img, d_amri, d_lab, plain_out, nh, sigma, beta, steps, warmup, seed,
                plain_ms, train_shape=(64, 96, 96), train_subjects=4, train_rounds=2,
                train_steps=60, topk=32, samples=27, margin=0.0, max_slabs=3, shape_id=-1,
                max_err=None, params=None):
    """Learned-pruned fusion on the resident inputs: the embedding MLP is
    trained on the GPU on `train_subjects` other seeded phantoms, then the
    pruned fusion is timed and compared with the plain output."""
    import numpy as np
    import torch

    import pseudoctimaging_amd as pct
    from pseudoctimaging_amd import learned
    dev = d_img.device
    n = d_amri.shape[0]
    if max_err is None:
        max_err = learned.DEFAULT_MAX_ERR
    tz, tx, ty = train_shape
    subjects = []
    for k in range(train_subjects):
        img, amri, _ = pct.synth_tissue(tx, ty, tz, n, 3, False, seed + 1000 + k)
        subjects.append((torch.from_numpy(img).to(dev), torch.from_numpy(amri).to(dev)))
    mu, inv_sd = learned.intensity_stats(d_img.cpu().numpy())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = []
    if params is None:
        params = learned.mlp_init(seed, mu, inv_sd)
        for rnd in range(train_rounds):
            for k, (ti, ta) in enumerate(subjects):
                params, trace = learned.mlp_train(ti, ta, nh, sigma, beta, params,
                                                  steps=train_steps, anchors=256, lr=3e-3,
                                                  seed=seed + 10 * rnd + k)
                losses.append([float(trace[0]), float(trace[-1])])
    train_s = time.perf_counter() - t0
    del subjects
    out = torch.empty_like(plain_out)
    kw = dict(topk=topk, samples=samples, margin=margin, max_slabs=max_slabs, shape_id=shape_id,
              max_err=max_err)
    frac = learned.fuse_pruned_device(d_img, d_amri, d_lab, out, nh, sigma, beta, params, **kw)
    for _ in range(warmup):
        learned.fuse_pruned_device(d_img, d_amri, d_lab, out, nh, sigma, beta, params,
                                   report_fallback=False, **kw)
    from pseudoctimaging_amd import _lib
    L = _lib.lib()
    L.pct_kernel_timing(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        learned.fuse_pruned_device(d_img, d_amri, d_lab, out, nh, sigma, beta, params,
                                   report_fallback=False, **kw)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    tot_ms, launches = _lib.c_double(), _lib.c_int()
    _lib.check(L.pct_kernel_timing_read(_lib.ctypes.byref(tot_ms), _lib.ctypes.byref(launches)))
    L.pct_kernel_timing(0)
    kernel_ms = tot_ms.value / max(1, launches.value)
    w, bx, bz, nw = PRUNED_SHAPES[shape_id if shape_id >= 0 else 13]
    kname = f"pruned_kernel<{w},{bx},{bz},{nw},{(nh - 1) // 2}>"
    roofline = {"bound": "valu_fp32", "kernel": kname, "kernel_ms": round(kernel_ms, 4),
                "launches": launches.value, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "achieved": None, "frac": None, "issue_frac": None}
    nz_, nx_, ny_ = plain_out.shape
    if (nz_, nx_, ny_, n, nh, topk, samples, max_slabs) == (112, 222, 222, 10, 11, 32, 27, 3):
        entry, why = profiled_counters(kname)
        if entry is not None:
            roofline.update(counter_roofline(entry, kernel_ms))
            roofline["counters_source"] = PROFILE
        else:
            roofline["counters_note"] = why
    a = out.cpu().numpy().astype(np.float64)
    b = plain_out.cpu().numpy().astype(np.float64)
    ok = np.isfinite(a) & np.isfinite(b)
    dhu = np.abs(4095.0 * (a[ok] - b[ok]) / 255.0)
    full = int(np.prod(plain_out.shape))
    return {"ms_per_step": ms, "voxels_per_s": full / (ms * 1e-3),
            "speedup_vs_plain": plain_ms / ms, "mae_hu_vs_plain": float(dhu.mean()),
            "max_err_hu_vs_plain": float(dhu.max()), "p99_err_hu_vs_plain":
            float(np.percentile(dhu, 99)),
            "nan_mismatch": int(np.count_nonzero(np.isnan(a) != np.isnan(b))),
            "fallback_frac": frac, "topk": topk, "lattice_samples": samples, "margin": margin,
            "max_err": max_err, "max_slabs": max_slabs, "params": params,
            "embedding": "30-64-64-16 MLP, BF16 MFMA inference, FP32 training on GPU",
            "train": {"subjects": train_subjects, "shape_zxy": list(train_shape),
                      "rounds": train_rounds, "steps_per_subject": train_steps,
                      "anchors_per_step": 256, "seconds": train_s,
                      "loss_first_last": losses},
            "timed": "embeddings of target and atlases + selection + pruned weighting "
                     "+ fallback + finalize, per call",
            "pruned_roofline": roofline}


def nan_fill_leg(seed, reps=3):
    """The post chain on a NaN-heavy full-size volume: the reference's random
    mode (uniform 10 U[0, 1) data, cpp:961-973) at 222x222x112, n=2, NH=3,
    sigma=0.33, where most voxels have no weight above DBL_EPSILON.  The
    device chain (pct_postprocess_dev: NaN list, wavefront-ordered fill,
    round, HU) against the host chain (pct_postprocess) on the same fused
    volume; both must agree bit for bit."""
    import numpy as np
    import torch

    import pseudoctimaging_amd as pct
    nx, ny, nz, n, nh, sigma = 222, 222, 112, 2, 3, 0.33
    img, amri, lab = pct.synth_uniform(nx, ny, nz, n, seed)
    pct.normalize_labels(lab)
    dev = torch.device("cuda:0")
    d = [torch.from_numpy(x).to(dev) for x in (img, amri, lab)]
    fused_d = torch.empty(img.shape, device=dev)
    pct.fuse_device(*d, fused_d, nh, sigma, 1.0)
    torch.cuda.synchronize()
    fused = fused_d.cpu().numpy()
    n_nan = int(np.isnan(fused).sum())
    t_dev = []
    for _ in range(reps):
        t = fused_d.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pct.postprocess_device(t, nh)
        torch.cuda.synchronize()
        t_dev.append(time.perf_counter() - t0)
    got = t.cpu().numpy()
    host = fused.copy()
    t0 = time.perf_counter()
    pct.postprocess(host, nh)
    t_host = time.perf_counter() - t0
    return {"shape_zxy": [nz, nx, ny], "n": n, "nh": nh, "sigma": sigma,
            "nan_voxels": n_nan, "nan_fraction": n_nan / fused.size,
            "device_chain_ms": float(np.median(t_dev)) * 1e3, "host_chain_ms": t_host * 1e3,
            "same_output": bool(np.array_equal(got.view(np.uint32), host.view(np.uint32))),
            "data": "random mode: uniform 10 U[0,1) target, atlas MRI and CT (seeded)"}


def end_to_end_leg(img, amri, lab, nh, sigma, beta, reps):
    """The region the reference's GPU program times around its SM call
    (src/GPU/PatchBasedPseudoCT.cu:2152-2188): host buffers in, three
    host-to-device copies, the fusion, the copy back and the post chain (NaN
    fill, round, HU) on the host.  The reference allocates device memory per
    call; pct_fuse_host keeps its device buffers between calls."""
    import numpy as np

    import pseudoctimaging_amd as pct
    nz, nx, ny = img.shape
    n = amri.shape[0]
    fused = pct.multi_patch_based_segmentation_gpu(img, amri, lab, nh, sigma, beta)   # warm-up
    pct.postprocess(fused, nh)
    t_fuse, t_post = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        fused = pct.multi_patch_based_segmentation_gpu(img, amri, lab, nh, sigma, beta)
        t1 = time.perf_counter()
        pct.postprocess(fused, nh)
        t2 = time.perf_counter()
        t_fuse.append(t1 - t0)
        t_post.append(t2 - t1)
    fuse_s, post_s = float(np.median(t_fuse)), float(np.median(t_post))
    # the same region in one call, the post chain on the device
    hu = pct.pseudo_ct_hu(img, amri, lab, nh, sigma, beta)   # warm-up
    t_one = []
    for _ in range(reps):
        t0 = time.perf_counter()
        hu = pct.pseudo_ct_hu(img, amri, lab, nh, sigma, beta)
        t_one.append(time.perf_counter() - t0)
    one_s = float(np.median(t_one))
    same = bool(np.array_equal(hu.view(np.uint32), fused.view(np.uint32)))
    # the reference's effective-bandwidth line (cu:2193-2196), on the whole region
    h = (nh - 1) // 2
    m = (nx - h) * (ny - h) * (nz - h)
    rb = 4.0 * m * (n * nh ** 3 + 2 * n * 27 * nh ** 3)
    wb = 4.0 * m
    total = fuse_s + post_s
    return {"fuse_host_ms": fuse_s * 1e3, "post_host_ms": post_s * 1e3,
            "total_ms": total * 1e3, "voxels_per_s": nx * ny * nz / total,
            "reference_effective_bandwidth_gbs": (rb + wb) / total / 1e9,
            "pseudo_ct_host_ms": one_s * 1e3, "pseudo_ct_host_same_output": same,
            "pcie_bytes": 4 * nx * ny * nz * (2 + 2 * n) + 4 * nx * ny * nz,
            "reps": reps, "statistic": "median",
            "region": "pct_fuse_host (3 H2D through pinned staging, overlapped with the "
                      "fusion; D2H; device buffers cached between calls, where the reference "
                      "allocates per call) + host NaN fill, round and HU rescale; "
                      "pseudo_ct_host_ms: the same region as pct_pseudo_ct_host, the post "
                      "chain on the device"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--shape", default="112,222,222", help="nz,nx,ny")
    ap.add_argument("--atlases", type=int, default=10)
    ap.add_argument("--nh", type=int, default=11)
    ap.add_argument("--sigma", type=float, default=2.5)
    ap.add_argument("--beta", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tile-shape", type=int, default=-1, metavar="ID",
                    help="run the plain leg with this tile shape (atlas group 1), no autotuning")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=96)
    ap.add_argument("--parity-voxels", type=int, default=2000)
    ap.add_argument("--no-batch", action="store_true", help="skip the uneven-batch leg")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-pruned leg")
    ap.add_argument("--no-end-to-end", action="store_true",
                    help="skip the host-buffer end-to-end leg")
    ap.add_argument("--no-nan-fill", action="store_true",
                    help="skip the NaN-heavy post-chain leg")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized float32 buffer three times (a plain copy "
                         "kernel of known size, to scale FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE "
              f"{world}); multi-GPU sharding of the fusion is a design note, not built.",
              file=sys.stderr)
        return 2

    # the CPUs this process may use, read before an OpenMP runtime binds the
    # main thread to one of them
    hostinfo = host_cpu_info()
    # the CPU baseline's OpenMP threads spread over those CPUs (read when the
    # OpenMP runtime starts, so before anything loads it)
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    import numpy as np
    import torch

    if args.copy_calibration > 0:
        nbytes = args.copy_calibration << 20
        a = torch.ones(nbytes // 4, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": nbytes, "bytes_written": nbytes,
                          "dispatches": 3}), flush=True)
        return 0

    import pseudoctimaging_amd as pct
    from pseudoctimaging_amd import _lib

    nz, nx, ny = (int(v) for v in args.shape.split(","))
    n, nh, sigma, beta = args.atlases, args.nh, args.sigma, args.beta
    r = (nh - 1) // 2 + 1
    n_int = interior(nz, r) * interior(nx, r) * interior(ny, r)
    cvp = n_int * n * nh ** 3
    full = nx * ny * nz

    log(f"generating {nx}x{ny}x{nz} n={n} tissue phantom (seed {args.seed})")
    t = time.perf_counter()
    img, amri, act = pct.synth_tissue(nx, ny, nz, n, 3, False, args.seed)
    lab = act.copy()
    pct.normalize_labels(lab)
    del act
    log(f"inputs ready in {time.perf_counter() - t:.1f}s")

    dev = torch.device("cuda:0")
    d_img = torch.from_numpy(img).to(dev)
    d_amri = torch.from_numpy(amri).to(dev)
    d_lab = torch.from_numpy(lab).to(dev)
    d_out = torch.empty(img.shape, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)
    torch.cuda.synchronize()

    tuned = None
    if args.tile_shape >= 0:
        pct.force_tile_config(args.tile_shape, 1)
        log(f"tile shape {args.tile_shape} forced")
    elif not args.no_autotune:
        log("autotuning tile configuration")
        tuned = pct.autotune(d_img, d_amri, d_lab, d_out, nh, sigma, beta, reps=1, stream=stream)
        log(f"tuned: {tuned}")
    cfg = pct.tile_config(img.shape, n, nh)

    for i in range(args.warmup):
        pct.fuse_device(d_img, d_amri, d_lab, d_out, nh, sigma, beta, pct.CHOICE_SM, stream)
    torch.cuda.synchronize()
    log(f"warmup {args.warmup} done; timing {args.steps} steps")

    L = _lib.lib()
    L.pct_kernel_timing(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        pct.fuse_device(d_img, d_amri, d_lab, d_out, nh, sigma, beta, pct.CHOICE_SM, stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    tot_ms, launches = _lib.c_double(), _lib.c_int()
    _lib.check(L.pct_kernel_timing_read(_lib.ctypes.byref(tot_ms), _lib.ctypes.byref(launches)))
    L.pct_kernel_timing(0)
    elapsed = t1 - t0
    ms_per_step = elapsed * 1e3 / args.steps
    kernel_ms = tot_ms.value / max(1, launches.value)
    value = full * args.steps / elapsed
    log(f"{ms_per_step:.3f} ms/step, kernel {kernel_ms:.3f} ms, {value:.4e} voxels/s")

    flops = FLOPS_PER_CVP * cvp
    effective = flops / (kernel_ms * 1e-3) / 1e12
    kname = "fuse_tiled_xt" if cfg.get("form", 0) == 1 else "fuse_tiled"
    timed_kernel = (f"{kname}<{cfg['seg_width']},{cfg['bx']},{cfg['bz']},{cfg['waves']},"
                    f"{(nh - 1) // 2}>")
    stamp = kernel_source_stamp()
    roofline = {
        "bound": "valu_fp32",
        "achieved": None,
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": None,
        "traffic": None,
        "issue_frac": None,
        "effective_tflops": round(effective, 3),
        "effective_frac": round(effective / PEAK_FP32_TFLOPS, 4),
        "kernel": timed_kernel,
        "kernel_ms": round(kernel_ms, 4),
        "algorithmic_flops_per_launch": flops,
        "cvp_per_launch": cvp,
        "algorithmic_bytes": 4 * full * (2 + 2 * n),
        "source_stamp": stamp,
    }
    if (nz, nx, ny, n, nh) == (112, 222, 222, 10, 11):
        # counters of another kernel (or of other sources) do not describe this launch
        prof, why = profiled_counters(timed_kernel)
        if prof is not None:
            roofline["traffic"] = prof.get("hbm_bytes_per_launch")
            roofline.update(counter_roofline(prof, kernel_ms))
            roofline["counters_source"] = PROFILE
        else:
            roofline["counters_note"] = why

    out = d_out.cpu().numpy()
    parity = None
    if args.parity_voxels > 0:
        import oracle
        rng = np.random.default_rng(args.seed)
        zi = rng.integers(r, nz - r, args.parity_voxels)
        xi = rng.integers(r, nx - r, args.parity_voxels)
        yi = rng.integers(r, ny - r, args.parity_voxels)
        flat = (zi * nx + xi) * ny + yi
        ref = oracle.voxels(img, amri, lab, nh, sigma, beta, flat)
        got = out.reshape(-1)[flat].astype(np.float64)
        dhu = np.abs(4095.0 * (got - ref) / 255.0)
        parity = {"voxels": int(args.parity_voxels), "max_dhu_vs_fp64": float(np.nanmax(dhu)),
                  "max_rel_vs_fp64": float(np.nanmax(np.abs(got - ref) / np.abs(ref))),
                  "nan_mismatch": int(np.count_nonzero(np.isnan(got) != np.isnan(ref)))}
        log(f"parity sample: {parity}")

    e2e = None
    if not args.no_end_to_end:
        log("end-to-end leg: host buffers, pinned staging, host post chain")
        e2e = end_to_end_leg(img, amri, lab, nh, sigma, beta, 3)
        log(f"end-to-end: {e2e['total_ms']:.1f} ms (fusion call {e2e['fuse_host_ms']:.1f} ms, "
            f"post {e2e['post_host_ms']:.1f} ms); pct_pseudo_ct_host "
            f"{e2e['pseudo_ct_host_ms']:.1f} ms, same output {e2e['pseudo_ct_host_same_output']}")

    if args.tile_shape >= 0:
        pct.force_tile_config(-1, 0)   # the other legs choose their own shapes
    nan_fill = None
    if not args.no_nan_fill:
        log("NaN-fill leg: random-mode volume at sigma 0.33, device vs host post chain")
        nan_fill = nan_fill_leg(args.seed)
        log(f"NaN fill: {nan_fill['nan_voxels']} NaN voxels; device chain "
            f"{nan_fill['device_chain_ms']:.1f} ms, host chain {nan_fill['host_chain_ms']:.1f} ms, "
            f"same output {nan_fill['same_output']}")

    learned_res = None
    if not args.no_learned:
        log("learned leg: training the embedding on 4 other phantoms, then pruned fusion")
        learned_res = learned_leg(d_img, d_amri, d_lab, d_out, nh, sigma, beta,
                                  max(2, args.steps // 2), 1, args.seed, ms_per_step)
        learned_res.pop("params")
        log(f"learned: {learned_res['ms_per_step']:.2f} ms/step, speedup "
            f"{learned_res['speedup_vs_plain']:.2f}x, MAE {learned_res['mae_hu_vs_plain']:.3f} HU, "
            f"fallback {learned_res['fallback_frac']:.3f}; pruned kernel "
            f"{learned_res['pruned_roofline']['kernel_ms']:.3f} ms, frac "
            f"{learned_res['pruned_roofline']['frac']}")

    batched = None
    if not args.no_batch:
        del d_img, d_amri, d_lab, d_out
        torch.cuda.empty_cache()
        log("batched leg: six uneven subjects in one graph launch")
        batched = batched_leg(max(2, args.steps // 2), 1, args.seed)
        log(f"batched: {batched['ms_per_batch']:.2f} ms per batch, "
            f"{batched['voxels_per_s']:.4e} voxels/s; one call per subject "
            f"{batched['sequential_tuned_ms']:.2f} ms")

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(16, hostinfo["affinity_cpus"])
        iz, ix = interior(nz, r), interior(nx, r)
        # the reference parallelises z with OpenMP's static schedule (cpp:1374):
        # a crop of one plane per thread is one round, and the full volume
        # takes ceil(iz / threads) rounds
        planes = min(threads, iz)
        rows = min(args.cpu_rows, ix)
        log(f"cpu baseline: {planes} planes x {rows} rows crop, {threads} threads, "
            f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}")
        kind, dt, vox_s = cpu_baseline(img, amri, lab, nh, sigma, beta, planes, rows, threads)
        rounds_full = math.ceil(iz / threads)
        rounds_crop = math.ceil(planes / threads)
        full_time = dt * (rounds_full / rounds_crop) * (ix / rows)
        cpu = {"value": full / full_time, "unit": "voxels/s", "cores": threads, "kind": kind,
               "sample": f"crop of the same inputs: {planes} interior z planes x {rows} "
                         f"interior x rows x all y ({vox_s} interior voxels, n={n}, NH={nh}) "
                         f"in {dt:.2f}s; the full volume is {rounds_full} rounds of "
                         f"{threads} planes under the static schedule and {ix / rows:.2f}x "
                         f"the rows",
               "seconds_full_volume": full_time,
               "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
               "host": hostinfo,
               "parallelism": f"OpenMP over z as the reference (cpp:1374), static schedule: "
                              f"at most {iz} threads have work at full size"}
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain multi-atlas patch fusion {nx}x{ny}x{nz}, n={n}, "
                               f"NH={nh}, patch 3, sigma={sigma}, beta={beta}",
                   "global_batch": 1, "parallelism": "single GPU",
                   "tile_config": cfg, "autotuned": tuned},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity_sample": parity,
        "end_to_end": e2e,
        "nan_fill": nan_fill,
        "learned_pruned": learned_res,
        "batched": batched,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
