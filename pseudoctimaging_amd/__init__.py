"""MI355X-native (gfx950) multi-atlas patch fusion for pseudo-CT synthesis.

The fusion runs in hand-written HIP kernels inside libpct.so (C ABI in
include/pct/pct.h); this package is the Python host interface over it.
"""
from .api import (CHOICE_DIRECT, CHOICE_GM, CHOICE_GM2, CHOICE_SM, PctError, autotune,  # noqa: F401
                  autotune_batch, batch_choice, check_volume_set, fuse_batch, set_batch_choice,
                  set_batch_config,
                  device_count, force_tile_config, fuse_device, fused_to_hu_raw,
                  mat_read, mat_write, matrixio_read, matrixio_write, matrixio_write_sidecar,
                  multi_patch_based_segmentation_gpu,
                  normalize_labels, normalize_labels_device, postprocess, postprocess_device,
                  problem, pseudo_ct, pseudo_ct_hu, regularization_step_device,
                  read_config, regularization_step, release, round_matrix, synth_tissue,
                  synth_uniform,
                  tile_config, to_hu, weight_divisor)

__all__ = [
    "CHOICE_DIRECT", "CHOICE_GM", "CHOICE_GM2", "CHOICE_SM", "PctError", "autotune", "autotune_batch",
    "batch_choice", "check_volume_set", "fuse_batch", "set_batch_choice", "set_batch_config", "device_count",
    "force_tile_config", "fuse_device", "fused_to_hu_raw", "mat_read", "mat_write",
    "matrixio_read", "matrixio_write", "matrixio_write_sidecar",
    "multi_patch_based_segmentation_gpu", "normalize_labels", "normalize_labels_device",
    "postprocess", "postprocess_device", "problem", "pseudo_ct", "pseudo_ct_hu", "read_config",
    "regularization_step", "regularization_step_device", "release", "round_matrix",
    "synth_tissue", "synth_uniform", "tile_config", "to_hu", "weight_divisor",
]
