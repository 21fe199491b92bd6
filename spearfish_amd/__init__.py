"""MI355X-native base-level alignment for spearfish: Myers bit-parallel edit distance and banded
affine-gap alignment as HIP kernels for gfx950 behind a C ABI (include/sf_align.h).

The Python layer mirrors crates/map/src/align.rs (see spearfish_amd.align); the HIP library
libsf_align.so must be built (python -m spearfish_amd.build) -- there is no CPU fallback.
"""
from .align import (NEG, USIZE_MAX, U32_MAX, Affine, Batch, Mode, PairSet, Pattern, affine, affine_batch,
                    affine_sorted, batch_distances, distance, distance_with, infix_distance, myers_batch)

__all__ = ["NEG", "USIZE_MAX", "U32_MAX", "Affine", "Batch", "Mode", "PairSet", "Pattern", "affine",
           "affine_batch", "affine_sorted", "batch_distances", "distance", "distance_with", "infix_distance",
           "myers_batch"]
