"""Orientation matching on the GPU (cmtip.alignment)."""
from . import errors
from .align import (Slicer, calc_eudist, compute_slices, generate_weights, key_d2, key_index, match_orientations,
                    match_orientations_batch, match_orientations_mult, reference_orientations)

__all__ = ["Slicer", "calc_eudist", "compute_slices", "generate_weights", "key_d2", "key_index",
           "match_orientations", "match_orientations_batch", "match_orientations_mult", "reference_orientations",
           "errors"]
