"""Phase retrieval of the autocorrelation on the GPU (cmtip.phasing)."""
from .phase import (ER, HIO, ER_loop, HIO_loop, Phaser, center_of_mass, create_support_, get_phaser, phase,
                    recenter, shrink_wrap, step_phase)

__all__ = ["ER", "HIO", "ER_loop", "HIO_loop", "Phaser", "center_of_mass", "create_support_", "get_phaser", "phase",
           "recenter", "shrink_wrap", "step_phase"]
