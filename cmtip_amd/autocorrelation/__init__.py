"""Autocorrelation solve on the GPU (cmtip.autocorrelation)."""
from .autocorrelation import (ACSolver, WOperator, core_problem_convolution, fourier_reg, gen_F_antisupport,
                              gen_nonuniform_positions, get_solver, setup_linops, solve_ac,
                              solve_ac_mult)

__all__ = ["ACSolver", "WOperator", "core_problem_convolution", "fourier_reg", "gen_F_antisupport",
           "gen_nonuniform_positions", "get_solver", "setup_linops", "solve_ac", "solve_ac_mult"]
