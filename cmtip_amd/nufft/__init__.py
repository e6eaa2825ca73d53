"""Type-2 (forward) and type-1 (adjoint) 3-D NUFFTs on the GPU (cmtip.nufft)."""
from .adjoint import adjoint_cpu, adjoint_gpu, adjoint_ones
from .forward import forward_cpu, forward_gpu, forward_real
from .plan import DEFAULT_EPS, NufftPlan, clear_plans, get_plan, spread_candidates

__all__ = ["forward_cpu", "forward_gpu", "forward_real", "adjoint_cpu", "adjoint_gpu", "adjoint_ones",
           "NufftPlan", "get_plan", "clear_plans", "spread_candidates", "DEFAULT_EPS"]
