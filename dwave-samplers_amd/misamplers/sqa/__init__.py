"""Path-integral annealing on the MI355X engine (mirrors dwave.samplers.sqa)."""
from .pimc_annealing import simulated_annealing
from .sampler import PathIntegralAnnealingSampler

__all__ = ["PathIntegralAnnealingSampler", "simulated_annealing"]
