"""Path-integral and rotor-model annealing on the MI355X engine (mirrors dwave.samplers.sqa)."""
from .pimc_annealing import simulated_annealing
from .sampler import PathIntegralAnnealingSampler, RotorModelAnnealingSampler

__all__ = ["PathIntegralAnnealingSampler", "RotorModelAnnealingSampler", "simulated_annealing"]
