"""Simulated annealing on the MI355X sweep engine (mirrors dwave.samplers.sa)."""
from .sampler import Neal, SimulatedAnnealingSampler, default_beta_range, _default_ising_beta_range
from .simulated_annealing import simulated_annealing

__all__ = ["Neal", "SimulatedAnnealingSampler", "default_beta_range", "simulated_annealing"]
