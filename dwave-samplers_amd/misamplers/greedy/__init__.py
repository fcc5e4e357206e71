"""Steepest descent on the MI355X engine (mirrors dwave.samplers.greedy)."""
from .composite import SteepestDescentComposite
from .descent import steepest_gradient_descent
from .sampler import SteepestDescentSampler, SteepestDescentSolver

__all__ = ["SteepestDescentComposite", "SteepestDescentSampler", "SteepestDescentSolver",
           "steepest_gradient_descent"]
