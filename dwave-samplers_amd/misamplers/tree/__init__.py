"""Exact inference by bucket elimination on the MI355X engine (mirrors dwave.samplers.tree)."""
from .samplers import TreeDecompositionSampler, TreeDecompositionSolver, TreePlan

__all__ = ["TreeDecompositionSampler", "TreeDecompositionSolver", "TreePlan"]
