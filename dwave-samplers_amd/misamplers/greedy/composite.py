"""``SteepestDescentComposite``: descend the samples of a child sampler.

Mirrors dwave/samplers/greedy/composite.py:27-82: sample the child, then run
steepest descent seeded with the child's samples.  With a misamplers child the
whole chain stays on the GPU device except the sample set hand-over.
"""
from __future__ import annotations

from .sampler import SteepestDescentSolver

__all__ = ["SteepestDescentComposite"]


class SteepestDescentComposite:
    def __init__(self, child_sampler):
        self.children = [child_sampler]
        self.parameters = child_sampler.parameters.copy()
        self.properties = dict(child_properties=child_sampler.properties.copy())

    @property
    def child(self):
        return self.children[0]

    def sample(self, bqm, **parameters):
        sampleset = self.child.sample(bqm, **parameters)
        return SteepestDescentSolver().sample(bqm, initial_states=sampleset)

    def sample_ising(self, h, J, **parameters):
        from ..dimodlite import BinaryQuadraticModel
        return self.sample(BinaryQuadraticModel.from_ising(h, J), **parameters)

    def sample_qubo(self, Q, **parameters):
        from ..dimodlite import BinaryQuadraticModel
        return self.sample(BinaryQuadraticModel.from_qubo(Q), **parameters)
