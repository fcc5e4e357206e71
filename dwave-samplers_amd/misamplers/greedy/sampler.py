"""Drop-in ``SteepestDescentSolver`` (alias ``SteepestDescentSampler``).

Mirrors dwave/samplers/greedy/sampler.py:31-318: the same ``sample`` signature
and defaults (:137-142), seed validation (:260-264), initial-state parsing,
``num_steps`` data vector, vartype round trip and ``timing`` info.
"""
from __future__ import annotations

from numbers import Integral
from time import perf_counter_ns

import numpy as np

from .. import dimodlite
from .descent import steepest_gradient_descent

__all__ = ["SteepestDescentSolver", "SteepestDescentSampler"]


class SteepestDescentSolver(dimodlite.Sampler, dimodlite.Initialized):
    """Steepest gradient descent on binary quadratic models, on MI355X."""

    parameters = None
    properties = None

    def __init__(self):
        self.parameters = {"num_reads": [], "initial_states": [],
                           "initial_states_generator": ["initial_states_generators"], "seed": [],
                           "large_sparse_opt": ["large_sparse_opt_values"]}
        self.properties = {"initial_states_generators": ("none", "tile", "random"),
                           "large_sparse_opt_values": (True, False)}

    def sample(self, bqm, num_reads=None, initial_states=None, initial_states_generator="random", seed=None,
               large_sparse_opt=False, **kwargs):
        """Find minima of a binary quadratic model (greedy/sampler.py:137-304)."""
        t_pre = perf_counter_ns()
        original_vartype = bqm.vartype
        if dimodlite.as_vartype(bqm.vartype) is not dimodlite.SPIN:
            bqm = bqm.change_vartype(dimodlite.SPIN if isinstance(bqm, dimodlite.BinaryQuadraticModel)
                                     else "SPIN", inplace=False)
        if not (seed is None or isinstance(seed, Integral)):
            raise TypeError("'seed' should be None or a positive 32-bit integer")
        if isinstance(seed, Integral) and not 0 <= seed <= 2**32 - 1:
            raise ValueError("'seed' should be an integer between 0 and 2**32 - 1 inclusive")
        parsed = self.parse_initial_states(bqm, num_reads=num_reads, initial_states=initial_states,
                                           initial_states_generator=initial_states_generator, seed=seed)
        num_reads = parsed.num_reads
        variables = parsed.initial_states.variables
        linear, (cs, ce, cw), offset = bqm.to_numpy_vectors(variable_order=variables)
        init = np.ascontiguousarray(parsed.initial_states.record.sample, dtype=np.int8)
        t_sample = perf_counter_ns()
        samples, energies, num_steps = steepest_gradient_descent(num_reads, linear, cs, ce, cw, init,
                                                                 large_sparse_opt)
        t_post = perf_counter_ns()
        result = dimodlite.SampleSet.from_samples((samples, variables), energy=energies + offset,
                                                  vartype=dimodlite.SPIN, num_steps=num_steps)
        result.change_vartype(original_vartype, inplace=True)
        result.info.update(dict(timing=dict(preprocessing_ns=t_sample - t_pre, sampling_ns=t_post - t_sample,
                                            postprocessing_ns=perf_counter_ns() - t_post)))
        return result


SteepestDescentSampler = SteepestDescentSolver
