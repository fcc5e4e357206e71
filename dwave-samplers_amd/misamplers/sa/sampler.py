"""Drop-in ``SimulatedAnnealingSampler`` running the HIP sweep engine.

Mirrors dwave/samplers/sa/sampler.py: same ``sample`` signature and defaults
(:156-169), the same validation errors (:331-426), the same schedule
construction (:373-426), ``info`` keys (:439-459), ``parameters`` and
``properties`` (:138-154), the alias ``Neal`` (:464) and the default beta range
(:467-604).  Sampling itself goes through ``simulated_annealing`` (the C ABI).
"""
from __future__ import annotations

import warnings
from collections import defaultdict
from time import perf_counter_ns
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np

from .. import dimodlite
from .._sampling import SA_RULES, anneal_schedule, check_sweeps_per_beta, spin_form, spin_inputs, spin_response
from .simulated_annealing import simulated_annealing

__all__ = ["Neal", "SimulatedAnnealingSampler", "default_beta_range", "_default_ising_beta_range"]


def build_beta_schedule(beta_range, num_sweeps, num_sweeps_per_beta, beta_schedule_type,
                        beta_schedule, linear, quadratic):
    """Validated (beta_schedule, beta_range) as sa/sampler.py:362-426 builds them."""
    check_sweeps_per_beta(num_sweeps_per_beta)
    return anneal_schedule(SA_RULES, beta_range, num_sweeps, num_sweeps_per_beta, beta_schedule_type,
                           beta_schedule, linear, quadratic)


class SimulatedAnnealingSampler(dimodlite.Sampler, dimodlite.Initialized):
    """Simulated annealing sampler for binary quadratic models, on MI355X.

    Same interface as ``dwave.samplers.SimulatedAnnealingSampler``; every read
    is an independent chain seeded from ``seed`` (see ``simulated_annealing``).
    """

    parameters = None
    properties = None

    def __init__(self):
        self.parameters = {"beta_range": [], "num_reads": [], "num_sweeps": [],
                           "num_sweeps_per_beta": [], "beta_schedule_type": ["beta_schedule_options"],
                           "beta_schedule": [], "seed": [], "interrupt_function": [],
                           "initial_states": [], "initial_states_generator": [],
                           "randomize_order": [], "proposal_acceptance_criteria": []}
        self.properties = {"beta_schedule_options": ("linear", "geometric", "custom")}

    def sample(self, bqm, beta_range: Optional[Union[List[float], Tuple[float, float]]] = None,
               num_reads: Optional[int] = None, num_sweeps: Optional[int] = None,
               num_sweeps_per_beta: int = 1, beta_schedule_type: str = "geometric",
               seed: Optional[int] = None, interrupt_function=None,
               beta_schedule: Optional[Union[Sequence[float], np.ndarray]] = None,
               initial_states=None, initial_states_generator: str = "random",
               randomize_order: bool = False, proposal_acceptance_criteria: str = "Metropolis",
               **kwargs):
        """Sample from ``bqm`` (sa/sampler.py:156-461)."""
        for k in kwargs:
            if k not in self.parameters:
                warnings.warn(f"Ignoring unknown kwarg: {k!r}", UserWarning, stacklevel=2)
        inputs = spin_inputs(self, bqm, seed, num_reads, initial_states, initial_states_generator,
                             interrupt_function)
        beta_schedule, beta_range = build_beta_schedule(
            beta_range, num_sweeps, num_sweeps_per_beta, beta_schedule_type, beta_schedule,
            inputs.bqm.linear, inputs.bqm.quadratic)
        t_sample = perf_counter_ns()
        samples, energies = simulated_annealing(
            inputs.num_reads, inputs.h, *inputs.couplers, num_sweeps_per_beta, beta_schedule, inputs.seed,
            inputs.states, randomize_order, proposal_acceptance_criteria, interrupt_function)
        t_post = perf_counter_ns()
        info = {"beta_range": beta_range, "beta_schedule_type": beta_schedule_type}
        return spin_response(inputs, samples, energies, info, t_sample, t_post)


Neal = SimulatedAnnealingSampler


def _default_ising_beta_range(h, J, max_single_qubit_excitation_rate=0.01, scale_T_with_N=True):
    """Default [hot_beta, cold_beta] from h and J (sa/sampler.py:467-600).

    hot: log(2) / (2 * max_i sum of |biases| on i) so every flip is taken with
    probability >= 1/2; cold: log(#minimal gaps / rate) / (2 * min nonzero
    |bias|); [0.1, 1] with a warning when every bias is zero.
    """
    if not 0 < max_single_qubit_excitation_rate < 1:
        raise ValueError("Targeted single qubit excitations rates must be in range (0,1)")
    sum_abs = defaultdict(int, {k: abs(v) for k, v in h.items()})
    min_abs = {k: v for k, v in sum_abs.items() if v != 0} if sum_abs else {}
    for (k1, k2), v in J.items():
        for k in (k1, k2):
            sum_abs[k] += abs(v)
            if v != 0:
                min_abs[k] = min(abs(v), min_abs[k]) if k in min_abs else abs(v)
    if not min_abs:
        warnings.warn("All bqm biases are zero (all energies are zero), this is likely a value error. "
                      "Temperature range is set arbitrarily to [0.1,1]. Metropolis-Hastings update is "
                      "non-ergodic.")
        return [0.1, 1]
    max_field = max(sum_abs.values(), default=0)
    hot_beta = 1 if max_field == 0 else np.log(2) / (2 * max_field)
    values = np.array(list(min_abs.values()), dtype=float)
    min_field = np.min(values)
    n_min = np.sum(min_field == values) if scale_T_with_N else 1
    cold_beta = np.log(n_min / max_single_qubit_excitation_rate) / (2 * min_field)
    return [hot_beta, cold_beta]


def default_beta_range(bqm):
    spin = spin_form(bqm)
    return _default_ising_beta_range(spin.linear, spin.quadratic)
