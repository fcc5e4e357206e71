"""Drop-in ``SimulatedAnnealingSampler`` running the HIP sweep engine.

Mirrors dwave/samplers/sa/sampler.py: same ``sample`` signature and defaults
(:156-169), the same validation errors (:331-426), the same schedule
construction (:373-426), ``info`` keys (:439-459), ``parameters`` and
``properties`` (:138-154), the alias ``Neal`` (:464) and the default beta range
(:467-604).  Sampling itself goes through ``simulated_annealing`` (the C ABI).
"""
from __future__ import annotations

import math
import warnings
from collections import defaultdict
from numbers import Integral
from time import perf_counter_ns
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np

from .. import dimodlite
from .simulated_annealing import simulated_annealing

__all__ = ["Neal", "SimulatedAnnealingSampler", "default_beta_range", "_default_ising_beta_range"]


def _ising_views(bqm):
    """(h dict, J dict) of a BQM's spin form (dimod or dimodlite)."""
    spin = bqm if dimodlite.as_vartype(bqm.vartype) is dimodlite.SPIN else bqm.change_vartype(
        dimodlite.SPIN if isinstance(bqm, dimodlite.BinaryQuadraticModel) else "SPIN", inplace=False)
    return spin.linear, spin.quadratic


def build_beta_schedule(beta_range, num_sweeps, num_sweeps_per_beta, beta_schedule_type,
                        beta_schedule, linear, quadratic):
    """Validated (beta_schedule, beta_range) as sa/sampler.py:362-426 builds them."""
    if not isinstance(num_sweeps_per_beta, Integral):
        raise TypeError(f"'num_sweeps_per_beta' should be a positive integer: value = {num_sweeps_per_beta}")
    if num_sweeps_per_beta < 1:
        raise ValueError(f"'num_sweeps_per_beta' should be a positive integer: value = {num_sweeps_per_beta}")
    if beta_schedule_type == "custom":
        if beta_schedule is None:
            raise ValueError("'beta_schedule' must be provided for beta_schedule_type = 'custom': value is None")
        try:
            beta_schedule = np.array(beta_schedule, dtype=float)
        except Exception:
            raise ValueError("beta_schedule cannot be case as a numpy array of dtype=float")
        if num_sweeps is not None and num_sweeps != len(beta_schedule) * num_sweeps_per_beta:
            raise ValueError("'num_sweeps' should be set to None, or a value consistent with "
                             "'beta_schedule' and 'num_sweeps_per_beta' for 'beta_schedule_type' = 'custom'")
        if beta_range is not None and (beta_range[0] != beta_schedule[0] or beta_range[-1] != beta_schedule[-1]):
            raise ValueError("'beta_range' should be set to None, or a value consistent with "
                             "'beta_schedule', for 'beta_schedule_type'='custom'.")
        if beta_schedule.size and np.min(beta_schedule) < 0:
            raise ValueError("'beta_schedule' cannot include negative values.")
        return beta_schedule, beta_range
    if beta_schedule is not None:
        raise ValueError("'beta_schedule' must be set to None for 'beta_schedule_type' not equal to 'custom'")
    if num_sweeps is None:
        num_sweeps = 1000
    num_betas, rem = divmod(num_sweeps, num_sweeps_per_beta)
    if rem > 0 or num_betas < 0:
        raise ValueError("'num_sweeps' must be a non-negative value divisible by 'num_sweeps_per_beta'.")
    if beta_range is None:
        beta_range = _default_ising_beta_range(linear, quadratic)
    elif len(beta_range) != 2 or min(beta_range) < 0:
        raise ValueError("'beta_range' should be a 2-tuple, or 2 element list of positive numbers. "
                         "The latter value is the target value.")
    if num_betas == 1:
        return np.array([beta_range[-1]], dtype=float), beta_range
    if beta_schedule_type == "linear":
        return np.linspace(*beta_range, num=num_betas), beta_range
    if beta_schedule_type == "geometric":
        if min(beta_range) <= 0:
            raise ValueError("'beta_range' must contain non-zero values for 'beta_schedule_type' = 'geometric'")
        return np.geomspace(*beta_range, num=num_betas), beta_range
    raise ValueError(f"Beta schedule type {beta_schedule_type} not implemented")


def validate_seed(seed):
    """sa/sampler.py:331-340: None -> random in [0, 2^31); else an Integral in range."""
    if seed is None:
        return int(np.random.randint(2**31))
    if not isinstance(seed, Integral):
        raise TypeError(f"'seed' should be None or an integer between 0 and 2^32 - 1: value = {seed}")
    if not (0 <= seed < 2**31):
        raise ValueError(f"'seed' should be an integer between 0 and 2^32 - 1: value = {seed}")
    return int(seed)


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
        t_pre = perf_counter_ns()
        original_vartype = bqm.vartype
        if dimodlite.as_vartype(bqm.vartype) is not dimodlite.SPIN:
            bqm = bqm.change_vartype(dimodlite.SPIN if isinstance(bqm, dimodlite.BinaryQuadraticModel)
                                     else "SPIN", inplace=False)
        seed = validate_seed(seed)
        parsed = self.parse_initial_states(bqm, num_reads=num_reads, initial_states=initial_states,
                                           initial_states_generator=initial_states_generator, seed=seed)
        num_reads = parsed.num_reads
        initial_states_array = np.ascontiguousarray(parsed.initial_states.record.sample, dtype=np.int8)
        variable_order = parsed.initial_states.variables
        ldata, (irow, icol, qdata), off = bqm.to_numpy_vectors(variable_order=variable_order)
        if interrupt_function and not callable(interrupt_function):
            raise TypeError("'interrupt_function' should be a callable")
        beta_schedule, beta_range = build_beta_schedule(
            beta_range, num_sweeps, num_sweeps_per_beta, beta_schedule_type, beta_schedule,
            bqm.linear, bqm.quadratic)
        t_sample = perf_counter_ns()
        samples, energies = simulated_annealing(
            num_reads, ldata, irow, icol, qdata, num_sweeps_per_beta, beta_schedule, seed,
            initial_states_array, randomize_order, proposal_acceptance_criteria, interrupt_function)
        t_post = perf_counter_ns()
        info = {"beta_range": beta_range, "beta_schedule_type": beta_schedule_type}
        response = dimodlite.SampleSet.from_samples((samples, variable_order), energy=energies + off,
                                                    info=info, vartype=dimodlite.SPIN)
        response.change_vartype(original_vartype, inplace=True)
        response.info.update(dict(timing=dict(preprocessing_ns=t_sample - t_pre,
                                              sampling_ns=t_post - t_sample,
                                              postprocessing_ns=perf_counter_ns() - t_post)))
        return response


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
    h, J = _ising_views(bqm)
    return _default_ising_beta_range(h, J)
