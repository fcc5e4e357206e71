"""Drop-in ``PathIntegralAnnealingSampler`` and ``RotorModelAnnealingSampler``
running the HIP path-integral and rotor-model engines.

The interface follows dwave/samplers/sqa/sampler.py: the same ``sample``
keywords and defaults (:119-145, :656-684), ``parameters`` / ``properties``
(:100-117), validation messages and their order, and ``info`` keys
(:529-545, :1084-1100).  The host steps (spin form, seed, initial states,
schedule, break buffers, response) are the shared ones of
``misamplers._sampling``.  ``num_trotter_slices`` (default 32) is an extension
of the path-integral sampler: the engine's imaginary time has that many slices.
"""
from __future__ import annotations

from time import perf_counter_ns
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np

from .. import dimodlite
from .._sampling import (PATH_RULES, ROTOR_RULES, anneal_schedule, break_buffers, check_interrupt,
                         check_sample_interval, check_seed, check_sweeps_per_beta, json_ready, spin_form,
                         spin_inputs, spin_response, transverse_schedule, SpinInputs)
from .pimc_annealing import simulated_annealing as path_annealing
from .rotormc_annealing import simulated_annealing as rotor_annealing

__all__ = ["PathIntegralAnnealingSampler", "RotorModelAnnealingSampler"]

_PARAMETERS = ("beta_range", "num_reads", "num_sweeps", "num_sweeps_per_beta", "beta_schedule_type", "seed",
               "interrupt_function", "initial_states", "initial_states_generator")


def _parameters():
    out = {k: [] for k in _PARAMETERS}
    out["beta_schedule_type"] = ["beta_schedule_options"]
    return out


def _path_info(beta_range, schedule_type, interval, statistics, keep_worldlines, num_breaks, buffer):
    """info of a path-integral response: statistics when sampled along the
    schedule; the final worldlines as break lists when kept, marked complete
    when the buffer held every break."""
    info = {"beta_range": beta_range, "beta_schedule_type": schedule_type}
    if interval > 0:
        info["statistics"] = statistics
    if keep_worldlines:
        total = int(np.sum(num_breaks))
        complete = buffer.size >= total
        info["num_breaks"] = num_breaks
        info["complete_break_specification"] = complete
        info["break_positions"] = buffer[:total] if complete else buffer
    return info


def _rotor_fields(trans_fields, num_vars: int) -> np.ndarray:
    """Per-rotor transverse fields: ones by default, a sequence of num_vars
    values, or a scalar for every rotor."""
    if trans_fields is None:
        return np.ones(num_vars)
    if isinstance(trans_fields, (list, tuple, np.ndarray)):
        if len(trans_fields) != num_vars:
            raise ValueError("bqm.num_variables and trans_fields lengths incompatible")
        return np.array(trans_fields)
    return trans_fields * np.ones(num_vars)


def _down(angles: np.ndarray, prng) -> np.ndarray:
    """Random spin projection of angles i (theta = pi i / 128): True where the
    spin is -1, which has probability (1 - cos theta) / 2
    (sqa/sampler.py:857-858, :1105-1109 project the same way)."""
    return np.cos(np.pi * angles / 128) < 1 - 2 * prng.random(size=angles.shape)


class PathIntegralAnnealingSampler(dimodlite.Sampler, dimodlite.Initialized):
    """Path-integral (transverse-field) annealing sampler for binary quadratic
    models, on MI355X.  Same interface as
    ``dwave.samplers.sqa.PathIntegralAnnealingSampler``."""

    parameters = None
    properties = None

    def __init__(self):
        self.parameters = _parameters()
        self.properties = {"beta_schedule_options": ("linear", "geometric", "custom")}

    def sample(self, bqm, *, beta_range: Optional[Union[List[float], Tuple[float, float]]] = None,
               num_reads: Optional[int] = None, num_sweeps: Optional[int] = None,
               num_sweeps_per_beta: int = 1, beta_schedule_type: str = "geometric",
               seed: Optional[int] = None, Hp_field: Optional[Union[Sequence[float], np.ndarray]] = None,
               Hd_field: Optional[Union[Sequence[float], np.ndarray]] = None, Gamma: float = 1,
               chain_coupler_strength: float = 1, qubits_per_chain: int = 1, qubits_per_update: int = 1,
               initial_states=None, initial_states_generator: str = "random",
               make_info_json_serializable: bool = False, project_states: Tuple[bool, bool] = (True, True),
               num_breaks: Optional[np.ndarray] = None, breaks_in: Optional[np.ndarray] = None,
               breaks_buffer_out: Optional[np.ndarray] = None, schedule_sample_interval: Optional[int] = None,
               interrupt_function=None, num_trotter_slices: int = 32, **kwargs):
        """Sample from ``bqm`` (sqa/sampler.py:119-569)."""
        inputs = spin_inputs(self, bqm, seed, num_reads, initial_states, initial_states_generator,
                             interrupt_function)
        check_sweeps_per_beta(num_sweeps_per_beta)
        hp, beta_range = anneal_schedule(PATH_RULES, beta_range, num_sweeps, num_sweeps_per_beta,
                                         beta_schedule_type, Hp_field, inputs.bqm.linear, inputs.bqm.quadratic)
        hd = transverse_schedule(Hd_field, len(hp))
        keep_worldlines = not project_states[1]
        num_breaks, breaks_in, buffer = break_buffers(inputs.states.shape, num_breaks, breaks_in,
                                                      breaks_buffer_out, keep_worldlines, hd, len(hp))
        interval = check_sample_interval(schedule_sample_interval, len(hp))

        t_sample = perf_counter_ns()
        samples, energies, statistics, num_breaks, buffer = path_annealing(
            inputs.num_reads, inputs.h, *inputs.couplers, num_sweeps_per_beta, hp, hd, Gamma,
            chain_coupler_strength, qubits_per_chain, qubits_per_update, inputs.seed, inputs.states,
            project_states[0], project_states[1], num_breaks, breaks_in, buffer, interval, interrupt_function,
            num_trotter_slices=num_trotter_slices)
        t_post = perf_counter_ns()

        info = _path_info(beta_range, beta_schedule_type, interval, statistics, keep_worldlines, num_breaks, buffer)
        if make_info_json_serializable:
            json_ready(info, ("statistics", "num_breaks", "break_positions"))
        return spin_response(inputs, samples, energies, info, t_sample, t_post)


class RotorModelAnnealingSampler(dimodlite.Sampler, dimodlite.Initialized):
    """Rotor-model (planar spin) annealing sampler for binary quadratic models,
    on MI355X.  Same interface as ``dwave.samplers.sqa.RotorModelAnnealingSampler``
    (sqa/sampler.py:572-1132); the anneal runs in the HIP rotor engine, read r
    on its own xorshift128+ stream seeded from the call's seed."""

    parameters = None
    properties = None

    def __init__(self):
        self.parameters = _parameters()
        self.properties = {"beta_schedule_options": ("linear", "geometric", "custom")}

    def _angles(self, spin, num_reads, initial_states, initial_states_generator, seed, project_in, prng):
        """(parsed inputs, num_reads, initial angles): an ndarray of initial
        states is taken as given (shape (num_reads, num_vars)); angles (uint8)
        pass through or are projected to spin angles, spins +-1 map to 0 / 128."""
        if type(initial_states) == np.ndarray:
            parsed = self.parse_initial_states(spin, num_reads=num_reads, initial_states=None,
                                               initial_states_generator=initial_states_generator, seed=seed)
            if num_reads is None:
                num_reads = initial_states.shape[0]
            if (num_reads, spin.num_variables) != initial_states.shape:
                raise ValueError("Shape of initial_states specified asndarray should match (num_reads, num_vars)")
            given = initial_states
        else:
            parsed = self.parse_initial_states(spin, num_reads=num_reads, initial_states=initial_states,
                                               initial_states_generator=initial_states_generator, seed=seed)
            num_reads = parsed.num_reads
            given = parsed.initial_states.record.sample
        if given.dtype == np.uint8:
            angles = (128 * _down(given, prng)) if project_in else given
        else:
            if not set(np.unique(given)).issubset({-1, 1}):
                raise ValueError("initial_states dtype is neither np.uint8 nor consistent with spins {-1, 1}")
            angles = 64 - 64 * given
        return parsed, num_reads, np.array(angles, dtype=np.uint8, order="C", copy=True)

    def sample(self, bqm, *, beta_range: Optional[Union[List[float], Tuple[float, float]]] = None,
               num_reads: Optional[int] = None, num_sweeps: Optional[int] = None,
               num_sweeps_per_beta: int = 1, beta_schedule_type: str = "geometric",
               seed: Optional[int] = None, Hp_field: Optional[Union[Sequence[float], np.ndarray]] = None,
               Hd_field: Optional[Union[Sequence[float], np.ndarray]] = None, initial_states=None,
               initial_states_generator: str = "random", make_info_json_serializable: bool = False,
               project_states: Tuple[bool, bool] = (False, True), schedule_sample_interval: Optional[int] = None,
               randomize_order: bool = False, proposal_acceptance_criteria: str = "MetropolisUniform",
               trans_fields=None, interrupt_function=None, **kwargs):
        """Sample from ``bqm`` (sqa/sampler.py:656-1132).  Samples are the
        final angles projected to spins at random (P(+1) = (1 + cos theta) / 2);
        energies are the cos-weighted classical energies of the angles plus the
        offset; ``info["rotor_states"]`` holds the angles when
        ``project_states[1]`` is False."""
        t_start = perf_counter_ns()
        original = bqm.vartype
        spin = spin_form(bqm)
        fields = _rotor_fields(trans_fields, spin.num_variables)
        seed = check_seed(seed)
        prng = np.random.default_rng(seed)  # reproducible projections
        parsed, num_reads, angles = self._angles(spin, num_reads, initial_states, initial_states_generator, seed,
                                                 project_states[0], prng)
        variables = parsed.initial_states.variables
        h, couplers, offset = spin.to_numpy_vectors(variable_order=variables)
        check_interrupt(interrupt_function)
        check_sweeps_per_beta(num_sweeps_per_beta)
        hp, beta_range = anneal_schedule(ROTOR_RULES, beta_range, num_sweeps, num_sweeps_per_beta,
                                         beta_schedule_type, Hp_field, spin.linear, spin.quadratic)
        hd = transverse_schedule(Hd_field, len(hp))
        interval = check_sample_interval(schedule_sample_interval, len(hp))

        t_sample = perf_counter_ns()
        final, energies, statistics = rotor_annealing(
            num_reads, h, *couplers, fields, num_sweeps_per_beta, hp, hd, seed, angles, randomize_order,
            proposal_acceptance_criteria, interval, interrupt_function)
        t_post = perf_counter_ns()

        info = {"beta_range": beta_range, "beta_schedule_type": beta_schedule_type}
        if interval > 0:
            info["statistics"] = statistics
        if not project_states[1]:
            info["rotor_states"] = final.copy()
        if make_info_json_serializable:
            json_ready(info, ("statistics", "rotor_states"))
        samples = (1 - 2 * _down(final, prng)).astype(np.int8)  # always projected to spins
        inputs = SpinInputs(spin, original, seed, num_reads, angles, variables, h, couplers, offset, t_start)
        return spin_response(inputs, samples, energies, info, t_sample, t_post)
