"""Drop-in ``PathIntegralAnnealingSampler`` running the HIP path-integral engine.

Mirrors dwave/samplers/sqa/sampler.py:36-569: the same ``sample`` signature and
defaults (:119-145), validation errors (:330-500), (Hp, Hd) schedule
construction (:372-458), break-buffer sizing heuristic (:460-495), ``info`` keys
(:529-545), ``parameters`` / ``properties`` (:100-117) and return conversion
(:547-569).  The keyword ``num_trotter_slices`` (default 32) is an extension:
the engine's imaginary time has that many slices.
"""
from __future__ import annotations

import warnings
from numbers import Integral
from time import perf_counter_ns
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np

from .. import dimodlite
from ..sa.sampler import _default_ising_beta_range
from .pimc_annealing import simulated_annealing as path_annealing
from .rotormc_annealing import simulated_annealing as rotor_annealing

__all__ = ["PathIntegralAnnealingSampler", "RotorModelAnnealingSampler"]


class PathIntegralAnnealingSampler(dimodlite.Sampler, dimodlite.Initialized):
    """Path-integral (transverse-field) annealing sampler for binary quadratic
    models, on MI355X.  Same interface as
    ``dwave.samplers.sqa.PathIntegralAnnealingSampler``."""

    parameters = None
    properties = None

    def __init__(self):
        self.parameters = {"beta_range": [], "num_reads": [], "num_sweeps": [], "num_sweeps_per_beta": [],
                           "beta_schedule_type": ["beta_schedule_options"], "seed": [],
                           "interrupt_function": [], "initial_states": [], "initial_states_generator": []}
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
        t_pre = perf_counter_ns()
        original_vartype = bqm.vartype
        if dimodlite.as_vartype(bqm.vartype) is not dimodlite.SPIN:
            bqm = bqm.change_vartype(dimodlite.SPIN if isinstance(bqm, dimodlite.BinaryQuadraticModel)
                                     else "SPIN", inplace=False)
        if seed is None:
            seed = int(np.random.randint(2**31))
        elif not isinstance(seed, Integral):
            raise TypeError("'seed' should be None or an integer between 0 and 2^32 - 1: value = {}".format(seed))
        elif not (0 <= seed < 2**31):
            raise ValueError("'seed' should be an integer between 0 and 2^32 - 1: value = {}".format(seed))
        parsed = self.parse_initial_states(bqm, num_reads=num_reads, initial_states=initial_states,
                                           initial_states_generator=initial_states_generator, seed=seed)
        num_reads = parsed.num_reads
        initial_states_array = np.ascontiguousarray(parsed.initial_states.record.sample, dtype=np.int8)
        variable_order = parsed.initial_states.variables
        ldata, (irow, icol, qdata), off = bqm.to_numpy_vectors(variable_order=variable_order)
        if interrupt_function and not callable(interrupt_function):
            raise TypeError("'interrupt_function' should be a callable")
        if not isinstance(num_sweeps_per_beta, Integral):
            raise TypeError("'num_sweeps_per_beta' should be a positive integer: value = {}".format(
                num_sweeps_per_beta))
        if num_sweeps_per_beta < 1:
            raise ValueError("'num_sweeps_per_beta' should be a positive integer: value = {}".format(
                num_sweeps_per_beta))

        if beta_schedule_type == "custom":
            if Hp_field is None:
                raise ValueError("'Hp_field' must be provided for beta_schedule_type = 'custom': value is None")
            try:
                Hp_field = np.array(Hp_field, dtype=float)
            except Exception:
                raise ValueError("Hp_field cannot be case as a numpy array of dtype=float")
            if num_sweeps is not None and num_sweeps != len(Hp_field) * num_sweeps_per_beta:
                raise ValueError("'num_sweeps' should be set to None, or a value consistent with 'Hp_field' "
                                 "and 'num_sweeps_per_beta' for 'beta_schedule_type' = 'custom': value = ")
            elif beta_range is not None and (beta_range[0] != Hp_field[0] or beta_range[-1] != Hp_field[-1]):
                raise ValueError("'beta_range' should be set to None, or a value consistent with 'Hp_field', "
                                 "for 'beta_schedule_type'='custom'.")
            elif np.min(Hp_field) < 0:
                raise ValueError("'Hp_field' cannot include negative values.")
        else:
            if Hp_field is not None:
                raise ValueError("'Hp_field' must be set to None for 'beta_schedule_type' not equal to 'custom'")
            elif num_sweeps is None:
                num_sweeps = 100
            num_betas, rem = divmod(num_sweeps, num_sweeps_per_beta)
            if rem > 0 or num_betas < 0:
                raise ValueError("'num_sweeps' must be a positive value divisible by 'num_sweeps_per_beta'.")
            if beta_range is None:
                beta_range = _default_ising_beta_range(bqm.linear, bqm.quadratic)
            elif len(beta_range) != 2 or min(beta_range) < 0:
                raise ValueError("'beta_range' should be a 2-tuple, or 2 element list of positive numbers. "
                                 "The latter value is the target value.")
            if num_betas == 1:
                Hp_field = np.array([beta_range[-1]], dtype=float)
            elif beta_schedule_type == "linear":
                Hp_field = np.linspace(*beta_range, num=num_betas)
            elif beta_schedule_type == "geometric":
                if min(beta_range) <= 0:
                    raise ValueError("'beta_range' must contain non-zero values for 'beta_schedule_type' = "
                                     "'geometric'")
                Hp_field = np.geomspace(*beta_range, num=num_betas)
            else:
                raise ValueError("Beta schedule type {} not implemented".format(beta_schedule_type))

        num_betas = len(Hp_field)
        if Hd_field is None:
            Hd_field = np.zeros(num_betas)
        else:
            Hd_field = np.array(Hd_field, dtype=float)
            if len(Hd_field) < num_betas:
                raise ValueError("Hd_field incompatible with Hp_field")
        if num_breaks is None:
            num_breaks = np.zeros(shape=initial_states_array.shape, dtype=np.intc)
        elif num_breaks.shape != initial_states_array.shape:
            raise ValueError("Path integral initialization is inconsistent")
        if breaks_in is None:
            breaks_in = np.zeros(max(1, num_breaks.sum()))
        breaks_in = np.ascontiguousarray(breaks_in, dtype=np.intc)
        if breaks_buffer_out is None:
            if not project_states[1]:
                # Heuristic buffer size of sqa/sampler.py:466-491.
                num_break_patterns = num_reads * bqm.num_variables
                if num_betas > 0:
                    out_requirement = int((1 + Hd_field[-1]) * (num_break_patterns + np.log(num_break_patterns))
                                          + 1000)
                else:
                    out_requirement = int(np.sum(num_breaks))
            else:
                out_requirement = 1
            breaks_buffer_out = np.zeros(out_requirement, dtype=np.intc)
        breaks_buffer_out = np.ascontiguousarray(breaks_buffer_out, dtype=np.intc)
        num_breaks = np.ascontiguousarray(num_breaks, dtype=np.intc)
        if schedule_sample_interval is None:
            schedule_sample_interval = 0
        elif schedule_sample_interval > len(Hp_field) or schedule_sample_interval < 1:
            raise ValueError("schedule_sample_interval is incompatible with Hp_field")

        t_sample = perf_counter_ns()
        samples, energies, statistics, num_breaks, breaks_buffer_out = path_annealing(
            num_reads, ldata, irow, icol, qdata, num_sweeps_per_beta, Hp_field, Hd_field, Gamma,
            chain_coupler_strength, qubits_per_chain, qubits_per_update, seed, initial_states_array,
            project_states[0], project_states[1], num_breaks, breaks_in, breaks_buffer_out,
            schedule_sample_interval, interrupt_function, num_trotter_slices=num_trotter_slices)
        t_post = perf_counter_ns()

        info = {"beta_range": beta_range, "beta_schedule_type": beta_schedule_type}
        if schedule_sample_interval > 0:
            info["statistics"] = statistics
        if not project_states[1]:
            info["num_breaks"] = num_breaks
            num_breaks_total = np.sum(num_breaks)
            if breaks_buffer_out.size >= num_breaks_total:
                info["complete_break_specification"] = True
                info["break_positions"] = breaks_buffer_out[:num_breaks_total]
            else:
                info["complete_break_specification"] = False
                info["break_positions"] = breaks_buffer_out
        if make_info_json_serializable:
            info["beta_range"] = np.array(info["beta_range"]).tolist()
            for key in ("statistics", "num_breaks", "break_positions"):
                if key in info:
                    info[key] = info[key].tolist()
        response = dimodlite.SampleSet.from_samples((samples, variable_order), energy=energies + off,
                                                    info=info, vartype=dimodlite.SPIN)
        response.change_vartype(original_vartype, inplace=True)
        response.info.update(timing=dict(preprocessing_ns=t_sample - t_pre, sampling_ns=t_post - t_sample,
                                         postprocessing_ns=perf_counter_ns() - t_post))
        return response


class RotorModelAnnealingSampler(dimodlite.Sampler, dimodlite.Initialized):
    """Rotor-model (planar spin) annealing sampler for binary quadratic models,
    on MI355X.  Same interface as ``dwave.samplers.sqa.RotorModelAnnealingSampler``
    (sqa/sampler.py:572-1132); the anneal runs in the HIP rotor engine, read r
    on its own xorshift128+ stream seeded from the call's seed."""

    parameters = None
    properties = None

    def __init__(self):
        self.parameters = {"beta_range": [], "num_reads": [], "num_sweeps": [], "num_sweeps_per_beta": [],
                           "beta_schedule_type": ["beta_schedule_options"], "seed": [],
                           "interrupt_function": [], "initial_states": [], "initial_states_generator": []}
        self.properties = {"beta_schedule_options": ("linear", "geometric", "custom")}

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
        t_pre = perf_counter_ns()
        original_vartype = bqm.vartype
        if dimodlite.as_vartype(bqm.vartype) is not dimodlite.SPIN:
            bqm = bqm.change_vartype(dimodlite.SPIN if isinstance(bqm, dimodlite.BinaryQuadraticModel)
                                     else "SPIN", inplace=False)
        if trans_fields is None:
            trans_fields = np.ones(bqm.num_variables)
        elif isinstance(trans_fields, (list, tuple, np.ndarray)):
            if len(trans_fields) != bqm.num_variables:
                raise ValueError("bqm.num_variables and trans_fields lengths incompatible")
            trans_fields = np.array(trans_fields)
        else:
            trans_fields = trans_fields * np.ones(bqm.num_variables)
        if seed is None:
            seed = int(np.random.randint(2**31))
        elif not isinstance(seed, Integral):
            raise TypeError("'seed' should be None or an integer between 0 and 2^32 - 1: value = {}".format(seed))
        elif not (0 <= seed < 2**31):
            raise ValueError("'seed' should be an integer between 0 and 2^32 - 1: value = {}".format(seed))
        prng = np.random.default_rng(seed)   # reproducible projections (:857-858)
        if type(initial_states) == np.ndarray:
            parsed = self.parse_initial_states(bqm, num_reads=num_reads, initial_states=None,
                                               initial_states_generator=initial_states_generator, seed=seed)
            if num_reads is None:
                num_reads = initial_states.shape[0]
            if (num_reads, bqm.num_variables) != initial_states.shape:
                raise ValueError("Shape of initial_states specified asndarray should match (num_reads, num_vars)")
        else:
            parsed = self.parse_initial_states(bqm, num_reads=num_reads, initial_states=initial_states,
                                               initial_states_generator=initial_states_generator, seed=seed)
            num_reads = parsed.num_reads
            initial_states = parsed.initial_states.record.sample
        variable_order = parsed.initial_states.variables
        if initial_states.dtype == np.uint8:
            if project_states[0]:
                initial_states_array = np.ascontiguousarray(
                    128 * (np.cos(np.pi * (initial_states / 128)) < 1 - 2 * prng.random(size=initial_states.shape)),
                    dtype=np.uint8)
            else:
                initial_states_array = np.ascontiguousarray(initial_states)
        else:
            if not set(np.unique(initial_states)).issubset({-1, 1}):
                raise ValueError("initial_states dtype is neither np.uint8 nor consistent with spins {-1, 1}")
            initial_states_array = np.ascontiguousarray(64 - 64 * initial_states, dtype=np.uint8)
        ldata, (irow, icol, qdata), off = bqm.to_numpy_vectors(variable_order=variable_order)
        if interrupt_function and not callable(interrupt_function):
            raise TypeError("'interrupt_function' should be a callable")
        if not isinstance(num_sweeps_per_beta, Integral):
            raise TypeError("'num_sweeps_per_beta' should be a positive integer: value = {}".format(
                num_sweeps_per_beta))
        elif num_sweeps_per_beta < 1:
            raise ValueError("'num_sweeps_per_beta' should be a positive integer: value = {}".format(
                num_sweeps_per_beta))

        if beta_schedule_type == "custom":
            if Hp_field is None:
                raise ValueError("'Hp_field' must be provided for beta_schedule_type = 'custom': value is None")
            try:
                Hp_field = np.array(Hp_field, dtype=float)
            except Exception:
                raise ValueError("Hp_field cannot be case as a numpy array of dtype=float")
            if num_sweeps is not None and num_sweeps != len(Hp_field) * num_sweeps_per_beta:
                raise ValueError("'num_sweeps' should be set to None, or a value consistent with 'Hp_field' "
                                 "and 'num_sweeps_per_beta' for 'beta_schedule_type' = 'custom': value = "
                                 f"{num_sweeps} != {Hp_field.size}x{num_sweeps_per_beta}")
            elif beta_range is not None and (beta_range[0] != Hp_field[0] or beta_range[-1] != Hp_field[-1]):
                raise ValueError("'beta_range' should be set to None, or a value consistent with 'Hp_field', "
                                 "for 'beta_schedule_type'='custom'.")
            elif np.min(Hp_field) < 0:
                raise ValueError("'Hp_field' cannot include negative values.")
        elif num_sweeps == 0:
            Hp_field = []
        else:
            if Hp_field is not None:
                raise ValueError("'Hp_field' must be set to None for 'beta_schedule_type' not equal to 'custom'")
            elif num_sweeps is None:
                num_sweeps = 100
            num_betas, rem = divmod(num_sweeps, num_sweeps_per_beta)
            if rem > 0 or num_betas < 0:
                raise ValueError("'num_sweeps' must be a positive value divisible by 'num_sweeps_per_beta'.")
            if beta_range is None:
                beta_range = _default_ising_beta_range(bqm.linear, bqm.quadratic)
            elif len(beta_range) != 2 or min(beta_range) < 0:
                raise ValueError("'beta_range' should be a 2-tuple, or 2 element list of positive numbers. "
                                 "The latter value is the target value.")
            if num_betas == 1:
                Hp_field = np.array([beta_range[-1]], dtype=float)
            elif beta_schedule_type == "linear":
                Hp_field = np.linspace(*beta_range, num=num_betas)
            elif beta_schedule_type == "geometric":
                if min(beta_range) <= 0:
                    raise ValueError("'beta_range' must contain non-zero values for 'beta_schedule_type' = "
                                     "'geometric'")
                Hp_field = np.geomspace(*beta_range, num=num_betas)
            else:
                raise ValueError("Beta schedule type {} not implemented".format(beta_schedule_type))
        num_betas = len(Hp_field)
        if Hd_field is None:
            Hd_field = np.zeros(num_betas)
        else:
            Hd_field = np.array(Hd_field, dtype=float)
            if len(Hd_field) < num_betas:
                raise ValueError("Hd_field incompatible with Hp_field")
        if schedule_sample_interval is None:
            schedule_sample_interval = 0
        elif schedule_sample_interval > len(Hp_field) or schedule_sample_interval < 1:
            raise ValueError("schedule_sample_interval is incompatible with Hp_field")

        t_sample = perf_counter_ns()
        samples, energies, statistics = rotor_annealing(
            num_reads, ldata, irow, icol, qdata, trans_fields, num_sweeps_per_beta, Hp_field, Hd_field, seed,
            initial_states_array, randomize_order, proposal_acceptance_criteria, schedule_sample_interval,
            interrupt_function)
        t_post = perf_counter_ns()

        info = {"beta_range": beta_range, "beta_schedule_type": beta_schedule_type}
        if schedule_sample_interval > 0:
            info["statistics"] = statistics
        if not project_states[1]:
            info["rotor_states"] = samples.copy()
        if make_info_json_serializable:
            info["beta_range"] = np.array(info["beta_range"]).tolist()
            for key in ("statistics", "rotor_states"):
                if key in info:
                    info[key] = info[key].tolist()
        # states are always projected to spins (:1105-1109)
        samples = np.array(1 - 2 * (np.cos(np.pi * samples / 128) < 1 - 2 * prng.random(size=samples.shape)),
                           dtype=np.int8)
        response = dimodlite.SampleSet.from_samples((samples, variable_order), energy=energies + off,
                                                    info=info, vartype=dimodlite.SPIN)
        response.change_vartype(original_vartype, inplace=True)
        response.info.update(timing=dict(preprocessing_ns=t_sample - t_pre, sampling_ns=t_post - t_sample,
                                         postprocessing_ns=perf_counter_ns() - t_post))
        return response
