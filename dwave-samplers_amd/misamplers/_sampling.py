"""Host-side steps shared by the drop-in annealing samplers.

``SimulatedAnnealingSampler``, ``PathIntegralAnnealingSampler`` and
``RotorModelAnnealingSampler`` prepare a call the same way: the model in spin
form, a checked seed, the initial states and problem arrays in one variable
order, an inverse-temperature schedule (given, or interpolated over a beta
range), and after sampling a response in the caller's vartype with timing
figures.  The pieces differ only in a few names and defaults, which
``ScheduleRules`` records per sampler.  Error messages and the order of the
checks are the reference samplers' (dwave/samplers/sa/sampler.py:321-426,
dwave/samplers/sqa/sampler.py:315-569 and :850-1132), since callers match on
them.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from numbers import Integral
from time import perf_counter_ns
from typing import Any, List, Optional, Tuple

import numpy as np

from . import dimodlite

__all__ = ["ScheduleRules", "SA_RULES", "PATH_RULES", "ROTOR_RULES", "check_seed", "check_interrupt",
           "check_sweeps_per_beta", "anneal_schedule", "transverse_schedule", "check_sample_interval",
           "spin_form", "SpinInputs", "spin_inputs", "break_buffers", "json_ready", "spin_response"]


# ---------------------------------------------------------------------------
# seeds and small argument checks
# ---------------------------------------------------------------------------

def check_seed(seed) -> int:
    """None draws a seed in [0, 2^31); anything else must be an integer in that
    range (the messages name 2^32 - 1, as the reference's do)."""
    if seed is None:
        return int(np.random.randint(2**31))
    if not isinstance(seed, Integral):
        raise TypeError(f"'seed' should be None or an integer between 0 and 2^32 - 1: value = {seed}")
    if not 0 <= seed < 2**31:
        raise ValueError(f"'seed' should be an integer between 0 and 2^32 - 1: value = {seed}")
    return int(seed)


def check_interrupt(fn) -> None:
    if fn and not callable(fn):
        raise TypeError("'interrupt_function' should be a callable")


def check_sweeps_per_beta(num_sweeps_per_beta) -> None:
    msg = f"'num_sweeps_per_beta' should be a positive integer: value = {num_sweeps_per_beta}"
    if not isinstance(num_sweeps_per_beta, Integral):
        raise TypeError(msg)
    if num_sweeps_per_beta < 1:
        raise ValueError(msg)


# ---------------------------------------------------------------------------
# inverse-temperature schedule
# ---------------------------------------------------------------------------

@dataclass(frozen=True)
class ScheduleRules:
    """How one sampler names and defaults its schedule.

    name: the keyword carrying a custom schedule ('beta_schedule' or 'Hp_field');
    default_sweeps: sweeps when num_sweeps is None;
    sweeps_word: how the divisibility message qualifies num_sweeps;
    mismatch_detail: the custom-length message ends with 'num_sweeps != len x per_beta';
    zero_sweeps_empty: num_sweeps == 0 (without a custom schedule) is an empty
        schedule before any other schedule check.
    """
    name: str
    default_sweeps: int
    sweeps_word: str
    mismatch_detail: bool = False
    zero_sweeps_empty: bool = False


SA_RULES = ScheduleRules("beta_schedule", 1000, "non-negative")
PATH_RULES = ScheduleRules("Hp_field", 100, "positive")
ROTOR_RULES = ScheduleRules("Hp_field", 100, "positive", mismatch_detail=True, zero_sweeps_empty=True)


def _custom_schedule(rules: ScheduleRules, schedule, beta_range, num_sweeps, num_sweeps_per_beta) -> np.ndarray:
    if schedule is None:
        raise ValueError(f"'{rules.name}' must be provided for beta_schedule_type = 'custom': value is None")
    try:
        values = np.array(schedule, dtype=float)
    except Exception:
        raise ValueError(f"{rules.name} cannot be case as a numpy array of dtype=float")
    if num_sweeps is not None and num_sweeps != len(values) * num_sweeps_per_beta:
        tail = f"{num_sweeps} != {values.size}x{num_sweeps_per_beta}" if rules.mismatch_detail else ""
        raise ValueError(f"'num_sweeps' should be set to None, or a value consistent with '{rules.name}' and "
                         f"'num_sweeps_per_beta' for 'beta_schedule_type' = 'custom': value = {tail}")
    if beta_range is not None and (beta_range[0] != values[0] or beta_range[-1] != values[-1]):
        raise ValueError(f"'beta_range' should be set to None, or a value consistent with '{rules.name}', for "
                         "'beta_schedule_type'='custom'.")
    if values.size and values.min() < 0:
        raise ValueError(f"'{rules.name}' cannot include negative values.")
    return values


def _interpolated(schedule_type: str, beta_range, num_betas: int) -> np.ndarray:
    lo, hi = beta_range[0], beta_range[-1]
    if num_betas == 1:
        return np.array([hi], dtype=float)  # one sweep in the target model
    if schedule_type == "linear":
        return np.linspace(lo, hi, num=num_betas)
    if schedule_type == "geometric":
        if min(beta_range) <= 0:
            raise ValueError("'beta_range' must contain non-zero values for 'beta_schedule_type' = 'geometric'")
        return np.geomspace(lo, hi, num=num_betas)
    raise ValueError(f"Beta schedule type {schedule_type} not implemented")


def anneal_schedule(rules: ScheduleRules, beta_range, num_sweeps, num_sweeps_per_beta, schedule_type, schedule,
                    linear, quadratic) -> Tuple[Any, Any]:
    """(schedule, beta_range): the given schedule checked, or num_sweeps /
    num_sweeps_per_beta values interpolated over beta_range (the default range
    of the spin-form biases when None)."""
    if schedule_type == "custom":
        return _custom_schedule(rules, schedule, beta_range, num_sweeps, num_sweeps_per_beta), beta_range
    if rules.zero_sweeps_empty and num_sweeps == 0:
        return np.zeros(0), beta_range
    if schedule is not None:
        raise ValueError(f"'{rules.name}' must be set to None for 'beta_schedule_type' not equal to 'custom'")
    if num_sweeps is None:
        num_sweeps = rules.default_sweeps
    num_betas, rem = divmod(num_sweeps, num_sweeps_per_beta)
    if rem > 0 or num_betas < 0:
        raise ValueError(f"'num_sweeps' must be a {rules.sweeps_word} value divisible by 'num_sweeps_per_beta'.")
    if beta_range is None:
        from .sa.sampler import _default_ising_beta_range
        beta_range = _default_ising_beta_range(linear, quadratic)
    elif len(beta_range) != 2 or min(beta_range) < 0:
        raise ValueError("'beta_range' should be a 2-tuple, or 2 element list of positive numbers. "
                         "The latter value is the target value.")
    return _interpolated(schedule_type, beta_range, num_betas), beta_range


def transverse_schedule(hd_field, num_steps: int) -> np.ndarray:
    """Hd per schedule step: zeros when not given, else at least num_steps values."""
    if hd_field is None:
        return np.zeros(num_steps)
    hd = np.array(hd_field, dtype=float)
    if len(hd) < num_steps:
        raise ValueError("Hd_field incompatible with Hp_field")
    return hd


def check_sample_interval(interval, num_steps: int) -> int:
    """0 (no statistics) when None, else 1..num_steps."""
    if interval is None:
        return 0
    if interval > num_steps or interval < 1:
        raise ValueError("schedule_sample_interval is incompatible with Hp_field")
    return int(interval)


# ---------------------------------------------------------------------------
# the problem in spin form, initial states and arrays in one variable order
# ---------------------------------------------------------------------------

def spin_form(bqm):
    """The model in spin form (a copy when it is binary)."""
    if dimodlite.as_vartype(bqm.vartype) is dimodlite.SPIN:
        return bqm
    return bqm.change_vartype(dimodlite.SPIN if isinstance(bqm, dimodlite.BinaryQuadraticModel) else "SPIN",
                              inplace=False)


@dataclass
class SpinInputs:
    """What a sampler hands to its core call, and what its response needs."""
    bqm: Any                 # spin form
    original_vartype: Any
    seed: int
    num_reads: int
    states: np.ndarray       # (num_reads, n) in variable order
    variables: List[Any]
    h: np.ndarray
    couplers: Tuple[np.ndarray, np.ndarray, np.ndarray]
    offset: float
    t_start: int


def spin_inputs(sampler, bqm, seed, num_reads, initial_states, initial_states_generator, interrupt_function,
                states_dtype=np.int8) -> SpinInputs:
    """Spin form, seed, parsed initial states and the problem arrays, checked in
    the reference's order (vartype, seed, initial states, interrupt)."""
    t_start = perf_counter_ns()
    original = bqm.vartype
    spin = spin_form(bqm)
    seed = check_seed(seed)
    parsed = sampler.parse_initial_states(spin, num_reads=num_reads, initial_states=initial_states,
                                          initial_states_generator=initial_states_generator, seed=seed)
    states = np.ascontiguousarray(parsed.initial_states.record.sample, dtype=states_dtype)
    variables = parsed.initial_states.variables
    h, couplers, offset = spin.to_numpy_vectors(variable_order=variables)
    check_interrupt(interrupt_function)
    return SpinInputs(spin, original, seed, parsed.num_reads, states, variables, h, couplers, offset, t_start)


# ---------------------------------------------------------------------------
# path-integral break buffers
# ---------------------------------------------------------------------------

def _break_buffer_length(num_reads: int, num_vars: int, hd_last: Optional[float], breaks_given: int) -> int:
    """Entries reserved for the output break list when the caller gives no
    buffer: about (1 + final Hd) breaks per worldline plus slack
    (dwave/samplers/sqa/sampler.py:460-491 sizes it the same way); with no
    schedule the input breaks pass through unchanged."""
    if hd_last is None:
        return int(breaks_given)
    worldlines = num_reads * num_vars
    return int((1 + hd_last) * (worldlines + math.log(worldlines)) + 1000)


def break_buffers(states_shape, num_breaks, breaks_in, breaks_buffer_out, keep_worldlines: bool, hd_field,
                  num_steps: int):
    """(num_breaks, breaks_in, breaks_buffer_out) as intc arrays: break counts
    per worldline (zeros by default; the shape of the initial states), input
    break positions, and the output buffer (one entry when worldlines are
    projected, else the default length above)."""
    if num_breaks is None:
        num_breaks = np.zeros(states_shape, dtype=np.intc)
    elif num_breaks.shape != tuple(states_shape):
        raise ValueError("Path integral initialization is inconsistent")
    if breaks_in is None:
        breaks_in = np.zeros(max(1, int(num_breaks.sum())))
    if breaks_buffer_out is None:
        length = 1
        if keep_worldlines:
            length = _break_buffer_length(states_shape[0], states_shape[1],
                                          float(hd_field[-1]) if num_steps > 0 else None, int(np.sum(num_breaks)))
        breaks_buffer_out = np.zeros(length, dtype=np.intc)
    return (np.ascontiguousarray(num_breaks, dtype=np.intc), np.ascontiguousarray(breaks_in, dtype=np.intc),
            np.ascontiguousarray(breaks_buffer_out, dtype=np.intc))


# ---------------------------------------------------------------------------
# response
# ---------------------------------------------------------------------------

def json_ready(info: dict, array_keys) -> dict:
    """info with beta_range and the named array entries as plain lists."""
    info["beta_range"] = np.array(info["beta_range"]).tolist()
    for key in array_keys:
        if key in info:
            info[key] = np.asarray(info[key]).tolist()
    return info


def spin_response(inputs: SpinInputs, samples, energies, info: dict, t_sample: int, t_post: int):
    """SampleSet of spin samples with the offset added, in the caller's vartype,
    with the three timing figures in info."""
    response = dimodlite.SampleSet.from_samples((samples, inputs.variables), energy=energies + inputs.offset,
                                                info=info, vartype=dimodlite.SPIN)
    response.change_vartype(inputs.original_vartype, inplace=True)
    response.info.update(timing=dict(preprocessing_ns=t_sample - inputs.t_start, sampling_ns=t_post - t_sample,
                                     postprocessing_ns=perf_counter_ns() - t_post))
    return response
