"""Core-function mirror of ``dwave.samplers.sa.simulated_annealing``.

Same signature, argument meaning and return value as the reference's Cython
binding (dwave/samplers/sa/simulated_annealing.pyx:45-189); the work is done
by the HIP engine through the C ABI (``misa_sa_sample``).

Difference in seeding, by design: the reference seeds one xorshift128+ stream
per call and lets it run on across reads (cpu_sa.cpp:283-284, :321-358), so
read r depends on every draw of reads 0..r-1.  Here reads run in parallel and
read r uses its own stream seeded with ``read_seeds(seed, num_samples)[r]``;
read r equals the reference called with ``num_samples=1`` and that seed.
"""
from __future__ import annotations

import ctypes
from numbers import Integral

import numpy as np

from .. import _lib

__all__ = ["simulated_annealing"]


def _proposal_code(name) -> int:
    low = name.lower() if isinstance(name, str) else name
    if low == "gibbs":
        return _lib.GIBBS
    if low == "metropolis":
        return _lib.METROPOLIS
    raise ValueError(f"Unknown proposal_acceptance_criteria: {name}")


def _interrupt_thunk(fn):
    """C callback that calls ``fn``; an exception counts as an interrupt
    (simulated_annealing.pyx:192-197)."""
    if fn is None:
        return _lib.INTERRUPT_CB()

    def call(_ctx):
        try:
            return 1 if fn() else 0
        except Exception:
            return 1

    return _lib.INTERRUPT_CB(call)


def simulated_annealing(num_samples, h, coupler_starts, coupler_ends,
                        coupler_weights, sweeps_per_beta, beta_schedule, seed,
                        states_numpy, randomize_order=False,
                        proposal_acceptance_criteria="Metropolis",
                        interrupt_function=None, *, read_seeds=None, mode=0,
                        lanes_per_read=0, stats=None):
    """Anneal ``num_samples`` reads of an Ising problem on the GPU.

    Args mirror ``simulated_annealing.pyx:45-50``; ``states_numpy`` (int8,
    C-contiguous, shape (num_samples, num_vars)) holds the initial states and
    is overwritten with the samples.  Keyword-only extensions: ``read_seeds``
    (explicit per-read seeds, uint64), ``mode`` (0 auto, 1 int8, 2 int16,
    3 real), ``lanes_per_read`` (0 auto) and ``stats`` (a dict filled with
    flip counts and kernel time).

    Returns ``(states_numpy[:num], energies[:num])`` with ``num`` the number of
    reads completed before ``interrupt_function`` returned True.
    """
    num_vars = len(h)
    if num_samples * num_vars == 0:
        return (np.empty((num_samples, num_vars), dtype=np.int8),
                np.zeros(num_samples, dtype=np.double))
    if not isinstance(states_numpy, np.ndarray) or states_numpy.dtype != np.int8:
        raise ValueError("Buffer dtype mismatch, expected 'int8_t'")
    if states_numpy.ndim != 2:
        raise ValueError("Buffer has wrong number of dimensions (expected 2)")
    if not states_numpy.flags.c_contiguous:
        raise ValueError("ndarray is not C-contiguous")
    if states_numpy.shape != (num_samples, num_vars):
        raise ValueError("states_numpy must have shape (num_samples, num_variables)")
    prop = _proposal_code(proposal_acceptance_criteria)
    order = _lib.RANDOM if randomize_order else _lib.SEQUENTIAL
    if read_seeds is None:
        if not isinstance(seed, Integral) or not (0 <= int(seed) < 2**64):
            raise OverflowError("seed must be an unsigned 64-bit integer")
        seeds = _lib.read_seeds(int(seed), num_samples)
    else:
        seeds = np.ascontiguousarray(read_seeds, dtype=np.uint64)
        if seeds.shape != (num_samples,):
            raise ValueError("read_seeds must have one seed per read")
    betas = np.ascontiguousarray(beta_schedule, dtype=np.float64).ravel()
    problem = _lib.Problem(h, coupler_starts, coupler_ends, coupler_weights)
    energies = np.empty(num_samples, dtype=np.float64)
    cb = _interrupt_thunk(interrupt_function)
    st = _lib.SaStats()
    num = _lib.check(_lib.lib.misa_sa_sample(
        problem.handle, _lib.ptr(states_numpy), _lib.ptr(energies), int(num_samples),
        int(sweeps_per_beta), _lib.ptr(betas), len(betas), _lib.ptr(seeds), order, prop,
        int(mode), int(lanes_per_read), cb, None, ctypes.byref(st) if stats is not None else None))
    if stats is not None:
        stats.update(flips=int(st.flips), flip_degree=int(st.flip_degree),
                     kernel_ms=float(st.kernel_ms), mode=int(st.mode),
                     lanes_per_read=int(st.lanes_per_read),
                     reads_per_block=int(st.reads_per_block),
                     state_in_lds=bool(st.state_in_lds),
                     kernel={1: "sa_int_kernel", 2: "sa_real_kernel", 4: "sa_seg_kernel",
                             5: "sa_chain_kernel", 6: "sa_dense_kernel"}.get(int(st.kernel), "sa_sweep_kernel"),
                     csr_in_lds=bool(st.csr_in_lds), waves_per_block=int(st.waves_per_block),
                     chain_form={1: "stepping", 2: "ring", 3: "few"}.get(int(st.chain_form)))
    problem.close()
    return states_numpy[:num], energies[:num]
