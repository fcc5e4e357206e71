"""Core-function mirror of ``dwave.samplers.greedy.descent``.

Same signature, argument meaning and return value as the reference's Cython
binding (dwave/samplers/greedy/descent.pyx:24-112); the descents run on the GPU
through the C ABI (``misa_descent_sample``), one wave per read.  Results are
those of the reference's dense solver (descent.cpp:117-187); its
``large_sparse_opt`` solver (:201-292) reaches the same states, so the flag
only selects nothing here.  Self-loops are constant energy (the reference's
bookkeeping for them can flip a variable back and forth forever).
"""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _lib

__all__ = ["steepest_gradient_descent"]


def steepest_gradient_descent(num_samples, linear_biases, coupler_starts, coupler_ends, coupler_weights,
                              states_numpy, large_sparse_opt=False, *, max_steps=0xFFFFFFFF, stats=None):
    """Descend ``num_samples`` reads of an Ising problem to local minima.

    Args mirror descent.pyx:24-67; ``states_numpy`` (int8, C-contiguous,
    (num_samples, num_vars), values +-1) holds the initial states and is
    overwritten.  Keyword-only extensions: ``max_steps`` (flips per read) and
    ``stats`` (a dict receiving the kernel time).

    Returns ``(samples, energies, num_steps)``.
    """
    num_vars = len(linear_biases)
    if num_samples == 0 or num_vars == 0:
        return (np.empty((num_samples, num_vars), dtype=np.int8), np.zeros(num_samples, dtype=np.double),
                np.zeros(num_samples, dtype=np.uint32))
    if not isinstance(states_numpy, np.ndarray) or states_numpy.dtype != np.int8:
        raise ValueError("Buffer dtype mismatch, expected 'int8_t'")
    if states_numpy.ndim != 2:
        raise ValueError("Buffer has wrong number of dimensions (expected 2)")
    if not states_numpy.flags.c_contiguous:
        raise ValueError("ndarray is not C-contiguous")
    if states_numpy.shape != (num_samples, num_vars):
        raise ValueError("states_numpy must have shape (num_samples, num_variables)")
    problem = _lib.Problem(linear_biases, coupler_starts, coupler_ends, coupler_weights)
    energies = np.empty(num_samples, dtype=np.float64)
    steps = np.empty(num_samples, dtype=np.uint32)
    ms = ctypes.c_double()
    try:
        _lib.check(_lib.lib.misa_descent_sample(problem.handle, _lib.ptr(states_numpy), _lib.ptr(energies),
                                                _lib.ptr(steps), int(num_samples), int(max_steps),
                                                ctypes.byref(ms)))
    finally:
        problem.close()
    if stats is not None:
        stats.update(kernel_ms=float(ms.value), kernel="descent_kernel")
    return states_numpy, energies, steps
