"""Core-function mirror of ``dwave.samplers.sqa.pimc_annealing``.

Same signature, argument meaning and return value as the reference's Cython
binding (dwave/samplers/sqa/pimc_annealing.pyx:58-280); the anneal runs in the
HIP path-integral engine through the C ABI (``misa_pimc_sample``).

Differences, by design:

* imaginary time is discrete, ``num_trotter_slices`` (keyword-only, default
  32) slices instead of the reference's 2^16 (localPIMC.h:91); both sample
  the same transverse-field Ising model in the limit of many slices;
* the reference shares one mt19937 stream (32-bit ``seed``) across reads in
  order (localPIMC.cpp:680-711); here reads run in parallel and read r draws
  from streams keyed by ``read_seeds(seed, num_samples)[r]``;
* variables are visited colour class by colour class instead of in uniform
  random order (localPIMC.cpp:134-139).

Break positions use the reference's format (sorted positions in [0, 2^16) per
spin, flat over reads and spins); a kink between slices k-1 and k is reported
at (2k-1) * 2^16 / (2P).
"""
from __future__ import annotations

import ctypes
import os
from numbers import Integral

import numpy as np

from .. import _lib

__all__ = ["simulated_annealing"]


def _interrupt_thunk(fn):
    """C callback calling ``fn``; an exception counts as an interrupt
    (pimc_annealing.pyx:282-287)."""
    if fn is None:
        return _lib.INTERRUPT_CB()

    def call(_ctx):
        try:
            return 1 if fn() else 0
        except Exception:
            return 1

    return _lib.INTERRUPT_CB(call)


def _check_array(a, dtype, ndim, name):
    if not isinstance(a, np.ndarray) or a.dtype != dtype:
        raise ValueError(f"Buffer dtype mismatch for {name}, expected '{np.dtype(dtype).name}'")
    if a.ndim != ndim:
        raise ValueError(f"Buffer has wrong number of dimensions (expected {ndim}, got {a.ndim})")
    if not a.flags.c_contiguous:
        raise ValueError("ndarray is not C-contiguous")


def simulated_annealing(num_samples, h, coupler_starts, coupler_ends, coupler_weights, sweeps_per_beta,
                        Hp_field, Hd_field, Gamma, chain_coupler_strength, qubits_per_chain,
                        qubits_per_update, seed, states_numpy, project_inputs, project_outputs,
                        num_breaks_numpy, breaks_in_numpy, breaks_buffer_out_numpy,
                        schedule_sample_interval, interrupt_function=None, *, read_seeds=None,
                        num_trotter_slices=32, stats=None):
    """Path-integral anneal of ``num_samples`` reads on the GPU.

    Arguments as in pimc_annealing.pyx:58-80.  ``states_numpy`` (int8,
    (num_samples, num_vars)) holds the initial states (or first slices, with
    ``project_inputs`` False) and is overwritten with the projected samples.
    ``num_breaks_numpy`` (intc, same shape) and ``breaks_buffer_out_numpy``
    (intc, 1-D) receive the final worldlines when ``project_outputs`` is
    False.  Keyword-only extensions: ``read_seeds`` (uint64 per read),
    ``num_trotter_slices`` (1..32) and ``stats`` (dict filled with launch facts
    and the kernel time).

    Returns ``(states[:num], energies[:num], statistics[:num], num_breaks[:num],
    breaks_buffer_out)``; ``num`` counts the reads completed before
    ``interrupt_function`` returned True.
    """
    num_vars = len(h)
    if num_samples * num_vars == 0:
        return (np.empty((num_samples, num_vars), dtype=np.int8), np.zeros(num_samples, dtype=np.double),
                np.empty(0, dtype=np.int8), np.empty((num_samples, num_vars), dtype=np.int8),
                np.empty(0, dtype=np.int8))
    _check_array(states_numpy, np.int8, 2, "states_numpy")
    _check_array(num_breaks_numpy, np.intc, 2, "num_breaks_numpy")
    _check_array(breaks_in_numpy, np.intc, 1, "breaks_in_numpy")
    _check_array(breaks_buffer_out_numpy, np.intc, 1, "breaks_buffer_out_numpy")
    if states_numpy.shape != (num_samples, num_vars):
        raise ValueError("states_numpy must have shape (num_samples, num_variables)")
    if num_breaks_numpy.shape != states_numpy.shape:
        raise ValueError("num_breaks_numpy must have the shape of states_numpy")
    hp = np.ascontiguousarray(Hp_field, dtype=np.float64).ravel()
    hd_all = np.ascontiguousarray(Hd_field, dtype=np.float64).ravel()
    if hd_all.size < hp.size:
        raise ValueError("Hd_field incompatible with Hp_field")
    hd = np.ascontiguousarray(hd_all[:hp.size])
    interval = int(schedule_sample_interval or 0)

    # Statistics array as pimc_annealing.pyx:206-220 allocates it.
    if interval:
        ncp = 1 + (hd_all.size - 1) // interval
        statistics = np.zeros((num_samples, ncp, num_vars), dtype=np.int8)
    else:
        statistics = np.zeros((num_samples, 1, 1), dtype=np.int8)
    ncol_run = (1 + (max(hp.size, 1) - 1) // interval) if interval else 0
    stats_buf = (np.zeros((num_samples, ncol_run, num_vars), dtype=np.int8)
                 if interval else np.zeros(1, dtype=np.int8))

    if read_seeds is None:
        if not isinstance(seed, Integral):
            raise TypeError("an integer is required for 'seed'")
        if not (0 <= int(seed) < 2**32):
            raise OverflowError("seed must be an unsigned 32-bit integer")
        s = int(seed)
        if s == 0:  # localPIMC.cpp:560-569: seed 0 asks for a random device
            s = int.from_bytes(os.urandom(8), "little")
        seeds = _lib.read_seeds(s, num_samples)
    else:
        seeds = np.ascontiguousarray(read_seeds, dtype=np.uint64)
        if seeds.shape != (num_samples,):
            raise ValueError("read_seeds must have one seed per read")

    problem = _lib.Problem(h, coupler_starts, coupler_ends, coupler_weights)
    energies = np.empty(num_samples, dtype=np.float64)
    cb = _interrupt_thunk(interrupt_function)
    st = _lib.PimcStats()
    try:
        num = _lib.check(_lib.lib.misa_pimc_sample(
            problem.handle, _lib.ptr(states_numpy), _lib.ptr(energies), int(num_samples),
            1 if project_inputs else 0, 1 if project_outputs else 0, _lib.ptr(num_breaks_numpy),
            _lib.ptr(breaks_in_numpy), breaks_in_numpy.size, _lib.ptr(breaks_buffer_out_numpy),
            breaks_buffer_out_numpy.size, int(sweeps_per_beta), _lib.ptr(hp), _lib.ptr(hd), hp.size,
            float(Gamma), float(chain_coupler_strength), int(qubits_per_chain), int(qubits_per_update),
            _lib.ptr(seeds), int(num_trotter_slices), _lib.ptr(stats_buf) if interval else None,
            interval, cb, None, ctypes.byref(st)))
    finally:
        problem.close()
    if interval:
        statistics[:, :ncol_run, :] = stats_buf
    if stats is not None:
        stats.update(kernel_ms=float(st.kernel_ms), site_updates=int(st.site_updates),
                     num_slices=int(st.num_slices), reads_per_block=int(st.reads_per_block),
                     threads_per_block=int(st.threads_per_block), num_classes=int(st.num_classes),
                     kind={1: "signed", 0: "table"}[int(st.kind)], lds_bytes=int(st.lds_bytes),
                     max_degree=int(st.max_degree), blocks=int(st.blocks), kernel="pimc_kernel")
    return (states_numpy[:num], energies[:num], statistics[:num], num_breaks_numpy[:num],
            breaks_buffer_out_numpy)
