"""Core-function mirror of ``dwave.samplers.sqa.rotormc_annealing``.

Same signature, argument meaning and return value as the reference's Cython
binding (dwave/samplers/sqa/rotormc_annealing.pyx:53-235); the anneal runs in
the HIP rotor engine through the C ABI (``misa_rotor_sample``).  States are
angles i in [0, 256) (theta = 2 pi i / 256) as uint8; energies are the
cos-weighted classical energies of the final angles.

Differences, by design:

* the reference shares one xorshift128+ stream across reads in order
  (cpu_rotormc.cpp:407-408, 449-558); here read r runs its own stream seeded
  with ``read_seeds(seed, num_samples)[r]``, i.e. read r equals the reference
  called with ``num_samples=1`` and that seed;
* statistics have the shape the Cython layer allocates,
  (num_samples, 1 + (len(Hd_field) - 1) // interval, num_vars); the reference
  core strides reads by len(Hp_field) // interval snapshots, which disagrees
  with that shape when the interval does not divide the schedule length, and
  leaves unwritten entries uninitialised (here they are zero).
"""
from __future__ import annotations

import ctypes
from numbers import Integral

import numpy as np

from .. import _lib

__all__ = ["simulated_annealing", "PROPOSALS"]

# cpu_rotormc.h:36-41 (enum order) and rotormc_annealing.pyx:196-205 (names)
PROPOSALS = {"GibbsNonErgodic": 0, "MetropolisNonErgodic": 1, "MetropolisUniform": 2, "MetropolisTF": 3}


def _interrupt_thunk(fn):
    """C callback calling ``fn``; an exception counts as an interrupt
    (rotormc_annealing.pyx:237-243)."""
    if fn is None:
        return _lib.INTERRUPT_CB()

    def call(_ctx):
        try:
            return 1 if fn() else 0
        except Exception:
            return 1

    return _lib.INTERRUPT_CB(call)


def simulated_annealing(num_samples, h, coupler_starts, coupler_ends, coupler_weights, trans_fields,
                        sweeps_per_beta, Hp_field, Hd_field, seed, states_numpy, randomize_order,
                        proposal_acceptance_criteria, schedule_sample_interval, interrupt_function=None, *,
                        read_seeds=None, stats=None):
    """Rotor-model anneal of ``num_samples`` reads on the GPU.

    Arguments as in rotormc_annealing.pyx:53-150: ``states_numpy`` (uint8,
    (num_samples, num_vars), C-contiguous) holds the initial angles and is
    overwritten with the final ones.  Keyword-only extensions: ``read_seeds``
    (uint64 per read) and ``stats`` (dict filled with launch facts, accepted
    proposals and the kernel time).

    Returns ``(states[:num], energies[:num], statistics[:num])``; ``num``
    counts the reads completed before ``interrupt_function`` returned True.
    """
    num_vars = len(h)
    if num_samples * num_vars == 0:
        return (np.empty((num_samples, num_vars), dtype=np.uint8), np.zeros(num_samples, dtype=np.double),
                np.empty(0, dtype=np.uint8))
    if not isinstance(states_numpy, np.ndarray) or states_numpy.dtype != np.uint8:
        raise ValueError("Buffer dtype mismatch, expected 'uint8_t'")
    if states_numpy.ndim != 2:
        raise ValueError("Buffer has wrong number of dimensions (expected 2)")
    if not states_numpy.flags.c_contiguous:
        raise ValueError("ndarray is not C-contiguous")
    if states_numpy.shape != (num_samples, num_vars):
        raise ValueError("states_numpy must have shape (num_samples, num_variables)")
    if proposal_acceptance_criteria not in PROPOSALS:
        raise ValueError("Unknown proposal_acceptance_criteria")
    proposal = PROPOSALS[proposal_acceptance_criteria]
    tf = np.ascontiguousarray(trans_fields, dtype=np.float64).ravel()
    if tf.size != num_vars:
        raise ValueError("trans_fields must have one value per variable")
    hp = np.ascontiguousarray(Hp_field, dtype=np.float64).ravel()
    hd_all = np.ascontiguousarray(Hd_field, dtype=np.float64).ravel()
    if hd_all.size < hp.size:
        raise ValueError("Hd_field incompatible with Hp_field")
    hd = np.ascontiguousarray(hd_all[:hp.size])
    interval = int(schedule_sample_interval or 0)
    if interval < 0:
        raise ValueError("schedule_sample_interval must be non-negative")

    # rotormc_annealing.pyx:171-176
    if interval:
        statistics = np.zeros((num_samples, 1 + (hd_all.size - 1) // interval, num_vars), dtype=np.uint8)
    else:
        statistics = np.zeros((num_samples, 1, 1), dtype=np.uint8)
    ncol_run = (1 + (hp.size - 1) // interval) if (interval and hp.size) else 0
    stats_buf = np.zeros((num_samples, ncol_run, num_vars), dtype=np.uint8) if ncol_run else None

    if read_seeds is None:
        if not isinstance(seed, Integral) or not (0 <= int(seed) < 2**64):
            raise OverflowError("seed must be an unsigned 64-bit integer")
        seeds = _lib.read_seeds(int(seed), num_samples)
    else:
        seeds = np.ascontiguousarray(read_seeds, dtype=np.uint64)
        if seeds.shape != (num_samples,):
            raise ValueError("read_seeds must have one seed per read")

    problem = _lib.Problem(h, coupler_starts, coupler_ends, coupler_weights)
    energies = np.empty(num_samples, dtype=np.float64)
    cb = _interrupt_thunk(interrupt_function)
    st = _lib.RotorStats()
    try:
        num = _lib.check(_lib.lib.misa_rotor_sample(
            problem.handle, _lib.ptr(states_numpy), _lib.ptr(energies), int(num_samples), _lib.ptr(tf),
            int(sweeps_per_beta), _lib.ptr(hp), _lib.ptr(hd), hp.size, _lib.ptr(seeds),
            1 if randomize_order else 0, proposal, _lib.ptr(stats_buf) if stats_buf is not None else None,
            interval if stats_buf is not None else 0, cb, None, ctypes.byref(st)))
    finally:
        problem.close()
    if stats_buf is not None:
        statistics[:, :ncol_run, :] = stats_buf
    if stats is not None:
        stats.update(kernel_ms=float(st.kernel_ms), accepted=int(st.accepted),
                     lanes_per_read=int(st.lanes_per_read), reads_per_block=int(st.reads_per_block),
                     threads_per_block=int(st.threads_per_block), csr_in_lds=int(st.csr_in_lds),
                     lds_bytes=int(st.lds_bytes), wave_per_read=int(st.wave_per_read),
                     run_kernel=int(st.run_kernel),
                     kernel=("rotor_run_kernel" if st.run_kernel else
                             "rotor_wave_kernel" if st.wave_per_read else "rotor_kernel"))
    return states_numpy[:num], energies[:num], statistics[:num]
