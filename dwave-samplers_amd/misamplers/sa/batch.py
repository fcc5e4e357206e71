"""Batched simulated annealing: many Ising problems of uneven size in one launch.

No reference counterpart (the reference samples one BQM per call,
dwave/samplers/sa/simulated_annealing.pyx:45).  Each problem has its own
couplers, reads, schedule and seed; per-read streams follow the same rule as
``simulated_annealing`` (read r of problem b is seeded with
``read_seeds(seeds[b], num_reads[b])[r]``), so a problem's samples equal the
plain engine's for that problem run in float64 mode.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .. import _lib
from .simulated_annealing import _proposal_code

__all__ = ["Batch", "simulated_annealing_batch"]


class Batch:
    """Owning handle of a ``misa_batch`` (the problems are kept alive here)."""

    def __init__(self, problems: Sequence[Tuple], num_reads, sweeps_per_beta, beta_schedules,
                 randomize_order=False, proposal_acceptance_criteria="Metropolis"):
        k = len(problems)
        self.problems = [p if isinstance(p, _lib.Problem) else _lib.Problem(*p) for p in problems]
        self.num_reads = np.broadcast_to(np.asarray(num_reads, dtype=np.int32), (k,)).copy()
        self.spb = np.broadcast_to(np.asarray(sweeps_per_beta, dtype=np.int32), (k,)).copy()
        self.betas = [np.ascontiguousarray(b, dtype=np.float64).ravel() for b in beta_schedules]
        if len(self.betas) != k:
            raise ValueError("one beta schedule per problem")
        self.nbetas = np.array([len(b) for b in self.betas], dtype=np.int32)
        handles = (ctypes.c_void_p * max(k, 1))(*[p.handle.value for p in self.problems])
        bptrs = (ctypes.c_void_p * max(k, 1))(*[b.ctypes.data if b.size else None for b in self.betas])
        out = ctypes.c_void_p()
        _lib.check(_lib.lib.misa_batch_create(
            k, ctypes.cast(handles, ctypes.c_void_p), _lib.ptr(self.num_reads), _lib.ptr(self.spb),
            ctypes.cast(bptrs, ctypes.c_void_p), _lib.ptr(self.nbetas),
            _lib.RANDOM if randomize_order else _lib.SEQUENTIAL, _proposal_code(proposal_acceptance_criteria),
            ctypes.byref(out)))
        self.handle = out
        self.state_offsets = [int(_lib.lib.misa_batch_state_offset(out, b)) for b in range(k)]
        self.read_offsets = [int(_lib.lib.misa_batch_read_offset(out, b)) for b in range(k)]
        self.total_reads = int(_lib.lib.misa_batch_total_reads(out))
        self.total_state = int(_lib.lib.misa_batch_total_state(out))

    def set_policy(self, policy: int):
        _lib.check(_lib.lib.misa_batch_set_policy(self.handle, int(policy)))

    def autotune(self, trial_sweeps: int = 8) -> int:
        return _lib.check(_lib.lib.misa_batch_autotune(self.handle, int(trial_sweeps)))

    def run_device(self, d_states: int, d_energies: int, d_seeds: int, d_counters: Optional[int] = None,
                   stream: Optional[int] = None, use_graph: bool = False):
        _lib.check(_lib.lib.misa_batch_run_device(
            self.handle, ctypes.c_void_p(d_states), ctypes.c_void_p(d_energies), ctypes.c_void_p(d_seeds),
            ctypes.c_void_p(d_counters or 0), ctypes.c_void_p(stream or 0), 1 if use_graph else 0))

    def set_values_device(self, d_h: int, d_J: int, stream: Optional[int] = None):
        """New biases from device memory (float32: every problem's h, then
        every problem's couplers in input order, problems concatenated), on
        ``stream`` (``misa_batch_set_values_device``)."""
        _lib.check(_lib.lib.misa_batch_set_values_device(self.handle, ctypes.c_void_p(d_h or 0),
                                                         ctypes.c_void_p(d_J or 0), ctypes.c_void_p(stream or 0)))

    def sample(self, states: np.ndarray, seeds: np.ndarray, stats: Optional[dict] = None):
        """Anneal the concatenated states in place; returns the energies."""
        if states.dtype != np.int8 or not states.flags.c_contiguous or states.size != self.total_state:
            raise ValueError("states must be a contiguous int8 array of all problems' reads")
        seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        if seeds.size != self.total_reads:
            raise ValueError("one seed per read")
        energies = np.empty(self.total_reads, dtype=np.float64)
        st = _lib.SaStats()
        _lib.check(_lib.lib.misa_batch_sample(self.handle, _lib.ptr(states), _lib.ptr(energies), _lib.ptr(seeds),
                                              ctypes.byref(st)))
        if stats is not None:
            stats.update(flips=int(st.flips), flip_degree=int(st.flip_degree), kernel_ms=float(st.kernel_ms))
        return energies

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib.misa_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def simulated_annealing_batch(problems: Sequence[Tuple], num_reads, sweeps_per_beta, beta_schedules,
                              seeds: Sequence[int], initial_states: Sequence[np.ndarray],
                              randomize_order=False, proposal_acceptance_criteria="Metropolis",
                              *, policy: Optional[int] = None, autotune: int = 0,
                              stats: Optional[dict] = None) -> List[Tuple[np.ndarray, np.ndarray]]:
    """Anneal each problem ``(h, coupler_starts, coupler_ends, coupler_weights)``
    with its own reads, schedule and seed in one launch.  Returns a list of
    ``(states, energies)`` per problem."""
    batch = Batch(problems, num_reads, sweeps_per_beta, beta_schedules, randomize_order,
                  proposal_acceptance_criteria)
    try:
        if policy is not None:
            batch.set_policy(policy)
        if autotune:
            batch.autotune(autotune)
        parts, seed_parts = [], []
        for b, p in enumerate(batch.problems):
            R, n = int(batch.num_reads[b]), p.num_vars
            st = np.ascontiguousarray(initial_states[b], dtype=np.int8)
            if st.shape != (R, n):
                raise ValueError(f"initial_states[{b}] must have shape ({R}, {n})")
            parts.append(st.ravel())
            seed_parts.append(_lib.read_seeds(int(seeds[b]), R))
        states = np.ascontiguousarray(np.concatenate(parts) if parts else np.zeros(0, np.int8))
        seed_all = np.concatenate(seed_parts) if seed_parts else np.zeros(0, np.uint64)
        energies = batch.sample(states, seed_all, stats)
        out = []
        for b, p in enumerate(batch.problems):
            R, n = int(batch.num_reads[b]), p.num_vars
            so, ro = batch.state_offsets[b], batch.read_offsets[b]
            out.append((states[so:so + R * n].reshape(R, n).copy(), energies[ro:ro + R].copy()))
        return out
    finally:
        batch.close()
