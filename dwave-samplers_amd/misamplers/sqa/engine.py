"""Device-buffer access to the path-integral engine (C ABI ``misa_pimc_plan_*``).

``PimcPlan`` holds the adjacency, colour classes and schedule tables of one
problem on the device; ``run_device`` anneals worldline words that already
live in device memory (for example torch tensors on ``cuda:0``) on a given
HIP stream, and ``project_device`` extracts slice 0 and its energy.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _lib

__all__ = ["PimcPlan", "colouring", "chain_colouring"]


def colouring(h, coupler_starts, coupler_ends, coupler_weights):
    """(number of classes, class of every variable) in the engine's update
    order; host-only."""
    prob = _lib.Problem(h, coupler_starts, coupler_ends, coupler_weights)
    out = np.zeros(max(prob.num_vars, 1), dtype=np.int32)
    try:
        nc = _lib.check(_lib.lib.misa_pimc_colouring(prob.handle, _lib.ptr(out)))
    finally:
        prob.close()
    return nc, out[:prob.num_vars]


def chain_colouring(h, coupler_starts, coupler_ends, coupler_weights, qubits_per_chain):
    """(number of classes, class of every chain) in the chain-update order of
    a plan with ``qubits_per_chain``; host-only."""
    prob = _lib.Problem(h, coupler_starts, coupler_ends, coupler_weights)
    out = np.zeros(max(prob.num_vars // int(qubits_per_chain), 1), dtype=np.int32)
    try:
        nc = _lib.check(_lib.lib.misa_pimc_chain_colouring(prob.handle, int(qubits_per_chain), _lib.ptr(out)))
    finally:
        prob.close()
    return nc, out[:prob.num_vars // int(qubits_per_chain)]


class PimcPlan:
    """``qubits_per_chain`` set: chain updates of that many consecutive qubits
    coupled by ``chain_coupler_strength`` (misa_pimc_plan_create_chains);
    otherwise site updates."""

    def __init__(self, h, coupler_starts, coupler_ends, coupler_weights, sweeps_per_beta, hp_field, hd_field,
                 num_slices=32, qubits_per_chain=None, chain_coupler_strength=0.0):
        self.problem = _lib.Problem(h, coupler_starts, coupler_ends, coupler_weights)
        self.hp = np.ascontiguousarray(hp_field, dtype=np.float64).ravel()
        self.hd = np.ascontiguousarray(hd_field, dtype=np.float64).ravel()
        if self.hd.size < self.hp.size:
            raise ValueError("Hd_field incompatible with Hp_field")
        self.hd = np.ascontiguousarray(self.hd[:self.hp.size])
        self.num_slices = int(num_slices)
        handle = ctypes.c_void_p()
        if qubits_per_chain is None:
            rc = _lib.lib.misa_pimc_plan_create(self.problem.handle, int(sweeps_per_beta), _lib.ptr(self.hp),
                                                _lib.ptr(self.hd), self.hp.size, self.num_slices,
                                                ctypes.byref(handle))
        else:
            rc = _lib.lib.misa_pimc_plan_create_chains(self.problem.handle, int(sweeps_per_beta),
                                                       _lib.ptr(self.hp), _lib.ptr(self.hd), self.hp.size,
                                                       self.num_slices, int(qubits_per_chain),
                                                       float(chain_coupler_strength), ctypes.byref(handle))
        if rc:
            self.problem.close()
            self.problem = None
        _lib.check(rc)
        self.handle = handle

    @property
    def num_vars(self) -> int:
        return self.problem.num_vars

    def stat_columns(self, interval: int) -> int:
        return 1 + (max(self.hp.size, 1) - 1) // interval if interval > 0 else 0

    def info(self, num_reads: int) -> dict:
        st = _lib.PimcStats()
        _lib.check(_lib.lib.misa_pimc_plan_info(self.handle, int(num_reads), ctypes.byref(st)))
        return dict(site_updates=int(st.site_updates), num_slices=int(st.num_slices),
                    reads_per_block=int(st.reads_per_block), threads_per_block=int(st.threads_per_block),
                    num_classes=int(st.num_classes),
                    kind={1: "signed", 0: "table", 3: "chain"}[int(st.kind)],
                    lds_bytes=int(st.lds_bytes), max_degree=int(st.max_degree), blocks=int(st.blocks))

    def run_device(self, d_worldlines: int, d_seeds: int, num_reads: int, d_statistics: int = 0,
                   interval: int = 0, stream: int = 0, d_flips: int = 0):
        """Enqueue the anneal of ``num_reads`` worldline rows (device pointers as
        ints); ``d_flips`` (u64 per read, zeroed by the caller) counts changed
        slice spins."""
        _lib.check(_lib.lib.misa_pimc_run_device(self.handle, ctypes.c_void_p(d_worldlines),
                                                 ctypes.c_void_p(d_seeds), int(num_reads),
                                                 ctypes.c_void_p(d_statistics or None), int(interval),
                                                 ctypes.c_void_p(d_flips or None), ctypes.c_void_p(stream or None)))

    def project_device(self, d_worldlines: int, num_reads: int, d_states: int = 0, d_energies: int = 0,
                       stream: int = 0):
        _lib.check(_lib.lib.misa_pimc_project_device(self.problem.handle, ctypes.c_void_p(d_worldlines),
                                                     int(num_reads), ctypes.c_void_p(d_states or None),
                                                     ctypes.c_void_p(d_energies or None),
                                                     ctypes.c_void_p(stream or None)))

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib.misa_pimc_plan_destroy(self.handle)
            self.handle = None
        if getattr(self, "problem", None) is not None:
            self.problem.close()
            self.problem = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
