"""Differentiable fixed-beta sampling of Ising models (PyTorch autograd).

For a model with energy E_theta(s) = sum_i h_i s_i + sum_c J_c s_u s_v and
free energy F = -(1/beta) log sum_s exp(-beta E_theta(s)),

    dF/dh_i = <s_i>,        dF/dJ_c = <s_u s_v>        (Boltzmann averages).

``FixedBetaSampler`` draws R samples per model with the engine's Gibbs sweeps
at constant beta (the reference's ``simulated_annealing`` with
proposal_acceptance_criteria="Gibbs" and a custom schedule [beta] * S,
dwave/samplers/sa/sampler.py:373-392; every read is bit-identical to that
call with num_samples=1 and the read's seed) and returns, per model, the mean
sample energy L(theta) = (1/R) sum_r E_theta(s_r) with the samples held fixed.
Its gradient is exactly the moment estimate of dF/dtheta above, computed on
the GPU by the moment kernel (``misa_moments_run``) in the backward pass.

Forward and backward run on ``torch.cuda.current_stream()`` with no host
copy and no host synchronisation: the batch, its plans and CSR are built once,
the float32 parameters are widened to float64 on the device
(``misa_batch_set_values_device``), the read seeds and random initial states
are drawn on the device (``misa_read_seeds_device``,
``misa_random_spins_device``), and the sweeps run on buffers the sampler
keeps, so a call allocates no device memory of its own beyond the returned
copies.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .sa.batch import Batch

__all__ = ["FixedBetaSampler", "MomentPlan"]


class MomentPlan:
    """Owning handle of a ``misa_moments_plan`` for fixed model sizes."""

    def __init__(self, num_vars: Sequence[int], num_edges: Sequence[int], num_reads: Sequence[int]):
        self.num_vars = np.ascontiguousarray(num_vars, dtype=np.int32)
        self.num_edges = np.ascontiguousarray(num_edges, dtype=np.int64)
        self.num_reads = np.ascontiguousarray(num_reads, dtype=np.int32)
        if not (len(self.num_vars) == len(self.num_edges) == len(self.num_reads)):
            raise ValueError("one size per model")
        out = ctypes.c_void_p()
        _lib.check(_lib.lib.misa_moments_plan_create(len(self.num_vars), _lib.ptr(self.num_vars),
                                                     _lib.ptr(self.num_edges), _lib.ptr(self.num_reads),
                                                     ctypes.byref(out)))
        self.handle = out

    def run(self, d_samples: int, d_edges: int, d_scale: int = 0, d_grad_h: int = 0, d_grad_J: int = 0,
            d_sums: int = 0, stream: int = 0):
        _lib.check(_lib.lib.misa_moments_run(self.handle, ctypes.c_void_p(d_samples or None),
                                             ctypes.c_void_p(d_edges or None), ctypes.c_void_p(d_scale or None),
                                             ctypes.c_void_p(d_grad_h or None), ctypes.c_void_p(d_grad_J or None),
                                             ctypes.c_void_p(d_sums or None), ctypes.c_void_p(stream or None)))

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib.misa_moments_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FixedBetaSampler:
    """Samples a fixed list of Ising models (structure fixed, values free) at
    constant beta and differentiates the mean sample energy.

    Args:
        num_vars: variables per model.
        edges: per model, an (m_b, 2) integer array of couplers (u, v).
        num_reads: reads per model (one value or one per model).
        num_sweeps: Gibbs sweeps per read.
        beta: inverse temperature.
        device: torch device (a gfx950 GPU).
    """

    def __init__(self, num_vars: Sequence[int], edges: Sequence, num_reads=4096, num_sweeps: int = 2000,
                 beta: float = 1.0, device="cuda:0"):
        import torch
        self.torch = torch
        self.device = torch.device(device)
        self.num_vars = [int(n) for n in num_vars]
        self.edges = [np.ascontiguousarray(e, dtype=np.int32).reshape(-1, 2) for e in edges]
        if len(self.edges) != len(self.num_vars):
            raise ValueError("one edge list per model")
        for n, e in zip(self.num_vars, self.edges):
            if e.size and (e.min() < 0 or e.max() >= n):
                raise ValueError("coupler indexes contain an invalid variable")
        B = len(self.num_vars)
        self.num_reads = [int(r) for r in np.broadcast_to(np.asarray(num_reads), (B,))]
        self.num_sweeps = int(num_sweeps)
        self.beta = float(beta)
        self.h_off = np.concatenate([[0], np.cumsum(self.num_vars)]).astype(np.int64)
        self.j_off = np.concatenate([[0], np.cumsum([len(e) for e in self.edges])]).astype(np.int64)
        self.plan = MomentPlan(self.num_vars, [len(e) for e in self.edges], self.num_reads)
        allp = np.concatenate(self.edges, axis=0) if B else np.zeros((0, 2), np.int32)
        self.d_edges = torch.from_numpy(np.ascontiguousarray(allp, dtype=np.int32)).to(self.device)
        self.last_samples = None
        # the batch of sweeps, built once on the coupler structure (values are
        # written on the device by every call)
        problems = [(np.zeros(n), e[:, 0], e[:, 1], np.zeros(len(e))) for n, e in zip(self.num_vars, self.edges)]
        self.batch = Batch(problems, self.num_reads, 1, [np.full(self.num_sweeps, self.beta)] * B,
                           proposal_acceptance_criteria="Gibbs")
        self.d_states = torch.empty(max(self.batch.total_state, 1), dtype=torch.int8, device=self.device)
        self.d_energies = torch.empty(max(self.batch.total_reads, 1), dtype=torch.float64, device=self.device)
        self.d_seeds = torch.empty(max(self.batch.total_reads, 1), dtype=torch.int64, device=self.device)

    def sample(self, h, J, seed: int = 0, initial_states: Optional[object] = None):
        """Sample all models.  ``h`` (sum n_b,) and ``J`` (sum m_b,) tensors
        (float32 values, models concatenated in order).  Returns
        ``(energies, samples)``: per-read float64 energies (sum R_b,) and int8
        samples (sum R_b n_b,), both device tensors, models concatenated.
        Without ``initial_states`` the reads start from
        ``misa_random_spins_device(seed)``."""
        torch = self.torch
        h32 = h.detach().to(self.device, torch.float32).contiguous()
        J32 = J.detach().to(self.device, torch.float32).contiguous()
        if h32.numel() != int(self.h_off[-1]) or J32.numel() != int(self.j_off[-1]):
            raise ValueError("h and J must hold every model's biases")
        total_state, total_reads = self.batch.total_state, self.batch.total_reads
        stream = torch.cuda.current_stream(self.device).cuda_stream
        self.batch.set_values_device(h32.data_ptr(), J32.data_ptr(), stream)
        _lib.check(_lib.lib.misa_read_seeds_device(ctypes.c_uint64(int(seed) & 0xFFFFFFFFFFFFFFFF), total_reads,
                                                   ctypes.c_void_p(self.d_seeds.data_ptr()),
                                                   ctypes.c_void_p(stream)))
        if initial_states is None:
            _lib.check(_lib.lib.misa_random_spins_device(ctypes.c_uint64(int(seed) & 0xFFFFFFFFFFFFFFFF),
                                                         total_state, ctypes.c_void_p(self.d_states.data_ptr()),
                                                         ctypes.c_void_p(stream)))
        else:
            init = initial_states.to(self.device, torch.int8).reshape(-1)
            if init.numel() != total_state:
                raise ValueError("initial_states must hold every model's reads")
            self.d_states[:total_state].copy_(init)
        self.batch.run_device(self.d_states.data_ptr(), self.d_energies.data_ptr(), self.d_seeds.data_ptr(),
                              stream=stream)
        return self.d_energies[:total_reads].clone(), self.d_states[:total_state].clone()

    def __call__(self, h, J, seed: int = 0, initial_states=None):
        """Per-model mean sample energy (float32, differentiable in h and J)
        and the samples (int8, models concatenated)."""
        out = _FixedBetaEnergy.apply(h, J, self, int(seed), initial_states)
        return out, self.last_samples

    def moments(self, samples, scale=None):
        """(grad_h, grad_J, sums) = scale_b * <s>, scale_b * <s s> and the exact
        int64 sums, for samples laid out as ``sample`` returns them."""
        torch = self.torch
        gh = torch.empty(int(self.h_off[-1]), dtype=torch.float32, device=self.device)
        gj = torch.empty(int(self.j_off[-1]), dtype=torch.float32, device=self.device)
        sums = torch.empty(int(self.h_off[-1] + self.j_off[-1]), dtype=torch.int64, device=self.device)
        sc = None
        if scale is not None:
            sc = scale.detach().to(self.device, torch.float32).contiguous()
        samples = samples.contiguous()
        self.plan.run(samples.data_ptr(), self.d_edges.data_ptr(), sc.data_ptr() if sc is not None else 0,
                      gh.data_ptr(), gj.data_ptr(), sums.data_ptr(),
                      torch.cuda.current_stream(self.device).cuda_stream)
        return gh, gj, sums

    def close(self):
        self.plan.close()
        self.batch.close()


def _model_means(sampler, energies):
    torch = sampler.torch
    means = []
    o = 0
    for R in sampler.num_reads:
        means.append(energies[o:o + R].mean() if R else torch.zeros((), dtype=energies.dtype, device=energies.device))
        o += R
    return torch.stack(means).to(torch.float32)


def _autograd_function():
    import torch

    class FixedBetaEnergy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, h, J, sampler, seed, initial_states):
            energies, samples = sampler.sample(h, J, seed, initial_states)
            sampler.last_samples = samples
            ctx.sampler = sampler
            ctx.save_for_backward(samples)
            return _model_means(sampler, energies)

        @staticmethod
        def backward(ctx, grad_out):
            (samples,) = ctx.saved_tensors
            gh, gj, _ = ctx.sampler.moments(samples, grad_out)
            return gh, gj, None, None, None

    return FixedBetaEnergy


class _Lazy:
    fn = None

    @classmethod
    def apply(cls, *args):
        if cls.fn is None:
            cls.fn = _autograd_function()
        return cls.fn.apply(*args)


_FixedBetaEnergy = _Lazy
