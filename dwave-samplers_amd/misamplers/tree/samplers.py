"""Mirror of ``dwave.samplers.tree.TreeDecompositionSampler``
(dwave/samplers/tree/samplers.py:186-359).

Exact Boltzmann sampling and marginals of a binary quadratic model by bucket
elimination: the host fixes the elimination order (min-fill when none is
given, samplers.py:319-321) and the bucket scopes; the HIP kernels evaluate
the log-domain tables, the downward messages, the marginals and the samples
(``misa_tree_*``).  Same parameters, properties, ``info`` keys and errors as
the reference; read r's sample comes from its own xorshift128+ stream seeded
from ``seed`` (the reference draws all samples from one mt19937).
"""
from __future__ import annotations

import ctypes
import os
from typing import Any, List, Optional

import numpy as np

from .. import _lib, dimodlite

__all__ = ["TreeDecompositionSampler", "TreeDecompositionSolver", "TreePlan"]


class TreePlan:
    """Owning handle of a ``misa_tree_plan`` (elimination order, bucket tables)."""

    def __init__(self, h, coupler_starts, coupler_ends, coupler_weights, beta=1.0, low=-1, elimination_order=None,
                 max_treewidth=25):
        self.problem = _lib.Problem(h, coupler_starts, coupler_ends, coupler_weights)
        self.n = len(self.problem.h)
        order = None
        if elimination_order is not None:
            order = np.ascontiguousarray(elimination_order, dtype=np.int32)
            if order.shape != (self.n,):
                raise ValueError("elimination order is not a permutation of the variables")
        out = ctypes.c_void_p()
        try:
            _lib.check(_lib.lib.misa_tree_plan_create(self.problem.handle, _lib.ptr(order) if order is not None else None,
                                                      float(beta), int(low), int(max_treewidth), ctypes.byref(out)))
        except Exception:
            self.problem.close()
            raise
        self.handle = out
        tw, entries, npairs = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int()
        self.order = np.zeros(self.n, dtype=np.int32)
        _lib.check(_lib.lib.misa_tree_plan_info(self.handle, ctypes.byref(tw), ctypes.byref(entries),
                                                ctypes.byref(npairs), _lib.ptr(self.order)))
        self.treewidth, self.table_entries = int(tw.value), int(entries.value)
        self.pairs_u = np.zeros(npairs.value, dtype=np.int32)
        self.pairs_v = np.zeros(npairs.value, dtype=np.int32)
        _lib.check(_lib.lib.misa_tree_pairs(self.handle, _lib.ptr(self.pairs_u), _lib.ptr(self.pairs_v)))

    def marginals(self, variables=True, pairs=True, stream=0):
        """(log Z, P(x_i = 1) per variable, P(x_u, x_v) per merged coupler as
        (low, low), (1, low), (low, 1), (1, 1))."""
        lz = ctypes.c_double()
        vp = np.zeros(self.n, dtype=np.float64) if variables else None
        pp = np.zeros((len(self.pairs_u), 4), dtype=np.float64) if pairs else None
        _lib.check(_lib.lib.misa_tree_marginals(self.handle, ctypes.byref(lz), _lib.ptr(vp) if vp is not None else None,
                                                _lib.ptr(pp) if pp is not None and pp.size else None,
                                                ctypes.c_void_p(stream or None)))
        return float(lz.value), vp, pp

    def last_stats(self):
        """Device time and work of the last marginals or samples call."""
        st = _lib.TreeStats()
        _lib.check(_lib.lib.misa_tree_last_stats(self.handle, ctypes.byref(st)))
        return st

    def recompute(self):
        """Evaluate the tables again on the next call."""
        _lib.check(_lib.lib.misa_tree_recompute(self.handle))

    def solve(self, max_solutions):
        """The min(max_solutions, 2^n) lowest-energy states, ascending by
        energy (without the problem's offset): (samples int8, energies)."""
        k = min(int(max_solutions), 2 ** min(self.n, 62))
        samples = np.zeros((max(k, 0), self.n), dtype=np.int8)
        energies = np.zeros(max(k, 0), dtype=np.float64)
        found = ctypes.c_int()
        _lib.check(_lib.lib.misa_tree_solve(self.handle, int(max_solutions), _lib.ptr(samples), _lib.ptr(energies),
                                            ctypes.byref(found)))
        return samples[:found.value], energies[:found.value]

    def sample(self, num_samples, read_seeds, stream=0):
        seeds = np.ascontiguousarray(read_seeds, dtype=np.uint64)
        out = np.zeros((num_samples, self.n), dtype=np.int8)
        _lib.check(_lib.lib.misa_tree_sample(self.handle, int(num_samples), _lib.ptr(seeds), _lib.ptr(out),
                                             ctypes.c_void_p(stream or None)))
        return out

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib.misa_tree_plan_destroy(self.handle)
            self.handle = None
        if getattr(self, "problem", None) is not None:
            self.problem.close()
            self.problem = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TreeDecompositionSampler(dimodlite.Sampler):
    """Tree decomposition-based sampler for binary quadratic models, on MI355X:
    exact samples of the Boltzmann distribution and its marginals."""

    parameters = {"num_reads": [], "elimination_order": ["max_treewidth"], "beta": [], "marginals": [],
                  "seed": []}
    properties = {"max_treewidth": 25}

    def __init__(self):
        self.parameters = dict(self.parameters)
        self.properties = dict(self.properties)

    def sample(self, bqm, num_reads: Optional[int] = 1, elimination_order: Optional[List[Any]] = None,
               beta: Optional[float] = 1.0, marginals: Optional[bool] = True, seed: Optional[int] = None):
        """Draw ``num_reads`` exact samples of P(x) ~ exp(-beta E(x)) and, with
        ``marginals``, the variable and interaction marginals
        (samplers.py:235-359).  ``info['log_partition_function']`` includes
        the energy offset."""
        vartype = dimodlite.as_vartype(bqm.vartype)
        low = -1 if vartype is dimodlite.SPIN else 0
        variables = list(bqm.variables)
        if not len(variables):
            info = {"log_partition_function": 0.0}
            if marginals:
                info["variable_marginals"] = {}
                info["interaction_marginals"] = {}
            samples = np.empty((num_reads, 0), dtype=np.int8)
            return dimodlite.SampleSet.from_samples((samples, []), vartype, energy=np.zeros(num_reads) + bqm.offset,
                                                    info=info)
        index = {v: i for i, v in enumerate(variables)}
        order = None
        if elimination_order is not None:
            if len(elimination_order) != len(variables) or set(elimination_order) != set(variables):
                raise ValueError("elimination_order must contain every variable of the bqm exactly once")
            order = [index[v] for v in elimination_order]
        ldata, (irow, icol, qdata), off = bqm.to_numpy_vectors(variable_order=variables)
        try:
            plan = TreePlan(ldata, irow, icol, qdata, beta, low, order, self.properties["max_treewidth"])
        except _lib.EngineError as e:
            if "treewidth" in str(e):
                raise ValueError(f"maximum treewidth of {self.properties['max_treewidth']} exceeded.") from e
            raise
        try:
            if seed is None or seed < 0:
                seed = int.from_bytes(os.urandom(8), "little")
            samples = plan.sample(int(num_reads), _lib.read_seeds(int(seed), int(num_reads)))
            lz, vp, pp = plan.marginals(variables=bool(marginals), pairs=bool(marginals))
            info = {"log_partition_function": lz - beta * off}
            if marginals:
                info["variable_marginals"] = {variables[plan.order[q]]: float(vp[plan.order[q]])
                                              for q in range(len(variables))}
                # as sample.hpp:113-144,205-225 and samplers.py:343-351 report them: pairs with
                # a nonzero bias, keyed (i, j) with i < j in variable order, in that order
                configs = (low, low), (1, low), (low, 1), (1, 1)
                weight = {(min(a, b), max(a, b)): 0.0 for a, b in zip(irow, icol)}
                for a, b, w in zip(irow, icol, qdata):
                    weight[(min(a, b), max(a, b))] += w
                inter = []
                for u, v, p in zip(plan.pairs_u, plan.pairs_v, pp):
                    i, j = min(u, v), max(u, v)
                    if weight[(i, j)] != 0.0:
                        inter.append((i, j, p if u == i else p[[0, 2, 1, 3]]))
                inter.sort(key=lambda t: (t[0], t[1]))
                info["interaction_marginals"] = {(variables[i], variables[j]): dict(zip(configs, map(float, p)))
                                                 for i, j, p in inter}
        finally:
            plan.close()
        energies = bqm.energies((samples, variables))
        return dimodlite.SampleSet.from_samples((samples, variables), vartype, energy=energies, info=info)


class TreeDecompositionSolver(dimodlite.Sampler):
    """Tree decomposition-based solver for binary quadratic models, on MI355X:
    the ``num_reads`` lowest-energy states in order of energy
    (dwave/samplers/tree/samplers.py:29-183).  The k-best bucket elimination
    runs in HIP kernels (``misa_tree_solve``); states of equal energy may come
    in another order than the reference's."""

    parameters = {"num_reads": [], "elimination_order": ["max_treewidth"]}
    properties = {"max_treewidth": 25}

    def __init__(self):
        self.parameters = dict(self.parameters)
        self.properties = dict(self.properties)

    def sample(self, bqm, num_reads: Optional[int] = 1, elimination_order: Optional[List[Any]] = None):
        """Find ground states of a binary quadratic model (samplers.py:113-183):
        the min(num_reads, 2^n) lowest-energy states; for more reads the
        states repeat, as ``num_occurrences``."""
        vartype = dimodlite.as_vartype(bqm.vartype)
        variables = list(bqm.variables)
        if not len(variables):
            samples = np.empty((num_reads, 0), dtype=np.int8)
            energies = bqm.energies((samples, []))
            return dimodlite.SampleSet.from_samples((samples, []), vartype, energy=energies)
        low = -1 if vartype is dimodlite.SPIN else 0
        max_samples = min(int(num_reads), 2 ** min(len(variables), 62))
        index = {v: i for i, v in enumerate(variables)}
        order = None
        if elimination_order is not None:
            if len(elimination_order) != len(variables) or set(elimination_order) != set(variables):
                raise ValueError("elimination_order must contain every variable of the bqm exactly once")
            order = [index[v] for v in elimination_order]
        ldata, (irow, icol, qdata), off = bqm.to_numpy_vectors(variable_order=variables)
        try:
            plan = TreePlan(ldata, irow, icol, qdata, 1.0, low, order, self.properties["max_treewidth"])
        except _lib.EngineError as e:
            if "treewidth" in str(e):
                raise ValueError(f"maximum treewidth of {self.properties['max_treewidth']} exceeded.") from e
            raise
        try:
            samples, energies = plan.solve(max_samples)
        finally:
            plan.close()
        energies = energies + off
        num_occurrences = np.ones(len(energies), dtype=np.intc)
        if num_reads > len(energies) and len(energies):
            q, r = divmod(int(num_reads), len(energies))
            num_occurrences *= q
            num_occurrences[:r] += 1
        return dimodlite.SampleSet.from_samples((samples, variables), vartype, energy=energies,
                                                num_occurrences=num_occurrences)
