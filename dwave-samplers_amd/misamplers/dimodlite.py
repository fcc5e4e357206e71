"""Small dimod-compatible data types used when dimod is not importable.

The reference samplers are dimod plugins (``dimod.Sampler``,
``dimod.Initialized``, ``dimod.BinaryQuadraticModel``, ``dimod.SampleSet``).
dimod is a third-party dependency (dimod==0.12.21, pyproject.toml:5,39) and is
absent here, so this module provides the subset of its behaviour the sampling
path touches: a BQM with SPIN/BINARY vartypes and ``to_numpy_vectors``, a
SampleSet with ``record``/``first``/``info``/``change_vartype``, the
``sample_ising``/``sample_qubo`` mixin and ``parse_initial_states`` with the
'none' / 'tile' / 'random' generators.  When dimod is importable the samplers
accept its objects too (duck typing on ``vartype``, ``linear``, ``quadratic``,
``offset``, ``variables``).

Behaviour of dimod that no reference test pins (the random initial-state
generator's stream and the coupler order of ``to_numpy_vectors``) is this
module's own: random states come from ``numpy.random.default_rng(seed)`` and
couplers keep insertion order.
"""
from __future__ import annotations

import enum
import numbers
from collections import namedtuple
from typing import Any, Dict, Iterable, List, Sequence, Tuple

import numpy as np


class Vartype(enum.Enum):
    SPIN = frozenset({-1, 1})
    BINARY = frozenset({0, 1})


SPIN = Vartype.SPIN
BINARY = Vartype.BINARY


def as_vartype(vartype) -> Vartype:
    if isinstance(vartype, Vartype):
        return vartype
    name = getattr(vartype, "name", vartype)
    if name in ("SPIN", "spin") or vartype == frozenset({-1, 1}):
        return SPIN
    if name in ("BINARY", "binary") or vartype == frozenset({0, 1}):
        return BINARY
    raise TypeError(f"unknown vartype {vartype!r}")


def _is_spin(vartype) -> bool:
    return as_vartype(vartype) is SPIN


class BinaryQuadraticModel:
    """Ising (SPIN) or QUBO (BINARY) model: E(s) = sum h_i s_i + sum J_ij s_i s_j + offset."""

    def __init__(self, linear=None, quadratic=None, offset=0.0, vartype=SPIN):
        if isinstance(linear, (str, Vartype)) and quadratic is None:
            vartype, linear = linear, None
        self.vartype = as_vartype(vartype)
        self.linear: Dict[Any, float] = {}
        self.quadratic: Dict[Tuple[Any, Any], float] = {}
        self.offset = float(offset)
        for v, b in (linear or {}).items():
            self.add_variable(v, b)
        for (u, v), b in (quadratic or {}).items():
            self.add_interaction(u, v, b)

    # -- construction -------------------------------------------------------
    @classmethod
    def empty(cls, vartype):
        return cls({}, {}, 0.0, vartype)

    @classmethod
    def from_ising(cls, h, J, offset=0.0):
        if not isinstance(h, dict):
            h = dict(enumerate(h))
        return cls(h, J, offset, SPIN)

    @classmethod
    def from_qubo(cls, Q, offset=0.0):
        bqm = cls({}, {}, offset, BINARY)
        for (u, v), b in Q.items():
            if u == v:
                bqm.add_variable(u, b)
            else:
                bqm.add_interaction(u, v, b)
        return bqm

    def add_variable(self, v, bias=0.0):
        self.linear[v] = self.linear.get(v, 0.0) + float(bias)

    def add_interaction(self, u, v, bias):
        if u == v:
            raise ValueError(f"{u!r} cannot have an interaction with itself")
        self.add_variable(u, 0.0)
        self.add_variable(v, 0.0)
        key = (u, v) if (u, v) in self.quadratic or (v, u) not in self.quadratic else (v, u)
        self.quadratic[key] = self.quadratic.get(key, 0.0) + float(bias)

    # -- views --------------------------------------------------------------
    @property
    def variables(self) -> List[Any]:
        return list(self.linear)

    @property
    def num_variables(self) -> int:
        return len(self.linear)

    def __len__(self):
        return len(self.linear)

    @property
    def spin(self):
        return self if self.vartype is SPIN else self.change_vartype(SPIN, inplace=False)

    @property
    def binary(self):
        return self if self.vartype is BINARY else self.change_vartype(BINARY, inplace=False)

    def copy(self):
        return BinaryQuadraticModel(dict(self.linear), dict(self.quadratic), self.offset, self.vartype)

    def change_vartype(self, vartype, inplace=True):
        vartype = as_vartype(vartype)
        if vartype is self.vartype:
            return self if inplace else self.copy()
        lin = {v: 0.0 for v in self.linear}
        quad: Dict[Tuple[Any, Any], float] = {}
        off = self.offset
        if vartype is SPIN:   # x = (s + 1) / 2
            for v, b in self.linear.items():
                lin[v] += b / 2.0
                off += b / 2.0
            for (u, v), b in self.quadratic.items():
                quad[(u, v)] = b / 4.0
                lin[u] += b / 4.0
                lin[v] += b / 4.0
                off += b / 4.0
        else:                  # s = 2x - 1
            for v, b in self.linear.items():
                lin[v] += 2.0 * b
                off -= b
            for (u, v), b in self.quadratic.items():
                quad[(u, v)] = 4.0 * b
                lin[u] -= 2.0 * b
                lin[v] -= 2.0 * b
                off += b
        new = BinaryQuadraticModel(lin, quad, off, vartype)
        if inplace:
            self.linear, self.quadratic, self.offset, self.vartype = new.linear, new.quadratic, new.offset, vartype
            return self
        return new

    def to_numpy_vectors(self, variable_order=None):
        order = list(self.linear) if variable_order is None else list(variable_order)
        index = {v: i for i, v in enumerate(order)}
        ldata = np.array([self.linear[v] for v in order], dtype=np.float64)
        irow = np.array([index[u] for (u, v) in self.quadratic], dtype=np.int32)
        icol = np.array([index[v] for (u, v) in self.quadratic], dtype=np.int32)
        qdata = np.array(list(self.quadratic.values()), dtype=np.float64)
        return ldata, (irow, icol, qdata), self.offset

    def energies(self, samples_like, dtype=np.float64):
        samples, labels = as_samples(samples_like)
        index = {v: i for i, v in enumerate(labels)}
        s = samples.astype(np.float64)
        e = np.full(s.shape[0], self.offset, dtype=np.float64)
        for v, b in self.linear.items():
            e += b * s[:, index[v]]
        for (u, v), b in self.quadratic.items():
            e += b * s[:, index[u]] * s[:, index[v]]
        return e

    def energy(self, sample):
        return float(self.energies(sample)[0])


def as_samples(samples_like) -> Tuple[np.ndarray, List[Any]]:
    """(samples 2-D int8 array, labels) from the dimod 'samples_like' forms."""
    if isinstance(samples_like, SampleSet):
        return samples_like.record.sample, list(samples_like.variables)
    if hasattr(samples_like, "record") and hasattr(samples_like, "variables"):
        return np.asarray(samples_like.record.sample), list(samples_like.variables)
    if isinstance(samples_like, tuple) and len(samples_like) == 2 and not isinstance(samples_like[0], numbers.Number):
        arr, labels = samples_like
        if isinstance(arr, dict):
            return as_samples(arr)
        arr = np.atleast_2d(np.asarray(arr))
        if arr.size == 0 and (arr.ndim != 2 or arr.shape[1] != len(labels)):
            arr = arr.reshape(0, len(labels))
        return arr.astype(np.int8), list(labels)
    if isinstance(samples_like, dict):
        labels = list(samples_like)
        return np.array([[samples_like[v] for v in labels]], dtype=np.int8), labels
    if isinstance(samples_like, (list, tuple)) and samples_like and isinstance(samples_like[0], dict):
        labels = list(samples_like[0])
        return np.array([[d[v] for v in labels] for d in samples_like], dtype=np.int8), labels
    arr = np.atleast_2d(np.asarray(samples_like))
    return arr.astype(np.int8), list(range(arr.shape[1]))


_Rec = namedtuple("_Rec", ["sample", "energy", "num_occurrences"])


class SampleSet:
    """Samples (rows), their energies and an ``info`` dict."""

    def __init__(self, samples: np.ndarray, variables: Sequence[Any], energy, vartype, info=None, vectors=None):
        samples = np.asarray(samples)
        self._samples = samples.astype(np.int8, copy=False)
        self._variables = list(variables)
        energy = np.asarray(energy, dtype=np.float64).reshape(-1)
        self._energy = energy
        self.vartype = as_vartype(vartype)
        self.info = dict(info or {})
        self._vectors = {k: np.asarray(v) for k, v in (vectors or {}).items()}

    @classmethod
    def from_samples(cls, samples_like, vartype, energy, info=None, **vectors):
        """Extra keyword arrays (one value per sample) become data vectors
        (dimod's ``SampleSet.from_samples(**vectors)``)."""
        samples, labels = as_samples(samples_like)
        energy = np.broadcast_to(np.asarray(energy, dtype=np.float64), (samples.shape[0],)).copy()
        vectors = {k: np.broadcast_to(np.asarray(v), (samples.shape[0],)).copy() for k, v in vectors.items()}
        return cls(samples, labels, energy, vartype, info, vectors)

    @classmethod
    def from_samples_bqm(cls, samples_like, bqm, **kw):
        samples, labels = as_samples(samples_like)
        return cls.from_samples((samples, labels), bqm.vartype, bqm.energies((samples, labels)), **kw)

    def _occurrences(self):
        occ = self._vectors.get("num_occurrences")
        return np.ones(len(self._energy), dtype=np.int64) if occ is None else occ

    @property
    def record(self):
        extra = {k: v for k, v in self._vectors.items() if k != "num_occurrences"}
        if not extra:
            return _Rec(self._samples, self._energy, self._occurrences())
        rec = namedtuple("_Rec", ["sample", "energy", "num_occurrences"] + list(extra))
        return rec(self._samples, self._energy, self._occurrences(), *extra.values())

    @property
    def variables(self):
        return list(self._variables)

    @property
    def data_vectors(self):
        out = {"energy": self._energy, "num_occurrences": self._occurrences()}
        out.update(self._vectors)
        return out

    def __len__(self):
        return self._samples.shape[0]

    def samples(self, n=None, sorted_by="energy"):
        idx = np.argsort(self._energy, kind="stable") if sorted_by == "energy" else np.arange(len(self))
        if n is not None:
            idx = idx[:n]
        return [dict(zip(self._variables, (int(x) for x in self._samples[i]))) for i in idx]

    def data(self, fields=None, sorted_by="energy"):
        Datum = namedtuple("Datum", ["sample", "energy", "num_occurrences"])
        idx = np.argsort(self._energy, kind="stable") if sorted_by == "energy" else np.arange(len(self))
        for i in idx:
            yield Datum(dict(zip(self._variables, (int(x) for x in self._samples[i]))), float(self._energy[i]), 1)

    @property
    def first(self):
        if not len(self):
            raise ValueError("empty SampleSet")
        return next(self.data())

    def lowest(self):
        if not len(self):
            return self
        keep = self._energy == self._energy.min()
        return SampleSet(self._samples[keep], self._variables, self._energy[keep], self.vartype, self.info,
                         {k: v[keep] for k, v in self._vectors.items()})

    def change_vartype(self, vartype, energy_offset=0.0, inplace=True):
        vartype = as_vartype(vartype)
        new = self if inplace else SampleSet(self._samples.copy(), self._variables, self._energy.copy(),
                                              self.vartype, self.info, self._vectors)
        if vartype is not new.vartype:
            if vartype is BINARY:
                new._samples = ((new._samples + 1) // 2).astype(np.int8)
            else:
                new._samples = (2 * new._samples - 1).astype(np.int8)
            new.vartype = vartype
        if energy_offset:
            new._energy = new._energy + energy_offset
        return new

    def __repr__(self):
        return f"SampleSet(num_rows={len(self)}, variables={self._variables!r}, vartype={self.vartype.name})"


ParsedInputs = namedtuple("ParsedInputs", ["initial_states", "initial_states_generator", "num_reads", "seed"])


class _InitialStates:
    """What ``parse_initial_states`` returns as ``initial_states``."""

    def __init__(self, samples, variables):
        self.record = _Rec(samples, np.zeros(samples.shape[0]), np.ones(samples.shape[0], dtype=np.int64))
        self.variables = list(variables)


class Initialized:
    """``parse_initial_states`` with dimod's 'none' / 'tile' / 'random' generators."""

    def parse_initial_states(self, bqm, initial_states=None, initial_states_generator="random",
                             num_reads=None, seed=None):
        if initial_states_generator not in ("none", "tile", "random"):
            raise ValueError(f"unknown initial_states_generator {initial_states_generator!r}")
        variables = list(bqm.variables)
        n = len(variables)
        if initial_states is None:
            samples, labels = np.empty((0, n), dtype=np.int8), variables
        else:
            samples, labels = as_samples(initial_states)
            if samples.shape[0] == 0 and samples.ndim == 2 and samples.shape[1] != len(labels):
                samples = samples.reshape(0, len(labels))
        if set(labels) != set(variables) or len(labels) != n:
            if samples.shape[0] or set(labels) != set(variables):
                raise ValueError("mismatch between variables in 'initial_states' and 'bqm'")
        if labels != variables and n:
            col = [labels.index(v) for v in variables]
            samples = samples[:, col]
        samples = np.asarray(samples, dtype=np.int8)
        samples = samples.reshape(samples.shape[0] if samples.ndim == 2 else (-1 if n else 0), n)
        vt = as_vartype(bqm.vartype)
        if samples.size and not np.isin(samples, list(vt.value)).all():
            raise ValueError("expected initial_states to match the bqm vartype")
        if num_reads is None:
            num_reads = samples.shape[0] if samples.shape[0] else 1
        if isinstance(num_reads, bool) or not isinstance(num_reads, numbers.Integral):
            raise TypeError("'num_reads' should be a positive integer")
        if num_reads < 1:
            raise ValueError("'num_reads' should be a positive integer")
        num_reads = int(num_reads)
        k = samples.shape[0]
        if num_reads < k:
            samples = samples[:num_reads]
        elif num_reads > k:
            if initial_states_generator == "none":
                raise ValueError("insufficient number of initial states given")
            if initial_states_generator == "tile":
                if k == 0:
                    raise ValueError("cannot tile an empty sample set of initial states")
                samples = np.tile(samples, ((num_reads + k - 1) // k, 1))[:num_reads]
            else:
                rng = np.random.default_rng(seed)
                extra = rng.integers(0, 2, size=(num_reads - k, n), dtype=np.int8)
                if vt is SPIN:
                    extra = (2 * extra - 1).astype(np.int8)
                samples = np.concatenate([samples, extra], axis=0)
        # a fresh array: the engines anneal it in place, and the caller's
        # initial states (often reused across calls) must stay as given
        return ParsedInputs(_InitialStates(np.array(samples, dtype=np.int8, order="C", copy=True), variables),
                            initial_states_generator, num_reads, seed)


class Sampler:
    """``sample_ising`` / ``sample_qubo`` on top of ``sample(bqm, **kw)``."""

    def sample_ising(self, h, J, **parameters):
        return self.sample(BinaryQuadraticModel.from_ising(h, J), **parameters)

    def sample_qubo(self, Q, **parameters):
        return self.sample(BinaryQuadraticModel.from_qubo(Q), **parameters)

    def remove_unknown_kwargs(self, **kwargs):
        return {k: v for k, v in kwargs.items() if k in self.parameters}
