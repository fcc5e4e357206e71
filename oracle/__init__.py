"""ORACLE — test infrastructure only.

ctypes bindings for the CPU checker libraries built by ``oracle/Makefile``:

* ``liboracle.so`` — the plain-C restatement of the reference simulated
  annealing core (``dwave/samplers/sa/src/cpu_sa.cpp``), see ``sa_oracle.c``.
* ``_ref/libref_sa.so`` — the reference core itself, compiled from its own
  source file with a thin driver (``ref_driver.cpp``); present wherever it was
  built in a container holding the reference tree.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this package, and only as the checker / CPU baseline.  The
product (``misamplers``) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ORACLE_SO = os.path.join(_HERE, "liboracle.so")
_REF_SO = os.path.join(_HERE, "_ref", "libref_sa.so")

_c_i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")
_c_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_c_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_c_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")

GIBBS, METROPOLIS = 0, 1
SEQUENTIAL, RANDOM = 0, 1

_lib = None
_ref = None


def build():
    """Compile the oracle (and the reference core, when its source is present)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_ORACLE_SO):
            build()
        lib = ctypes.CDLL(_ORACLE_SO)
        lib.orc_sa_sample.restype = ctypes.c_int
        lib.orc_sa_sample.argtypes = [
            ctypes.c_int, _c_f64p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
            _c_i32p, _c_i32p, _c_f64p, ctypes.c_int, _c_f64p, ctypes.c_int,
            _c_u64p, ctypes.c_int64, _c_i8p, _c_f64p, ctypes.c_int, ctypes.c_int,
            ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        lib.orc_sa_sample_stream.restype = ctypes.c_int
        lib.orc_sa_sample_stream.argtypes = [
            ctypes.c_int, _c_f64p, ctypes.c_int64, _c_i32p, _c_i32p, _c_f64p,
            ctypes.c_int, _c_f64p, ctypes.c_int, ctypes.c_uint64, ctypes.c_int64,
            _c_i8p, _c_f64p, ctypes.c_int, ctypes.c_int]
        lib.orc_read_seeds.restype = None
        lib.orc_read_seeds.argtypes = [ctypes.c_uint64, ctypes.c_int64, _c_u64p]
        lib.orc_xorshift_next.restype = ctypes.c_uint64
        lib.orc_xorshift_next.argtypes = [_c_u64p]
        _lib = lib
    return _lib


def have_ref() -> bool:
    return os.path.exists(_REF_SO)


def _load_ref():
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(_REF_SO)
        lib.ref_sa_call.restype = ctypes.c_int
        lib.ref_sa_call.argtypes = [
            _c_i8p, _c_f64p, ctypes.c_int, ctypes.c_int, _c_f64p, ctypes.c_int,
            _c_i32p, _c_i32p, _c_f64p, ctypes.c_int, _c_f64p, ctypes.c_int,
            ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
        lib.ref_sa_read_by_read.restype = ctypes.c_int
        lib.ref_sa_read_by_read.argtypes = [
            _c_i8p, _c_f64p, ctypes.c_int, ctypes.c_int, _c_f64p, ctypes.c_int,
            _c_i32p, _c_i32p, _c_f64p, ctypes.c_int, _c_f64p, ctypes.c_int,
            _c_u64p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        _ref = lib
    return _ref


def _prep(h, cs, ce, cw, betas):
    return (np.ascontiguousarray(h, dtype=np.float64),
            np.ascontiguousarray(cs, dtype=np.int32),
            np.ascontiguousarray(ce, dtype=np.int32),
            np.ascontiguousarray(cw, dtype=np.float64),
            np.ascontiguousarray(betas, dtype=np.float64))


def read_seeds(seed: int, num_reads: int) -> np.ndarray:
    out = np.empty(num_reads, dtype=np.uint64)
    _load().orc_read_seeds(ctypes.c_uint64(seed), num_reads, out)
    return out


def xorshift_stream(s0: int, s1: int, count: int) -> list:
    st = np.array([s0, s1], dtype=np.uint64)
    return [int(_load().orc_xorshift_next(st)) for _ in range(count)]


def sa_sample(h, cs, ce, cw, sweeps_per_beta, betas, seeds, states,
              order=SEQUENTIAL, proposal=METROPOLIS, nthreads=1,
              return_counts=False):
    """Oracle read-by-read SA.  ``states`` (R, n) int8 is copied, not modified."""
    h, cs, ce, cw, betas = _prep(h, cs, ce, cw, betas)
    st = np.array(states, dtype=np.int8, order="C", copy=True)
    R = st.shape[0]
    n = len(h)
    if st.ndim != 2 or st.shape[1] != n:
        raise ValueError("states must be (num_reads, num_vars)")
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    en = np.zeros(R, dtype=np.float64)
    flips = np.zeros(R, dtype=np.uint64)
    degs = np.zeros(R, dtype=np.uint64)
    rc = _load().orc_sa_sample(
        n, h, len(cs), len(ce), len(cw), cs, ce, cw, int(sweeps_per_beta), betas,
        len(betas), seeds, R, st, en, int(order), int(proposal), int(nthreads),
        flips.ctypes.data, degs.ctypes.data)
    if rc == -1:
        raise ValueError("coupler vectors have mismatched lengths")
    if rc == -2:
        raise ValueError("coupler indexes contain an invalid variable")
    if return_counts:
        return st, en, flips, degs
    return st, en


def sa_sample_stream(h, cs, ce, cw, sweeps_per_beta, betas, seed, states,
                     order=SEQUENTIAL, proposal=METROPOLIS):
    h, cs, ce, cw, betas = _prep(h, cs, ce, cw, betas)
    st = np.array(states, dtype=np.int8, order="C", copy=True)
    en = np.zeros(st.shape[0], dtype=np.float64)
    rc = _load().orc_sa_sample_stream(len(h), h, len(cs), cs, ce, cw,
                                      int(sweeps_per_beta), betas, len(betas),
                                      ctypes.c_uint64(seed), st.shape[0], st, en,
                                      int(order), int(proposal))
    if rc:
        raise ValueError("coupler indexes contain an invalid variable")
    return st, en


def ref_sa_read_by_read(h, cs, ce, cw, sweeps_per_beta, betas, seeds, states,
                        order=SEQUENTIAL, proposal=METROPOLIS, nthreads=1):
    """The compiled reference core, read by read (num_samples=1, seeds[r])."""
    h, cs, ce, cw, betas = _prep(h, cs, ce, cw, betas)
    st = np.array(states, dtype=np.int8, order="C", copy=True)
    en = np.zeros(st.shape[0], dtype=np.float64)
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    rc = _load_ref().ref_sa_read_by_read(st, en, st.shape[0], len(h), h, len(cs),
                                         cs, ce, cw, int(sweeps_per_beta), betas,
                                         len(betas), seeds, int(order), int(proposal),
                                         int(nthreads))
    if rc < 0:
        raise RuntimeError("reference core raised")
    return st, en


def ref_sa_call(h, cs, ce, cw, sweeps_per_beta, betas, seed, states,
                order=SEQUENTIAL, proposal=METROPOLIS):
    """The compiled reference core, one call over all reads (one RNG stream)."""
    h, cs, ce, cw, betas = _prep(h, cs, ce, cw, betas)
    st = np.array(states, dtype=np.int8, order="C", copy=True)
    en = np.zeros(st.shape[0], dtype=np.float64)
    rc = _load_ref().ref_sa_call(st, en, st.shape[0], len(h), h, len(cs), cs, ce,
                                 cw, int(sweeps_per_beta), betas, len(betas),
                                 ctypes.c_uint64(seed), int(order), int(proposal))
    if rc < 0:
        raise RuntimeError("reference core raised")
    return st, en
