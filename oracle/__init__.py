"""ORACLE — test infrastructure only.

ctypes bindings for the CPU checker libraries built by ``oracle/Makefile``:

* ``liboracle.so`` — the plain-C restatement of the reference simulated
  annealing core (``dwave/samplers/sa/src/cpu_sa.cpp``), see ``sa_oracle.c``.
* ``_ref/libref_sa.so`` — the reference core itself, compiled from its own
  source file with a thin driver (``ref_driver.cpp``); present wherever it was
  built in a container holding the reference tree.
* the discrete-slice path-integral restatement (``pimc_oracle.c``, in
  ``liboracle.so``) and ``_ref/libref_pimc.so``, the reference's continuous-time
  core (``dwave/samplers/sqa/src/localPIMC.cpp``) behind ``ref_pimc_driver.cpp``.

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
_REF_PIMC_SO = os.path.join(_HERE, "_ref", "libref_pimc.so")
_REF_GD_SO = os.path.join(_HERE, "_ref", "libref_greedy.so")
_REF_ROTOR_SO = os.path.join(_HERE, "_ref", "libref_rotor.so")
_c_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")

_c_i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")
_c_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_c_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_c_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_c_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")

GIBBS, METROPOLIS = 0, 1
SEQUENTIAL, RANDOM = 0, 1

_lib = None
_ref = None
_ref_pimc = None
_ref_gd = None
_ref_rotor = None


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
        lib.orc_state_energy.restype = ctypes.c_double
        lib.orc_state_energy.argtypes = [ctypes.c_int, _c_i8p, _c_f64p, ctypes.c_int64,
                                         _c_i32p, _c_i32p, _c_f64p]
        lib.orc_pimc_sample.restype = ctypes.c_int
        lib.orc_pimc_sample.argtypes = [
            ctypes.c_int, _c_f64p, ctypes.c_int64, _c_i32p, _c_i32p, _c_f64p,
            ctypes.c_int, ctypes.c_int, _c_f64p, _c_f64p, ctypes.c_int, ctypes.c_int,
            _c_u64p, _c_u32p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_int]
        lib.orc_pimc_sample_chains.restype = ctypes.c_int
        lib.orc_pimc_sample_chains.argtypes = [
            ctypes.c_int, _c_f64p, ctypes.c_int64, _c_i32p, _c_i32p, _c_f64p,
            ctypes.c_int, ctypes.c_int, _c_f64p, _c_f64p, ctypes.c_int, ctypes.c_int,
            _c_u64p, _c_u32p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
            ctypes.c_double, ctypes.c_void_p, ctypes.c_int]
        lib.orc_pimc_colouring.restype = ctypes.c_int
        lib.orc_pimc_colouring.argtypes = [ctypes.c_int, ctypes.c_int64, _c_i32p, _c_i32p, _c_i32p]
        lib.orc_pimc_step_consts.restype = None
        lib.orc_pimc_step_consts.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(ctypes.c_float),
                                             ctypes.POINTER(ctypes.c_int)]
        lib.orc_descent.restype = ctypes.c_int
        lib.orc_descent.argtypes = [ctypes.c_int, _c_f64p, ctypes.c_int64, _c_i32p, _c_i32p, _c_f64p,
                                    ctypes.c_int, _c_i8p, _c_f64p, _c_u32p, ctypes.c_uint32, ctypes.c_int]
        lib.orc_rotor_tables.restype = None
        lib.orc_rotor_tables.argtypes = [_c_f64p, _c_f64p]
        lib.orc_rotor_sample.restype = ctypes.c_int
        lib.orc_rotor_sample.argtypes = [ctypes.c_int, _c_f64p, ctypes.c_int64, _c_i32p, _c_i32p, _c_f64p,
                                         _c_f64p, ctypes.c_int, ctypes.c_int, _c_f64p, _c_f64p, ctypes.c_int,
                                         _c_u64p, ctypes.c_int, ctypes.c_int, _c_u8p, _c_f64p, ctypes.c_void_p,
                                         ctypes.c_int, ctypes.c_int]
        lib.orc_pimc_exp2f.restype = ctypes.c_float
        lib.orc_pimc_exp2f.argtypes = [ctypes.c_float]
        lib.orc_pimc_log2u.restype = ctypes.c_float
        lib.orc_pimc_log2u.argtypes = [ctypes.c_uint32]
        lib.orc_pimc_xoshiro.restype = ctypes.c_uint32
        lib.orc_pimc_xoshiro.argtypes = [_c_u32p, ctypes.c_int, _c_u32p]
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


def state_energy(state, h, cs, ce, cw) -> float:
    """Energy in the reference's summation order (cpu_sa.cpp:210-227,
    localPIMC.cpp:56-73)."""
    h, cs, ce, cw, _ = _prep(h, cs, ce, cw, [])
    s = np.ascontiguousarray(state, dtype=np.int8)
    return float(_load().orc_state_energy(len(h), s, h, len(cs), cs, ce, cw))


# ---------------------------------------------------------------------------
# Discrete-slice path-integral Monte Carlo
# ---------------------------------------------------------------------------

def pimc_sample(h, cs, ce, cw, sweeps_per_beta, hp, hd, seeds, worldlines,
                num_slices=32, interval=0, nthreads=1, return_flips=False, colour=None):
    """Oracle path-integral anneal.  ``worldlines`` (R, n) uint32, bit k = slice
    k spin +1, is copied.  Returns ``(worldlines, statistics)`` (plus per-read
    changed-slice counts with ``return_flips``); statistics is
    (R, 1 + (len(hp)-1)//interval, n) int8 when ``interval`` > 0, else None.
    ``colour`` (class per variable, any proper colouring) sets the update
    order; None uses the label-order greedy colouring."""
    h, cs, ce, cw, hp = _prep(h, cs, ce, cw, hp)
    hd = np.ascontiguousarray(hd, dtype=np.float64)
    W = np.array(worldlines, dtype=np.uint32, order="C", copy=True)
    R, n = W.shape[0], len(h)
    if W.ndim != 2 or W.shape[1] != n:
        raise ValueError("worldlines must be (num_reads, num_vars)")
    if len(hd) < len(hp):
        raise ValueError("hd shorter than hp")
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    stats = None
    sp = None
    if interval > 0:
        ncol = 1 + (len(hp) - 1) // interval if len(hp) else 1
        stats = np.zeros((R, ncol, n), dtype=np.int8)
        sp = stats.ctypes.data
    flips = np.zeros(R, dtype=np.uint64)
    col = None
    if colour is not None:
        colour = np.ascontiguousarray(colour, dtype=np.int32)
        if colour.shape != (n,):
            raise ValueError("colour must have one class per variable")
        col = colour.ctypes.data
    rc = _load().orc_pimc_sample(n, h, len(cs), cs, ce, cw, R, int(sweeps_per_beta), hp, hd,
                                 len(hp), int(num_slices), seeds, W, sp, int(interval),
                                 flips.ctypes.data, col, int(nthreads))
    if rc == -2:
        raise ValueError("coupler indexes contain an invalid variable")
    if rc:
        raise ValueError("bad arguments")
    if return_flips:
        return W, stats, flips
    return W, stats


def pimc_sample_chains(h, cs, ce, cw, sweeps_per_beta, hp, hd, seeds, worldlines, qubits_per_chain,
                       chain_coupler_strength, num_slices=32, interval=0, nthreads=1, return_flips=False,
                       chain_colour=None):
    """Oracle path-integral anneal with chain updates: chains of
    ``qubits_per_chain`` consecutive qubits joined by the implicit coupling
    ``chain_coupler_strength`` (localPIMC.cpp:385-558).  Arguments and returns
    as ``pimc_sample``; ``chain_colour`` (class per chain) sets the update
    order, None the label-order greedy colouring of the chain graph."""
    h, cs, ce, cw, hp = _prep(h, cs, ce, cw, hp)
    hd = np.ascontiguousarray(hd, dtype=np.float64)
    W = np.array(worldlines, dtype=np.uint32, order="C", copy=True)
    R, n = W.shape[0], len(h)
    if W.ndim != 2 or W.shape[1] != n:
        raise ValueError("worldlines must be (num_reads, num_vars)")
    if len(hd) < len(hp):
        raise ValueError("hd shorter than hp")
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    stats = None
    sp = None
    if interval > 0:
        ncol = 1 + (len(hp) - 1) // interval if len(hp) else 1
        stats = np.zeros((R, ncol, n), dtype=np.int8)
        sp = stats.ctypes.data
    flips = np.zeros(R, dtype=np.uint64)
    col = None
    if chain_colour is not None:
        chain_colour = np.ascontiguousarray(chain_colour, dtype=np.int32)
        if chain_colour.shape != (n // int(qubits_per_chain),):
            raise ValueError("chain_colour must have one class per chain")
        col = chain_colour.ctypes.data
    rc = _load().orc_pimc_sample_chains(n, h, len(cs), cs, ce, cw, R, int(sweeps_per_beta), hp, hd,
                                        len(hp), int(num_slices), seeds, W, sp, int(interval),
                                        flips.ctypes.data, int(qubits_per_chain),
                                        float(chain_coupler_strength), col, int(nthreads))
    if rc == -2:
        raise ValueError("coupler indexes contain an invalid variable")
    if rc:
        raise ValueError("bad arguments")
    if return_flips:
        return W, stats, flips
    return W, stats


def pimc_colouring(n, cs, ce):
    cs = np.ascontiguousarray(cs, dtype=np.int32)
    ce = np.ascontiguousarray(ce, dtype=np.int32)
    out = np.zeros(max(n, 1), dtype=np.int32)
    nc = _load().orc_pimc_colouring(n, len(cs), cs, ce, out)
    if nc < 0:
        raise ValueError("coupler indexes contain an invalid variable")
    return nc, out[:n]


def pimc_step_consts(hp, hd, P):
    bs, gi, md = ctypes.c_double(), ctypes.c_float(), ctypes.c_int()
    _load().orc_pimc_step_consts(float(hp), float(hd), int(P), ctypes.byref(bs),
                                 ctypes.byref(gi), ctypes.byref(md))
    return bs.value, gi.value, md.value


def pimc_exp2f(y: float) -> float:
    return float(_load().orc_pimc_exp2f(float(y)))


def pimc_log2u(v: int) -> float:
    return float(_load().orc_pimc_log2u(int(v)))


def pimc_xoshiro(state, count: int) -> np.ndarray:
    st = np.array(state, dtype=np.uint32)
    out = np.zeros(count, dtype=np.uint32)
    _load().orc_pimc_xoshiro(st, int(count), out)
    return out


def have_ref_pimc() -> bool:
    return os.path.exists(_REF_PIMC_SO)


def _load_ref_pimc():
    global _ref_pimc
    if _ref_pimc is None:
        lib = ctypes.CDLL(_REF_PIMC_SO)
        lib.ref_pimc_call.restype = ctypes.c_int
        lib.ref_pimc_call.argtypes = [
            _c_i8p, _c_f64p, ctypes.c_int, ctypes.c_int, _c_f64p, ctypes.c_int,
            _c_i32p, _c_i32p, _c_f64p, ctypes.c_int, _c_f64p, _c_f64p, ctypes.c_int,
            _c_u32p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        _ref_pimc = lib
    return _ref_pimc


def ref_pimc_call(h, cs, ce, cw, sweeps_per_beta, hp, hd, states, seeds,
                  interval=0, nthreads=1):
    """The compiled reference continuous-time core with classical states in and
    out; reads are split into ``len(seeds)`` chunks, one call (and one 32-bit
    seed) each.  Returns ``(states, energies, statistics or None)``."""
    h, cs, ce, cw, hp = _prep(h, cs, ce, cw, hp)
    hd = np.ascontiguousarray(hd, dtype=np.float64)
    st = np.array(states, dtype=np.int8, order="C", copy=True)
    R, n = st.shape
    en = np.zeros(R, dtype=np.float64)
    seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
    stats = None
    sp = None
    if interval > 0:
        stats = np.zeros((R, 1 + (len(hp) - 1) // interval, n), dtype=np.int8)
        sp = stats.ctypes.data
    rc = _load_ref_pimc().ref_pimc_call(st, en, R, n, h, len(cs), cs, ce, cw,
                                        int(sweeps_per_beta), hp, hd, len(hp), seeds,
                                        len(seeds), sp, int(interval), int(nthreads))
    if rc < 0:
        raise RuntimeError("reference core raised")
    return st, en, stats


# ---------------------------------------------------------------------------
# Steepest descent
# ---------------------------------------------------------------------------

def descent(h, cs, ce, cw, states, nthreads=1, max_steps=0xFFFFFFFF):
    """Oracle steepest descent (dense solver; self-loops are constants).
    ``states`` (R, n) int8 is copied.  Returns ``(states, energies, num_steps)``."""
    h, cs, ce, cw, _ = _prep(h, cs, ce, cw, [])
    st = np.array(states, dtype=np.int8, order="C", copy=True)
    R = st.shape[0]
    en = np.zeros(R, dtype=np.float64)
    steps = np.zeros(R, dtype=np.uint32)
    rc = _load().orc_descent(len(h), h, len(cs), cs, ce, cw, R, st, en, steps, int(max_steps), int(nthreads))
    if rc:
        raise ValueError("coupler indexes contain an invalid variable")
    return st, en, steps


def have_ref_greedy() -> bool:
    return os.path.exists(_REF_GD_SO)


def ref_descent(h, cs, ce, cw, states, large_sparse_opt=False, nthreads=1):
    """The compiled reference steepest_gradient_descent (descent.cpp)."""
    global _ref_gd
    if _ref_gd is None:
        lib = ctypes.CDLL(_REF_GD_SO)
        lib.ref_descent.restype = ctypes.c_int
        lib.ref_descent.argtypes = [_c_i8p, _c_f64p, _c_u32p, ctypes.c_int, ctypes.c_int, _c_f64p,
                                    ctypes.c_int, _c_i32p, _c_i32p, _c_f64p, ctypes.c_int, ctypes.c_int]
        _ref_gd = lib
    h, cs, ce, cw, _ = _prep(h, cs, ce, cw, [])
    st = np.array(states, dtype=np.int8, order="C", copy=True)
    R = st.shape[0]
    en = np.zeros(R, dtype=np.float64)
    steps = np.zeros(R, dtype=np.uint32)
    rc = _ref_gd.ref_descent(st, en, steps, R, len(h), h, len(cs), cs, ce, cw,
                             1 if large_sparse_opt else 0, int(nthreads))
    if rc:
        raise RuntimeError("reference core raised")
    return st, en, steps


# ---------------------------------------------------------------------------
# Rotor-model annealing
# ---------------------------------------------------------------------------

ROTOR_PROPOSALS = {"GibbsNonErgodic": 0, "MetropolisNonErgodic": 1, "MetropolisUniform": 2, "MetropolisTF": 3}


def rotor_tables():
    """(cos, |sin|) of the 256 angles 2 pi i / 256, as the reference tabulates them."""
    c = np.zeros(256, dtype=np.float64)
    sn = np.zeros(256, dtype=np.float64)
    _load().orc_rotor_tables(c, sn)
    return c, sn


def _rotor_args(h, cs, ce, cw, tf, Hp, Hd, seeds, states, interval):
    h, cs, ce, cw, _ = _prep(h, cs, ce, cw, [])
    tf = np.ascontiguousarray(tf, dtype=np.float64)
    Hp = np.ascontiguousarray(Hp, dtype=np.float64)
    Hd = np.ascontiguousarray(Hd, dtype=np.float64)
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    st = np.array(states, dtype=np.uint8, order="C", copy=True)
    R, n = st.shape[0], len(h)
    nf = len(Hp)
    cnt = (1 + (nf - 1) // interval) if (interval and nf) else 0
    stats = np.zeros((R, cnt, n), dtype=np.uint8) if cnt else None
    return h, cs, ce, cw, tf, Hp, Hd, seeds, st, stats


def rotor_sample(h, cs, ce, cw, tf, sweeps_per_beta, Hp, Hd, seeds, states, randomize_order=False,
                 proposal="MetropolisUniform", interval=0, nthreads=1):
    """Oracle rotor annealing, one read per seed.  ``states`` (R, n) uint8
    angles is copied.  Returns ``(states, energies, statistics)`` (statistics
    (R, 1 + (len(Hp) - 1) // interval, n) uint8, or None)."""
    h, cs, ce, cw, tf, Hp, Hd, seeds, st, stats = _rotor_args(h, cs, ce, cw, tf, Hp, Hd, seeds, states, interval)
    R = st.shape[0]
    en = np.zeros(R, dtype=np.float64)
    rc = _load().orc_rotor_sample(len(h), h, len(cs), cs, ce, cw, tf, R, int(sweeps_per_beta), Hp, Hd, len(Hp),
                                  seeds, 1 if randomize_order else 0, ROTOR_PROPOSALS[proposal], st, en,
                                  stats.ctypes.data if stats is not None else None, int(interval), int(nthreads))
    if rc:
        raise ValueError("coupler indexes contain an invalid variable")
    return st, en, stats


def have_ref_rotor() -> bool:
    return os.path.exists(_REF_ROTOR_SO)


def ref_rotor(h, cs, ce, cw, tf, sweeps_per_beta, Hp, Hd, seeds, states, randomize_order=False,
              proposal="MetropolisUniform", interval=0, nthreads=1):
    """The compiled reference rotor core (cpu_rotormc.cpp), one call per read
    with num_samples = 1 and that read's seed.  Same returns as rotor_sample."""
    global _ref_rotor
    if _ref_rotor is None:
        lib = ctypes.CDLL(_REF_ROTOR_SO)
        lib.ref_rotor_call.restype = ctypes.c_int
        lib.ref_rotor_call.argtypes = [_c_u8p, _c_f64p, ctypes.c_int, ctypes.c_int, _c_f64p, ctypes.c_int,
                                       _c_i32p, _c_i32p, _c_f64p, _c_f64p, ctypes.c_int, _c_f64p, _c_f64p,
                                       ctypes.c_int, _c_u64p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_int, ctypes.c_int]
        _ref_rotor = lib
    h, cs, ce, cw, tf, Hp, Hd, seeds, st, stats = _rotor_args(h, cs, ce, cw, tf, Hp, Hd, seeds, states, interval)
    R = st.shape[0]
    en = np.zeros(R, dtype=np.float64)
    rc = _ref_rotor.ref_rotor_call(st, en, R, len(h), h, len(cs), cs, ce, cw, tf, int(sweeps_per_beta), Hp, Hd,
                                   len(Hp), seeds, 1 if randomize_order else 0, ROTOR_PROPOSALS[proposal],
                                   stats.ctypes.data if stats is not None else None, int(interval), int(nthreads))
    if rc:
        raise RuntimeError("reference core raised")
    return st, en, stats
