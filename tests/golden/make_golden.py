"""Write the SA golden vectors from the compiled reference core.

Runs in a container that holds the reference tree: ``make -C oracle`` compiles
dwave/samplers/sa/src/cpu_sa.cpp into oracle/_ref/libref_sa.so, and this
script calls it read by read (num_samples=1, per-read seed) on small seeded
problems covering the edge cases the reference tests exercise.  Each case is
saved as ``sa_<name>.npz`` (inputs plus expected states and energies); the
fixtures are data only and are checked by tests/test_oracle.py and
tests/test_sa_gpu.py.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def lattice3d(L):
    cs, ce = [], []
    for x in range(L):
        for y in range(L):
            for z in range(L):
                v = x * L * L + y * L + z
                for dx, dy, dz in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
                    u = ((x + dx) % L) * L * L + ((y + dy) % L) * L + (z + dz) % L
                    cs.append(v)
                    ce.append(u)
    return np.array(cs, np.int32), np.array(ce, np.int32)


def gnp(rng, n, p):
    pairs = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < p]
    cs = np.array([a for a, _ in pairs], np.int32)
    ce = np.array([b for _, b in pairs], np.int32)
    return cs, ce


def cases():
    rng = np.random.default_rng(20241203)
    out = {}
    # integer +-1 lattice (int8 mode), sequential Metropolis, geometric schedule
    cs, ce = lattice3d(4)
    out["lattice4_int_metro"] = dict(h=np.zeros(64), cs=cs, ce=ce, cw=rng.choice([-1.0, 1.0], len(cs)),
                                     betas=np.geomspace(0.1, 5.0, 60), spb=1, order=0, prop=1, R=24)
    # integer biases with fields needing int16 (|field| > 63)
    cs, ce = gnp(rng, 30, 0.5)
    out["gnp30_int16_metro"] = dict(h=rng.integers(-20, 21, 30).astype(float), cs=cs, ce=ce,
                                    cw=rng.integers(-9, 10, len(cs)).astype(float),
                                    betas=np.geomspace(0.01, 1.0, 40), spb=2, order=0, prop=1, R=16)
    # real-valued G(n, 0.5), Metropolis and Gibbs, sequential and random order
    cs, ce = gnp(rng, 40, 0.5)
    base = dict(h=rng.uniform(-1, 1, 40), cs=cs, ce=ce, cw=rng.uniform(-1, 1, len(cs)),
                betas=np.geomspace(0.2, 6.0, 50), spb=1, R=16)
    for order in (0, 1):
        for prop in (0, 1):
            out[f"gnp40_real_o{order}_p{prop}"] = dict(base, order=order, prop=prop)
    # integer Gibbs and random order on the lattice
    cs, ce = lattice3d(3)
    out["lattice3_int_gibbs_random"] = dict(h=rng.integers(-1, 2, 27).astype(float), cs=cs, ce=ce,
                                            cw=rng.choice([-1.0, 1.0], len(cs)),
                                            betas=np.linspace(0.1, 3.0, 30), spb=3, order=1, prop=0, R=12)
    # self-loops and repeated couplers (the core accepts them, test_simulated_annealing.py:98-100)
    n = 7
    cs = np.array([u for u in range(n) for v in range(u, n)], np.int32)
    ce = np.array([v for u in range(n) for v in range(u, n)], np.int32)
    cs = np.concatenate([cs, cs[:5]]).astype(np.int32)
    ce = np.concatenate([ce, ce[:5]]).astype(np.int32)
    for tag, cw in (("int", -np.ones(len(cs))), ("real", rng.uniform(-1, 1, len(cs)))):
        for prop in (0, 1):
            out[f"selfloops_{tag}_p{prop}"] = dict(h=-np.ones(n) if tag == "int" else rng.uniform(-1, 1, n),
                                                   cs=cs, ce=ce, cw=cw, betas=np.linspace(0.01, 3, 40), spb=1,
                                                   order=0, prop=prop, R=10)
    # beta = 0 and beta = inf in the schedule (thresholds +inf and 0)
    cs, ce = gnp(rng, 12, 0.6)
    out["beta0_inf_int"] = dict(h=rng.integers(-1, 2, 12).astype(float), cs=cs, ce=ce,
                                cw=rng.choice([-1.0, 1.0], len(cs)), betas=np.array([0.0, 0.5, np.inf, 1.0, np.inf]),
                                spb=2, order=0, prop=1, R=8)
    out["beta0_inf_real_gibbs"] = dict(h=rng.uniform(-1, 1, 12), cs=cs, ce=ce, cw=rng.uniform(-1, 1, len(cs)),
                                       betas=np.array([0.0, 0.5, np.inf, 1.0]), spb=2, order=0, prop=0, R=8)
    # single variable, no couplers; and null BQM (all zeros)
    out["single_var"] = dict(h=np.array([0.7]), cs=np.zeros(0, np.int32), ce=np.zeros(0, np.int32),
                             cw=np.zeros(0), betas=np.linspace(0.1, 2, 10), spb=1, order=0, prop=1, R=6)
    out["null_bqm"] = dict(h=np.zeros(9), cs=np.zeros(0, np.int32), ce=np.zeros(0, np.int32), cw=np.zeros(0),
                           betas=np.geomspace(0.1, 1, 7), spb=1, order=0, prop=1, R=5)
    return out


def main():
    if not oracle.have_ref():
        oracle.build()
    if not oracle.have_ref():
        raise SystemExit("reference core not built (needs the reference tree)")
    for name, c in cases().items():
        n = len(c["h"])
        R = c["R"]
        rng = np.random.default_rng(zlib.crc32(name.encode()))
        init = (2 * rng.integers(0, 2, (R, n)) - 1).astype(np.int8)
        seeds = oracle.read_seeds(1000 + R, R)
        seeds[0] = 0  # seed 0 maps to the all-ones state (cpu_sa.cpp:283)
        st, en = oracle.ref_sa_read_by_read(c["h"], c["cs"], c["ce"], c["cw"], c["spb"], c["betas"], seeds,
                                            init, c["order"], c["prop"])
        np.savez(os.path.join(HERE, f"sa_{name}.npz"), h=np.asarray(c["h"], np.float64), cs=c["cs"], ce=c["ce"],
                 cw=np.asarray(c["cw"], np.float64), betas=np.asarray(c["betas"], np.float64),
                 spb=np.int64(c["spb"]), order=np.int64(c["order"]), prop=np.int64(c["prop"]),
                 seeds=seeds, init=init, states=st, energies=en)
        print(name, n, R, en[:4])


if __name__ == "__main__":
    main()
