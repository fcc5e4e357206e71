"""GPU parity of the steepest-descent kernel through the C ABI: bit-exact
states, energies and step counts against the oracle (itself bit-exact against
the reference's descent.cpp), plus the reference's sampler and composite tests."""
import numpy as np
import pytest

import oracle
from conftest import lattice3d, random_regular
from misamplers import (BinaryQuadraticModel, SimulatedAnnealingSampler, SteepestDescentComposite,
                        SteepestDescentSolver)
from misamplers.greedy.descent import steepest_gradient_descent

pytestmark = pytest.mark.gpu


def _check(h, cs, ce, cw, st0, max_steps=0xFFFFFFFF):
    print(f"case n={len(h)} m={len(cs)} R={st0.shape[0]}", flush=True)
    st = np.array(st0, dtype=np.int8, order="C", copy=True)
    stats = {}
    g = steepest_gradient_descent(st.shape[0], h, cs, ce, cw, st, max_steps=max_steps, stats=stats)
    o = oracle.descent(h, cs, ce, cw, st0, nthreads=8, max_steps=max_steps)
    for x, y in zip(g, o):
        np.testing.assert_array_equal(x, y)
    return g


def test_lattice_integer():
    rng = np.random.default_rng(1)
    cs, ce = lattice3d(16)
    cw = rng.choice([-1.0, 1.0], len(cs))
    st0 = rng.choice(np.array([-1, 1], np.int8), size=(256, 4096))
    _, _, steps = _check(np.zeros(4096), cs, ce, cw, st0)
    assert steps.mean() > 100


def test_real_dense_rows_over_64():
    rng = np.random.default_rng(2)
    n = 150
    pairs = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < 0.7]
    cs = np.array([a for a, _ in pairs], np.int32)
    ce = np.array([b for _, b in pairs], np.int32)
    _check(rng.normal(size=n), cs, ce, rng.normal(size=len(cs)), rng.choice(np.array([-1, 1], np.int8),
                                                                              size=(100, n)))


def test_repeats_selfloops_and_ragged_sizes():
    rng = np.random.default_rng(3)
    for n in (1, 2, 63, 65, 130):
        pairs = [(u, v) for u in range(n) for v in range(u, n) if rng.random() < 0.1]
        pairs += pairs[:5]
        cs = np.array([a for a, _ in pairs], np.int32)
        ce = np.array([b for _, b in pairs], np.int32)
        _check(rng.normal(size=n), cs, ce, rng.normal(size=len(cs)),
               rng.choice(np.array([-1, 1], np.int8), size=(37, n)))


def test_many_chunks_and_step_cap():
    rng = np.random.default_rng(4)
    cs, ce = random_regular(5000, 3, rng)
    cw = rng.normal(size=len(cs))
    h = rng.normal(0, 0.2, 5000)
    st0 = rng.choice(np.array([-1, 1], np.int8), size=(12, 5000))
    _check(h, cs, ce, cw, st0)
    _, _, steps = _check(h, cs, ce, cw, st0, max_steps=50)
    assert np.all(steps == 50)


def test_hub_rows_over_64_with_repeats_and_many_chunks():
    """A variable with more than 64 neighbours sends the kernel to the CSR rows;
    its repeated coupler and self-loop make that row an ordered (one-lane) row,
    and n > 4096 keeps the chunk minima in local memory."""
    rng = np.random.default_rng(5)
    n = 4500
    cs, ce = random_regular(n, 3, rng)
    hub = rng.choice(np.arange(1, n), 90, replace=False)
    cs = np.concatenate([cs, np.zeros(90, np.int32), [0, 0], [hub[3]]]).astype(np.int32)
    ce = np.concatenate([ce, hub, [hub[7], 0], [0]]).astype(np.int32)
    cw = rng.normal(size=len(cs))
    st0 = rng.choice(np.array([-1, 1], np.int8), size=(10, n))
    _check(rng.normal(0, 0.3, n), cs, ce, cw, st0)


def test_energy_sum_dyadic_and_rounding_integers():
    """Dyadic h and J (multiples of 1/8) sum exactly in any order; integers near
    2^50 do not (the sequential float64 sum rounds), so their energies must
    follow the reference's order."""
    rng = np.random.default_rng(6)
    cs, ce = random_regular(600, 4, rng)
    st0 = rng.choice(np.array([-1, 1], np.int8), size=(20, 600))
    _check(rng.integers(-40, 40, 600) / 8.0, cs, ce, rng.integers(-40, 40, len(cs)) / 8.0, st0)
    big = float(2**50)
    h = rng.integers(-3, 4, 600) * big + rng.integers(-9, 10, 600)
    cw = rng.integers(-3, 4, len(cs)) * big + rng.integers(-9, 10, len(cs))
    st, en, _ = _check(h, cs, ce, cw, st0)
    hi, wi = [int(x) for x in h], [int(x) for x in cw]
    exact = [sum(hi[v] * int(row[v]) for v in range(600)) +
             sum(int(row[a]) * wi[c] * int(row[b]) for c, (a, b) in enumerate(zip(cs, ce))) for row in st]
    assert any(float(x) != e for x, e in zip(exact, en))  # the sequential sum did round


def test_sampler_small_convex_ising_and_qubo():
    """tests/test_greedy_sampler.py:91-119."""
    ss = SteepestDescentSolver().sample_ising({0: 2, 1: 2}, {(0, 1): -1}, initial_states=([1, 1], [0, 1]))
    assert ss.first.sample == {0: -1, 1: -1} and ss.first.energy == -5
    assert ss.record.num_steps[0] == 2
    ss = SteepestDescentSolver().sample_qubo({(0, 0): 2, (1, 1): 2, (0, 1): -1}, initial_states=([1, 1], [0, 1]))
    assert ss.first.sample == {0: 0, 1: 0} and ss.first.energy == 0


def test_sampler_reads_and_seed():
    bqm = BinaryQuadraticModel({i: 0.1 * i for i in range(8)}, {(i, i + 1): -1.0 for i in range(7)}, 0.5, "SPIN")
    a = SteepestDescentSolver().sample(bqm, num_reads=20, seed=7)
    b = SteepestDescentSolver().sample(bqm, num_reads=20, seed=7)
    np.testing.assert_array_equal(a.record.sample, b.record.sample)
    assert len(a) == 20 and "timing" in a.info
    for d in a.data():
        assert d.energy == pytest.approx(bqm.energy(d.sample))


def test_composite_descends_sa_samples():
    rng = np.random.default_rng(5)
    cs, ce = random_regular(64, 3, rng)
    h = {i: float(rng.normal(0, 0.3)) for i in range(64)}
    J = {(int(u), int(v)): float(rng.normal()) for u, v in zip(cs, ce)}
    comp = SteepestDescentComposite(SimulatedAnnealingSampler())
    ss = comp.sample_ising(h, J, num_reads=16, num_sweeps=20, seed=3)
    child = SimulatedAnnealingSampler().sample_ising(h, J, num_reads=16, num_sweeps=20, seed=3)
    assert len(ss) == 16
    assert ss.record.energy.min() <= child.record.energy.min() + 1e-9
