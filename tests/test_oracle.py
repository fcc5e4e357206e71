"""The oracle (C restatement of cpu_sa.cpp) pinned against the reference's
golden vectors, the compiled reference itself, and closed-form cases."""
import numpy as np
import pytest

import oracle
from conftest import golden_cases, lattice3d

CASES = golden_cases()


def test_golden_present():
    assert len(CASES) >= 10


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_golden(name):
    c = CASES[name]
    st, en = oracle.sa_sample(c["h"], c["cs"], c["ce"], c["cw"], int(c["spb"]), c["betas"], c["seeds"],
                              c["init"], int(c["order"]), int(c["prop"]))
    np.testing.assert_array_equal(st, c["states"])
    np.testing.assert_array_equal(en, c["energies"])  # bit-exact, real values included


def _xorshift_py(s0, s1, count):
    """xorshift128+ from its published recurrence, in Python integers."""
    M = (1 << 64) - 1
    out = []
    for _ in range(count):
        x, y = s0, s1
        s0 = y
        x ^= (x << 23) & M
        s1 = x ^ y ^ (x >> 17) ^ (y >> 26)
        out.append((s1 + y) & M)
    return out


def test_xorshift_known_answers():
    for s0, s1 in ((1, 0), ((1 << 64) - 1, 0), (0x123456789ABCDEF, 0xFEDCBA987654321)):
        assert oracle.xorshift_stream(s0, s1, 50) == _xorshift_py(s0, s1, 50)


def test_splitmix64_known_vector():
    # published splitmix64 outputs for state 0
    got = [int(x) for x in oracle.read_seeds(0, 3)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_zero_sweeps_identity():
    rng = np.random.default_rng(3)
    cs, ce = lattice3d(3)
    init = (2 * rng.integers(0, 2, (5, 27)) - 1).astype(np.int8)
    st, en = oracle.sa_sample(np.zeros(27), cs, ce, np.ones(len(cs)), 1, [], oracle.read_seeds(1, 5), init)
    np.testing.assert_array_equal(st, init)
    e = np.array([sum(init[r, a] * init[r, b] for a, b in zip(cs, ce)) for r in range(5)], float)
    np.testing.assert_array_equal(en, e)


def test_null_bqm_sequential_metropolis_flips_every_sweep():
    # tests/test_simulated_annealing_sampler.py:678-700: with all biases zero every
    # dE is 0 <= 0, so every spin flips on every sweep: after S sweeps s = (-1)^S s0.
    rng = np.random.default_rng(4)
    init = (2 * rng.integers(0, 2, (4, 10)) - 1).astype(np.int8)
    for S in (1, 2, 7):
        st, _ = oracle.sa_sample(np.zeros(10), [], [], [], 1, np.ones(S), oracle.read_seeds(2, 4), init)
        np.testing.assert_array_equal(st, init * (-1) ** S)


def test_greedy_limit_independent_spins():
    # beta = inf: threshold 0, only strictly downhill flips (sampler tests :790-845)
    h = np.array([1.0, -2.0, 0.5, -0.25])
    init = np.array([[1, 1, 1, 1], [-1, -1, -1, -1]], np.int8)
    st, en = oracle.sa_sample(h, [], [], [], 1, [np.inf], oracle.read_seeds(9, 2), init)
    np.testing.assert_array_equal(st, np.tile(-np.sign(h).astype(np.int8), (2, 1)))
    np.testing.assert_array_equal(en, [-np.abs(h).sum()] * 2)


def test_gibbs_acceptance_rate_single_spin():
    # one free spin with field h: Gibbs flips with probability 1/(1+exp(beta*dE));
    # the stationary P(s=+1) = e^{-beta h}/(2 cosh beta h)
    h, beta, R = np.array([0.3]), 1.7, 20000
    init = np.ones((R, 1), np.int8)
    st, _ = oracle.sa_sample(h, [], [], [], 50, [beta], oracle.read_seeds(11, R), init, proposal=oracle.GIBBS)
    p_up = np.exp(-beta * 0.3) / (2 * np.cosh(beta * 0.3))
    frac = (st[:, 0] == 1).mean()
    assert abs(frac - p_up) < 4 * np.sqrt(p_up * (1 - p_up) / R)


def test_bad_indices_and_lengths():
    with pytest.raises(ValueError, match="invalid variable"):
        oracle.sa_sample([0, 0], [0], [2], [1.0], 1, [1.0], oracle.read_seeds(1, 1), np.ones((1, 2), np.int8))
    with pytest.raises(ValueError, match="mismatched"):
        oracle.sa_sample([0, 0], [0], [1, 1], [1.0], 1, [1.0], oracle.read_seeds(1, 1), np.ones((1, 2), np.int8))


@pytest.mark.skipif(not oracle.have_ref(), reason="reference core not built here")
@pytest.mark.parametrize("order,prop", [(0, 1), (0, 0), (1, 1), (1, 0)])
def test_oracle_matches_reference_random_problems(order, prop):
    rng = np.random.default_rng(100 + 2 * order + prop)
    for trial in range(4):
        n = int(rng.integers(2, 50))
        pairs = [(u, v) for u in range(n) for v in range(u, n) if rng.random() < 0.3]
        cs = np.array([a for a, _ in pairs], np.int32)
        ce = np.array([b for _, b in pairs], np.int32)
        if trial % 2:
            cw, h = rng.uniform(-1, 1, len(pairs)), rng.uniform(-1, 1, n)
        else:
            cw, h = rng.choice([-1.0, 1.0], len(pairs)), rng.integers(-2, 3, n).astype(float)
        betas = np.geomspace(0.05, 8, 30)
        init = (2 * rng.integers(0, 2, (8, n)) - 1).astype(np.int8)
        seeds = oracle.read_seeds(trial, 8)
        a = oracle.sa_sample(h, cs, ce, cw, 2, betas, seeds, init, order, prop, nthreads=4)
        b = oracle.ref_sa_read_by_read(h, cs, ce, cw, 2, betas, seeds, init, order, prop)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
        c = oracle.sa_sample_stream(h, cs, ce, cw, 2, betas, 77, init, order, prop)
        d = oracle.ref_sa_call(h, cs, ce, cw, 2, betas, 77, init, order, prop)
        np.testing.assert_array_equal(c[0], d[0])
        np.testing.assert_array_equal(c[1], d[1])
