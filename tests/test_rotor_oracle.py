"""The rotor-model oracle (oracle/rotor_oracle.c), pinned on the CPU:
bit-exact against the reference's own cpu_rotormc.cpp compiled into
oracle/_ref (every proposal and order, self-loops, repeated couplers, the
transverse-field window's integer corner cases, snapshots), and against
mathematics that does not depend on the reference: the null-model
reflection, the RNG -> proposal path replayed with Python integers, the exact
stationary law of one rotor under both angle proposals, and the Gibbs law of
independent reflected spins."""
import numpy as np
import pytest
from scipy import stats as sst

import oracle

PROPOSALS = list(oracle.ROTOR_PROPOSALS)
M64 = (1 << 64) - 1


def _problem(rng, n, p, selfloops=False, repeats=False):
    pairs = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < p]
    if selfloops:
        pairs += [(0, 0), (n - 1, n - 1)]
    if repeats and pairs:
        pairs += pairs[:3]
    cs = np.array([a for a, _ in pairs], np.int32)
    ce = np.array([b for _, b in pairs], np.int32)
    return rng.normal(size=n), cs, ce, rng.normal(size=len(cs)), rng.uniform(0.2, 1.5, n)


def _same(a, b):
    for x, y in zip(a, b):
        if x is None or y is None:
            assert x is None and y is None
        else:
            np.testing.assert_array_equal(x, y)


needs_ref = pytest.mark.skipif(not oracle.have_ref_rotor(), reason="compiled reference core not built")


@needs_ref
@pytest.mark.parametrize("proposal", PROPOSALS)
@pytest.mark.parametrize("randomize", [False, True])
def test_oracle_matches_compiled_reference(proposal, randomize):
    rng = np.random.default_rng(11)
    h, cs, ce, cw, tf = _problem(rng, 40, 0.15, selfloops=True, repeats=True)
    Hp = np.geomspace(0.05, 4.0, 25)
    Hd = np.linspace(3.0, 0.0, 25)
    seeds = rng.integers(0, 2**63, 24, dtype=np.uint64)
    seeds[0] = 0   # seed 0 means the all-ones state
    st0 = rng.integers(0, 256, (24, 40), dtype=np.uint8)
    a = oracle.rotor_sample(h, cs, ce, cw, tf, 2, Hp, Hd, seeds, st0, randomize, proposal, interval=4, nthreads=8)
    b = oracle.ref_rotor(h, cs, ce, cw, tf, 2, Hp, Hd, seeds, st0, randomize, proposal, interval=4, nthreads=8)
    _same(a, b)
    assert a[2].shape == (24, 7, 40)   # 1 + (25 - 1) // 4 snapshots (not a multiple)


@needs_ref
def test_tf_window_corner_cases_match_reference():
    """rat = Hd / Hp below 1 (windows of every width), negative Hd (negative
    window, wrapped integer arithmetic), Hp = 0 (rat infinite, and NaN with
    Hd = 0), and a zero transverse field."""
    rng = np.random.default_rng(12)
    h, cs, ce, cw, tf = _problem(rng, 24, 0.2)
    tf[:5] = 0.0
    Hp = np.array([1.0, 1.0, 1.0, 2.0, 0.0, 0.0, 3.0, 1.0, 1.0, 5.0, 1.0, 1.0])
    Hd = np.array([0.0, 1 / 300, 0.3, 1.9, 1.0, 0.0, -0.7, -3.0, 0.999, 4.99, -1e12, 1e-300])
    seeds = rng.integers(1, 2**63, 20, dtype=np.uint64)
    st0 = rng.integers(0, 256, (20, 24), dtype=np.uint8)
    for randomize in (False, True):
        a = oracle.rotor_sample(h, cs, ce, cw, tf, 3, Hp, Hd, seeds, st0, randomize, "MetropolisTF", interval=1)
        b = oracle.ref_rotor(h, cs, ce, cw, tf, 3, Hp, Hd, seeds, st0, randomize, "MetropolisTF", interval=1)
        _same(a, b)


@needs_ref
def test_single_variable_and_no_couplers_match_reference():
    rng = np.random.default_rng(13)
    for n in (1, 5):
        h = rng.normal(size=n)
        e = np.zeros(0, np.int32)
        seeds = rng.integers(1, 2**63, 7, dtype=np.uint64)
        st0 = rng.integers(0, 256, (7, n), dtype=np.uint8)
        for prop in PROPOSALS:
            a = oracle.rotor_sample(h, e, e, np.zeros(0), np.ones(n), 5, [0.5, 2.0], [1.0, 0.1], seeds, st0,
                                    True, prop)
            b = oracle.ref_rotor(h, e, e, np.zeros(0), np.ones(n), 5, [0.5, 2.0], [1.0, 0.1], seeds, st0,
                                 True, prop)
            _same(a, b)


def test_tables():
    c, s = oracle.rotor_tables()
    th = 2 * np.pi * np.arange(256) / 256
    np.testing.assert_allclose(c, np.cos(th), atol=2e-16)
    np.testing.assert_allclose(s, np.abs(np.sin(th)), atol=2e-16)
    assert c[0] == 1.0 and c[128] == -1.0 and s[0] == 0.0 and s[64] == 1.0


def test_null_model_reflections_are_deterministic():
    """tests/test_rotor_annealing_sampler.py:687-712: with no fields every
    reflection has dE = 0 and Metropolis accepts it, so each sweep adds 128."""
    rng = np.random.default_rng(14)
    n = 100
    e = np.zeros(0, np.int32)
    st0 = rng.integers(0, 256, (3, n), dtype=np.uint8)
    for sweeps in (1, 2, 3):
        st, en, _ = oracle.rotor_sample(np.zeros(n), e, e, np.zeros(0), np.ones(n), sweeps, [1.0], [0.0],
                                        np.array([5, 6, 7], np.uint64), st0, False, "MetropolisNonErgodic")
        np.testing.assert_array_equal(st, (st0.astype(int) + 128 * sweeps) % 256)


def _xs_stream(seed):
    s0, s1 = (seed if seed else M64), 0
    while True:
        x, y = s0, s1
        s0 = y
        x = (x ^ (x << 23)) & M64
        s1 = x ^ y ^ (x >> 17) ^ (y >> 26)
        yield (s1 + y) & M64


def test_uniform_proposals_follow_the_rng_exactly():
    """With no fields and no transverse field every MetropolisUniform proposal
    has log-measure change 0 and is accepted, so after one sweep variable v
    holds (draw 32 + v) % 256 of the read's xorshift128+ stream (after the 32
    burn-in draws); with random order variable (draw) % n is visited."""
    n = 37
    e = np.zeros(0, np.int32)
    seeds = np.array([0, 1, 123456789, 2**63 + 5], np.uint64)
    st0 = np.zeros((4, n), np.uint8)
    st, _, _ = oracle.rotor_sample(np.zeros(n), e, e, np.zeros(0), np.zeros(n), 1, [1.0], [0.0], seeds, st0,
                                   False, "MetropolisUniform")
    for r, sd in enumerate(seeds):
        g = _xs_stream(int(sd))
        draws = [next(g) for _ in range(32 + n)]
        np.testing.assert_array_equal(st[r], [d % 256 for d in draws[32:]])
    st, _, _ = oracle.rotor_sample(np.zeros(n), e, e, np.zeros(0), np.zeros(n), 1, [1.0], [0.0], seeds, st0,
                                   True, "MetropolisUniform")
    for r, sd in enumerate(seeds):
        g = _xs_stream(int(sd))
        for _ in range(32):
            next(g)
        want = np.zeros(n, np.int64)
        for _ in range(n):
            v = next(g) % n
            want[v] = next(g) % 256
        np.testing.assert_array_equal(st[r], want)


@pytest.mark.parametrize("proposal,hd", [("MetropolisUniform", 0.8), ("MetropolisTF", 0.8), ("MetropolisTF", 3.0)])
def test_one_rotor_stationary_law(proposal, hd):
    """One rotor with field h and transverse field t at (Hp, Hd): the chain
    leaves P(i) proportional to exp(-Hp h cos_i + Hd t |sin_i|) invariant
    (both angle proposals are symmetric).  Chi-square of the final angles of
    many independent reads against that law, binned into 16 arcs."""
    c, s = oracle.rotor_tables()
    h, t, Hp = 0.7, 1.3, 1.5
    law = np.exp(-Hp * h * c + hd * t * s)
    law /= law.sum()
    R = 4000
    e = np.zeros(0, np.int32)
    seeds = np.arange(1, R + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    st, _, _ = oracle.rotor_sample([h], e, e, np.zeros(0), [t], 60, [Hp], [hd], seeds,
                                   np.zeros((R, 1), np.uint8), False, proposal, nthreads=8)
    obs = np.bincount(st[:, 0] // 16, minlength=16)
    exp_ = R * law.reshape(16, 16).sum(1)
    assert sst.chisquare(obs, exp_).pvalue > 1e-3


def test_gibbs_reflection_law_of_independent_spins():
    """tests/test_rotor_annealing_sampler.py:972-1010: independent spins under
    sequential GibbsNonErgodic from the poles take +1 (angle 0) with
    probability exp(-Hp h) / (2 cosh(Hp h)) after one sweep at the last Hp."""
    rng = np.random.default_rng(2023)
    n, R = 4, 4000
    h = -rng.random(n)
    e = np.zeros(0, np.int32)
    seeds = rng.integers(1, 2**63, R, dtype=np.uint64)
    st, _, _ = oracle.rotor_sample(h, e, e, np.zeros(0), np.ones(n), 1, [1.0, 2.0], [0.0, 0.0], seeds,
                                   np.zeros((R, n), np.uint8), False, "GibbsNonErgodic", nthreads=8)
    assert set(np.unique(st)) <= {0, 128}
    p = np.exp(-2.0 * h) / (2 * np.cosh(2.0 * h))
    k = (st == 0).sum(0)
    lo, hi = R * p - 4 * np.sqrt(R * p * (1 - p)), R * p + 4 * np.sqrt(R * p * (1 - p))
    assert np.all((k > lo) & (k < hi))


def test_energy_is_cos_weighted_in_reference_order():
    rng = np.random.default_rng(15)
    h, cs, ce, cw, tf = _problem(rng, 12, 0.4)
    st0 = rng.integers(0, 256, (5, 12), dtype=np.uint8)
    st, en, _ = oracle.rotor_sample(h, cs, ce, cw, tf, 0, [], [], np.arange(5, dtype=np.uint64), st0)
    np.testing.assert_array_equal(st, st0)   # no schedule steps: identity
    c, _ = oracle.rotor_tables()
    for r in range(5):
        e = 0.0
        for v in range(12):
            e += c[st[r, v]] * h[v]
        for k in range(len(cs)):
            e += c[st[r, cs[k]]] * cw[k] * c[st[r, ce[k]]]
        assert e == en[r]


def test_invalid_coupler_index():
    with pytest.raises(ValueError):
        oracle.rotor_sample([0.0, 0.0], [0], [2], [1.0], [1.0, 1.0], 1, [1.0], [0.0], [1],
                            np.zeros((1, 2), np.uint8))
