"""CPU checks of the path-integral oracle (oracle/pimc_oracle.c) against what
the mathematics fixes: the exact transfer matrix of the P-slice Trotter model,
the quantum density matrix as P grows, the classical Boltzmann law at Hd = 0
or P = 1, and the f32 helpers against libm.  The reference's own
continuous-time core (compiled into oracle/_ref) is checked against the same
quantum law, which ties the two discretisations to one model."""
import math

import numpy as np
import pytest
from scipy import stats as sps

import oracle
from conftest import lattice3d, random_regular
from oracle import pimc_exact as px

# A frustrated 5-spin instance with fields (odd cycle 0-1-2-3-4-0 plus a chord).
H5 = np.array([0.3, -0.2, 0.45, -0.1, 0.05])
CS5 = [0, 1, 2, 3, 0, 1]
CE5 = [1, 2, 3, 4, 4, 3]
CW5 = [0.7, -1.0, 0.5, -0.8, 1.1, 0.3]


def _masks(P):
    return np.uint32(0xFFFFFFFF) if P == 32 else np.uint32((1 << P) - 1)


def _random_worldlines(rng, R, n, P):
    W = rng.integers(0, 2**32, size=(R, n), dtype=np.uint64).astype(np.uint32)
    return W & _masks(P)


def _popcount(a):
    a = a.astype(np.uint64)
    c = np.zeros(a.shape, dtype=np.int64)
    for k in range(32):
        c += ((a >> np.uint64(k)) & np.uint64(1)).astype(np.int64)
    return c


def _kinks(W, P):
    rot = ((W << np.uint32(1)) | (W >> np.uint32(P - 1))) & _masks(P)
    return _popcount(W ^ rot)


def _chi2_pvalue(idx, prob, R):
    counts = np.bincount(idx, minlength=len(prob)).astype(float)
    exp = prob * R
    keep = exp > 5
    # pool the sparse cells into one
    obs = np.append(counts[keep], counts[~keep].sum())
    ex = np.append(exp[keep], exp[~keep].sum())
    if ex[-1] == 0:
        obs, ex = obs[:-1], ex[:-1]
    chi2 = ((obs - ex) ** 2 / ex).sum()
    return sps.chi2.sf(chi2, len(obs) - 1)


# ---------------------------------------------------------------------------
# f32 helpers and tables
# ---------------------------------------------------------------------------

def test_exp2f_polynomial():
    ys = np.concatenate([np.linspace(-92.0, 92.0, 4001), [-0.5, 0.0, 1e-7, 0.999999, 1.0, 63.99]])
    for y in ys:
        y32 = float(np.float32(y))
        got = oracle.pimc_exp2f(y32)
        assert abs(got / 2.0 ** y32 - 1.0) < 2e-7, y32


def test_log2u_polynomial():
    vs = np.unique(np.concatenate([np.arange(1, 2000), np.geomspace(1, 2**24, 3000).astype(np.int64),
                                   [2**23, 2**24 - 1, 2**24]]))
    for v in vs:
        got = oracle.pimc_log2u(int(v))
        want = math.log2(int(v)) - 24.0
        assert abs(got - want) < 4e-7 * max(1.0, abs(want)), v
    assert oracle.pimc_log2u(2**24) == 0.0


def test_step_consts():
    for P in (1, 8, 32):
        bs, gi, md = oracle.pimc_step_consts(2.5, 0.0, P)
        assert md == 0 and bs == 2.5 / P
        bs, gi, md = oracle.pimc_step_consts(1.0, -1.0, P)
        assert md == 0
        bs, gi, md = oracle.pimc_step_consts(1.0, 1e5, P)
        assert md == 2
        for hd in (1e-3, 0.3, 3.0, 30.0):
            bs, gi, md = oracle.pimc_step_consts(1.0, hd, P)
            if math.tanh(hd / P) == 1.0:  # every bond cut
                assert md == 2
                continue
            assert md == 1
            # P(gap >= 1) = 2^(1/ginv) must be the keep probability 1 - tanh(Hd/P).
            keep = 2.0 ** (1.0 / gi)
            assert keep == pytest.approx(1.0 - math.tanh(hd / P), rel=1e-6)
    assert oracle.pimc_step_consts(1.0, 1e-300, 32)[2] == 0  # no representable cut law


def _xoshiro128p(state, count):
    """xoshiro128+ from its published recurrence (Blackman and Vigna), on
    Python integers."""
    s = list(state)
    M = 0xFFFFFFFF
    out = []
    for _ in range(count):
        out.append((s[0] + s[3]) & M)
        t = (s[1] << 9) & M
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = ((s[3] << 11) | (s[3] >> 21)) & M
    return out


def test_xoshiro_recurrence_and_bits():
    st = [0x12345678, 0x9ABCDEF0, 0x0F1E2D3C, 0x4B5A6978]
    out = oracle.pimc_xoshiro(st, 200000)
    assert out[:64].tolist() == _xoshiro128p(st, 64)
    # upper 24 bits (the ones the sampler uses) are uniform
    bits = ((out[:, None] >> np.arange(8, 32, dtype=np.uint32)[None, :]) & 1).mean(axis=0)
    assert np.all(np.abs(bits - 0.5) < 5 * 0.5 / math.sqrt(len(out)))
    assert len(np.unique(out)) > 199990


# ---------------------------------------------------------------------------
# Colouring
# ---------------------------------------------------------------------------

def test_colouring_is_proper_and_greedy():
    rng = np.random.default_rng(3)
    cs, ce = random_regular(200, 6, rng)
    nc, col = oracle.pimc_colouring(200, cs, ce)
    assert np.all(col[cs] != col[ce])
    nbrs = [set() for _ in range(200)]
    for u, v in zip(cs, ce):
        nbrs[u].add(v)
        nbrs[v].add(u)
    for v in range(200):
        lower = {col[u] for u in nbrs[v] if u < v}
        assert col[v] == min(set(range(nc + 1)) - lower)
    assert nc == col.max() + 1


def test_colouring_lattice_bipartite():
    cs, ce = lattice3d(6)
    nc, col = oracle.pimc_colouring(216, cs, ce)
    assert nc == 2


def test_colouring_ignores_selfloops():
    nc, col = oracle.pimc_colouring(3, [0, 1, 1], [0, 1, 2])
    assert nc == 2 and col[0] == 0


# ---------------------------------------------------------------------------
# Equilibrium against exact laws
# ---------------------------------------------------------------------------

def _equilibrate(h, cs, ce, cw, hp, hd, P, R, S, seed, rng, classical_init=False):
    seeds = oracle.read_seeds(seed, R)
    W0 = _random_worldlines(rng, R, len(h), P)
    if classical_init:
        W0 = np.where(W0 & 1, _masks(P), np.uint32(0)).astype(np.uint32)
    W, _ = oracle.pimc_sample(h, cs, ce, cw, 1, np.full(S, hp), np.full(S, hd), seeds, W0,
                              num_slices=P, nthreads=8)
    return W


@pytest.mark.parametrize("hp,hd,P", [(1.0, 1.0, 4), (2.0, 0.5, 8), (0.5, 3.0, 4), (1.5, 1.5, 32),
                                     (3.0, 2.0, 5), (1.0, 40.0, 3)])
def test_trotter_equilibrium(hp, hd, P):
    rng = np.random.default_rng(11)
    R = 40000
    W = _equilibrate(H5, CS5, CE5, CW5, hp, hd, P, R, 40, 7, rng)
    tm = px.TrotterModel(H5, CS5, CE5, CW5, hp, hd, P)
    st = np.where(W & 1, 1, -1)
    assert _chi2_pvalue(px.state_index(st), tm.slice_distribution(), R) > 1e-4
    # kink probability per site and bond
    k = _kinks(W, P) / P
    z = (k.mean(axis=0) - tm.kink_probability()) / (k.std(axis=0) / math.sqrt(R) + 1e-12)
    assert np.max(np.abs(z)) < 5.0


def test_trotter_equilibrium_other_slices():
    """Every slice has the same law (imaginary-time translation invariance)."""
    rng = np.random.default_rng(12)
    P, R = 6, 30000
    W = _equilibrate(H5, CS5, CE5, CW5, 1.2, 1.3, P, R, 40, 8, rng)
    p = px.TrotterModel(H5, CS5, CE5, CW5, 1.2, 1.3, P).slice_distribution()
    for k in (2, 5):
        st = np.where((W >> np.uint32(k)) & 1, 1, -1)
        assert _chi2_pvalue(px.state_index(st), p, R) > 1e-4


def test_classical_limits():
    rng = np.random.default_rng(13)
    R = 40000
    p = px.boltzmann(H5, CS5, CE5, CW5, 1.3)
    # Hd = 0 with constant worldlines: heat bath at beta = Hp
    W = _equilibrate(H5, CS5, CE5, CW5, 1.3, 0.0, 32, R, 30, 9, rng, classical_init=True)
    assert np.all((W == 0) | (W == 0xFFFFFFFF))
    assert _chi2_pvalue(px.state_index(np.where(W & 1, 1, -1)), p, R) > 1e-4
    # P = 1: the transverse field drops out of the single-slice model
    W = _equilibrate(H5, CS5, CE5, CW5, 1.3, 2.0, 1, R, 30, 10, rng)
    assert _chi2_pvalue(px.state_index(np.where(W & 1, 1, -1)), p, R) > 1e-4


def test_many_slices_approach_quantum():
    rng = np.random.default_rng(14)
    R = 40000
    hp, hd = 1.0, 1.0
    W = _equilibrate(H5, CS5, CE5, CW5, hp, hd, 32, R, 40, 11, rng)
    q = px.QuantumModel(H5, CS5, CE5, CW5, hp, hd)
    tm = px.TrotterModel(H5, CS5, CE5, CW5, hp, hd, 32)
    # Trotter error at P = 32 is far below the sampling error here ...
    assert np.abs(tm.slice_distribution() - q.slice_distribution()).max() < 2e-3
    st = np.where(W & 1, 1, -1).astype(float)
    zz = st.T @ st / R
    assert np.abs(zz - q.zz()).max() < 0.03


@pytest.mark.skipif(not oracle.have_ref_pimc(), reason="reference PIMC core not built")
def test_reference_core_samples_quantum_law():
    """The reference's continuous-time worldlines (localPIMC.cpp) sample
    exp(-Hp Hz + Hd sum X): Hp is the inverse temperature and Hd the scaled
    transverse field, the mapping this engine uses."""
    rng = np.random.default_rng(15)
    R = 30000
    for hp, hd in ((1.0, 1.0), (0.5, 3.0)):
        st0 = rng.choice([-1, 1], size=(R, 5)).astype(np.int8)
        st, en, _ = oracle.ref_pimc_call(H5, CS5, CE5, CW5, 1, np.full(40, hp), np.full(40, hd), st0,
                                         np.arange(1, 33, dtype=np.uint32), nthreads=8)
        q = px.QuantumModel(H5, CS5, CE5, CW5, hp, hd).slice_distribution()
        assert _chi2_pvalue(px.state_index(st), q, R) > 1e-4
        for r in range(20):
            assert en[r] == oracle.state_energy(st[r], H5, CS5, CE5, CW5)


# ---------------------------------------------------------------------------
# Determinism, identity, statistics
# ---------------------------------------------------------------------------

def test_zero_steps_identity():
    rng = np.random.default_rng(16)
    W0 = _random_worldlines(rng, 10, 5, 32)
    W, _ = oracle.pimc_sample(H5, CS5, CE5, CW5, 1, [], [], oracle.read_seeds(1, 10), W0)
    np.testing.assert_array_equal(W, W0)
    W, _ = oracle.pimc_sample(H5, CS5, CE5, CW5, 0, [1.0] * 5, [1.0] * 5, oracle.read_seeds(1, 10), W0)
    np.testing.assert_array_equal(W, W0)


def test_seed_determinism():
    rng = np.random.default_rng(17)
    W0 = _random_worldlines(rng, 6, 5, 32)
    a, _ = oracle.pimc_sample(H5, CS5, CE5, CW5, 1, [1.0] * 4, [1.0] * 4, oracle.read_seeds(5, 6), W0)
    b, _ = oracle.pimc_sample(H5, CS5, CE5, CW5, 1, [1.0] * 4, [1.0] * 4, oracle.read_seeds(5, 6), W0,
                              nthreads=3)
    c, _ = oracle.pimc_sample(H5, CS5, CE5, CW5, 1, [1.0] * 4, [1.0] * 4, oracle.read_seeds(6, 6), W0)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, c)


def test_statistics_are_slice0_snapshots():
    rng = np.random.default_rng(18)
    n, R, S, interval = 5, 4, 10, 3
    W0 = _random_worldlines(rng, R, n, 32)
    hp, hd = np.linspace(0.1, 2, S), np.linspace(2, 0, S)
    seeds = oracle.read_seeds(9, R)
    W, stats = oracle.pimc_sample(H5, CS5, CE5, CW5, 2, hp, hd, seeds, W0, interval=interval)
    assert stats.shape == (R, 1 + (S - 1) // interval, n)
    for col, k in enumerate(range(0, S, interval)):
        Wk, _ = oracle.pimc_sample(H5, CS5, CE5, CW5, 2, hp[:k + 1], hd[:k + 1], seeds, W0)
        np.testing.assert_array_equal(stats[:, col], np.where(Wk & 1, 1, -1))


def test_selfloops_are_constants():
    """A self-loop shifts every energy by J: same dynamics as without it."""
    rng = np.random.default_rng(19)
    W0 = _random_worldlines(rng, 8, 5, 32)
    seeds = oracle.read_seeds(3, 8)
    a, _ = oracle.pimc_sample(H5, CS5, CE5, CW5, 1, [1.0] * 6, [1.5] * 6, seeds, W0)
    b, _ = oracle.pimc_sample(H5, CS5 + [2, 4], CE5 + [2, 4], CW5 + [5.0, -3.0], 1, [1.0] * 6, [1.5] * 6,
                              seeds, W0)
    np.testing.assert_array_equal(a, b)


def test_bad_index_rejected():
    with pytest.raises(ValueError):
        oracle.pimc_sample(H5, [0, 7], [1, 2], [1.0, 1.0], 1, [1.0], [1.0], oracle.read_seeds(1, 1),
                           np.zeros((1, 5), np.uint32))
