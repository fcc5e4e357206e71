"""CPU checks of the path-integral oracle's chain updates
(orc_pimc_sample_chains, the discrete form of localPIMC.cpp:385-558) against
what the mathematics fixes: the transfer matrix of the P-slice Trotter model
with the chain couplings added (ferromagnetic chains: Swendsen-Wang joins are
exact), the model without them (Jchain >= 0 never joins and adds no field),
qubits outside every chain left as they are, and determinism."""
import math

import numpy as np
import pytest
from scipy import stats as sps

import oracle
from oracle import pimc_exact as px

# Two chains of three qubits (0-1-2, 3-4-5) with couplers between the chains
# only (the reference's contract: chain couplers are not in the coupler list).
H6 = np.array([0.3, -0.2, 0.45, -0.1, 0.05, 0.25])
CS6 = [0, 1, 2, 0, 2]
CE6 = [3, 4, 5, 5, 3]
CW6 = [0.7, -1.0, 0.5, 0.4, -0.6]


def _mask(P):
    return np.uint32(0xFFFFFFFF) if P == 32 else np.uint32((1 << P) - 1)


def _chi2_pvalue(idx, prob, R):
    counts = np.bincount(idx, minlength=len(prob)).astype(float)
    exp = prob * R
    keep = exp > 5
    obs = np.append(counts[keep], counts[~keep].sum())
    ex = np.append(exp[keep], exp[~keep].sum())
    if ex[-1] == 0:
        obs, ex = obs[:-1], ex[:-1]
    return sps.chi2.sf(((obs - ex) ** 2 / ex).sum(), len(obs) - 1)


def _popcount(a):
    a = a.astype(np.uint64)
    c = np.zeros(a.shape, dtype=np.int64)
    for k in range(32):
        c += ((a >> np.uint64(k)) & np.uint64(1)).astype(np.int64)
    return c


def _run(Q, jc, hp, hd, P, R, S, seed, h=H6, cs=CS6, ce=CE6, cw=CW6, W0=None, nthreads=8):
    rng = np.random.default_rng(seed)
    if W0 is None:
        W0 = (rng.integers(0, 2**32, size=(R, len(h)), dtype=np.uint64).astype(np.uint32) & _mask(P))
    W, _ = oracle.pimc_sample_chains(h, cs, ce, cw, 1, np.full(S, hp), np.full(S, hd), oracle.read_seeds(seed, R),
                                     W0, Q, jc, num_slices=P, nthreads=nthreads)
    return W0, W


@pytest.mark.parametrize("jc,hp,hd,P", [(-1.0, 1.0, 1.0, 4), (-0.6, 2.0, 0.5, 8), (-1.5, 0.5, 3.0, 5),
                                        (-0.8, 1.5, 1.5, 32), (-2.0, 1.0, 40.0, 3)])
def test_ferromagnetic_chains_sample_the_trotter_model(jc, hp, hd, P):
    """Ferromagnetic chains: the stationary law on every slice is that of the
    P-slice model of the couplers plus the chain couplings."""
    R = 40000
    _, W = _run(3, jc, hp, hd, P, R, 40, 21)
    ccs, cce, ccw = px.chain_couplers(6, 3, jc)
    tm = px.TrotterModel(H6, CS6 + ccs, CE6 + cce, CW6 + ccw, hp, hd, P)
    st = np.where(W & 1, 1, -1)
    assert _chi2_pvalue(px.state_index(st), tm.slice_distribution(), R) > 1e-4
    rot = ((W << np.uint32(1)) | (W >> np.uint32(P - 1))) & _mask(P)
    k = _popcount(W ^ rot) / P
    z = (k.mean(axis=0) - tm.kink_probability()) / (k.std(axis=0) / math.sqrt(R) + 1e-12)
    assert np.max(np.abs(z)) < 5.0


def test_chains_of_two_and_six():
    """Other chain lengths: three chains of two, one chain of six."""
    R, P, hp, hd, jc = 40000, 6, 1.2, 1.3, -0.9
    for Q in (2, 6):
        _, W = _run(Q, jc, hp, hd, P, R, 40, 22 + Q)
        ccs, cce, ccw = px.chain_couplers(6, Q, jc)
        # couplers inside a chain of 2 would break the contract: (0,3) etc. join chains {0,1},{2,3},{4,5}
        inside = [(u // Q) == (v // Q) for u, v in zip(CS6, CE6)]
        cs = [u for u, i in zip(CS6, inside) if not i]
        ce = [v for v, i in zip(CE6, inside) if not i]
        cw = [w for w, i in zip(CW6, inside) if not i]
        _, W = _run(Q, jc, hp, hd, P, R, 40, 22 + Q, cs=cs, ce=ce, cw=cw)
        tm = px.TrotterModel(H6, cs + ccs, ce + cce, cw + ccw, hp, hd, P)
        assert _chi2_pvalue(px.state_index(np.where(W & 1, 1, -1)), tm.slice_distribution(), R) > 1e-4


@pytest.mark.parametrize("jc", [0.0, 0.7])
def test_non_ferromagnetic_chain_coupling_is_inert(jc):
    """Jchain >= 0: pNotJoin >= 1 never joins and the coupling adds no field,
    so the chains sample the couplers' model alone (as localPIMC.cpp:385-452
    does)."""
    R, P, hp, hd = 40000, 8, 1.0, 1.0
    _, W = _run(3, jc, hp, hd, P, R, 40, 31)
    tm = px.TrotterModel(H6, CS6, CE6, CW6, hp, hd, P)
    assert _chi2_pvalue(px.state_index(np.where(W & 1, 1, -1)), tm.slice_distribution(), R) > 1e-4


def test_single_qubit_chains_sample_the_trotter_model():
    R, P, hp, hd = 40000, 4, 1.0, 1.0
    _, W = _run(1, -5.0, hp, hd, P, R, 40, 41)
    tm = px.TrotterModel(H6, CS6, CE6, CW6, hp, hd, P)
    assert _chi2_pvalue(px.state_index(np.where(W & 1, 1, -1)), tm.slice_distribution(), R) > 1e-4


def test_classical_limit_of_strong_chains():
    """Hd = 0 and constant worldlines: each chain's joins at Jchain -> -inf
    make it one spin; the law is Boltzmann at beta = Hp of the couplers plus
    chain couplings."""
    R, hp, jc = 40000, 0.8, -1.0
    rng = np.random.default_rng(51)
    W0 = np.where(rng.integers(0, 2, size=(R, 6)) == 1, np.uint32(0xFFFFFFFF), np.uint32(0))
    _, W = _run(3, jc, hp, 0.0, 32, R, 40, 51, W0=W0)
    assert np.all((W == 0) | (W == 0xFFFFFFFF))
    ccs, cce, ccw = px.chain_couplers(6, 3, jc)
    p = px.boltzmann(H6, CS6 + ccs, CE6 + cce, CW6 + ccw, hp)
    assert _chi2_pvalue(px.state_index(np.where(W & 1, 1, -1)), p, R) > 1e-4


def test_qubits_outside_chains_are_never_updated():
    """n = 7, Q = 3: chains {0,1,2}, {3,4,5}; qubit 6 is in no chain
    (localPIMC.cpp:146 updates chains c < n / Q only)."""
    h = np.append(H6, 0.4)
    W0, W = _run(3, -1.0, 1.0, 1.0, 32, 64, 10, 61, h=h, cs=CS6 + [6], ce=CE6 + [0], cw=CW6 + [0.9])
    np.testing.assert_array_equal(W[:, 6], W0[:, 6])
    assert not np.array_equal(W[:, :6], W0[:, :6])


def test_chain_determinism_and_identity():
    rng = np.random.default_rng(71)
    W0 = rng.integers(0, 2**32, size=(6, 6), dtype=np.uint64).astype(np.uint32)
    seeds = oracle.read_seeds(5, 6)
    run = lambda sd, nt, S=4, spb=1: oracle.pimc_sample_chains(H6, CS6, CE6, CW6, spb, [1.0] * S, [1.0] * S, sd,
                                                               W0, 3, -1.0, nthreads=nt)[0]
    a, b = run(seeds, 1), run(seeds, 3)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, run(oracle.read_seeds(6, 6), 1))
    np.testing.assert_array_equal(run(seeds, 1, S=0), W0)
    np.testing.assert_array_equal(run(seeds, 1, spb=0), W0)
    with pytest.raises(ValueError):
        oracle.pimc_sample_chains(H6, [0, 9], [1, 2], [1.0, 1.0], 1, [1.0], [1.0], seeds[:1], W0[:1], 3, -1.0)
