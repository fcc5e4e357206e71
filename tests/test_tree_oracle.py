"""The exact-inference oracle (oracle/tree_exact.py) pinned on the CPU: the
strip transfer matrix against brute-force enumeration (spin and binary), both
against the open chain's closed form Z = 2 (2 cosh(beta J))^(n-1), and the
numpy bucket elimination against enumeration (random graphs with repeated
couplers and self-loops, random orders) and the transfer matrix."""
import numpy as np
import pytest

from oracle.tree_exact import bucket_eliminate, enumerate_exact, strip_exact, strip_problem


@pytest.mark.parametrize("low", [-1, 0])
@pytest.mark.parametrize("w,L", [(1, 9), (3, 4), (4, 3)])
def test_strip_matches_enumeration(low, w, L):
    rng = np.random.default_rng(w * 10 + L)
    h, Jv, Jh, (hf, cs, ce, cw) = strip_problem(w, L, rng, low)
    lz, p1, pv, ph = strip_exact(h, Jv, Jh, beta=0.7, low=low)
    lz2, p12, pairs = enumerate_exact(hf, cs, ce, cw, beta=0.7, low=low)
    assert lz == pytest.approx(lz2, rel=1e-12, abs=1e-12)
    np.testing.assert_allclose(p1.T.ravel(), p12, atol=1e-12)
    want = np.concatenate([pv.transpose(1, 0, 2).reshape(-1, 4), ph.transpose(1, 0, 2).reshape(-1, 4)])
    np.testing.assert_allclose(want, pairs, atol=1e-12)


def test_open_chain_closed_form():
    n, J, beta = 12, 0.8, 1.3
    lz, p1, _ = enumerate_exact(np.zeros(n), np.arange(n - 1), np.arange(1, n), np.full(n - 1, J), beta=beta)
    assert lz == pytest.approx(np.log(2) + (n - 1) * np.log(2 * np.cosh(beta * J)), rel=1e-13)
    np.testing.assert_allclose(p1, 0.5, atol=1e-13)
    h = np.zeros((1, 200))
    lz2, p12, _, ph = strip_exact(h, np.zeros((0, 200)), np.full((1, 199), J), beta=beta)
    assert lz2 == pytest.approx(np.log(2) + 199 * np.log(2 * np.cosh(beta * J)), rel=1e-13)
    # nearest-neighbour correlation of the open chain with E = J s s:
    # <s s'> = -tanh(beta J)
    corr = ph[0, :, 0] + ph[0, :, 3] - ph[0, :, 1] - ph[0, :, 2]
    np.testing.assert_allclose(corr, -np.tanh(beta * J), atol=1e-12)


def _merged(cs, ce):
    """First-appearance index of each coupler's pair (self-loops: None)."""
    seen, idx = {}, []
    for u, v in zip(cs, ce):
        if u == v:
            idx.append(None)
            continue
        key = (min(u, v), max(u, v))
        seen.setdefault(key, len(seen))
        idx.append(seen[key])
    return idx


@pytest.mark.parametrize("low", [-1, 0])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_bucket_elimination_matches_enumeration(low, seed):
    rng = np.random.default_rng(40 + seed)
    n = 13
    pairs = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < 0.3]
    pairs += [(4, 4), (7, 2), (2, 7)]             # a self-loop and a repeated coupler, reversed
    cs = np.array([a for a, _ in pairs])
    ce = np.array([b for _, b in pairs])
    h, cw = rng.normal(size=n), rng.normal(size=len(cs))
    order = list(rng.permutation(n))
    lz, p1, cells = bucket_eliminate(h, cs, ce, cw, order, beta=0.9, low=low)
    lz2, p12, pairs2 = enumerate_exact(h, cs, ce, cw, beta=0.9, low=low)
    assert lz == pytest.approx(lz2, rel=1e-12, abs=1e-12)
    np.testing.assert_allclose(p1, p12, atol=1e-12)
    for k, (m, (u, v)) in enumerate(zip(_merged(cs, ce), pairs)):
        if m is None:
            continue
        got = cells[m]
        first = [(a, b) for a, b in pairs if a != b and (min(a, b), max(a, b)) == (min(u, v), max(u, v))][0]
        if first != (u, v):                      # cells are ordered (first u, first v)
            got = got[[0, 2, 1, 3]]
        np.testing.assert_allclose(got, pairs2[k], atol=1e-12)
    assert bucket_eliminate(h, cs, ce, cw, order, beta=0.9, low=low, marginals=False) == pytest.approx(lz, rel=1e-14)


def test_bucket_elimination_matches_transfer_matrix():
    rng = np.random.default_rng(44)
    h, Jv, Jh, (hf, cs, ce, cw) = strip_problem(6, 30, rng, scale=0.8)
    lz, p1, cells = bucket_eliminate(hf, cs, ce, cw, list(range(180)), beta=1.2)
    lz2, p12, pv, ph = strip_exact(h, Jv, Jh, beta=1.2)
    assert lz == pytest.approx(lz2, rel=1e-12)
    np.testing.assert_allclose(p1, p12.T.ravel(), atol=1e-12)
    want = np.concatenate([pv.transpose(1, 0, 2).reshape(-1, 4), ph.transpose(1, 0, 2).reshape(-1, 4)])
    np.testing.assert_allclose(cells, want, atol=1e-12)


@pytest.mark.parametrize("low", [-1, 0])
def test_kbest_elimination_against_enumeration(low):
    # min-sum elimination with k-best lists equals the k smallest of all 2^n
    # energies (self-loops, a repeated reversed coupler, isolated variables)
    from oracle.tree_exact import kbest_bruteforce, kbest_eliminate
    rng = np.random.default_rng(91 + low)
    for n, p in ((1, 1.0), (4, 0.6), (9, 0.35), (11, 0.2)):
        pairs = [(u, v) for u in range(n) for v in range(u, n) if rng.random() < p]
        if pairs and pairs[-1][0] != pairs[-1][1]:
            pairs.append((pairs[-1][1], pairs[-1][0]))
        cs = [a for a, _ in pairs]
        ce = [b for _, b in pairs]
        h, cw = rng.normal(size=n), rng.normal(size=len(cs))
        for k in (1, 2, 7, 2 ** n + 3):
            ref = kbest_bruteforce(h, cs, ce, cw, low, k, offset=0.5)[0]
            got = kbest_eliminate(h, cs, ce, cw, low, k, offset=0.5)
            assert len(got) == min(k, 2 ** n)
            np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)
