"""The exact-inference oracle (oracle/tree_exact.py) pinned on the CPU: the
strip transfer matrix against brute-force enumeration (spin and binary), and
both against the open chain's closed form Z = 2 (2 cosh(beta J))^(n-1)."""
import numpy as np
import pytest

from oracle.tree_exact import enumerate_exact, strip_exact, strip_problem


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
