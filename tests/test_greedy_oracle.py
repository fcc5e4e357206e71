"""CPU checks of the steepest-descent oracle (oracle/greedy_oracle.c): bit-exact
against the reference's own descent.cpp (compiled into oracle/_ref) in both of
its solvers, the reference's known answer, and the local-minimum property; plus
host-side argument handling of the drop-in core function (no GPU needed)."""
import numpy as np
import pytest

import oracle
from conftest import lattice3d, random_regular
from misamplers.greedy.descent import steepest_gradient_descent


def _random(rng, n, p, integer, repeats=0):
    pairs = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < p]
    pairs += pairs[:repeats]
    cs = np.array([a for a, _ in pairs], np.int32)
    ce = np.array([b for _, b in pairs], np.int32)
    if integer:
        return rng.integers(-3, 4, n).astype(float), cs, ce, rng.integers(-3, 4, len(cs)).astype(float)
    return rng.normal(size=n), cs, ce, rng.normal(size=len(cs))


def _flip_energies(h, cs, ce, cw, s):
    f = np.array(h, dtype=float).copy()
    for u, v, w in zip(cs, ce, cw):
        if u != v:
            f[u] += s[v] * w
            f[v] += s[u] * w
    return -2 * s * f


@pytest.mark.skipif(not oracle.have_ref_greedy(), reason="reference descent core not built")
@pytest.mark.parametrize("integer,repeats", [(True, 0), (False, 0), (True, 3), (False, 4)])
def test_oracle_matches_reference(integer, repeats):
    rng = np.random.default_rng(10 + repeats + integer)
    for _ in range(6):
        n = int(rng.integers(1, 80))
        h, cs, ce, cw = _random(rng, n, 0.15, integer, repeats)
        st0 = rng.choice(np.array([-1, 1], np.int8), size=(13, n))
        a = oracle.descent(h, cs, ce, cw, st0, nthreads=4)
        for ls in (False, True):
            b = oracle.ref_descent(h, cs, ce, cw, st0, large_sparse_opt=ls, nthreads=4)
            for x, y in zip(a, b):
                np.testing.assert_array_equal(x, y)


@pytest.mark.skipif(not oracle.have_ref_greedy(), reason="reference descent core not built")
def test_oracle_matches_reference_lattice():
    rng = np.random.default_rng(3)
    cs, ce = lattice3d(8)
    cw = rng.choice([-1.0, 1.0], len(cs))
    st0 = rng.choice(np.array([-1, 1], np.int8), size=(40, 512))
    a = oracle.descent(np.zeros(512), cs, ce, cw, st0, nthreads=8)
    b = oracle.ref_descent(np.zeros(512), cs, ce, cw, st0, nthreads=8)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_known_answer_small_convex():
    """tests/test_greedy.py:26-51: global minimum (-1, -1), energy -5."""
    st, en, steps = oracle.descent([2, 2], [0], [1], [-1], np.array([[1, 1]], np.int8))
    np.testing.assert_array_equal(st, [[-1, -1]])
    np.testing.assert_array_equal(en, [-5])
    assert steps[0] == 2


def test_results_are_local_minima_and_energy_consistent():
    rng = np.random.default_rng(4)
    h, cs, ce, cw = _random(rng, 60, 0.2, False)
    st0 = rng.choice(np.array([-1, 1], np.int8), size=(20, 60))
    st, en, steps = oracle.descent(h, cs, ce, cw, st0)
    for r in range(20):
        assert np.all(_flip_energies(h, cs, ce, cw, st[r].astype(float)) >= -1e-9)
        assert en[r] == pytest.approx(oracle.state_energy(st[r], h, cs, ce, cw), abs=1e-12)
        assert en[r] <= oracle.state_energy(st0[r], h, cs, ce, cw) + 1e-12
    # zero steps exactly when the start is already a local minimum
    st2, en2, steps2 = oracle.descent(h, cs, ce, cw, st)
    np.testing.assert_array_equal(st2, st)
    assert np.all(steps2 == 0)


def test_selfloops_are_constants_and_terminate():
    """A positive self-loop makes the reference's dense bookkeeping cycle; the
    oracle (like the engine) treats it as the constant it is."""
    h, cs, ce, cw = [0.5, -0.25, 0.0], [0, 0, 1, 2], [0, 1, 2, 2], [3.0, -1.0, 0.5, -2.0]
    st0 = np.array([[1, 1, 1], [-1, 1, -1]], np.int8)
    st, en, steps = oracle.descent(h, cs, ce, cw, st0, max_steps=1000)
    st_ns, en_ns, _ = oracle.descent(h, [0, 1], [1, 2], [-1.0, 0.5], st0)
    np.testing.assert_array_equal(st, st_ns)
    np.testing.assert_allclose(en, en_ns + 3.0 - 2.0)
    assert np.all(steps < 1000)


def test_core_function_host_checks():
    st, en, k = steepest_gradient_descent(0, [0.0, 1.0], [0], [1], [1.0], np.empty((0, 2), np.int8))
    assert st.shape == (0, 2) and en.shape == (0,) and k.dtype == np.uint32
    with pytest.raises(ValueError):
        steepest_gradient_descent(2, [0.0, 1.0], [0], [1], [1.0], np.ones((2, 2), np.int16))
    with pytest.raises(ValueError, match="coupler indexes"):
        steepest_gradient_descent(2, [0.0, 1.0], [0], [5], [1.0], np.ones((2, 2), np.int8))
    with pytest.raises(ValueError, match="spins"):
        steepest_gradient_descent(2, [0.0, 1.0], [0], [1], [1.0], np.zeros((2, 2), np.int8))
