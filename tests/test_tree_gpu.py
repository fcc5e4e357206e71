"""GPU tests of exact inference by bucket elimination (misa_tree_*): log Z,
variable and pair marginals against brute-force enumeration and the strip
transfer matrix (oracle/tree_exact.py), exact samples against the marginals,
the numpy bucket elimination on models beyond both (a 10 x 48 strip, random
3-regular graphs), the reference's TreeDecompositionSampler tests
(tests/test_tree_sampler.py) restated, and the moments of fixed-beta sampling checked against exact
marginals on a model larger than enumeration reaches."""
import numpy as np
import pytest

from conftest import random_regular
from misamplers import BinaryQuadraticModel, TreeDecompositionSampler
from misamplers.tree import TreePlan
from oracle.tree_exact import bucket_eliminate, enumerate_exact, strip_exact, strip_problem

pytestmark = pytest.mark.gpu

TOL = 1e-10   # float64 log-domain sums in another order than the oracle's


def _pairs_in_plan_order(plan, cs, ce, pairs):
    key = {(min(u, v), max(u, v)): k for k, (u, v) in enumerate(zip(cs, ce))}
    idx = [key[(min(u, v), max(u, v))] for u, v in zip(plan.pairs_u, plan.pairs_v)]
    out = pairs[idx].copy()
    swap = [(u, v) != (cs[k], ce[k]) for u, v, k in zip(plan.pairs_u, plan.pairs_v, idx)]
    out[swap] = out[swap][:, [0, 2, 1, 3]]
    return out


@pytest.mark.parametrize("low", [-1, 0])
def test_random_graph_against_enumeration(low):
    rng = np.random.default_rng(31)
    n = 18
    pairs = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < 0.2]
    cs = np.array([a for a, _ in pairs], np.int32)
    ce = np.array([b for _, b in pairs], np.int32)
    h, cw = rng.normal(size=n), rng.normal(size=len(cs))
    for beta in (0.3, 1.0, 2.5):
        plan = TreePlan(h, cs, ce, cw, beta=beta, low=low)
        print(f"case enumeration n={n} low={low} beta={beta} width={plan.treewidth}", flush=True)
        lz, p1, pp = plan.marginals()
        lz2, p12, pairs2 = enumerate_exact(h, cs, ce, cw, beta=beta, low=low)
        assert lz == pytest.approx(lz2, rel=TOL, abs=TOL)
        np.testing.assert_allclose(p1, p12, atol=TOL)
        np.testing.assert_allclose(pp, _pairs_in_plan_order(plan, cs, ce, pairs2), atol=TOL)
        plan.close()


@pytest.mark.parametrize("low", [-1, 0])
@pytest.mark.parametrize("w,L", [(3, 60), (5, 40)])
def test_strips_against_transfer_matrix(low, w, L):
    rng = np.random.default_rng(w * 100 + L)
    h, Jv, Jh, (hf, cs, ce, cw) = strip_problem(w, L, rng, low, scale=0.7)
    plan = TreePlan(hf, cs, ce, cw, beta=1.1, low=low)
    print(f"case strip {w}x{L} low={low} width={plan.treewidth}", flush=True)
    assert plan.treewidth <= w
    lz, p1, pp = plan.marginals()
    lz2, p12, pv, ph = strip_exact(h, Jv, Jh, beta=1.1, low=low)
    assert lz == pytest.approx(lz2, rel=TOL)
    np.testing.assert_allclose(p1, p12.T.ravel(), atol=TOL)
    want = np.concatenate([pv.transpose(1, 0, 2).reshape(-1, 4), ph.transpose(1, 0, 2).reshape(-1, 4)])
    np.testing.assert_allclose(pp, _pairs_in_plan_order(plan, cs, ce, want), atol=TOL)
    plan.close()


@pytest.mark.parametrize("case", ["strip", "regular_spin", "regular_binary"])
def test_larger_models_against_numpy_elimination(case):
    """Single-bucket levels (the strip's chain of buckets, column order) and
    levels of many buckets (random graphs, min-fill) against the numpy
    elimination in the same order."""
    rng = np.random.default_rng(35)
    low = 0 if case == "regular_binary" else -1
    if case == "strip":
        h, Jv, Jh, (hf, cs, ce, cw) = strip_problem(10, 48, rng, scale=0.6)
        order = np.arange(480)
    else:
        n = 90
        cs, ce = random_regular(n, 3, rng)
        cs, ce = np.asarray(cs, np.int32), np.asarray(ce, np.int32)
        hf, cw = rng.normal(size=n), rng.normal(size=len(cs))
        order = None
    plan = TreePlan(hf, cs, ce, cw, beta=0.8, low=low, elimination_order=order)
    print(f"case {case} n={plan.n} width={plan.treewidth}", flush=True)
    lz, p1, pp = plan.marginals()
    st = plan.last_stats()
    if case == "strip":
        assert st.forward_levels == 480 and st.downward_levels == 479
    else:
        assert st.forward_levels < plan.n
    lz2, p12, cells = bucket_eliminate(hf, cs, ce, cw, list(plan.order), beta=0.8, low=low)
    assert lz == pytest.approx(lz2, rel=TOL)
    np.testing.assert_allclose(p1, p12, atol=TOL)
    np.testing.assert_allclose(pp, cells, atol=TOL)
    plan.recompute()   # the captured passes replay to the same values
    lz3, p13, pp3 = plan.marginals()
    assert lz3 == lz
    np.testing.assert_array_equal(p13, p1)
    np.testing.assert_array_equal(pp3, pp)
    plan.close()


def test_given_order_equals_min_fill_and_errors():
    rng = np.random.default_rng(32)
    h, Jv, Jh, (hf, cs, ce, cw) = strip_problem(3, 10, rng)
    a = TreePlan(hf, cs, ce, cw)
    b = TreePlan(hf, cs, ce, cw, elimination_order=np.arange(30)[::-1])
    la, pa, qa = a.marginals()
    lb, pb, qb = b.marginals()
    assert la == pytest.approx(lb, rel=TOL)
    np.testing.assert_allclose(pa, pb, atol=TOL)
    np.testing.assert_allclose(qa, qb, atol=TOL)
    a.close()
    b.close()
    with pytest.raises(ValueError):
        TreePlan(hf, cs, ce, cw, elimination_order=np.zeros(30, np.int32))
    n = 30   # a complete graph has width n - 1
    full = [(u, v) for u in range(n) for v in range(u + 1, n)]
    with pytest.raises(ValueError):
        TreeDecompositionSampler().sample(BinaryQuadraticModel({}, {p: 1.0 for p in full}, 0.0, "SPIN"))


def test_samples_follow_the_marginals():
    rng = np.random.default_rng(33)
    h, Jv, Jh, (hf, cs, ce, cw) = strip_problem(4, 25, rng, scale=0.6)
    plan = TreePlan(hf, cs, ce, cw, beta=1.0)
    _, p1, pp = plan.marginals()
    R = 40000
    seeds = np.arange(1, R + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    x = plan.sample(R, seeds)
    assert set(np.unique(x)) <= {-1, 1}
    f1 = (x == 1).mean(0)
    sig = np.sqrt(p1 * (1 - p1) / R) + 1e-9
    assert np.all(np.abs(f1 - p1) < 5 * sig)
    fu = np.stack([((x[:, u] == a) & (x[:, v] == b)).mean() for u, v in zip(plan.pairs_u, plan.pairs_v)
                   for a, b in ((-1, -1), (1, -1), (-1, 1), (1, 1))]).reshape(-1, 4)
    assert np.all(np.abs(fu - pp) < 5 * np.sqrt(pp * (1 - pp) / R) + 1e-9)
    np.testing.assert_array_equal(plan.sample(50, seeds[:50]), x[:50])   # same seeds, same samples
    plan.close()


# ---- the reference's tests (tests/test_tree_sampler.py) ----

BQMS = [
    BinaryQuadraticModel.from_ising({}, {("a", "b"): .69, ("b", "c"): 1, ("a", "c"): .5}),
    BinaryQuadraticModel.from_qubo({("a", "b"): .69, ("b", "c"): 1, ("a", "c"): .5}),
    BinaryQuadraticModel.from_qubo({("a", "b"): .69, ("b", "c"): 1}),
    BinaryQuadraticModel({"a": 6.0, 0: 1}, {("a", 0): -3, (0, "c"): 10}, 0, "BINARY"),
]


def _exact(bqm):
    vs = list(bqm.variables)
    low = -1 if bqm.vartype.name == "SPIN" else 0
    import itertools
    X = np.array(list(itertools.product([low, 1], repeat=len(vs))))
    return vs, X, bqm.energies((X, vs))


@pytest.mark.parametrize("k", range(len(BQMS)))
def test_reference_marginals(k):
    """tests/test_tree_sampler.py:120-241 (log Z, variable and interaction
    marginals analytic and empirical)."""
    bqm = BQMS[k]
    vs, X, E = _exact(bqm)
    for beta in (.5, 1., 1.5, 2):
        ss = TreeDecompositionSampler().sample(bqm, num_reads=10, beta=beta)
        lz = ss.info["log_partition_function"]
        assert lz == pytest.approx(np.log(np.sum(np.exp(-beta * E))))
        Z = np.exp(lz)
        for v, p in ss.info["variable_marginals"].items():
            assert p == pytest.approx(np.sum(np.exp(-beta * E[X[:, vs.index(v)] > 0])) / Z)
        for (u, v), combos in ss.info["interaction_marginals"].items():
            for (a, b), p in combos.items():
                sel = (X[:, vs.index(u)] == a) & (X[:, vs.index(v)] == b)
                assert p == pytest.approx(np.sum(np.exp(-beta * E[sel])) / Z)
        big = TreeDecompositionSampler().sample(bqm, num_reads=22500, beta=beta, seed=k)
        for v, p in ss.info["variable_marginals"].items():
            assert abs(p - np.mean(big.record.sample[:, vs.index(v)] > 0)) < 0.05
    ss = TreeDecompositionSampler().sample(bqm, num_reads=10)
    assert ss.info["log_partition_function"] == pytest.approx(np.log(np.sum(np.exp(-E))))


def test_interaction_keys_as_reference():
    """sample.hpp:113-144,205-225 and samplers.py:343-351: interaction marginals only
    for pairs with a nonzero bias, keyed (u, v) with u before v in variable
    order, in that order; cells (u, v) = (low, low), (1, low), (low, 1), (1, 1)."""
    bqm = BinaryQuadraticModel({"x": 0.5, "a": 0.0, "b": 0.2},
                               {("b", "x"): 1.0, ("a", "x"): 0.0, ("a", "b"): -0.7}, 0.25, "SPIN")
    ss = TreeDecompositionSampler().sample(bqm, num_reads=2, beta=0.9)
    assert list(ss.info["interaction_marginals"]) == [("x", "b"), ("a", "b")]
    vs, X, E = _exact(bqm)
    Z = np.sum(np.exp(-0.9 * E))
    assert ss.info["log_partition_function"] == pytest.approx(np.log(Z))
    for (u, v), combos in ss.info["interaction_marginals"].items():
        assert list(combos) == [(-1, -1), (1, -1), (-1, 1), (1, 1)]
        for (a, b), p in combos.items():
            sel = (X[:, vs.index(u)] == a) & (X[:, vs.index(v)] == b)
            assert p == pytest.approx(np.sum(np.exp(-0.9 * E[sel])) / Z)


def test_reference_empty_and_properties():
    """tests/test_tree_sampler.py:28-60."""
    s = TreeDecompositionSampler()
    assert s.properties == {"max_treewidth": 25}
    assert set(s.parameters) == {"num_reads", "elimination_order", "beta", "marginals", "seed"}
    ss = s.sample(BinaryQuadraticModel({}, {}, 0.0, "SPIN"), num_reads=3)
    assert len(ss) == 3 and ss.info["log_partition_function"] == 0.0
    ss = s.sample(BinaryQuadraticModel({"a": 1.0}, {}, 0.5, "SPIN"), num_reads=4, marginals=False, seed=1)
    assert "variable_marginals" not in ss.info and len(ss) == 4


def test_fixed_beta_moments_against_exact_marginals():
    """The moments that differentiable fixed-beta sampling returns as its
    gradient, on a 64-spin 4 x 16 strip (beyond exact enumeration), agree with
    the exact <s_i> and <s_u s_v> of the same model within Monte Carlo error."""
    import torch
    from misamplers.fixed_beta import FixedBetaSampler
    rng = np.random.default_rng(34)
    h, Jv, Jh, (hf, cs, ce, cw) = strip_problem(4, 16, rng, scale=0.4)
    plan = TreePlan(hf, cs, ce, cw, beta=1.0)
    _, p1, pp = plan.marginals()
    plan.close()
    R = 4096
    fb = FixedBetaSampler([64], [np.stack([cs, ce], 1)], num_reads=R, num_sweeps=300, beta=1.0)
    th = torch.tensor(hf, dtype=torch.float32, device="cuda:0", requires_grad=True)
    tj = torch.tensor(cw, dtype=torch.float32, device="cuda:0", requires_grad=True)
    L, _ = fb(th, tj, seed=9)
    L.sum().backward()
    fb.close()
    # the engine samples the float32 parameters; the exact marginals use the
    # float64 ones (a relative change of ~1e-7, far below the tolerance)
    m1 = 2 * p1 - 1
    pair_key = {(min(u, v), max(u, v)): k for k, (u, v) in enumerate(zip(plan.pairs_u, plan.pairs_v))}
    m2 = np.array([pp[pair_key[(min(u, v), max(u, v))]] @ np.array([1, -1, -1, 1]) for u, v in zip(cs, ce)])
    gh, gj = th.grad.cpu().numpy(), tj.grad.cpu().numpy()
    assert np.all(np.abs(gh - m1) < 5 * np.sqrt((1 - m1 ** 2) / R) + 1e-3)
    assert np.all(np.abs(gj - m2) < 5 * np.sqrt((1 - m2 ** 2) / R) + 1e-3)


# ---- k lowest-energy states (TreeDecompositionSolver) -----------------------

def _check_solutions(samples, energies, h, cs, ce, cw, low, ref, atol):
    # the k lowest energies, each the energy of its own distinct state
    from oracle.tree_exact import kbest_bruteforce  # noqa: F401  (same energy formula)
    assert len(energies) == len(ref)
    np.testing.assert_allclose(energies, ref, rtol=0, atol=atol)
    assert np.all(np.diff(energies) >= 0)
    X = samples.astype(np.float64)
    assert set(np.unique(samples)) <= {low, 1}
    E = X @ np.asarray(h, dtype=np.float64)
    for u, v, w in zip(cs, ce, cw):
        E = E + w * X[:, u] * X[:, v]
    np.testing.assert_allclose(E, energies, rtol=0, atol=atol)
    assert len({bytes(r) for r in samples}) == len(samples)


@pytest.mark.parametrize("low", [-1, 0])
@pytest.mark.parametrize("integer", [True, False])
def test_kbest_against_enumeration(low, integer):
    from oracle.tree_exact import kbest_bruteforce
    rng = np.random.default_rng(5 + low + 3 * integer)
    for n, p in ((1, 1.0), (2, 1.0), (7, 0.5), (12, 0.3), (14, 0.15)):
        pairs = [(u, v) for u in range(n) for v in range(u, n) if rng.random() < p]
        if pairs and pairs[-1][0] != pairs[-1][1]:
            pairs.append((pairs[-1][1], pairs[-1][0]))  # repeated, reversed
        cs = np.array([a for a, _ in pairs], np.int32)
        ce = np.array([b for _, b in pairs], np.int32)
        if integer:
            h, cw = rng.integers(-2, 3, n).astype(float), rng.integers(-2, 3, len(cs)).astype(float)
        else:
            h, cw = rng.normal(size=n), rng.normal(size=len(cs))
        plan = TreePlan(h, cs, ce, cw, low=low)
        for k in ((1, 5, 2 ** n, 2 ** n + 9) if n <= 7 else (1, 5, 64)):
            print(f"case kbest n={n} low={low} integer={integer} k={k} width={plan.treewidth}", flush=True)
            samples, energies = plan.solve(k)
            ref = kbest_bruteforce(h, cs, ce, cw, low, k)[0]
            _check_solutions(samples, energies, h, cs, ce, cw, low, ref, 0.0 if integer else 1e-9)
        plan.close()


@pytest.mark.parametrize("low", [-1, 0])
def test_kbest_beyond_enumeration(low):
    # a 4 x 15 strip and a random 3-regular graph on 40 variables against the
    # oracle's k-best elimination
    from oracle.tree_exact import kbest_eliminate
    rng = np.random.default_rng(17 - low)
    cases = [strip_problem(4, 15, rng, low=low)[3]]
    c2, e2 = random_regular(40, 3, rng)
    cases.append((rng.normal(size=40), c2, e2, rng.normal(size=len(c2))))
    for hh, a, b, ww in cases:
        plan = TreePlan(hh, a, b, ww, low=low)
        print(f"case kbest-large n={len(hh)} low={low} width={plan.treewidth}", flush=True)
        samples, energies = plan.solve(6)
        ref = kbest_eliminate(hh, a, b, ww, low, 6)
        _check_solutions(samples, energies, hh, a, b, ww, low, ref, 1e-9)
        plan.close()


def test_solver_reference_cases():
    # the reference's TreeDecompositionSolver tests (tests/test_tree_solver.py)
    from misamplers import TreeDecompositionSolver
    import itertools
    solver = TreeDecompositionSolver()
    assert solver.properties == {"max_treewidth": 25}
    assert set(solver.parameters) == {"num_reads", "elimination_order"}
    # empty
    bqm = BinaryQuadraticModel.empty("SPIN")
    ss = solver.sample(bqm, num_reads=10)
    assert len(ss) == 10
    # single variable, spin and binary
    ss = solver.sample(BinaryQuadraticModel.from_ising({"a": -1}, {}), num_reads=1)
    assert list(ss.samples()) == [{"a": 1}]
    ss = solver.sample(BinaryQuadraticModel.from_qubo({(0, 0): 1}), num_reads=1)
    assert ss.samples()[0] == {0: 0}
    # single interaction
    bqm = BinaryQuadraticModel.from_ising({"a": -1}, {("a", "b"): 1})
    ss = solver.sample(bqm, num_reads=1)
    assert list(ss.samples()) == [{"a": 1, "b": -1}]
    np.testing.assert_allclose(ss.record.energy, bqm.energies((ss.record.sample, ss.variables)))
    # chain of 100: two ground states, then one frustrated edge
    bqm = BinaryQuadraticModel.from_ising({}, {(v, v + 1): -1 for v in range(99)})
    ss = solver.sample(bqm, num_reads=3)
    assert len(ss) == 3
    np.testing.assert_allclose(ss.record.energy, bqm.energies((ss.record.sample, ss.variables)))
    g0, g1, ex = ss.samples()
    assert len(set(g0.values())) == 1 and len(set(g1.values())) == 1
    assert set(g0.values()) | set(g1.values()) == {-1, 1}
    assert sum(ex[u] != ex[v] for u, v in bqm.quadratic) == 1
    # clique of 20 (treewidth 19)
    bqm = BinaryQuadraticModel.from_qubo({pair: -1 for pair in itertools.combinations(range(20), 2)})
    ss = solver.sample(bqm, num_reads=2)
    ground, excited = ss.samples()
    assert set(ground.values()) == {1}
    assert sum(excited.values()) == len(excited) - 1
    # offset
    bqm = BinaryQuadraticModel.from_qubo({pair: -1 for pair in itertools.combinations(range(4), 2)})
    bqm.offset += -5
    ss = solver.sample(bqm, num_reads=2)
    np.testing.assert_allclose(ss.record.energy, bqm.energies((ss.record.sample, ss.variables)))
    # more reads than states: repeats counted
    bqm = BinaryQuadraticModel.from_qubo({(0, 0): -1, (0, 1): 1})
    ss = solver.sample(bqm, num_reads=101)
    assert len(ss) == 4 and int(np.sum(ss.record.num_occurrences)) == 101
