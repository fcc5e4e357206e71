"""Chain-run kernel (int8 words, sequential order): bit-exact against the
oracle and against the pipelined integer kernel (MISA_CHAIN=0) on 3D and 2D
lattices, sparse irregular graphs with self-loops and repeated couplers,
path graphs (every consecutive pair coupled), problems shorter than one run,
and read counts that leave block slots empty; each of its three forms
(generator stepped in each run, output rings with per-lane comparisons,
output rings with one ballot per distinct threshold), with run-major slot
records (rows of at most 6 entries) and with per-variable records."""
import os

import numpy as np
import pytest

import oracle
from conftest import lattice3d
from misamplers.sa.simulated_annealing import simulated_annealing

pytestmark = pytest.mark.gpu
PROP = {0: "Gibbs", 1: "Metropolis"}


def run(h, cs, ce, cw, spb, betas, seeds, init, prop, chain, ring=True, few=True, slots=True):
    """chain: MISA_CHAIN=1 (take the chain-run kernel whenever it can run) or
    0; ring: the variant with per-read output rings (MISA_CHAIN_RING unset)
    or the one stepping the generator inside each run (MISA_CHAIN_RING=0);
    few: the ring variant may take the few-threshold form (MISA_CHAIN_FEW
    unset) or keeps the per-lane comparisons (MISA_CHAIN_FEW=0); slots: rows
    of at most 6 entries read run-major slot records (MISA_CHAIN_SLOTS unset)
    or per-variable records and run bounds (MISA_CHAIN_SLOTS=0)."""
    env = {"MISA_CHAIN": "1" if chain else "0", "MISA_CHAIN_RING": None if ring else "0",
           "MISA_CHAIN_FEW": None if few else "0", "MISA_CHAIN_SLOTS": None if slots else "0"}
    old = {k: os.environ.get(k) for k in env}
    try:
        for k, v in env.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
        st = np.array(init, dtype=np.int8, copy=True)
        stats = {}
        out = simulated_annealing(st.shape[0], h, cs, ce, cw, spb, betas, 0, st, False, PROP[prop],
                                  read_seeds=seeds, stats=stats)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return out, stats


def lattice2d(L):
    cs, ce = [], []
    for x in range(L):
        for y in range(L):
            v = x * L + y
            cs += [v, v]
            ce += [((x + 1) % L) * L + y, x * L + (y + 1) % L]
    return np.array(cs, np.int32), np.array(ce, np.int32)


def sparse_irregular(n, rng, maxdeg=7):
    """Random couplers with every degree <= maxdeg, plus a self-loop and a
    repeated coupler (each counts toward the degree of its ends)."""
    deg = np.zeros(n, int)
    cs, ce = [], []
    for _ in range(3 * n):
        u, v = (int(x) for x in rng.integers(0, n, 2))
        if u == v or deg[u] >= maxdeg - 2 or deg[v] >= maxdeg - 2:
            continue
        cs.append(u)
        ce.append(v)
        deg[u] += 1
        deg[v] += 1
    cs += [3, cs[0]]
    ce += [3, ce[0]]
    return np.array(cs, np.int32), np.array(ce, np.int32)


def check(h, cs, ce, cw, R, betas, prop, seed, spb=1, oracle_check=True, ring=True, few=True, form=None,
          slots=True):
    """form: the chain form the launch must report ("stepping", "ring" or
    "few"); by default "stepping" without rings and either ring form with
    them."""
    rng = np.random.default_rng(seed)
    n = len(h)
    init = (2 * rng.integers(0, 2, (R, n)) - 1).astype(np.int8)
    seeds = oracle.read_seeds(seed, R)
    (st, en), stats = run(h, cs, ce, cw, spb, betas, seeds, init, prop, chain=True, ring=ring, few=few, slots=slots)
    assert stats["kernel"] == "sa_chain_kernel", stats
    assert stats["lanes_per_read"] == 16
    if form is not None:
        assert stats["chain_form"] == form, stats
    else:
        assert stats["chain_form"] in (("ring", "few") if ring else ("stepping",)), stats
    if not few:
        assert stats["chain_form"] != "few"
    (st0, en0), stats0 = run(h, cs, ce, cw, spb, betas, seeds, init, prop, chain=False)
    assert stats0["kernel"] != "sa_chain_kernel"
    np.testing.assert_array_equal(st, st0)
    np.testing.assert_array_equal(en, en0)
    assert (stats["flips"], stats["flip_degree"]) == (stats0["flips"], stats0["flip_degree"])
    if oracle_check:
        ref = oracle.sa_sample(h, cs, ce, cw, spb, betas, seeds, init, 0, prop, nthreads=8, return_counts=True)
        np.testing.assert_array_equal(st, ref[0])
        np.testing.assert_array_equal(en, ref[1])
        assert stats["flips"] == int(ref[2].sum())
        assert stats["flip_degree"] == int(ref[3].sum())


@pytest.mark.parametrize("ring", [True, False])
@pytest.mark.parametrize("prop", [1, 0])
@pytest.mark.parametrize("R", [40, 16, 1])
def test_lattice3d(prop, R, ring):
    rng = np.random.default_rng(R + 10 * prop)
    cs, ce = lattice3d(6)
    n = 216
    h = rng.integers(-1, 2, n).astype(float)
    cw = rng.choice([-1.0, 1.0], len(cs))
    check(h, cs, ce, cw, R, np.geomspace(0.1, 5.0, 40), prop, seed=100 + R, spb=2, ring=ring)


def test_ring_refills_over_a_long_anneal():
    """Enough sweeps that every read refills its output ring many times
    (several thousand draws per read) with reads drawing at different rates."""
    rng = np.random.default_rng(9)
    cs, ce = lattice3d(4)
    h = rng.integers(-1, 2, 64).astype(float)
    cw = rng.choice([-1.0, 1.0], len(cs))
    check(h, cs, ce, cw, 48, np.geomspace(0.05, 3.0, 150), 1, seed=44, spb=3)


@pytest.mark.parametrize("prop", [1, 0])
def test_lattice2d_and_weights(prop):
    rng = np.random.default_rng(3 + prop)
    cs, ce = lattice2d(9)
    h = rng.integers(-3, 4, 81).astype(float)
    cw = rng.integers(-5, 6, len(cs)).astype(float)
    check(h, cs, ce, cw, 37, np.geomspace(0.05, 3.0, 30), prop, seed=7)


@pytest.mark.parametrize("prop", [1, 0])
def test_sparse_irregular_with_selfloop_and_repeat(prop):
    rng = np.random.default_rng(11 + prop)
    n = 300
    cs, ce = sparse_irregular(n, rng)
    h = rng.integers(-2, 3, n).astype(float)
    cw = rng.integers(-3, 4, len(cs)).astype(float)
    check(h, cs, ce, cw, 50, np.geomspace(0.1, 4.0, 25), prop, seed=21)


@pytest.mark.parametrize("ring", [True, False])
@pytest.mark.parametrize("n", [2, 3, 14, 15, 16, 17, 31, 46])
def test_paths_and_short_problems(n, ring):
    """A path couples every consecutive pair (runs of 15, each member patched
    by its predecessor); n <= 15 is a single run; a closing coupler (n-1, 0)
    never joins a run."""
    rng = np.random.default_rng(n)
    cs = np.arange(n - 1, dtype=np.int32)
    ce = cs + 1
    if n > 2:
        cs = np.append(cs, n - 1).astype(np.int32)
        ce = np.append(ce, 0).astype(np.int32)
    h = rng.integers(-1, 2, n).astype(float)
    cw = rng.choice([-2.0, -1.0, 1.0, 2.0], len(cs))
    check(h, cs, ce, cw, 24, np.geomspace(0.2, 6.0, 30), 1, seed=n, ring=ring)


@pytest.mark.parametrize("slots", [True, False])
@pytest.mark.parametrize("form", ["few", "ring", "stepping"])
def test_full_lattice_against_pipelined_kernel(form, slots):
    """The L=16 lattice (4096 spins) of the throughput workload, 64 reads,
    a short schedule: identical to the pipelined integer kernel in each form
    (the few-threshold form is the default here: h = 0 and J = +-1 leave the
    draws dE in {4, 8, 12})."""
    rng = np.random.default_rng(20241203)
    cs, ce = lattice3d(16)
    cw = rng.choice([-1.0, 1.0], len(cs))
    h = np.zeros(4096)
    check(h, cs, ce, cw, 64, np.geomspace(0.1, 5.0, 12), 1, seed=5, oracle_check=False,
          ring=form != "stepping", few=form == "few", form=form, slots=slots)


def ring_problem(n):
    """A closed path: every variable has the same two neighbours' worth of
    field values."""
    cs = np.arange(n, dtype=np.int32)
    return cs, ((cs + 1) % n).astype(np.int32)


# (problem, proposal, form the default launch takes): a beta's drawing words
# have at most three distinct thresholds for the few-threshold form
FEW_CASES = [
    ("lattice3d", 1, "few"),    # h = 0, J = +-1, Metropolis: dE in {4, 8, 12}
    ("lattice3d", 0, "ring"),   # Gibbs draws for dE in {-12, ..., 12}: 7 thresholds
    ("lattice2d", 1, "few"),    # dE in {4, 8}
    ("lattice2d", 0, "ring"),   # dE in {-8, ..., 8}: 5 thresholds
    ("ring", 0, "few"),         # closed path, J = +-1: Gibbs dE in {-4, 0, 4}
    ("ring", 1, "few"),         # Metropolis dE = 4 only
    ("lattice3d_h", 1, "ring"),  # h in {-1, 0, 1}: dE in {2, ..., 14}
]


@pytest.mark.parametrize("case,prop,form", FEW_CASES)
def test_few_threshold_form(case, prop, form):
    """The default launch takes the few-threshold form exactly when every
    beta has at most three distinct thresholds over the words a variable can
    hold; both ring forms equal the pipelined kernel and the oracle, with the
    reads drawing from rings that refill many times."""
    rng = np.random.default_rng(len(case) + 7 * prop)
    if case.startswith("lattice3d"):
        cs, ce = lattice3d(6)
        n = 216
    elif case == "lattice2d":
        cs, ce = lattice2d(12)
        n = 144
    else:
        n = 61
        cs, ce = ring_problem(n)
    cw = rng.choice([-1.0, 1.0], len(cs))
    h = rng.integers(-1, 2, n).astype(float) if case.endswith("_h") else np.zeros(n)
    betas = np.geomspace(0.05, 6.0, 60)
    check(h, cs, ce, cw, 40, betas, prop, seed=31 + prop, spb=3, form=form)
    check(h, cs, ce, cw, 40, betas, prop, seed=31 + prop, spb=3, form=form, slots=False)
    if form == "few":
        check(h, cs, ce, cw, 40, betas, prop, seed=31 + prop, spb=3, few=False, form="ring")


def test_dense_rows_take_the_pipelined_kernel():
    """Rows of more than 7 entries are not taken by the chain-run kernel."""
    rng = np.random.default_rng(2)
    n = 60
    cs = np.array([u for u in range(n) for v in range(u + 1, min(n, u + 5))], np.int32)
    ce = np.array([v for u in range(n) for v in range(u + 1, min(n, u + 5))], np.int32)
    h = np.zeros(n)
    cw = rng.choice([-1.0, 1.0], len(cs))
    init = np.ones((8, n), np.int8)
    (_, _), stats = run(h, cs, ce, cw, 1, np.geomspace(0.1, 3, 5), oracle.read_seeds(1, 8), init, 1, chain=True)
    assert stats["kernel"] == "sa_int_kernel"


def test_short_runs_take_the_pipelined_kernel_by_default():
    """With MISA_CHAIN unset, a problem whose runs average fewer than 9
    variables (here every v is coupled to v+2, so runs are 2 long) keeps the
    pipelined integer kernel; the L=16 lattice (runs of 15 and 16 along its
    z-columns) takes the chain-run kernel."""
    rng = np.random.default_rng(4)
    n = 64
    cs = np.arange(n - 2, dtype=np.int32)
    ce = cs + 2
    cw = rng.choice([-1.0, 1.0], len(cs))
    init = np.ones((8, n), np.int8)
    old = os.environ.pop("MISA_CHAIN", None)
    try:
        st = init.copy()
        stats = {}
        simulated_annealing(8, np.zeros(n), cs, ce, cw, 1, np.geomspace(0.1, 3, 5), 0, st, False, "Metropolis",
                            read_seeds=oracle.read_seeds(1, 8), stats=stats)
        assert stats["kernel"] == "sa_int_kernel"
        cs3, ce3 = lattice3d(16)
        st = np.ones((8, 4096), np.int8)
        simulated_annealing(8, np.zeros(4096), cs3, ce3, rng.choice([-1.0, 1.0], len(cs3)), 1,
                            np.geomspace(0.1, 3, 5), 0, st, False, "Metropolis", read_seeds=oracle.read_seeds(1, 8),
                            stats=stats)
        assert stats["kernel"] == "sa_chain_kernel"
    finally:
        if old is not None:
            os.environ["MISA_CHAIN"] = old


def test_large_lattice_takes_the_stepping_variant():
    """An 80 x 80 periodic lattice (6400 spins): 16 bytes of words per spin
    leave no room for the output rings, so the default launch takes the
    variant that steps the generator in every lane; identical to the
    pipelined integer kernel."""
    rng = np.random.default_rng(80)
    cs, ce = lattice2d(80)
    n = 6400
    h = rng.integers(-1, 2, n).astype(float)
    cw = rng.choice([-1.0, 1.0], len(cs))
    R = 40
    init = (2 * rng.integers(0, 2, (R, n)) - 1).astype(np.int8)
    seeds = oracle.read_seeds(3, R)
    betas = np.geomspace(0.1, 4.0, 6)
    old = os.environ.pop("MISA_CHAIN", None)
    try:
        st = init.copy()
        stats = {}
        en = simulated_annealing(R, h, cs, ce, cw, 1, betas, 0, st, False, "Metropolis", read_seeds=seeds,
                                 stats=stats)[1]
        assert stats["kernel"] == "sa_chain_kernel"
    finally:
        if old is not None:
            os.environ["MISA_CHAIN"] = old
    (st0, en0), stats0 = run(h, cs, ce, cw, 1, betas, seeds, init, 1, chain=False)
    np.testing.assert_array_equal(st, st0)
    np.testing.assert_array_equal(en, en0)
    ref = oracle.sa_sample(h, cs, ce, cw, 1, betas, seeds[:4], init[:4], 0, 1, nthreads=4)
    np.testing.assert_array_equal(st[:4], ref[0])
    np.testing.assert_array_equal(en[:4], ref[1])
