"""Dense float64 kernel (n <= 128, sequential order, no repeated coupler or
self-loop): bit-exact against the oracle and against the LDS float64 kernels
(MISA_DENSE=0) on dense and sparse random problems, at the register
boundaries (n = 2, 64, 65, 128), with integer-valued problems run in float64
mode, with read counts that leave waves of the last block empty and that fill
16 waves per block, with flip counters, and across several betas per sweep
count; problems with a repeated coupler or a self-loop keep the other
kernels."""
import os

import numpy as np
import pytest

import oracle
from misamplers.sa.simulated_annealing import simulated_annealing

pytestmark = pytest.mark.gpu
PROP = {0: "Gibbs", 1: "Metropolis"}
REAL = 3


def gnp(n, p, rng, integer=False):
    iu, ju = np.triu_indices(n, 1)
    keep = rng.random(len(iu)) < p
    cs, ce = iu[keep].astype(np.int32), ju[keep].astype(np.int32)
    if integer:
        h = rng.integers(-3, 4, n).astype(float)
        cw = rng.integers(-3, 4, len(cs)).astype(float)
    else:
        h = rng.uniform(-1, 1, n)
        cw = rng.uniform(-1, 1, len(cs))
    return h, cs, ce, cw


def run(h, cs, ce, cw, spb, betas, seeds, init, prop, dense):
    old = os.environ.get("MISA_DENSE")
    os.environ["MISA_DENSE"] = "1" if dense else "0"
    try:
        st = np.array(init, dtype=np.int8, copy=True)
        stats = {}
        out = simulated_annealing(st.shape[0], h, cs, ce, cw, spb, betas, 0, st, False, PROP[prop],
                                  read_seeds=seeds, mode=REAL, stats=stats)
    finally:
        if old is None:
            os.environ.pop("MISA_DENSE", None)
        else:
            os.environ["MISA_DENSE"] = old
    return out, stats


def check(h, cs, ce, cw, R, betas, prop, seed, spb=1, oracle_reads=None, other=True):
    rng = np.random.default_rng(seed)
    n = len(h)
    init = (2 * rng.integers(0, 2, (R, n)) - 1).astype(np.int8)
    seeds = oracle.read_seeds(seed, R)
    (st, en), stats = run(h, cs, ce, cw, spb, betas, seeds, init, prop, dense=True)
    assert stats["kernel"] == "sa_dense_kernel", stats
    k = R if oracle_reads is None else oracle_reads
    ref = oracle.sa_sample(h, cs, ce, cw, spb, betas, seeds[:k], init[:k], 0, prop, nthreads=8,
                           return_counts=True)
    np.testing.assert_array_equal(st[:k], ref[0])
    np.testing.assert_array_equal(en[:k], ref[1])
    if oracle_reads is None:
        assert stats["flips"] == int(ref[2].sum())
        assert stats["flip_degree"] == int(ref[3].sum())
    if other:
        (st0, en0), stats0 = run(h, cs, ce, cw, spb, betas, seeds, init, prop, dense=False)
        assert stats0["kernel"] != "sa_dense_kernel"
        np.testing.assert_array_equal(st, st0)
        np.testing.assert_array_equal(en, en0)
    return st, en


@pytest.mark.parametrize("prop", [1, 0])
@pytest.mark.parametrize("n", [2, 17, 64, 65, 100, 128])
def test_dense_random_problems(n, prop):
    rng = np.random.default_rng(n + 1000 * prop)
    h, cs, ce, cw = gnp(n, 0.5, rng)
    check(h, cs, ce, cw, 45, np.geomspace(0.05, 8.0, 25), prop, seed=n, spb=2)


@pytest.mark.parametrize("prop", [1, 0])
def test_sparse_and_integer_valued(prop):
    """A sparse problem (every v has few couplers, many zero entries in each
    dense row) and an integer-valued one in float64 mode."""
    rng = np.random.default_rng(7 + prop)
    h, cs, ce, cw = gnp(90, 0.04, rng)
    check(h, cs, ce, cw, 33, np.geomspace(0.1, 5.0, 30), prop, seed=3)
    h, cs, ce, cw = gnp(120, 0.3, rng, integer=True)
    check(h, cs, ce, cw, 33, np.geomspace(0.1, 5.0, 30), prop, seed=4)


def test_gnp_workload_all_reads():
    """G(100, 0.5) with uniform [-1, 1] biases, 1000 reads, the sampler's
    default beta range at 100 sweeps: every read against the oracle."""
    from misamplers.sa.sampler import _default_ising_beta_range
    rng = np.random.default_rng(2024)
    h, cs, ce, cw = gnp(100, 0.5, rng)
    lin = {i: h[i] for i in range(100)}
    quad = {(int(u), int(v)): w for u, v, w in zip(cs, ce, cw)}
    b0, b1 = _default_ising_beta_range(lin, quad)
    check(h, cs, ce, cw, 1000, np.geomspace(b0, b1, 100), 1, seed=11, other=False)


@pytest.mark.parametrize("R", [1, 257, 4100])
def test_read_counts(R):
    """One read; a last block with empty waves; 16 waves per block."""
    rng = np.random.default_rng(R)
    h, cs, ce, cw = gnp(70, 0.5, rng)
    check(h, cs, ce, cw, R, np.geomspace(0.2, 4.0, 8), 1, seed=R, oracle_reads=min(R, 300), other=R < 1000)


def test_repeats_and_selfloops_keep_other_kernels():
    rng = np.random.default_rng(5)
    h, cs, ce, cw = gnp(60, 0.5, rng)
    for extra_cs, extra_ce in (([cs[0]], [ce[0]]), ([3], [3])):
        cs2 = np.concatenate([cs, np.array(extra_cs, np.int32)])
        ce2 = np.concatenate([ce, np.array(extra_ce, np.int32)])
        cw2 = np.concatenate([cw, [0.5]])
        init = np.ones((8, 60), np.int8)
        seeds = oracle.read_seeds(1, 8)
        (st, en), stats = run(h, cs2, ce2, cw2, 1, np.geomspace(0.1, 3, 10), seeds, init, 1, dense=True)
        assert stats["kernel"] != "sa_dense_kernel"
        ref = oracle.sa_sample(h, cs2, ce2, cw2, 1, np.geomspace(0.1, 3, 10), seeds, init, 0, 1)
        np.testing.assert_array_equal(st, ref[0])
        np.testing.assert_array_equal(en, ref[1])


def test_default_takes_dense_for_dense_rows_only():
    """With MISA_DENSE unset the kernel is taken when rows hold at least 16
    entries on average."""
    rng = np.random.default_rng(8)
    old = os.environ.pop("MISA_DENSE", None)
    try:
        for p, want in ((0.5, True), (0.05, False)):
            h, cs, ce, cw = gnp(100, p, rng)
            st = np.ones((8, 100), np.int8)
            stats = {}
            simulated_annealing(8, h, cs, ce, cw, 1, np.geomspace(0.1, 3, 5), 0, st, False, "Metropolis",
                                read_seeds=oracle.read_seeds(1, 8), stats=stats)
            assert (stats["kernel"] == "sa_dense_kernel") == want, (p, stats)
    finally:
        if old is not None:
            os.environ["MISA_DENSE"] = old
