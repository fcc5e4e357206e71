"""GPU parity of the HIP sweep engine against the oracle and the reference's
golden vectors, through the C ABI.  Bit-exact for every case: integer modes
by construction; real mode uses the same float64 operation order and the
device exp, and the cases here are checked to agree exactly."""
import os

import numpy as np
import pytest

import oracle
from conftest import golden_cases, lattice3d, random_regular
from misamplers import _lib
from misamplers.sa.simulated_annealing import simulated_annealing

pytestmark = pytest.mark.gpu
CASES = golden_cases()
PROP = {0: "Gibbs", 1: "Metropolis"}


def run_gpu(h, cs, ce, cw, spb, betas, seeds, init, order, prop, mode=0, lanes=0, stats=None, interrupt=None):
    st = np.array(init, dtype=np.int8, order="C", copy=True)
    return simulated_annealing(st.shape[0], h, cs, ce, cw, spb, betas, 0, st, bool(order), PROP[prop],
                               interrupt, read_seeds=seeds, mode=mode, lanes_per_read=lanes, stats=stats)


def test_device_present():
    assert _lib.lib.misa_device_count() >= 1


@pytest.mark.parametrize("name", sorted(CASES))
def test_golden(name):
    c = CASES[name]
    st, en = run_gpu(c["h"], c["cs"], c["ce"], c["cw"], int(c["spb"]), c["betas"], c["seeds"], c["init"],
                     int(c["order"]), int(c["prop"]))
    np.testing.assert_array_equal(st, c["states"])
    np.testing.assert_array_equal(en, c["energies"])


def _random_problem(rng, n, p, integer, big=False, selfloops=False):
    pairs = [(u, v) for u in range(n) for v in range(u if selfloops else u + 1, n) if rng.random() < p]
    cs = np.array([a for a, _ in pairs], np.int32)
    ce = np.array([b for _, b in pairs], np.int32)
    if integer:
        hi = 12 if big else 1
        return rng.integers(-hi, hi + 1, n).astype(float), cs, ce, rng.integers(-hi, hi + 1, len(cs)).astype(float)
    return rng.uniform(-1, 1, n), cs, ce, rng.uniform(-1, 1, len(cs))


@pytest.mark.parametrize("integer,big,mode", [(True, False, 0), (True, True, 0), (False, False, 0),
                                              (True, False, 3), (True, False, 2)])
@pytest.mark.parametrize("order,prop", [(0, 1), (0, 0), (1, 1), (1, 0)])
def test_parity_random_problems(integer, big, mode, order, prop):
    rng = np.random.default_rng(hash((integer, big, mode, order, prop)) % 2**32)
    for n, p in ((37, 0.4), (130, 0.05)):
        h, cs, ce, cw = _random_problem(rng, n, p, integer, big)
        R = 70
        init = (2 * rng.integers(0, 2, (R, n)) - 1).astype(np.int8)
        seeds = oracle.read_seeds(int(rng.integers(2**31)), R)
        betas = np.geomspace(0.05, 4.0, 25)
        ref = oracle.sa_sample(h, cs, ce, cw, 2, betas, seeds, init, order, prop, nthreads=8,
                               return_counts=True)
        stats = {}
        st, en = run_gpu(h, cs, ce, cw, 2, betas, seeds, init, order, prop, mode=mode, stats=stats)
        np.testing.assert_array_equal(st, ref[0])
        np.testing.assert_array_equal(en, ref[1])
        assert stats["flips"] == int(ref[2].sum())
        assert stats["flip_degree"] == int(ref[3].sum())


@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16, 32, 64])
def test_lanes_per_read_do_not_change_results(lanes):
    rng = np.random.default_rng(lanes)
    h, cs, ce, cw = _random_problem(rng, 60, 0.3, integer=False)
    R = 50
    init = (2 * rng.integers(0, 2, (R, 60)) - 1).astype(np.int8)
    seeds = oracle.read_seeds(5, R)
    betas = np.geomspace(0.1, 3.0, 20)
    ref = oracle.sa_sample(h, cs, ce, cw, 1, betas, seeds, init, 0, 1, nthreads=8)
    st, en = run_gpu(h, cs, ce, cw, 1, betas, seeds, init, 0, 1, lanes=lanes)
    np.testing.assert_array_equal(st, ref[0])
    np.testing.assert_array_equal(en, ref[1])


@pytest.mark.parametrize("mode", [0, 3])
def test_selfloops_and_repeated_couplers(mode):
    rng = np.random.default_rng(7)
    h, cs, ce, cw = _random_problem(rng, 20, 0.5, integer=True, selfloops=True)
    cs, ce, cw = np.concatenate([cs, cs[:9]]), np.concatenate([ce, ce[:9]]), np.concatenate([cw, cw[:9]])
    R = 40
    init = (2 * rng.integers(0, 2, (R, 20)) - 1).astype(np.int8)
    seeds = oracle.read_seeds(8, R)
    for prop in (0, 1):
        ref = oracle.sa_sample(h, cs, ce, cw, 1, np.linspace(0.01, 3, 30), seeds, init, 0, prop)
        st, en = run_gpu(h, cs, ce, cw, 1, np.linspace(0.01, 3, 30), seeds, init, 0, prop, mode=mode, lanes=16)
        np.testing.assert_array_equal(st, ref[0])
        np.testing.assert_array_equal(en, ref[1])


def test_state_in_device_memory(monkeypatch):
    monkeypatch.setenv("MISA_STATE_IN_DEVICE_MEMORY", "1")
    rng = np.random.default_rng(9)
    for integer in (True, False):
        h, cs, ce, cw = _random_problem(rng, 90, 0.1, integer=integer)
        R = 33
        init = (2 * rng.integers(0, 2, (R, 90)) - 1).astype(np.int8)
        seeds = oracle.read_seeds(10, R)
        ref = oracle.sa_sample(h, cs, ce, cw, 1, np.geomspace(0.1, 3, 30), seeds, init, 0, 1)
        stats = {}
        st, en = run_gpu(h, cs, ce, cw, 1, np.geomspace(0.1, 3, 30), seeds, init, 0, 1, stats=stats)
        assert stats["state_in_lds"] is False
        np.testing.assert_array_equal(st, ref[0])
        np.testing.assert_array_equal(en, ref[1])


@pytest.mark.parametrize("order,prop", [(0, 1), (0, 0), (1, 1)])
def test_int_kernel_csr_in_device_memory(monkeypatch, order, prop):
    monkeypatch.setenv("MISA_CSR_IN_DEVICE_MEMORY", "1")
    rng = np.random.default_rng(19 + order + prop)
    h, cs, ce, cw = _random_problem(rng, 120, 0.08, integer=True)
    R = 45
    init = (2 * rng.integers(0, 2, (R, 120)) - 1).astype(np.int8)
    seeds = oracle.read_seeds(21, R)
    ref = oracle.sa_sample(h, cs, ce, cw, 1, np.geomspace(0.1, 3, 30), seeds, init, order, prop)
    st, en = run_gpu(h, cs, ce, cw, 1, np.geomspace(0.1, 3, 30), seeds, init, order, prop)
    np.testing.assert_array_equal(st, ref[0])
    np.testing.assert_array_equal(en, ref[1])


@pytest.mark.parametrize("device_csr", [True, False])
@pytest.mark.parametrize("order,prop", [(0, 1), (0, 0), (1, 1)])
def test_real_kernel_pipelined_rows(monkeypatch, device_csr, order, prop):
    # float64 kernel, adjacency in device memory or staged in LDS (sequential
    # order pipelines the rows, fields and spins of the next steps): tiny,
    # serial (self-loops, repeated couplers), sparse and long-row problems
    if device_csr:
        monkeypatch.setenv("MISA_CSR_IN_DEVICE_MEMORY", "1")
    monkeypatch.setenv("MISA_SEG", "0")  # this test is about the pipelined loop, not the run kernel
    monkeypatch.setenv("MISA_DENSE", "0")  # nor the dense kernel
    rng = np.random.default_rng(41 + order + 2 * prop)
    cases = [(1, 1.0, True, 0), (2, 1.0, False, 0), (3, 1.0, True, 0), (120, 0.08, False, 0),
             (40, 0.3, False, 0), (60, 0.9, True, 8)]
    for n, p, selfloops, lanes in cases:
        h, cs, ce, cw = _random_problem(rng, n, p, integer=False, selfloops=selfloops)
        if n == 40:  # repeated couplers, reversed
            cs, ce = np.concatenate([cs, ce[:5]]), np.concatenate([ce, cs[:5]])
            cw = np.concatenate([cw, rng.uniform(-1, 1, 5)])
        R = 40
        init = (2 * rng.integers(0, 2, (R, n)) - 1).astype(np.int8)
        seeds = oracle.read_seeds(int(rng.integers(2**31)), R)
        betas = np.geomspace(0.1, 3, 20)
        ref = oracle.sa_sample(h, cs, ce, cw, 2, betas, seeds, init, order, prop)
        stats = {}
        st, en = run_gpu(h, cs, ce, cw, 2, betas, seeds, init, order, prop, mode=3, lanes=lanes, stats=stats)
        assert stats["kernel"] == "sa_real_kernel" and stats["csr_in_lds"] is (not device_csr), (n, stats)
        np.testing.assert_array_equal(st, ref[0], err_msg=f"n={n}")
        np.testing.assert_array_equal(en, ref[1], err_msg=f"n={n}")


@pytest.mark.parametrize("device_csr", [False, True])
@pytest.mark.parametrize("prop", [1, 0])
@pytest.mark.parametrize("policy", [0, 1, 2, 3, 6, 13])
def test_run_kernel(monkeypatch, device_csr, prop, policy):
    # float64 run kernel (consecutive non-adjacent variables decided together),
    # forced on every problem: runs of one (dense, serial rows with self-loops
    # and repeated couplers), long runs (sparse), rows longer than a read's
    # lanes, ragged n; every launch shape the policies give (8..64 lanes per
    # read), with and without flip counters
    monkeypatch.setenv("MISA_SEG", "2")
    monkeypatch.setenv("MISA_SEG_POLICY", str(policy))
    if device_csr:
        monkeypatch.setenv("MISA_CSR_IN_DEVICE_MEMORY", "1")
    rng = np.random.default_rng(97 + prop + 2 * policy + 8 * device_csr)
    cases = [(1, 1.0, True, 7), (2, 1.0, False, 3), (5, 0.5, True, 40), (130, 0.03, False, 64), (300, 0.02, True, 200),
             (70, 0.7, False, 33), (40, 0.3, False, 130), (1000, 0.006, False, 1000), (1500, 0.01, False, 300)]
    for n, p, selfloops, R in cases:
        h, cs, ce, cw = _random_problem(rng, n, p, integer=False, selfloops=selfloops)
        if n == 40:  # repeated couplers, reversed
            cs, ce = np.concatenate([cs, ce[:5]]), np.concatenate([ce, cs[:5]])
            cw = np.concatenate([cw, rng.uniform(-1, 1, 5)])
        init = (2 * rng.integers(0, 2, (R, n)) - 1).astype(np.int8)
        seeds = oracle.read_seeds(int(rng.integers(2**31)), R)
        betas = np.geomspace(0.05, 4, 25)
        ref = oracle.sa_sample(h, cs, ce, cw, 2, betas, seeds, init, 0, prop, nthreads=8, return_counts=True)
        print("run kernel", n, R, flush=True)
        stats = {}
        st, en = run_gpu(h, cs, ce, cw, 2, betas, seeds, init, 0, prop, mode=3, stats=stats)
        assert stats["kernel"] == "sa_seg_kernel", (n, stats)
        if device_csr:
            assert stats["csr_in_lds"] is False
        np.testing.assert_array_equal(st, ref[0], err_msg=f"n={n}")
        np.testing.assert_array_equal(en, ref[1], err_msg=f"n={n}")
        assert stats["flips"] == int(ref[2].sum()) and stats["flip_degree"] == int(ref[3].sum()), n
        st2, en2 = run_gpu(h, cs, ce, cw, 2, betas, seeds, init, 0, prop, mode=3)  # without counters
        np.testing.assert_array_equal(st2, ref[0])
        np.testing.assert_array_equal(en2, ref[1])


def test_run_kernel_rows_beyond_a_wave(monkeypatch):
    # rows longer than 64 entries do not take the run kernel; the float64
    # kernel they fall back to gives the same samples
    monkeypatch.setenv("MISA_SEG", "2")
    rng = np.random.default_rng(5)
    h, cs, ce, cw = _random_problem(rng, 90, 0.95, integer=False)
    init = (2 * rng.integers(0, 2, (20, 90)) - 1).astype(np.int8)
    seeds = oracle.read_seeds(3, 20)
    betas = np.geomspace(0.1, 3, 15)
    stats = {}
    st, en = run_gpu(h, cs, ce, cw, 1, betas, seeds, init, 0, 1, mode=3, stats=stats)
    assert stats["kernel"] == "sa_real_kernel"
    ref = oracle.sa_sample(h, cs, ce, cw, 1, betas, seeds, init, 0, 1)
    np.testing.assert_array_equal(st, ref[0])
    np.testing.assert_array_equal(en, ref[1])


def test_long_rows_exceed_lanes():
    # rows longer than the lanes of a read group take the loop path
    rng = np.random.default_rng(23)
    h, cs, ce, cw = _random_problem(rng, 80, 0.9, integer=True)
    R = 30
    init = (2 * rng.integers(0, 2, (R, 80)) - 1).astype(np.int8)
    seeds = oracle.read_seeds(22, R)
    ref = oracle.sa_sample(h, cs, ce, cw, 1, np.geomspace(0.01, 1, 20), seeds, init, 0, 1)
    for lanes in (4, 16, 64):
        st, en = run_gpu(h, cs, ce, cw, 1, np.geomspace(0.01, 1, 20), seeds, init, 0, 1, lanes=lanes)
        np.testing.assert_array_equal(st, ref[0])
        np.testing.assert_array_equal(en, ref[1])


def test_zero_sweeps_identity_and_energy():
    rng = np.random.default_rng(11)
    h, cs, ce, cw = _random_problem(rng, 50, 0.2, integer=False)
    init = (2 * rng.integers(0, 2, (12, 50)) - 1).astype(np.int8)
    st, en = run_gpu(h, cs, ce, cw, 1, [], oracle.read_seeds(1, 12), init, 0, 1)
    np.testing.assert_array_equal(st, init)
    ref = oracle.sa_sample(h, cs, ce, cw, 1, [], oracle.read_seeds(1, 12), init)
    np.testing.assert_array_equal(en, ref[1])


@pytest.mark.parametrize("stop_after", [1, 5])
def test_interrupt_returns_prefix(stop_after):
    n = 5
    cs = [u for u in range(n) for v in range(u, n)]
    ce = [v for u in range(n) for v in range(u, n)]
    count = [0]

    def stop():
        count[0] += 1
        return count[0] >= stop_after

    init = np.ones((100, n), np.int8)
    st, en = simulated_annealing(100, [-1] * n, cs, ce, [-1] * len(cs), 1, np.linspace(0.01, 3, 100), 1,
                                 init, interrupt_function=stop)
    assert st.shape == (stop_after, n) and en.shape == (stop_after,)


def test_interrupt_exception_counts_as_stop():
    def boom():
        raise NotImplementedError

    st, en = simulated_annealing(10, [-1, -1], [0], [1], [-1.0], 1, [1.0], 1, np.ones((10, 2), np.int8),
                                 interrupt_function=boom)
    assert st.shape == (1, 2)


def test_ferromagnet_ground_state_with_selfloops():
    # tests/test_simulated_annealing.py:158-179
    n, R = 5, 100
    cs = [u for u in range(n) for v in range(u, n)]
    ce = [v for u in range(n) for v in range(u, n)]
    init = (2 * np.random.RandomState(1234).randint(2, size=(R, n)).astype(np.int8) - 1)
    st, en = simulated_annealing(R, [-1] * n, cs, ce, [-1] * len(cs), 1, np.linspace(0.01, 3, 1000), 1, init)
    assert [1] * n in st.tolist()
    assert abs(np.mean(en) - (-(n + 3) * n / 2)) <= 2


def test_lattice_l16_full_size_parity_subset():
    """The 3D +-J L=16 lattice (4096 spins) with the full 2000-sweep geometric
    schedule: a subset of reads bit-exact against the oracle; all 4096 reads
    checked for spin validity and energy consistency."""
    cs, ce = lattice3d(16)
    rng = np.random.default_rng(2024)
    cw = rng.choice([-1.0, 1.0], len(cs))
    h = np.zeros(4096)
    betas = np.geomspace(0.1, 5.0, 2000)
    R = 4096
    init = (2 * rng.integers(0, 2, (R, 4096)) - 1).astype(np.int8)
    seeds = oracle.read_seeds(77, R)
    stats = {}
    st, en = run_gpu(h, cs, ce, cw, 1, betas, seeds, init, 0, 1, stats=stats)
    assert stats["mode"] == _lib.MODE_INT8
    assert set(np.unique(st)) <= {-1, 1}
    s = st.astype(np.int64)
    e = (s[:, cs] * s[:, ce] * cw.astype(np.int64)).sum(axis=1)
    np.testing.assert_array_equal(en, e.astype(float))
    assert en.min() < -6000  # L=16 +-J ground state is about -1.78 per spin
    pick = np.linspace(0, R - 1, 32).astype(int)
    ref = oracle.sa_sample(h, cs, ce, cw, 1, betas, seeds[pick], init[pick], 0, 1, nthreads=16)
    np.testing.assert_array_equal(st[pick], ref[0])
    np.testing.assert_array_equal(en[pick], ref[1])


@pytest.mark.parametrize("integer", [True, False])
def test_counting_does_not_change_samples(integer):
    # without a statistics dict the kernels keep no flip counters; the
    # samples must be the same either way, and equal to the oracle's
    cs, ce = lattice3d(6)
    rng = np.random.default_rng(77 + integer)
    n = 216
    cw = rng.choice([-1.0, 1.0], len(cs)) if integer else rng.normal(0, 1, len(cs))
    h = np.zeros(n) if integer else rng.normal(0, 0.2, n)
    R = 48
    init = (2 * rng.integers(0, 2, (R, n)) - 1).astype(np.int8)
    seeds = oracle.read_seeds(5, R)
    betas = np.geomspace(0.1, 4.0, 40)
    stats = {}
    st1, en1 = run_gpu(h, cs, ce, cw, 1, betas, seeds, init, 0, 1, stats=stats)
    st2, en2 = run_gpu(h, cs, ce, cw, 1, betas, seeds, init, 0, 1, stats=None)
    assert stats["flips"] > 0 and stats["flip_degree"] >= stats["flips"]
    np.testing.assert_array_equal(st1, st2)
    np.testing.assert_array_equal(en1, en2)
    ref = oracle.sa_sample(h, cs, ce, cw, 1, betas, seeds, init, 0, 1, nthreads=8, return_counts=True)
    np.testing.assert_array_equal(st2, ref[0])
    np.testing.assert_array_equal(en2, ref[1])
    assert stats["flips"] == int(np.sum(ref[2]))
    assert stats["flip_degree"] == int(np.sum(ref[3]))


@pytest.mark.parametrize("prop", [1, 0])
def test_int_kernel_regular_graphs(monkeypatch, prop):
    # every row of the same length takes the regular-graph loop of the integer
    # kernel (row starts computed, couplings to v+1 and v+2 from the row
    # info): random 3- and 5-regular graphs, and one with a self-loop on
    # every variable (rows of 3 + 2 entries); MISA_CHAIN=0 keeps these sparse
    # graphs off the chain-run kernel
    monkeypatch.setenv("MISA_CHAIN", "0")
    rng = np.random.default_rng(61 + prop)
    for n, d, loops in ((120, 3, False), (90, 5, False), (64, 3, True)):
        cs, ce = random_regular(n, d, rng)
        if loops:
            cs = np.concatenate([cs, np.arange(n)])
            ce = np.concatenate([ce, np.arange(n)])
        cw = rng.integers(-2, 3, len(cs)).astype(float)
        h = rng.integers(-1, 2, n).astype(float)
        R = 40
        init = (2 * rng.integers(0, 2, (R, n)) - 1).astype(np.int8)
        seeds = oracle.read_seeds(int(rng.integers(2**31)), R)
        betas = np.geomspace(0.1, 3.0, 30)
        stats = {}
        st, en = run_gpu(h, cs, ce, cw, 1, betas, seeds, init, 0, prop, stats=stats)
        assert stats["kernel"] == "sa_int_kernel", stats
        ref = oracle.sa_sample(h, cs, ce, cw, 1, betas, seeds, init, 0, prop, nthreads=8)
        np.testing.assert_array_equal(st, ref[0], err_msg=f"n={n} d={d}")
        np.testing.assert_array_equal(en, ref[1], err_msg=f"n={n} d={d}")


def test_lds_bandwidth_microbenchmark():
    """The LDS peak measured by misa_lds_bandwidth is positive, above what 128 B
    per clock and CU at a 1.5 GHz clock would give (the sweep kernels' LDS
    traffic is priced against this figure), and below 256 CUs x 256 B per
    clock at 2.4 GHz (with 5 % slack)."""
    from misamplers import _lib
    gbs, ms = _lib.lds_bandwidth(4096)
    print(f"LDS read bandwidth {gbs:.0f} GB/s in {ms:.2f} ms", flush=True)
    assert 256 * 128 * 1.5 < gbs < 256 * 256 * 2.4 * 1.05 and ms > 0.0
