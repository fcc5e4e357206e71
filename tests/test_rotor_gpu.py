"""GPU tests of the rotor-model engine through the C ABI: bit-exact angles,
energies and snapshots against the oracle (itself bit-exact against the
reference's cpu_rotormc.cpp), plus the reference's sampler tests
(tests/test_rotor_annealing_sampler.py) restated for the mirror."""
import json

import numpy as np
import pytest

import oracle
from conftest import random_regular
from misamplers import BinaryQuadraticModel, RotorModelAnnealingSampler, SampleSet
from misamplers.sqa.rotormc_annealing import simulated_annealing as rotor_core

pytestmark = pytest.mark.gpu

PROPOSALS = ["GibbsNonErgodic", "MetropolisNonErgodic", "MetropolisUniform", "MetropolisTF"]


def _problem(rng, n, p, selfloops=False, repeats=False):
    pairs = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < p]
    if selfloops:
        pairs += [(0, 0), (n - 1, n - 1)]
    if repeats and pairs:
        pairs += pairs[:3]
    cs = np.array([a for a, _ in pairs], np.int32)
    ce = np.array([b for _, b in pairs], np.int32)
    return rng.normal(size=n), cs, ce, rng.normal(size=len(cs)), rng.uniform(0.2, 1.5, n)


def _check(h, cs, ce, cw, tf, spb, Hp, Hd, st0, randomize, proposal, interval=0, seeds=None):
    R, n = st0.shape
    print(f"case n={n} m={len(cs)} R={R} {proposal} random={randomize} interval={interval}", flush=True)
    if seeds is None:
        seeds = np.arange(1, R + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    st = np.array(st0, dtype=np.uint8, order="C", copy=True)
    info = {}
    g = rotor_core(R, h, cs, ce, cw, tf, spb, Hp, Hd, 0, st, randomize, proposal, interval,
                   read_seeds=seeds, stats=info)
    o = oracle.rotor_sample(h, cs, ce, cw, tf, spb, Hp, Hd, seeds, st0, randomize, proposal, interval, nthreads=8)
    np.testing.assert_array_equal(g[0], o[0])
    np.testing.assert_array_equal(g[1], o[1])
    if interval:
        np.testing.assert_array_equal(g[2][:, :o[2].shape[1]], o[2])
    return g, info


@pytest.mark.parametrize("proposal", PROPOSALS)
@pytest.mark.parametrize("randomize", [False, True])
def test_parity_with_serial_rows_and_snapshots(proposal, randomize):
    rng = np.random.default_rng(21)
    h, cs, ce, cw, tf = _problem(rng, 90, 0.08, selfloops=True, repeats=True)
    Hp = np.geomspace(0.1, 4.0, 30)
    Hd = np.linspace(3.0, 0.05, 30)
    st0 = rng.integers(0, 256, (150, 90), dtype=np.uint8)
    _, info = _check(h, cs, ce, cw, tf, 2, Hp, Hd, st0, randomize, proposal, interval=4)
    assert info["csr_in_lds"] == 1


@pytest.mark.parametrize("proposal", PROPOSALS)
def test_parity_ragged_sizes(proposal):
    rng = np.random.default_rng(22)
    for n in (1, 2, 63, 65, 130):
        h, cs, ce, cw, tf = _problem(rng, n, 0.1, selfloops=True, repeats=True)
        st0 = rng.integers(0, 256, (37, n), dtype=np.uint8)
        for randomize in (False, True):
            _check(h, cs, ce, cw, tf, 3, [0.5, 1.0, 2.0], [1.0, 0.4, 0.0], st0, randomize, proposal, interval=1)


@pytest.mark.parametrize("proposal", ["MetropolisNonErgodic", "MetropolisUniform"])
def test_parity_adjacency_in_device_memory(proposal):
    """A problem whose adjacency does not fit LDS (more than 65535 variables
    is not needed: 16000 couplers of f64 + u16 already exceed it)."""
    rng = np.random.default_rng(23)
    n = 3000
    cs, ce = random_regular(n, 12, rng)
    st0 = rng.integers(0, 256, (70, n), dtype=np.uint8)
    _, info = _check(rng.normal(size=n), cs, ce, rng.normal(size=len(cs)), np.ones(n), 1, [0.3, 1.5], [1.0, 0.2],
                     st0, False, proposal)
    assert info["csr_in_lds"] == 0


@pytest.mark.parametrize("proposal", ["MetropolisUniform", "MetropolisTF"])
@pytest.mark.parametrize("kernel", ["wave", "lane_groups"])
def test_parity_sequential_angle_kernels(proposal, kernel, monkeypatch):
    """Sequential angle proposals run one wave per read when every row has at
    most 64 entries, else (or with MISA_ROTOR_WAVE=0) on lane groups; both
    bit-exact.  The problem has a row of exactly 64 entries, an isolated
    variable, self-loops and repeated couplers; a row of 65 entries sends the
    same problem to the lane groups."""
    if kernel == "lane_groups":
        monkeypatch.setenv("MISA_ROTOR_WAVE", "0")
    rng = np.random.default_rng(26)
    n = 120
    pairs = [(u, v) for u in range(1, n) for v in range(u + 1, n) if rng.random() < 0.05]
    pairs = [(a, b) for a, b in pairs if a != 7 and b != 7 and a != 11 and b != 11]
    pairs += [(7, u) for u in range(12, 76)]            # row 7: 64 entries (variable 11 stays isolated)
    pairs += [(3, 3), (5, 9), (5, 9), (9, 5)]
    cs = np.array([a for a, _ in pairs], np.int32)
    ce = np.array([b for _, b in pairs], np.int32)
    cw = rng.normal(size=len(cs))
    h, tf = rng.normal(size=n), rng.uniform(0.2, 1.5, n)
    st0 = rng.integers(0, 256, (70, n), dtype=np.uint8)
    Hp, Hd = np.geomspace(0.2, 3.0, 12), np.linspace(2.5, 0.1, 12)
    _, info = _check(h, cs, ce, cw, tf, 2, Hp, Hd, st0, False, proposal, interval=3)
    assert (info["lanes_per_read"] == 64) == (kernel == "wave")
    cs2, ce2 = np.append(cs, 7), np.append(ce, 80)      # row 7: 65 entries
    _, info = _check(h, cs2, ce2, np.append(cw, 0.5), tf, 1, Hp[:4], Hd[:4], st0, False, proposal)
    assert info["lanes_per_read"] < 64


def _path(n):
    return np.arange(n - 1, dtype=np.int32), np.arange(1, n, dtype=np.int32)


RUN_CASES = {"regular4": 30, "regular5": 31, "regular8": 36, "rows16": 32, "serial_rows": 33, "path": 34,
             "regular6": 35}


@pytest.mark.parametrize("proposal", ["MetropolisUniform", "MetropolisTF"])
@pytest.mark.parametrize("case", list(RUN_CASES))
def test_parity_run_kernel(case, proposal, monkeypatch):
    """Angle proposals in sequential order on rows of at most 16 entries run
    in runs of pairwise non-adjacent variables (rotor_run_kernel; the
    transverse-field window draws one output per proposal while Hd >= Hp and
    two after, and the schedule crosses over): bit-exact
    against the oracle and against the wave kernel (MISA_ROTOR_RUNS=0) on
    rows of exactly 4, 6 and 8 entries (the row-length-exact forms), rows of
    5 and of 14 entries (read past their end), rows with self-loops and
    repeated couplers, a path (every run one variable) and the path-integral
    workload's 6-regular problem; snapshots included."""
    rng = np.random.default_rng(RUN_CASES[case])
    if case in ("regular4", "regular5", "regular8"):
        n = 300
        cs, ce = random_regular(n, int(case[-1]), rng)
    elif case == "rows16":
        n = 257
        cs, ce = random_regular(n, 14, rng)
    elif case == "serial_rows":
        n = 150
        _, cs, ce, _, _ = _problem(rng, n, 0.04, selfloops=True, repeats=True)
    elif case == "path":
        n = 97
        cs, ce = _path(n)
    else:
        n = 2048
        cs, ce = random_regular(n, 6, rng)
    h, cw, tf = rng.uniform(-0.5, 0.5, n), rng.choice([-1.0, 1.0], len(cs)), rng.uniform(0.3, 1.5, n)
    R = 256 if case == "regular6" else 90
    S = 40 if case == "regular6" else 60
    st0 = rng.integers(0, 256, (R, n), dtype=np.uint8)
    Hp, Hd = np.linspace(0.05, 6.0, S), np.linspace(6.0, 0.05, S)
    g, info = _check(h, cs, ce, cw, tf, 1, Hp, Hd, st0, False, proposal, interval=7)
    assert info["kernel"] == "rotor_run_kernel", info
    monkeypatch.setenv("MISA_ROTOR_RUNS", "0")
    st = np.array(st0, copy=True)
    info2 = {}
    seeds = np.arange(1, R + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    w = rotor_core(R, h, cs, ce, cw, tf, 1, Hp, Hd, 0, st, False, proposal, 7, read_seeds=seeds,
                   stats=info2)
    assert info2["kernel"] == "rotor_wave_kernel", info2
    np.testing.assert_array_equal(g[0], w[0])
    np.testing.assert_array_equal(g[1], w[1])
    np.testing.assert_array_equal(g[2], w[2])
    assert info["accepted"] == info2["accepted"]


def test_parity_tf_window_corner_cases():
    rng = np.random.default_rng(24)
    h, cs, ce, cw, tf = _problem(rng, 40, 0.15)
    tf[:5] = 0.0
    Hp = np.array([1.0, 1.0, 1.0, 2.0, 0.0, 0.0, 3.0, 1.0, 1.0, 5.0, 1.0, 1.0])
    Hd = np.array([0.0, 1 / 300, 0.3, 1.9, 1.0, 0.0, -0.7, -3.0, 0.999, 4.99, -1e12, 1e-300])
    st0 = rng.integers(0, 256, (64, 40), dtype=np.uint8)
    for randomize in (False, True):
        _check(h, cs, ce, cw, tf, 3, Hp, Hd, st0, randomize, "MetropolisTF", interval=5)


def test_parity_lattice_like_many_reads():
    rng = np.random.default_rng(25)
    n = 512
    cs, ce = random_regular(n, 6, rng)
    cw = rng.choice([-1.0, 1.0], len(cs))
    st0 = rng.integers(0, 256, (600, n), dtype=np.uint8)
    for proposal in ("GibbsNonErgodic", "MetropolisTF"):
        _check(rng.uniform(-0.5, 0.5, n), cs, ce, cw, np.ones(n), 1, np.linspace(0.1, 3, 10),
               np.linspace(3, 0, 10), st0, False, proposal)


def test_core_interrupt_returns_a_prefix():
    rng = np.random.default_rng(26)
    h, cs, ce, cw, tf = _problem(rng, 20, 0.2)
    st = rng.integers(0, 256, (5000, 20), dtype=np.uint8)
    calls = []

    def stop():
        calls.append(1)
        return len(calls) >= 3

    out = rotor_core(5000, h, cs, ce, cw, tf, 1, [1.0], [0.5], 7, st, False, "MetropolisUniform", 0, stop)
    assert len(out[0]) == 3 and len(out[1]) == 3


def test_core_argument_errors():
    st = np.zeros((2, 2), np.uint8)
    with pytest.raises(ValueError):
        rotor_core(2, [0.0, 0.0], [0], [1], [1.0], [1.0, 1.0], 1, [1.0], [0.0], 1, st, False, "Sideways", 0)
    with pytest.raises(ValueError):
        rotor_core(2, [0.0, 0.0], [0], [1], [1.0], [1.0, 1.0], 1, [1.0], [0.0], 1, st.astype(np.int8), False,
                   "MetropolisUniform", 0)
    with pytest.raises(ValueError):
        rotor_core(2, [0.0, 0.0], [0], [2], [1.0], [1.0, 1.0], 1, [1.0], [0.0], 1, st, False, "MetropolisUniform", 0)
    with pytest.raises(ValueError):
        rotor_core(2, [0.0, 0.0], [0, 1], [1], [1.0], [1.0, 1.0], 1, [1.0], [0.0], 1, st, False,
                   "MetropolisUniform", 0)


# ---- the reference's sampler tests (tests/test_rotor_annealing_sampler.py) ----

SEED = 2023


def _binomial(p, n, k=3):
    mu, sig = n * p, np.sqrt(n * p * (1 - p))
    return mu - k * sig, mu + k * sig


def test_timing_and_basic_response():
    """:27-55, :170-199."""
    ss = RotorModelAnnealingSampler().sample_ising({"a": 0.5, "b": -0.5}, {("a", "b"): -1}, num_reads=10, seed=1)
    assert len(ss) == 10 and set(ss.info["timing"]) == {"preprocessing_ns", "sampling_ns", "postprocessing_ns"}
    assert all(v > 0 for v in ss.info["timing"].values())
    qs = RotorModelAnnealingSampler().sample_qubo({(0, 0): 1, (0, 1): -1}, num_reads=4, seed=1)
    assert set(np.unique(qs.record.sample)) <= {0, 1}


def test_seed_and_empty_problem():
    """:221-276."""
    bqm = BinaryQuadraticModel({i: 0.1 * i for i in range(10)}, {(i, i + 1): -1.0 for i in range(9)}, 0.0, "SPIN")
    a = RotorModelAnnealingSampler().sample(bqm, num_reads=20, seed=88, project_states=(False, False))
    b = RotorModelAnnealingSampler().sample(bqm, num_reads=20, seed=88, project_states=(False, False))
    np.testing.assert_array_equal(a.info["rotor_states"], b.info["rotor_states"])
    np.testing.assert_array_equal(a.record.sample, b.record.sample)
    c = RotorModelAnnealingSampler().sample(bqm, num_reads=20, seed=89, project_states=(False, False))
    assert not np.array_equal(a.info["rotor_states"], c.info["rotor_states"])
    empty = RotorModelAnnealingSampler().sample(BinaryQuadraticModel({}, {}, 0.0, "SPIN"), num_reads=3)
    assert len(empty) == 3 and empty.record.sample.shape == (3, 0)
    with pytest.raises(ValueError):
        RotorModelAnnealingSampler().sample(bqm, seed=-1)
    with pytest.raises(TypeError):
        RotorModelAnnealingSampler().sample(bqm, seed=1.5)


def test_zero_sweeps_is_identity_on_angles():
    """:426-451: num_sweeps=0 returns the initial angles and their energies."""
    rng = np.random.default_rng(3)
    bqm = BinaryQuadraticModel({i: float(rng.normal()) for i in range(6)},
                               {(i, i + 1): float(rng.normal()) for i in range(5)}, 0.25, "SPIN")
    init = rng.integers(0, 256, (4, 6), dtype=np.uint8)
    ss = RotorModelAnnealingSampler().sample(bqm, num_sweeps=0, initial_states=init, num_reads=4,
                                             project_states=(False, False), seed=5)
    np.testing.assert_array_equal(ss.info["rotor_states"], init)
    c, _ = oracle.rotor_tables()
    ang = ss.info["rotor_states"]
    for r in range(4):
        e = sum(c[ang[r, v]] * bqm.linear[v] for v in range(6))
        e += sum(c[ang[r, u]] * w * c[ang[r, v]] for (u, v), w in bqm.quadratic.items())
        assert ss.record.energy[r] == pytest.approx(e + 0.25, abs=1e-12)


def test_metropolis_ergodicity_breaking():
    """:687-712."""
    n = 100
    init = 1 - 2 * np.random.default_rng(0).integers(2, size=n)
    bqm = BinaryQuadraticModel({i: 0 for i in range(n)}, {}, 0.0, "SPIN")
    initial_states = SampleSet.from_samples_bqm({i: int(init[i]) for i in range(n)}, bqm)
    for num_sweeps in (1, 2):
        ss = RotorModelAnnealingSampler().sample(bqm, initial_states=initial_states, num_reads=1,
                                                 num_sweeps_per_beta=num_sweeps, beta_schedule_type="custom",
                                                 Hp_field=[1], proposal_acceptance_criteria="MetropolisNonErgodic")
        assert np.all(ss.record.sample == (-1) ** num_sweeps * init)


def test_central_limits_random_updates():
    """:714-802."""
    n = 10000
    bqm = BinaryQuadraticModel({i: 0 for i in range(n)}, {}, 0.0, "SPIN")
    init = SampleSet.from_samples_bqm({i: 1 for i in range(n)}, bqm)
    common = dict(initial_states=init, num_reads=1, seed=SEED, num_sweeps=1, beta_range=[0.1, 1])
    lo, hi = _binomial(0.5, n)
    s = RotorModelAnnealingSampler().sample(bqm, proposal_acceptance_criteria="GibbsNonErgodic", **common)
    assert lo < np.sum(s.record.sample == 1) < hi
    lo, hi = _binomial(np.cosh(1) * np.exp(-1), n)
    s = RotorModelAnnealingSampler().sample(bqm, randomize_order=True,
                                            proposal_acceptance_criteria="MetropolisNonErgodic", **common)
    assert lo < np.sum(s.record.sample == 1) < hi
    lo, hi = _binomial(1 / np.e + (1 - 1 / np.e) * 0.5, n)
    s = RotorModelAnnealingSampler().sample(bqm, randomize_order=True,
                                            proposal_acceptance_criteria="GibbsNonErgodic", **common)
    assert lo < np.sum(s.record.sample == 1) < hi
    bqm1 = BinaryQuadraticModel({i: 1 for i in range(n)}, {}, 0.0, "SPIN")
    p = np.exp(-1.5) / (2 * np.cosh(1.5))
    lo, hi = _binomial(p, n)
    s = RotorModelAnnealingSampler().sample(bqm1, initial_states=init, num_reads=1, seed=SEED,
                                            proposal_acceptance_criteria="GibbsNonErgodic",
                                            beta_schedule_type="custom", Hp_field=[1, 1.5], num_sweeps_per_beta=1)
    assert lo < np.sum(s.record.sample == 1) < hi


def test_greedy_limit_independent_spins():
    """:804-861."""
    n = 10000
    h = np.random.default_rng(1).normal(size=n)
    bqm = BinaryQuadraticModel({i: float(h[i]) for i in range(n)}, {}, 0.0, "SPIN")
    ground = np.array([-int(np.sign(h[i])) for i in range(n)])
    gs = SampleSet.from_samples_bqm({i: int(ground[i]) for i in range(n)}, bqm)
    sky = SampleSet.from_samples_bqm({i: -int(ground[i]) for i in range(n)}, bqm)
    lo, hi = _binomial(1 - np.exp(-1), n)
    for randomize in (False, True):
        for prop in ("MetropolisNonErgodic", "GibbsNonErgodic"):
            kw = dict(num_reads=1, num_sweeps=1, randomize_order=randomize, proposal_acceptance_criteria=prop,
                      beta_schedule_type="custom", Hp_field=[float("inf")], seed=SEED)
            s = RotorModelAnnealingSampler().sample(bqm, initial_states=gs, **kw)
            assert np.all(s.record.sample == ground)
            s = RotorModelAnnealingSampler().sample(bqm, initial_states=sky, **kw)
            if randomize:
                assert lo < np.sum(s.record.sample == ground) < hi
            else:
                assert np.all(s.record.sample == ground)


def test_independent_angles_and_tf_window():
    """:863-970."""
    prng = np.random.default_rng(SEED)
    R, n = 4, 64
    bqm = BinaryQuadraticModel({i: 0 for i in range(n)}, {}, 0.0, "SPIN")
    uid = prng.integers(256, size=(R, n), dtype=np.uint8)
    lo, hi = _binomial(1 - np.exp(-(R * n) / 256), R * n)
    for randomize in (False, True):
        for prop in PROPOSALS:
            s = RotorModelAnnealingSampler().sample(bqm, initial_states=uid.copy(), randomize_order=randomize,
                                                    proposal_acceptance_criteria=prop, beta_range=[0.1, 1],
                                                    project_states=(False, False), seed=SEED)
            assert lo < len(np.unique(s.info["rotor_states"])) < hi
    delta = 64 * np.ones((R, n), dtype=np.uint8)
    s = RotorModelAnnealingSampler().sample(bqm, initial_states=delta, proposal_acceptance_criteria="MetropolisUniform",
                                            beta_range=[0.1, 1], project_states=(False, False), seed=SEED,
                                            num_sweeps=1)
    assert lo < len(np.unique(s.info["rotor_states"])) < hi
    root, R = 32, 1000
    delta = root * np.ones((R, n), dtype=np.uint8)
    for gamma in (0, 1, 4, 16, 123, 255, 256):
        hd = (gamma + prng.random()) / 256
        s = RotorModelAnnealingSampler().sample(bqm, initial_states=delta.copy(),
                                                proposal_acceptance_criteria="MetropolisTF",
                                                beta_schedule_type="custom", Hp_field=[1], Hd_field=[hd],
                                                project_states=(False, False), trans_fields=[0] * n,
                                                num_sweeps_per_beta=1, num_sweeps=1, seed=SEED + gamma)
        us = set(np.unique(s.info["rotor_states"]).tolist())
        osi = 0 if hd == 0 else (1 if hd < 1 / 256 else int(hd * 256 + 1) // 2)
        if hd < 1:
            assert us <= {(root + i) % 256 for i in range(-osi, osi + 1)}
        # circular deviation from the start; with no fields every proposal is
        # accepted, and the window's boundary probability p is chosen so that
        # the mean |deviation| is (Hd / Hp) * 64 (cpu_rotormc.cpp:218-230)
        d = np.abs((np.asarray(s.info["rotor_states"], dtype=int) - root + 128) % 256 - 128)
        if hd < 1:
            assert abs(d.mean() - hd * 64) < 5 * d.std() / np.sqrt(d.size) + 1e-12, (gamma, d.mean(), hd * 64)


def test_equilibrated_statistics_classical_limit():
    """:972-1010."""
    n, R = 4, 1000
    prng = np.random.default_rng(SEED)
    v = prng.random(size=n)
    bqm = BinaryQuadraticModel({i: -v[i] for i in range(n)}, {}, 0.0, "SPIN")
    s = RotorModelAnnealingSampler().sample(bqm, num_reads=R, seed=SEED, proposal_acceptance_criteria="GibbsNonErgodic",
                                            beta_schedule_type="custom", Hp_field=[1, 2], num_sweeps_per_beta=1)
    k = np.sum(s.record.sample == 1, axis=0)
    for i in range(n):
        p = np.exp(-2 * bqm.linear[i]) / (2 * np.cosh(2 * bqm.linear[i]))
        lo, hi = _binomial(p, R)
        assert lo < k[i] < hi


def test_equilibrated_statistics_quantum_limit():
    """:1012-1054."""
    n, R, S = 4, 1000, 1000
    bqm = BinaryQuadraticModel({i: 0 for i in range(n)}, {}, 0.0, "SPIN")
    sin_states = np.abs(np.sin(np.arange(256) / 256 * 2 * np.pi))
    P = np.exp(1.0 * (sin_states - 1))
    P /= P.sum()
    EX = np.sum(P * sin_states)
    sig = np.sqrt(np.sum(P * sin_states ** 2) - EX ** 2)
    s = RotorModelAnnealingSampler().sample(bqm, num_reads=R, seed=SEED, beta_schedule_type="custom",
                                            Hp_field=[2] * S, Hd_field=[1] * S, num_sweeps_per_beta=1,
                                            proposal_acceptance_criteria="MetropolisUniform",
                                            project_states=(True, False))
    stat = np.mean(np.abs(np.sin(2 * np.pi * s.info["rotor_states"] / 256)))
    assert abs(stat - EX) < 3 * sig / np.sqrt(R * n)


def test_tf_equals_uniform_when_window_covers_the_circle():
    """:1056-1074, as intended: with Hd / Hp >= 1 the transverse-field
    proposal is the uniform one, draw for draw, so the runs are identical."""
    bqm = BinaryQuadraticModel({0: 0.1}, {(0, 1): 0.1}, 0.0, "SPIN")
    out = {}
    for prop in ("MetropolisUniform", "MetropolisTF"):
        s = RotorModelAnnealingSampler().sample(bqm, num_reads=10, seed=SEED, beta_schedule_type="custom",
                                                Hp_field=[1, 1], Hd_field=[1.5, 1], num_sweeps_per_beta=1,
                                                proposal_acceptance_criteria=prop, project_states=(True, False))
        out[prop] = s.info["rotor_states"]
    np.testing.assert_array_equal(out["MetropolisUniform"], out["MetropolisTF"])


def test_equilibrated_statistics_pair():
    """:1076-1160."""
    J, h, tf = 0.8, [-0.1, 0.2], [1.1, 0.5]
    theta = np.arange(256) / 128 * np.pi
    c, s_ = np.cos(theta), np.abs(np.sin(theta))
    Hd, Hp = [1.1] * 100, [1] * 100
    EnQ = -tf[0] * s_[:, None] - tf[1] * s_[None, :]
    EnC = h[0] * c[:, None] + h[1] * c[None, :] + J * c[:, None] * c[None, :]
    w = np.exp(-Hd[-1] * (EnQ - EnQ.max()) - Hp[-1] * (EnC - EnC.max()))
    Z = w.sum()
    stats = {"mag1": c[:, None] + 0 * c[None, :], "mag2": 0 * c[:, None] + c[None, :],
             "corr": c[:, None] * c[None, :], "align1": s_[:, None] + 0 * s_[None, :],
             "align2": 0 * s_[:, None] + s_[None, :]}
    R = 1000
    for prop in ("MetropolisUniform", "MetropolisTF"):
        out = RotorModelAnnealingSampler().sample_ising({0: h[0], 1: h[1]}, {(0, 1): J}, num_reads=R, seed=SEED,
                                                        beta_schedule_type="custom", Hp_field=Hp, trans_fields=tf,
                                                        Hd_field=Hd, project_states=(False, False),
                                                        proposal_acceptance_criteria=prop, num_sweeps_per_beta=1)
        a = out.info["rotor_states"]
        for key, st in stats.items():
            mu = np.sum(w * st) / Z
            sig = np.sqrt(np.sum(w * st * st) / Z - mu * mu)
            val = np.mean(st[a[:, 0], a[:, 1]])
            assert abs(val - mu) < 3 * sig / np.sqrt(R), (prop, key)


def test_serializable_and_statistics_shape():
    """:1162-1207."""
    for ser in (False, True):
        r = RotorModelAnnealingSampler().sample_ising({0: 1}, {}, beta_range=np.array([0.1, 2]),
                                                      make_info_json_serializable=ser)
        try:
            json.dumps(r.info)
            ok = True
        except TypeError:
            ok = False
        assert ok == ser
    for ssi in (1, 2):
        r = RotorModelAnnealingSampler().sample_ising({i: 1 for i in range(10)}, {}, num_sweeps=4, num_reads=3,
                                                      schedule_sample_interval=ssi)
        assert r.info["statistics"].shape == (3, 4 // ssi, 10)
        assert r.info["statistics"].dtype == np.uint8


def test_sampler_argument_errors():
    bqm = BinaryQuadraticModel({0: 1.0, 1: 0.0}, {(0, 1): 1.0}, 0.0, "SPIN")
    smp = RotorModelAnnealingSampler()
    with pytest.raises(ValueError):
        smp.sample(bqm, trans_fields=[1.0])
    with pytest.raises(ValueError):
        smp.sample(bqm, beta_schedule_type="custom")
    with pytest.raises(ValueError):
        smp.sample(bqm, beta_schedule_type="custom", Hp_field=[1, 2], Hd_field=[1])
    with pytest.raises(ValueError):
        smp.sample(bqm, schedule_sample_interval=1000, num_sweeps=10)
    with pytest.raises(ValueError):
        smp.sample(bqm, initial_states=np.zeros((3, 5), np.uint8))
    with pytest.raises(ValueError):
        smp.sample(bqm, proposal_acceptance_criteria="Sideways")
