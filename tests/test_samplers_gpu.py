"""The drop-in samplers on the GPU: ``SimulatedAnnealingSampler`` and
``PathIntegralAnnealingSampler`` called exactly as a user of the reference
calls them.

* Drop-in parity: ``.sample(bqm, initial_states=..., seed=...)`` equals the
  oracle read by read on the arrays the sampler hands to the engine (BINARY
  and SPIN models, default and custom schedules, offsets).
* The reference's sampler-level tests, restated:
  tests/test_simulated_annealing_sampler.py:528-656 (job shop with a linear
  schedule; 3D +-J L=12 lattice below -3000) and :659-845 (TestCoreSpinUpdate:
  ergodicity breaking, central limits with seed 2023 at 3 sigma, greedy
  limit); tests/test_path_annealing_sampler.py:377-512 (job shop; L=12 lattice
  below -2900).
"""
import itertools

import numpy as np
import pytest

import oracle
from misamplers import (BinaryQuadraticModel, PathIntegralAnnealingSampler, SampleSet,
                        SimulatedAnnealingSampler)
from misamplers import _lib
from misamplers.sa.sampler import default_beta_range
from misamplers.sqa.engine import colouring

pytestmark = pytest.mark.gpu

SEED = 2023  # tests/test_simulated_annealing_sampler.py:663


def _random_bqm(rng, n, p, vartype, integer):
    pairs = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < p]
    if integer:
        lin = {i: float(rng.integers(-2, 3)) for i in range(n)}
        quad = {e: float(rng.choice([-1, 1])) for e in pairs}
    else:
        lin = {i: rng.uniform(-1, 1) for i in range(n)}
        quad = {e: rng.uniform(-1, 1) for e in pairs}
    return BinaryQuadraticModel(lin, quad, 0.75, vartype)


# ---------------------------------------------------------------------------
# drop-in parity against the oracle
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("vartype,integer,schedule", [("SPIN", True, "geometric"), ("SPIN", False, "linear"),
                                                      ("BINARY", False, "geometric"), ("SPIN", False, "custom")])
def test_sa_sample_equals_oracle_read_by_read(vartype, integer, schedule):
    rng = np.random.default_rng(11 + integer)
    n, R, seed = 90, 40, 77
    bqm = _random_bqm(rng, n, 0.2, vartype, integer)
    # initial states are given in the spin form the sampler converts the model to
    init = rng.choice([-1, 1], size=(R, n)).astype(np.int8)
    kw = dict(num_reads=R, seed=seed, initial_states=(init, list(range(n))))
    if schedule == "custom":
        kw.update(beta_schedule_type="custom", beta_schedule=np.geomspace(0.2, 4.0, 150))
    else:
        kw.update(beta_schedule_type=schedule, num_sweeps=150)
    ss = SimulatedAnnealingSampler().sample(bqm, **kw)

    # what the engine receives: the spin form in variable order, per-read seeds
    spin = bqm.change_vartype("SPIN", inplace=False)
    h, (cs, ce, cw), off = spin.to_numpy_vectors(variable_order=list(range(n)))
    if schedule == "custom":
        betas = kw["beta_schedule"]
    else:
        lo, hi = default_beta_range(bqm)
        betas = np.linspace(lo, hi, 150) if schedule == "linear" else np.geomspace(lo, hi, 150)
    ref_st, ref_en = oracle.sa_sample(h, cs, ce, cw, 1, betas, _lib.read_seeds(seed, R), init)
    rec = ss.record
    got = rec.sample if vartype == "SPIN" else (2 * rec.sample.astype(np.int16) - 1).astype(np.int8)
    # the sampler keeps read order in its record
    np.testing.assert_array_equal(got, ref_st)
    np.testing.assert_allclose(rec.energy, ref_en + off, rtol=0, atol=1e-12)
    assert set(ss.info) >= {"beta_range", "beta_schedule_type", "timing"}


def test_pimc_sample_equals_oracle_read_by_read():
    rng = np.random.default_rng(5)
    n, R, seed, P = 60, 24, 91, 32
    bqm = _random_bqm(rng, n, 0.1, "SPIN", False)
    init = rng.choice([-1, 1], size=(R, n)).astype(np.int8)
    hp, hd = np.linspace(0.1, 3.0, 40), np.linspace(3.0, 0.05, 40)
    ss = PathIntegralAnnealingSampler().sample(bqm, num_reads=R, seed=seed, initial_states=(init, list(range(n))),
                                               beta_schedule_type="custom", Hp_field=hp, Hd_field=hd)
    h, (cs, ce, cw), off = bqm.to_numpy_vectors(variable_order=list(range(n)))
    W0 = np.where(init > 0, np.uint32(0xFFFFFFFF), np.uint32(0)).astype(np.uint32)
    Wo, _ = oracle.pimc_sample(h, cs, ce, cw, 1, hp, hd, _lib.read_seeds(seed, R), W0, num_slices=P,
                               colour=colouring(h, cs, ce, cw)[1])
    want = np.where(Wo & 1, 1, -1).astype(np.int8)
    np.testing.assert_array_equal(ss.record.sample, want)
    want_en = np.array([oracle.state_energy(s, h, cs, ce, cw) for s in want]) + off
    np.testing.assert_array_equal(ss.record.energy, want_en)


# ---------------------------------------------------------------------------
# tests/test_simulated_annealing_sampler.py, restated
# ---------------------------------------------------------------------------

JSS_LINEAR = {"b_0,0": -2.0, "b_0,1": -2.0, "b_0,2": -2.0, "b_0,3": -2.0, "b_1,0": 0.125, "b_1,1": -1.5,
              "b_1,2": 0.0, "g_0,0": -1.875, "g_0,1": -1.5, "g_0,2": 0.0, "o_0,0": -2.0, "o_0,1": -2.0,
              "o_0,2": -2.0, "o_1,0": 0.03125, "o_1,1": -1.875, "o_1,2": -1.5, "o_1,3": 0.0}
JSS_QUADRATIC = {("b_0,0", "g_0,0"): 4, ("b_0,1", "b_0,0"): 4.0, ("b_0,1", "g_0,0"): 2, ("b_0,2", "b_0,0"): 4.0,
                 ("b_0,2", "b_0,1"): 4.0, ("b_0,2", "b_1,2"): 2, ("b_0,2", "g_0,1"): 2, ("b_0,2", "g_0,2"): 4,
                 ("b_0,3", "b_0,0"): 4.0, ("b_0,3", "b_0,1"): 4.0, ("b_0,3", "b_0,2"): 4.0, ("b_0,3", "b_1,2"): 2,
                 ("b_0,3", "g_0,2"): 2, ("b_1,1", "b_0,1"): 2, ("b_1,1", "b_0,2"): 2, ("b_1,1", "b_0,3"): 2,
                 ("b_1,1", "b_1,2"): 4.0, ("g_0,1", "b_0,1"): 4, ("g_0,1", "g_0,0"): 4.0, ("g_0,2", "g_0,0"): 4.0,
                 ("g_0,2", "g_0,1"): 4.0, ("o_0,0", "o_1,1"): 2, ("o_0,1", "o_0,0"): 4.0, ("o_0,1", "o_1,1"): 2,
                 ("o_0,1", "o_1,2"): 2, ("o_0,2", "o_0,0"): 4.0, ("o_0,2", "o_0,1"): 4.0, ("o_0,2", "o_1,1"): 2,
                 ("o_1,2", "o_0,2"): 2, ("o_1,2", "o_1,1"): 4.0, ("o_1,3", "o_0,2"): 2, ("o_1,3", "o_1,1"): 4.0,
                 ("o_1,3", "o_1,2"): 4.0}
JSS_OPTIMUM = {"b_0,0": 1, "b_0,1": 0, "b_0,2": 0, "b_0,3": 0, "b_1,0": 0, "b_1,1": 1, "b_1,2": 0, "g_0,0": 0,
               "g_0,1": 1, "g_0,2": 0, "o_0,0": 1, "o_0,1": 0, "o_0,2": 0, "o_1,0": 0, "o_1,1": 0, "o_1,2": 1,
               "o_1,3": 0}


def _jss_bqm():
    return BinaryQuadraticModel(JSS_LINEAR, JSS_QUADRATIC, 9.0, "BINARY")


def _cubic_lattice_J(N=12):
    """tests/test_simulated_annealing_sampler.py:632-643 (the same edges and draws)."""
    def edges():
        for x, y, z in itertools.product(range(N), repeat=3):
            u = x, y, z
            yield u, ((x + 1) % N, y, z)
            yield u, (x, (y + 1) % N, z)
            yield u, (x, y, (z + 1) % N)
    np_rand = np.random.RandomState(128)
    return {e: np_rand.choice((-1, 1)) for e in edges()}


def test_sa_job_shop_scheduling_with_linear():
    bqm = _jss_bqm()
    optimal_energy = bqm.energy((np.array([[JSS_OPTIMUM[v] for v in bqm.variables]]), bqm.variables))
    assert optimal_energy == 0.5
    response = SimulatedAnnealingSampler().sample(bqm, beta_schedule_type="linear", num_reads=10)
    _, response_energy, _ = next(response.data())
    assert response_energy < optimal_energy + 0.1


def test_sa_cubic_lattice_with_geometric():
    response = SimulatedAnnealingSampler().sample_ising({}, _cubic_lattice_J(12), beta_schedule_type="geometric",
                                                        num_reads=10)
    _, response_energy, _ = next(response.data())
    assert response_energy < -3000, f"response_energy, {response_energy}, exceeds threshold"


def _interval(p, num_samples, k=3):
    mu = num_samples * p
    sig = np.sqrt(num_samples * p * (1 - p))
    return mu - k * sig, mu + k * sig


def test_sa_metropolis_ergodicity_breaking():
    num_vars = 100
    init_vector = 1 - 2 * np.random.randint(2, size=num_vars)
    bqm = BinaryQuadraticModel.from_ising({i: 0 for i in range(num_vars)}, {})
    initial_states = SampleSet.from_samples_bqm({i: int(init_vector[i]) for i in range(num_vars)}, bqm)
    for num_sweeps in range(3):
        response = SimulatedAnnealingSampler().sample(bqm, initial_states=initial_states, num_reads=1,
                                                      num_sweeps=num_sweeps, beta_range=[0.1, 1])
        assert np.all(response.record.sample == (-1) ** num_sweeps * init_vector)


def test_sa_central_limits_random_updates():
    num_vars = 10000
    bqm = BinaryQuadraticModel.from_ising({i: 0 for i in range(num_vars)}, {})
    initial_states = SampleSet.from_samples_bqm({i: 1 for i in range(num_vars)}, bqm)
    sampler = SimulatedAnnealingSampler()
    common = dict(initial_states=initial_states, num_reads=1, seed=SEED, num_sweeps=1, beta_range=[0.1, 1])
    # Gibbs, sequential order: every spin 50:50
    lo, hi = _interval(0.5, num_vars)
    stat = np.sum(sampler.sample(bqm, proposal_acceptance_criteria="Gibbs", **common).record.sample == 1)
    assert lo < stat < hi
    # Metropolis, random order: flipped iff picked an odd number of times
    lo, hi = _interval(np.cosh(1) * np.exp(-1), num_vars)
    stat = np.sum(sampler.sample(bqm, randomize_order=True, **common).record.sample == 1)
    assert lo < stat < hi
    # Gibbs, random order: 50:50 on the spins picked at least once
    lo, hi = _interval(1 / np.exp(1) + (1 - 1 / np.exp(1)) * 0.5, num_vars)
    stat = np.sum(sampler.sample(bqm, randomize_order=True, proposal_acceptance_criteria="Gibbs",
                                 **common).record.sample == 1)
    assert lo < stat < hi
    # Gibbs with a field: +1 with probability e^-b / (2 cosh b) at the final beta
    bqm = BinaryQuadraticModel.from_ising({i: 1 for i in range(num_vars)}, {})
    betas = [1, 1.5]
    lo, hi = _interval(np.exp(-betas[-1]) / (2 * np.cosh(betas[-1])), num_vars)
    response = sampler.sample(bqm, initial_states=initial_states, num_reads=1, seed=SEED,
                              proposal_acceptance_criteria="Gibbs", beta_schedule_type="custom",
                              beta_schedule=betas, num_sweeps_per_beta=1)
    stat = np.sum(response.record.sample == 1)
    assert lo < stat < hi


def test_sa_greedy_limit_independent_spins():
    num_vars = 10000
    init_vector = np.random.normal(size=num_vars)
    bqm = BinaryQuadraticModel.from_ising({i: init_vector[i] for i in range(num_vars)}, {})
    ground = np.array([-int(np.sign(bqm.linear[i])) for i in range(num_vars)])
    ground_state = SampleSet.from_samples_bqm({i: int(ground[i]) for i in range(num_vars)}, bqm)
    sky_state = SampleSet.from_samples_bqm({i: -int(ground[i]) for i in range(num_vars)}, bqm)
    lo, hi = _interval(1 - np.exp(-1), num_vars)
    for randomize_order in (False, True):
        for crit in ("Metropolis", "Gibbs"):
            kw = dict(num_reads=1, num_sweeps=1, randomize_order=randomize_order, proposal_acceptance_criteria=crit,
                      beta_schedule_type="custom", beta_schedule=[float("inf")], seed=SEED)
            r = SimulatedAnnealingSampler().sample(bqm, initial_states=ground_state, **kw)
            assert np.all(r.record.sample == ground)
            r = SimulatedAnnealingSampler().sample(bqm, initial_states=sky_state, **kw)
            if not randomize_order:
                assert np.all(r.record.sample == ground)
            else:
                stat = np.sum(r.record.sample == ground)
                assert lo < stat < hi


# ---------------------------------------------------------------------------
# tests/test_path_annealing_sampler.py, restated
# ---------------------------------------------------------------------------

def test_pimc_job_shop_scheduling_with_linear():
    bqm = _jss_bqm()
    optimal_energy = bqm.energy((np.array([[JSS_OPTIMUM[v] for v in bqm.variables]]), bqm.variables))
    response = PathIntegralAnnealingSampler().sample(bqm, beta_schedule_type="linear", num_reads=10)
    _, response_energy, _ = next(response.data())
    assert response_energy < optimal_energy + 2


def test_pimc_cubic_lattice_with_geometric():
    response = PathIntegralAnnealingSampler().sample_ising({}, _cubic_lattice_J(12), beta_schedule_type="geometric",
                                                           num_reads=10)
    _, response_energy, _ = next(response.data())
    assert response_energy < -2900, f"response_energy, {response_energy}, exceeds threshold, -2900"
