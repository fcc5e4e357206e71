"""Sampler-boundary formats without dimod (host logic, CPU): initial-state
parsing with the 'none' / 'tile' / 'random' generators and the soft num_reads
rule, restated from the reference's sampler tests
(tests/test_simulated_annealing_sampler.py:335-429), and BINARY <-> SPIN
conversion checked by exhaustive enumeration."""
import itertools
import warnings

import numpy as np
import pytest

from misamplers.dimodlite import BinaryQuadraticModel, Initialized, SampleSet


def _triangle():
    return BinaryQuadraticModel.from_ising({}, {("a", "b"): -1, ("b", "c"): 1, ("a", "c"): 1})


def _init(bqm):
    return SampleSet.from_samples_bqm([{"a": 1, "b": 1, "c": 1}, {"a": -1, "b": -1, "c": -1}], bqm)


def _count(bqm, **kw):
    p = Initialized().parse_initial_states(bqm, **kw)
    st = p.initial_states.record.sample
    assert st.shape == (p.num_reads, 3) and st.dtype == np.int8
    assert np.isin(st, [-1, 1]).all()
    return p.num_reads, st


def test_initial_states_generator():
    """tests/test_simulated_annealing_sampler.py:335-404."""
    bqm = _triangle()
    init = _init(bqm)
    fixed = np.array([[1, 1, 1], [-1, -1, -1]], np.int8)
    # 2 fixed initial states, 8 random (implicit and explicit 'random')
    for kw in ({}, {"initial_states_generator": "random"}):
        k, st = _count(bqm, initial_states=init, num_reads=10, **kw)
        assert k == 10
        np.testing.assert_array_equal(st[:2], fixed)
    # all random
    assert _count(bqm, initial_states_generator="random", num_reads=10)[0] == 10
    assert _count(bqm, num_reads=10)[0] == 10
    # zero-length initial states in tuple format, extended by random samples
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        assert _count(bqm, initial_states=(np.empty((0, 3)), ["a", "b", "c"]), num_reads=10)[0] == 10
    # explicit None: one random initial state
    assert _count(bqm, initial_states=None)[0] == 1
    # truncation to num_reads, whatever the generator
    for gen in ("none", "tile", "random"):
        k, st = _count(bqm, initial_states=init, initial_states_generator=gen, num_reads=1)
        assert k == 1
        np.testing.assert_array_equal(st, fixed[:1])
    # 2 fixed initial states repeated 5 times
    k, st = _count(bqm, initial_states=init, initial_states_generator="tile", num_reads=10)
    np.testing.assert_array_equal(st, np.tile(fixed, (5, 1)))
    # cannot tile empty states; not enough states for 'none'
    with pytest.raises(ValueError):
        _count(bqm, initial_states_generator="tile", num_reads=10)
    with pytest.raises(ValueError):
        _count(bqm, initial_states_generator="none", num_reads=3)
    # initial states on other variables than the bqm's
    other = SampleSet.from_samples({"a": 1, "b": 1}, vartype="SPIN", energy=0)
    with pytest.raises(ValueError):
        _count(bqm, initial_states=other)
    with pytest.raises(ValueError):
        _count(bqm, initial_states_generator="sideways")


def test_soft_num_reads():
    """tests/test_simulated_annealing_sampler.py:406-429."""
    bqm = _triangle()
    init = _init(bqm)
    assert _count(bqm)[0] == 1
    assert _count(bqm, initial_states_generator="random")[0] == 1
    assert _count(bqm, initial_states=init)[0] == 2
    assert _count(bqm, initial_states=init, num_reads=1)[0] == 1
    assert _count(bqm, initial_states=init, num_reads=3)[0] == 3
    assert _count(bqm, num_reads=4)[0] == 4


def test_initial_states_follow_bqm_variable_order_and_vartype():
    bqm = _triangle()
    k, st = _count(bqm, initial_states=(np.array([[1, -1, -1]]), ["c", "a", "b"]))
    np.testing.assert_array_equal(st, [[-1, -1, 1]])
    with pytest.raises(ValueError):   # binary values for a spin model
        _count(bqm, initial_states=(np.array([[0, 1, 1]]), ["a", "b", "c"]))
    qubo = BinaryQuadraticModel({"a": 1.0, "b": 0.0, "c": 0.0}, {("a", "b"): 1.0}, 0.0, "BINARY")
    p = Initialized().parse_initial_states(qubo, num_reads=50, seed=3)
    assert set(np.unique(p.initial_states.record.sample)) <= {0, 1}
    q = Initialized().parse_initial_states(qubo, num_reads=50, seed=3)
    np.testing.assert_array_equal(p.initial_states.record.sample, q.initial_states.record.sample)


def test_empty_model_initial_states():
    """An empty model still yields num_reads (empty) initial states
    (tests/test_rotor_annealing_sampler.py:221-235)."""
    empty = BinaryQuadraticModel({}, {}, 0.0, "SPIN")
    p = Initialized().parse_initial_states(empty, num_reads=3)
    assert p.num_reads == 3 and p.initial_states.record.sample.shape == (3, 0)
    p = Initialized().parse_initial_states(empty)
    assert p.num_reads == 1 and p.initial_states.record.sample.shape == (1, 0)


@pytest.mark.parametrize("vartype", ["BINARY", "SPIN"])
def test_vartype_conversion_preserves_every_energy(vartype):
    """E_binary(x) = E_spin(2x - 1) on all 2^5 states, both directions, with
    offset; the conversion is exact for dyadic biases."""
    rng = np.random.default_rng(0)
    lin = {v: float(rng.integers(-8, 8)) / 4 for v in "abcde"}
    quad = {(u, v): float(rng.integers(-8, 8)) / 4 for u, v in itertools.combinations("abcde", 2)
            if rng.random() < 0.6}
    bqm = BinaryQuadraticModel(lin, quad, 1.25, vartype)
    other = bqm.change_vartype("SPIN" if vartype == "BINARY" else "BINARY", inplace=False)
    assert bqm.change_vartype(vartype, inplace=False).quadratic == bqm.quadratic
    labels = list("abcde")
    xs = np.array(list(itertools.product([0, 1], repeat=5)))
    ss = 2 * xs - 1
    e_b = (bqm if vartype == "BINARY" else other).energies((xs, labels))
    e_s = (other if vartype == "BINARY" else bqm).energies((ss, labels))
    np.testing.assert_array_equal(e_b, e_s)
    back = other.change_vartype(vartype, inplace=False)
    np.testing.assert_array_equal(back.energies((xs if vartype == "BINARY" else ss, labels)),
                                  bqm.energies((xs if vartype == "BINARY" else ss, labels)))


def test_parsed_initial_states_do_not_alias_the_callers():
    """The engines anneal the parsed array in place; the caller's initial
    states (a SampleSet reused across calls, an array) must stay as given."""
    bqm = _triangle()
    init = _init(bqm)
    before = init.record.sample.copy()
    p = Initialized().parse_initial_states(bqm, initial_states=init)
    p.initial_states.record.sample[:] = 0
    np.testing.assert_array_equal(init.record.sample, before)
    arr = np.array([[1, -1, 1]], np.int8)
    p = Initialized().parse_initial_states(bqm, initial_states=(arr, ["a", "b", "c"]))
    p.initial_states.record.sample[:] = 0
    np.testing.assert_array_equal(arr, [[1, -1, 1]])
