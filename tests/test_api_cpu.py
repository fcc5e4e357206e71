"""Host-side checks that need no GPU: the C-ABI library loads and exports
every symbol include/misamplers.h declares, argument errors map to the
reference's messages, and the Python layers validate like the reference."""
import ctypes
import os
import re
import warnings

import numpy as np
import pytest

import misamplers
from misamplers import _lib
from misamplers.sa import sampler as sa_sampler
from misamplers.sa.simulated_annealing import simulated_annealing
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "misamplers.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(misa_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert "misa_sa_sample" in names and "misa_sa_run_device" in names
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in names:
        assert hasattr(lib, name), name


def test_read_seeds_match_oracle():
    for seed in (0, 1, 2**31 - 1, 123456789):
        np.testing.assert_array_equal(_lib.read_seeds(seed, 100), oracle.read_seeds(seed, 100))


def test_random_spins_definition():
    """+1 exactly when the top bit of splitmix64 output i of
    (seed XOR 0xD1B54A32D192ED03) is set (Python integers, not the library)."""
    M = (1 << 64) - 1

    def splitmix(seed, i):
        z = (seed + (i + 1) * 0x9E3779B97F4A7C15) & M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    for seed in (0, 7, 2**63 + 11):
        got = _lib.random_spins(seed, 300)
        want = [1 if splitmix(seed ^ 0xD1B54A32D192ED03, i) >> 63 else -1 for i in range(300)]
        np.testing.assert_array_equal(got, want)
    s = _lib.random_spins(3, 100000)
    assert abs(int(s.astype(np.int64).sum())) < 5 * 316  # balanced


def test_problem_errors_match_reference_messages():
    with pytest.raises(ValueError, match="coupler indexes contain an invalid variable"):
        _lib.Problem([0, 0], [0], [2], [1.0])
    with pytest.raises(ValueError, match="coupler indexes contain an invalid variable"):
        _lib.Problem([0, 0], [-1], [1], [1.0])
    with pytest.raises(ValueError, match="coupler vectors have mismatched lengths"):
        _lib.Problem([0, 0], [0, 1], [1], [1.0])


def test_auto_mode_selection():
    assert _lib.Problem([0, 1, -1], [0, 1], [1, 2], [1.0, -2.0]).auto_mode == _lib.MODE_INT8
    assert _lib.Problem([100, 0], [0], [1], [3.0]).auto_mode == _lib.MODE_INT16
    assert _lib.Problem([0.5, 0], [0], [1], [3.0]).auto_mode == _lib.MODE_REAL
    assert _lib.Problem([5000, 0], [0], [1], [3.0]).auto_mode == _lib.MODE_REAL


def test_empty_inputs_need_no_device():
    s, e = simulated_annealing(0, [0.0, 1.0], [0], [1], [1.0], 1, [1.0], 5, np.empty((0, 2), np.int8))
    assert s.shape == (0, 2) and e.shape == (0,)
    s, e = simulated_annealing(3, [], [], [], [], 1, [1.0], 5, np.empty((3, 0), np.int8))
    assert s.shape == (3, 0) and np.array_equal(e, np.zeros(3))


def test_core_argument_validation():
    st = np.ones((2, 2), np.int8)
    with pytest.raises(ValueError, match="Unknown proposal_acceptance_criteria"):
        simulated_annealing(2, [0, 0], [0], [1], [1.0], 1, [1.0], 1, st, proposal_acceptance_criteria="x")
    with pytest.raises(ValueError):
        simulated_annealing(2, [0, 0], [0], [1], [1.0], 1, [1.0], 1, np.ones((2, 2), np.int16))


def test_seed_validation():
    s = misamplers.SimulatedAnnealingSampler()
    for bad in (3.5, float("inf"), "string", [], {}):
        with pytest.raises(TypeError):
            s.sample_ising({}, {}, seed=bad)
    for bad in (-1, -100, 2**65):
        with pytest.raises(ValueError):
            s.sample_ising({}, {}, seed=bad)


def test_schedule_validation():
    s = misamplers.SimulatedAnnealingSampler()
    h, J = {"a": 1}, {("a", "b"): 1}
    with pytest.raises(ValueError):
        s.sample_ising(h, J, beta_schedule_type="custom")
    with pytest.raises(ValueError):
        s.sample_ising(h, J, beta_schedule_type="custom", beta_schedule=[-1, 1])
    with pytest.raises(ValueError):
        s.sample_ising(h, J, beta_schedule_type="custom", beta_schedule=[1, 2], num_sweeps=3)
    with pytest.raises(ValueError):
        s.sample_ising(h, J, beta_schedule=[1, 2])
    with pytest.raises(ValueError):
        s.sample_ising(h, J, num_sweeps=10, num_sweeps_per_beta=3)
    with pytest.raises(ValueError):
        s.sample_ising(h, J, beta_range=[0, 1], beta_schedule_type="geometric")
    with pytest.raises(ValueError):
        s.sample_ising(h, J, beta_schedule_type="exotic")
    with pytest.raises(TypeError):
        s.sample_ising(h, J, num_sweeps_per_beta=1.5)
    with pytest.raises(ValueError):
        s.sample_ising(h, J, num_reads=0)
    with pytest.raises(TypeError):
        s.sample_ising(h, J, num_reads=3.5)


def test_schedules_closed_form():
    lin, quad = {"a": 1.0}, {}
    b, _ = sa_sampler.build_beta_schedule([0.5, 2.0], 10, 2, "linear", None, lin, quad)
    np.testing.assert_array_equal(b, np.linspace(0.5, 2.0, 5))
    b, _ = sa_sampler.build_beta_schedule([0.5, 2.0], 10, 2, "geometric", None, lin, quad)
    np.testing.assert_array_equal(b, np.geomspace(0.5, 2.0, 5))
    b, _ = sa_sampler.build_beta_schedule([0.5, 2.0], 3, 3, "geometric", None, lin, quad)
    np.testing.assert_array_equal(b, [2.0])


def test_default_beta_range_closed_forms():
    f = sa_sampler._default_ising_beta_range
    # one variable, h = 1: hot = ln2/2, cold = ln(1/0.01)/2
    hot, cold = f({"a": 1.0}, {})
    assert hot == pytest.approx(np.log(2) / 2) and cold == pytest.approx(np.log(100) / 2)
    # one edge J = -3 (+ h = 0): each spin sees 3: hot = ln2/6; two minimal gaps: ln(2/0.01)/6
    hot, cold = f({"a": 0.0, "b": 0.0}, {("a", "b"): -3.0})
    assert hot == pytest.approx(np.log(2) / 6) and cold == pytest.approx(np.log(200) / 6)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        assert f({"a": 0.0}, {}) == [0.1, 1]
        assert len(w) == 1
    # reference test inequalities (tests/test_simulated_annealing_sampler.py:483-524)
    h = {i: 0.5 * (-1) ** i for i in range(6)}
    J = {(i, i + 1): 1.0 for i in range(5)}
    b1 = f(h, J, scale_T_with_N=True)
    b2 = f(h, J, scale_T_with_N=False)
    assert b1[1] >= b2[1] and b1[0] == b2[0]
    b3 = f(h, J, max_single_qubit_excitation_rate=0.001)
    assert b3[1] > b1[1]
    with pytest.raises(ValueError):
        f(h, J, max_single_qubit_excitation_rate=1.5)


def test_dimodlite_vartype_round_trip():
    bqm = misamplers.BinaryQuadraticModel({"a": 1.0, "b": -2.0}, {("a", "b"): 3.0}, 0.5, "BINARY")
    spin = bqm.change_vartype("SPIN", inplace=False)
    xs = np.array([[0, 0], [0, 1], [1, 0], [1, 1]])
    np.testing.assert_allclose(bqm.energies((xs, ["a", "b"])), spin.energies((2 * xs - 1, ["a", "b"])))


def test_tree_solver_api_without_device():
    # TreeDecompositionSolver's surface (tree/samplers.py:29-183): parameters,
    # properties, the empty model and the elimination-order check need no GPU
    from misamplers import BinaryQuadraticModel, TreeDecompositionSolver
    solver = TreeDecompositionSolver()
    assert solver.properties == {"max_treewidth": 25}
    assert set(solver.parameters) == {"num_reads", "elimination_order"}
    ss = solver.sample(BinaryQuadraticModel.empty("SPIN"), num_reads=10)
    assert len(ss) == 10 and np.all(ss.record.energy == 0.0)
    ss = solver.sample(BinaryQuadraticModel({}, {}, 1.5, "BINARY"), num_reads=3)
    assert np.all(ss.record.energy == 1.5)
    bqm = BinaryQuadraticModel.from_ising({"a": 1.0, "b": -1.0}, {("a", "b"): 0.5})
    with pytest.raises(ValueError):
        solver.sample(bqm, elimination_order=["a"])
