"""Host-side checks of the path-integral boundary that need no GPU: the
engine's colouring (update order), argument validation with the reference's
error types, and the empty-problem early return (pimc_annealing.pyx:197-203)."""
import numpy as np
import pytest

from conftest import lattice3d, random_regular
from misamplers import _lib
from misamplers.sqa.engine import colouring
from misamplers.sqa.pimc_annealing import simulated_annealing


def _proper(col, cs, ce):
    cs, ce = np.asarray(cs), np.asarray(ce)
    keep = cs != ce
    return np.all(col[cs[keep]] != col[ce[keep]])


def test_colouring_proper_and_balanced_on_regular_graph():
    rng = np.random.default_rng(4096)
    cs, ce = random_regular(2048, 6, rng)
    nc, col = colouring(np.zeros(2048), cs, ce, np.ones(len(cs)))
    assert _proper(col, cs, ce)
    sizes = np.bincount(col)
    assert len(sizes) == nc and sizes.max() - sizes.min() <= 1 and nc <= 6


def test_colouring_lattice_two_classes():
    cs, ce = lattice3d(8)
    nc, col = colouring(np.zeros(512), cs, ce, np.ones(len(cs)))
    assert nc == 2 and _proper(col, cs, ce) and np.bincount(col).tolist() == [256, 256]


def test_colouring_dense_and_selfloops():
    n = 40
    pairs = [(u, v) for u in range(n) for v in range(u, n)]  # complete graph plus self-loops
    cs = [a for a, _ in pairs]
    ce = [b for _, b in pairs]
    nc, col = colouring(np.zeros(n), cs, ce, np.ones(len(cs)))
    assert nc == n and _proper(col, cs, ce)
    nc, col = colouring(np.zeros(3), [], [], [])
    assert nc == 1 and col.tolist() == [0, 0, 0]


def _args(R=3, n=4, **kw):
    a = dict(num_samples=R, h=[0.0] * n, coupler_starts=[0], coupler_ends=[1], coupler_weights=[1.0],
             sweeps_per_beta=1, Hp_field=np.ones(3), Hd_field=np.ones(3), Gamma=1, chain_coupler_strength=1,
             qubits_per_chain=1, qubits_per_update=1, seed=1, states_numpy=np.ones((R, n), np.int8),
             project_inputs=True, project_outputs=True, num_breaks_numpy=np.zeros((R, n), np.intc),
             breaks_in_numpy=np.zeros(1, np.intc), breaks_buffer_out_numpy=np.zeros(1, np.intc),
             schedule_sample_interval=0)
    a.update(kw)
    return a


def test_empty_problem_returns_early():
    st, en, stats, nb, bo = simulated_annealing(**_args(R=0))
    assert st.shape == (0, 4) and en.shape == (0,) and stats.size == 0
    st, en, _, _, _ = simulated_annealing(**_args(R=3, n=0, coupler_starts=[], coupler_ends=[],
                                                  coupler_weights=[], states_numpy=np.ones((3, 0), np.int8),
                                                  num_breaks_numpy=np.zeros((3, 0), np.intc)))
    assert st.shape == (3, 0) and np.all(en == 0)


def test_buffer_types_checked():
    with pytest.raises(ValueError):
        simulated_annealing(**_args(states_numpy=np.ones((3, 4), np.int16)))
    with pytest.raises(ValueError):
        simulated_annealing(**_args(num_breaks_numpy=np.zeros((3, 4), np.int64)))
    with pytest.raises(ValueError):
        simulated_annealing(**_args(breaks_in_numpy=np.zeros((1, 1), np.intc)))
    with pytest.raises(ValueError):
        simulated_annealing(**_args(states_numpy=np.ones((2, 4), np.int8)))
    with pytest.raises(ValueError):
        simulated_annealing(**_args(Hd_field=np.ones(2)))


def test_seed_range():
    with pytest.raises(OverflowError):
        simulated_annealing(**_args(seed=2**32))
    with pytest.raises(OverflowError):
        simulated_annealing(**_args(seed=-1))


def test_engine_rejections_before_device_work():
    with pytest.raises(ValueError, match="coupler indexes"):
        simulated_annealing(**_args(coupler_starts=[0], coupler_ends=[9]))
    with pytest.raises(ValueError, match="coupler vectors"):
        simulated_annealing(**_args(coupler_starts=[0, 1], coupler_ends=[1]))
    with pytest.raises(ValueError, match="spins"):
        simulated_annealing(**_args(states_numpy=np.zeros((3, 4), np.int8)))
    # localPIMC.cpp:101: qubitsPerChain >= qubitsPerUpdate
    with pytest.raises(ValueError, match="qubits_per_chain must be at least qubits_per_update"):
        simulated_annealing(**_args(qubits_per_update=2))
    with pytest.raises(_lib.EngineError, match="chain updates support"):
        simulated_annealing(**_args(qubits_per_chain=40, qubits_per_update=2))
