"""GPU parity of the path-integral engine, through the C ABI.

* Bit-exact against the oracle restatement (oracle/pimc_oracle.c) on the same
  seeds and initial worldlines: every worldline word, every statistics
  snapshot, and energies in the reference's summation order.
* The reference's own core-function tests (tests/test_path_annealing.py)
  restated against the drop-in ``simulated_annealing``.
* Equilibrium against the exact transfer matrix (statistical).
* The break-list format and buffer discipline of localPIMC.cpp:169-199.
"""
import math

import numpy as np
import pytest

import oracle
from conftest import lattice3d, random_regular
from misamplers import PathIntegralAnnealingSampler, _lib
from misamplers.sqa.engine import PimcPlan, colouring
from misamplers.sqa.pimc_annealing import simulated_annealing
from oracle import pimc_exact as px

pytestmark = pytest.mark.gpu


def _mask(P):
    return np.uint32(0xFFFFFFFF) if P == 32 else np.uint32((1 << P) - 1)


def run_engine(h, cs, ce, cw, spb, hp, hd, seeds, W0, P=32, interval=0):
    """Device run of the plan on torch buffers; returns (W, stats, states, energies)."""
    import torch
    plan = PimcPlan(h, cs, ce, cw, spb, hp, hd, P)
    R, n = W0.shape
    dev = torch.device("cuda:0")
    dW = torch.from_numpy(W0.view(np.int32).copy()).to(dev)
    dS = torch.from_numpy(np.ascontiguousarray(seeds, dtype=np.uint64).view(np.int64).copy()).to(dev)
    ncol = plan.stat_columns(interval)
    dstat = torch.zeros((R, max(ncol, 1), n), dtype=torch.int8, device=dev)
    dst = torch.zeros((R, n), dtype=torch.int8, device=dev)
    den = torch.zeros(R, dtype=torch.float64, device=dev)
    dfl = torch.zeros(R, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    plan.run_device(dW.data_ptr(), dS.data_ptr(), R, dstat.data_ptr() if interval else 0, interval, stream,
                    d_flips=dfl.data_ptr())
    plan.project_device(dW.data_ptr(), R, dst.data_ptr(), den.data_ptr(), stream)
    torch.cuda.synchronize()
    info = plan.info(R)
    info["flips"] = dfl.cpu().numpy().view(np.uint64)
    out = (dW.cpu().numpy().view(np.uint32), dstat.cpu().numpy()[:, :ncol] if interval else None,
           dst.cpu().numpy(), den.cpu().numpy())
    plan.close()
    return out + (info,)


def check_exact(h, cs, ce, cw, spb, hp, hd, R, P, seed, interval=0, rng=None, classical=False):
    rng = rng or np.random.default_rng(seed)
    n = len(h)
    W0 = rng.integers(0, 2**32, size=(R, n), dtype=np.uint64).astype(np.uint32) & _mask(P)
    if classical:
        W0 = np.where(W0 & 1, _mask(P), np.uint32(0)).astype(np.uint32)
    seeds = oracle.read_seeds(seed, R)
    print(f"case n={n} m={len(cs)} R={R} P={P} steps={len(hp)} spb={spb} interval={interval}", flush=True)
    W, stats, st, en, info = run_engine(h, cs, ce, cw, spb, hp, hd, seeds, W0, P, interval)
    nc, col = colouring(h, cs, ce, cw)
    assert info["num_classes"] == nc
    Wo, so, fo = oracle.pimc_sample(h, cs, ce, cw, spb, hp, hd, seeds, W0, num_slices=P, interval=interval,
                                    nthreads=8, return_flips=True, colour=col)
    np.testing.assert_array_equal(W, Wo)
    np.testing.assert_array_equal(info["flips"], fo)
    if interval:
        np.testing.assert_array_equal(stats, so)
    np.testing.assert_array_equal(st, np.where(Wo & 1, 1, -1).astype(np.int8))
    for r in range(R):
        assert en[r] == oracle.state_energy(st[r], h, cs, ce, cw)
    return W, info


def test_device_present():
    assert _lib.lib.misa_device_count() >= 1


def test_exact_signed_regular():
    rng = np.random.default_rng(1)
    cs, ce = random_regular(40, 3, rng)
    cw = rng.choice([-1.0, 1.0], len(cs))
    h = rng.uniform(-0.5, 0.5, 40)
    _, info = check_exact(h, cs, ce, cw, 2, np.linspace(0.1, 3, 30), np.linspace(3, 0, 30), 37, 32, 11, rng=rng)
    assert info["kind"] == "signed"


def test_exact_table_with_selfloops_and_repeats():
    rng = np.random.default_rng(2)
    n = 50
    pairs = [(u, v) for u in range(n) for v in range(u, n) if rng.random() < 0.08]
    pairs += pairs[:5]  # repeated couplers
    cs = np.array([a for a, _ in pairs], np.int32)
    ce = np.array([b for _, b in pairs], np.int32)
    cw = rng.normal(0, 1, len(cs))
    h = rng.normal(0, 0.3, n)
    hd = np.concatenate([np.linspace(4, 0, 20), [0.0, 1e3, 1e-30, 2.0]])
    hp = np.linspace(0.2, 2.0, len(hd))
    _, info = check_exact(h, cs, ce, cw, 1, hp, hd, 20, 7, 12, rng=rng)
    assert info["kind"] == "table"


def test_exact_high_degree_tail():
    rng = np.random.default_rng(3)
    n = 30
    pairs = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < 0.6]
    cs = np.array([a for a, _ in pairs], np.int32)
    ce = np.array([b for _, b in pairs], np.int32)
    for cw, kind in ((rng.uniform(-1, 1, len(cs)), "table"), (np.where(rng.random(len(cs)) < 0.5, -0.75, 0.75),
                                                                 "signed")):
        _, info = check_exact(rng.uniform(-1, 1, n), cs, ce, cw, 1, np.linspace(0.1, 2, 15),
                              np.linspace(2, 0.1, 15), 9, 16, 13, rng=rng)
        assert info["kind"] == kind and info["max_degree"] > 8


def test_exact_single_slice_and_statistics():
    rng = np.random.default_rng(4)
    cs, ce = random_regular(24, 4, rng)
    cw = rng.normal(0, 1, len(cs))
    check_exact(rng.normal(0, 1, 24), cs, ce, cw, 1, np.linspace(0.1, 2, 12), np.linspace(1, 0, 12), 10, 1, 14,
                interval=3, rng=rng)
    check_exact(rng.normal(0, 1, 24), cs, ce, cw, 3, np.linspace(0.1, 2, 12), np.linspace(1, 0, 12), 70, 32, 15,
                interval=1, rng=rng)


def test_exact_lattice_many_reads():
    rng = np.random.default_rng(5)
    cs, ce = lattice3d(6)
    cw = rng.choice([-1.0, 1.0], len(cs))
    _, info = check_exact(np.zeros(216), cs, ce, cw, 1, np.linspace(0, 4, 40), np.linspace(6, 0, 40), 300, 32,
                          16, rng=rng)
    assert info["num_classes"] == 2


def test_exact_full_size_subset():
    """The benchmark problem (n=2048, 6-regular, J=+-1, h~U[-0.5,0.5], P=32,
    Hd 10->0, Hp 0->10 over 1000 steps) on a handful of reads."""
    rng = np.random.default_rng(2048)
    cs, ce = random_regular(2048, 6, rng)
    cw = rng.choice([-1.0, 1.0], len(cs))
    h = rng.uniform(-0.5, 0.5, 2048)
    check_exact(h, cs, ce, cw, 1, np.linspace(0, 10, 1000), np.linspace(10, 0, 1000), 6, 32, 17, rng=rng,
                classical=True)


def test_equilibrium_matches_transfer_matrix():
    from scipy import stats as sps
    h = np.array([0.3, -0.2, 0.45, -0.1, 0.05])
    cs, ce, cw = [0, 1, 2, 3, 0, 1], [1, 2, 3, 4, 4, 3], [0.7, -1.0, 0.5, -0.8, 1.1, 0.3]
    R, P = 40000, 8
    rng = np.random.default_rng(6)
    W0 = rng.integers(0, 2**32, size=(R, 5), dtype=np.uint64).astype(np.uint32) & _mask(P)
    W, _, st, _, _ = run_engine(h, cs, ce, cw, 1, np.full(40, 2.0), np.full(40, 0.5), oracle.read_seeds(3, R), W0, P)
    p = px.TrotterModel(h, cs, ce, cw, 2.0, 0.5, P).slice_distribution()
    counts = np.bincount(px.state_index(st), minlength=32)
    chi2 = ((counts - p * R) ** 2 / (p * R)).sum()
    assert sps.chi2.sf(chi2, 31) > 1e-4


# ---------------------------------------------------------------------------
# The reference's core-function tests (tests/test_path_annealing.py)
# ---------------------------------------------------------------------------

def _fm_problem(num_variables=10, num_samples=10, num_sweeps=100):
    """tests/test_path_annealing.py:94-143."""
    h = [-1] * num_variables
    cs, ce, cw = zip(*((u, v, -1) for u in range(num_variables) for v in range(u, num_variables)))
    initial = 2 * np.random.RandomState(1234).randint(2, size=(num_samples, num_variables)).astype(np.int8) - 1
    return [num_samples, h, cs, ce, cw, 1, np.linspace(0.01, 3, num=num_sweeps), np.linspace(3, 0, num=num_sweeps),
            1, 1, 1, 1, 1, initial, True, True, np.zeros(initial.shape, dtype=np.intc), np.empty(1, dtype=np.intc),
            np.empty(1, dtype=np.intc), 0]


def test_submit_problem():
    n, R = 5, 10
    out = simulated_annealing(*_fm_problem(n, R))
    assert len(out) == 5
    samples, energies = out[0], out[1]
    assert isinstance(samples, np.ndarray) and samples.shape == (R, n)
    assert isinstance(energies, np.ndarray) and energies.shape == (R,)


def test_good_results():
    n = 5
    samples, energies, _, _, _ = simulated_annealing(*_fm_problem(num_variables=n))
    assert [1] * n in samples.tolist()
    assert abs(np.mean(energies) - (-(n + 3) * n / 2)) <= 2


def test_seed():
    n, S, R = 100, 5, 100
    h = [0] * n
    cs, ce, cw = zip(*((u, v, 1) for u in range(n) for v in range(u, n)))
    init = 2 * np.random.RandomState(1234).randint(1, size=(R, n)).astype(np.int8) - 1
    prev = []
    for seed in (1, 40, 235, 152436, 3462354, 92352355):
        args = lambda: [R, h, cs, ce, cw, 1, np.linspace(0.3, 0.4, num=S), np.ones(S), 1, 1, 1, 1, seed,
                        np.copy(init), True, True, np.zeros((R, n), dtype=np.intc), np.empty(1, dtype=np.intc),
                        np.empty(1, dtype=np.intc), 0]
        s0 = simulated_annealing(*args())[0]
        s1 = simulated_annealing(*args())[0]
        assert np.array_equal(s0, s1)
        for p in prev:
            assert not np.array_equal(s0, p)
        prev.append(s0)


def test_immediate_interrupt():
    samples, energies, _, _, _ = simulated_annealing(*_fm_problem(5), interrupt_function=lambda: True)
    assert samples.shape == (1, 5) and energies.shape == (1,)


def test_5_interrupt():
    count = [1]

    def stop():
        if count[0] >= 5:
            return True
        count[0] += 1
        return False

    samples, energies, _, _, _ = simulated_annealing(*_fm_problem(5), interrupt_function=stop)
    assert samples.shape == (5, 5) and energies.shape == (5,)


def test_initial_states_zero_sweeps():
    n, R = 100, 100
    cs, ce, cw = zip(*((u, v, 1) for u in range(n) for v in range(u, n)))
    init = 2 * np.random.RandomState(1234).randint(1, size=(R, n)).astype(np.int8) - 1
    samples = simulated_annealing(R, [0] * n, cs, ce, cw, 1, np.linspace(0.3, 0.4, num=0),
                                  np.linspace(0.4, 0.3, num=0), 1, 2, 1, 1, 1234567890, np.copy(init), True, True,
                                  np.zeros((R, n), dtype=np.intc), np.empty(1, dtype=np.intc),
                                  np.empty(1, dtype=np.intc), 0)[0]
    assert np.array_equal(init, samples)


def test_chain_updates_need_long_enough_chains():
    """qubits_per_update != 1 selects chain updates, which need
    qubits_per_chain >= qubits_per_update (localPIMC.cpp:101)."""
    args = _fm_problem(5)
    args[11] = 2
    with pytest.raises(ValueError, match="qubits_per_chain must be at least qubits_per_update"):
        simulated_annealing(*args)


# ---------------------------------------------------------------------------
# Break lists (localPIMC.cpp:169-199)
# ---------------------------------------------------------------------------

def _breaks_of(W, P):
    """Per-spin sorted break positions of worldline words (kink between
    slices k-1 and k at (2k-1) * 2^16 / (2P))."""
    out = []
    for w in W.ravel():
        w = int(w)
        out.append([((2 * k - 1) * 65536) // (2 * P) for k in range(1, P + 1)
                    if ((w >> (k - 1)) & 1) != ((w >> (k % P)) & 1)])
    return out


def _read_breaks(lists, R, n, size):
    """localPIMC.cpp:188-199 with the caller's cursor advance (:703-706)."""
    buf = np.zeros(size, dtype=np.intc)
    nb = np.zeros((R, n), dtype=np.intc)
    cursor = 0
    for r in range(R):
        cap = size - cursor
        wp = cursor
        tot = 0
        for i in range(n):
            b = lists[r * n + i]
            nb[r, i] = len(b)
            if cap >= len(b):
                buf[wp:wp + len(b)] = b
                wp += len(b)
                cap -= len(b)
            tot += len(b)
        cursor += tot
    return nb, buf


def _quantum_run(h, cs, ce, cw, hp, hd, init, seed, P, num_breaks=None, breaks_in=None, out_size=None):
    R, n = init.shape
    if num_breaks is None:
        num_breaks = np.zeros((R, n), dtype=np.intc)
        breaks_in = np.zeros(1, dtype=np.intc)
        proj_in = True
    else:
        proj_in = False
    out = np.zeros(out_size if out_size is not None else R * n * P + 1, dtype=np.intc)
    return simulated_annealing(R, h, cs, ce, cw, 1, hp, hd, 1, 1, 1, 1, seed, np.copy(init), proj_in, False,
                               np.copy(num_breaks), np.ascontiguousarray(breaks_in, dtype=np.intc), out, 0,
                               num_trotter_slices=P)


def test_break_lists_round_trip():
    rng = np.random.default_rng(7)
    cs, ce = random_regular(30, 3, rng)
    cw = rng.normal(0, 1, len(cs))
    h = rng.normal(0, 0.5, 30)
    init = rng.choice([-1, 1], size=(12, 30)).astype(np.int8)
    for P in (32, 5):
        st, en, _, nb, buf = _quantum_run(h, cs, ce, cw, np.full(10, 1.0), np.full(10, 2.0), init, 5, P)
        total = int(nb.sum())
        assert total > 0
        pos = buf[:total]
        assert np.all((pos >= 0) & (pos < 65536))
        # feeding the worldlines back with zero steps returns them unchanged
        st2, en2, _, nb2, buf2 = _quantum_run(h, cs, ce, cw, [], [], st, 9, P, nb, pos)
        np.testing.assert_array_equal(st2, st)
        np.testing.assert_array_equal(nb2, nb)
        np.testing.assert_array_equal(buf2[:total], pos)
        np.testing.assert_array_equal(en2, en)


def test_break_lists_match_worldlines_and_buffer_discipline():
    rng = np.random.default_rng(8)
    cs, ce = random_regular(20, 3, rng)
    cw = rng.normal(0, 1, len(cs))
    h = rng.normal(0, 0.5, 20)
    R, P = 6, 32
    seeds = oracle.read_seeds(4, R)
    init = rng.choice([-1, 1], size=(R, 20)).astype(np.int8)
    W0 = np.where(init > 0, _mask(P), np.uint32(0)).astype(np.uint32)
    Wo, _ = oracle.pimc_sample(h, cs, ce, cw, 1, np.full(8, 1.0), np.full(8, 3.0), seeds, W0, num_slices=P,
                               colour=colouring(h, cs, ce, cw)[1])
    lists = _breaks_of(Wo, P)
    total = sum(len(b) for b in lists)
    for size in (total + 3, total, total // 2, 1):
        out = np.zeros(size, dtype=np.intc)
        st, en, _, nb, buf = simulated_annealing(
            R, h, cs, ce, cw, 1, np.full(8, 1.0), np.full(8, 3.0), 1, 1, 1, 1, 0, np.copy(init), True, False,
            np.zeros((R, 20), dtype=np.intc), np.zeros(1, dtype=np.intc), out, 0, read_seeds=seeds)
        want_nb, want_buf = _read_breaks(lists, R, 20, size)
        np.testing.assert_array_equal(nb, want_nb)
        np.testing.assert_array_equal(buf, want_buf)
        np.testing.assert_array_equal(st, np.where(Wo & 1, 1, -1))


def test_quantum_input_matches_worldline_input():
    """Initial break lists select the same worldlines as the words they encode."""
    rng = np.random.default_rng(9)
    cs, ce = random_regular(16, 3, rng)
    cw = rng.normal(0, 1, len(cs))
    h = rng.normal(0, 0.5, 16)
    R, P = 5, 32
    W0 = rng.integers(0, 2**32, size=(R, 16), dtype=np.uint64).astype(np.uint32)
    lists = _breaks_of(W0, P)
    nb = np.array([len(b) for b in lists], dtype=np.intc).reshape(R, 16)
    flat = np.array([p for b in lists for p in b], dtype=np.intc)
    seeds = oracle.read_seeds(6, R)
    hp, hd = np.full(6, 1.0), np.full(6, 1.0)
    st, en, _, nb_out, buf = simulated_annealing(
        R, h, cs, ce, cw, 1, hp, hd, 1, 1, 1, 1, 0, np.where(W0 & 1, 1, -1).astype(np.int8), False, False, nb, flat,
        np.zeros(R * 16 * P, dtype=np.intc), 0, read_seeds=seeds)
    Wo, _ = oracle.pimc_sample(h, cs, ce, cw, 1, hp, hd, seeds, W0, num_slices=P,
                               colour=colouring(h, cs, ce, cw)[1])
    want = _breaks_of(Wo, P)
    np.testing.assert_array_equal(nb_out.ravel(), [len(b) for b in want])
    np.testing.assert_array_equal(buf[:int(nb_out.sum())], [p for b in want for p in b])


# ---------------------------------------------------------------------------
# Sampler
# ---------------------------------------------------------------------------

def test_sampler_docstring_example():
    sampler = PathIntegralAnnealingSampler()
    h = {"a": 0.0, "b": 0.0, "c": 0.0}
    J = {("a", "b"): 1.0, ("b", "c"): 1.0, ("a", "c"): 1.0}
    ss = sampler.sample_ising(h, J, num_reads=10, seed=3)
    assert ss.first.energy == -1.0
    assert set(ss.info) >= {"beta_range", "beta_schedule_type", "timing"}


def test_sampler_quantum_info_and_statistics():
    from misamplers import BinaryQuadraticModel
    bqm = BinaryQuadraticModel({0: 0.5, 1: -0.5, 2: 0.1}, {(0, 1): -1.0, (1, 2): 0.5}, 0.25, "BINARY")
    sampler = PathIntegralAnnealingSampler()
    ss = sampler.sample(bqm, num_reads=7, seed=11, beta_schedule_type="custom", Hp_field=np.linspace(0.1, 2, 9),
                        Hd_field=np.linspace(2, 0.1, 9), project_states=(True, False), schedule_sample_interval=4,
                        make_info_json_serializable=True)
    assert ss.vartype.name == "BINARY"
    assert np.array(ss.info["statistics"]).shape == (7, 3, 3)
    assert ss.info["complete_break_specification"] is True
    assert len(ss.info["break_positions"]) == int(np.sum(ss.info["num_breaks"]))
    for datum in ss.data():
        assert datum.energy == pytest.approx(bqm.energy(datum.sample))
