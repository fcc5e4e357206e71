"""Batched launch: every problem's samples equal the plain engine's (float64
mode) and the oracle's, for every geometry policy and with hipGraph replay."""
import numpy as np
import pytest

import oracle
from conftest import random_regular
from misamplers import _lib
from misamplers.sa.batch import Batch, simulated_annealing_batch
from misamplers.sa.simulated_annealing import simulated_annealing

pytestmark = pytest.mark.gpu


def make_batch(rng, sizes, reads, integer=False):
    problems, schedules, inits, seeds = [], [], [], []
    for i, (n, R) in enumerate(zip(sizes, reads)):
        d = 3 if n < 20 else 15
        if (n * d) % 2:
            d -= 1
        cs, ce = random_regular(n, d, rng)
        if integer:
            h, cw = rng.integers(-1, 2, n).astype(float), rng.choice([-1.0, 1.0], len(cs))
        else:
            h, cw = rng.normal(0, 0.1, n), rng.normal(0, 1, len(cs))
        problems.append((h, cs, ce, cw))
        schedules.append(np.geomspace(0.05 + 0.01 * i, 3.0, 12 + i % 5))
        inits.append((2 * rng.integers(0, 2, (R, n)) - 1).astype(np.int8))
        seeds.append(1000 + 7 * i)
    return problems, schedules, inits, seeds


@pytest.mark.parametrize("integer", [False, True])
@pytest.mark.parametrize("order,prop", [(0, 1), (0, 0), (1, 1)])
def test_batch_equals_plain_and_oracle(integer, order, prop):
    rng = np.random.default_rng(31 + integer + 2 * order + 4 * prop)
    sizes = [10, 64, 130, 400, 1000, 2500]
    reads = [3, 64, 17, 40, 9, 5]
    problems, schedules, inits, seeds = make_batch(rng, sizes, reads, integer)
    names = {0: "Gibbs", 1: "Metropolis"}
    out = simulated_annealing_batch(problems, reads, 1, schedules, seeds, inits, bool(order), names[prop])
    for b, (h, cs, ce, cw) in enumerate(problems):
        st = inits[b].copy()
        ps, pe = simulated_annealing(reads[b], h, cs, ce, cw, 1, schedules[b], seeds[b], st, bool(order),
                                     names[prop], mode=_lib.MODE_REAL)
        np.testing.assert_array_equal(out[b][0], ps)
        np.testing.assert_array_equal(out[b][1], pe)
        if sizes[b] <= 400:
            ref = oracle.sa_sample(h, cs, ce, cw, 1, schedules[b], _lib.read_seeds(seeds[b], reads[b]), inits[b],
                                   order, prop, nthreads=8)
            np.testing.assert_array_equal(out[b][0], ref[0])
            np.testing.assert_array_equal(out[b][1], ref[1])


def test_policies_agree():
    rng = np.random.default_rng(5)
    problems, schedules, inits, seeds = make_batch(rng, [40, 300, 2000, 4096], [33, 64, 8, 4])
    base = None
    for pol in range(16):
        out = simulated_annealing_batch(problems, [33, 64, 8, 4], 1, schedules, seeds, inits, policy=pol)
        if base is None:
            base = out
        for (s0, e0), (s1, e1) in zip(base, out):
            np.testing.assert_array_equal(s0, s1)
            np.testing.assert_array_equal(e0, e1)


def test_autotune_then_graph_replay():
    import torch
    rng = np.random.default_rng(6)
    sizes, reads = [64, 500, 1500], [64, 32, 16]
    problems, schedules, inits, seeds = make_batch(rng, sizes, reads)
    batch = Batch(problems, reads, 1, schedules)
    pol = batch.autotune(4)
    assert 0 <= pol <= 15
    init = np.concatenate([i.ravel() for i in inits])
    seed_all = np.concatenate([_lib.read_seeds(s, r) for s, r in zip(seeds, reads)])
    ref_states = init.copy()
    ref_en = batch.sample(ref_states, seed_all)
    dev = torch.device("cuda:0")
    d_init = torch.from_numpy(init).to(dev)
    d_st = torch.empty_like(d_init)
    d_en = torch.empty(batch.total_reads, dtype=torch.float64, device=dev)
    d_sd = torch.from_numpy(seed_all.view(np.int64)).to(dev)
    stream = torch.cuda.Stream()
    for rep in range(3):  # capture once, replay; eager host work between replays
        with torch.cuda.stream(stream):
            d_st.copy_(d_init)
            batch.run_device(d_st.data_ptr(), d_en.data_ptr(), d_sd.data_ptr(), stream=stream.cuda_stream,
                             use_graph=True)
        stream.synchronize()
        np.testing.assert_array_equal(d_st.cpu().numpy(), ref_states)
        np.testing.assert_array_equal(d_en.cpu().numpy(), ref_en)
    batch.close()


def test_buffer_sets_back_to_back_on_a_side_stream():
    """Two buffer sets run one after the other on a side stream with no host
    synchronisation between them: the second call rebuilds the work items
    while the first launch may still run, and both results must equal the
    host-buffer runs."""
    import torch
    rng = np.random.default_rng(8)
    sizes, reads = [80, 700, 1500], [40, 24, 12]
    problems, schedules, inits, seeds = make_batch(rng, sizes, reads)
    batch = Batch(problems, reads, 1, schedules)
    dev = torch.device("cuda:0")
    sets = []
    for k in range(2):
        init = np.concatenate([(2 * rng.integers(0, 2, i.shape) - 1).astype(np.int8).ravel() for i in inits])
        seed_all = np.concatenate([_lib.read_seeds(s + 17 * k, r) for s, r in zip(seeds, reads)])
        ref = init.copy()
        ref_en = batch.sample(ref, seed_all)
        d_st = torch.from_numpy(init).to(dev)
        d_en = torch.empty(batch.total_reads, dtype=torch.float64, device=dev)
        d_sd = torch.from_numpy(seed_all.view(np.int64)).to(dev)
        sets.append((ref, ref_en, d_st, d_en, d_sd))
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        for _, _, d_st, d_en, d_sd in sets:
            batch.run_device(d_st.data_ptr(), d_en.data_ptr(), d_sd.data_ptr(), stream=stream.cuda_stream)
    stream.synchronize()
    for ref, ref_en, d_st, d_en, _ in sets:
        np.testing.assert_array_equal(d_st.cpu().numpy(), ref)
        np.testing.assert_array_equal(d_en.cpu().numpy(), ref_en)
    batch.close()


def test_oversized_problem_is_rejected():
    rng = np.random.default_rng(7)
    n = 30000
    cs, ce = random_regular(n, 4, rng)
    with pytest.raises(_lib.EngineError, match="too large"):
        simulated_annealing_batch([(np.zeros(n), cs, ce, np.ones(len(cs)))], [2], 1, [np.ones(2)], [1],
                                  [np.ones((2, n), np.int8)])
