"""Every benchmark configuration at its full size on the GPU, against the oracle.

The problems come from ``bench.py``'s own generators, so these are the exact
workloads the throughput figures are quoted on:

* G(100, 0.5) real-valued, 1000 reads x 1000 sweeps, default geometric beta
  range: every read bit-exact against the oracle (states and energies).
* 512 BQMs (n even in [64, 4096], 15-regular, h ~ N(0, 0.1^2), J ~ N(0, 1)),
  256 reads each x 1000 sweeps, one batched launch (autotuned, hipGraph):
  reads 0, 127 and 255 of every BQM bit-exact against the oracle; every read a
  +-1 state whose reported energy equals its recomputed energy.
* Differentiable fixed-beta sampling, 64 x K16 + 8 x (n = 1024, 15-regular),
  4096 reads x 2000 Gibbs sweeps at beta = 1: four reads of every model
  bit-exact against the oracle; the backward sums equal the exact integer
  reduction of all samples.
"""
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from misamplers import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
import bench  # noqa: E402  (the benchmark's workload generators)

pytestmark = pytest.mark.gpu

WORKERS = 16


def _energies(h, cs, ce, cw, states):
    """Energies of every row of ``states`` (numpy order, float64)."""
    s = states.astype(np.float64)
    return s @ h + (s[:, cs] * s[:, ce]) @ cw


def test_gnp_full_config_every_read_bit_exact():
    from misamplers.sa.simulated_annealing import simulated_annealing
    w = bench.workload_gnp()
    assert w["R"] == 1000 and w["S"] == 1000 and w["n"] == 100
    seeds = _lib.read_seeds(12345, w["R"])
    print("gnp: GPU", flush=True)
    stats = {}
    st, en = simulated_annealing(w["R"], w["h"], w["cs"], w["ce"], w["cw"], 1, w["betas"], 0, w["init"].copy(),
                                 read_seeds=seeds, stats=stats)
    assert stats["mode"] == _lib.MODE_REAL
    print("gnp: oracle", flush=True)
    ref_st, ref_en = oracle.sa_sample(w["h"], w["cs"], w["ce"], w["cw"], 1, w["betas"], seeds, w["init"],
                                      nthreads=WORKERS)
    np.testing.assert_array_equal(st, ref_st)
    np.testing.assert_array_equal(en, ref_en)


def test_batch_full_config_against_oracle():
    import torch
    from misamplers.sa.batch import Batch
    w = bench.workload_batch()
    R, S, sizes = w["R"], w["S"], w["sizes"]
    assert len(sizes) == 512 and R == 256 and S == 1000
    print("batch: build + autotune", flush=True)
    batch = Batch(w["problems"], R, 1, w["schedules"])
    batch.autotune(4)
    init = np.concatenate([x.ravel() for x in w["inits"]])
    seeds = np.concatenate([_lib.read_seeds(s, R) for s in w["seeds"]])
    dev = torch.device("cuda:0")
    d_st = torch.from_numpy(init).to(dev)
    d_en = torch.empty(batch.total_reads, dtype=torch.float64, device=dev)
    d_sd = torch.from_numpy(seeds.view(np.int64)).to(dev)
    stream = torch.cuda.Stream()
    print("batch: GPU launch", flush=True)
    batch.run_device(d_st.data_ptr(), d_en.data_ptr(), d_sd.data_ptr(), stream=stream.cuda_stream, use_graph=True)
    stream.synchronize()
    st_all, en_all = d_st.cpu().numpy(), d_en.cpu().numpy()
    batch.close()

    picks = np.array([0, 127, 255])
    print("batch: oracle on 3 reads of every BQM", flush=True)

    def check(b):
        h, cs, ce, cw = w["problems"][b]
        n = int(sizes[b])
        so, ro = batch.state_offsets[b], batch.read_offsets[b]
        st = st_all[so:so + R * n].reshape(R, n)
        en = en_all[ro:ro + R]
        assert np.all(np.abs(st) == 1), f"BQM {b}: a spin is not +-1"
        scale = np.abs(h).sum() + np.abs(cw).sum()
        np.testing.assert_allclose(en, _energies(h, cs, ce, cw, st), rtol=0, atol=1e-10 * scale,
                                   err_msg=f"BQM {b}: energy differs from the recomputed one")
        ref_st, ref_en = oracle.sa_sample(h, cs, ce, cw, 1, w["schedules"][b], seeds[ro + picks],
                                          w["inits"][b][picks])
        np.testing.assert_array_equal(st[picks], ref_st, err_msg=f"BQM {b} (n={n}): states differ from the oracle")
        np.testing.assert_array_equal(en[picks], ref_en, err_msg=f"BQM {b} (n={n}): energies differ from the oracle")
        return b

    with ThreadPoolExecutor(WORKERS) as ex:
        done = list(ex.map(check, range(len(sizes))))
    assert done == list(range(len(sizes)))


def test_fixed_beta_full_config_against_oracle():
    import torch
    from misamplers.fixed_beta import FixedBetaSampler
    w = bench.workload_moments()
    nv, edges, R, S = w["nv"], w["edges"], w["R"], w["S"]
    assert len(nv) == 72 and R == 4096 and S == 2000
    dev = torch.device("cuda:0")
    fb = FixedBetaSampler(nv, edges, num_reads=R, num_sweeps=S, beta=1.0)
    th = torch.tensor(w["h"], device=dev, requires_grad=True)
    tj = torch.tensor(w["J"], device=dev, requires_grad=True)
    seed = 2024
    print("fixed beta: forward", flush=True)
    L, samples = fb(th, tj, seed=seed)
    print("fixed beta: backward", flush=True)
    L.sum().backward()
    _, _, sums = fb.moments(samples)
    torch.cuda.synchronize()
    smp = samples.cpu().numpy()
    sums = sums.cpu().numpy()
    fb.close()

    # the forward's inputs, as FixedBetaSampler.sample draws them
    total_reads = R * len(nv)
    init_all = _lib.random_spins(seed, R * sum(nv))
    seeds = _lib.read_seeds(seed, total_reads)
    hoff = np.concatenate([[0], np.cumsum(nv)])
    joff = np.concatenate([[0], np.cumsum([len(e) for e in edges])])
    soff = R * hoff
    picks = np.array([0, 1365, 2730, 4095])
    betas = np.ones(S)
    print("fixed beta: oracle on 4 reads of every model", flush=True)

    def check(b):
        n, e = nv[b], edges[b]
        h = w["h"][hoff[b]:hoff[b + 1]].astype(np.float64)
        J = w["J"][joff[b]:joff[b + 1]].astype(np.float64)
        st = smp[soff[b]:soff[b + 1]].reshape(R, n)
        init = init_all[soff[b]:soff[b + 1]].reshape(R, n)
        ref_st, _ = oracle.sa_sample(h, e[:, 0], e[:, 1], J, 1, betas, seeds[b * R + picks], init[picks],
                                     oracle.SEQUENTIAL, oracle.GIBBS)
        np.testing.assert_array_equal(st[picks], ref_st, err_msg=f"model {b} (n={n}) differs from the oracle")
        # exact sums of the backward pass: sum_r s_i, then sum_r s_u s_v
        s64 = st.astype(np.int64)
        o = hoff[b] + joff[b]
        np.testing.assert_array_equal(sums[o:o + n], s64.sum(0))
        np.testing.assert_array_equal(sums[o + n:o + n + len(e)], (s64[:, e[:, 0]] * s64[:, e[:, 1]]).sum(0))
        return b

    with ThreadPoolExecutor(WORKERS) as ex:
        done = list(ex.map(check, range(len(nv))))
    assert done == list(range(len(nv)))
