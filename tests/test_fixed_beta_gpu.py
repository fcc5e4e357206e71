"""GPU checks of differentiable fixed-beta sampling (misamplers.fixed_beta).

* Forward: every read bit-identical to the oracle's Gibbs sweeps at constant
  beta (the reference's simulated_annealing with a custom schedule).
* Backward: the moment kernel equals the exact integer reduction of the same
  samples, and float32 gradients equal its scaled value exactly.
* Statistics: K16 moments against exact Boltzmann enumeration.
"""
import math

import numpy as np
import pytest

import oracle
from conftest import random_regular
from misamplers import _lib
from misamplers.fixed_beta import FixedBetaSampler
from oracle import pimc_exact as px

pytestmark = pytest.mark.gpu


def _models(rng, specs):
    """specs: list of ('k', n) complete graphs or ('r', n, d) regular graphs."""
    nv, edges, hs, js = [], [], [], []
    for sp in specs:
        if sp[0] == "k":
            n = sp[1]
            e = np.array([(u, v) for u in range(n) for v in range(u + 1, n)], np.int32).reshape(-1, 2)
        elif sp[0] == "r":
            n = sp[1]
            cs, ce = random_regular(n, sp[2], rng)
            e = np.stack([cs, ce], 1).astype(np.int32)
        else:  # ('x', n, edges)
            n, e = sp[1], np.asarray(sp[2], np.int32).reshape(-1, 2)
        nv.append(n)
        edges.append(e)
        hs.append(rng.normal(0, 0.5, n))
        js.append(rng.normal(0, 0.5, len(e)))
    return nv, edges, np.concatenate(hs).astype(np.float32), np.concatenate(js).astype(np.float32)


def _split(x, sizes):
    return np.split(np.asarray(x), np.cumsum(sizes)[:-1])


def test_forward_bit_exact_and_backward_exact():
    import torch
    rng = np.random.default_rng(1)
    nv, edges, h, J = _models(rng, [("k", 16), ("x", 7, [(0, 1), (1, 2), (2, 2), (3, 6), (0, 1)]),
                                    ("r", 64, 3), ("x", 5, [])])
    R, S, beta, seed = 100, 30, 0.8, 77
    fb = FixedBetaSampler(nv, edges, num_reads=R, num_sweeps=S, beta=beta)
    th = torch.tensor(h, device="cuda:0", requires_grad=True)
    tj = torch.tensor(J, device="cuda:0", requires_grad=True)
    print("forward", flush=True)
    L, samples = fb(th, tj, seed=seed)
    g = torch.tensor([1.0, -2.0, 0.5, 3.0], device="cuda:0")
    print("backward", flush=True)
    (L * g).sum().backward()
    torch.cuda.synchronize()
    S_all = samples.cpu().numpy()
    # initial states and seeds as the sampler draws them
    init = _lib.random_spins(seed, S_all.size)
    seeds = _lib.read_seeds(seed, R * len(nv))
    so = 0
    L_np = L.detach().cpu().numpy()
    gh, gj = th.grad.cpu().numpy(), tj.grad.cpu().numpy()
    for b, (n, e) in enumerate(zip(nv, edges)):
        hb = _split(h, nv)[b].astype(np.float64)
        jb = _split(J, [len(x) for x in edges])[b].astype(np.float64)
        st, en = oracle.sa_sample(hb, e[:, 0], e[:, 1], jb, 1, np.full(S, beta), seeds[b * R:(b + 1) * R],
                                  init[so:so + R * n].reshape(R, n), oracle.SEQUENTIAL, oracle.GIBBS, nthreads=8)
        got = S_all[so:so + R * n].reshape(R, n)
        np.testing.assert_array_equal(got, st)
        assert np.isclose(L_np[b], en.mean(), rtol=1e-6, atol=1e-6)
        s = got.astype(np.int64)
        sum_h = s.sum(0)
        sum_j = (s[:, e[:, 0]] * s[:, e[:, 1]]).sum(0) if len(e) else np.zeros(0, np.int64)
        invR = np.float32(1.0) / np.float32(R)
        want_h = np.float32(g[b].item()) * (sum_h.astype(np.float32) * invR)
        want_j = np.float32(g[b].item()) * (sum_j.astype(np.float32) * invR)
        np.testing.assert_array_equal(_split(gh, nv)[b], want_h)
        np.testing.assert_array_equal(_split(gj, [len(x) for x in edges])[b], want_j)
        so += R * n
    fb.close()


def test_moment_sums_exact_large():
    """Exact int64 sums on the benchmark's sparse shape (n=1024, 15-regular)."""
    import torch
    rng = np.random.default_rng(2)
    cs, ce = random_regular(1024, 15, rng)
    e = np.stack([cs, ce], 1).astype(np.int32)
    R = 4096 + 37  # a partial last tile: one tile per block
    k16 = np.array([(u, v) for u in range(16) for v in range(u + 1, 16)])
    # the second n = 1024 model has the benchmark's 4096 reads
    fb = FixedBetaSampler([1024, 1024, 16], [e, e, k16], num_reads=[R, 4096, 333])
    s1 = rng.choice(np.array([-1, 1], np.int8), size=(R, 1024))
    s3 = np.where(rng.random((4096, 1024)) < 0.5 + 0.45 * rng.random(1024), 1, -1).astype(np.int8)
    s2 = rng.choice(np.array([-1, 1], np.int8), size=(333, 16))
    samples = torch.from_numpy(np.concatenate([s1.ravel(), s3.ravel(), s2.ravel()])).to("cuda:0")
    gh, gj, sums = fb.moments(samples)
    torch.cuda.synchronize()
    sums = sums.cpu().numpy()
    want = np.concatenate([s1.astype(np.int64).sum(0), (s1[:, cs].astype(np.int64) * s1[:, ce]).sum(0),
                           s3.astype(np.int64).sum(0), (s3[:, cs].astype(np.int64) * s3[:, ce]).sum(0),
                           s2.astype(np.int64).sum(0), (s2[:, k16[:, 0]].astype(np.int64) * s2[:, k16[:, 1]]).sum(0)])
    np.testing.assert_array_equal(sums, want)
    fb.close()


def test_moment_sums_exact_shapes():
    """Sizes that take every path of the moment kernel: variables not a
    multiple of 16 (byte loads), one to four tiles per block, partial tiles,
    models with no edges or no reads, blocks of up to 16 tiles of small
    models (edges staged in LDS up to 1024 of them, read from memory above),
    and blocks of one tile of n >= 256 variables (four waves each) on the
    staged path (the first three models start at 16-byte aligned sample
    offsets) and on the byte-load path (n = 300)."""
    import torch
    rng = np.random.default_rng(4)
    specs = [(1, 0, 1), (17, 30, 63), (33, 60, 65), (48, 100, 130), (256, 700, 200), (5, 4, 0), (64, 0, 64),
             (2000, 9000, 70),  # more edges than the 4096 a block loads up front
             (16, 120, 1024), (16, 1500, 1100), (12, 66, 700), (16, 1024, 300)]
    # aligned: staged rows, one tile per block (a partial last tile, whole
    # tiles, two tiles); then byte loads of whole tiles
    specs = [(512, 3000, 250), (512, 3000, 256), (300, 800, 384), (256, 700, 128)] + specs
    nv, edges, reads, parts, want = [], [], [], [], []
    for n, m, R in specs:
        e = rng.integers(0, n, size=(m, 2)).astype(np.int32)
        s = rng.choice(np.array([-1, 1], np.int8), size=(R, n))
        nv.append(n)
        edges.append(e)
        reads.append(R)
        parts.append(s.ravel())
        s64 = s.astype(np.int64)
        want.append(np.concatenate([s64.sum(0), (s64[:, e[:, 0]] * s64[:, e[:, 1]]).sum(0)]))
    print("moment shapes", specs, flush=True)
    fb = FixedBetaSampler(nv, edges, num_reads=reads)
    samples = torch.from_numpy(np.concatenate(parts)).to("cuda:0")
    gh, gj, sums = fb.moments(samples)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(sums.cpu().numpy(), np.concatenate(want))
    fb.close()


def test_forward_new_values_every_call():
    """The batch is built once: every call writes its own h and J on the
    device, draws its seeds and initial states there, and equals the oracle;
    given initial states are taken as they are."""
    import torch
    rng = np.random.default_rng(9)
    nv, edges, h, J = _models(rng, [("r", 64, 3), ("x", 7, [(0, 1), (1, 2), (2, 2), (3, 6), (0, 1), (6, 0)]),
                                    ("k", 16)])
    R, S, beta = 50, 20, 1.3
    fb = FixedBetaSampler(nv, edges, num_reads=R, num_sweeps=S, beta=beta)
    hoff = np.concatenate([[0], np.cumsum(nv)])
    joff = np.concatenate([[0], np.cumsum([len(e) for e in edges])])
    for call in range(3):
        hc = (h * (1 + call) - 0.1 * call).astype(np.float32)
        Jc = (J[::-1].copy() if call % 2 else J).astype(np.float32)
        init = None
        if call == 2:
            init = rng.choice(np.array([-1, 1], np.int8), size=R * sum(nv))
        print(f"call {call}", flush=True)
        en, smp = fb.sample(torch.tensor(hc, device="cuda:0"), torch.tensor(Jc, device="cuda:0"), seed=100 + call,
                            initial_states=None if init is None else torch.from_numpy(init).to("cuda:0"))
        torch.cuda.synchronize()
        en, smp = en.cpu().numpy(), smp.cpu().numpy()
        start = _lib.random_spins(100 + call, R * sum(nv)) if init is None else init
        seeds = _lib.read_seeds(100 + call, R * len(nv))
        for b, (n, e) in enumerate(zip(nv, edges)):
            so = R * hoff[b]
            st, eo = oracle.sa_sample(hc[hoff[b]:hoff[b + 1]].astype(np.float64), e[:, 0], e[:, 1],
                                      Jc[joff[b]:joff[b + 1]].astype(np.float64), 1, np.full(S, beta),
                                      seeds[b * R:(b + 1) * R], start[so:so + R * n].reshape(R, n),
                                      oracle.SEQUENTIAL, oracle.GIBBS, nthreads=8)
            np.testing.assert_array_equal(smp[so:so + R * n].reshape(R, n), st)
            np.testing.assert_array_equal(en[b * R:(b + 1) * R], eo)
    fb.close()


def test_device_seed_and_spin_generators():
    import torch
    d = torch.empty(1000, dtype=torch.int64, device="cuda:0")
    s = torch.empty(777, dtype=torch.int8, device="cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    import ctypes
    _lib.check(_lib.lib.misa_read_seeds_device(ctypes.c_uint64(2**63 + 5), 1000, ctypes.c_void_p(d.data_ptr()),
                                               ctypes.c_void_p(stream)))
    _lib.check(_lib.lib.misa_random_spins_device(ctypes.c_uint64(12345), 777, ctypes.c_void_p(s.data_ptr()),
                                                 ctypes.c_void_p(stream)))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d.cpu().numpy().view(np.uint64), _lib.read_seeds(2**63 + 5, 1000))
    np.testing.assert_array_equal(s.cpu().numpy(), _lib.random_spins(12345, 777))


def test_k16_moments_match_enumeration():
    import torch
    rng = np.random.default_rng(3)
    nv, edges, h, J = _models(rng, [("k", 16), ("k", 16)])
    R, beta = 4096, 1.0
    fb = FixedBetaSampler(nv, edges, num_reads=R, num_sweeps=200, beta=beta)
    th = torch.tensor(h, device="cuda:0", requires_grad=True)
    tj = torch.tensor(J, device="cuda:0", requires_grad=True)
    L, samples = fb(th, tj, seed=5)
    L.sum().backward()
    gh, gj = th.grad.cpu().numpy(), tj.grad.cpu().numpy()
    S_all = samples.cpu().numpy()
    for b in range(2):
        e = edges[b]
        hb = h[16 * b:16 * (b + 1)].astype(np.float64)
        jb = J[120 * b:120 * (b + 1)].astype(np.float64)
        p = px.boltzmann(hb, e[:, 0], e[:, 1], jb, beta)
        Ssp = px.spins(16).astype(float)
        mh = p @ Ssp
        mj = p @ (Ssp[:, e[:, 0]] * Ssp[:, e[:, 1]])
        s = S_all[R * 16 * b:R * 16 * (b + 1)].reshape(R, 16).astype(float)
        sd_h = np.sqrt(np.maximum(1 - mh ** 2, 1e-3) / R)
        sd_j = np.sqrt(np.maximum(1 - mj ** 2, 1e-3) / R)
        assert np.max(np.abs(gh[16 * b:16 * (b + 1)] - mh) / sd_h) < 5.5
        assert np.max(np.abs(gj[120 * b:120 * (b + 1)] - mj) / sd_j) < 5.5
        assert abs(L[b].item() - float(p @ px.energies(hb, e[:, 0], e[:, 1], jb))) < 0.05 * 16
    fb.close()
