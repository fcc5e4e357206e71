"""GPU parity of the path-integral engine's chain updates (qubits_per_update
!= 1; the reference's chainUpdate, localPIMC.cpp:385-558), through the C ABI.

* Bit-exact against the oracle restatement (orc_pimc_sample_chains) on the
  same seeds, initial worldlines and chain colouring: every worldline word,
  flip count, statistics snapshot and energy.
* Equilibrium against the exact transfer matrix with the chain couplings.
* The drop-in ``PathIntegralAnnealingSampler`` with chain updates, read by
  read against the oracle, and its argument checks.
"""
import numpy as np
import pytest
from scipy import stats as sps

import oracle
from misamplers import BinaryQuadraticModel, PathIntegralAnnealingSampler, _lib
from misamplers.sqa.engine import PimcPlan, chain_colouring
from oracle import pimc_exact as px

pytestmark = pytest.mark.gpu


def _mask(P):
    return np.uint32(0xFFFFFFFF) if P == 32 else np.uint32((1 << P) - 1)


def _inter_chain_graph(rng, n, Q, p, integer=False):
    """Couplers between different chains only (chain couplers are implicit)."""
    cs, ce = [], []
    for u in range(n):
        for v in range(u + 1, n):
            if u // Q != v // Q and rng.random() < p:
                cs.append(u)
                ce.append(v)
    cw = rng.choice([-1.0, 1.0], size=len(cs)) if integer else rng.uniform(-1, 1, size=len(cs))
    return rng.uniform(-1, 1, size=n), cs, ce, list(cw)


def run_chains(h, cs, ce, cw, spb, hp, hd, seeds, W0, P, Q, jc, interval=0):
    import torch
    plan = PimcPlan(h, cs, ce, cw, spb, hp, hd, P, qubits_per_chain=Q, chain_coupler_strength=jc)
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
    out = (dW.cpu().numpy().view(np.uint32), dstat.cpu().numpy()[:, :ncol] if interval else None,
           dst.cpu().numpy(), den.cpu().numpy(), dfl.cpu().numpy().view(np.uint64), info)
    plan.close()
    return out


def check_exact(h, cs, ce, cw, spb, hp, hd, R, P, Q, jc, seed, interval=0, classical=False):
    rng = np.random.default_rng(seed)
    n = len(h)
    W0 = rng.integers(0, 2**32, size=(R, n), dtype=np.uint64).astype(np.uint32) & _mask(P)
    if classical:
        W0 = np.where(W0 & 1, _mask(P), np.uint32(0)).astype(np.uint32)
    seeds = oracle.read_seeds(seed, R)
    print(f"chain case n={n} m={len(cs)} R={R} P={P} Q={Q} jc={jc} steps={len(hp)} spb={spb} "
          f"interval={interval}", flush=True)
    W, stats, st, en, flips, info = run_chains(h, cs, ce, cw, spb, hp, hd, seeds, W0, P, Q, jc, interval)
    nc, col = chain_colouring(h, cs, ce, cw, Q)
    assert info["kind"] == "chain" and info["num_classes"] == nc
    Wo, so, fo = oracle.pimc_sample_chains(h, cs, ce, cw, spb, hp, hd, seeds, W0, Q, jc, num_slices=P,
                                           interval=interval, nthreads=8, return_flips=True, chain_colour=col)
    np.testing.assert_array_equal(W, Wo)
    np.testing.assert_array_equal(flips, fo)
    if interval:
        np.testing.assert_array_equal(stats, so)
    np.testing.assert_array_equal(st, np.where(Wo & 1, 1, -1).astype(np.int8))
    for r in range(min(R, 16)):
        assert en[r] == oracle.state_energy(st[r], h, cs, ce, cw)
    return W0, W


@pytest.mark.parametrize("n,Q,P,jc,p,integer", [
    (6, 3, 32, -1.0, 0.5, False),
    (64, 4, 16, -0.8, 0.08, False),
    (64, 4, 32, -1.5, 0.08, True),
    (90, 3, 8, -0.5, 0.05, False),
    (50, 4, 32, -1.0, 0.1, False),      # qubits 48, 49 are in no chain
    (40, 1, 32, -1.0, 0.15, False),     # chains of one qubit
    (64, 16, 32, -0.7, 0.2, False),     # 512 segments at most per chain update
    (96, 32, 16, -0.9, 0.1, False),     # the longest chains
    (60, 4, 12, 0.6, 0.1, False),       # antiferromagnetic chain coupling never joins
])
def test_chain_updates_bit_exact(n, Q, P, jc, p, integer):
    rng = np.random.default_rng(n * 100 + Q * 10 + P)
    h, cs, ce, cw = _inter_chain_graph(rng, n, Q, p, integer)
    hp, hd = np.linspace(0.1, 3.0, 12), np.linspace(3.0, 0.05, 12)
    W0, W = check_exact(h, cs, ce, cw, 1, hp, hd, 48, P, Q, jc, 1000 + n)
    nch = n // Q
    np.testing.assert_array_equal(W[:, nch * Q:], W0[:, nch * Q:])


def test_chain_updates_cut_laws_and_statistics():
    """Hd = 0 (kinks only), tanh(Hd/P) = 1 (every bond cut), two sweeps per
    step and statistics every third step."""
    rng = np.random.default_rng(3)
    h, cs, ce, cw = _inter_chain_graph(rng, 48, 4, 0.1)
    hp = np.linspace(0.2, 2.5, 10)
    hd = np.array([0.0, 0.0, 1.0, 1e9, 2.0, 0.5, 0.0, 1e9, 3.0, 0.1])
    check_exact(h, cs, ce, cw, 2, hp, hd, 40, 32, 4, -1.2, 7, interval=3)
    check_exact(h, cs, ce, cw, 1, hp, np.zeros(10), 40, 32, 4, -1.2, 8, classical=True)


def test_chain_updates_reads_beyond_one_block():
    rng = np.random.default_rng(4)
    h, cs, ce, cw = _inter_chain_graph(rng, 32, 4, 0.15)
    hp, hd = np.full(6, 1.0), np.full(6, 1.0)
    check_exact(h, cs, ce, cw, 1, hp, hd, 700, 16, 4, -1.0, 9)


def _chi2_pvalue(idx, prob, R):
    counts = np.bincount(idx, minlength=len(prob)).astype(float)
    exp = prob * R
    keep = exp > 5
    obs = np.append(counts[keep], counts[~keep].sum())
    ex = np.append(exp[keep], exp[~keep].sum())
    if ex[-1] == 0:
        obs, ex = obs[:-1], ex[:-1]
    return sps.chi2.sf(((obs - ex) ** 2 / ex).sum(), len(obs) - 1)


def test_chain_equilibrium_transfer_matrix():
    h = np.array([0.3, -0.2, 0.45, -0.1, 0.05, 0.25])
    cs, ce, cw = [0, 1, 2, 0, 2], [3, 4, 5, 5, 3], [0.7, -1.0, 0.5, 0.4, -0.6]
    R, P, hp, hd, jc = 40000, 8, 1.0, 1.2, -0.9
    rng = np.random.default_rng(5)
    W0 = rng.integers(0, 2**32, size=(R, 6), dtype=np.uint64).astype(np.uint32) & _mask(P)
    W, *_ = run_chains(h, cs, ce, cw, 1, np.full(40, hp), np.full(40, hd), oracle.read_seeds(5, R), W0, P, 3, jc)
    ccs, cce, ccw = px.chain_couplers(6, 3, jc)
    tm = px.TrotterModel(h, cs + ccs, ce + cce, cw + ccw, hp, hd, P)
    assert _chi2_pvalue(px.state_index(np.where(W & 1, 1, -1)), tm.slice_distribution(), R) > 1e-4


def test_sampler_chain_updates_equal_oracle_read_by_read():
    rng = np.random.default_rng(6)
    n, R, seed, Q, P, jc = 64, 24, 93, 4, 32, -1.3
    h, cs, ce, cw = _inter_chain_graph(rng, n, Q, 0.08)
    bqm = BinaryQuadraticModel({i: h[i] for i in range(n)}, {(u, v): w for u, v, w in zip(cs, ce, cw)}, 0.5, "SPIN")
    init = rng.choice([-1, 1], size=(R, n)).astype(np.int8)
    hp, hd = np.linspace(0.1, 3.0, 30), np.linspace(3.0, 0.05, 30)
    ss = PathIntegralAnnealingSampler().sample(bqm, num_reads=R, seed=seed, initial_states=(init, list(range(n))),
                                               beta_schedule_type="custom", Hp_field=hp, Hd_field=hd,
                                               qubits_per_chain=Q, qubits_per_update=Q, chain_coupler_strength=jc)
    hv, (bcs, bce, bcw), off = bqm.to_numpy_vectors(variable_order=list(range(n)))
    W0 = np.where(init > 0, np.uint32(0xFFFFFFFF), np.uint32(0)).astype(np.uint32)
    col = chain_colouring(hv, bcs, bce, bcw, Q)[1]
    Wo, _ = oracle.pimc_sample_chains(hv, bcs, bce, bcw, 1, hp, hd, _lib.read_seeds(seed, R), W0, Q, jc,
                                      num_slices=P, chain_colour=col)
    want = np.where(Wo & 1, 1, -1).astype(np.int8)
    np.testing.assert_array_equal(ss.record.sample, want)
    want_en = np.array([oracle.state_energy(s, hv, bcs, bce, bcw) for s in want]) + off
    np.testing.assert_array_equal(ss.record.energy, want_en)


def test_sampler_chain_argument_checks():
    bqm = BinaryQuadraticModel({0: 0.1, 1: -0.2, 2: 0.3, 3: 0.0}, {(0, 2): 1.0}, 0.0, "SPIN")
    with pytest.raises(ValueError, match="qubits_per_chain must be at least qubits_per_update"):
        PathIntegralAnnealingSampler().sample(bqm, num_reads=2, qubits_per_chain=2, qubits_per_update=4)
    with pytest.raises(_lib.EngineError, match="chain updates support"):
        PathIntegralAnnealingSampler().sample(bqm, num_reads=2, qubits_per_chain=33, qubits_per_update=2)
