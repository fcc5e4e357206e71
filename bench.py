#!/usr/bin/env python3
"""Spin-update throughput of the MI355X simulated-annealing engine.

Workload (one "step" = one full anneal of one batch, inputs resident in HBM):
3D cubic +-J lattice L=16 (4096 spins, periodic, 12288 couplers, J in {-1,+1}
equiprobable, h=0), 4096 reads, 2000 sweeps, geometric beta 0.1 -> 5.0,
sequential Metropolis; 3.36e10 spin-update attempts per step.

Prints one JSON line with the throughput, a roofline entry for the sweep kernel
(algorithmic bytes per launch / kernel time) and the reference CPU core timed
on this host's cores on a bounded sample of the same reads (bit-exactness of
those reads is checked and reported).

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dwave-samplers_amd"))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
# LDS: 256 CUs x 256 B/clk (128-bit reads) x 2.4 GHz = 157.3 TB/s; the engine's own
# microbenchmark (misa_lds_bandwidth) measures what the device delivers, reported beside it
LDS_PEAK_GBS = 256 * 256 * 2.4


_LDS_MEASURED = []


def lds_peak_measured():
    """The device's LDS read bandwidth in GB/s, measured once per run with the
    engine's own microbenchmark (misa_lds_bandwidth)."""
    if not _LDS_MEASURED:
        from misamplers import _lib
        _LDS_MEASURED.append(_lib.lds_bandwidth(4096)[0])
    return _LDS_MEASURED[0]


def lds_fields(achieved):
    """Fractions of the spec LDS peak and of the measured one."""
    meas = lds_peak_measured()
    return {"lds_peak": LDS_PEAK_GBS, "lds_frac": achieved / LDS_PEAK_GBS, "lds_peak_measured": meas,
            "lds_frac_measured": achieved / meas}


def on_chip_roofline(achieved, kernel, kern_ms, alg_bytes):
    """Roofline entry for a kernel whose algorithmic bytes are LDS accesses
    (per-read state resident in LDS): priced against the LDS peak, the spec
    figure and the device's measured one."""
    out = {"bound": "lds", "achieved": achieved, "peak": LDS_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / LDS_PEAK_GBS, "traffic": None, "kernel": kernel, "kernel_ms": kern_ms,
           "alg_bytes_per_launch": alg_bytes, "hbm_frac_equivalent": achieved / HBM_PEAK_GBS}
    out.update(lds_fields(achieved))
    return out


def profile_figures(name):
    """Machine-readable ``key: value`` lines of a committed rocprofv3 summary
    under profiles/ (counter figures that a run cannot collect itself)."""
    out = {}
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            t = line.strip()
            if ":" in t and t.split(":", 1)[0].replace("_", "").isalnum() and " " not in t.split(":", 1)[0]:
                k, v = t.split(":", 1)
                try:
                    out[k] = float(v)
                except ValueError:
                    pass
    except OSError:
        pass
    return out


def lattice3d(L):
    cs, ce = [], []
    for x in range(L):
        for y in range(L):
            for z in range(L):
                v = x * L * L + y * L + z
                for dx, dy, dz in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
                    u = ((x + dx) % L) * L * L + ((y + dy) % L) * L + (z + dz) % L
                    cs.append(v)
                    ce.append(u)
    return np.array(cs, np.int32), np.array(ce, np.int32)


def workload(num_reads=4096, num_sweeps=2000, L=16):
    """3D +-J lattice, geometric beta 0.1 -> 5.0 (integer biases)."""
    rng = np.random.default_rng(20241203)
    cs, ce = lattice3d(L)
    cw = rng.choice([-1.0, 1.0], len(cs))
    n = L ** 3
    h = np.zeros(n)
    betas = np.geomspace(0.1, 5.0, num_sweeps)
    init = (2 * rng.integers(0, 2, (num_reads, n), dtype=np.int8) - 1).astype(np.int8)
    return dict(h=h, cs=cs, ce=ce, cw=cw, betas=betas, init=init, n=n, R=num_reads, S=num_sweeps,
                name=f"SA 3D +-J cubic lattice L={L} ({n} spins), {num_reads} reads x {num_sweeps} sweeps, "
                     "geometric beta 0.1->5.0, sequential Metropolis")


def workload_gnp(num_reads=1000, num_sweeps=1000, n=100, p=0.5):
    """G(n, p) with h, J ~ U[-1, 1] float64, default geometric beta range."""
    sys.path.insert(0, os.path.join(ROOT, "dwave-samplers_amd"))
    from misamplers.sa.sampler import _default_ising_beta_range
    rng = np.random.default_rng(20241204)
    pairs = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < p]
    cs = np.array([a for a, _ in pairs], np.int32)
    ce = np.array([b for _, b in pairs], np.int32)
    h = rng.uniform(-1, 1, n)
    cw = rng.uniform(-1, 1, len(pairs))
    hot, cold = _default_ising_beta_range(dict(enumerate(h)), {(int(a), int(b)): w for a, b, w in zip(cs, ce, cw)})
    betas = np.geomspace(hot, cold, num_sweeps)
    init = (2 * rng.integers(0, 2, (num_reads, n), dtype=np.int8) - 1).astype(np.int8)
    return dict(h=h, cs=cs, ce=ce, cw=cw, betas=betas, init=init, n=n, R=num_reads, S=num_sweeps,
                name=f"SA G({n},{p}) real-valued h,J~U[-1,1], {num_reads} reads x {num_sweeps} sweeps, "
                     "default geometric beta range, sequential Metropolis")


def random_regular(n, d, rng):
    """Edges of a random d-regular simple graph (configuration model; invalid
    pairs re-drawn until none remain)."""
    stubs = np.repeat(np.arange(n), d)
    rng.shuffle(stubs)
    pairs = stubs.reshape(-1, 2)
    for _ in range(1000):
        a, b = np.minimum(pairs[:, 0], pairs[:, 1]), np.maximum(pairs[:, 0], pairs[:, 1])
        _, first = np.unique(a.astype(np.int64) * n + b, return_index=True)
        bad = np.ones(len(pairs), bool)
        bad[first] = False
        bad |= a == b
        if not bad.any():
            return a.astype(np.int32), b.astype(np.int32)
        sel = np.unique(np.concatenate([np.flatnonzero(bad), rng.choice(len(pairs), int(bad.sum()), replace=False)]))
        flat = pairs[sel].ravel()
        rng.shuffle(flat)
        pairs[sel] = flat.reshape(-1, 2)
    raise RuntimeError("could not build a simple regular graph")


def default_beta_range_np(h, cs, ce, cw, rate=0.01):
    """Vectorised restatement of the sampler's default beta range (same
    arithmetic as misamplers.sa.sampler._default_ising_beta_range)."""
    n = len(h)
    sum_abs = np.abs(h).astype(float).copy()
    np.add.at(sum_abs, cs, np.abs(cw))
    np.add.at(sum_abs, ce, np.abs(cw))
    big = np.inf
    min_abs = np.where(h != 0, np.abs(h), big)
    nz = cw != 0
    np.minimum.at(min_abs, cs[nz], np.abs(cw[nz]))
    np.minimum.at(min_abs, ce[nz], np.abs(cw[nz]))
    vals = min_abs[np.isfinite(min_abs)]
    hot = np.log(2) / (2 * sum_abs.max())
    mn = vals.min()
    cold = np.log(np.sum(vals == mn) / rate) / (2 * mn)
    return hot, cold


def workload_batch(num_bqms=512, num_reads=256, num_sweeps=1000):
    """512 BQMs, n uniform over even integers in [64, 4096], random 15-regular
    graphs, h ~ N(0, 0.1^2), J ~ N(0, 1) float64, default geometric schedule per BQM."""
    rng = np.random.default_rng(20241205)
    sizes = 2 * rng.integers(32, 2049, num_bqms)
    problems, schedules, inits, seeds = [], [], [], []
    for i, n in enumerate(sizes):
        cs, ce = random_regular(int(n), 15, rng)
        h = rng.normal(0, 0.1, int(n))
        cw = rng.normal(0, 1, len(cs))
        hot, cold = default_beta_range_np(h, cs, ce, cw)
        problems.append((h, cs, ce, cw))
        schedules.append(np.geomspace(hot, cold, num_sweeps))
        inits.append((2 * rng.integers(0, 2, (num_reads, int(n)), dtype=np.int8) - 1).astype(np.int8))
        seeds.append(100 + i)
    return dict(problems=problems, schedules=schedules, inits=inits, seeds=seeds, sizes=sizes, R=num_reads,
                S=num_sweeps, name=f"batched SA: {num_bqms} BQMs, n uniform even in [64,4096], 15-regular, "
                                   f"h~N(0,0.1^2), J~N(0,1), {num_reads} reads each x {num_sweeps} sweeps, "
                                   "default geometric schedule per BQM, sequential Metropolis")


def host_cpu():
    """Model name and logical CPUs of the host (the CPU baseline's machine)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count()}


def host_cores():
    """Every host core this process may run on (the CPU baseline's threads)."""
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except Exception:
        return max(1, os.cpu_count() or 1)


def cpu_baseline(w, seeds, gpu_states, gpu_energies, target_s=12.0):
    """Reference core (oracle/_ref) or the oracle port, read by read, on a
    bounded sample of the same reads; checks those reads against the GPU."""
    import oracle
    kind = "reference" if oracle.have_ref() else "port"
    run = oracle.ref_sa_read_by_read if kind == "reference" else oracle.sa_sample
    cores = host_cores()
    n, S = w["n"], w["S"]
    # calibrate with one read per thread, then size the sample to ~target_s
    k = cores
    t0 = time.perf_counter()
    run(w["h"], w["cs"], w["ce"], w["cw"], 1, w["betas"], seeds[:k], w["init"][:k], 0, 1, nthreads=cores)
    t1 = time.perf_counter() - t0
    rounds = max(1, int(target_s / max(t1, 1e-3)))
    k = min(w["R"], cores * rounds)
    pick = np.linspace(0, w["R"] - 1, k).astype(int)
    t0 = time.perf_counter()
    st, en = run(w["h"], w["cs"], w["ce"], w["cw"], 1, w["betas"], seeds[pick], w["init"][pick], 0, 1,
                 nthreads=cores)
    dt = time.perf_counter() - t0
    exact = bool(np.array_equal(st, gpu_states[pick]) and np.array_equal(en, gpu_energies[pick]))
    attempts = k * S * n
    return {"value": attempts / dt, "unit": "spin-update attempts/s", "cores": cores, "kind": kind,
            "sample": f"{k} of {w['R']} reads x {S} sweeps x {n} spins ({attempts:.3g} attempts, {dt:.1f} s); "
                      f"GPU bit-exact on these reads: {exact}",
            "bit_exact": exact}


def workload_pimc(num_reads=1024, num_steps=1000, n=2048, d=6, P=32):
    """Path-integral annealing: random 6-regular graph on 2048 spins, J = +-1
    equiprobable, h ~ U[-0.5, 0.5], Hd 10 -> 0 and Hp 0 -> 10 linearly over
    1000 steps of one sweep, 32 Trotter slices, random classical initial states."""
    rng = np.random.default_rng(4096)
    cs, ce = random_regular(n, d, rng)
    cw = rng.choice([-1.0, 1.0], len(cs))
    h = rng.uniform(-0.5, 0.5, n)
    init = rng.choice(np.array([-1, 1], np.int8), size=(num_reads, n))
    return {"n": n, "R": num_reads, "S": num_steps, "P": P, "h": h, "cs": cs, "ce": ce, "cw": cw,
            "hp": np.linspace(0.0, 10.0, num_steps), "hd": np.linspace(10.0, 0.0, num_steps), "init": init,
            "name": f"path-integral annealing: random {d}-regular n={n} (m={len(cs)}), J=+-1, h~U[-0.5,0.5], "
                    f"{P} Trotter slices, Hd 10->0 / Hp 0->10 over {num_steps} steps x 1 sweep, "
                    f"{num_reads} reads"}


def pimc_cpu_baseline(w, gpu_energies, target_s=15.0):
    """The reference's continuous-time core (oracle/_ref/libref_pimc.so) on a
    bounded sample of the reads (one call per read, one thread per call), or
    the discrete oracle port when the reference library is absent."""
    import oracle
    from scipy import stats as sps
    cores = host_cores()
    kind = "reference" if oracle.have_ref_pimc() else "port"
    n, S = w["n"], w["S"]

    def run(k):
        pick = np.linspace(0, w["R"] - 1, k).astype(int)
        t0 = time.perf_counter()
        if kind == "reference":
            _, en, _ = oracle.ref_pimc_call(w["h"], w["cs"], w["ce"], w["cw"], 1, w["hp"], w["hd"], w["init"][pick],
                                            np.arange(1, k + 1, dtype=np.uint32), nthreads=cores)
        else:
            P = w["P"]
            mask = np.uint32(0xFFFFFFFF) if P == 32 else np.uint32((1 << P) - 1)
            W0 = np.where(w["init"][pick] > 0, mask, np.uint32(0)).astype(np.uint32)
            W, _ = oracle.pimc_sample(w["h"], w["cs"], w["ce"], w["cw"], 1, w["hp"], w["hd"],
                                      oracle.read_seeds(7, k), W0, num_slices=P, nthreads=cores)
            st = np.where(W & 1, 1, -1).astype(np.int8)
            en = np.array([oracle.state_energy(x, w["h"], w["cs"], w["ce"], w["cw"]) for x in st])
        return time.perf_counter() - t0, en

    t1, _ = run(cores)
    rounds = max(1, int(target_s / max(t1, 1e-3)))
    k = min(w["R"], cores * rounds)
    dt, en = run(k)
    updates = k * S * n
    ks = sps.ks_2samp(en, gpu_energies)
    return {"value": updates / dt, "unit": "worldline (site) updates/s", "cores": cores, "kind": kind,
            "sample": f"{k} of {w['R']} reads x {S} steps x {n} sites ({updates:.3g} site updates, {dt:.1f} s), "
                      "continuous imaginary time (2^16 resolution), one call and one seed per read",
            "energy_report": {"reference_mean": float(np.mean(en)), "reference_min": float(np.min(en)),
                              "gpu_mean": float(np.mean(gpu_energies)), "gpu_min": float(np.min(gpu_energies)),
                              "ks_statistic": float(ks.statistic), "ks_pvalue": float(ks.pvalue)}}


def run_pimc(args):
    """Discrete-slice path-integral anneal: one step = the full schedule for
    all reads (sweep kernel + projection / energies), worldlines resident."""
    import torch
    from misamplers.sqa.engine import PimcPlan
    from misamplers import _lib
    w = workload_pimc(args.reads if args.reads != 4096 else 1024, args.sweeps if args.sweeps != 2000 else 1000,
                      P=args.slices)
    R, n, S, P = w["R"], w["n"], w["S"], w["P"]
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    plan = PimcPlan(w["h"], w["cs"], w["ce"], w["cw"], 1, w["hp"], w["hd"], P)
    mask = np.uint32(0xFFFFFFFF) if P == 32 else np.uint32((1 << P) - 1)
    W0 = np.where(w["init"] > 0, mask, np.uint32(0)).astype(np.uint32)
    seeds = _lib.read_seeds(4321, R)
    dW0 = torch.from_numpy(W0.view(np.int32)).to(dev)
    dW = torch.empty_like(dW0)
    dS = torch.from_numpy(seeds.view(np.int64)).to(dev)
    dst = torch.empty((R, n), dtype=torch.int8, device=dev)
    den = torch.empty(R, dtype=torch.float64, device=dev)
    dfl = torch.zeros(R, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()
    kev = []

    def step(record):
        dW.copy_(dW0)
        dfl.zero_()
        if record:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
        plan.run_device(dW.data_ptr(), dS.data_ptr(), R, 0, 0, stream.cuda_stream, d_flips=dfl.data_ptr())
        if record:
            e1.record(stream)
            kev.append((e0, e1))
        plan.project_device(dW.data_ptr(), R, dst.data_ptr(), den.data_ptr(), stream.cuda_stream)

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        ev[i][0].record(stream)
        step(True)
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    step_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in kev]))
    info = plan.info(R)
    flips = int(dfl.cpu().numpy().sum())
    attempts = R * S * n * P
    site_updates = R * S * n
    dbar = 2.0 * len(w["cs"]) / n
    alg_bytes = attempts * (11.0 + dbar) + flips
    achieved = alg_bytes / (kern_ms / 1e3) / 1e9
    gpu_energies = den.cpu().numpy()
    out = {
        "metric": "slice-spin update attempts/s on 1 GPU (path-integral annealing)",
        "value": attempts / (ms / 1e3), "unit": "slice-spin update attempts/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": w["name"], "num_reads": R, "num_steps": S, "num_spins": n, "num_slices": P,
                   "site_updates_per_s": site_updates / (ms / 1e3), "step_ms_events": step_ms,
                   "reads_per_block": info["reads_per_block"], "threads_per_block": info["threads_per_block"],
                   "colour_classes": info["num_classes"], "lds_bytes": info["lds_bytes"], "kind": info["kind"],
                   "mean_energy": float(np.mean(gpu_energies)), "min_energy": float(np.min(gpu_energies))},
        # Worldlines, adjacency and fields are LDS-resident: the algorithmic bytes
        # are LDS bytes (HBM carries only the initial and final worldlines).
        "roofline": dict(on_chip_roofline(achieved, "pimc_kernel", kern_ms, alg_bytes), flips_per_launch=flips),
    }
    if not args.no_cpu_baseline:
        cb = pimc_cpu_baseline(w, gpu_energies)
        cb["speedup_site_updates"] = (site_updates / (ms / 1e3)) / cb["value"]
        cb.update(host_cpu())
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)
    plan.close()


def workload_moments(num_reads=4096, num_sweeps=2000):
    """Differentiable fixed-beta sampling: 64 complete graphs K16 (m=120) and
    8 random 15-regular graphs on 1024 spins (m=7680); h, J ~ N(0, 0.5^2)
    float32; beta = 1, Gibbs, sequential order; 4096 reads per model."""
    rng = np.random.default_rng(5)
    k16 = np.array([(u, v) for u in range(16) for v in range(u + 1, 16)], np.int32)
    nv, edges = [16] * 64, [k16] * 64
    for _ in range(8):
        cs, ce = random_regular(1024, 15, rng)
        nv.append(1024)
        edges.append(np.stack([cs, ce], 1).astype(np.int32))
    h = rng.normal(0, 0.5, sum(nv)).astype(np.float32)
    J = rng.normal(0, 0.5, sum(len(e) for e in edges)).astype(np.float32)
    return {"nv": nv, "edges": edges, "h": h, "J": J, "R": num_reads, "S": num_sweeps,
            "name": f"fixed-beta Gibbs (beta=1, sequential) on 64 x K16 + 8 x 15-regular n=1024, "
                    f"h,J~N(0,0.5^2) float32, {num_reads} reads x {num_sweeps} sweeps per model; "
                    "backward: spin moments <s_i>, <s_u s_v>"}


def run_moments(args):
    """One step = forward (batched fixed-beta Gibbs sweeps) + backward (moment
    kernel), through the torch autograd op."""
    import torch
    from misamplers.fixed_beta import FixedBetaSampler
    w = workload_moments(args.reads if args.reads != 4096 else 4096, args.sweeps if args.sweeps != 2000 else 2000)
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    fb = FixedBetaSampler(w["nv"], w["edges"], num_reads=w["R"], num_sweeps=w["S"], beta=1.0)
    th = torch.tensor(w["h"], device=dev, requires_grad=True)
    tj = torch.tensor(w["J"], device=dev, requires_grad=True)
    stream = torch.cuda.current_stream()
    fwd_ms, bwd_ms, step_ms = [], [], []

    def step(i):
        th.grad = None
        tj.grad = None
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record(stream)
        L, samples = fb(th, tj, seed=1000 + i)
        e1.record(stream)
        L.sum().backward()
        e2.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1), e1.elapsed_time(e2), samples

    for i in range(args.warmup):
        step(i)
    t0 = time.perf_counter()
    for i in range(args.steps):
        f, b, samples = step(args.warmup + i)
        fwd_ms.append(f)
        bwd_ms.append(b)
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # backward kernel alone, timed over repeated launches of the moment plan
    gh = torch.empty(sum(w["nv"]), dtype=torch.float32, device=dev)
    gj = torch.empty(sum(len(e) for e in w["edges"]), dtype=torch.float32, device=dev)
    for _ in range(3):
        fb.plan.run(samples.data_ptr(), fb.d_edges.data_ptr(), 0, gh.data_ptr(), gj.data_ptr(), 0, stream.cuda_stream)
    ka, kb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 50
    ka.record(stream)
    for _ in range(reps):
        fb.plan.run(samples.data_ptr(), fb.d_edges.data_ptr(), 0, gh.data_ptr(), gj.data_ptr(), 0, stream.cuda_stream)
    kb.record(stream)
    torch.cuda.synchronize()
    moment_ms = ka.elapsed_time(kb) / reps
    R, S = w["R"], w["S"]
    n_tot = sum(w["nv"])
    m_tot = sum(len(e) for e in w["edges"])
    attempts = R * S * n_tot
    moment_bytes = R * n_tot + 8 * m_tot + 4 * (n_tot + m_tot)
    achieved = moment_bytes / (moment_ms / 1e3) / 1e9
    out = {
        "metric": "spin-update attempts/s on 1 GPU (differentiable fixed-beta sampling, forward)",
        "value": attempts / (float(np.mean(fwd_ms)) / 1e3), "unit": "spin-update attempts/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": w["name"], "num_models": len(w["nv"]), "reads_per_model": R, "sweeps": S,
                   "forward_ms": float(np.mean(fwd_ms)), "backward_ms": float(np.mean(bwd_ms)),
                   "attempts_per_step": attempts},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "moments_partial_kernel + moments_reduce_kernel",
                     "kernel_ms": moment_ms, "alg_bytes_per_launch": moment_bytes},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = moments_cpu_baseline(w, samples.cpu().numpy(), 1000 + args.warmup + args.steps - 1)
        out["cpu_baseline"]["speedup"] = out["value"] / out["cpu_baseline"]["value"]
        out["cpu_baseline"].update(host_cpu())
    print(json.dumps(out), flush=True)
    fb.close()


def moments_cpu_baseline(w, gpu_samples, seed, target_s=12.0):
    """Reference SA core (Gibbs, constant beta = 1) read by read on a bounded
    sample of the sparse models' reads, checked bit-exact against the GPU."""
    import oracle
    from misamplers import _lib
    kind = "reference" if oracle.have_ref() else "port"
    run = oracle.ref_sa_read_by_read if kind == "reference" else oracle.sa_sample
    cores = host_cores()
    R, S = w["R"], w["S"]
    b = 64  # first sparse model
    n = w["nv"][b]
    e = w["edges"][b]
    hoff = sum(w["nv"][:b])
    joff = sum(len(x) for x in w["edges"][:b])
    h = w["h"][hoff:hoff + n].astype(np.float64)
    J = w["J"][joff:joff + len(e)].astype(np.float64)
    total_state = R * sum(w["nv"])
    init_all = _lib.random_spins(seed, total_state)  # as FixedBetaSampler draws them
    so = R * sum(w["nv"][:b])
    init = init_all[so:so + R * n].reshape(R, n)
    seeds = _lib.read_seeds(seed, R * len(w["nv"]))[b * R:(b + 1) * R]
    betas = np.ones(S)
    t0 = time.perf_counter()
    run(h, e[:, 0], e[:, 1], J, 1, betas, seeds[:cores], init[:cores], 0, 0, nthreads=cores)
    t1 = time.perf_counter() - t0
    k = min(R, cores * max(1, int(target_s / max(t1, 1e-3))))
    t0 = time.perf_counter()
    st, _ = run(h, e[:, 0], e[:, 1], J, 1, betas, seeds[:k], init[:k], 0, 0, nthreads=cores)
    dt = time.perf_counter() - t0
    exact = bool(np.array_equal(st, gpu_samples[so:so + R * n].reshape(R, n)[:k]))
    return {"value": k * S * n / dt, "unit": "spin-update attempts/s", "cores": cores, "kind": kind,
            "sample": f"{k} of {R} reads of one n=1024 15-regular model x {S} Gibbs sweeps at beta=1 "
                      f"({k * S * n:.3g} attempts, {dt:.1f} s); GPU bit-exact on these reads: {exact}",
            "bit_exact": exact}


def run_descent(args):
    """Steepest descent (greedy) of 4096 random states of the 3D +-J L=16 lattice
    to local minima: one step = one launch over all reads, states resident."""
    import torch
    from misamplers import _lib
    w = workload(args.reads, 1)
    R, n = w["R"], w["n"]
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    prob = _lib.Problem(w["h"], w["cs"], w["ce"], w["cw"])
    init_d = torch.from_numpy(w["init"]).to(dev)
    st_d = torch.empty_like(init_d)
    en_d = torch.empty(R, dtype=torch.float64, device=dev)
    k_d = torch.empty(R, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()

    def step():
        st_d.copy_(init_d)
        _lib.check(_lib.lib.misa_descent_run_device(prob.handle, ctypes.c_void_p(st_d.data_ptr()),
                                                    ctypes.c_void_p(en_d.data_ptr()), ctypes.c_void_p(k_d.data_ptr()),
                                                    R, 0xFFFFFFFF, ctypes.c_void_p(stream.cuda_stream)))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        ev[i][0].record(stream)
        step()
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    steps = k_d.cpu().numpy().astype(np.int64)
    flips = int(steps.sum())
    deg = 2.0 * len(w["cs"]) / n
    # reference dense solver per flip: scan n float64 flip energies, then update
    # the flipped variable and its neighbours (29 B per neighbour, 9 B own)
    alg_bytes = flips * (8.0 * n + 9.0 + 29.0 * deg) + R * n * 9.0
    achieved = alg_bytes / (kern_ms / 1e3) / 1e9
    out = {"metric": "descended reads/s on 1 GPU (steepest descent)", "value": R / (ms / 1e3),
           "unit": "reads/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"steepest descent of {R} random states, 3D +-J cubic lattice L=16 (4096 spins)",
                      "flips_per_step": flips, "flips_per_s": flips / (ms / 1e3), "mean_flips_per_read": flips / R},
           "roofline": on_chip_roofline(achieved, "descent_kernel", kern_ms, alg_bytes)}
    if not args.no_cpu_baseline:
        import oracle
        kind = "reference" if oracle.have_ref_greedy() else "port"
        run = (lambda st: oracle.ref_descent(w["h"], w["cs"], w["ce"], w["cw"], st, nthreads=host_cores())) \
            if kind == "reference" else (lambda st: oracle.descent(w["h"], w["cs"], w["ce"], w["cw"], st,
                                                                   nthreads=host_cores()))
        k = min(R, 64 * host_cores())
        t0 = time.perf_counter()
        cst, cen, ck = run(w["init"][:k])
        dt = time.perf_counter() - t0
        exact = bool(np.array_equal(cst, st_d.cpu().numpy()[:k]) and np.array_equal(cen, en_d.cpu().numpy()[:k])
                     and np.array_equal(ck.astype(np.int64), steps[:k]))
        out["cpu_baseline"] = {"value": k / dt, "unit": "reads/s", "cores": host_cores(), "kind": kind,
                               "sample": f"{k} of {R} reads ({dt:.1f} s), dense solver; GPU bit-exact on these "
                                         f"reads: {exact}", "bit_exact": exact}
        out["cpu_baseline"]["speedup"] = out["value"] / out["cpu_baseline"]["value"]
        out["cpu_baseline"].update(host_cpu())
    print(json.dumps(out), flush=True)
    prob.close()


def batch_leg(args, steps, warmup, cpu=True):
    """The batched workload (512 uneven BQMs): one autotuned batched launch
    replayed from a hipGraph, against plain per-BQM launches of the same
    float64 engine on a size-stratified subset.  Returns the record."""
    import torch
    from misamplers import _lib
    from misamplers.sa.batch import Batch
    w = workload_batch(args.bqms, args.reads if args.reads != 4096 else 256,
                       args.sweeps if args.sweeps != 2000 else 1000)
    R, S, sizes = w["R"], w["S"], w["sizes"]
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    batch = Batch(w["problems"], R, 1, w["schedules"])
    policy = batch.autotune(4)
    init = np.concatenate([x.ravel() for x in w["inits"]])
    seeds = np.concatenate([_lib.read_seeds(s, R) for s in w["seeds"]])
    d_init = torch.from_numpy(init).to(dev)
    d_st = torch.empty_like(d_init)
    d_en = torch.empty(batch.total_reads, dtype=torch.float64, device=dev)
    d_sd = torch.from_numpy(seeds.view(np.int64)).to(dev)
    d_ct = torch.zeros((batch.total_reads, 2), dtype=torch.int64, device=dev)  # flips, sum of degrees
    stream = torch.cuda.Stream()
    attempts = int(R * S * int(np.sum(sizes)))

    def step():
        with torch.cuda.stream(stream):
            d_st.copy_(d_init)
            batch.run_device(d_st.data_ptr(), d_en.data_ptr(), d_sd.data_ptr(), d_counters=d_ct.data_ptr(),
                             stream=stream.cuda_stream, use_graph=True)

    for _ in range(warmup):
        step()
    stream.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        ev[i][0].record(stream)
        step()
        ev[i][1].record(stream)
    stream.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    bst, ben = d_st.cpu().numpy(), d_en.cpu().numpy()

    # plain: one launch per BQM (its own plan), on a size-stratified subset
    order = np.argsort(sizes)
    subset = order[np.linspace(0, len(order) - 1, min(args.plain_subset, len(order))).astype(int)]
    plain_time, plain_att, identical = 0.0, 0, True
    for b in subset:
        h, cs, ce, cw = w["problems"][b]
        prob = _lib.Problem(h, cs, ce, cw)
        plan = ctypes.c_void_p()
        betas = np.ascontiguousarray(w["schedules"][b])
        _lib.check(_lib.lib.misa_sa_plan_create(prob.handle, 1, _lib.ptr(betas), len(betas), _lib.SEQUENTIAL,
                                                _lib.METROPOLIS, _lib.MODE_REAL, 0, ctypes.byref(plan)))
        n = int(sizes[b])
        so, ro = batch.state_offsets[b], batch.read_offsets[b]
        p_st = d_init[so:so + R * n].clone()
        p_en = torch.empty(R, dtype=torch.float64, device=dev)
        p_sd = d_sd[ro:ro + R].clone()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        _lib.check(_lib.lib.misa_sa_run_device(plan, ctypes.c_void_p(p_st.data_ptr()),
                                               ctypes.c_void_p(p_en.data_ptr()), ctypes.c_void_p(p_sd.data_ptr()),
                                               R, None, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        e1.record()
        torch.cuda.synchronize()
        plain_time += e0.elapsed_time(e1) / 1e3
        plain_att += R * S * n
        identical &= bool(np.array_equal(p_st.cpu().numpy(), bst[so:so + R * n]) and
                          np.array_equal(p_en.cpu().numpy(), ben[ro:ro + R]))
        _lib.lib.misa_sa_plan_destroy(plan)
        prob.close()
    plain_rate = plain_att / plain_time
    value = attempts / (ms / 1e3)
    ct = d_ct.cpu().numpy()
    flips, degsum = int(ct[:, 0].sum()), int(ct[:, 1].sum())
    alg_bytes = 9.0 * attempts + 9.0 * flips + 29.0 * degsum
    achieved = alg_bytes / (kern_ms / 1e3) / 1e9
    roof = dict(on_chip_roofline(achieved, "sa_batch_kernel", kern_ms, alg_bytes), flips_per_launch=flips)
    prof = profile_figures("sa_batch_kernel_c3.md")
    for k in ("valu_issue_frac", "valu_insts_per_attempt", "wait_frac", "waves_per_cu"):
        if k in prof:
            roof[k] = prof[k]
    if prof:
        roof["counter_source"] = "profiles/sa_batch_kernel_c3.md (rocprofv3 SQ counters)"
    out = {
        "value": value, "unit": "spin-update attempts/s", "steps": steps, "warmup": warmup, "ms_per_step": ms,
        "workload": w["name"], "num_bqms": len(sizes), "total_reads": int(batch.total_reads),
        "attempts_per_step": attempts, "policy": int(policy), "graph": True, "kernel_ms": kern_ms,
        "plain": {"value": plain_rate, "subset_bqms": int(len(subset)), "subset_attempts": int(plain_att),
                  "subset_seconds": plain_time},
        "batched_vs_plain": value / plain_rate, "identical_to_plain": identical,
        "roofline": roof,
    }
    if cpu and not args.no_cpu_baseline:
        out["cpu_baseline"] = batch_cpu_baseline(w, subset, batch, bst, ben)
        out["cpu_baseline"]["speedup"] = value / out["cpu_baseline"]["value"]
        out["cpu_baseline"].update(host_cpu())
    batch.close()
    return out


def run_batch(args):
    """The batched workload on its own line (--config batch)."""
    rec = batch_leg(args, args.steps, args.warmup)
    out = {"metric": "spin-update attempts/s on 1 GPU (batched SA)", "value": rec["value"],
           "unit": "spin-update attempts/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": rec["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64 fields, u64 rng", "data": "synthetic",
           "config": {k: rec[k] for k in ("workload", "num_bqms", "total_reads", "attempts_per_step", "policy",
                                          "graph", "kernel_ms")},
           "plain": rec["plain"], "batched_vs_plain": rec["batched_vs_plain"],
           "identical_to_plain": rec["identical_to_plain"], "roofline": rec["roofline"]}
    if "cpu_baseline" in rec:
        out["cpu_baseline"] = rec["cpu_baseline"]
    print(json.dumps(out), flush=True)


def batch_cpu_baseline(w, subset, batch, gpu_states, gpu_energies, target_s=12.0):
    """The reference core (oracle/_ref, or the oracle's port) read by read on a
    size-stratified subset of the BQMs, k reads of each, on every host core;
    the rate is attempts/s over that sample (not extrapolated), and the
    sampled reads are checked bit for bit against the batched GPU run."""
    import oracle
    from misamplers import _lib
    kind = "reference" if oracle.have_ref() else "port"
    run = oracle.ref_sa_read_by_read if kind == "reference" else oracle.sa_sample
    cores = host_cores()
    R, S = w["R"], w["S"]

    def sample(k):
        t, att, exact = 0.0, 0, True
        for b in subset:
            h, cs, ce, cw = w["problems"][b]
            n = len(h)
            sd = _lib.read_seeds(w["seeds"][b], R)[:k]
            t0 = time.perf_counter()
            st, en = run(h, cs, ce, cw, 1, w["schedules"][b], sd, w["inits"][b][:k], 0, 1, nthreads=cores)
            t += time.perf_counter() - t0
            att += k * S * n
            so, ro = batch.state_offsets[b], batch.read_offsets[b]
            exact &= bool(np.array_equal(st, gpu_states[so:so + k * n].reshape(k, n)) and
                          np.array_equal(en, gpu_energies[ro:ro + k]))
        return t, att, exact

    k = cores
    t, att, exact = sample(k)
    if t < target_s / 3:
        k = min(R, max(k, int(k * target_s / max(t, 1e-3)) // cores * cores))
        t, att, exact = sample(k)
    return {"value": att / t, "unit": "spin-update attempts/s", "cores": cores, "kind": kind,
            "sample": f"{k} reads of each of {len(subset)} size-stratified BQMs x {S} sweeps ({att:.3g} attempts, "
                      f"{t:.1f} s); GPU bit-exact on these reads: {exact}", "bit_exact": exact}


def run_rotor(args):
    """Rotor-model annealing (MetropolisUniform, the sampler's default) of the
    path-integral workload's problem: one step = the full schedule for all
    reads, angles resident in HBM."""
    import torch
    from misamplers import _lib
    w = workload_pimc(args.reads if args.reads != 4096 else 1024, args.sweeps if args.sweeps != 2000 else 1000)
    R, n, S = w["R"], w["n"], w["S"]
    proposal = "MetropolisUniform"
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    prob = _lib.Problem(w["h"], w["cs"], w["ce"], w["cw"])
    tf = np.ones(n)
    hp, hd = np.ascontiguousarray(w["hp"]), np.ascontiguousarray(w["hd"])
    plan = ctypes.c_void_p()
    _lib.check(_lib.lib.misa_rotor_plan_create(prob.handle, _lib.ptr(tf), 1, _lib.ptr(hp), _lib.ptr(hd), S, 0, 2,
                                               ctypes.byref(plan)))
    info = _lib.RotorStats()
    _lib.check(_lib.lib.misa_rotor_plan_info(plan, ctypes.byref(info)))
    init = (64 - 64 * w["init"].astype(np.int16)).astype(np.uint8)   # spins to the poles, as the sampler does
    seeds = _lib.read_seeds(4321, R)
    d0 = torch.from_numpy(init).to(dev)
    dst = torch.empty_like(d0)
    den = torch.empty(R, dtype=torch.float64, device=dev)
    dS = torch.from_numpy(seeds.view(np.int64)).to(dev)
    dacc = torch.zeros(R, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()

    def step():
        dst.copy_(d0)
        _lib.check(_lib.lib.misa_rotor_run_device(plan, ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(den.data_ptr()),
                                                  ctypes.c_void_p(dS.data_ptr()), R, None, 0,
                                                  ctypes.c_void_p(dacc.data_ptr()), ctypes.c_void_p(stream.cuda_stream)))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        ev[i][0].record(stream)
        step()
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    attempts = R * S * n
    accepted = int(dacc.cpu().numpy().sum())
    dbar = 2.0 * len(w["cs"]) / n
    # per attempt at reference widths: own angle, h and t (f64), per neighbour
    # index (i32) + weight (f64) + angle, four table lookups (f64); per
    # acceptance one angle write
    alg_bytes = attempts * (17.0 + 13.0 * dbar + 32.0) + accepted
    achieved = alg_bytes / (kern_ms / 1e3) / 1e9
    gpu_states = dst.cpu().numpy()
    gpu_energies = den.cpu().numpy()
    out = {
        "metric": "rotor update attempts/s on 1 GPU (rotor-model annealing)",
        "value": attempts / (ms / 1e3), "unit": "rotor update attempts/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "rotor-model annealing (" + proposal + ") of the " + w["name"].split(": ", 1)[1]
                   .replace("32 Trotter slices, ", ""), "num_reads": R, "num_steps": S, "num_vars": n,
                   "accepted": accepted, "lanes_per_read": info.lanes_per_read,
                   "reads_per_block": info.reads_per_block, "threads_per_block": info.threads_per_block,
                   "csr_in_lds": info.csr_in_lds, "lds_bytes": info.lds_bytes,
                   "wave_per_read": info.wave_per_read, "run_kernel": info.run_kernel,
                   "mean_energy": float(np.mean(gpu_energies))},
        "roofline": on_chip_roofline(achieved, "rotor_run_kernel" if info.run_kernel else
                                     "rotor_wave_kernel" if info.wave_per_read else "rotor_kernel",
                                     kern_ms, alg_bytes),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = rotor_cpu_baseline(w, tf, init, seeds, gpu_states, gpu_energies, proposal)
        out["cpu_baseline"]["speedup"] = out["value"] / out["cpu_baseline"]["value"]
        out["cpu_baseline"].update(host_cpu())
    print(json.dumps(out), flush=True)
    _lib.lib.misa_rotor_plan_destroy(plan)
    prob.close()


def rotor_cpu_baseline(w, tf, init, seeds, gpu_states, gpu_energies, proposal, target_s=12.0):
    """The reference rotor core (oracle/_ref/libref_rotor.so) on a bounded
    sample of the reads, one call per read on every host core, checked bit for
    bit against the GPU on those reads; the oracle's C restatement when the
    compiled reference is absent."""
    import oracle
    cores = host_cores()
    use_ref = oracle.have_ref_rotor()
    fn = oracle.ref_rotor if use_ref else oracle.rotor_sample
    n, S = w["n"], w["S"]
    k = cores
    t0 = time.perf_counter()
    st, en, _ = fn(w["h"], w["cs"], w["ce"], w["cw"], tf, 1, w["hp"], w["hd"], seeds[:k], init[:k], False, proposal,
                   0, nthreads=cores)
    dt = time.perf_counter() - t0
    if dt < target_s / 3:
        k2 = min(len(seeds), max(k, int(k * target_s / max(dt, 1e-3)) // cores * cores))
        if k2 > k:
            t0 = time.perf_counter()
            st, en, _ = fn(w["h"], w["cs"], w["ce"], w["cw"], tf, 1, w["hp"], w["hd"], seeds[:k2], init[:k2], False,
                           proposal, 0, nthreads=cores)
            dt = time.perf_counter() - t0
            k = k2
    exact = bool(np.array_equal(st, gpu_states[:k]) and np.array_equal(en, gpu_energies[:k]))
    return {"value": k * S * n / dt, "unit": "rotor update attempts/s", "cores": cores,
            "kind": "reference" if use_ref else "port",
            "sample": f"{k} of {len(seeds)} reads ({dt:.1f} s), one read per thread; "
                      f"GPU bit-exact on these reads: {exact}", "bit_exact": exact}


def workload_tree(w=16, L=64, beta=1.0):
    """A 16 x 64 strip (1024 spins; variable (r, l) -> r + 16 l; couplers
    within columns, then along rows), h, J ~ N(0, 0.5^2) float64, beta 1,
    eliminated column by column (treewidth 16)."""
    rng = np.random.default_rng(2024)
    h = rng.normal(0, 0.5, size=(w, L))
    cs, ce = [], []
    for l in range(L):
        for r in range(w - 1):
            cs.append(r + w * l)
            ce.append(r + 1 + w * l)
    for l in range(L - 1):
        for r in range(w):
            cs.append(r + w * l)
            ce.append(r + w * (l + 1))
    cw = rng.normal(0, 0.5, size=len(cs))
    return {"h": h.T.ravel().copy(), "cs": np.array(cs, np.int32), "ce": np.array(ce, np.int32), "cw": cw,
            "beta": beta, "order": np.arange(w * L, dtype=np.int32), "n": w * L,
            "name": f"exact log Z and all variable and pair marginals (bucket elimination) of a {w} x {L} strip "
                    f"({w * L} spins, {len(cs)} couplers, h, J ~ N(0, 0.5^2) float64), beta {beta}, column-by-column "
                    f"elimination order (treewidth {w})"}


def run_tree(args):
    """Exact inference by bucket elimination: one step = the upward pass, log
    Z, the downward pass and every variable and pair marginal of the strip,
    tables resident in HBM (the plan is built once)."""
    import torch
    from misamplers.tree import TreePlan
    w = workload_tree()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    plan = TreePlan(w["h"], w["cs"], w["ce"], w["cw"], beta=w["beta"], low=-1, elimination_order=w["order"])
    stream = torch.cuda.current_stream()

    def step():
        plan.recompute()
        return plan.marginals(stream=stream.cuda_stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    fwd, marg = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        lz, p1, pp = step()
        st = plan.last_stats()
        fwd.append(st.forward_ms)
        marg.append(st.marginals_ms)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    fwd_ms, marg_ms = float(np.mean(fwd)), float(np.mean(marg))
    # dominant launches: the upward pass, one launch per level; per launch its
    # share of the pass's algorithmic bytes (T and mu written, children's mu read)
    levels = st.forward_levels
    kern_ms = fwd_ms / levels
    alg_bytes = st.forward_bytes / levels
    achieved = st.forward_bytes / (fwd_ms / 1e3) / 1e9
    out = {
        "metric": "exact inference passes/s on 1 GPU (bucket elimination: log Z + all marginals)",
        "value": 1e3 / ms, "unit": "passes/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": w["name"], "treewidth": plan.treewidth, "table_entries": int(st.table_entries),
                   "upward_launches": levels, "downward_launches": st.downward_levels,
                   "marginal_blocks": int(st.marginal_blocks), "upward_ms": fwd_ms, "downward_and_marginals_ms": marg_ms,
                   "entries_per_s": st.table_entries / (ms / 1e3), "log_z": lz},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "tree_forward_value_kernel",
                     "kernel_ms": kern_ms, "alg_bytes_per_launch": alg_bytes},
    }
    if not args.no_cpu_baseline:
        import oracle.tree_exact as te
        t0 = time.perf_counter()
        clz, cp1, ccells = te.bucket_eliminate(w["h"], w["cs"], w["ce"], w["cw"], list(w["order"]), beta=w["beta"])
        dt = time.perf_counter() - t0
        dz = abs(clz - lz) / abs(clz)
        dm = max(float(np.abs(cp1 - p1).max()), float(np.abs(ccells - pp).max()))
        out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "passes/s", "cores": 1, "kind": "port",
                               "sample": f"one full pass ({dt:.1f} s), numpy bucket elimination in the same order; "
                                         f"GPU vs CPU: log Z rel. diff {dz:.1e}, max marginal diff {dm:.1e}",
                               "agree_1e-10": bool(dz < 1e-10 and dm < 1e-10)}
        out["cpu_baseline"]["speedup"] = out["value"] / out["cpu_baseline"]["value"]
        out["cpu_baseline"].update(host_cpu())
    print(json.dumps(out), flush=True)
    plan.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--reads", type=int, default=4096)
    ap.add_argument("--sweeps", type=int, default=2000)
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--config", choices=["lattice", "gnp", "batch", "pimc", "moments", "descent", "rotor", "tree"],
                    default="lattice",
                    help="lattice: 3D +-J L=16 (headline); gnp: G(100,0.5) real-valued; batch: 512 uneven BQMs; "
                         "pimc: path-integral annealing, n=2048 6-regular, 32 slices; moments: differentiable "
                         "fixed-beta sampling of 64 K16 + 8 sparse models (forward + backward); descent: "
                         "steepest descent of 4096 random states on the L=16 lattice; rotor: rotor-model annealing of "
                         "the path-integral problem; tree: exact log Z and marginals of a 16 x 64 strip")
    ap.add_argument("--slices", type=int, default=32)
    ap.add_argument("--bqms", type=int, default=512)
    ap.add_argument("--plain-subset", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="default line only: leave out the batched sub-record (one warmup + one timed step "
                         "of the 512-BQM workload, whatever --steps says)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (reads are independent "
              "chains: each GPU would run the same kernel on its share of the reads)", file=sys.stderr)
        sys.exit(2)

    if args.config == "batch":
        return run_batch(args)
    if args.config == "pimc":
        return run_pimc(args)
    if args.config == "moments":
        return run_moments(args)
    if args.config == "descent":
        return run_descent(args)
    if args.config == "rotor":
        return run_rotor(args)
    if args.config == "tree":
        return run_tree(args)

    import torch
    from misamplers import _lib

    if args.config == "gnp":
        w = workload_gnp(args.reads if args.reads != 4096 else 1000, args.sweeps if args.sweeps != 2000 else 1000)
    else:
        w = workload(args.reads, args.sweeps)
    R, n, S = w["R"], w["n"], w["S"]
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    prob = _lib.Problem(w["h"], w["cs"], w["ce"], w["cw"])
    plan = ctypes.c_void_p()
    betas = np.ascontiguousarray(w["betas"])
    _lib.check(_lib.lib.misa_sa_plan_create(prob.handle, 1, _lib.ptr(betas), len(betas), _lib.SEQUENTIAL,
                                            _lib.METROPOLIS, _lib.MODE_AUTO, args.lanes, ctypes.byref(plan)))
    seeds = _lib.read_seeds(12345, R)
    init_d = torch.from_numpy(w["init"]).to(dev)
    states_d = torch.empty_like(init_d)
    energies_d = torch.empty(R, dtype=torch.float64, device=dev)
    seeds_d = torch.from_numpy(seeds.view(np.int64)).to(dev)
    counters_d = torch.zeros((R, 2), dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()

    def step(count=False):
        states_d.copy_(init_d)
        _lib.check(_lib.lib.misa_sa_run_device(plan, ctypes.c_void_p(states_d.data_ptr()),
                                               ctypes.c_void_p(energies_d.data_ptr()),
                                               ctypes.c_void_p(seeds_d.data_ptr()), R,
                                               ctypes.c_void_p(counters_d.data_ptr() if count else 0),
                                               ctypes.c_void_p(stream.cuda_stream)))

    # The timed launches keep no flip counters, as the sampler's own runs do.
    # The flips (for the algorithmic bytes) come from one counted run of the
    # same inputs: the first warmup run, or one after the timed region; the
    # samples are the same with and without counting (checked below).
    counted_energies = None
    for i in range(args.warmup):
        step(count=(i == 0))
        if i == 0:
            counted_energies = energies_d.clone()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        ev[i][0].record(stream)
        step()
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = [a.elapsed_time(b) for a, b in ev]
    ms = wall * 1e3 / args.steps
    attempts = R * S * n
    info = _lib.SaStats()
    _lib.lib.misa_sa_plan_info(plan, ctypes.byref(info))
    timed_energies = energies_d.clone()
    if counted_energies is None:
        step(count=True)
        counted_energies = energies_d.clone()
        torch.cuda.synchronize()
    counted_same = bool(torch.equal(counted_energies, timed_energies))

    cnt = counters_d.cpu().numpy().astype(np.uint64)
    flips, degsum = int(cnt[:, 0].sum()), int(cnt[:, 1].sum())
    alg_bytes = 9 * attempts + 9 * flips + 29 * degsum
    kern_s = float(np.mean(step_ms)) / 1e3
    achieved = alg_bytes / kern_s / 1e9
    gpu_states = states_d.cpu().numpy()
    gpu_energies = energies_d.cpu().numpy()

    roof = {"bound": "lds", "achieved": achieved, "peak": LDS_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / LDS_PEAK_GBS, "traffic": None,
            "kernel": {1: "sa_int_kernel", 2: "sa_real_kernel", 4: "sa_seg_kernel", 5: "sa_chain_kernel",
                       6: "sa_dense_kernel"}.get(int(info.kernel), "sa_sweep_kernel"),
            "kernel_ms": kern_s * 1e3, "alg_bytes_per_launch": alg_bytes, "flips_per_launch": flips,
            "flips_from_counted_run": True, "counted_run_identical": counted_same,
            "hbm_frac_equivalent": achieved / HBM_PEAK_GBS}
    roof.update(lds_fields(achieved))
    summary = None
    if args.config == "lattice" and (R, S, n) == (4096, 2000, 4096):
        # counter figures of this exact launch from its committed rocprofv3 summary:
        # HBM bytes (FETCH_SIZE + WRITE_SIZE, scaled by the copy calibration) and the
        # issue profile (SQ counters), the binding limit of this kernel
        summary = {"sa_chain_kernel": "sa_chain_kernel_lattice_L16.md",
                   "sa_int_kernel": "sa_int_kernel_lattice_L16.md"}.get(roof["kernel"])
    elif args.config == "gnp" and (R, S, n) == (1000, 1000, 100) and roof["kernel"] == "sa_dense_kernel":
        summary = "sa_dense_kernel_gnp100.md"
    if summary:
        prof = profile_figures(summary)
        if "traffic_bytes_per_launch" in prof:
            roof["traffic"] = int(prof["traffic_bytes_per_launch"])
        for k in ("valu_issue_frac", "valu_insts_per_step", "cycles_per_step"):
            if k in prof:
                roof[k] = prof[k]
        roof["counter_source"] = f"profiles/{summary} (rocprofv3 --pmc)"
    out = {
        "metric": "spin-update attempts/s on 1 GPU (plain vs batched) + % of LDS/HBM roofline"
                  if args.config == "lattice" else "spin-update attempts/s on 1 GPU (plain SA)",
        "value": attempts / (ms / 1e3),
        "unit": "spin-update attempts/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {1: "int8", 2: "int16", 3: "f64"}.get(int(info.mode), "?") + " fields, u64 rng",
        "data": "synthetic",
        "config": {"workload": w["name"],
                   "num_reads": R, "num_sweeps": S, "num_spins": n,
                   "lanes_per_read": int(info.lanes_per_read), "reads_per_block": int(info.reads_per_block),
                   "state_in_lds": bool(info.state_in_lds), "csr_in_lds": bool(info.csr_in_lds),
                   "waves_per_block": int(info.waves_per_block)},
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, seeds, gpu_states, gpu_energies)
        out["cpu_baseline"]["speedup"] = out["value"] / out["cpu_baseline"]["value"]
        out["cpu_baseline"].update(host_cpu())
    _lib.lib.misa_sa_plan_destroy(plan)
    prob.close()
    if args.config == "lattice" and not args.no_batched:
        # the batched half of the metric: one warmup and one timed step of the
        # 512-BQM workload (a step takes tens of seconds)
        args.reads, args.sweeps = 4096, 2000
        out["batched"] = batch_leg(args, 1, 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
