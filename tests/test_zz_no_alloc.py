"""misa_sa_run_device_noalloc (the allocation-free device run) against
misa_sa_run_device and the oracle: two calls in a row, the second with fewer
reads, for float64 and integer plans, with the per-read state in LDS and in
the entry point's static scratch (the general kernel), and the size limit."""
import ctypes

import numpy as np
import pytest

import oracle
from misamplers import _lib

pytestmark = pytest.mark.gpu


def _problem(rng, n, integer):
    pairs = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < 0.04]
    cs = np.array([a for a, _ in pairs], np.int32)
    ce = np.array([b for _, b in pairs], np.int32)
    if integer:
        return rng.integers(-1, 2, n).astype(float), cs, ce, rng.choice([-1.0, 1.0], len(cs))
    return rng.uniform(-1, 1, n), cs, ce, rng.uniform(-1, 1, len(cs))


def _run(fn, plan, init, seeds, counters):
    import torch
    dev = torch.device("cuda:0")
    st = torch.from_numpy(init.copy()).to(dev)
    en = torch.empty(init.shape[0], dtype=torch.float64, device=dev)
    sd = torch.from_numpy(seeds.view(np.int64)).to(dev)
    ct = torch.zeros((init.shape[0], 2), dtype=torch.int64, device=dev) if counters else None
    stream = torch.cuda.current_stream().cuda_stream
    _lib.check(fn(plan, ctypes.c_void_p(st.data_ptr()), ctypes.c_void_p(en.data_ptr()),
                  ctypes.c_void_p(sd.data_ptr()), init.shape[0], ctypes.c_void_p(ct.data_ptr() if counters else 0),
                  ctypes.c_void_p(stream)))
    torch.cuda.synchronize()
    return st.cpu().numpy(), en.cpu().numpy(), (ct.cpu().numpy() if counters else None)


@pytest.mark.parametrize("in_device_memory", [True, False])
@pytest.mark.parametrize("integer", [False, True])
def test_noalloc_equals_run_device(monkeypatch, in_device_memory, integer):
    if in_device_memory:
        monkeypatch.setenv("MISA_STATE_IN_DEVICE_MEMORY", "1")
    rng = np.random.default_rng(7 + 2 * integer + in_device_memory)
    n = 300
    h, cs, ce, cw = _problem(rng, n, integer)
    prob = _lib.Problem(h, cs, ce, cw)
    betas = np.ascontiguousarray(np.geomspace(0.1, 3, 40))
    plan = ctypes.c_void_p()
    _lib.check(_lib.lib.misa_sa_plan_create(prob.handle, 1, _lib.ptr(betas), len(betas), _lib.SEQUENTIAL,
                                            _lib.METROPOLIS, _lib.MODE_AUTO, 0, ctypes.byref(plan)))
    try:
        for R in (96, 37):  # two calls in a row, the second smaller
            init = (2 * rng.integers(0, 2, (R, n)) - 1).astype(np.int8)
            seeds = oracle.read_seeds(int(rng.integers(2**31)), R)
            print("no-alloc run", integer, in_device_memory, R, flush=True)
            a = _run(_lib.lib.misa_sa_run_device, plan, init, seeds, True)
            b = _run(_lib.lib.misa_sa_run_device_noalloc, plan, init, seeds, True)
            info = _lib.SaStats()
            _lib.lib.misa_sa_plan_info(plan, ctypes.byref(info))
            assert bool(info.state_in_lds) is (not in_device_memory)
            for x, y in zip(a, b):
                np.testing.assert_array_equal(x, y)
            ref = oracle.sa_sample(h, cs, ce, cw, 1, betas, seeds, init, nthreads=8)
            np.testing.assert_array_equal(b[0], ref[0])
            np.testing.assert_array_equal(b[1], ref[1])
    finally:
        _lib.lib.misa_sa_plan_destroy(plan)
        prob.close()


def test_noalloc_rejects_state_beyond_the_scratch(monkeypatch):
    monkeypatch.setenv("MISA_STATE_IN_DEVICE_MEMORY", "1")
    n, R = 40000, 200  # 9 bytes per variable and read: 72 MB > 64 MiB
    rng = np.random.default_rng(3)
    h = rng.uniform(-1, 1, n)
    cs = np.arange(n - 1, dtype=np.int32)
    ce = cs + 1
    prob = _lib.Problem(h, cs, ce, rng.uniform(-1, 1, n - 1))
    betas = np.ascontiguousarray(np.geomspace(0.1, 3, 2))
    plan = ctypes.c_void_p()
    _lib.check(_lib.lib.misa_sa_plan_create(prob.handle, 1, _lib.ptr(betas), len(betas), _lib.SEQUENTIAL,
                                            _lib.METROPOLIS, _lib.MODE_REAL, 0, ctypes.byref(plan)))
    try:
        init = np.ones((R, n), np.int8)
        with pytest.raises(_lib.EngineError, match="exceeds the allocation-free scratch"):
            _run(_lib.lib.misa_sa_run_device_noalloc, plan, init, oracle.read_seeds(1, R), False)
    finally:
        _lib.lib.misa_sa_plan_destroy(plan)
        prob.close()
