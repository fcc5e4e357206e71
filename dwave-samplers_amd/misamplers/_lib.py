"""ctypes binding of the engine library ``libmisamplers.so`` (include/misamplers.h).

The library is the only compute path: if it is missing, importing this module
raises ImportError with the build command, and there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmisamplers.so")

MISA_OK = 0
MISA_ERR_ARG = -1
MISA_ERR_LENGTHS = -2
MISA_ERR_INDEX = -3
MISA_ERR_DEVICE = -4
MISA_ERR_UNSUPPORTED = -5

MODE_AUTO, MODE_INT8, MODE_INT16, MODE_REAL = 0, 1, 2, 3
GIBBS, METROPOLIS = 0, 1
SEQUENTIAL, RANDOM = 0, 1

INTERRUPT_CB = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p)


class SaStats(ctypes.Structure):
    _fields_ = [("flips", ctypes.c_uint64),
                ("flip_degree", ctypes.c_uint64),
                ("kernel_ms", ctypes.c_double),
                ("mode", ctypes.c_int),
                ("lanes_per_read", ctypes.c_int),
                ("reads_per_block", ctypes.c_int),
                ("state_in_lds", ctypes.c_int),
                ("kernel", ctypes.c_int),
                ("csr_in_lds", ctypes.c_int),
                ("waves_per_block", ctypes.c_int),
                ("chain_form", ctypes.c_int)]


class RotorStats(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double),
                ("accepted", ctypes.c_int64),
                ("lanes_per_read", ctypes.c_int),
                ("reads_per_block", ctypes.c_int),
                ("threads_per_block", ctypes.c_int),
                ("csr_in_lds", ctypes.c_int),
                ("lds_bytes", ctypes.c_int),
                ("wave_per_read", ctypes.c_int),
                ("run_kernel", ctypes.c_int)]


class TreeStats(ctypes.Structure):
    _fields_ = [("forward_ms", ctypes.c_double),
                ("marginals_ms", ctypes.c_double),
                ("forward_levels", ctypes.c_int),
                ("downward_levels", ctypes.c_int),
                ("marginal_blocks", ctypes.c_int64),
                ("table_entries", ctypes.c_int64),
                ("forward_bytes", ctypes.c_double),
                ("downward_bytes", ctypes.c_double),
                ("marginal_bytes", ctypes.c_double)]


class PimcStats(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double),
                ("site_updates", ctypes.c_int64),
                ("num_slices", ctypes.c_int),
                ("reads_per_block", ctypes.c_int),
                ("threads_per_block", ctypes.c_int),
                ("num_classes", ctypes.c_int),
                ("kind", ctypes.c_int),
                ("lds_bytes", ctypes.c_int),
                ("max_degree", ctypes.c_int),
                ("blocks", ctypes.c_int)]



def _share_torch_hip_runtime():
    """PyTorch-ROCm ships its own HIP and HSA runtime libraries.  Loading them
    first (global symbols) makes the engine bind to the same runtime, so a
    process that uses both torch tensors and the engine holds one HIP runtime
    instead of two.  Without torch the engine uses the system ROCm runtime."""
    try:
        import importlib.util
        spec = importlib.util.find_spec("torch")
        if spec is None or not spec.submodule_search_locations:
            return
        libdir = os.path.join(list(spec.submodule_search_locations)[0], "lib")
        for name in ("libhsa-runtime64.so", "libamdhip64.so"):
            path = os.path.join(libdir, name)
            if os.path.exists(path):
                ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    except OSError:
        pass


_share_torch_hip_runtime()

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"misamplers engine library not found at {LIB_PATH}; build it with "
        "`make -C dwave-samplers_amd` (hipcc, gfx950)")

lib = ctypes.CDLL(LIB_PATH)

_vp = ctypes.c_void_p
lib.misa_last_error.restype = ctypes.c_char_p
lib.misa_version.restype = ctypes.c_char_p
lib.misa_device_count.restype = ctypes.c_int
lib.misa_read_seeds.restype = None
lib.misa_read_seeds.argtypes = [ctypes.c_uint64, ctypes.c_int64, _vp]
lib.misa_read_seeds_device.restype = ctypes.c_int
lib.misa_read_seeds_device.argtypes = [ctypes.c_uint64, ctypes.c_int64, _vp, _vp]
lib.misa_random_spins.restype = None
lib.misa_random_spins.argtypes = [ctypes.c_uint64, ctypes.c_int64, _vp]
lib.misa_random_spins_device.restype = ctypes.c_int
lib.misa_random_spins_device.argtypes = [ctypes.c_uint64, ctypes.c_int64, _vp, _vp]
lib.misa_batch_set_values_device.restype = ctypes.c_int
lib.misa_batch_set_values_device.argtypes = [_vp, _vp, _vp, _vp]
lib.misa_problem_create.restype = ctypes.c_int
lib.misa_problem_create.argtypes = [ctypes.c_int, _vp, ctypes.c_int64, _vp, ctypes.c_int64, _vp,
                                    ctypes.c_int64, _vp, ctypes.POINTER(_vp)]
lib.misa_problem_destroy.restype = None
lib.misa_problem_destroy.argtypes = [_vp]
lib.misa_problem_auto_mode.restype = ctypes.c_int
lib.misa_problem_auto_mode.argtypes = [_vp]
lib.misa_problem_max_degree.restype = ctypes.c_int
lib.misa_problem_max_degree.argtypes = [_vp]
lib.misa_sa_sample.restype = ctypes.c_int
lib.misa_sa_sample.argtypes = [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp,
                               ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               INTERRUPT_CB, _vp, ctypes.POINTER(SaStats)]
lib.misa_sa_plan_create.restype = ctypes.c_int
lib.misa_sa_plan_create.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]
lib.misa_sa_plan_destroy.restype = None
lib.misa_sa_plan_destroy.argtypes = [_vp]
lib.misa_sa_plan_info.restype = ctypes.c_int
lib.misa_sa_plan_info.argtypes = [_vp, ctypes.POINTER(SaStats)]
lib.misa_sa_run_device.restype = ctypes.c_int
lib.misa_sa_run_device.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_int, _vp, _vp]
lib.misa_sa_run_device_noalloc.restype = ctypes.c_int
lib.misa_sa_run_device_noalloc.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_int, _vp, _vp]

_pp = ctypes.POINTER(ctypes.c_void_p)
lib.misa_batch_create.restype = ctypes.c_int
lib.misa_batch_create.argtypes = [ctypes.c_int, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _pp]
lib.misa_batch_destroy.restype = None
lib.misa_batch_destroy.argtypes = [_vp]
for _f in ("misa_batch_state_offset", "misa_batch_read_offset"):
    getattr(lib, _f).restype = ctypes.c_int64
    getattr(lib, _f).argtypes = [_vp, ctypes.c_int]
for _f in ("misa_batch_total_reads", "misa_batch_total_state"):
    getattr(lib, _f).restype = ctypes.c_int64
    getattr(lib, _f).argtypes = [_vp]
lib.misa_batch_set_policy.restype = ctypes.c_int
lib.misa_batch_set_policy.argtypes = [_vp, ctypes.c_int]
lib.misa_batch_autotune.restype = ctypes.c_int
lib.misa_batch_autotune.argtypes = [_vp, ctypes.c_int]
lib.misa_batch_run_device.restype = ctypes.c_int
lib.misa_batch_run_device.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int]
lib.misa_batch_sample.restype = ctypes.c_int
lib.misa_batch_sample.argtypes = [_vp, _vp, _vp, _vp, ctypes.POINTER(SaStats)]

lib.misa_pimc_plan_create.restype = ctypes.c_int
lib.misa_pimc_plan_create.argtypes = [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, _pp]
lib.misa_pimc_plan_create_chains.restype = ctypes.c_int
lib.misa_pimc_plan_create_chains.argtypes = [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_double, _pp]
lib.misa_pimc_plan_destroy.restype = None
lib.misa_pimc_plan_destroy.argtypes = [_vp]
lib.misa_pimc_colouring.restype = ctypes.c_int
lib.misa_pimc_colouring.argtypes = [_vp, _vp]
lib.misa_pimc_chain_colouring.restype = ctypes.c_int
lib.misa_pimc_chain_colouring.argtypes = [_vp, ctypes.c_int, _vp]
lib.misa_pimc_plan_info.restype = ctypes.c_int
lib.misa_pimc_plan_info.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(PimcStats)]
lib.misa_pimc_run_device.restype = ctypes.c_int
lib.misa_pimc_run_device.argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp]
lib.misa_pimc_project_device.restype = ctypes.c_int
lib.misa_pimc_project_device.argtypes = [_vp, _vp, ctypes.c_int, _vp, _vp, _vp]
lib.misa_pimc_sample.restype = ctypes.c_int
lib.misa_pimc_sample.argtypes = [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp,
                                 ctypes.c_int64, _vp, ctypes.c_int64, ctypes.c_int, _vp, _vp,
                                 ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                 ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.c_int, INTERRUPT_CB, _vp,
                                 ctypes.POINTER(PimcStats)]

lib.misa_descent_run_device.restype = ctypes.c_int
lib.misa_descent_run_device.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_uint32, _vp]
lib.misa_descent_sample.restype = ctypes.c_int
lib.misa_descent_sample.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_uint32,
                                    ctypes.POINTER(ctypes.c_double)]
lib.misa_rotor_plan_create.restype = ctypes.c_int
lib.misa_rotor_plan_create.argtypes = [_vp, _vp, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       _pp]
lib.misa_rotor_plan_destroy.restype = None
lib.misa_rotor_plan_destroy.argtypes = [_vp]
lib.misa_rotor_plan_info.restype = ctypes.c_int
lib.misa_rotor_plan_info.argtypes = [_vp, ctypes.POINTER(RotorStats)]
lib.misa_rotor_run_device.restype = ctypes.c_int
lib.misa_rotor_run_device.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp]
lib.misa_rotor_sample.restype = ctypes.c_int
lib.misa_rotor_sample.argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp,
                                  ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, INTERRUPT_CB, _vp,
                                  ctypes.POINTER(RotorStats)]
lib.misa_tree_plan_create.restype = ctypes.c_int
lib.misa_tree_plan_create.argtypes = [_vp, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, _pp]
lib.misa_tree_plan_destroy.restype = None
lib.misa_tree_plan_destroy.argtypes = [_vp]
lib.misa_tree_plan_info.restype = ctypes.c_int
lib.misa_tree_plan_info.argtypes = [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int64),
                                    ctypes.POINTER(ctypes.c_int), _vp]
lib.misa_tree_recompute.restype = ctypes.c_int
lib.misa_tree_recompute.argtypes = [_vp]
lib.misa_tree_pairs.restype = ctypes.c_int
lib.misa_tree_pairs.argtypes = [_vp, _vp, _vp]
lib.misa_tree_marginals.restype = ctypes.c_int
lib.misa_tree_marginals.argtypes = [_vp, ctypes.POINTER(ctypes.c_double), _vp, _vp, _vp]
lib.misa_lds_bandwidth.restype = ctypes.c_int
lib.misa_lds_bandwidth.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
lib.misa_tree_last_stats.restype = ctypes.c_int
lib.misa_tree_last_stats.argtypes = [_vp, ctypes.POINTER(TreeStats)]
lib.misa_tree_sample.restype = ctypes.c_int
lib.misa_tree_sample.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp]
lib.misa_tree_solve.restype = ctypes.c_int
lib.misa_tree_solve.argtypes = [_vp, ctypes.c_int, _vp, _vp, ctypes.POINTER(ctypes.c_int)]
lib.misa_moments_plan_create.restype = ctypes.c_int
lib.misa_moments_plan_create.argtypes = [ctypes.c_int, _vp, _vp, _vp, _pp]
lib.misa_moments_plan_destroy.restype = None
lib.misa_moments_plan_destroy.argtypes = [_vp]
lib.misa_moments_run.restype = ctypes.c_int
lib.misa_moments_run.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]


class EngineError(RuntimeError):
    """A non-argument failure reported by the engine (device or limits)."""


def check(rc: int) -> int:
    """Map a negative return code to the reference's exception types."""
    if rc >= 0:
        return rc
    msg = (lib.misa_last_error() or b"").decode()
    if rc in (MISA_ERR_ARG, MISA_ERR_LENGTHS, MISA_ERR_INDEX):
        raise ValueError(msg)
    raise EngineError(f"misamplers error {rc}: {msg}")


def ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data) if a.size else ctypes.c_void_p(0)


def read_seeds(seed: int, num_reads: int) -> np.ndarray:
    """Per-read xorshift128+ seeds: output r of splitmix64 started at ``seed``."""
    out = np.empty(max(int(num_reads), 0), dtype=np.uint64)
    if out.size:
        lib.misa_read_seeds(ctypes.c_uint64(int(seed) & 0xFFFFFFFFFFFFFFFF), out.size, ptr(out))
    return out


def random_spins(seed: int, count: int) -> np.ndarray:
    """The +-1 spins ``misa_random_spins_device`` draws on the device for
    ``seed``, computed on the host."""
    out = np.empty(max(int(count), 0), dtype=np.int8)
    if out.size:
        lib.misa_random_spins(ctypes.c_uint64(int(seed) & 0xFFFFFFFFFFFFFFFF), out.size, ptr(out))
    return out


class Problem:
    """Owning handle of a ``misa_problem`` (h plus couplers in input order)."""

    def __init__(self, h, coupler_starts, coupler_ends, coupler_weights):
        self.h = np.ascontiguousarray(h, dtype=np.float64).ravel()
        self.cs = np.ascontiguousarray(coupler_starts, dtype=np.int32).ravel()
        self.ce = np.ascontiguousarray(coupler_ends, dtype=np.int32).ravel()
        self.cw = np.ascontiguousarray(coupler_weights, dtype=np.float64).ravel()
        handle = ctypes.c_void_p()
        check(lib.misa_problem_create(len(self.h), ptr(self.h), len(self.cs), ptr(self.cs),
                                      len(self.ce), ptr(self.ce), len(self.cw), ptr(self.cw),
                                      ctypes.byref(handle)))
        self.handle = handle

    @property
    def num_vars(self) -> int:
        return len(self.h)

    @property
    def auto_mode(self) -> int:
        return lib.misa_problem_auto_mode(self.handle)

    @property
    def max_degree(self) -> int:
        return lib.misa_problem_max_degree(self.handle)

    def close(self):
        if getattr(self, "handle", None):
            lib.misa_problem_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lds_bandwidth(iters=4096):
    """(GB/s, kernel ms) of the device's LDS read bandwidth (misa_lds_bandwidth)."""
    gbs, ms = ctypes.c_double(), ctypes.c_double()
    check(lib.misa_lds_bandwidth(int(iters), ctypes.byref(gbs), ctypes.byref(ms)))
    return gbs.value, ms.value
