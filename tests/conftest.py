import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "dwave-samplers_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def golden_cases(prefix="sa_"):
    out = {}
    for path in sorted(glob.glob(os.path.join(GOLDEN, prefix + "*.npz"))):
        with np.load(path, allow_pickle=False) as z:
            out[os.path.basename(path)[len(prefix):-4]] = {k: z[k] for k in z.files}
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


def random_regular(n, d, rng):
    """Edges of a random d-regular simple graph on n nodes (n*d even): the
    configuration model with invalid pairs re-drawn until none remain."""
    stubs = np.repeat(np.arange(n), d)
    rng.shuffle(stubs)
    pairs = stubs.reshape(-1, 2)
    for _ in range(1000):
        a, b = np.minimum(pairs[:, 0], pairs[:, 1]), np.maximum(pairs[:, 0], pairs[:, 1])
        key = a.astype(np.int64) * n + b
        _, first = np.unique(key, return_index=True)
        bad = np.ones(len(pairs), bool)
        bad[first] = False
        bad |= a == b
        if not bad.any():
            return a.astype(np.int32), b.astype(np.int32)
        idx = np.flatnonzero(bad)
        other = rng.choice(len(pairs), size=len(idx), replace=False)
        sel = np.unique(np.concatenate([idx, other]))
        flat = pairs[sel].ravel()
        rng.shuffle(flat)
        pairs[sel] = flat.reshape(-1, 2)
    raise RuntimeError("could not build a simple regular graph")
