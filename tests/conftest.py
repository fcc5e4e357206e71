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
