"""ORACLE — test infrastructure only (tests/ may import it; the engine never
does).  Exact Boltzmann quantities for checking bucket elimination
(the reference's tree decomposition sampler, dwave/samplers/tree/samplers.py:
186-359, whose engine needs the dimod C++ headers and cannot be built here):

* ``enumerate_exact``: brute force over all 2^n states (n <= 22): log Z, the
  variable marginals P(x_i = 1) and the pair marginals in the order
  (low, low), (1, low), (low, 1), (1, 1);
* ``strip_exact``: a w x L strip (rows r, columns l) with fields, couplers
  along rows (column l to l+1) and within columns (row r to r+1), by a
  forward-backward transfer matrix over the 2^w column states (log domain).
"""
from __future__ import annotations

import itertools

import numpy as np


def _lse(a, axis=None):
    m = np.max(a, axis=axis, keepdims=True)
    out = m + np.log(np.sum(np.exp(a - m), axis=axis, keepdims=True))
    return np.squeeze(out, axis=axis) if axis is not None else float(out.reshape(()))


def enumerate_exact(h, cs, ce, cw, beta=1.0, low=-1, offset=0.0):
    h = np.asarray(h, dtype=np.float64)
    n = len(h)
    X = np.array(list(itertools.product([low, 1], repeat=n)), dtype=np.float64)
    E = X @ h + offset
    for u, v, w in zip(cs, ce, cw):
        E = E + w * X[:, u] * X[:, v]
    lw = -beta * E
    lz = _lse(lw)
    p = np.exp(lw - lz)
    p1 = np.array([p[X[:, i] == 1].sum() for i in range(n)])
    pairs = []
    for u, v in zip(cs, ce):
        pairs.append([p[(X[:, u] == a) & (X[:, v] == b)].sum() for a, b in ((low, low), (1, low), (low, 1), (1, 1))])
    return lz, p1, np.array(pairs).reshape(-1, 4)


def strip_problem(w, L, rng, low=-1, scale=1.0):
    """A random w x L strip: variable (r, l) -> r + w * l; couplers within
    columns first, then along rows, in that order."""
    h = rng.normal(0, scale, size=(w, L))
    Jv = rng.normal(0, scale, size=(max(w - 1, 0), L))
    Jh = rng.normal(0, scale, size=(w, max(L - 1, 0)))
    cs, ce, cw = [], [], []
    for l in range(L):
        for r in range(w - 1):
            cs.append(r + w * l)
            ce.append(r + 1 + w * l)
            cw.append(Jv[r, l])
    for l in range(L - 1):
        for r in range(w):
            cs.append(r + w * l)
            ce.append(r + w * (l + 1))
            cw.append(Jh[r, l])
    return h, Jv, Jh, (h.T.ravel(), np.array(cs, np.int32), np.array(ce, np.int32), np.array(cw))


def strip_exact(h, Jv, Jh, beta=1.0, low=-1):
    """log Z, P(x=1) as a (w, L) array, and the pair marginals of the column
    couplers (w-1, L, 4) and row couplers (w, L-1, 4)."""
    w, L = h.shape
    S = 1 << w
    X = np.array([[1.0 if (s >> r) & 1 else float(low) for r in range(w)] for s in range(S)])  # (S, w)
    phi = np.zeros((L, S))
    for l in range(L):
        phi[l] = -beta * (X @ h[:, l])
        for r in range(w - 1):
            phi[l] += -beta * Jv[r, l] * X[:, r] * X[:, r + 1]
    psi = np.zeros((max(L - 1, 0), S, S))
    for l in range(L - 1):
        psi[l] = -beta * ((X * Jh[:, l]) @ X.T)
    alpha = np.zeros((L, S))
    back = np.zeros((L, S))
    alpha[0] = phi[0]
    for l in range(1, L):
        alpha[l] = phi[l] + _lse(alpha[l - 1][:, None] + psi[l - 1], axis=0)
    for l in range(L - 2, -1, -1):
        back[l] = _lse(psi[l] + (phi[l + 1] + back[l + 1])[None, :], axis=1)
    lz = _lse(alpha[L - 1])
    node = np.exp(alpha + back - lz)  # (L, S)
    p1 = np.array([[node[l][X[:, r] == 1].sum() for l in range(L)] for r in range(w)])
    cfg = ((low, low), (1, low), (low, 1), (1, 1))
    pv = np.zeros((max(w - 1, 0), L, 4))
    for l in range(L):
        for r in range(w - 1):
            pv[r, l] = [node[l][(X[:, r] == a) & (X[:, r + 1] == b)].sum() for a, b in cfg]
    ph = np.zeros((w, max(L - 1, 0), 4))
    for l in range(L - 1):
        joint = np.exp(alpha[l][:, None] + psi[l] + (phi[l + 1] + back[l + 1])[None, :] - lz)  # (S, S)
        for r in range(w):
            ph[r, l] = [joint[np.ix_(X[:, r] == a, X[:, r] == b)].sum() for a, b in cfg]
    return lz, p1, pv, ph
