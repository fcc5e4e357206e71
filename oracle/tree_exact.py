"""ORACLE — test infrastructure only (tests/ may import it; the engine never
does).  Exact Boltzmann quantities for checking bucket elimination
(the reference's tree decomposition sampler, dwave/samplers/tree/samplers.py:
186-359, whose engine needs the dimod C++ headers and cannot be built here):

* ``enumerate_exact``: brute force over all 2^n states (n <= 22): log Z, the
  variable marginals P(x_i = 1) and the pair marginals in the order
  (low, low), (1, low), (low, 1), (1, 1);
* ``bucket_eliminate``: bucket elimination with numpy tables, one array axis
  per scope variable: log Z, and with the downward pass every variable and
  pair marginal (the CPU timing of the bench, and a check where neither
  enumeration nor a transfer matrix reaches);
* ``strip_exact``: a w x L strip (rows r, columns l) with fields, couplers
  along rows (column l to l+1) and within columns (row r to r+1), by a
  forward-backward transfer matrix over the 2^w column states (log domain);
* ``kbest_bruteforce``: the k lowest energies of all 2^n states, sorted (the
  definition of the solver's output, dwave/samplers/tree/samplers.py:29-183);
* ``kbest_eliminate``: the same k energies by variable elimination with
  sorted lists of at most k values per factor entry (min-sum with k-best
  lists), for problems too large to enumerate; checked against
  ``kbest_bruteforce`` in the tests.
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


def bucket_eliminate(h, cs, ce, cw, order, beta=1.0, low=-1, marginals=True):
    """Bucket elimination in ``order``.  Bucket q eliminates order[q]; its table
    T_q has one axis per scope variable (the eliminated one first, then its
    neighbours at elimination time by elimination position) and sums its own
    terms and its children's messages mu_c (log-sum-exp of T_c over axis 0).
    The downward message of a child is the log-sum-exp of its parent's
    T_p + delta_p over the parent axes outside the child's separator, minus
    mu_c; T_q + delta_q is then the log of the unnormalised scope marginal.
    Repeated couplers are added; self-loops are constants (spin) or fields
    (binary).  Returns log Z, or (log Z, P(x_i = 1) per variable, the four
    cells (low, low), (1, low), (low, 1), (1, 1) per merged coupler in
    first-appearance order) when ``marginals``."""
    from scipy.special import logsumexp
    n = len(h)
    h = np.array(h, dtype=np.float64)
    const = 0.0
    pw, first = {}, []
    for u, v, w in zip(cs, ce, cw):
        u, v = int(u), int(v)
        if u == v:
            if low == -1:
                const += w
            else:
                h[u] += w
            continue
        key = (min(u, v), max(u, v))
        if key not in pw:
            first.append((u, v))
            pw[key] = 0.0
        pw[key] += w
    at = {v: q for q, v in enumerate(order)}
    adj = [set() for _ in range(n)]
    for (u, v) in pw:
        adj[u].add(v)
        adj[v].add(u)
    vals = np.array([low, 1.0])
    pending = [[] for _ in range(n)]   # (scope, table) factors by bucket
    for (u, v), w in pw.items():
        pending[min(at[u], at[v])].append(((u, v), -beta * w * np.multiply.outer(vals, vals)))
    for v in range(n):
        pending[at[v]].append(((v,), -beta * h[v] * vals))
    scopes, T, mu, parent, kids = [], [], [], [], [[] for _ in range(n)]
    for q, v in enumerate(order):
        nb = sorted(adj[v], key=lambda x: at[x])
        scope = [v] + nb
        for i in range(len(nb)):
            for j in range(i + 1, len(nb)):
                adj[nb[i]].add(nb[j])
                adj[nb[j]].add(nb[i])
        for u in nb:
            adj[u].discard(v)
        pos = {x: i for i, x in enumerate(scope)}
        t = np.zeros((2,) * len(scope))
        for sc, tab in pending[q]:
            # the factor's axes moved to their scope positions, the rest broadcast
            perm = sorted(range(len(sc)), key=lambda i: pos[sc[i]])
            shape = [1] * len(scope)
            for i in perm:
                shape[pos[sc[i]]] = 2
            t = t + np.transpose(tab, perm).reshape(shape)
        m = np.logaddexp(t[0], t[1])
        scopes.append(scope)
        T.append(t)
        mu.append(m)
        parent.append(at[nb[0]] if nb else -1)
        if nb:
            pending[at[nb[0]]].append((tuple(nb), m))
            kids[at[nb[0]]].append(q)
    lz = float(sum(float(mu[q]) for q in range(n) if parent[q] < 0)) - beta * const
    if not marginals:
        return lz
    delta = [None] * n
    for q in range(n - 1, -1, -1):
        if parent[q] < 0:
            delta[q] = np.zeros(())
            continue
        p = parent[q]
        belief = T[p] + delta[p]            # delta_p broadcasts over axis 0
        keep = set(scopes[q][1:])
        out = tuple(i for i, x in enumerate(scopes[p]) if x not in keep)
        delta[q] = logsumexp(belief, axis=out) - mu[q]
    beliefs = {}

    def belief(q):                             # normalised scope marginal
        if q not in beliefs:
            b = T[q] + delta[q]
            beliefs[q] = np.exp(b - logsumexp(b))
        return beliefs[q]

    p1 = np.zeros(n)
    for q, v in enumerate(order):
        p1[v] = belief(q)[1].sum()
    cells = np.zeros((len(first), 4))
    for k, (u, v) in enumerate(first):
        q = min(at[u], at[v])
        b = belief(q)
        iu, iv = scopes[q].index(u), scopes[q].index(v)
        other = tuple(i for i in range(b.ndim) if i not in (iu, iv))
        m2 = b.sum(axis=other)                 # axes (min(iu, iv), max(iu, iv))
        if iu > iv:
            m2 = m2.T                          # rows u, columns v
        cells[k] = [m2[0, 0], m2[1, 0], m2[0, 1], m2[1, 1]]
    return lz, p1, cells


def kbest_bruteforce(h, cs, ce, cw, low=-1, k=1, offset=0.0):
    """(the k smallest energies ascending, every state's energy, the states)."""
    h = np.asarray(h, dtype=np.float64)
    n = len(h)
    X = np.array(list(itertools.product([low, 1], repeat=n)), dtype=np.float64)
    E = X @ h + offset
    for u, v, w in zip(cs, ce, cw):
        E = E + w * X[:, u] * X[:, v]
    return np.sort(E)[:k], E, X


def _kbest_sum(a, b, k):
    """The k smallest a_i + b_j of two value lists, ascending."""
    return sorted(x + y for x in a for y in b)[:k]


def kbest_eliminate(h, cs, ce, cw, low=-1, k=1, offset=0.0):
    """The k smallest energies by min-sum elimination with k-best lists.

    Factors map an assignment of their scope (a tuple of 0 / 1 per scope
    variable, 1 = the high value) to a sorted list of at most k partial
    energies.  Eliminating v combines the factors holding v (k-best sums)
    for both values of v and keeps the k smallest of the two lists; variables
    go in greedy min-degree order (any order gives the same energies)."""
    h = np.asarray(h, dtype=np.float64)
    n = len(h)
    val = (low, 1)
    factors = []
    for v in range(n):
        factors.append(((v,), {(0,): [h[v] * low], (1,): [h[v]]}))
    for u, v, w in zip(cs, ce, cw):
        u, v = int(u), int(v)
        if u == v:
            factors.append(((u,), {(x,): [w * val[x] * val[x]] for x in (0, 1)}))
        else:
            factors.append(((u, v), {(x, y): [w * val[x] * val[y]] for x in (0, 1) for y in (0, 1)}))
    remaining = set(range(n))
    const_lists = [[offset]]
    while remaining:
        # fewest distinct neighbours among the factors holding it
        def degree(v):
            nb = set()
            for sc, _ in factors:
                if v in sc:
                    nb.update(sc)
            return len(nb)
        v = min(remaining, key=lambda x: (degree(x), x))
        remaining.discard(v)
        holding = [f for f in factors if v in f[0]]
        factors = [f for f in factors if v not in f[0]]
        scope = sorted(set().union(*[set(sc) for sc, _ in holding]) - {v})
        table = {}
        for assign in itertools.product((0, 1), repeat=len(scope)):
            env = dict(zip(scope, assign))
            best = []
            for xv in (0, 1):
                env[v] = xv
                cur = [0.0]
                for sc, t in holding:
                    cur = _kbest_sum(cur, t[tuple(env[x] for x in sc)], k)
                best = sorted(best + cur)[:k]
            table[tuple(assign)] = best
        if scope:
            factors.append((tuple(scope), table))
        else:
            const_lists.append(table[()])
    cur = [0.0]
    for lst in const_lists:
        cur = _kbest_sum(cur, lst, k)
    return np.array(cur)
