"""ORACLE — test infrastructure only.

Exact equilibrium of the transverse-field Ising model on a few spins, the
mathematical anchor of the path-integral sampler's statistical parity:

* the P-slice Trotter model the engine samples,
      Z_P = Tr[(exp(-(Hp/P) Hz) exp((Hd/P) sum_i X_i))^P],
  by its transfer matrix (exact for every P);
* the quantum model both the engine (as P grows) and the reference's
  continuous-time worldlines (localPIMC.cpp, 2^16 slices) approximate,
      rho = exp(-Hp Hz + Hd sum_i X_i) / Z,
  by dense diagonalisation;
* the classical Boltzmann law exp(-Hp E(s)) (Hd = 0).

Hz = sum_i h_i Z_i + sum_c J_c Z_u Z_v (localPIMC.cpp:18-20 with the fields
already scaled, sqa/sampler.py:196-206).  Basis states are integers whose bit i
is set when spin i is +1, the engine's worldline bit convention.
"""
from __future__ import annotations

import numpy as np


def spins(n: int) -> np.ndarray:
    """(2^n, n) array of +-1 spins, row s has bit i of s -> spin i."""
    idx = np.arange(1 << n)[:, None]
    return np.where((idx >> np.arange(n)[None, :]) & 1, 1, -1).astype(np.int64)


def energies(h, cs, ce, cw) -> np.ndarray:
    h = np.asarray(h, dtype=np.float64)
    S = spins(len(h)).astype(np.float64)
    E = S @ h
    for u, v, w in zip(cs, ce, cw):
        E = E + w * S[:, u] * S[:, v]
    return E


def _hamming(n):
    idx = np.arange(1 << n)
    x = idx[:, None] ^ idx[None, :]
    d = np.zeros_like(x)
    for i in range(n):
        d += (x >> i) & 1
    return d


class TrotterModel:
    """Transfer matrix of the P-slice model at fixed (Hp, Hd)."""

    def __init__(self, h, cs, ce, cw, hp: float, hd: float, P: int):
        self.n = n = len(h)
        self.P = P
        E = energies(h, cs, ce, cw)
        eps = hd / P
        dh = _hamming(n)
        # T[s, s'] = prod_i (cosh eps if equal else sinh eps), normalised by cosh^n.
        T = np.tanh(eps) ** dh if eps > 0 else (dh == 0).astype(np.float64)
        w = np.exp(-(hp / P) * (E - E.min()) / 2.0)
        self.D2 = w
        self.M = w[:, None] * T * w[None, :]
        self.T = T
        self.dh = dh
        self.MP1 = np.linalg.matrix_power(self.M, P - 1)
        self.MP = self.MP1 @ self.M
        self.Z = np.trace(self.MP)

    def slice_distribution(self) -> np.ndarray:
        """Probability of each basis state on one slice."""
        return np.diag(self.MP) / self.Z

    def mean_z(self) -> np.ndarray:
        return self.slice_distribution() @ spins(self.n)

    def zz(self) -> np.ndarray:
        S = spins(self.n).astype(np.float64)
        p = self.slice_distribution()
        return (S * p[:, None]).T @ S

    def kink_probability(self) -> np.ndarray:
        """P(s_i differs between two adjacent slices), per site."""
        out = np.zeros(self.n)
        for i in range(self.n):
            flip = ((np.arange(1 << self.n)[:, None] ^ np.arange(1 << self.n)[None, :]) >> i) & 1
            Mi = self.M * flip
            out[i] = np.trace(self.MP1 @ Mi) / self.Z
        return out


class QuantumModel:
    """rho = exp(-Hp Hz + Hd sum X) / Z by diagonalisation."""

    def __init__(self, h, cs, ce, cw, hp: float, hd: float):
        self.n = n = len(h)
        E = energies(h, cs, ce, cw)
        dh = _hamming(n)
        H = np.diag(hp * E) - hd * (dh == 1).astype(np.float64)
        vals, vecs = np.linalg.eigh(H)
        wts = np.exp(-(vals - vals.min()))
        self.rho = (vecs * wts[None, :]) @ vecs.T / wts.sum()

    def slice_distribution(self) -> np.ndarray:
        return np.diag(self.rho).copy()

    def zz(self) -> np.ndarray:
        S = spins(self.n).astype(np.float64)
        p = self.slice_distribution()
        return (S * p[:, None]).T @ S


def chain_couplers(n: int, qubits_per_chain: int, strength: float):
    """(cs, ce, cw) of the implicit chain couplings of chain updates: qubits
    c*Q + k and c*Q + k + 1 for every chain c < n // Q (localPIMC.cpp:385-390
    couples consecutive qubits of a chain with Jchain)."""
    Q = qubits_per_chain
    cs = [c * Q + k for c in range(n // Q) for k in range(Q - 1)]
    return cs, [u + 1 for u in cs], [float(strength)] * len(cs)


def boltzmann(h, cs, ce, cw, beta: float) -> np.ndarray:
    E = energies(h, cs, ce, cw)
    w = np.exp(-beta * (E - E.min()))
    return w / w.sum()


def state_index(states) -> np.ndarray:
    """Basis index of +-1 state rows (bit i set when spin i is +1)."""
    s = np.asarray(states)
    return ((s > 0).astype(np.int64) << np.arange(s.shape[1])[None, :]).sum(axis=1)
