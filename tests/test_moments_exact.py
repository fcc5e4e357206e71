"""CPU checks of the mathematics behind differentiable fixed-beta sampling:
with E = sum_i h_i s_i + sum_c J_c s_u s_v and F = -(1/beta) log Z, exact
enumeration on 10 spins gives dF/dh_i = <s_i> and dF/dJ_c = <s_u s_v> by
central differences of log Z, and the surrogate (1/R) sum_r E(s_r) over fixed
samples has exactly the sample moments as gradient."""
import numpy as np

from oracle import pimc_exact as px


def _logZ(h, cs, ce, cw, beta):
    E = px.energies(h, cs, ce, cw)
    a = -beta * E
    return a.max() + np.log(np.exp(a - a.max()).sum())


def test_free_energy_gradients_are_moments():
    rng = np.random.default_rng(0)
    n = 10
    pairs = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < 0.5]
    cs = np.array([a for a, _ in pairs])
    ce = np.array([b for _, b in pairs])
    h = rng.normal(0, 0.5, n)
    J = rng.normal(0, 0.5, len(cs))
    for beta in (0.7, 1.3):
        p = px.boltzmann(h, cs, ce, J, beta)
        S = px.spins(n).astype(float)
        mh = p @ S
        mj = p @ (S[:, cs] * S[:, ce])
        F = lambda hh, jj: -_logZ(hh, cs, ce, jj, beta) / beta
        eps = 1e-5
        for i in range(n):
            d = np.zeros(n)
            d[i] = eps
            assert abs((F(h + d, J) - F(h - d, J)) / (2 * eps) - mh[i]) < 1e-7
        for c in range(len(cs)):
            d = np.zeros(len(cs))
            d[c] = eps
            assert abs((F(h, J + d) - F(h, J - d)) / (2 * eps) - mj[c]) < 1e-7


def test_sample_energy_surrogate_gradient():
    rng = np.random.default_rng(1)
    n, R = 6, 50
    cs, ce = np.array([0, 1, 2, 3, 4]), np.array([1, 2, 3, 4, 5])
    s = rng.choice([-1, 1], size=(R, n)).astype(float)
    h = rng.normal(size=n)
    J = rng.normal(size=5)
    L = lambda hh, jj: np.mean(s @ hh + (s[:, cs] * s[:, ce]) @ jj)
    eps = 1e-6
    for i in range(n):
        d = np.zeros(n)
        d[i] = eps
        assert abs((L(h + d, J) - L(h - d, J)) / (2 * eps) - s[:, i].mean()) < 1e-8
    for c in range(5):
        d = np.zeros(5)
        d[c] = eps
        assert abs((L(h, J + d) - L(h, J - d)) / (2 * eps) - (s[:, cs[c]] * s[:, ce[c]]).mean()) < 1e-8
