"""The single-precision acceptance tests of the SA kernels (sa_real_block.h
metropolis_accept / gibbs_accept, the same margins in sa_dense_kernel.hip)
against the reference's double expressions (cpu_sa.cpp:154, :163).

The float path is emulated in numpy float32, with the hardware exp2's
result moved by up to one ulp either way.  Draws x are taken uniformly and
right at the decision boundary, where a margin that is too thin, a swapped
margin or a missing + 1 in xh + 1 shows up as a disagreement.  A decision
the float path does not settle goes to the double expression in the kernels,
so only settled decisions are compared.
"""
import numpy as np

F = np.float32
TWO64 = 18446744073709551616.0
LOG2E = F(1.44269504)
HI = F(1.00006104)  # 1 + 2^-14 (rounded to float)
LO = F(0.99993896)  # 1 - 2^-14


def exp2_hw(z, rng):
    """float32 exp2 with an error of up to one ulp either way."""
    e = np.exp2(z.astype(np.float64)).astype(F)
    step = rng.integers(-1, 2, size=e.shape)
    e = np.where(step > 0, np.nextafter(e, F(np.inf)), e)
    return np.where(step < 0, np.nextafter(e, F(0)), e).astype(F)


def metropolis_settled(t, x, rng, lo=LO, hi=HI, one=F(1.0)):
    xh = (x >> np.uint64(40)).astype(np.uint32).astype(F)
    a = exp2_hw(t.astype(F) * LOG2E, rng) * F(16777216.0)
    acc = a * lo > xh + one
    rej = a * hi < xh
    return acc, rej


def gibbs_settled(t, x, rng, lo=LO, hi=HI, one=F(1.0)):
    xh = (x >> np.uint64(40)).astype(np.uint32).astype(F)
    g = F(1.0) + exp2_hw(t.astype(F) * LOG2E, rng)
    flip = (xh + one) * g * hi < F(16777216.0)
    none = xh * g * lo > F(16777216.0)
    return flip, none


def boundary_draws(p, rng, spread):
    """x near p 2^64 (p in (0, 1]): the boundary moved by a relative offset of
    up to spread, and by a few units, clipped to the 64-bit range."""
    base = p * TWO64
    v = base * (1.0 + rng.uniform(-spread, spread, size=p.shape)) + rng.integers(-3, 4, size=p.shape)
    v = np.clip(np.floor(v), 0, TWO64 - 4096)
    return v.astype(np.uint64)


def check(kind, n, rng, lo=LO, hi=HI, one=F(1.0)):
    if kind == "metropolis":
        t = -rng.uniform(0, 44.37, size=n)
        t[: n // 8] = -rng.uniform(0, 1e-3, size=n // 8)  # exp(t) near 1
        p = np.exp(t)
    else:
        t = rng.uniform(-60, 44.37, size=n)
        p = 1.0 / (1.0 + np.exp(t))
    x = np.concatenate([rng.integers(0, 2**63, size=n, dtype=np.uint64) * np.uint64(2) + np.uint64(1),
                        boundary_draws(p, rng, 1e-2), boundary_draws(p, rng, 1e-4), boundary_draws(p, rng, 2e-6)])
    t = np.concatenate([t, t, t, t])
    xd = x.astype(np.float64)  # (double)x, round to nearest as the reference
    if kind == "metropolis":
        exact = np.exp(t) * TWO64 > xd
        yes, no = metropolis_settled(t, x, rng, lo, hi, one)
    else:
        exact = TWO64 > xd * (1.0 + np.exp(t))
        yes, no = gibbs_settled(t, x, rng, lo, hi, one)
    assert not np.any(yes & no)
    wrong = np.count_nonzero(yes & ~exact) + np.count_nonzero(no & exact)
    settled_uniform = np.count_nonzero((yes | no)[:n]) / n
    return wrong, settled_uniform


def test_metropolis_margin_agrees_with_double():
    rng = np.random.default_rng(11)
    wrong, settled = check("metropolis", 400_000, rng)
    assert wrong == 0
    assert settled > 0.999  # the double expression is the rare case


def test_gibbs_margin_agrees_with_double():
    rng = np.random.default_rng(12)
    wrong, settled = check("gibbs", 400_000, rng)
    assert wrong == 0
    assert settled > 0.999


def test_planted_mistakes_disagree():
    """The boundary draws can tell: without the margin, or with x taken as
    xh 2^40 (the + 1 dropped), both tests settle draws the wrong way."""
    rng = np.random.default_rng(13)
    for kind in ("metropolis", "gibbs"):
        assert check(kind, 400_000, rng, F(1.0), F(1.0))[0] > 0
        assert check(kind, 400_000, rng, one=F(0.0))[0] > 0
