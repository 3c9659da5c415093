"""Pins the FP64 gradient oracle (oracle_logless_float_grad, the adjoint of the LoglessPairHMM
recurrence with continuous Phred parameters) three independent ways:
  * central finite differences of the FP64 forward at a 1e-3 Phred step;
  * torch.autograd of an independent anti-diagonal torch restatement of the forward;
  * torch.autograd.gradcheck of that torch restatement on 32 sampled pairs;
plus the closed-form check of an exact-match read, where d log10 L / d e_i ~ -1 / ((1 - e_i) ln 10).
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle


def _pair(rng, H, R, sub=0.05):
    hap = rng.choice(list(b"ACGT"), H).astype(np.uint8)
    st = int(rng.integers(0, max(1, H - R)))
    read = np.resize(hap[st:st + R], R).copy()
    flip = rng.random(R) < sub
    read[flip] = rng.choice(list(b"ACGT"), int(flip.sum())).astype(np.uint8)
    q = (rng.uniform(10, 40, R), rng.uniform(30, 50, R), rng.uniform(30, 50, R), rng.uniform(8, 15, R))
    return hap, read, q


def _fd(hap, read, q, which, k, step=1e-3):
    qs = [x.copy() for x in q]
    qs[which][k] += step
    fp = oracle.logless_float(hap, read, *qs)
    qs[which][k] -= 2 * step
    fm = oracle.logless_float(hap, read, *qs)
    return (fp - fm) / (2 * step)


@pytest.mark.parametrize("seed", range(4))
def test_adjoint_matches_finite_differences(seed):
    rng = np.random.default_rng(seed)
    hap, read, q = _pair(rng, int(rng.integers(20, 60)), int(rng.integers(5, 30)))
    g = oracle.logless_float_grad(hap, read, *q)
    names = ("qb", "qi", "qd", "qc")
    for which, name in enumerate(names):
        fd = np.array([_fd(hap, read, q, which, k) for k in range(len(read))])
        err = np.linalg.norm(g[name] - fd) / max(np.linalg.norm(fd), 1e-300)
        assert err <= 1e-3, (name, err)
    # forward value equals the FP64 forward
    assert abs(g["log10_l"] - oracle.logless_float(hap, read, *q)) <= 1e-12


@pytest.mark.parametrize("seed", range(3))
def test_adjoint_matches_torch_autograd(seed):
    rng = np.random.default_rng(10 + seed)
    hap, read, q = _pair(rng, int(rng.integers(20, 50)), int(rng.integers(5, 25)))
    g = oracle.logless_float_grad(hap, read, *q)
    t = [torch.tensor(x, dtype=torch.float64, requires_grad=True) for x in q]
    L = oracle.torch_log10_likelihood(hap, read, *t)
    assert abs(L.item() - g["log10_l"]) <= 1e-12
    L.backward()
    for name, tt in zip(("qb", "qi", "qd", "qc"), t):
        err = np.max(np.abs(tt.grad.numpy() - g[name])) / max(np.max(np.abs(g[name])), 1e-300)
        assert err <= 1e-10, (name, err)


def test_gradcheck_torch_reference_on_32_pairs():
    """torch.autograd.gradcheck of the FP64 CPU reference on 32 sampled pairs (fast mode: one
    random projection of the Jacobian against FP64 finite differences)."""
    rng = np.random.default_rng(2024)
    for k in range(32):
        hap, read, q = _pair(rng, int(rng.integers(8, 30)), int(rng.integers(3, 12)))
        t = tuple(torch.tensor(x, dtype=torch.float64, requires_grad=True) for x in q)
        f = lambda *a: oracle.torch_log10_likelihood(hap, read, *a)  # noqa: E731
        assert torch.autograd.gradcheck(f, t, eps=1e-4, atol=1e-6, rtol=1e-4, fast_mode=True), k


def test_exact_match_closed_form():
    """A read equal to the haplotype with very high gap penalties: the match path dominates and
    d log10 L / d e_i = -1 / ((1 - e_i) ln 10) at every position."""
    hap = np.frombuffer(b"ACGTTGCAAGTCCATG" * 3, np.uint8).copy()
    read = hap.copy()
    R = len(read)
    q = (np.full(R, 25.0), np.full(R, 90.0), np.full(R, 90.0), np.full(R, 90.0))
    g = oracle.logless_float_grad(hap, read, *q)
    e = 10 ** (-2.5)
    want = -1.0 / ((1 - e) * math.log(10))
    assert np.max(np.abs(g["eb"] / want - 1.0)) <= 1e-6


def test_continuous_forward_equals_logless_at_integer_phred():
    """At integer Phred the continuous model (MM = 1 - e_ins - e_del) is LoglessPairHMM itself: the
    Jacobian log-sum table behind matchToMatchProb is exact at integer Phred differences
    (MathUtils.java:406-423, PairHMMModel.java:373-388).  200 random pairs with |q_ins - q_del| < 80
    agree to 1e-10 with the pinned FP64 Logless restatement."""
    rng = np.random.default_rng(77)
    worst = 0.0
    for _ in range(200):
        H, R = int(rng.integers(10, 120)), int(rng.integers(3, 90))
        hap, read, _ = _pair(rng, H, R)
        qb = rng.integers(6, 41, R)
        qi = rng.integers(20, 60, R)
        qd = rng.integers(20, 60, R)
        qc = rng.integers(5, 20, R)
        want = oracle.logless(hap, read, qb, qi, qd, qc)
        got = oracle.logless_float(hap, read, *(x.astype(np.float64) for x in (qb, qi, qd, qc)))
        worst = max(worst, abs(got - want))
    assert worst <= 1e-10, worst


def test_continuous_forward_at_the_table_drop_rule():
    """|q_ins - q_del| >= 80: matchToMatchProb drops the smaller error term (MathUtils.java:409,479),
    so Java's MM exceeds the continuous 1 - e_ins - e_del by e_small on those rows.  Everything else
    coincides, so 0 < log10 L_java - log10 L_float <= sum over rows of e_small / (MM ln 10)."""
    rng = np.random.default_rng(78)
    for _ in range(20):
        H, R = int(rng.integers(20, 80)), int(rng.integers(5, 40))
        hap, read, _ = _pair(rng, H, R)
        qb = rng.integers(10, 41, R)
        qi = rng.integers(5, 15, R)
        qd = qi + rng.integers(80, 100, R)
        qc = rng.integers(5, 20, R)
        java = oracle.logless(hap, read, qb, qi, qd, qc)
        cont = oracle.logless_float(hap, read, *(x.astype(np.float64) for x in (qb, qi, qd, qc)))
        e_small = 10.0 ** (-qd / 10.0)
        mm = 1.0 - 10.0 ** (-qi / 10.0) - e_small
        bound = float(np.sum(e_small / mm)) / math.log(10)
        assert 0.0 < java - cont <= bound * (1 + 1e-6) + 1e-13, (java - cont, bound)
        # one step inside the rule (|q_ins - q_del| = 79) nothing is dropped and the two agree
        assert abs(oracle.logless(hap, read, qb, qi, qi + 79, qc)
                   - oracle.logless_float(hap, read, *(x.astype(np.float64) for x in (qb, qi, qi + 79, qc)))) <= 1e-10
