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
