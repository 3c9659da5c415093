"""GPU parity of the forward + backward (adjoint) kernel through the C-ABI against the FP64 gradient
oracle (itself pinned by finite differences, torch autograd and gradcheck in test_oracle_grad.py).

Bars: log10 L within 1e-4 (FP32 cells) / 1e-9 (FP64 kernel); per-read gradient relative error
(norm-wise, over the four parameter kinds) <= 1e-3 against the FP64 oracle and against FP64 central
finite differences at a 1e-3 Phred step; <= 1e-7 for the FP64 kernel.
"""
import numpy as np
import pytest
import torch

from gatk_amd import synth
from gatk_amd.autograd import GradEngine, GradProblem, pairhmm_log10_likelihood
from gatk_amd.batch import pack_regions
from oracle import oracle

pytestmark = pytest.mark.gpu

NAMES = ("qb", "qi", "qd", "qc")


def _problem(seed, n_regions=3, hap_len=(20, 360), read_len=(5, 200), sub=0.03):
    rng = np.random.default_rng(seed)
    regions = []
    for _ in range(n_regions):
        W = int(rng.integers(*hap_len))
        ref = rng.choice(list(b"ACGT"), W).astype(np.uint8)
        haps = []
        for _ in range(int(rng.integers(1, 4))):
            h = ref.copy()
            pos = rng.integers(0, W, 2)
            h[pos] = rng.choice(list(b"ACGT"), 2).astype(np.uint8)
            haps.append(h)
        reads = []
        for _ in range(int(rng.integers(1, 4))):
            R = int(rng.integers(*read_len))
            st = int(rng.integers(0, max(1, W - R)))
            b = np.resize(ref[st:st + R], R).copy()
            flip = rng.random(R) < sub
            b[flip] = rng.choice(list(b"ACGT"), int(flip.sum())).astype(np.uint8)
            z = np.zeros(R, np.uint8)
            reads.append((b, z, z, z, z))
        regions.append((haps, reads))
    batch = pack_regions(regions)
    n = len(batch.read_bases)
    q = {"qb": rng.uniform(10, 40, n), "qi": rng.uniform(30, 50, n), "qd": rng.uniform(30, 50, n),
         "qc": rng.uniform(8, 15, n)}
    q = {k: v.astype(np.float32) for k, v in q.items()}  # the kernel reads float32 Phred values
    return batch, q


def _oracle(batch, q, weights=None):
    n = len(batch.read_bases)
    out = np.zeros(batch.n_pairs)
    grads = {k: np.zeros(n) for k in NAMES + ("eb",)}
    for p in range(batch.n_pairs):
        r, h = batch.pair_read[p], batch.pair_hap[p]
        o, R = int(batch.read_off[r]), int(batch.read_len[r])
        ho, H = int(batch.hap_off[h]), int(batch.hap_len[h])
        sl = slice(o, o + R)
        g = oracle.logless_float_grad(batch.hap_bases[ho:ho + H], batch.read_bases[sl],
                                      *(q[k][sl].astype(np.float64) for k in NAMES))
        out[p] = g["log10_l"]
        w = 1.0 if weights is None else weights[p]
        for k in NAMES + ("eb",):
            grads[k][sl] += w * g[k]
    return out, grads


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _per_read_err(batch, got, want):
    worst = 0.0
    for r in range(len(batch.read_len)):
        o, R = int(batch.read_off[r]), int(batch.read_len[r])
        a = np.concatenate([got[k][o:o + R] for k in NAMES])
        b = np.concatenate([want[k][o:o + R] for k in NAMES])
        worst = max(worst, _rel(a, b))
    return worst


def _run(batch, q, use_double=False, weights=None):
    prob = GradProblem.from_batch(batch)
    tq = {k: torch.from_numpy(v).cuda() for k, v in q.items()}
    w = None if weights is None else torch.from_numpy(weights).cuda()
    with GradEngine(use_double=use_double) as eng:
        out, grads = eng.run(prob, tq["qb"], tq["qi"], tq["qd"], tq["qc"], weight=w, want_eb=True)
        st = eng.last_stats
    return out.cpu().numpy(), {k: v.cpu().numpy() for k, v in grads.items()}, st


@pytest.mark.parametrize("seed", [0, 1])
def test_fp32_adjoint_matches_oracle(seed):
    batch, q = _problem(seed)
    want_out, want = _oracle(batch, q)
    out, got, _ = _run(batch, q)
    assert np.max(np.abs(out - want_out)) <= 1e-4
    assert _per_read_err(batch, got, want) <= 1e-3
    assert _rel(got["eb"], want["eb"]) <= 1e-3


def test_fp64_adjoint_matches_oracle():
    batch, q = _problem(7)
    want_out, want = _oracle(batch, q)
    out, got, st = _run(batch, q, use_double=True)
    assert st.n_rescued == batch.n_pairs
    assert np.max(np.abs(out - want_out)) <= 1e-9
    assert _per_read_err(batch, got, want) <= 1e-7


def test_weighted_sum_of_pairs():
    batch, q = _problem(3)
    w = np.random.default_rng(0).uniform(-2, 2, batch.n_pairs)
    _, want = _oracle(batch, q, weights=w)
    _, got, _ = _run(batch, q, weights=w)
    assert _per_read_err(batch, got, want) <= 1e-3


def _single_pair_regions(seed, n, hap_len, read_len, sub=0.03):
    """n regions of one read and one haplotype each: every read belongs to exactly one pair, so a
    read's gradient is that pair's gradient and compares to the pair's own finite differences."""
    rng = np.random.default_rng(seed)
    regions = []
    for _ in range(n):
        W = int(rng.integers(*hap_len))
        hap = rng.choice(list(b"ACGT"), W).astype(np.uint8)
        R = int(rng.integers(*read_len))
        st = int(rng.integers(0, max(1, W - R)))
        b = np.resize(hap[st:st + R], R).copy()
        flip = rng.random(R) < sub
        b[flip] = rng.choice(list(b"ACGT"), int(flip.sum())).astype(np.uint8)
        z = np.zeros(R, np.uint8)
        regions.append(([hap], [(b, z, z, z, z)]))
    batch = pack_regions(regions)
    n_b = len(batch.read_bases)
    q = {"qb": rng.uniform(10, 40, n_b), "qi": rng.uniform(30, 50, n_b), "qd": rng.uniform(30, 50, n_b),
         "qc": rng.uniform(8, 15, n_b)}
    return batch, {k: v.astype(np.float32) for k, v in q.items()}


def test_against_finite_differences():
    """Single-pair regions against FP64 central differences of the forward (1e-3 Phred step)."""
    batch, q = _single_pair_regions(11, 4, hap_len=(40, 120), read_len=(20, 90))
    assert all(np.sum(batch.pair_read == r) == 1 for r in range(len(batch.read_len)))
    _, got, _ = _run(batch, q)
    compared = 0
    for p in range(batch.n_pairs):
        r, h = batch.pair_read[p], batch.pair_hap[p]
        o, R = int(batch.read_off[r]), int(batch.read_len[r])
        ho, H = int(batch.hap_off[h]), int(batch.hap_len[h])
        hap, read = batch.hap_bases[ho:ho + H], batch.read_bases[o:o + R]
        qs = [q[k][o:o + R].astype(np.float64) for k in NAMES]
        fd = []
        for which in range(4):
            col = []
            for k in range(R):
                a = [x.copy() for x in qs]
                a[which][k] += 1e-3
                fp = oracle.logless_float(hap, read, *a)
                a[which][k] -= 2e-3
                col.append((fp - oracle.logless_float(hap, read, *a)) / 2e-3)
            fd.append(np.array(col))
        g = np.concatenate([got[k][o:o + R] for k in NAMES])
        assert _rel(g, np.concatenate(fd)) <= 1e-3, p
        compared += 1
    assert compared >= 3


def test_batched_uneven_regions_fwd_bwd():
    """Forward + adjoint on pairs shaped like the 2,000-region differentiation workload: regions of
    the batched_uneven generator (haplotypes 150-600 bp, reads 100-250 bp), qb = the base
    qualities, qi = qd = 45, qc = 10.  A few reads per region, each against every kept haplotype,
    so per-read gradient sums cover several pairs; all against the FP64 oracle adjoint."""
    full = synth.workload("batched_uneven", n_regions=3, seed=3)
    regions = []
    for k in range(full.n_regions):
        reads, haps = full.region(k)
        regions.append((haps[:6], reads[:3]))
    batch = pack_regions(regions)
    assert batch.n_pairs >= 20 and int(batch.hap_len.max()) >= 150
    n = len(batch.read_bases)
    q = {"qb": batch.read_quals.astype(np.float32), "qi": np.full(n, 45.0, np.float32),
         "qd": np.full(n, 45.0, np.float32), "qc": np.full(n, 10.0, np.float32)}
    want_out, want = _oracle(batch, q)
    out, got, st = _run(batch, q)
    assert np.max(np.abs(out - want_out)) <= 1e-4
    assert _per_read_err(batch, got, want) <= 1e-3
    assert _rel(got["eb"], want["eb"]) <= 1e-3


def test_precision_stress_pairs_go_to_fp64():
    """Long, low-quality reads on STR-rich haplotypes: pairs whose scaled likelihood is below the
    FP32 adjoint range are recomputed by the FP64 kernel; results still match the oracle."""
    b = synth.workload("precision_stress", n_regions=1, seed=9)
    keep_reads = np.arange(min(3, len(b.read_len)))
    keep_haps = np.arange(min(2, len(b.hap_len)))
    reads, haps = b.region(0)
    # Haplotypes cut to 700 bp and reads to 300 bp keep the FP64 kernel's LDS slice within 64 KB.
    batch = pack_regions([([haps[h][:700] for h in keep_haps],
                           [tuple(x[:300] for x in reads[r]) for r in keep_reads])])
    rng = np.random.default_rng(4)
    n = len(batch.read_bases)
    q = {"qb": batch.read_quals.astype(np.float32) + rng.uniform(0, 0.5, n).astype(np.float32),
         "qi": np.full(n, 45.0, np.float32), "qd": np.full(n, 45.0, np.float32), "qc": np.full(n, 10.0, np.float32)}
    want_out, want = _oracle(batch, q)
    out, got, st = _run(batch, q)
    assert st.n_rescued > 0
    assert np.max(np.abs(out - want_out)) <= 1e-4
    assert _per_read_err(batch, got, want) <= 1e-3


def test_low_likelihood_pairs_keep_fp32_adjoint():
    """Pairs with log10 L between the forward's FP32 rescue threshold (L * 2^120 = 1e-28, i.e.
    log10 L ~ -64.1) and -47 keep their adjoint in FP32; pairs below the threshold go to the FP64
    kernel, exactly as in the forward.  Both kinds match the oracle."""
    batch, q = _problem(31, n_regions=3, hap_len=(250, 360), read_len=(150, 220), sub=0.12)
    want_out, want = _oracle(batch, q)
    below = want_out + 120 * np.log10(2.0) < -28
    window = ~below & (want_out < -47)
    assert window.sum() >= 3 and below.sum() >= 3
    out, got, st = _run(batch, q)
    assert st.n_rescued == int(below.sum())
    assert np.max(np.abs(out - want_out)) <= 1e-4
    assert _per_read_err(batch, got, want) <= 1e-3
    assert _rel(got["eb"], want["eb"]) <= 1e-3


def test_autograd_function():
    batch, q = _problem(5, n_regions=2)
    w = np.random.default_rng(1).uniform(0.5, 1.5, batch.n_pairs)
    _, want = _oracle(batch, q, weights=w)
    prob = GradProblem.from_batch(batch)
    tq = {k: torch.from_numpy(v).cuda().requires_grad_(True) for k, v in q.items()}
    with GradEngine() as eng:
        L = pairhmm_log10_likelihood(prob, tq["qb"], tq["qi"], tq["qd"], tq["qc"], eng)
        (L * torch.from_numpy(w).cuda()).sum().backward()
    got = {k: tq[k].grad.double().cpu().numpy() for k in NAMES}
    assert _per_read_err(batch, got, want) <= 1e-3


def test_gradcheck_fp64_kernel():
    """torch.autograd.gradcheck of the HIP FP64 kernel itself (fast mode) on a small problem."""
    batch, q = _problem(21, n_regions=1, hap_len=(15, 40), read_len=(4, 20))
    prob = GradProblem.from_batch(batch)
    tq = tuple(torch.from_numpy(q[k].astype(np.float64)).cuda().requires_grad_(True) for k in NAMES)
    with GradEngine(use_double=True) as eng:
        f = lambda *a: pairhmm_log10_likelihood(prob, *a, eng)  # noqa: E731
        # Phred inputs reach the kernel as float32, so the step is 1e-2 (a 1e-3 step would be
        # resolved to only ~1e-3 relative at Q ~ 30).
        # FP64 atomics accumulate positions in a run-dependent order: allow last-bit differences.
        assert torch.autograd.gradcheck(f, tq, eps=1e-2, atol=1e-5, rtol=1e-2, fast_mode=True, nondet_tol=1e-10)
