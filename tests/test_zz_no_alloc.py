"""pairhmm_forward_backward_noalloc (fixed device scratch, no allocation in the call) against
pairhmm_forward_backward on the same inputs: two calls in a row, the second smaller, in FP32 (with
the FP64 recompute of low-likelihood pairs) and FP64.  log10 L must be identical (each pair is
computed by one wave with the same code); per-read gradient sums may differ in the last bits only
(FP64 atomics accumulate pairs in a schedule-dependent order), so they are held to 1e-12 relative.
"""
import ctypes

import numpy as np
import pytest
import torch

from gatk_amd import _native, synth
from gatk_amd.autograd import GradEngine, GradProblem
from gatk_amd.batch import pack_regions

pytestmark = pytest.mark.gpu

NAMES = ("qb", "qi", "qd", "qc", "eb")


def _problem(n_regions, reads_per_region, haps_per_region, seed):
    full = synth.workload("batched_uneven", n_regions=n_regions, seed=seed)
    regions = []
    for k in range(full.n_regions):
        reads, haps = full.region(k)
        regions.append((haps[:haps_per_region], reads[:reads_per_region]))
    batch = pack_regions(regions)
    n = len(batch.read_bases)
    rng = np.random.default_rng(seed)
    q = {"qb": batch.read_quals.astype(np.float32) + rng.uniform(0, 1, n).astype(np.float32),
         "qi": rng.uniform(35, 50, n).astype(np.float32), "qd": rng.uniform(35, 50, n).astype(np.float32),
         "qc": rng.uniform(8, 12, n).astype(np.float32)}
    return batch, q


def _run(batch, q, use_double, no_alloc):
    prob = GradProblem.from_batch(batch)
    tq = {k: torch.from_numpy(v).cuda() for k, v in q.items()}
    with GradEngine(use_double=use_double, no_alloc=no_alloc) as eng:
        out, grads = eng.run(prob, tq["qb"], tq["qi"], tq["qd"], tq["qc"], want_eb=True)
        st = eng.last_stats
    return out.cpu().numpy(), {k: v.cpu().numpy() for k, v in grads.items()}, st


def _per_read_rel(batch, a, b):
    worst = 0.0
    for r in range(len(batch.read_len)):
        o, R = int(batch.read_off[r]), int(batch.read_len[r])
        x = np.concatenate([a[k][o:o + R] for k in NAMES])
        y = np.concatenate([b[k][o:o + R] for k in NAMES])
        worst = max(worst, float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300)))
    return worst


@pytest.mark.parametrize("use_double", [False, True])
def test_noalloc_matches_forward_backward_two_calls(use_double):
    calls = [_problem(6, 6, 10, 41), _problem(2, 3, 4, 42)]  # the second call is smaller
    assert calls[1][0].n_pairs < calls[0][0].n_pairs
    print(f"no_alloc: use_double={use_double}, pairs {[b.n_pairs for b, _ in calls]}", flush=True)
    for batch, q in calls:
        want_out, want_g, want_st = _run(batch, q, use_double, no_alloc=False)
        got_out, got_g, got_st = _run(batch, q, use_double, no_alloc=True)
        assert np.array_equal(got_out, want_out)
        assert _per_read_rel(batch, got_g, want_g) <= 1e-12
        assert got_st.n_rescued == want_st.n_rescued
        assert got_st.cells == want_st.cells


def test_noalloc_rejects_checkpoints_beyond_the_fixed_scratch():
    """A declared read length whose checkpoints for one workgroup exceed the 256 MiB array is a
    clear BAD_ARG, raised before any adjoint kernel runs."""
    batch, q = _problem(1, 2, 2, 43)
    prob = GradProblem.from_batch(batch)
    prob.max_read_len = 200_000
    tq = {k: torch.from_numpy(v).cuda() for k, v in q.items()}
    with GradEngine(use_double=True, no_alloc=True) as eng:
        with pytest.raises(ValueError, match="fixed scratch"):
            eng.run(prob, tq["qb"], tq["qi"], tq["qd"], tq["qc"])
