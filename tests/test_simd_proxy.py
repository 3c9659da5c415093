"""The GKL-class CPU proxy (oracle/simd_proxy.c), timed by the bench beside the FP64 port, agrees with
the FP64 oracle to the engine's own FP32 tolerance, and is independent of the thread count."""
import numpy as np

from gatk_amd import synth
from oracle import oracle


def test_proxy_matches_fp64_oracle():
    b = synth.workload("batched_uneven", n_regions=3, seed=5)
    ref = oracle.logless_batch(b, n_threads=4)
    got, _ = oracle.simd_proxy_batch(b, n_threads=4)
    fin = np.isfinite(ref)
    assert np.array_equal(fin, np.isfinite(got))
    assert np.max(np.abs(got[fin] - ref[fin])) <= 1e-4
    one, _ = oracle.simd_proxy_batch(b, n_threads=1)
    assert np.array_equal(one, got)


def test_proxy_rescues_underflowing_pairs():
    b = synth.workload("precision_stress", n_regions=1, seed=2)
    sel = np.arange(min(b.n_pairs, 400))
    ref = oracle.logless_batch(b, n_threads=4, pairs=sel)
    got, n_rescued = oracle.simd_proxy_batch(b, n_threads=4, pairs=sel)
    assert n_rescued > 0
    fin = np.isfinite(ref)
    assert np.max(np.abs(got[fin] - ref[fin])) <= 1e-4
