"""GPU parity of the likelihood post-processing (pairhmm_normalize_and_filter_likelihoods,
postprocess.hip) against the oracle, bit for bit: FP64 with the same operation order."""
import math

import numpy as np
import pytest
import torch

from gatk_amd.postprocess import LikelihoodPostProcessor, log10_mismapping_cap
from oracle import postprocess_oracle as pp

pytestmark = pytest.mark.gpu


def _case(seed, n_reads, n_alleles):
    rng = np.random.default_rng(seed)
    lk = -rng.exponential(15.0, (n_reads, n_alleles))
    lk[rng.random(lk.shape) < 0.05] = -np.inf
    lk[rng.random(n_reads) < 0.02] = -np.inf          # whole rows without a finite value
    lens = rng.integers(0, 300, n_reads).astype(np.int32)
    off = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    quals = rng.integers(0, 64, int(lens.sum())).astype(np.uint8)
    quals[rng.random(len(quals)) < 0.01] = 255
    return lk, lens, off, quals


@pytest.mark.parametrize("seed,n_alleles,ref_index,cap,symmetric,dynamic", [
    (0, 2, 0, -4.5, False, False), (1, 5, 0, -4.5, True, False), (2, 8, 3, -1.0, False, True),
    (3, 1, 0, -4.5, False, True), (4, 16, -1, -math.inf, False, False), (5, 3, 2, -0.001, True, True),
    (6, 128, 0, -4.5, False, True)])
def test_matches_oracle(seed, n_alleles, ref_index, cap, symmetric, dynamic):
    n_reads = 3000
    lk, lens, off, quals = _case(seed, n_reads, n_alleles)
    print(f"[postprocess] seed={seed} A={n_alleles} ref={ref_index} cap={cap} sym={symmetric} dyn={dynamic}",
          flush=True)
    d_lk = torch.from_numpy(lk.copy()).cuda()
    d_len, d_off, d_q = (torch.from_numpy(x).cuda() for x in (lens, off, quals))
    th = torch.empty(n_reads, dtype=torch.float64, device="cuda")
    with LikelihoodPostProcessor() as p:
        keep, _ = p.normalize_and_filter(d_lk, d_len, ref_index, cap, symmetric, d_q, d_off, 0.02, dynamic, 1.0,
                                         threshold=th)
    torch.cuda.synchronize()
    rq = [quals[off[r]:off[r] + lens[r]] for r in range(n_reads)]
    norm, want_keep, want_th = pp.postprocess(lk, rq, cap, symmetric, ref_index, dynamic, 0.02, 1.0)
    assert np.array_equal(d_lk.cpu().numpy(), norm)
    assert np.array_equal(th.cpu().numpy(), want_th)
    assert np.array_equal(keep.cpu().numpy().astype(bool), want_keep)


def test_engine_default_cap_and_argument_errors():
    assert log10_mismapping_cap(45) == 45 * -0.1
    assert log10_mismapping_cap(-1) == -math.inf
    lk = torch.zeros((4, 2), dtype=torch.float64, device="cuda")
    ln = torch.full((4,), 10, dtype=torch.int32, device="cuda")
    with LikelihoodPostProcessor() as p:
        with pytest.raises(ValueError):
            p.normalize_and_filter(lk, ln, 0, 0.0)        # cap must be negative (AlleleLikelihoods.java:417)
        with pytest.raises(ValueError):
            p.normalize_and_filter(lk, ln, 5, -4.5)       # reference index out of range
        keep, _ = p.normalize_and_filter(lk, ln, 0, -4.5)
    assert keep.cpu().tolist() == [1, 1, 1, 1]
