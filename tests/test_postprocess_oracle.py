"""Pins the likelihood post-processing oracle (oracle/postprocess_oracle.py) with the reference's own
tests and closed forms."""
import math

import numpy as np
import pytest

from oracle import postprocess_oracle as pp


def _random_lk(rng, n_reads, n_alleles, p_inf=0.1):
    lk = -rng.exponential(20.0, (n_reads, n_alleles))
    lk[rng.random(lk.shape) < p_inf] = -np.inf
    return lk


@pytest.mark.parametrize("seed", range(5))
def test_normalize_cap_worst_lk_symmetric(seed):
    """AlleleLikelihoodsUnitTest.testNormalizeCapWorstLK (:357-387)."""
    rng = np.random.default_rng(seed)
    lk = _random_lk(rng, 50, int(rng.integers(2, 6)))
    lk[3] = -np.inf
    got = pp.normalize_likelihoods(lk, -0.001, True, 0)
    for r in range(lk.shape[0]):
        best = lk[r].max()
        want = lk[r] if best == -np.inf else np.maximum(lk[r], best - 0.001)
        assert np.array_equal(got[r], want)


@pytest.mark.parametrize("ref_index", [-1, 0, 1, 2])
def test_normalize_excludes_the_reference_when_asymmetric(ref_index):
    """searchBestAllele with canBeReference false skips the reference allele (:512-523)."""
    rng = np.random.default_rng(10 + ref_index)
    lk = _random_lk(rng, 40, 4, p_inf=0.0)
    cap = -4.5
    got = pp.normalize_likelihoods(lk, cap, False, ref_index)
    for r in range(lk.shape[0]):
        alts = [a for a in range(4) if a != ref_index]
        best = lk[r, alts].max()
        assert np.array_equal(got[r], np.maximum(lk[r], best + cap))


def test_normalize_no_op_cases():
    lk = np.array([[-3.0, -50.0], [-np.inf, -np.inf]])
    assert np.array_equal(pp.normalize_likelihoods(lk, -np.inf, True, 0), lk)
    one = np.array([[-7.0], [-100.0]])
    assert np.array_equal(pp.normalize_likelihoods(one, -1.0, True, 0), one)
    with pytest.raises(ValueError):
        pp.normalize_likelihoods(lk, 0.0, True, 0)
    with pytest.raises(ValueError):
        pp.normalize_likelihoods(lk, float("nan"), True, 0)


def test_filter_poorly_modeled_reads_fixed_threshold():
    """AlleleLikelihoodsUnitTest.testFilterPoorlyModeledReads (:192-216): even reads are bad."""
    rng = np.random.default_rng(3)
    n, A = 21, 3
    lk = -rng.uniform(0.0, 50.0, (n, A))
    lk[0::2] = -rng.uniform(101.0, 300.0, (len(lk[0::2]), A))
    keep = pp.keep_reads(lk, [-100.0] * n)
    assert keep.sum() == n // 2
    assert np.array_equal(np.nonzero(keep)[0], np.arange(1, n, 2))


@pytest.mark.parametrize("L,rate,cap,want", [(100, 0.02, True, -8.0), (100, 0.02, False, -8.0),
                                             (250, 0.02, True, -8.0), (250, 0.02, False, -20.0),
                                             (10, 0.02, True, -4.0), (0, 0.02, False, 0.0),
                                             (151, 0.02, False, -16.0)])
def test_log10_min_true_likelihood(L, rate, cap, want):
    """maxErrors = ceil(L * rate), capped at 2 when asked; times log10 Q40 per base (-4)."""
    assert pp.log10_min_true_likelihood(L, rate, cap) == want


@pytest.mark.parametrize("q", [0, 1, 2, 17, 30, 40, 41, 60, 255])
def test_dynamic_threshold_constant_quality(q):
    """A read of L bases of quality q: -0.1 (L mean_q + K sqrt(L var_q)), qualities <= 1 using the
    first entry and > 40 the last."""
    L, K = 150, 1.0
    e = 0 if q <= 1 else min(40, q) - 1
    mean, var = pp.DYNAMIC_TABLE[e]
    got = pp.dynamic_read_qual_threshold([q] * L, K)
    assert got == pytest.approx(-0.1 * (L * mean + K * math.sqrt(L * var)), rel=1e-12)


def test_dynamic_threshold_takes_the_lower_value():
    """daynamicLog10MinLiklihoodModel (:101-113): min of the dynamic threshold and the uncapped one."""
    low_q = [5] * 150    # dynamic threshold far below -ceil(3) * 4
    high_q = [40] * 150  # dynamic threshold near 0
    t_low = pp.read_threshold(low_q, True, 0.02, 1.0)
    t_high = pp.read_threshold(high_q, True, 0.02, 1.0)
    assert t_low == pp.dynamic_read_qual_threshold(low_q, 1.0) < -12.0
    assert t_high == pp.log10_min_true_likelihood(150, 0.02, False) == -12.0
    assert pp.read_threshold(high_q, False, 0.02, 1.0) == -8.0


def test_table_is_monotone_in_mean():
    means = [m for m, _ in pp.DYNAMIC_TABLE]
    assert all(a > b for a, b in zip(means, means[1:]))
    assert len(pp.DYNAMIC_TABLE) == 40
