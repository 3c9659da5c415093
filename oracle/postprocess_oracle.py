"""Test-only restatement of the likelihood post-processing that follows the PairHMM in
PairHMMLikelihoodCalculationEngine.computeReadLikelihoods (PairHMMLikelihoodCalculationEngine.java:
200-201): the mismapping-rate cap (AlleleLikelihoods.normalizeLikelihoods) and the poorly-modelled
read filter (ReadLikelihoodCalculationEngine.filterPoorlyModeledEvidence), in plain Python.

This is checker code: only tests may import it.  The product implementation is
gatk_amd/csrc/postprocess.hip behind pairhmm_normalize_and_filter_likelihoods.

Pinned by (tests/test_postprocess_oracle.py):
  * AlleleLikelihoodsUnitTest.testNormalizeCapWorstLK (:357-387): with symmetric normalisation every
    value becomes max(value, best - 0.001), rows whose best is -inf unchanged;
  * AlleleLikelihoodsUnitTest.testFilterPoorlyModeledReads (:192-216): a fixed -100 threshold removes
    exactly the reads whose best likelihood is below it;
  * closed forms of log10MinTrueLikelihood and of the dynamic threshold for constant-quality reads
    (L * mean + K * sqrt(L * var), times -0.1), and the reference-exclusion rule of searchBestAllele.
"""
from __future__ import annotations

import math
from typing import Optional, Sequence

import numpy as np

# ReadLikelihoodCalculationEngine.dynamicReadQualThreshLookupTable (:161-176): (mean, variance) for
# base qualities 1..40.  Values are the reference's table.
DYNAMIC_TABLE = (
    (5.996842844, 0.196616587), (5.870018422, 1.388545569), (5.401558531, 5.641990128),
    (4.818940919, 10.33176216), (4.218758304, 14.25799688), (3.646319832, 17.02880749),
    (3.122346753, 18.64537883), (2.654731979, 19.27521677), (2.244479156, 19.13584613),
    (1.88893867, 18.43922003), (1.583645342, 17.36842261), (1.3233807, 16.07088712),
    (1.102785365, 14.65952563), (0.916703025, 13.21718577), (0.760361881, 11.80207947),
    (0.629457387, 10.45304833), (0.520175654, 9.194183767), (0.42918208, 8.038657241),
    (0.353590663, 6.991779595), (0.290923699, 6.053379213), (0.23906788, 5.219610436),
    (0.196230431, 4.484302033), (0.160897421, 3.839943445), (0.131795374, 3.27839108),
    (0.1078567, 2.791361596), (0.088189063, 2.370765375), (0.072048567, 2.008921719),
    (0.058816518, 1.698687797), (0.047979438, 1.433525748), (0.039111985, 1.207526336),
    (0.031862437, 1.015402928), (0.025940415, 0.852465956), (0.021106532, 0.714585285),
    (0.017163711, 0.598145851), (0.013949904, 0.500000349), (0.011332027, 0.41742159),
    (0.009200898, 0.348056286), (0.007467036, 0.289881373), (0.006057179, 0.241163527),
    (0.004911394, 0.200422214),
)
MAXIMUM_DYNAMIC_QUAL_THRESHOLD_ENTRY_BASEQ = 40   # :178
LOG10_QUAL_PER_BASE = -4.0                        # :125


def best_likelihood(row: Sequence[float], ref_index: int, can_be_reference: bool) -> float:
    """AlleleLikelihoods.searchBestAllele (AlleleLikelihoods.java:505-533) without priorities: the
    best likelihood, the reference allele skipped unless can_be_reference."""
    n = len(row)
    if n == 0 or (n == 1 and ref_index == 0 and not can_be_reference):
        return -math.inf
    best_i = 0 if (can_be_reference or ref_index != 0) else 1
    best = float(row[best_i])
    for a in range(best_i + 1, n):
        if not can_be_reference and ref_index == a:
            continue
        if float(row[a]) > best:
            best = float(row[a])
    return best


def normalize_likelihoods(lk: np.ndarray, cap: float, symmetric: bool, ref_index: int) -> np.ndarray:
    """normalizeLikelihoods (:416-458) on a read-major [reads, alleles] matrix (a copy)."""
    if not (cap < 0.0) or math.isnan(cap):
        raise ValueError("the minimum reference likelihood fall must be negative")
    out = np.array(lk, dtype=np.float64, copy=True)
    if cap == -math.inf or out.shape[1] <= 1:
        return out
    for r in range(out.shape[0]):
        worst = best_likelihood(out[r], ref_index, symmetric) + cap
        for a in range(out.shape[1]):
            if out[r, a] < worst:
                out[r, a] = worst
    return out


def log10_min_true_likelihood(read_length: int, max_error_per_base: float, cap_likelihoods: bool) -> float:
    """log10MinTrueLikelihood (ReadLikelihoodCalculationEngine.java:118-127)."""
    n = math.ceil(read_length * max_error_per_base)
    errors = min(2.0, float(n)) if cap_likelihoods else float(n)
    return errors * LOG10_QUAL_PER_BASE


def dynamic_read_qual_threshold(quals: Sequence[int], scale: float) -> float:
    """calculateLog10DynamicReadQualThreshold (:137-155): sums in read order, then * -0.1."""
    s_mean = 0.0
    s_var = 0.0
    for q in quals:
        bq = int(q) & 0xFF
        e = 0 if bq <= 1 else min(MAXIMUM_DYNAMIC_QUAL_THRESHOLD_ENTRY_BASEQ, bq) - 1
        s_mean += DYNAMIC_TABLE[e][0]
        s_var += DYNAMIC_TABLE[e][1]
    return (s_mean + scale * math.sqrt(s_var)) * -0.1


def read_threshold(quals: Sequence[int], dynamic: bool, max_error_per_base: float, scale: float) -> float:
    """filterPoorlyModeledEvidence's per-read threshold (:98-116)."""
    if dynamic:
        t = log10_min_true_likelihood(len(quals), max_error_per_base, False)
        d = dynamic_read_qual_threshold(quals, scale)
        return d if d < t else t
    return log10_min_true_likelihood(len(quals), max_error_per_base, True)


def keep_reads(lk: np.ndarray, thresholds: Sequence[float]) -> np.ndarray:
    """AlleleLikelihoods.filterPoorlyModeledEvidence (:1351-1375): a read stays unless its best
    likelihood over all alleles (maximumLikelihoodOverAllAlleles, :1198-1208) is below its threshold."""
    keep = np.ones(lk.shape[0], dtype=bool)
    for r in range(lk.shape[0]):
        m = -math.inf
        for a in range(lk.shape[1]):
            if lk[r, a] > m:
                m = lk[r, a]
        keep[r] = not (m < thresholds[r])
    return keep


def postprocess(lk: np.ndarray, read_quals: Sequence[Sequence[int]], cap: float, symmetric: bool, ref_index: int,
                dynamic: bool, max_error_per_base: float, scale: float):
    """Both steps in the engine's order: (normalized matrix, keep flags, thresholds)."""
    norm = normalize_likelihoods(lk, cap, symmetric, ref_index)
    th = [read_threshold(q, dynamic, max_error_per_base, scale) for q in read_quals]
    return norm, keep_reads(norm, th), np.asarray(th)
