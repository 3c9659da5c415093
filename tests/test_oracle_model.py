"""Pins the oracle's model tables: QualityUtils caches, the Jacobian log-sum table and
PairHMMModel's matchToMatch / qualToTransProbs, against closed forms
(PairHMMModelUnitTest.java:19-109, 186-295) and brute force over every byte quality."""
import math

import numpy as np
import pytest

from oracle import oracle

TOL = 1e-9  # PairHMMModelUnitTest.java:19
TEST_QUALS = [0, 1, 2, 5, 10, 13, 17, 20, 23, 27, 30, 43, 57, 70, 100, 200, 254]  # :237
INS_QUALS = [30, 45, 20, 10, 5, 60, 123]  # :289-293
GAP_QUALS = [10, 20, 5]


def test_qual_to_error_prob_closed_form():
    for q in range(255):
        assert oracle.qual_to_error_prob(q) == pytest.approx(10.0 ** (-q / 10.0), rel=1e-15)
        assert oracle.qual_to_prob(q) == pytest.approx(1.0 - 10.0 ** (-q / 10.0), abs=1e-16)
    assert math.isnan(oracle.qual_to_error_prob(255))  # outside the 255-entry cache


def test_match_to_match_closed_form_grid():
    """PairHMMModelUnitTest.testDoubleQualToProb: max(0, 1 - e1 - e2) at 1e-9, both orders."""
    for i, q1 in enumerate(TEST_QUALS):
        for q2 in TEST_QUALS[i:]:
            e = max(0.0, 1.0 - (10 ** (-0.1 * q1) + 10 ** (-0.1 * q2)))
            for a, b in ((q1, q2), (q2, q1)):
                assert abs(oracle.match_to_match_prob(a, b) - e) <= TOL
                if e > 0:
                    assert abs(oracle.match_to_match_prob_log10(a, b) - math.log10(min(1.0, e))) <= TOL


def test_match_to_match_table_drop_rule_brute_force():
    """Every (ins, del) byte pair: the Jacobian table is exact at integer Phred and drops the
    smaller error term once 0.1*|dq| >= 8.0 (MathUtils.java:409, 479), decided by the double
    arithmetic (-0.1*min) - (-0.1*max) exactly as Java computes it."""
    dropped = 0
    for mx in range(255):
        for mn in range(mx + 1):
            got = oracle.match_to_match_prob(mn, mx)
            diff = (-0.1 * mn) - (-0.1 * mx)
            if diff < 8.0:
                want = 1.0 - (10 ** (-0.1 * mn) + 10 ** (-0.1 * mx))
            else:
                want = 1.0 - 10 ** (-0.1 * mn)
                dropped += 1
            assert abs(got - max(want, 0.0)) <= 1e-13, (mn, mx, got, want)
    assert dropped > 0
    # above MAX_QUAL the value is computed directly (PairHMMModel.java:386)
    assert abs(oracle.match_to_match_prob(30, 255) - (1.0 - 10 ** -3.0)) <= 1e-12


def test_jacobian_table():
    for k in [0, 1, 10, 1234, 40000, 79999, 80000]:
        d = k * 1e-4
        assert abs(oracle.jacobian_table_get(d) - math.log10(1 + 10 ** -d)) <= 1e-12
    assert oracle.approximate_log10_sum_log10(-3.0, -math.inf) == -3.0
    assert oracle.approximate_log10_sum_log10(-math.inf, -3.0) == -3.0
    a, b = -2.3, -4.1
    assert oracle.approximate_log10_sum_log10(a, b) == oracle.approximate_log10_sum_log10(b, a)
    assert abs(oracle.approximate_log10_sum_log10(a, b) - math.log10(10 ** a + 10 ** b)) <= 1e-4


def _qual_to_probs(ins, dele, gap):
    """PairHMMModelUnitTest.qualsToProbs (:186-200)."""
    return np.array([oracle.match_to_match_prob(ins, dele), 1 - 10 ** (-gap / 10), 10 ** (-ins / 10),
                     10 ** (-gap / 10), 10 ** (-dele / 10), 10 ** (-gap / 10)])


def test_qual_to_trans_probs_grid():
    for ins in INS_QUALS:
        for dele in INS_QUALS:
            for gap in GAP_QUALS:
                got = oracle.qual_to_trans_probs(ins, dele, gap)
                assert np.max(np.abs(got - _qual_to_probs(ins, dele, gap))) <= TOL
                # closed form of the match-to-match entry
                assert abs(got[0] - max(0, 1 - 10 ** (-ins / 10) - 10 ** (-dele / 10))) <= TOL


def test_qual_to_trans_probs_rejects_negative_bytes():
    for bad in (128, 200, 255):
        with pytest.raises(oracle.OracleError):
            oracle.qual_to_trans_probs(bad, 30, 10)
        with pytest.raises(oracle.OracleError):
            oracle.qual_to_trans_probs(30, 30, bad)
