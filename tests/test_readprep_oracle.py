"""The read-preparation oracle (PCR indel error model, minimum-quality caps) against the reference's
own assertions and closed forms.  CPU only."""
import math

import numpy as np
import pytest

from oracle import readprep_oracle as rp

# GATKVariantContextUtilsUnitTest.java:923-944 (two-array overload)
FNR_SIMPLE = [
    ("AT", "GATAT", False, 2), ("AT", "GATAT", True, 0), ("A", "ATATG", True, 1), ("AT", "ATATG", True, 2),
    ("CCC", "CCCCCCCC", True, 2), ("CCC", "CCCCCCCC", False, 2), ("ATG", "ATGATGATGATG", True, 4),
    ("G", "ATGATGATGATG", True, 0), ("T", "T", True, 1), ("AT", "ATGATGATCATG", True, 1),
    ("CCCCCCCC", "CCC", True, 0), ("AT", "AT", True, 1), ("AT", "", True, 0), ("ATG", "ATGATGATGATG", False, 4),
    ("G", "ATGATGATGATG", False, 1), ("T", "T", False, 1), ("AT", "ATGATGATCATG", False, 0),
    ("CCCCCCCC", "CCC", False, 0), ("AT", "AT", False, 1), ("AT", "", False, 0),
]
# GATKVariantContextUtilsUnitTest.java:949-962 (sub-array overload)
FNR_FULL = [
    ("XXXATG", 3, 3, "ATGATGATGATGYYY", 0, 12, True, 4), ("GGGG", 0, 1, "GGGGATGATGATGATG", 4, 12, True, 0),
    ("T", 0, 1, "TTTTT", 0, 1, True, 1), ("AT", 0, 2, "AT", 0, 0, True, 0), ("AT", 0, 2, "AT", 1, 0, True, 0),
    ("AT", 0, 2, "", 0, 0, True, 0), ("XXXAT", 3, 2, "XXXGATAT", 4, 4, False, 2), ("AT", 0, 2, "GATAT", 0, 5, False, 2),
    ("ATG", 0, 3, "ATGATGATGATG", 0, 12, False, 4), ("ATG", 0, 3, "ATGATGATGATGATG", 3, 12, False, 4),
    ("G", 0, 1, "ATGATGATGATG", 0, 12, False, 1), ("G", 0, 1, "ATGATGATGATGATG", 0, 12, False, 1),
]


@pytest.mark.parametrize("unit,test,leading,want", FNR_SIMPLE)
def test_find_number_of_repetitions(unit, test, leading, want):
    assert rp.find_number_of_repetitions_simple(unit.encode(), test.encode(), leading) == want


@pytest.mark.parametrize("unit,uo,ul,test,to,tl,leading,want", FNR_FULL)
def test_find_number_of_repetitions_full(unit, uo, ul, test, to, tl, leading, want):
    assert rp.find_number_of_repetitions(unit.encode(), uo, ul, test.encode(), to, tl, leading) == want


def test_tandem_repeat_closed_forms():
    for L in (1, 2, 5, 19, 20, 21, 40):
        read = b"A" * L
        assert [rp.find_tandem_repeat_units(read, i) for i in range(L)] == [min(L, 20)] * L
    # ReadLikelihoodCalculationEngine.java:236-240: TTCTT(C)CCC -> (C)4 at the bracketed C
    assert rp.find_tandem_repeat_units(b"TTCTTCCCC", 5) == 4


def test_error_model_adjusted_qual_formula():
    for rate in (1.0, 2.0, 3.0):
        for rl in range(21):
            x = 40.0 - math.exp(rl / (rate * math.pi)) + 1.0
            want = max(10, int(x + 0.5) if x > 0 else int(x - 0.5))
            assert rp.error_model_adjusted_qual(rl, rate) == want
    assert rp.error_model_adjusted_qual(0, 3.0) == 40  # exp(0) = 1
    assert rp.error_model_adjusted_qual(20, 1.0) == 10  # floor


@pytest.mark.parametrize("unit", ["A", "AC", "ACG", "ACGT"])
@pytest.mark.parametrize("n", [1, 2, 3, 5, 10, 15])
def test_reference_pcr_case(unit, n):
    """PairHMMLikelihoodCalculationEngineUnitTest.java:101-138 (CONSERVATIVE, all qualities 40)."""
    read = (unit * n).encode()
    ins, dels = rp.apply_pcr_error_model(read, [40] * len(read), [40] * len(read), "CONSERVATIVE")
    for i in range(1, len(read)):
        want = rp.error_model_adjusted_qual(rp.find_tandem_repeat_units(read, i - 1), 3.0)
        assert ins[i - 1] == want and dels[i - 1] == want
    if len(read) > 0:
        assert ins[-1] == 40 and dels[-1] == 40  # the last position is never adjusted


def test_pcr_none_is_identity():
    read = b"ACACACAGGGG"
    q = list(range(20, 20 + len(read)))
    assert rp.apply_pcr_error_model(read, q, q, "NONE") == (q, q)


def test_cap_minimum_qualities_java_bytes():
    quals = [2, 5, 6, 30, 60, 200]
    ins = [0, 5, 6, 45, 128, 255]
    dels = [1, 6, 7, 45, 127, 130]
    q, i, d = rp.cap_minimum_read_qualities(quals, ins, dels, mapping_quality=50, base_q_threshold=18,
                                            disable_cap_to_mapq=False)
    assert q == [6, 6, 6, 30, 50, 50]
    # bytes >= 128 are negative in Java, so they fall below MIN_USABLE_Q_SCORE and become 6
    assert i == [6, 6, 6, 45, 6, 6]
    assert d == [6, 6, 7, 45, 127, 6]
    q2, _, _ = rp.cap_minimum_read_qualities(quals, ins, dels, mapping_quality=50, base_q_threshold=18,
                                             disable_cap_to_mapq=True)
    assert q2 == [6, 6, 6, 30, 60, 6]  # 200 is -56 as a Java byte
