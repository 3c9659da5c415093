"""Pins the Smith-Waterman oracle (oracle/sw_oracle.c) with every case of the reference's
SmithWatermanAlignerAbstractUnitTest.java, run the way the reference's aligners run them (the exact
substring shortcut first, SmithWatermanJavaAligner.java:61-67), and with an invariant: the score at
the chosen end point equals the score of the CIGAR the traceback emits."""
import numpy as np
import pytest

from gatk_amd.smithwaterman import (ALIGNMENT_TO_BEST_HAPLOTYPE_SW_PARAMETERS, NEW_SW_PARAMETERS, ORIGINAL_DEFAULT,
                                    STANDARD_NGS, SWParameters, last_index_of)
from oracle import oracle


def java_align(ref: bytes, alt: bytes, p: SWParameters, strategy: str):
    """SmithWatermanJavaAligner.align: the substring shortcut for SOFTCLIP / IGNORE, else the full
    alignment of the oracle."""
    if strategy in ("SOFTCLIP", "IGNORE"):
        k = last_index_of(ref, alt)
        if k != -1:
            return f"{len(alt)}M", k
    return oracle.sw_align(ref, alt, (p.match_value, p.mismatch_penalty, p.gap_open_penalty, p.gap_extend_penalty),
                           strategy)


LONG_REF = ("ATAGAAAATAGTTTTTGGAAATATGGGTGAAGAGACATCTCCTCTTATGGAAAAAGGGATTCTAGAATTTAACAATAAATATTCCCAACTTTCCCC"
            "AAGGCTTTAAAATCTACCTTGAAGGAGCAGCTGATGTATTTCTAGAACAGACTTAGGTGTCTTGGTGTGGCCTGTAAAGAGATACTGTCTTTCTCTTTT"
            "GAGTGTAAGAGAGAAAGGACAGTCTACTCAATAAAGAGTGCTGGGAAAACTGAATATCCACACACAGAATAATAAAACTAGATCCTATCTCTCACCAT"
            "ATACAAAGATCAACTCAAAACAAATTAAAGACCTAAATGTAAGACAAGAAATTATAAAACTACTAGAAAAAAACACAAGGGAAATGCTTCAGGACATT"
            "GGC")

# (reference, read, parameters, strategy, expected offset, expected CIGAR) --
# SmithWatermanAlignerAbstractUnitTest.java:130-282
CASES = [
    ("AAAGGACTGACTG", "ACTGACTGACTG", ORIGINAL_DEFAULT, "SOFTCLIP", 1, "12M"),                       # :130-139
    ("AAAGACTACTG", "AACGGACACTG", SWParameters(50, -100, -220, -12), "SOFTCLIP", 1, "2M2I3M1D4M"),   # :141-157
    ("AAAGACTACTG", "AACGGACACTG", SWParameters(200, -50, -300, -22), "SOFTCLIP", 0, "11M"),
    ("AAACCCCC", "CCCCCGGG", ORIGINAL_DEFAULT, "SOFTCLIP", 3, "5M3S"),                                # :159-168
    ("TGTGTGTGTGTGTGACAGAGAGAGAGAGAGAGAGAGAGAGAGAGA",
     "ACAGAGAGAGAGAGAGAGAGAGAGAGAGAGAGAGAGAGAGAGAGAGAGAGA", STANDARD_NGS, "SOFTCLIP", 14, "31M20S"),  # :170-179
    ("AAACCCCC", "CCCCC", ORIGINAL_DEFAULT, "SOFTCLIP", 3, "5M"),                                     # :239-256
    ("AAACCCCC", "CCCCC", ORIGINAL_DEFAULT, "INDEL", 0, "3D5M"),
    ("AAACCCCC", "CCCCC", ORIGINAL_DEFAULT, "LEADING_INDEL", 0, "3D5M"),
    ("AAACCCCC", "CCCCC", ORIGINAL_DEFAULT, "IGNORE", 3, "5M"),
    (LONG_REF, "AAAAAAA", ORIGINAL_DEFAULT, "SOFTCLIP", 359, "7M"),                                   # :267-281
    (LONG_REF, "AAAAAAA", ORIGINAL_DEFAULT, "INDEL", 0, "1M358D6M29D"),
    (LONG_REF, "AAAAAAA", ORIGINAL_DEFAULT, "LEADING_INDEL", 0, "1M1D6M"),
    (LONG_REF, "AAAAAAA", ORIGINAL_DEFAULT, "IGNORE", 359, "7M"),
]


@pytest.mark.parametrize("ref,read,params,strategy,offset,cigar", CASES)
def test_reference_cases(ref, read, params, strategy, offset, cigar):
    assert java_align(ref.encode(), read.encode(), params, strategy) == (cigar, offset)


def test_identical_alignments_with_differing_flank_lengths():
    """testForIdenticalAlignmentsWithDifferingFlankLengths (:181-236): the indels sit in the same
    places whatever the flank length."""
    padded_ref = ("GCGTCGCAGTCTTAAGGCCCCGCCTTTTCAGACAGCTTCCGCTGGGCCTGGGCCGCTGCGGGGCGGTCACGGCCCCTTTAAGCCTGAGCCCCGCCC"
                  "CCTGGCTCCCCGCCCCCTCTTCTCCCCTCCCCCAAGCCAGCACCTGGTGCCCCGGCGGGTCGTGCGGCGCGGCGCTCCGCGGTGAGCGCCTGACC"
                  "CCGAGGGGGCCCGGGGCCGCGTCCCTGGGCCCTCCCCACCCTTGCGGTGGCCTCGCGGGTCCCAGGGGCGGGGCTGGAGCGGCAGCAGGGCCGGG"
                  "GAGATGGGCGGTGGGGAGCGCGGGAGGGACCGGGCCGAGCCGGGGGAAGGGCTCCGGTGACT")
    padded_hap = ("GCGTCGCAGTCTTAAGGCCCCGCCTTTTCAGACAGCTTCCGCTGGGCCTGGGCCGCTGCGGGGCGGTCACGGCCCCTTTAAGCCTGAGCCCCGCCC"
                  "CCTGGCTCCCCGCCCCCTCTTCTCCCCTCCCCCAAGCCAGCACCTGGTGCCCCGGCGGGTCGTGCGGCGCGGCGCTCCGCGGTGAGCGCCTGACC"
                  "CCGA--GGGCC---------------GGGCCCTCCCCACCCTTGCGGTGGCCTCGCGGGTCCCAGGGGCGGGGCTGGAGCGGCAGCAGGGCCGGG"
                  "GAGATGGGCGGTGGGGAGCGCGGGAGGGACCGGGCCGAGCCGGGGGAAGGGCTCCGGTGACT").replace("-", "")
    short_ref = ("CTTTAAGCCTGAGCCCCGCCCCCTGGCTCCCCGCCCCCTCTTCTCCCCTCCCCCAAGCCAGCACCTGGTGCCCCGGCGGGTCGTGCGGCGCGGCGCT"
                 "CCGCGGTGAGCGCCTGACCCCGAGGGGGCCCGGGGCCGCGTCCCTGGGCCCTCCCCACCCTTGCGGTGGCCTCGCGGGTCCCAGGGGCGGGGCTGG"
                 "AGCGGCAGCAGGGCCGGGGAGATGGGCGGTGGGGAGCGCGGGAGGGA")
    short_hap = ("CTTTAAGCCTGAGCCCCGCCCCCTGGCTCCCCGCCCCCTCTTCTCCCCTCCCCCAAGCCAGCACCTGGTGCCCCGGCGGGTCGTGCGGCGCGGCGCT"
                 "CCGCGGTGAGCGCCTGACCCCGA---------GGGCC--------GGGCCCTCCCCACCCTTGCGGTGGCCTCGCGGGTCCCAGGGGCGGGGCTGG"
                 "AGCGGCAGCAGGGCCGGGGAGATGGGCGGTGGGGAGCGCGGGAGGGA").replace("-", "")
    pad = "NNNNNNNNNN"
    a = java_align((pad + padded_ref + pad).encode(), (pad + padded_hap + pad).encode(), NEW_SW_PARAMETERS, "SOFTCLIP")
    b = java_align((pad + short_ref + pad).encode(), (pad + short_hap + pad).encode(), NEW_SW_PARAMETERS, "SOFTCLIP")

    def elements(c):
        out, num = [], ""
        for ch in c[0]:
            if ch.isdigit():
                num += ch
            else:
                out.append((int(num), ch))
                num = ""
        return out

    ea, eb = elements(a), elements(b)
    assert len(ea) == len(eb)
    for (la, oa), (lb, ob) in zip(ea, eb):
        if oa == "M" and ob == "M":
            continue
        assert (la, oa) == (lb, ob)


def _cigar_score(ref, alt, cigar, offset, p: SWParameters):
    """Score of an alignment read off its CIGAR: matches / mismatches, and open + (k-1) extend per gap."""
    i, j, score = offset, 0, 0
    num = ""
    for ch in cigar:
        if ch.isdigit():
            num += ch
            continue
        k = int(num)
        num = ""
        if ch == "S":
            j += k
        elif ch == "M":
            for _ in range(k):
                score += p.match_value if ref[i] == alt[j] else p.mismatch_penalty
                i += 1
                j += 1
        elif ch == "I":
            score += p.gap_open_penalty + (k - 1) * p.gap_extend_penalty
            j += k
        elif ch == "D":
            score += p.gap_open_penalty + (k - 1) * p.gap_extend_penalty
            i += k
    return score, i, j


@pytest.mark.parametrize("seed", range(20))
@pytest.mark.parametrize("params", [ORIGINAL_DEFAULT, STANDARD_NGS, NEW_SW_PARAMETERS,
                                    ALIGNMENT_TO_BEST_HAPLOTYPE_SW_PARAMETERS])
def test_softclip_score_equals_matrix_end_point(seed, params):
    """SOFTCLIP: the CIGAR's own score equals the matrix score at the end point the traceback started
    from (row offset + consumed reference bases, column = read bases minus the trailing clip)."""
    rng = np.random.default_rng(seed)
    hap = bytes(np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, int(rng.integers(40, 200)))])
    s = int(rng.integers(0, len(hap) - 30))
    read = bytearray(hap[s:s + int(rng.integers(20, 30 + len(hap) // 4))])
    for _ in range(int(rng.integers(0, 4))):  # substitutions and indels
        k = int(rng.integers(0, len(read)))
        r = rng.random()
        if r < 0.4:
            read[k] = b"ACGT"[int(rng.integers(0, 4))]
        elif r < 0.7:
            del read[k:k + int(rng.integers(1, 4))]
        else:
            read[k:k] = bytes(np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, int(rng.integers(1, 4)))])
    read = bytes(read) or b"A"
    pp = (params.match_value, params.mismatch_penalty, params.gap_open_penalty, params.gap_extend_penalty)
    cigar, off, sc, _ = oracle.sw_align(hap, read, pp, "SOFTCLIP", matrices=True)
    score, i_end, j_end = _cigar_score(hap, read, cigar, off, params)
    trailing = int(cigar[:-1].rsplit("M", 1)[-1].rsplit("I", 1)[-1].rsplit("D", 1)[-1]) if cigar.endswith("S") else 0
    assert j_end == len(read)
    assert sc[i_end, len(read) - trailing] == score
