"""Test-only restatement of HaplotypeCaller's read-quality preparation (the PCR indel error model and
the minimum-quality caps) in plain Python.

This is checker code: only tests may import it.  The product implementation is the device code in
gatk_amd/csrc/readprep.hip behind pairhmm_apply_pcr_error_model / pairhmm_cap_read_qualities.

Pinned by (tests/test_readprep_oracle.py):
  * every findNumberOfRepetitions assertion of GATKVariantContextUtilsUnitTest.java:923-962;
  * the reference's own PCR test (PairHMMLikelihoodCalculationEngineUnitTest.java:101-138): reads made
    of one repeat unit, all insertion/deletion qualities 40, must come out as
    getErrorModelAdjustedQual(findTandemRepeatUnits(read, i - 1)) at every position;
  * closed forms: a homopolymer of length L has repeat length min(L, 20) at every position; the
    source's own worked example TTCTT(C)CCC gives (C)4 (ReadLikelihoodCalculationEngine.java:236-240);
  * getErrorModelAdjustedQual's formula at every repeat length and rate factor, and Java's signed
    byte comparison in setToFixedValueIfTooLow (a quality byte >= 128 is negative in Java).
"""
from __future__ import annotations

import math
from typing import List, Sequence

MAX_STR_UNIT_LENGTH = 8     # ReadLikelihoodCalculationEngine.java:25
MAX_REPEAT_LENGTH = 20      # ReadLikelihoodCalculationEngine.java:26
MIN_ADJUSTED_QSCORE = 10    # PairHMMLikelihoodCalculationEngine.java:32
INITIAL_QSCORE = 40.0       # PairHMMLikelihoodCalculationEngine.java:35
MIN_USABLE_Q_SCORE = 6      # QualityUtils.java:37
PCR_RATE_FACTOR = {"NONE": 0.0, "HOSTILE": 1.0, "AGGRESSIVE": 2.0, "CONSERVATIVE": 3.0}  # :53-61


def find_number_of_repetitions(unit: bytes, u_off: int, u_len: int, test: bytes, t_off: int, t_len: int,
                               leading: bool) -> int:
    """GATKVariantContextUtils.findNumberOfRepetitions (GATKVariantContextUtils.java:976-1010)."""
    if t_len == 0:
        return 0
    diff = t_len - u_len
    want = unit[u_off:u_off + u_len]
    n = 0
    if leading:
        start = 0
        while start <= diff:
            if test[t_off + start:t_off + start + u_len] != want:
                return n
            n += 1
            start += u_len
        return n
    start = diff
    while start >= 0:
        if test[t_off + start:t_off + start + u_len] != want:
            return n
        n += 1
        start -= u_len
    return n


def find_number_of_repetitions_simple(unit: bytes, test: bytes, leading: bool) -> int:
    """The two-array overload (GATKVariantContextUtils.java:949-957)."""
    if len(unit) == 0:
        raise ValueError("empty repeatUnit")
    if len(test) == 0:
        return 0
    return find_number_of_repetitions(unit, 0, len(unit), test, 0, len(test), leading)


def find_tandem_repeat_units(read: bytes, offset: int) -> int:
    """ReadLikelihoodCalculationEngine.findTandemRepeatUnits (:193-253); returns the repeat length."""
    max_bw = 0
    best_bw = read[offset:offset + 1]
    for s in range(1, MAX_STR_UNIT_LENGTH + 1):
        if offset + 1 - s < 0:
            break
        max_bw = find_number_of_repetitions(read, offset - s + 1, s, read, 0, offset + 1, False)
        if max_bw > 1:
            best_bw = read[offset - s + 1:offset + 1]
            break
    max_rl = max_bw
    if offset < len(read) - 1:
        best_fw = read[offset + 1:offset + 2]
        max_fw = 0
        for s in range(1, MAX_STR_UNIT_LENGTH + 1):
            if offset + s + 1 > len(read):
                break
            max_fw = find_number_of_repetitions(read, offset + 1, s, read, offset + 1, len(read) - offset - 1, True)
            if max_fw > 1:
                best_fw = read[offset + 1:offset + s + 1]
                break
        if best_fw == best_bw:
            max_rl = max_bw + max_fw
        else:
            max_bw = find_number_of_repetitions(best_fw, 0, len(best_fw), read, 0, offset + 1, False)
            max_rl = max_fw + max_bw
    return min(max_rl, MAX_REPEAT_LENGTH)


def java_fast_round(d: float) -> int:
    """MathUtils.fastRound (MathUtils.java:428-430): (int) (d +/- 0.5), truncating toward zero."""
    return int(d + 0.5) if d > 0.0 else int(d - 0.5)


def error_model_adjusted_qual(repeat_length: int, rate_factor: float) -> int:
    """PairHMMLikelihoodCalculationEngine.getErrorModelAdjustedQual (:356-358)."""
    return max(MIN_ADJUSTED_QSCORE,
               java_fast_round(INITIAL_QSCORE - math.exp(repeat_length / (rate_factor * math.pi)) + 1.0)) & 0xFF


def pcr_cache(rate_factor: float) -> List[int]:
    """initializePCRErrorModel's cache (:339-353)."""
    return [error_model_adjusted_qual(i, rate_factor) for i in range(MAX_REPEAT_LENGTH + 1)]


def apply_pcr_error_model(read: bytes, ins: Sequence[int], dels: Sequence[int], model: str):
    """applyPCRErrorModel (:360-371): positions 0 .. L-2 are capped by the repeat length at them."""
    ins, dels = [int(x) & 0xFF for x in ins], [int(x) & 0xFF for x in dels]
    if model == "NONE":
        return ins, dels
    cache = pcr_cache(PCR_RATE_FACTOR[model])
    for i in range(1, len(read)):
        rl = find_tandem_repeat_units(read, i - 1)
        ins[i - 1] = min(ins[i - 1] & 0xFF, cache[rl])
        dels[i - 1] = min(dels[i - 1] & 0xFF, cache[rl])
    return ins, dels


def _signed(b: int) -> int:
    b = int(b) & 0xFF
    return b - 256 if b >= 128 else b


def cap_minimum_read_qualities(quals, ins, dels, mapping_quality: int, base_q_threshold: int,
                               disable_cap_to_mapq: bool):
    """capMinimumReadQualities + setToFixedValueIfTooLow (:307-320), with Java's signed byte compare."""
    quals, ins, dels = ([int(x) & 0xFF for x in a] for a in (quals, ins, dels))
    for i in range(len(quals)):
        if not disable_cap_to_mapq:
            quals[i] = min(quals[i] & 0xFF, mapping_quality) & 0xFF
        quals[i] = MIN_USABLE_Q_SCORE if _signed(quals[i]) < _signed(base_q_threshold) else quals[i] & 0xFF
        ins[i] = MIN_USABLE_Q_SCORE if _signed(ins[i]) < MIN_USABLE_Q_SCORE else ins[i] & 0xFF
        dels[i] = MIN_USABLE_Q_SCORE if _signed(dels[i]) < MIN_USABLE_Q_SCORE else dels[i] & 0xFF
    return quals, ins, dels
