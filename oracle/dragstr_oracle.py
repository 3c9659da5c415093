"""DRAGstr STR-context oracle (test infrastructure only; never imported by the product path).

Pure-Python restatement, for small inputs, of the reference's definitions:
  * repeat count of period p at position i -- the brute force of
    DragstrReadSTRAnalyzerTest.calculateExpectedRepeatLength (DragstrReadSTRAnalyzerTest.java:126-171),
    which the reference test asserts equal to DragstrReadSTRAnalyzer (:81-105);
  * best period / repeats -- DragstrReadSTRAnalyzerTest.calculateBestPeriodAndRepeat (:107-124):
    the most repeats, shorter period on ties;
  * imputation -- DragstrPairHMMInputScoreImputator.impute (DragstrPairHMMInputScoreImputator.java:53-70)
    with DragstrParams.lookup (DragstrParams.java:194-200);
  * the DRAGstr parameter text format -- DragstrParamUtils.parse / print (DragstrParamUtils.java:117-200).

Pinned by: the reference's empirical test case (DragstrReadSTRAnalyzerTest.java:62-66, held in
tests/golden/dragstr_empirical_case.txt) and the worked examples in DragstrReadSTRAnalyzer's class
documentation (DragstrReadSTRAnalyzer.java:22-43).  The reference's larger generated case file
(src/test/resources/large/dragstr-str-read-period-repeat-test-cases-with-induced-strs.txt.gz) is a
git-lfs pointer in the checkout, so it is not available.
"""
from __future__ import annotations

import math
from typing import List, Sequence, Tuple


def expected_repeat_length(seq: bytes, position: int, period: int) -> int:
    """DragstrReadSTRAnalyzerTest.java:129-171."""
    n = len(seq)
    if period > n:
        return 0
    start = max(0, position - period + 1)
    best = 0
    end = start + period
    while start <= position + 1 and end <= n:
        unit = seq[start:end]
        forward = 1
        off = end
        while off <= n - period:
            if seq[off:off + period] != unit:
                break
            forward += 1
            off += period
        backward = 0
        off = start - period
        while off >= 0:
            if seq[off:off + period] != unit:
                break
            backward += 1
            off -= period
        cand = forward + backward
        if position < period - 1 and start == 0:
            cand = max(1, cand - 1)
        best = max(best, cand)
        start += 1
        end += 1
    return best


def best_period_and_repeat(seq: bytes, position: int, max_period: int) -> Tuple[int, int]:
    """DragstrReadSTRAnalyzerTest.java:112-124."""
    bp, br = 1, expected_repeat_length(seq, position, 1)
    for p in range(2, max_period + 1):
        c = expected_repeat_length(seq, position, p)
        if c > br:
            bp, br = p, c
    return bp, br


def analyze(seq: bytes, max_period: int) -> Tuple[List[int], List[int]]:
    per, rep = [], []
    for i in range(len(seq)):
        p, r = best_period_and_repeat(seq, i, max_period)
        per.append(p)
        rep.append(r)
    return per, rep


def lookup(table: Sequence[Sequence[float]], period: int, repeats: int) -> float:
    """DragstrParams.java:194-200 (period and repeats clamp to the table's last row / column)."""
    max_p, max_r = len(table), len(table[0])
    pi = period - 1 if period < max_p else max_p - 1
    ri = repeats - 1 if repeats < max_r else max_r - 1
    return table[pi][ri]


def java_round(x: float) -> int:
    """Math.round(double): floor(x + 0.5)."""
    return int(math.floor(x + 0.5))


def impute(seq: bytes, gop_table, gcp_table) -> Tuple[List[int], List[int]]:
    """DragstrPairHMMInputScoreImputator.java:53-70: (gop, gcp) per position as unsigned bytes."""
    n = len(seq)
    per, rep = analyze(seq, len(gop_table))
    gop, gcp = [0] * n, [0] * n
    for i in range(n - 1):
        gop[i] = java_round(min(40.0, lookup(gop_table, per[i], rep[i]))) & 0xFF
        gcp[i] = java_round(lookup(gcp_table, per[i], rep[i])) & 0xFF
    if n:
        gop[n - 1], gcp[n - 1] = 45, 10
    return gop, gcp


def parse_params_text(text: str):
    """DragstrParamUtils.java:117-172: header of repeat numbers 1..R, then named tables 'GOP:', 'GCP:',
    'API:' of P rows each.  Returns (gop, gcp, api) as lists of rows."""
    lines = text.splitlines()
    k = 0
    while k < len(lines) and lines[k].startswith("#"):
        k += 1
    header = [int(x) for x in lines[k].split()]
    if header != list(range(1, len(header) + 1)):
        raise ValueError("header must list repeat lengths 1..R")
    tables, name, rows = {}, None, []
    for line in lines[k + 1:]:
        if not line.strip():
            continue
        if line.rstrip().endswith(":"):
            if name is not None:
                tables[name] = rows
            name, rows = line.rstrip()[:-1], []
        else:
            vals = [float(x) for x in line.split()]
            if len(vals) != len(header):
                raise ValueError("row length does not match the header")
            rows.append(vals)
    if name is not None:
        tables[name] = rows
    return tables["GOP"], tables["GCP"], tables["API"]
