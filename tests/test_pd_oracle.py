"""Pins the oracle's partially determined PairHMM (LoglessPDPairHMM.java:34-194).

The reference's fixture for this path (src/test/resources/large/expected.PDHMM.hmmresults.txt, read by
PDPairHMMLikelihoodCalculationEngineUnitTest.java:33-64) is a git-lfs pointer in this checkout, so the
restatement is pinned by reductions to the already pinned LoglessPairHMM restatement:

* no PD bits: bit for bit LoglessPairHMM (every cell takes the NORMAL branch);
* SNP alternatives (isBasePDMatching, :175-194): a column whose alternative set is {b} scores like a
  haplotype carrying b there for reads that never show the haplotype's own base; the full set ACGT
  scores like an 'N' haplotype base;
* a deletion span (DEL_START on its first base, DEL_END on its last, PartiallyDeterminedHaplotype.java:
  196-201) lets a read of X + Y align across X + S + Y as if S were absent: the result equals
  LoglessPairHMM on X + Y shifted by log10(|XY| / |XSY|) (the initial condition is 1 / haplotype
  length), to 1e-12 -- including a read that needs a deletion right after the span (the AFTER_DEL
  deletion operands);
* the state variable is declared outside the row loop (:63), so a DEL_START on the haplotype's last
  base leaves every later row INSIDE_DEL from its first column: the branch never picks up the value
  before the span and the result is exactly LoglessPairHMM on the full haplotype;
* every PD cell dominates its LoglessPairHMM cell (max(branch, main) >= main), so PD >= LoglessPairHMM.

The insertion rule at a deletion end (:79-84) only adds I(R, s) (a read ending in an insertion right
after X) a second time to the final sum; it is pinned in sign and bound only (0 < shift <= log10 2).
"""
import math

import numpy as np
import pytest

from oracle import oracle

SNP, DEL_START, DEL_END, PD_A, PD_C, PD_G, PD_T = 1, 2, 4, 8, 16, 32, 64
ALT_BIT = {ord("A"): PD_A, ord("C"): PD_C, ord("G"): PD_G, ord("T"): PD_T}
BASES = np.frombuffer(b"ACGT", np.uint8)


def _rand_seq(rng, n, alphabet=BASES):
    return bytes(alphabet[rng.integers(0, len(alphabet), n)])


def _quals(rng, n, hi=False):
    if hi:
        return [40] * n, [45] * n, [45] * n, [10] * n
    return (list(rng.integers(10, 41, n)), list(rng.integers(20, 46, n)), list(rng.integers(20, 46, n)),
            list(rng.integers(5, 16, n)))


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("tristate", [True, False])
def test_no_pd_bits_is_logless_bit_for_bit(seed, tristate):
    rng = np.random.default_rng(seed)
    hap = _rand_seq(rng, int(rng.integers(20, 200)), np.frombuffer(b"ACGTN", np.uint8))
    read = _rand_seq(rng, int(rng.integers(1, 120)), np.frombuffer(b"ACGTN", np.uint8))
    q, i, d, g = _quals(rng, len(read))
    a = oracle.pd_logless(hap, bytes(len(hap)), read, q, i, d, g, tristate=tristate)
    b = oracle.logless(hap, read, q, i, d, g, tristate=tristate)
    assert a == b


@pytest.mark.parametrize("seed", range(6))
def test_snp_single_alternative_is_substituted_haplotype(seed):
    rng = np.random.default_rng(100 + seed)
    # Haplotype of G/T with SNP columns offering A or C; reads of A/C never show a G/T base.
    hap = bytearray(_rand_seq(rng, 120, np.frombuffer(b"GT", np.uint8)))
    pd = bytearray(len(hap))
    sub = bytearray(hap)
    for j in rng.choice(len(hap), 12, replace=False):
        alt = b"AC"[int(rng.integers(0, 2))]
        pd[j] = SNP | ALT_BIT[alt]
        sub[j] = alt
    read = _rand_seq(rng, 60, np.frombuffer(b"AC", np.uint8))
    q, i, d, g = _quals(rng, len(read))
    assert oracle.pd_logless(bytes(hap), bytes(pd), read, q, i, d, g) == oracle.logless(bytes(sub), read, q, i, d, g)


@pytest.mark.parametrize("seed", range(4))
def test_snp_full_alternative_set_is_n_base(seed):
    rng = np.random.default_rng(200 + seed)
    hap = bytearray(_rand_seq(rng, 150))
    pd = bytearray(len(hap))
    sub = bytearray(hap)
    for j in rng.choice(len(hap), 10, replace=False):
        pd[j] = SNP | PD_A | PD_C | PD_G | PD_T
        sub[j] = ord("N")
    read = _rand_seq(rng, 80)
    q, i, d, g = _quals(rng, len(read))
    assert oracle.pd_logless(bytes(hap), bytes(pd), read, q, i, d, g) == oracle.logless(bytes(sub), read, q, i, d, g)


def test_lower_case_read_base_matches_snp_set_only():
    """'a' != 'A' as bytes (:163), but isBasePDMatching takes 'a' for A (:181-183)."""
    hap = b"CCCCACCCC"
    pd_none = bytes(9)
    pd_snp = bytearray(9)
    pd_snp[4] = SNP | PD_G | PD_A
    read = b"CCaCC"
    q, i, d, g = [30] * 5, [40] * 5, [40] * 5, [10] * 5
    plain = oracle.pd_logless(hap, pd_none, read, q, i, d, g)
    snp = oracle.pd_logless(hap, bytes(pd_snp), read, q, i, d, g)
    assert plain == oracle.logless(hap, read, q, i, d, g)
    assert snp > plain + 2.0  # the 'a' now matches the SNP column


def test_snp_with_unexpected_read_base_is_an_error():
    """isBasePDMatching throws for a read base outside ACGTacgt on a SNP column (:195-196)."""
    hap = b"ACGTACGT"
    pd = bytearray(8)
    pd[3] = SNP | PD_A
    read = b"ACGX"
    with pytest.raises(oracle.OracleError):
        oracle.pd_logless(hap, bytes(pd), read, [30] * 4, [40] * 4, [40] * 4, [10] * 4)
    # A read 'N' or a matching base short-circuits before isBasePDMatching (:163).
    oracle.pd_logless(hap, bytes(pd), b"ACGN", [30] * 4, [40] * 4, [40] * 4, [10] * 4)
    oracle.pd_logless(hap, bytes(pd), b"ACGT", [30] * 4, [40] * 4, [40] * 4, [10] * 4)


def _span(rng, lx=80, ls=12, ly=80):
    X, S, Y = _rand_seq(rng, lx), _rand_seq(rng, ls), _rand_seq(rng, ly)
    hap = X + S + Y
    pd = bytearray(len(hap))
    pd[lx] |= DEL_START
    pd[lx + ls - 1] |= DEL_END
    return X, S, Y, hap, pd


def _shifted_xy(X, Y, hap, read, q, i, d, g):
    return oracle.logless(X + Y, read, q, i, d, g) + math.log10(len(X + Y) / len(hap))


@pytest.mark.parametrize("seed", range(5))
@pytest.mark.parametrize("kind", ["across", "deletion_after_span", "insertion_inside_read"])
def test_deletion_span_reduces_to_haplotype_without_span(seed, kind):
    rng = np.random.default_rng(300 + seed)
    X, S, Y, hap, pd = _span(rng)
    if kind == "across":
        read = (X + Y)[50:120]
    elif kind == "deletion_after_span":
        read = X[40:] + Y[2:40]
    else:
        read = X[40:] + b"T" + Y[:30]
    q, i, d, g = _quals(rng, len(read), hi=True)
    pd_l = oracle.pd_logless(hap, bytes(pd), read, q, i, d, g)
    assert pd_l == pytest.approx(_shifted_xy(X, Y, hap, read, q, i, d, g), abs=1e-12)
    assert oracle.logless(hap, read, q, i, d, g) < pd_l - 3.0  # the span costs the plain model dearly


@pytest.mark.parametrize("seed", range(4))
def test_trailing_deletion_start_disables_the_branch(seed):
    """currentState survives the row end (:63): INSIDE_DEL from column 1 of every later row."""
    rng = np.random.default_rng(400 + seed)
    X, S, Y, hap, pd = _span(rng)
    pd[-1] |= DEL_START
    read = (X + Y)[50:120]
    q, i, d, g = _quals(rng, len(read), hi=True)
    assert oracle.pd_logless(hap, bytes(pd), read, q, i, d, g) == oracle.logless(hap, read, q, i, d, g)


@pytest.mark.parametrize("seed", range(4))
def test_deletion_end_insertion_rule_sign_and_bound(seed):
    rng = np.random.default_rng(500 + seed)
    X, Y = _rand_seq(rng, 80), _rand_seq(rng, 80)
    S = b"CA" * 6  # the read's inserted tail must not align to the span itself
    hap = X + S + Y
    pd = bytearray(len(hap))
    pd[80] |= DEL_START
    pd[80 + len(S) - 1] |= DEL_END
    read = X[40:] + b"TTG"
    q, i, d, g = _quals(rng, len(read), hi=True)
    shift = oracle.pd_logless(hap, bytes(pd), read, q, i, d, g) - _shifted_xy(X, Y, hap, read, q, i, d, g)
    assert 1e-6 < shift <= math.log10(2.0) + 1e-12


@pytest.mark.parametrize("seed", range(8))
def test_pd_dominates_logless(seed):
    rng = np.random.default_rng(600 + seed)
    H = int(rng.integers(30, 200))
    hap = _rand_seq(rng, H)
    pd = bytearray(H)
    for _ in range(int(rng.integers(1, 5))):
        s = int(rng.integers(0, H))
        e = min(H - 1, s + int(rng.integers(0, 6)))
        pd[s] |= DEL_START
        pd[e] |= DEL_END
    read = _rand_seq(rng, int(rng.integers(5, 100)))
    q, i, d, g = _quals(rng, len(read))
    assert oracle.pd_logless(hap, bytes(pd), read, q, i, d, g) >= oracle.logless(hap, read, q, i, d, g)


def test_batch_matches_single_pairs():
    rng = np.random.default_rng(7)
    reads = []
    for _ in range(5):
        r = _rand_seq(rng, int(rng.integers(1, 60)))
        reads.append((r,) + tuple(np.asarray(x, np.uint8) for x in _quals(rng, len(r))))
    haps = []
    for _ in range(4):
        X, S, Y, hap, pd = _span(rng, 30, 4, 30)
        pd[5] |= SNP | PD_C
        haps.append((hap, bytes(pd)))
    out = oracle.pd_logless_batch(reads, haps, n_threads=4)
    for r, rd in enumerate(reads):
        for h, (hap, pd) in enumerate(haps):
            assert out[r, h] == oracle.pd_logless(hap, pd, *rd)
