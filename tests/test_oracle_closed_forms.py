"""Pins the oracle's recurrences with the reference's closed-form tests
(PairHMMUnitTest.java, tristate correction off as at :33-38) and with brute-force enumeration of
every alignment path on tiny inputs (sum over paths == DP value)."""
import itertools
import math

import numpy as np
import pytest

from oracle import oracle

BASES = b"ACGT"
CONTEXT = b"ACGTAATGACGATTGCA"                              # PairHMMUnitTest.java:55
LEFT_FLANK = b"GATTTATCATCGAGTCTGC"                         # :56
RIGHT_FLANK = b"CATGGATCGTTATCAGCTATCTCGAGGGATTCACTTAACAGTTTTA"  # :57
MASSIVE_QUAL = 100                                          # :31


def _logless(hap, read, q, i, d, g, tristate=False):
    return oracle.logless(hap, read, q, i, d, g, tristate=tristate)


def _exact(hap, read, q, i, d, g, tristate=False):
    return oracle.log10_pairhmm(hap, read, q, i, d, g, exact=True, tristate=tristate)


def _original(hap, read, q, i, d, g, tristate=False):
    return oracle.log10_pairhmm(hap, read, q, i, d, g, exact=False, tristate=tristate)


class Basic:
    """BasicLikelihoodTestProvider (PairHMMUnitTest.java:49-150)."""

    def __init__(self, ref, read, base_q, ins_q, del_q, expected_q, gcp, left=False, right=False):
        self.ref, self.read = ref, read
        self.base_q, self.ins_q, self.del_q, self.gcp = base_q, ins_q, del_q, gcp
        self.expected_q = expected_q
        self.hap = (LEFT_FLANK if left else b"") + CONTEXT + ref + CONTEXT + (RIGHT_FLANK if right else b"")
        self.read_ctx = CONTEXT + read + CONTEXT

    def expected(self):  # :88-91
        return self.expected_q / -10.0 + 0.03 + math.log10(1.0 / len(self.hap))

    def quals(self, q, do_gop, anchor):  # :124-144
        n = len(self.read_ctx)
        if not anchor:
            return np.full(n, q, np.uint8)
        a = np.full(n, MASSIVE_QUAL, np.uint8)
        if do_gop:
            a[len(CONTEXT)] = q
        else:
            a[len(CONTEXT):len(CONTEXT) + len(self.read)] = q
        return a

    def run(self, fn, anchor):
        return fn(self.hap, self.read_ctx, self.quals(self.base_q, False, anchor), self.quals(self.ins_q, True, anchor),
                  self.quals(self.del_q, True, anchor), self.quals(self.gcp, False, anchor))


def _basic_cases(extensive):
    bq = [10, 20, 30, 40, 50] if extensive else [30]
    iq = [20, 30, 40, 50] if extensive else [40]
    gc = [8, 10, 20] if extensive else [10]
    sizes = [2, 3, 4, 5, 7, 8, 9, 10, 20, 30, 35] if extensive else [2]
    for b, i, g in itertools.product(bq, iq, gc):
        for rb, qb in itertools.product(BASES, BASES):
            yield Basic(bytes([rb]), bytes([qb]), b, i, i, 0 if rb == qb else b, g)
        for size in sizes:
            for base in BASES:
                for ins in (True, False):
                    small, big = bytes([base]), bytes([base]) * size
                    ref, read = (small, big) if ins else (big, small)
                    for left, right in ((False, False), (True, False), (False, True), (True, True)):
                        yield Basic(ref, read, b, i, i, i + (size - 2) * g, g, left, right)


def test_basic_likelihoods():
    """testBasicLikelihoods (:227-244): theoretical expectation within 0.2 for every HMM; ORIGINAL
    within 1e-3 of EXACT; LOGLESS within 0.2 of EXACT."""
    n = 0
    for cfg in _basic_cases(extensive=True):
        exact = cfg.run(_exact, True)
        for fn, tol in ((_exact, 1e-9), (_original, 1e-3), (_logless, 0.2)):
            v = cfg.run(fn, True)
            assert abs(v - cfg.expected()) <= 0.2, (cfg.ref, cfg.read, fn.__name__)
            assert abs(v - exact) <= tol
            assert v <= 0.0
        n += 1
    assert n == 60 * (16 + 11 * 4 * 2 * 4)


def test_optimized_likelihoods():
    """testOptimizedLikelihoods (:246-258, data :186-225): random ref/read, flat qualities."""
    rng = np.random.default_rng(7)
    sizes = [3, 20, 50, 90, 160]
    count = underflowed = 0
    for b, i, g in itertools.product([10, 30, 40, 60], [20, 40, 60], [10, 20, 30]):
        for rs, qs in itertools.product(sizes, sizes):
            ref = bytes(rng.choice(list(BASES), rs).astype(np.uint8))
            read = bytes(rng.choice(list(BASES), qs).astype(np.uint8))
            left, right = bool(rng.integers(2)), bool(rng.integers(2))
            cfg = Basic(ref, read, b, i, i, 0, g, left, right)
            exact = cfg.run(_exact, False)
            assert abs(cfg.run(_original, False) - exact) <= 1e-3
            logless = cfg.run(_logless, False)
            if exact > -300.0:
                assert abs(logless - exact) <= 0.2
            else:
                # 2^1020-scaled doubles underflow below log10 L ~ -323 (denormals included), so
                # the Logless value is -inf there, exactly as Java's doubles give.
                assert logless == -math.inf or abs(logless - exact) <= 0.2
                underflowed += logless == -math.inf
            count += 1
    assert count == 36 * 25


def test_logless_underflows_to_minus_infinity_like_java_doubles():
    """Below log10 L ~ -323 the 2^1020-scaled double sum underflows to 0, so LoglessPairHMM
    returns -inf (Math.log10(0)) while the log-space EXACT implementation stays finite."""
    n = 120
    q = lambda v: np.full(n, v, np.uint8)  # noqa: E731
    args = (b"C" * 150, b"A" * n, q(60), q(60), q(60), q(60))
    exact = oracle.log10_pairhmm(*args, exact=True)
    assert exact < -400
    assert oracle.logless(*args) == -math.inf


@pytest.mark.parametrize("centered", [True, False])
def test_mismatch_in_every_position(centered):
    """testMismatchInEveryPositionInTheRead[WithCenteredHaplotype] (:260-314): closed form, 1e-2."""
    hap = b"TTCTCTTCTGTTGTGGCTGGTT"
    match_q, mis_q, indel_q, offset = 90, 20, 80, 2
    n = len(hap) - (2 * offset if centered else offset)
    gop = np.full(n, indel_q, np.uint8)
    for k in range(n):
        quals = np.full(n, match_q, np.uint8)
        quals[k] = mis_q
        mread = bytearray(hap[offset:offset + n])
        mread[k] = ord("T") if mread[k] == ord("C") else ord("C")
        v = _logless(hap, bytes(mread), quals, gop, gop, gop)
        expected = math.log10(1.0 / len(hap) * (1 - 10 ** (-match_q / 10)) ** (n - 1) * 10 ** (-mis_q / 10))
        assert abs(v - expected) <= 1e-2


def _expected_matching(read_len, ref_len, base_q, ins_q):
    """getExpectedMatchingLogLikelihood (:379-388)."""
    init = abs(ref_len - read_len + 1) / ref_len
    if read_len < ref_len:
        return math.log10(init * (1 - 10 ** (-base_q / 10)) ** read_len)
    return math.log10(init * (1 - 10 ** (-base_q / 10)) ** ref_len * (10 ** (-ins_q / 10)) ** (read_len - ref_len))


@pytest.mark.parametrize("read_size,ref_size", [(r, h) for r in (1, 2, 5, 10) for h in (1, 2, 5, 10) if h > r])
def test_all_matching_read(read_size, ref_size):
    """testAllMatchingRead (:366-377): closed form within 1e-3, for all three HMMs."""
    read, ref = b"A" * read_size, b"A" * ref_size
    q = lambda v, n=read_size: np.full(n, v, np.uint8)  # noqa: E731
    for fn in (_logless, _exact, _original):
        d = fn(ref, read, q(20), q(100), q(100), q(100))
        assert abs(d - _expected_matching(read_size, ref_size, 20, 100)) <= 1e-3


@pytest.mark.parametrize("read_size", [1, 2, 5, 10])
def test_likelihoods_non_positive(read_size):
    """testMultipleReadMatchesInHaplotype / testReadSameAsHaplotype (:328-364): log10 L <= 0."""
    read = b"A" * read_size
    q = lambda v: np.full(read_size, v, np.uint8)  # noqa: E731
    for ref_size in (1, 2, 5, 10):
        ref = b"CC" + b"A" * ref_size + b"GGA"
        for fn in (_logless, _exact, _original):
            assert fn(ref, read, q(20), q(37), q(37), q(10)) <= 0.0
    for fn in (_logless, _exact, _original):
        assert fn(read, read, q(20), q(37), q(37), q(10)) <= 0.0


def test_likelihoods_from_haplotypes_closed_form():
    """testLikelihoodsFromHaplotypes (:410-437): read 10 x 'A', haplotype 20 x 'A', base Q 20,
    GOP/GCP 100 -> closed form within 1e-3."""
    q = lambda v: np.full(10, v, np.uint8)  # noqa: E731
    v = _logless(b"A" * 20, b"A" * 10, q(20), q(100), q(100), q(100))
    assert abs(v - _expected_matching(10, 20, 20, 100)) <= 1e-3


def test_really_big_reads():
    """testReallyBigReads (:390-408): up to 800 x 2000 completes with a finite result."""
    read1, ref1 = b"ACCAAGTAGTCACCGT", b"ACCAAGTAGTCACCGTAACG"
    read, ref = read1 * 50, ref1 * 100
    n = len(read)
    v = oracle.logless(ref, read, np.full(n, 30, np.uint8), np.full(n, 40, np.uint8), np.full(n, 40, np.uint8),
                       np.full(n, 10, np.uint8))
    assert math.isfinite(v) and v <= 0


def test_empty_read_is_minus_infinity():
    """R = 0: the last-row sum is over row 0 (M = I = 0), so log10(0) = -inf (LoglessPairHMM.java:62-67)."""
    e = np.zeros(0, np.uint8)
    assert oracle.logless(b"ACGT", b"", e, e, e, e) == -math.inf


def test_find_first_position_where_haplotypes_differ():
    """testFindFirstPositionWhereHaplotypesDiffer (:524-553)."""
    for n1 in range(10, 30):
        for n2 in range(10, 50):
            mn = min(n1, n2)
            for site in range(max(n1, n2) + 1):
                for one_is_diff in (True, False):
                    h1, h2 = bytearray(b"A" * n1), bytearray(b"A" * n2)
                    tgt = h1 if one_is_diff else h2
                    if site < len(tgt):
                        tgt[site] = ord("C")
                        expected = min(site, mn)
                    else:
                        expected = mn
                    assert oracle.first_position_haplotypes_differ(bytes(h1), bytes(h2)) == expected


# ---- brute force over alignment paths ----------------------------------------------------

def _path_sum(hap, read, q, ins, dele, gcp, tristate):
    """Sum over every HMM path of its probability (docs/pair_hmm.tex:80-103): paths start in
    D(0, c0) with weight 1/H for c0 = 0..H, and end in M(R, c) or I(R, c), c >= 1."""
    H, R = len(hap), len(read)
    e = lambda v: 10.0 ** (-int(v) / 10.0)  # noqa: E731
    MM = [None] + [oracle.match_to_match_prob(ins[r], dele[r]) for r in range(R)]
    MI = [None] + [e(ins[r]) for r in range(R)]
    MD = [None] + [e(dele[r]) for r in range(R)]
    GM = [None] + [1 - e(gcp[r]) for r in range(R)]
    GG = [None] + [e(gcp[r]) for r in range(R)]

    def prior(r, c):
        x, y = read[r - 1], hap[c - 1]
        if x == y or x == ord("N") or y == ord("N"):
            return 1 - e(q[r - 1])
        return e(q[r - 1]) / (3.0 if tristate else 1.0)

    total = 0.0

    def walk(state, r, c, w):
        nonlocal total
        if r == R and state in "MI" and c >= 1:
            total += w
        if state == "D" and r == 0:
            if r + 1 <= R and c + 1 <= H:
                walk("M", r + 1, c + 1, w * GM[r + 1] * prior(r + 1, c + 1))
            return
        if state == "M":
            if r + 1 <= R and c + 1 <= H:
                walk("M", r + 1, c + 1, w * MM[r + 1] * prior(r + 1, c + 1))
            if r + 1 <= R:
                walk("I", r + 1, c, w * MI[r + 1])
            if c + 1 <= H:
                walk("D", r, c + 1, w * MD[r])
        elif state == "I":
            if r + 1 <= R and c + 1 <= H:
                walk("M", r + 1, c + 1, w * GM[r + 1] * prior(r + 1, c + 1))
            if r + 1 <= R:
                walk("I", r + 1, c, w * GG[r + 1])
        elif state == "D":
            if r + 1 <= R and c + 1 <= H:
                walk("M", r + 1, c + 1, w * GM[r + 1] * prior(r + 1, c + 1))
            if c + 1 <= H:
                walk("D", r, c + 1, w * GG[r])

    for c0 in range(H + 1):
        walk("D", 0, c0, 1.0 / H)
    return total


@pytest.mark.parametrize("seed", range(12))
def test_dp_equals_sum_over_paths(seed):
    rng = np.random.default_rng(100 + seed)
    H, R = int(rng.integers(1, 5)), int(rng.integers(1, 5))
    hap = bytes(rng.choice(list(b"ACGTN"), H, p=[0.23, 0.23, 0.23, 0.23, 0.08]).astype(np.uint8))
    read = bytes(rng.choice(list(b"ACGTN"), R, p=[0.23, 0.23, 0.23, 0.23, 0.08]).astype(np.uint8))
    q = rng.integers(2, 40, R).astype(np.uint8)
    ins = rng.integers(5, 50, R).astype(np.uint8)
    dele = rng.integers(5, 50, R).astype(np.uint8)
    gcp = rng.integers(3, 30, R).astype(np.uint8)
    tristate = bool(seed % 2)
    want = math.log10(_path_sum(hap, read, q, ins, dele, gcp, tristate))
    got = oracle.logless(hap, read, q, ins, dele, gcp, tristate=tristate)
    assert abs(got - want) <= 1e-12
    assert abs(oracle.log10_pairhmm(hap, read, q, ins, dele, gcp, exact=True, tristate=tristate) - want) <= 1e-12
