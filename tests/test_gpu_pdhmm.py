"""GPU parity of the partially determined PairHMM (pairhmm_pd_compute_likelihoods, pdhmm.hip)
against the FP64 oracle (LoglessPDPairHMM restatement) on the same seeded inputs.

Tolerance: both sides are FP64; the device fuses multiply-adds, so results agree to rounding, held
here at |dlog10 L| <= 1e-9."""
import io

import numpy as np
import pytest

from gatk_amd.pairhmm import (HipPairHmm, LikelihoodMatrix, PairHMMNativeArguments, Read, ReadDataHolder,
                              StandardPairHMMInputScoreImputator)
from gatk_amd.pairhmm import HaplotypeDataHolder as PlainHaplotype
from gatk_amd.pdhmm import (DEL_END, DEL_START, SNP, A, C, G, T, HaplotypeDataHolder, HipLoglessPairPDHMM,
                            HipPDHmm, PDAllele, parse_pd_results_line, pd_bytes_for_alleles)
from oracle import oracle

pytestmark = pytest.mark.gpu
TOL = 1e-9
ALT = {ord("A"): A, ord("C"): C, ord("G"): G, ord("T"): T}


def _rand_seq(rng, n, alphabet=b"ACGT"):
    a = np.frombuffer(alphabet, np.uint8)
    return bytes(a[rng.integers(0, len(a), n)])


def _pd_haplotype(rng, H, n_snp, n_del, trailing_start=False, n_bases=0):
    hap = bytearray(_rand_seq(rng, H))
    for j in rng.choice(H, min(n_bases, H), replace=False):
        hap[j] = ord("N")
    pd = bytearray(H)
    for j in rng.choice(H, min(n_snp, H), replace=False):
        alts = [b for b in b"ACGT" if b != hap[j]]
        pick = rng.choice(alts, int(rng.integers(1, 4)), replace=False)
        pd[j] |= SNP
        for b in pick:
            pd[j] |= ALT[int(b)]
    for _ in range(n_del):
        s = int(rng.integers(0, H))
        e = min(H - 1, s + int(rng.integers(0, 8)))
        pd[s] |= DEL_START
        pd[e] |= DEL_END
    if trailing_start:
        pd[-1] |= DEL_START
    return bytes(hap), bytes(pd)


def _read(rng, R, src=None):
    if src is not None and len(src) >= R and R > 0:
        s = int(rng.integers(0, len(src) - R + 1))
        bases = bytearray(src[s:s + R])
        for k in rng.choice(R, max(1, R // 30), replace=False):
            bases[k] = b"ACGT"[int(rng.integers(0, 4))]
        bases = bytes(bases)
    else:
        bases = _rand_seq(rng, R)
    return (np.frombuffer(bases, np.uint8).copy(), rng.integers(6, 41, R).astype(np.uint8),
            rng.integers(20, 46, R).astype(np.uint8), rng.integers(20, 46, R).astype(np.uint8),
            rng.integers(5, 16, R).astype(np.uint8))


def _device(reads, haps):
    hmm = HipPDHmm()
    assert hmm.load()
    hmm.initialize()
    try:
        out = np.empty(len(reads) * len(haps))
        hmm.compute_likelihoods([ReadDataHolder(*r) for r in reads],
                                [HaplotypeDataHolder(np.frombuffer(h, np.uint8), np.frombuffer(p, np.uint8))
                                 for h, p in haps], out)
    finally:
        hmm.done()
    return out.reshape(len(reads), len(haps))


def _check(reads, haps, tag):
    print(f"[pdhmm] {tag}: reads={len(reads)} haps={len(haps)} "
          f"R<={max(len(r[0]) for r in reads)} H<={max(len(h) for h, _ in haps)}", flush=True)
    got = _device(reads, haps)
    want = oracle.pd_logless_batch(reads, haps, n_threads=8)
    fin = np.isfinite(want)
    assert np.array_equal(fin, np.isfinite(got)), tag
    assert np.array_equal(want[~fin], got[~fin]), tag
    err = np.abs(got[fin] - want[fin])
    assert err.max(initial=0.0) <= TOL, f"{tag}: max |dlog10L| = {err.max()}"


@pytest.mark.parametrize("seed", range(4))
def test_region_matches_oracle(seed):
    """One active region: PD haplotypes with SNP sets, deletion spans and 'N' bases; reads drawn from
    them (lengths across 1, 63, 64, 65 and several stripes)."""
    rng = np.random.default_rng(seed)
    haps = [_pd_haplotype(rng, int(rng.integers(120, 500)), int(rng.integers(0, 6)), int(rng.integers(0, 4)),
                          n_bases=int(rng.integers(0, 3))) for _ in range(6)]
    lens = [1, 2, 63, 64, 65, 127, 128, 129, 150, 151, 250] + list(rng.integers(1, 300, 9))
    reads = [_read(rng, int(R), haps[k % len(haps)][0]) for k, R in enumerate(lens)]
    _check(reads, haps, f"region seed={seed}")


def test_edge_shapes():
    """Haplotypes of 1..3 bases, single-base deletions (DEL_START|DEL_END), a trailing DEL_START
    (INSIDE_DEL carried into every later row), a DEL_END on the last base, SNPs on the first base."""
    rng = np.random.default_rng(11)
    haps = [(b"A", bytes([SNP | C])), (b"AC", bytes([DEL_START | DEL_END, 0])), (b"ACG", bytes([0, 0, DEL_END])),
            _pd_haplotype(rng, 200, 3, 2, trailing_start=True), _pd_haplotype(rng, 64, 0, 1),
            _pd_haplotype(rng, 65, 2, 0), _pd_haplotype(rng, 300, 0, 6)]
    hp = bytearray(haps[-1][1])
    hp[0] |= SNP | G
    haps.append((haps[-1][0], bytes(hp)))
    reads = [_read(rng, R, haps[3][0]) for R in (1, 3, 40, 64, 100, 190)]
    _check(reads, haps, "edge shapes")


def test_dense_deletion_spans():
    """Many overlapping deletion spans (0-24 per haplotype, 1-9 bases), a span opening at the first
    base, DEL_END on the first base and spans ending on the last one, with reads over several
    stripes: blocks with and without a branch column alternate often, so a branch value left
    behind on a block without one would surface."""
    rng = np.random.default_rng(16)
    haps = [_pd_haplotype(rng, int(rng.integers(40, 400)), int(rng.integers(0, 5)), int(rng.integers(0, 25)),
                          trailing_start=bool(k % 3 == 0)) for k in range(8)]
    hp = bytearray(haps[0][1])
    hp[0] |= DEL_START
    haps[0] = (haps[0][0], bytes(hp))
    hp = bytearray(haps[1][1])
    hp[0] |= DEL_END
    hp[-1] |= DEL_END
    haps[1] = (haps[1][0], bytes(hp))
    reads = [_read(rng, int(R), haps[k % len(haps)][0])
             for k, R in enumerate([1, 64, 65, 200, 320] + list(rng.integers(1, 330, 11)))]
    _check(reads, haps, "dense deletion spans")


def test_long_haplotypes_and_reads():
    rng = np.random.default_rng(12)
    haps = [_pd_haplotype(rng, H, 8, 5) for H in (1000, 1500, 2000)]
    reads = [_read(rng, R, haps[k % 3][0]) for k, R in enumerate((300, 450, 600, 777))]
    _check(reads, haps, "long")


def test_haplotype_length_limit():
    """One wave's LDS holds 20 bytes per column plus pads: 8,060 bases fit (and match the oracle),
    8,061 are rejected as the header states."""
    rng = np.random.default_rng(17)
    hap = _pd_haplotype(rng, 8060, 4, 3)
    reads = [_read(rng, R, hap[0]) for R in (1, 70)]
    _check(reads, [hap], "8,060-base haplotype")
    with pytest.raises(ValueError):
        _device(reads, [_pd_haplotype(rng, 8061, 0, 0)])


def test_no_pd_bits_matches_the_fp64_drop_in():
    """Without PD bits the PD kernel is the plain LoglessPairHMM: compare with the engine's own FP64
    drop-in (use_double_precision) on the same inputs."""
    rng = np.random.default_rng(13)
    haps = [(_rand_seq(rng, int(rng.integers(50, 400))), None) for _ in range(5)]
    haps = [(h, bytes(len(h))) for h, _ in haps]
    reads = [_read(rng, int(R), haps[k % 5][0]) for k, R in enumerate(rng.integers(1, 250, 12))]
    pd = _device(reads, haps)
    plain = HipPairHmm()
    assert plain.load()
    plain.initialize(PairHMMNativeArguments(use_double_precision=True))
    try:
        out = np.empty(len(reads) * len(haps))
        plain.compute_likelihoods([ReadDataHolder(*r) for r in reads],
                                  [PlainHaplotype(np.frombuffer(h, np.uint8)) for h, _ in haps], out)
    finally:
        plain.done()
    assert np.abs(pd - out.reshape(pd.shape)).max() <= TOL


def test_empty_read_and_bad_base():
    rng = np.random.default_rng(14)
    hap, pd = _pd_haplotype(rng, 50, 2, 1)
    empty = tuple(np.zeros(0, np.uint8) for _ in range(5))
    short = _read(rng, 10)
    got = _device([empty, short], [(hap, pd)])
    assert got[0, 0] == -np.inf  # no rows: the final sum over row 0 is 0
    assert got[0, 0] == oracle.pd_logless(hap, pd, *empty)
    assert abs(got[1, 0] - oracle.pd_logless(hap, pd, *short)) <= TOL
    bad = list(_read(rng, 5))
    bad[0] = np.frombuffer(b"ACXGT", np.uint8).copy()
    snp_pd = bytearray(50)
    snp_pd[10] = SNP | A
    with pytest.raises(ValueError):
        _device([tuple(bad)], [(hap, bytes(snp_pd))])
    # Without SNP columns an unexpected base is an ordinary mismatch, as in the reference.
    plain_pd = bytes(50)
    got = _device([tuple(bad)], [(hap, plain_pd)])
    assert abs(got[0, 0] - oracle.pd_logless(hap, plain_pd, *bad)) <= TOL


def test_vector_pd_hmm_mask_and_results_file():
    """VectorLoglessPairPDHMM.computeLog10Likelihoods: -inf where the read misses the determined
    extent (:125-135); the results file's records parse back (the reference test's reconstitution,
    PDPairHMMLikelihoodCalculationEngineUnitTest.java:45-54) and re-score to the written value."""
    rng = np.random.default_rng(15)
    ref_allele, alt_allele = b"ACGTT", b"A"
    X, Y = _rand_seq(rng, 100), _rand_seq(rng, 100)
    pdb = pd_bytes_for_alleles(ref_allele, alt_allele)
    hap = X + ref_allele[:1] + ref_allele[1:] + Y
    pd = bytes(101) + pdb + bytes(100)
    assert len(pd) == len(hap)
    alleles = [PDAllele(hap, pd, (1100, 1105)), PDAllele(X + ref_allele + Y, bytes(len(hap)), (1300, 1301))]
    reads = []
    for k in range(4):
        r = _read(rng, 80, hap)
        reads.append(Read(r[0], r[1], r[2], r[3]))
    spans = [(1050, 1129), (1080, 1159), (1220, 1299), (1500, 1579)]
    lk = LikelihoodMatrix([a.bases for a in alleles], len(reads))
    buf = io.StringIO()
    with HipLoglessPairPDHMM() as hmm:
        hmm.set_and_initialize_debug_output_stream(buf)
        hmm.compute_log10_likelihoods(lk, alleles, reads, StandardPairHMMInputScoreImputator(10), 0, spans)
    expect_inf = [[False, False, True, True], [True, True, False, True]]
    for a in range(2):
        for r in range(4):
            assert np.isneginf(lk.get(a, r)) == expect_inf[a][r], (a, r)
    lines = buf.getvalue().split("\n")
    assert lines[0].startswith("# hap-bases\thap-pd-bases")
    assert len(lines) == 1 + 8
    for line in lines[1:]:
        h, p, rd, val = parse_pd_results_line(line)
        if np.isneginf(val):
            continue
        want = oracle.pd_logless(h, p, rd.read_bases, rd.read_quals, rd.insertion_gop, rd.deletion_gop,
                                 rd.overall_gcp)
        assert abs(val - want) <= 5e-7 * max(1.0, abs(want))  # %e keeps 7 significant digits
