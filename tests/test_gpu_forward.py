"""GPU parity of the gfx950 forward kernels against the FP64 oracle and the reference's golden
vectors, through the C-ABI (drop-in path and batch path).

Tolerances: FP32 cells with FP64 rescue must be within |d log10 L| <= 1e-4 of the FP64 oracle (the
project bar) and within 1e-5 on the reference's 104 VectorPairHMMUnitTest rows (its own bar,
VectorPairHMMUnitTest.java:100); the FP64 kernel must be within 1e-9 of the oracle.
"""
import math
import threading

import numpy as np
import pytest

from gatk_amd import _native, synth
from gatk_amd.batch import BatchEngine, pack_regions
from gatk_amd.pairhmm import (HaplotypeDataHolder, HipLoglessPairHMM, HipPairHmm, LikelihoodMatrix,
                              PairHMMNativeArguments, Read, ReadDataHolder, StandardPairHMMInputScoreImputator)
from oracle import oracle

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-4
FP64_TOL = 1e-9


@pytest.fixture(scope="module")
def engine():
    with BatchEngine() as eng:
        yield eng


@pytest.fixture(scope="module")
def engine64():
    with BatchEngine(PairHMMNativeArguments(1, True)) as eng:
        yield eng


def _binding(use_double=False):
    b = HipPairHmm()
    assert b.load(None)
    b.initialize(PairHMMNativeArguments(1, use_double))
    return b


class _Imputator:
    def __init__(self, ins, dels, gcp):
        self.v = (ins, dels, gcp)

    def impute(self, read):
        from gatk_amd.pairhmm import PairHMMInputScoreImputation
        return PairHMMInputScoreImputation(*self.v)


@pytest.mark.parametrize("use_double", [False, True])
def test_vector_pairhmm_testdata(testdata_rows, use_double):
    """VectorPairHMMUnitTest.testLikelihoodsFromHaplotypesForAvailableImplementations
    (VectorPairHMMUnitTest.java:30-106): one initialize + computeLog10Likelihoods per row, 1e-5."""
    hmm = HipLoglessPairHMM(PairHMMNativeArguments(1, use_double))
    try:
        worst = 0.0
        for row in testdata_rows:
            hmm.initialize([row["hap"]], None, 0, 0)
            read = Read(row["read"], row["qual"], row["ins"], row["dels"])
            hmm.compute_log10_likelihoods(LikelihoodMatrix([row["hap"]]), [read],
                                          _Imputator(row["ins"], row["dels"], row["gcp"]))
            la = hmm.get_log_likelihood_array()
            worst = max(worst, abs(la[0] - row["expected"]))
        assert worst <= 1e-5
    finally:
        hmm.close()


def _single_pair_batch(rows):
    regions = [([r["hap"]], [(np.frombuffer(r["read"], np.uint8), r["qual"], r["ins"], r["dels"], r["gcp"])])
               for r in rows]
    return pack_regions(regions)


def test_hmmresults_fp32_and_fp64(engine, engine64, hmmresults):
    rows = hmmresults["Java"]
    b = _single_pair_batch(rows)
    want = oracle.logless_batch(b, n_threads=8)
    out = np.empty(b.n_pairs)
    engine.upload(b).forward(out=out)
    assert np.max(np.abs(out - want)) <= FP32_TOL
    assert np.max(np.abs(out - np.array([r["expected"] for r in rows]))) <= FP32_TOL
    out64 = np.empty(b.n_pairs)
    engine64.upload(b).forward(mode=_native.RUN_FP64, out=out64)
    assert np.max(np.abs(out64 - want)) <= FP64_TOL
    texts = ["%e" % v for v in out64]
    assert sum(t == r["text"] for t, r in zip(texts, rows)) >= len(rows) - 1


def test_plain_small_drop_in(hmmresults):
    """The plain_small workload (64 reads x 16 haplotypes) through computeLikelihoods."""
    b = synth.workload("plain_small")
    reads, haps = b.region(0)
    want = oracle.logless_batch(b, n_threads=8).reshape(len(reads), len(haps))
    for use_double, tol in ((False, FP32_TOL), (True, FP64_TOL)):
        binding = _binding(use_double)
        try:
            out = np.empty(len(reads) * len(haps))
            binding.compute_likelihoods([ReadDataHolder(*r) for r in reads], [HaplotypeDataHolder(h) for h in haps], out)
        finally:
            binding.done()
        assert np.max(np.abs(out.reshape(len(reads), len(haps)) - want)) <= tol


def test_matrix_allele_order_is_a_permutation_of_the_initialize_order():
    """VectorLoglessPairHMM.computeLog10Likelihoods (VectorLoglessPairHMM.java:140-153): the engine
    writes read-major results in the order of the haplotypes given to initialize, and each matrix
    allele takes the value at its own haplotype's index there, whatever the matrix's allele order."""
    b = synth.workload("plain_small")
    reads, haps = b.region(0)
    reads, haps = reads[:12], haps[:7]
    want = np.array([[oracle.logless(h, *r) for h in haps] for r in reads])  # [read, init order]
    perm = np.random.default_rng(0).permutation(len(haps))
    assert (perm != np.arange(len(haps))).all()
    hmm = HipLoglessPairHMM(PairHMMNativeArguments(1, False))
    try:
        hmm.initialize([bytes(h) for h in haps], None, 0, 0)
        matrix = LikelihoodMatrix([bytes(haps[i]) for i in perm], len(reads))
        hmm.compute_log10_likelihoods(matrix, [Read(r[0], r[1], r[2], r[3]) for r in reads], _InOrder(reads))
        la = hmm.get_log_likelihood_array().reshape(len(reads), len(haps))
    finally:
        hmm.close()
    assert np.max(np.abs(la - want)) <= FP32_TOL  # engine output: initialize order
    for a, i in enumerate(perm):
        for r in range(len(reads)):
            assert matrix.get(a, r) == la[r, i]
    assert np.max(np.abs(matrix.values - want[:, perm].T)) <= FP32_TOL


class _InOrder:
    """Hands out each read's own insertion / deletion / GCP planes in call order (impute is called
    once per read, in list order: VectorLoglessPairHMM.java:120-129)."""

    def __init__(self, reads):
        self.reads, self.k = reads, 0

    def impute(self, read):
        from gatk_amd.pairhmm import PairHMMInputScoreImputation
        r = self.reads[self.k]
        self.k += 1
        assert bytes(r[0]) == bytes(read.bases)
        return PairHMMInputScoreImputation(r[2], r[3], r[4])


def _edge_region(rng):
    hap = rng.choice(list(b"ACGT"), 300).astype(np.uint8)
    haps = [hap, hap[:1], hap[:20], hap[:64], hap[:65], np.frombuffer(b"ACGTNNACGTacgt" * 5, np.uint8)]
    reads = []
    for R in (0, 1, 2, 20, 63, 64, 65, 100, 127, 128, 129, 200, 250):
        start = int(rng.integers(0, 300 - R + 1)) if R <= 300 else 0
        bases = hap[start:start + R].copy()
        if R > 5:
            bases[R // 2] = ord("N")
        if R > 10:
            bases[3] = ord("a")
        reads.append((bases, rng.integers(6, 41, R).astype(np.uint8), rng.integers(20, 60, R).astype(np.uint8),
                      rng.integers(20, 60, R).astype(np.uint8), rng.integers(5, 20, R).astype(np.uint8)))
    return haps, reads


def test_edge_cases_match_oracle():
    """Empty / 1-base reads, reads shorter and longer than the haplotype, stripe boundaries at
    63/64/65/127/128/129 rows, 'N' wildcards in read and haplotype, lowercase bytes."""
    rng = np.random.default_rng(5)
    haps, reads = _edge_region(rng)
    b = pack_regions([(haps, reads)])
    want = oracle.logless_batch(b, n_threads=8)
    for use_double, tol in ((False, FP32_TOL), (True, FP64_TOL)):
        binding = _binding(use_double)
        try:
            out = np.empty(len(reads) * len(haps))
            binding.compute_likelihoods([ReadDataHolder(*r) for r in reads], [HaplotypeDataHolder(h) for h in haps], out)
        finally:
            binding.done()
        inf = np.isinf(want)
        assert np.array_equal(np.isinf(out), inf)
        assert np.all(out[inf] == -math.inf)  # the empty read
        assert np.max(np.abs(out[~inf] - want[~inf])) <= tol
        assert np.all(out <= 0)


def test_fp64_kernel_every_read_length_to_130(engine64):
    """The FP64 kernel against the oracle at 1e-9 for every read length
    1-130 (rows ending on every lane of a first and second stripe) on haplotypes of 3, 10, 40 and 70
    bases, whose column counts put a read's last rows in different 8-step blocks."""
    rng = np.random.default_rng(5)
    regions = []
    for R in range(1, 131):
        haps = [rng.choice(list(b"ACGT"), H).astype(np.uint8) for H in (3, 10, 40, 70)]
        bases = rng.choice(list(b"ACGT"), R).astype(np.uint8)
        regions.append((haps, [(bases, rng.integers(20, 41, R).astype(np.uint8), np.full(R, 45, np.uint8),
                                np.full(R, 45, np.uint8), np.full(R, 10, np.uint8))]))
    b = pack_regions(regions)
    want = oracle.logless_batch(b, n_threads=8)
    dev = engine64.upload(b)
    out = np.empty(b.n_pairs)
    dev.forward(mode=_native.RUN_FP64, out=out)
    dev.free()
    assert np.max(np.abs(out - want)) <= FP64_TOL


def test_fp64_kernel_random_shapes(engine64):
    """The FP64 kernel against the oracle at 1e-9 on 240 random pairs: reads of 1-600 bases, haplotypes
    of 1-1,500 (several column classes and stripe counts), base qualities 2-40."""
    rng = np.random.default_rng(23)
    regions = []
    for _ in range(60):
        haps = [rng.choice(list(b"ACGT"), int(rng.integers(1, 1501))).astype(np.uint8) for _ in range(2)]
        reads = []
        for _ in range(2):
            R = int(rng.integers(1, 601))
            reads.append((rng.choice(list(b"ACGT"), R).astype(np.uint8), rng.integers(2, 41, R).astype(np.uint8),
                          rng.integers(20, 60, R).astype(np.uint8), rng.integers(20, 60, R).astype(np.uint8),
                          rng.integers(5, 20, R).astype(np.uint8)))
        regions.append((haps, reads))
    b = pack_regions(regions)
    want = oracle.logless_batch(b, n_threads=8)
    dev = engine64.upload(b)
    out = np.empty(b.n_pairs)
    dev.forward(mode=_native.RUN_FP64, out=out)
    dev.free()
    fin = np.isfinite(want)
    assert np.array_equal(np.isfinite(out), fin)
    assert np.max(np.abs(out[fin] - want[fin])) <= FP64_TOL


def _repeat_pairs(read_unit, ref_unit, copies):
    """Reads and haplotypes of repeated units with the qualities of PairHMMUnitTest's big-read case
    (base 30, insertion 40, deletion 40, gap continuation 10)."""
    pairs = []
    for n_read, n_ref in copies:
        read = np.frombuffer(read_unit * n_read, np.uint8).copy()
        R = len(read)
        pairs.append((np.frombuffer(ref_unit * n_ref, np.uint8).copy(),
                      (read, np.full(R, 30, np.uint8), np.full(R, 40, np.uint8), np.full(R, 40, np.uint8),
                       np.full(R, 10, np.uint8))))
    return pairs


def test_really_big_reads():
    """PairHMMUnitTest.testReallyBigReads (PairHMMUnitTest.java:420-457): a 16-base read unit and a
    20-base haplotype unit repeated 1/2/10/20/50 and 1/2/10/20/100 times, every case with more
    haplotype copies than read copies (reads up to 800 bases, haplotypes up to 2,000).  The
    reference test only asks that nothing fails; here each pair, through the drop-in in FP32 (with
    the FP64 rescue) and in FP64, also matches the oracle."""
    copies = [(r, h) for r in (1, 2, 10, 20, 50) for h in (1, 2, 10, 20, 100) if h > r]
    pairs = _repeat_pairs(b"ACCAAGTAGTCACCGT", b"ACCAAGTAGTCACCGTAACG", copies)
    b = pack_regions([([hap], [read]) for hap, read in pairs])
    want = oracle.logless_batch(b, n_threads=8)
    assert np.all(np.isfinite(want))
    for use_double, tol in ((False, FP32_TOL), (True, FP64_TOL)):
        binding = _binding(use_double)
        try:
            got = np.empty(len(pairs))
            for k, (hap, read) in enumerate(pairs):
                out = np.empty(1)
                binding.compute_likelihoods([ReadDataHolder(*read)], [HaplotypeDataHolder(hap)], out)
                got[k] = out[0]
        finally:
            binding.done()
        assert np.max(np.abs(got - want)) <= tol


def _wide_regions(rng, hap_lens, read_lens=(100, 250, 400)):
    regions = []
    for H in hap_lens:
        ref = rng.choice(list(b"ACGT"), H + 5).astype(np.uint8)
        haps = [ref[:H], np.concatenate([ref[:H // 2], ref[H // 2 + 3:H + 3]])]
        reads = []
        for R in read_lens:
            st = int(rng.integers(0, H - R))
            bases = ref[st:st + R].copy()
            flip = rng.random(R) < 0.02
            bases[flip] = rng.choice(list(b"ACGT"), int(flip.sum())).astype(np.uint8)
            reads.append((bases, rng.integers(15, 41, R).astype(np.uint8), rng.integers(35, 50, R).astype(np.uint8),
                          rng.integers(35, 50, R).astype(np.uint8), np.full(R, 10, np.uint8)))
        regions.append((haps, reads))
    return regions


def test_haplotypes_beyond_the_fp32_classes(engine):
    """Haplotypes longer than the FP32 kernels' largest LDS class (4,832 bases) run in the FP64
    kernel, whose LDS holds only the haplotype, up to 32,608 bases.  In a batch they sit in their
    own slices next to FP32 ones: their pairs carry the FP64 kernel's values (1e-9 of the oracle),
    the rest stay FP32 (1e-4).  The drop-in takes such a region in FP64 too.  32,609 bases fail
    cleanly."""
    rng = np.random.default_rng(17)
    wide = _wide_regions(rng, (4900, 9000, 20000, 32608))
    normal = _wide_regions(rng, (300, 600), read_lens=(120, 200))
    b = pack_regions(normal + wide)
    want = oracle.logless_batch(b, n_threads=8)
    out = np.empty(b.n_pairs)
    dev = engine.upload(b)
    st = dev.forward(mode=_native.RUN_PERSISTENT_EAGER, out=out)
    dev.free()
    n_normal = int(b.region_pair_off[len(normal)])
    assert st.n_routed_fp64 >= b.n_pairs - n_normal
    assert np.max(np.abs(out[n_normal:] - want[n_normal:])) <= FP64_TOL
    assert np.max(np.abs(out[:n_normal] - want[:n_normal])) <= FP32_TOL
    binding = _binding(False)
    try:
        for k in range(len(normal), len(normal) + len(wide)):
            reads, haps = b.region(k)
            p0, p1 = int(b.region_pair_off[k]), int(b.region_pair_off[k + 1])
            got = np.empty(p1 - p0)
            binding.compute_likelihoods([ReadDataHolder(*r) for r in reads], [HaplotypeDataHolder(h) for h in haps],
                                        got)
            assert np.max(np.abs(got - want[p0:p1])) <= FP64_TOL
        too_long = rng.choice(list(b"ACGT"), 32609).astype(np.uint8)
        read = (too_long[:50].copy(), np.full(50, 30, np.uint8), np.full(50, 40, np.uint8), np.full(50, 40, np.uint8),
                np.full(50, 10, np.uint8))
        with pytest.raises(ValueError):
            binding.compute_likelihoods([ReadDataHolder(*read)], [HaplotypeDataHolder(too_long)], np.empty(1))
    finally:
        binding.done()


def test_bad_arguments_raise():
    binding = _binding()
    try:
        ok = np.full(4, 30, np.uint8)
        hap = [HaplotypeDataHolder(np.frombuffer(b"ACGTACGT", np.uint8))]
        out = np.empty(1)
        for bad in ("qual", "ins", "gcp"):
            arrs = [np.frombuffer(b"ACGT", np.uint8), ok.copy(), ok.copy(), ok.copy(), ok.copy()]
            if bad == "qual":
                arrs[1][2] = 255
            elif bad == "ins":
                arrs[2][0] = 128
            else:
                arrs[4][3] = 200
            with pytest.raises(ValueError):
                binding.compute_likelihoods([ReadDataHolder(*arrs)], hap, out)
        with pytest.raises(ValueError):
            binding.compute_likelihoods([ReadDataHolder(np.frombuffer(b"ACGT", np.uint8), ok, ok, ok, ok)],
                                        [HaplotypeDataHolder(np.zeros(0, np.uint8))], out)
    finally:
        binding.done()


def test_batched_uneven_persistent_graph(engine):
    b = synth.workload("batched_uneven", n_regions=24, seed=11)
    dev = engine.upload(b)
    outs = []
    for _ in range(3):
        o = np.empty(b.n_pairs)
        dev.forward(mode=_native.RUN_PERSISTENT, out=o)
        outs.append(o)
    eager = np.empty(b.n_pairs)
    st = dev.forward(mode=_native.RUN_PERSISTENT_EAGER, out=eager)
    assert st.fp32_kernel_ms > 0
    for o in outs[1:] + [eager]:
        assert np.array_equal(o, outs[0])
    rng = np.random.default_rng(0)
    sample = np.sort(rng.choice(b.n_pairs, size=min(4000, b.n_pairs), replace=False))
    want = oracle.logless_batch(b, n_threads=8, pairs=sample)
    assert np.max(np.abs(outs[0][sample] - want)) <= FP32_TOL
    dev.free()


def _schedule(lib, ctx, stripes):
    import ctypes
    t = _native.Tuning(4, 100, 0, stripes)
    assert lib.pairhmm_set_tuning(ctx, ctypes.byref(t)) == 0


@pytest.mark.parametrize("stripes", [0, 1], ids=["packed", "aligned"])
def test_precision_stress_rescue(engine64, stripes):
    """STR-rich haplotypes, low base qualities, long reads: many pairs underflow FP32 and are
    recomputed in FP64 (both stripe schedules); the combined result stays within 1e-4 of the oracle."""
    b = synth.workload("precision_stress", n_regions=2, seed=3)
    rng = np.random.default_rng(1)
    sample = np.sort(rng.choice(b.n_pairs, size=min(300, b.n_pairs), replace=False))
    want = oracle.logless_batch(b, n_threads=8, pairs=sample)
    out = np.empty(b.n_pairs)
    with BatchEngine() as eng:
        _schedule(eng.lib, eng.ctx, stripes)
        dev = eng.upload(b)
        st = dev.forward(mode=_native.RUN_PERSISTENT_EAGER, out=out)
        dev.free()
    assert st.n_rescued > 0 and st.n_routed_fp64 > 0
    finite = np.isfinite(want)
    assert np.array_equal(np.isfinite(out[sample]), finite)
    assert np.max(np.abs(out[sample][finite] - want[finite])) <= FP32_TOL
    out64 = np.empty(b.n_pairs)
    engine64.upload(b).forward(mode=_native.RUN_FP64, out=out64)
    assert np.max(np.abs(out64[sample][finite] - want[finite])) <= FP64_TOL


@pytest.mark.parametrize("stripes", [0, 1], ids=["packed", "aligned"])
def test_rescued_pairs_equal_fp64_kernel(stripes):
    """Pairs the FP32 pass hands to the FP64 kernel carry the FP64 kernel's value bit for bit (the
    per-class rescue launches over per-slice lists); pairs it keeps are FP32 values within 1e-4 of
    the FP64 kernel."""
    b = synth.workload("precision_stress", n_regions=2, seed=5)
    with BatchEngine() as eng:
        _schedule(eng.lib, eng.ctx, stripes)
        dev = eng.upload(b)
        out = np.empty(b.n_pairs)
        st = dev.forward(mode=_native.RUN_PERSISTENT_EAGER, out=out)
        out64 = np.empty(b.n_pairs)
        dev.forward(mode=_native.RUN_FP64, out=out64)
        dev.free()
    # FP64 pairs: those the FP32 pass flagged (FP32 sum L * 2^120 below 1e-28) and those the batch
    # sent there directly because their read's expected log10 emission is below -90.
    rescued = (out + 120 * np.log10(2.0) < -28.0) | _predicted_direct(b)
    assert st.n_rescued == int(rescued.sum()) and 0 < st.n_rescued < b.n_pairs
    np.testing.assert_array_equal(out[rescued], out64[rescued])
    assert np.max(np.abs(out[~rescued] - out64[~rescued])) <= FP32_TOL


@pytest.mark.parametrize("stripes", [0, 1], ids=["packed", "aligned"])
def test_drop_in_routes_like_the_batch(stripes):
    """computeLikelihoods per region on STR-rich regions, under either stripe schedule: pairs whose
    read predicts an FP32 underflow go straight to the FP64 kernel here too, so they carry the batch
    path's FP64 values bit for bit; every other pair is within the FP32 tolerance of the batch."""
    b = synth.workload("precision_stress", n_regions=3, seed=7)
    with BatchEngine() as eng:
        _schedule(eng.lib, eng.ctx, stripes)
        dev = eng.upload(b)
        want = np.empty(b.n_pairs)
        st = dev.forward(mode=_native.RUN_PERSISTENT_EAGER, out=want)
        dev.free()
    binding = _binding(False)
    _schedule(_native.load_library(), binding.context, stripes)
    got = np.empty(b.n_pairs)
    try:
        for k in range(len(b.region_pair_off) - 1):
            reads, haps = b.region(k)
            p0, p1 = int(b.region_pair_off[k]), int(b.region_pair_off[k + 1])
            out = np.empty(p1 - p0)
            binding.compute_likelihoods([ReadDataHolder(*r) for r in reads], [HaplotypeDataHolder(h) for h in haps],
                                        out)
            got[p0:p1] = out
    finally:
        binding.done()
    direct = _predicted_direct(b)
    assert 0 < int(direct.sum()) < b.n_pairs
    assert st.n_routed_fp64 == int(direct.sum())
    np.testing.assert_array_equal(got[direct], want[direct])
    assert np.max(np.abs(got - want)) <= FP32_TOL


def _predicted_direct(b):
    """Pairs whose read scores sum((1-e) log10(1-e) + e log10(e/3)) < -90 over its base qualities,
    each term rounded to 2^-40 fixed point and summed exactly."""
    score_of_q = np.array([math.floor(((1 - e) * math.log10(1 - e) + e * math.log10(e / 3) if e < 1 else -1.0)
                                      * 2.0 ** 40 + 0.5) for e in (10.0 ** (-q / 10.0) for q in range(256))],
                          dtype=np.int64)
    read_score = np.array([int(score_of_q[b.read_quals[o:o + n]].sum())
                           for o, n in zip(b.read_off.tolist(), b.read_len.tolist())])
    return read_score[b.pair_read] < -90 * 2 ** 40


def test_concurrent_contexts():
    """Independent contexts on two threads (Spark runs many engines per JVM,
    HaplotypeCallerSpark.java:172-183) give the same answers as one context."""
    b = synth.workload("plain_small")
    reads, haps = b.region(0)
    rh = [ReadDataHolder(*r) for r in reads]
    hh = [HaplotypeDataHolder(h) for h in haps]
    results, errors = {}, []

    def work(k):
        try:
            binding = _binding()
            out = np.empty(len(rh) * len(hh))
            for _ in range(5):
                binding.compute_likelihoods(rh, hh, out)
            binding.done()
            results[k] = out
        except Exception as e:  # pragma: no cover - reported below
            errors.append(e)

    ts = [threading.Thread(target=work, args=(k,)) for k in range(3)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors
    assert np.array_equal(results[0], results[1]) and np.array_equal(results[0], results[2])
