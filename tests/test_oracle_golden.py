"""Pins the FP64 oracle to the reference's committed golden vectors.

* pairhmm-testdata.txt, 104 rows, |d| <= 1e-5 (VectorPairHMMUnitTest.java:54-101)
* expected.Java.hmmresults.txt, 284 real HaplotypeCaller pairs: exact "%e" text match, the check
  HaplotypeCallerIntegrationTest.java:2193-2244 applies to the Java implementation
* expected.Exact / expected.Original: the Log10PairHMM restatements, exact text match
* expected.AVX: GKL's FP32+FP64 output, which agrees with Java to 3e-6
"""
import numpy as np

from oracle import oracle


def test_testdata_rows_within_1e5(testdata_rows):
    assert len(testdata_rows) == 104
    errs = [abs(oracle.logless(r["hap"], r["read"], r["qual"], r["ins"], r["dels"], r["gcp"]) - r["expected"])
            for r in testdata_rows]
    assert max(errs) <= 1e-5


def test_java_hmmresults_exact_text(hmmresults):
    rows = hmmresults["Java"]
    assert len(rows) == 284
    for r in rows:
        v = oracle.logless(r["hap"], r["read"], r["qual"], r["ins"], r["dels"], r["gcp"])
        assert "%e" % v == r["text"], (r["read"], v, r["text"])


def test_exact_hmmresults_exact_text(hmmresults):
    for r in hmmresults["Exact"]:
        v = oracle.log10_pairhmm(r["hap"], r["read"], r["qual"], r["ins"], r["dels"], r["gcp"], exact=True)
        assert "%e" % v == r["text"]


def test_original_hmmresults_exact_text(hmmresults):
    for r in hmmresults["Original"]:
        v = oracle.log10_pairhmm(r["hap"], r["read"], r["qual"], r["ins"], r["dels"], r["gcp"], exact=False)
        assert "%e" % v == r["text"]


def test_avx_hmmresults_close(hmmresults):
    errs = [abs(oracle.logless(r["hap"], r["read"], r["qual"], r["ins"], r["dels"], r["gcp"]) - r["expected"])
            for r in hmmresults["AVX"]]
    assert max(errs) <= 3e-6


def test_golden_files_share_inputs(hmmresults):
    keys = [tuple((r["hap"], r["read"])) for r in hmmresults["Java"]]
    for k in ("AVX", "Original", "Exact"):
        assert [tuple((r["hap"], r["read"])) for r in hmmresults[k]] == keys


def test_batch_matches_single_pairs(hmmresults):
    """oracle_logless_batch (OpenMP over pairs) equals the single-pair path bit for bit."""
    from gatk_amd.batch import pack_regions

    rows = hmmresults["Java"][:40]
    regions = [([r["hap"]], [(np.frombuffer(r["read"], np.uint8), r["qual"], r["ins"], r["dels"], r["gcp"])])
               for r in rows]
    b = pack_regions(regions)
    got = oracle.logless_batch(b, n_threads=4)
    want = [oracle.logless(r["hap"], r["read"], r["qual"], r["ins"], r["dels"], r["gcp"]) for r in rows]
    assert np.array_equal(got, np.array(want))
