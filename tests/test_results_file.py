"""The HMM results-file format (PairHMM.java:364-385) written by gatk_amd.pairhmm: Java's "%e" and the
record layout.  CPU only; the GPU test writes a whole file through the engine
(tests/test_gpu_readprep.py::test_results_file_matches_reference)."""
import math

import numpy as np

from gatk_amd.pairhmm import RESULTS_HEADER, ReadDataHolder, java_format_e, phred_to_fastq, results_line
from oracle import oracle


def test_format_matches_reference_text_on_oracle_values(hmmresults):
    """The FP64 oracle's values formatted as Java does reproduce every reference row's text."""
    for kind in ("Java",):
        for r in hmmresults[kind]:
            v = oracle.logless(r["hap"], r["read"], r["qual"], r["ins"], r["dels"], r["gcp"])
            assert java_format_e(v) == r["text"]


def test_format_special_values():
    assert java_format_e(float("-inf")) == "-Infinity"
    assert java_format_e(float("inf")) == "Infinity"
    assert java_format_e(float("nan")) == "NaN"
    assert java_format_e(0.0) == "0.000000e+00"
    assert java_format_e(-0.0) == "-0.000000e+00"
    assert java_format_e(-1.0) == "-1.000000e+00"
    assert java_format_e(-123.4567891) == "-1.234568e+02"
    assert java_format_e(1e-300) == "1.000000e-300"
    assert java_format_e(9.9999995) == "1.000000e+01"  # carries into the exponent
    # Java rounds the shortest decimal digits HALF_UP: 1.0000015 prints ...02, where C's printf
    # rounds the exact binary value 1.00000149999999998762...
    assert java_format_e(1.0000015) == "1.000002e+00"
    assert "%e" % 1.0000015 == "1.000001e+00"


def test_record_layout():
    z = np.zeros(3, np.uint8)
    rd = ReadDataHolder(np.frombuffer(b"ACG", np.uint8), np.array([30, 31, 32], np.uint8), z + 45, z + 44, z + 10)
    line = results_line(b"ACGT", rd, -1.5)
    assert line == "\nACGT ACG ?@A NNN MMM +++ -1.500000e+00"
    assert RESULTS_HEADER == "# hap-bases read-bases read-qual read-ins-qual read-del-qual gcp expected-result"
    assert phred_to_fastq([0, 93]) == "!~"
    assert math.isfinite(float("-1.500000e+00"))
