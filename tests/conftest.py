import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (run with -m gpu)")


def phred33(s: str, minimum: int = 0) -> np.ndarray:
    """Decode a Phred+33 string; VectorPairHMMUnitTest.normalize (VectorPairHMMUnitTest.java:108-118)."""
    a = np.frombuffer(s.encode("latin-1"), dtype=np.uint8).astype(np.int16) - 33
    return np.where(a < minimum, minimum, a).astype(np.uint8)


def read_pairhmm_testdata():
    """Rows of src/test/resources/pairhmm-testdata.txt decoded as VectorPairHMMUnitTest.java:59-69
    does (base quals clamped to >= 6)."""
    rows = []
    with open(os.path.join(GOLDEN, "pairhmm-testdata.txt")) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            t = line.split()
            rows.append(dict(hap=t[0].encode(), read=t[1].encode(), qual=phred33(t[2], 6), ins=phred33(t[3]),
                             dels=phred33(t[4]), gcp=phred33(t[5]), expected=float(t[6])))
    return rows


def read_hmmresults(kind: str):
    """Rows of expected.<kind>.hmmresults.txt (format PairHMM.java:367, 380)."""
    rows = []
    with open(os.path.join(GOLDEN, f"expected.{kind}.hmmresults.txt")) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            t = line.split()
            rows.append(dict(hap=t[0].encode(), read=t[1].encode(), qual=phred33(t[2]), ins=phred33(t[3]),
                             dels=phred33(t[4]), gcp=phred33(t[5]), expected=float(t[6]), text=t[6]))
    return rows


@pytest.fixture(scope="session")
def testdata_rows():
    return read_pairhmm_testdata()


@pytest.fixture(scope="session")
def hmmresults():
    return {k: read_hmmresults(k) for k in ("Java", "AVX", "Original", "Exact")}
