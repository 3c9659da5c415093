"""DRAGstr STR context and GOP/GCP imputation (host code in the native library) against the oracle.

The oracle restates the reference test's brute force (DragstrReadSTRAnalyzerTest.java:107-171) and
is itself pinned here by the reference's empirical case (:62-66) and the worked examples of the
analyzer's class documentation (DragstrReadSTRAnalyzer.java:22-43).
"""
import os

import numpy as np
import pytest

from gatk_amd import dragstr
from gatk_amd.batch import pack_regions
from oracle import dragstr_oracle as ora

HERE = os.path.dirname(__file__)


def _empirical_case():
    with open(os.path.join(HERE, "golden", "dragstr_empirical_case.txt")) as f:
        lines = [ln.strip() for ln in f if not ln.startswith("#")]
    seq = lines[0].encode()
    per = [int(x, 16) for x in lines[1].split()]
    rep = [int(x, 16) for x in lines[2].split()]
    return seq, per, rep


def test_oracle_pinned_by_reference_empirical_case():
    seq, per, rep = _empirical_case()
    p, r = ora.analyze(seq, 8)
    assert p[: len(per)] == per
    assert [min(20, x) for x in r[: len(rep)]] == rep


@pytest.mark.parametrize("seq,per,rep", [
    (b"AAATATATATAAA", "1222222222211", "3444444444433"),   # DragstrReadSTRAnalyzer.java:24-26
    (b"ATATATATATATAT", "22222222222222", "67777777777777"),  # :38-41
])
def test_oracle_pinned_by_class_doc_examples(seq, per, rep):
    p, r = ora.analyze(seq, 8)
    assert p == [int(c) for c in per]
    assert r == [int(c) for c in rep]


def test_oracle_first_position_discount():
    # DragstrReadSTRAnalyzer.java:230-234: ^CACATG -> periods 1 2 2 2 1 1, repeats 1 2 2 2 1 1 (the
    # comment's "12221" lists the first five repeat counts)
    p, r = ora.analyze(b"CACATG", 6)
    assert p[:5] == [1, 2, 2, 2, 1]
    assert r[:5] == [1, 2, 2, 2, 1]


def _random_str_seq(rng, n):
    b = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, n)].copy()
    for pos in range(n):
        if rng.random() < 1.5 / n:
            L = int(rng.integers(1, 11))
            reps = int(rng.integers(2, 12))
            unit = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, L)]
            for k in range(reps):
                o = pos + k * L
                if o >= n - L:
                    break
                b[o:o + L] = unit
    return bytes(b)


def test_library_matches_oracle_on_sequences():
    rng = np.random.default_rng(131)
    fixed = [b"TGATTTGCTCTGTCTGCTGCTGCTGCCTTCAGTAGGGTTGCACGCCTGGGCACGCCTGGAAT", b"AGTATACTGAT",
             b"GTCTATATATATTTTAATTAATTAATTAATTAAATATATTTTCTGCTGCCTTTTGGAT", b"AAAAA", b"A",
             b"ACGTAGATCTGTAGCACTATCGAGC"]  # DragstrReadSTRAnalyzerTest.java:176-182
    seqs = fixed + [_random_str_seq(rng, int(rng.integers(5, 200))) for _ in range(60)]
    for s in seqs:
        mp = max(5, len(s) // 4)
        a = dragstr.DragstrReadSTRAnalyzer(s, mp)
        p, r = ora.analyze(s, mp)
        assert a.period.tolist() == p, s
        assert a.repeats.tolist() == r, s


def test_library_matches_oracle_empirical_case():
    seq, per, rep = _empirical_case()
    a = dragstr.DragstrReadSTRAnalyzer(seq, 8)
    assert a.period[: len(per)].tolist() == per
    assert np.minimum(20, a.repeats[: len(rep)]).tolist() == rep
    with pytest.raises(IndexError):
        a.most_repeated_period(len(seq))


def test_params_file_round_trip_and_default():
    path = os.path.join(HERE, "golden", "example.dragstr-params.txt")
    p = dragstr.DragstrParams.read(path)
    g, c, a = ora.parse_params_text(open(path).read())
    assert p.gop_table.tolist() == g and p.gcp_table.tolist() == c and p.api_table.tolist() == a
    assert (p.max_period, p.max_repeats) == (8, 20)
    assert p.gop(1, 7) == 41.5 and p.gop(9, 50) == p.gop(8, 20) == 24.5  # lookup clamps
    q = dragstr.DragstrParams.parse(p.format(source="test"))
    np.testing.assert_array_equal(q.gop_table, p.gop_table)
    np.testing.assert_array_equal(q.api_table, p.api_table)
    d = dragstr.DragstrParams.default()
    assert d.gcp(3, 1) == 3.33 and d.gcp(1, 20) == 10.0 and d.gop(1, 1) == 45.0 and d.api(8, 20) == 4.0
    with pytest.raises(dragstr.BadInput):
        dragstr.DragstrParams.parse("1 3\nGOP:\n1 2\n")
    with pytest.raises(dragstr.BadInput):
        dragstr.DragstrParams.parse("1 2\nGOP:\n1 2\nGCP:\n1 2\n")  # API missing


def test_imputation_matches_oracle():
    params = dragstr.DragstrParams.read(os.path.join(HERE, "golden", "example.dragstr-params.txt"))
    rng = np.random.default_rng(7)
    reads = []
    for _ in range(40):
        s = _random_str_seq(rng, int(rng.integers(1, 160)))
        n = len(s)
        b = np.frombuffer(s, dtype=np.uint8)
        z = np.zeros(n, dtype=np.uint8)
        reads.append((b, z + 30, z + 45, z + 45, z + 10))
    batch = pack_regions([([np.frombuffer(b"ACGT" * 50, dtype=np.uint8)], reads)])
    imp = dragstr.DragstrPairHMMInputScoreImputator(params)
    gop, gcp = imp.impute_batch(batch)
    ctx = dragstr.contexts(batch, params.max_period, params.max_repeats)
    for k, rd in enumerate(reads):
        o, n = int(batch.read_off[k]), int(batch.read_len[k])
        eg, ec = ora.impute(rd[0].tobytes(), params.gop_table.tolist(), params.gcp_table.tolist())
        assert gop[o:o + n].tolist() == eg
        assert gcp[o:o + n].tolist() == ec
        per, rep = ora.analyze(rd[0].tobytes(), params.max_period)
        want = [(min(p, 8) - 1) * 20 + min(r, 20) - 1 for p, r in zip(per, rep)]
        want[-1] = -1
        assert ctx[o:o + n].tolist() == want
    g1, c1 = imp.impute(reads[0][0].tobytes())
    assert list(g1) == ora.impute(reads[0][0].tobytes(), params.gop_table.tolist(), params.gcp_table.tolist())[0]
