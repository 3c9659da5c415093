"""GPU parity of the device read-quality preparation (pairhmm_apply_pcr_error_model,
pairhmm_cap_read_qualities) against the read-prep oracle, bit for bit, and of the HMM results file
written through the engine against the reference's expected.Java.hmmresults.txt."""
import io
import os

import numpy as np
import pytest
import torch

from gatk_amd import synth
from gatk_amd.pairhmm import (HipLoglessPairHMM, LikelihoodMatrix, PairHMMNativeArguments, Read,
                              StandardPairHMMInputScoreImputator)
from gatk_amd.readprep import PCRErrorModel, ReadPrep
from oracle import readprep_oracle as rp

pytestmark = pytest.mark.gpu


def _reads(seed, n=300):
    """STR-rich reads (motifs 1-8 bp, long homopolymers) plus the reference test's repeat reads."""
    rng = np.random.default_rng(seed)
    out = []
    for unit in ("A", "AC", "ACG", "ACGT"):
        for k in (1, 2, 3, 5, 10, 15):
            out.append((unit * k).encode())
    out += [b"TTCTTCCCC", b"A", b"AT", b"A" * 45, b"AC" * 30 + b"G"]
    for _ in range(n):
        parts = []
        while sum(map(len, parts)) < int(rng.integers(20, 300)):
            if rng.random() < 0.5:
                motif = bytes(rng.choice(list(b"ACGT"), int(rng.integers(1, 9))))
                parts.append(motif * int(rng.integers(1, 25)))
            else:
                parts.append(bytes(rng.choice(list(b"ACGT"), int(rng.integers(1, 12)))))
        out.append(b"".join(parts)[:300])
    return out


def _pack(reads, rng):
    lens = np.array([len(r) for r in reads], np.int32)
    off = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    bases = np.frombuffer(b"".join(reads), np.uint8).copy()
    n = len(bases)
    quals = rng.integers(0, 256, n).astype(np.uint8)   # includes Java-negative bytes
    ins = rng.integers(0, 256, n).astype(np.uint8)
    dels = rng.integers(0, 256, n).astype(np.uint8)
    mapq = rng.integers(0, 256, len(reads)).astype(np.int32)
    return bases, off, lens, quals, ins, dels, mapq


@pytest.mark.parametrize("model", list(PCRErrorModel))
def test_pcr_error_model_matches_oracle(model):
    rng = np.random.default_rng(int(model) + 1)
    reads = _reads(int(model))
    bases, off, lens, quals, ins, dels, mapq = _pack(reads, rng)
    print(f"[readprep] model={model.name} reads={len(reads)} bases={len(bases)}", flush=True)
    t = lambda a: torch.from_numpy(a).cuda()  # noqa: E731
    d_ins, d_del = t(ins.copy()), t(dels.copy())
    with ReadPrep() as prep:
        prep.apply_pcr_error_model(t(bases), t(off), t(lens), d_ins, d_del, model)
    got_i, got_d = d_ins.cpu().numpy(), d_del.cpu().numpy()
    for r, read in enumerate(reads):
        o, L = int(off[r]), int(lens[r])
        wi, wd = rp.apply_pcr_error_model(read, ins[o:o + L], dels[o:o + L], model.name)
        assert list(got_i[o:o + L]) == wi, (model.name, read)
        assert list(got_d[o:o + L]) == wd, (model.name, read)


@pytest.mark.parametrize("disable", [False, True])
def test_cap_read_qualities_matches_oracle(disable):
    rng = np.random.default_rng(11)
    reads = _reads(5, n=100)
    bases, off, lens, quals, ins, dels, mapq = _pack(reads, rng)
    t = lambda a: torch.from_numpy(a).cuda()  # noqa: E731
    d_q, d_i, d_d = t(quals.copy()), t(ins.copy()), t(dels.copy())
    with ReadPrep() as prep:
        prep.cap_read_qualities(t(off), t(lens), d_q, d_i, d_d, mapping_quality=None if disable else t(mapq),
                                base_quality_threshold=18, disable_cap_to_mapq=disable)
    for r in range(len(reads)):
        o, L = int(off[r]), int(lens[r])
        wq, wi, wd = rp.cap_minimum_read_qualities(quals[o:o + L], ins[o:o + L], dels[o:o + L], int(mapq[r]), 18, disable)
        assert list(d_q.cpu().numpy()[o:o + L]) == wq
        assert list(d_i.cpu().numpy()[o:o + L]) == wi
        assert list(d_d.cpu().numpy()[o:o + L]) == wd


def test_modify_read_qualities_on_a_batch():
    """modifyReadQualities order (PCR model, then caps) on a synthetic batch's read planes."""
    b = synth.workload("plain_small")
    rng = np.random.default_rng(3)
    mapq = rng.integers(20, 61, len(b.read_len)).astype(np.int32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    q, i, d = t(b.read_quals.copy()), t(b.read_ins.copy()), t(b.read_del.copy())
    with ReadPrep() as prep:
        prep.modify_read_qualities(t(b.read_bases), t(b.read_off), t(b.read_len.astype(np.int32)), q, i, d,
                                   PCRErrorModel.AGGRESSIVE, mapping_quality=t(mapq))
    for r in range(len(b.read_len)):
        o, L = int(b.read_off[r]), int(b.read_len[r])
        read = bytes(b.read_bases[o:o + L])
        wi, wd = rp.apply_pcr_error_model(read, b.read_ins[o:o + L], b.read_del[o:o + L], "AGGRESSIVE")
        wq, wi, wd = rp.cap_minimum_read_qualities(b.read_quals[o:o + L], wi, wd, int(mapq[r]), 18, False)
        assert list(q.cpu().numpy()[o:o + L]) == wq
        assert list(i.cpu().numpy()[o:o + L]) == wi
        assert list(d.cpu().numpy()[o:o + L]) == wd


def test_results_file_matches_reference(hmmresults):
    """Every golden row through HipLoglessPairHMM (FP64) with the results writer attached: the file
    has the reference's header and records; the %e value text may differ in the last digit on at most
    one row, as the FP64 kernel sums the last row in another order."""
    rows = hmmresults["Java"]
    buf = io.StringIO()
    with HipLoglessPairHMM(PairHMMNativeArguments(1, True)) as hmm:
        hmm.set_and_initialize_debug_output_stream(buf)
        for r in rows:
            hmm.initialize([r["hap"]])
            m = LikelihoodMatrix([r["hap"]])
            read = Read(np.frombuffer(r["read"], np.uint8), r["qual"], r["ins"], r["dels"])
            hmm.compute_log10_likelihoods(m, [read], _ConstGcp(r["gcp"]))
    lines = buf.getvalue().split("\n")
    with open(os.path.join(os.path.dirname(__file__), "golden", "expected.Java.hmmresults.txt")) as f:
        want = f.read().rstrip("\n").split("\n")
    assert lines[0] == want[0]
    assert len(lines) == len(want)
    head = lambda s: s.rsplit(" ", 1)[0]  # noqa: E731
    assert all(head(a) == head(b) for a, b in zip(lines[1:], want[1:]))
    assert sum(a != b for a, b in zip(lines[1:], want[1:])) <= 1


class _ConstGcp(StandardPairHMMInputScoreImputator):
    """The golden rows carry their own GCP string; impute it as given."""

    def __init__(self, gcp):
        super().__init__(10)
        self.gcp = np.asarray(gcp, np.uint8)

    def impute(self, read):
        imp = super().impute(read)
        imp.gap_continuation_penalties = self.gcp
        return imp


@pytest.mark.parametrize("seed", [0, 1])
def test_dragstr_impute_device_matches_host(seed):
    """The device DRAGstr imputer against the host imputer (itself pinned by tests/test_dragstr.py
    against DragstrReadSTRAnalyzerTest's brute force), bit for bit, on STR-rich reads, with the
    default tables and with a random 8 x 20 table set."""
    from gatk_amd.dragstr import DragstrParams, DragstrPairHMMInputScoreImputator

    rng = np.random.default_rng(seed)
    reads = _reads(10 + seed)
    bases, off, lens, *_ = _pack(reads, rng)

    class _B:
        read_bases, read_off, read_len = bases, off, lens

    tables = [DragstrParams.default()]
    g = rng.uniform(5, 50, (8, 20))
    c = rng.uniform(2, 20, (8, 20))
    tables.append(DragstrParams(g, c, np.zeros((8, 20)), name="random"))
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    with ReadPrep() as prep:
        for params in tables:
            want_gop, want_gcp = DragstrPairHMMInputScoreImputator(params).impute_batch(_B)
            d_gop = torch.zeros(len(bases), dtype=torch.uint8, device="cuda")
            d_gcp = torch.zeros(len(bases), dtype=torch.uint8, device="cuda")
            prep.dragstr_impute(params, t(bases), t(off), t(lens), d_gop, d_gcp)
            np.testing.assert_array_equal(d_gop.cpu().numpy(), want_gop)
            np.testing.assert_array_equal(d_gcp.cpu().numpy(), want_gcp)
