"""GPU parity of the double-buffered region stream (pairhmm_stream_*) against the batch forward.

Chunks submitted through the stream run the same FP32 + FP64-rescue kernels as the batch path, so
every streamed likelihood must equal the batch result for the same pairs bit for bit; a sample is
also checked against the FP64 oracle at |d log10 L| <= 1e-4.
"""
import numpy as np
import pytest

from gatk_amd import synth
from gatk_amd.batch import BatchEngine
from gatk_amd.stream import RegionStream, run_streaming
from oracle import oracle

pytestmark = pytest.mark.gpu


def _chunks(n_chunks, per_chunk, seed):
    return [synth.generate(synth.BATCHED_UNEVEN, per_chunk, seed + k, mix=(synth.PRECISION_STRESS, 0.1))
            for k in range(n_chunks)]


def test_stream_matches_batch_forward():
    chunks = _chunks(5, 6, 7100)
    with BatchEngine() as eng:
        want = []
        for c in chunks:
            dev = eng.upload(c)
            out = np.empty(c.n_pairs)
            dev.forward(out=out)
            dev.free()
            want.append(out)
        got = [np.full(c.n_pairs, np.nan) for c in chunks]
        with RegionStream(eng) as rs:
            for c, o in zip(chunks, got):
                rs.submit(c, o)
            st = rs.finish()
    for w, g in zip(want, got):
        np.testing.assert_array_equal(g, w)
    assert st.cells == sum(c.cells() for c in chunks)
    # one chunk against the FP64 oracle
    ref = oracle.logless_batch(chunks[0], n_threads=8)
    assert np.max(np.abs(got[0] - ref)) <= 1e-4


def test_stream_reuse_and_empty_chunk():
    chunks = _chunks(3, 4, 7200)
    empty = chunks[0].subset_regions([])
    with BatchEngine() as eng, RegionStream(eng) as rs:
        outs = []
        for rnd in range(2):  # a finished stream starts again
            o = [np.full(c.n_pairs, np.nan) for c in chunks]
            rs.submit(empty, np.empty(0))
            for c, x in zip(chunks, o):
                rs.submit(c, x)
            st = rs.finish()
            assert st.cells == sum(c.cells() for c in chunks)
            outs.append(o)
    for a, b in zip(*outs):
        assert np.all(np.isfinite(a))
        np.testing.assert_array_equal(a, b)


def test_run_streaming_driver():
    res, outs = run_streaming(40, chunk_regions=10, workers=2, seed=7300, keep_outputs=True)
    assert res.pairs == sum(len(o) for o in outs)
    assert all(np.all(np.isfinite(o) & (o <= 0)) for o in outs)
    assert res.cells > 0 and res.wall_s > 0
