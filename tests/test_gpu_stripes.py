"""The FP32 pass's two stripe schedules against the FP64 oracle.

Packed stripes stream the rows of consecutive pairs through a wave's 64-row stripes (a pair's rows
plus three transition rows); aligned stripes give every pair whole stripes.  Both must hold
|d log10 L| <= 1e-4 on shapes that exercise the packed transitions: short reads (several pairs'
rows in one stripe, lanes left idle past the second pair), reads ending at 61/62/63/64 rows (the
summing row at a stripe's last lanes and the next pair's first rows at lane 0), empty reads, 'N'
haplotype bytes, and haplotypes holding the bytes the packed schedule reserves (0 and 1), which
run aligned whatever the tuning says.  Packed results are identical from run to run, whatever
the wave processed before (the side pass reads only the plane slots the previous stripe wrote),
and the packed kernel's two loop bodies (8 or 16 wave-steps) agree bit for bit.
"""
import numpy as np
import pytest

from gatk_amd import _native, synth
from gatk_amd.autotune import Schedule
from gatk_amd.batch import BatchEngine, pack_regions
from oracle import oracle

pytestmark = pytest.mark.gpu


def _region(rng, read_lens, hap_len=300, n_haps=5, hap_n=False, reserved=False):
    ref = rng.choice(list(b"ACGT"), hap_len + 20).astype(np.uint8)
    haps = []
    for k in range(n_haps):
        h = ref[k:k + hap_len - 3 * k].copy()
        if hap_n and k == 1:
            h[10:14] = ord("N")
        if reserved and k == 2:
            h[5] = 1
        haps.append(h)
    reads = []
    for R in read_lens:
        st = int(rng.integers(0, max(1, hap_len - R)))
        b = np.resize(ref[st:st + R], R).copy()
        flip = rng.random(R) < 0.02
        b[flip] = rng.choice(list(b"ACGT"), int(flip.sum())).astype(np.uint8)
        reads.append((b, rng.integers(10, 41, R).astype(np.uint8), rng.integers(30, 50, R).astype(np.uint8),
                      rng.integers(30, 50, R).astype(np.uint8), rng.integers(8, 15, R).astype(np.uint8)))
    return haps, reads


def _batch(seed):
    rng = np.random.default_rng(seed)
    regions = [
        _region(rng, [1, 2, 3, 5, 9, 17, 30]),             # several pairs per stripe
        _region(rng, [61, 62, 63, 64, 65, 125, 126, 128]),  # stripe-boundary endings
        _region(rng, [0, 40, 0, 150], hap_n=True),         # empty reads, 'N' haplotype bytes
        _region(rng, [100, 170, 250], hap_len=600, n_haps=7),
    ]
    return pack_regions(regions)


@pytest.mark.parametrize("stripes", [0, 1], ids=["packed", "aligned"])
def test_stripe_schedules_match_oracle(stripes):
    b = _batch(3)
    want = oracle.logless_batch(b, n_threads=8)
    with BatchEngine() as eng:
        Schedule(stripe_schedule=stripes).apply(eng)
        dev = eng.upload(b)
        outs = []
        for _ in range(2):
            o = np.empty(b.n_pairs)
            dev.forward(mode=_native.RUN_PERSISTENT_EAGER, out=o)
            outs.append(o)
        dev.free()
    inf = np.isinf(want)
    assert inf.any() and np.array_equal(np.isinf(outs[0]), inf)
    assert np.max(np.abs(outs[0][~inf] - want[~inf])) <= 1e-4
    assert np.array_equal(outs[0], outs[1])


def test_packed_loop_bodies_agree_bit_for_bit():
    """The packed kernel's 8- and 16-step loop bodies compute every cell with the same operations;
    the longer body only adds steps past the haplotype's end.  With one pair per queue chunk (a small
    batch) every pair's rows fall on the same lanes, so both bodies, and the per-class automatic
    choice, give the same bits, within 1e-4 of the oracle.  The 1,000-base haplotypes fall in an LDS
    class where the automatic choice is the 16-step body."""
    rng = np.random.default_rng(17)
    b = pack_regions([_region(rng, [1, 9, 61, 62, 63, 64, 65, 130]),
                      _region(rng, [0, 70, 150], hap_n=True),
                      _region(rng, [90, 200, 251], hap_len=1000, n_haps=4)])
    want = oracle.logless_batch(b, n_threads=8)
    outs = {}
    with BatchEngine() as eng:
        for unroll in (8, 16, 0):
            Schedule(stripe_schedule=0, step_unroll=unroll).apply(eng)
            dev = eng.upload(b)
            o = np.empty(b.n_pairs)
            dev.forward(mode=_native.RUN_PERSISTENT_EAGER, out=o)
            dev.free()
            outs[unroll] = o
    inf = np.isinf(want)
    assert inf.any() and np.array_equal(np.isinf(outs[8]), inf)
    assert np.max(np.abs(outs[8][~inf] - want[~inf])) <= 1e-4
    assert np.array_equal(outs[8].view(np.int64), outs[16].view(np.int64))
    assert np.array_equal(outs[8].view(np.int64), outs[0].view(np.int64))


def test_reserved_haplotype_bytes_run_aligned():
    """Haplotype bytes 0 or 1 switch a batch to aligned stripes: its results equal the aligned
    schedule's bit for bit, with the packed tuning in force."""
    rng = np.random.default_rng(9)
    b = pack_regions([_region(rng, [20, 90, 140], reserved=True)])
    assert (b.hap_bases == 1).any()
    outs = []
    for stripes in (0, 1):
        with BatchEngine() as eng:
            Schedule(stripe_schedule=stripes).apply(eng)
            dev = eng.upload(b)
            o = np.empty(b.n_pairs)
            dev.forward(mode=_native.RUN_PERSISTENT_EAGER, out=o)
            dev.free()
            outs.append(o)
    assert np.array_equal(outs[0], outs[1])
    want = oracle.logless_batch(b, n_threads=8)
    assert np.max(np.abs(outs[0] - want)) <= 1e-4


def test_packed_and_aligned_agree_on_batched_uneven():
    b = synth.workload("batched_uneven", n_regions=30, seed=5)
    outs = []
    for stripes in (0, 1):
        with BatchEngine() as eng:
            Schedule(stripe_schedule=stripes).apply(eng)
            dev = eng.upload(b)
            o = np.empty(b.n_pairs)
            dev.forward(out=o)
            dev.free()
            outs.append(o)
    assert np.max(np.abs(outs[0] - outs[1])) <= 1e-5
    rng = np.random.default_rng(1)
    idx = np.sort(rng.choice(b.n_pairs, 2000, replace=False))
    want = oracle.logless_batch(b, n_threads=8, pairs=idx)
    assert np.max(np.abs(outs[0][idx] - want)) <= 1e-4


@pytest.mark.parametrize("queue_order", [0, 1])
def test_packed_stripes_ending_at_lane_62_in_a_row(queue_order):
    """Reads of 60 then 61 bases on one haplotype: the first pair's rows end at lane 62 of a fresh
    stripe and every later pair's at lane 62 of the next, so each following pair has its constant
    row at lane 63 and its zero row at lane 0 of the next stripe."""
    rng = np.random.default_rng(11)
    haps, reads = _region(rng, [60, 61, 61, 61, 61, 100], n_haps=1)
    b = pack_regions([(haps, reads)])
    want = oracle.logless_batch(b, n_threads=8)
    with BatchEngine() as eng:
        Schedule(stripe_schedule=0, queue_order=queue_order).apply(eng)
        dev = eng.upload(b)
        o = np.empty(b.n_pairs)
        dev.forward(mode=_native.RUN_PERSISTENT_EAGER, out=o)
        dev.free()
    assert np.max(np.abs(o - want)) <= 1e-4


def test_packed_results_do_not_depend_on_the_waves_history():
    """Submission order puts pairs of very different haplotype lengths one after another, so a
    stripe often runs longer than the one before it.  Three passes give the same bits and the same
    rescue list size, and the FP32 pass flags about as many pairs as the aligned schedule."""
    b = synth.workload("batched_uneven", n_regions=1000, seed=1)
    rescued = {}
    with BatchEngine() as eng:
        for stripes in (0, 1):
            Schedule(stripe_schedule=stripes, queue_order=1).apply(eng)
            dev = eng.upload(b)
            outs, counts = [], []
            for _ in range(3):
                o = np.empty(b.n_pairs)
                counts.append(dev.forward(mode=_native.RUN_PERSISTENT_EAGER, out=o).n_rescued)
                outs.append(o)
            dev.free()
            assert counts[0] == counts[1] == counts[2], (stripes, counts)
            assert np.array_equal(outs[0].view(np.int64), outs[1].view(np.int64))
            assert np.array_equal(outs[0].view(np.int64), outs[2].view(np.int64))
            rescued[stripes] = counts[0]
    print(f"rescued: packed {rescued[0]}, aligned {rescued[1]}")
    assert abs(rescued[0] - rescued[1]) <= 0.05 * rescued[1] + 10


def test_tuning_round_trip_and_bad_values():
    """pairhmm_set_tuning stores what pairhmm_get_tuning returns and rejects values outside each
    knob's range (the loop body takes 0, 8 or 16), leaving the previous tuning in force."""
    import ctypes
    with BatchEngine() as eng:
        good = _native.Tuning(2, 150, 1, 0, 16)
        assert eng.lib.pairhmm_set_tuning(eng.ctx, ctypes.byref(good)) == 0
        for bad in (_native.Tuning(2, 150, 1, 0, 12), _native.Tuning(2, 150, 1, 2, 0),
                    _native.Tuning(9, 150, 1, 0, 0), _native.Tuning(2, 150, 3, 0, 0)):
            assert eng.lib.pairhmm_set_tuning(eng.ctx, ctypes.byref(bad)) == _native.PAIRHMM_ERR_BAD_ARG
        got = _native.Tuning()
        assert eng.lib.pairhmm_get_tuning(eng.ctx, ctypes.byref(got)) == 0
        assert (got.waves_per_block, got.grid_percent, got.queue_order, got.stripe_schedule, got.step_unroll) == \
            (2, 150, 1, 0, 16)
