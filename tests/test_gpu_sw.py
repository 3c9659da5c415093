"""GPU parity of the Smith-Waterman kernel (pairhmm_sw_align / pairhmm_sw_align_batch, sw.hip)
against the oracle: CIGAR text and alignment offset identical (integer arithmetic)."""
import numpy as np
import pytest

from gatk_amd.smithwaterman import (ALIGNMENT_TO_BEST_HAPLOTYPE_SW_PARAMETERS, NEW_SW_PARAMETERS, ORIGINAL_DEFAULT,
                                    STANDARD_NGS, HipSmithWatermanBinding, SWNativeAlignerWrapper,
                                    SWOverhangStrategy)
from oracle import oracle
from test_sw_oracle import CASES

pytestmark = pytest.mark.gpu
PRESETS = [ORIGINAL_DEFAULT, STANDARD_NGS, NEW_SW_PARAMETERS, ALIGNMENT_TO_BEST_HAPLOTYPE_SW_PARAMETERS]


@pytest.fixture(scope="module")
def binding():
    b = HipSmithWatermanBinding()
    assert b.load()
    yield b
    b.close()


def _p(p):
    return (p.match_value, p.mismatch_penalty, p.gap_open_penalty, p.gap_extend_penalty)


@pytest.mark.parametrize("ref,read,params,strategy,offset,cigar", CASES)
def test_reference_cases(binding, ref, read, params, strategy, offset, cigar):
    print(f"[sw] case {strategy} {len(ref)}x{len(read)}", flush=True)
    w = SWNativeAlignerWrapper(binding)
    a = w.align(ref.encode(), read.encode(), params, SWOverhangStrategy[strategy])
    assert (a.cigar, a.alignment_offset) == (cigar, offset)
    # The full alignment too (no substring shortcut), against the oracle.
    assert binding.align(ref.encode(), read.encode(), params, SWOverhangStrategy[strategy]) == \
        oracle.sw_align(ref.encode(), read.encode(), _p(params), strategy)


def _read_from(rng, hap, n):
    s = int(rng.integers(0, max(1, len(hap) - n)))
    read = bytearray(hap[s:s + n])
    for _ in range(int(rng.integers(0, 6))):
        k = int(rng.integers(0, max(1, len(read))))
        r = rng.random()
        if r < 0.5:
            if read:
                read[k % len(read)] = b"ACGT"[int(rng.integers(0, 4))]
        elif r < 0.75:
            del read[k:k + int(rng.integers(1, 6))]
        else:
            read[k:k] = bytes(np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, int(rng.integers(1, 6)))])
    return bytes(read) or b"C"


@pytest.mark.parametrize("strategy", list(SWOverhangStrategy))
@pytest.mark.parametrize("preset", range(4))
def test_batch_matches_oracle(binding, strategy, preset):
    """Reads of 1-300 bases (stripe boundaries 63/64/65/128/129 included) against haplotypes of
    1-700 bases, with substitutions and indels, plus unrelated pairs."""
    rng = np.random.default_rng(100 * preset + int(strategy))
    params = PRESETS[preset]
    refs, alts = [], []
    lens = [1, 2, 63, 64, 65, 127, 128, 129, 150, 151, 250, 300]
    for k in range(160):
        H = int(rng.integers(1, 700))
        hap = bytes(np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, H)])
        n = lens[k % len(lens)] if k < 48 else int(rng.integers(1, 300))
        alt = _read_from(rng, hap, n) if k % 5 else bytes(np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, n)])
        refs.append(hap)
        alts.append(alt)
    print(f"[sw] batch strategy={strategy.name} preset={preset} pairs={len(refs)}", flush=True)
    got = binding.align_batch(refs, alts, params, strategy)
    for r, a, g in zip(refs, alts, got):
        assert g == oracle.sw_align(r, a, _p(params), strategy.name), (len(r), len(a))


def test_long_sequences(binding):
    """Haplotype-to-reference shapes: 1-3 kb against each other (several 64-column stripes)."""
    rng = np.random.default_rng(7)
    refs, alts = [], []
    for H in (1000, 2000, 3000):
        ref = bytes(np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, H)])
        refs.append(ref)
        alts.append(_read_from(rng, ref, H - 50))
    got = binding.align_batch(refs, alts, NEW_SW_PARAMETERS, SWOverhangStrategy.SOFTCLIP)
    for r, a, g in zip(refs, alts, got):
        assert g == oracle.sw_align(r, a, _p(NEW_SW_PARAMETERS), "SOFTCLIP")


def test_many_element_cigars(binding):
    """Reads with many indels, interleaved with plain reads: CIGARs of 1 to about 80 elements, so
    the batch mixes pairs read back from the compact first-16-element copy with pairs whose full
    element lists come back one by one."""
    rng = np.random.default_rng(23)
    refs, alts, n_el = [], [], []
    for k in range(120):
        H = int(rng.integers(300, 900))
        hap = bytes(np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, H)])
        read = bytearray(hap[20:H - 20])
        sites = np.arange(10, len(read) - 10, 12)
        n_indels = 0 if k % 3 == 0 else min(int(rng.integers(4, 40)), len(sites))
        for pos in sorted(rng.choice(sites, n_indels, replace=False), reverse=True):
            if rng.random() < 0.5:
                del read[pos:pos + int(rng.integers(1, 4))]
            else:
                read[pos:pos] = bytes(np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, int(rng.integers(1, 4)))])
        refs.append(hap)
        alts.append(bytes(read))
    for strategy in SWOverhangStrategy:
        print(f"[sw] many-element CIGARs strategy={strategy.name}", flush=True)
        got = binding.align_batch(refs, alts, ALIGNMENT_TO_BEST_HAPLOTYPE_SW_PARAMETERS, strategy)
        assert got == binding.align_batch_buffers(refs, alts, ALIGNMENT_TO_BEST_HAPLOTYPE_SW_PARAMETERS, strategy)
        for r, a, g in zip(refs, alts, got):
            assert g == oracle.sw_align(r, a, _p(ALIGNMENT_TO_BEST_HAPLOTYPE_SW_PARAMETERS), strategy.name)
            n_el.append(sum(ch.isalpha() for ch in g[0]))
    n_el = np.array(n_el)
    assert (n_el <= 16).any() and (n_el > 16).sum() >= 40 and n_el.max() >= 40, np.bincount(n_el)


def test_empty_sequence_is_an_error(binding):
    with pytest.raises(ValueError):
        binding.align(b"", b"ACGT", ORIGINAL_DEFAULT, SWOverhangStrategy.SOFTCLIP)
    with pytest.raises(ValueError):
        binding.align(b"ACGT", b"", ORIGINAL_DEFAULT, SWOverhangStrategy.INDEL)


def test_packed_and_per_pair_buffers_agree(binding):
    """pairhmm_sw_align_batch_packed (one text, end offsets) and pairhmm_sw_align_batch (one buffer
    per pair) return the same CIGARs and offsets."""
    rng = np.random.default_rng(11)
    refs, alts = [], []
    for k in range(300):
        H = int(rng.integers(1, 600))
        hap = bytes(np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, H)])
        refs.append(hap)
        alts.append(_read_from(rng, hap, int(rng.integers(1, 250))))
    for strategy in SWOverhangStrategy:
        print(f"[sw] packed vs buffers strategy={strategy.name}", flush=True)
        assert binding.align_batch(refs, alts, STANDARD_NGS, strategy) == \
            binding.align_batch_buffers(refs, alts, STANDARD_NGS, strategy)


def test_packed_text_too_small_is_an_error(binding):
    """A CIGAR text buffer that cannot hold every CIGAR (and the closing 0) is rejected after the
    launch, with no write past its end."""
    import ctypes
    refs, alts = [b"ACGTACGTAC", b"AAAACCCCGG"], [b"ACGTTCGTAC", b"CCCC"]
    n, rl, al, ro, ao, rb, ab = binding._pack(refs, alts)
    text = np.full(8, 0x7F, np.uint8)  # "10M" fits, the second CIGAR does not
    ends = np.zeros(n, np.int64)
    offs = np.zeros(n, np.int32)
    p = ORIGINAL_DEFAULT.native()
    st = binding._lib.pairhmm_sw_align_batch_packed(binding._ctx, n, rb.ctypes.data, ro.ctypes.data, rl.ctypes.data,
                                                    ab.ctypes.data, ao.ctypes.data, al.ctypes.data, ctypes.byref(p),
                                                    int(SWOverhangStrategy.SOFTCLIP), text.ctypes.data, 5,
                                                    ends.ctypes.data, offs.ctypes.data)
    assert st != 0
    assert (text[5:] == 0x7F).all()
