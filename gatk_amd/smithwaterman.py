"""Smith-Waterman on the gfx950 engine: the drop-in for GKL's SWAlignerNativeBinding.

``HipSmithWatermanBinding`` mirrors the binding (load / align / close) over ``pairhmm_sw_align``;
``SWNativeAlignerWrapper`` mirrors the reference's wrapper around any binding
(src/main/java/org/broadinstitute/hellbender/utils/smithwaterman/SWNativeAlignerWrapper.java:19-60),
including its exact-substring shortcut; ``align_batch`` scores many pairs in one launch
(``pairhmm_sw_align_batch``).  The alignment itself always runs in sw.hip.
"""
from __future__ import annotations

import ctypes
import enum
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native
from .pairhmm import _as_u8, _raise_for


@dataclass(frozen=True)
class SWParameters:
    match_value: int
    mismatch_penalty: int
    gap_open_penalty: int
    gap_extend_penalty: int

    def native(self) -> _native.SWParameters:
        return _native.SWParameters(self.match_value, self.mismatch_penalty, self.gap_open_penalty,
                                    self.gap_extend_penalty)


# SmithWatermanAlignmentConstants.java:30-68
ORIGINAL_DEFAULT = SWParameters(3, -1, -4, -3)
STANDARD_NGS = SWParameters(25, -50, -110, -6)
NEW_SW_PARAMETERS = SWParameters(200, -150, -260, -11)
ALIGNMENT_TO_BEST_HAPLOTYPE_SW_PARAMETERS = SWParameters(10, -15, -30, -5)


class SWOverhangStrategy(enum.IntEnum):
    SOFTCLIP = 0
    INDEL = 1
    LEADING_INDEL = 2
    IGNORE = 3


@dataclass(frozen=True)
class SmithWatermanAlignment:
    cigar: str
    alignment_offset: int


def last_index_of(reference: bytes, query: bytes) -> int:
    """Utils.lastIndexOf (Utils.java:1124-1138): the last start of query in reference, or -1."""
    return bytes(reference).rfind(bytes(query))


def _cigar_cap(ref_len, alt_len):
    """Bytes that always hold an alignment's CIGAR text and its NUL: at most ref + alt + 4 elements
    (every element but the clips consumes a base), each at most the longer sequence's digits + 1."""
    longest = np.maximum(ref_len, alt_len)
    digits = 1 + (longest >= 10).astype(np.int32) + (longest >= 100) + (longest >= 1000) + (longest >= 10000)
    return (np.asarray(ref_len) + alt_len + 4) * (digits + 1) + 1


class HipSmithWatermanBinding:
    """SWAlignerNativeBinding over the gfx950 engine."""

    def __init__(self) -> None:
        self._lib = None
        self._ctx = ctypes.c_void_p()

    def load(self, tmp_dir=None) -> bool:
        self._lib = _native.load_library()
        if not self._lib.pairhmm_load():
            return False
        a = _native.NativeArguments(1, 0)
        _raise_for(self._lib.pairhmm_initialize(ctypes.byref(a), ctypes.byref(self._ctx)), "initialize")
        return True

    def align(self, ref, alt, parameters: SWParameters, strategy: SWOverhangStrategy) -> Tuple[str, int]:
        r, a = np.ascontiguousarray(_as_u8(ref)), np.ascontiguousarray(_as_u8(alt))
        cap = int(_cigar_cap(len(r), len(a)))
        buf = ctypes.create_string_buffer(cap)
        off = ctypes.c_int32()
        p = parameters.native()
        st = self._lib.pairhmm_sw_align(self._ctx, r.ctypes.data, len(r), a.ctypes.data, len(a), ctypes.byref(p),
                                        int(strategy), buf, cap, ctypes.byref(off))
        _raise_for(st, "align")
        return buf.value.decode("ascii"), off.value

    @staticmethod
    def _pack(refs, alts):
        n = len(refs)
        if n != len(alts):
            raise ValueError("refs and alts differ in length")
        rl = np.fromiter(map(len, refs), np.int32, n)
        al = np.fromiter(map(len, alts), np.int32, n)
        ro = np.concatenate([[0], np.cumsum(rl, dtype=np.int64)[:-1]]).astype(np.int64)
        ao = np.concatenate([[0], np.cumsum(al, dtype=np.int64)[:-1]]).astype(np.int64)
        # bytes.join takes any buffer (bytes, bytearray, uint8 arrays) without a per-item copy; the
        # library only reads the joined bases, so the read-only views are passed as they are.
        rb = np.frombuffer(b"".join(refs), np.uint8) if n else np.zeros(0, np.uint8)
        ab = np.frombuffer(b"".join(alts), np.uint8) if n else np.zeros(0, np.uint8)
        return n, rl, al, ro, ao, rb, ab

    def align_batch(self, refs: Sequence[bytes], alts: Sequence[bytes], parameters: SWParameters,
                    strategy: SWOverhangStrategy) -> List[Tuple[str, int]]:
        """Every pair in one launch (pairhmm_sw_align_batch_packed): the CIGARs come back as one text
        with end offsets, decoded once."""
        n, rl, al, ro, ao, rb, ab = self._pack(refs, alts)
        if n == 0:
            return []
        cap = int(_cigar_cap(rl, al).astype(np.int64).sum()) + 1
        text = np.empty(cap, np.uint8)  # only the used prefix is written (and paged in)
        ends = np.empty(n, np.int64)
        offs = np.empty(n, np.int32)
        p = parameters.native()
        st = self._lib.pairhmm_sw_align_batch_packed(self._ctx, n, rb.ctypes.data, ro.ctypes.data, rl.ctypes.data,
                                                     ab.ctypes.data, ao.ctypes.data, al.ctypes.data, ctypes.byref(p),
                                                     int(strategy), text.ctypes.data, cap, ends.ctypes.data,
                                                     offs.ctypes.data)
        _raise_for(st, "align_batch")
        e = ends.tolist()
        txt = text[:e[-1]].tobytes().decode("ascii")
        return [(txt[a:b], f) for a, b, f in zip([0] + e[:-1], e, offs.tolist())]

    def align_batch_buffers(self, refs: Sequence[bytes], alts: Sequence[bytes], parameters: SWParameters,
                            strategy: SWOverhangStrategy) -> List[Tuple[str, int]]:
        """The same through pairhmm_sw_align_batch: one NUL-terminated CIGAR buffer per pair."""
        n, rl, al, ro, ao, rb, ab = self._pack(refs, alts)
        if n == 0:
            return []
        caps = _cigar_cap(rl, al).astype(np.int32)
        coff = np.concatenate([[0], np.cumsum(caps, dtype=np.int64)[:-1]]).astype(np.int64)
        cig = np.empty(int(caps.sum()), np.uint8)  # each record is NUL-terminated by the library
        offs = np.zeros(n, np.int32)
        p = parameters.native()
        st = self._lib.pairhmm_sw_align_batch(self._ctx, n, rb.ctypes.data, ro.ctypes.data, rl.ctypes.data,
                                              ab.ctypes.data, ao.ctypes.data, al.ctypes.data, ctypes.byref(p),
                                              int(strategy), cig.ctypes.data, coff.ctypes.data, caps.ctypes.data,
                                              offs.ctypes.data)
        _raise_for(st, "align_batch")
        base = cig.ctypes.data
        return [(ctypes.string_at(base + o).decode("ascii"), f) for o, f in zip(coff.tolist(), offs.tolist())]

    def close(self) -> None:
        if self._lib is not None and self._ctx:
            self._lib.pairhmm_done(self._ctx)
            self._ctx = ctypes.c_void_p()


class SWNativeAlignerWrapper:
    """SWNativeAlignerWrapper (SWNativeAlignerWrapper.java:19-60) around the gfx950 binding."""

    def __init__(self, binding: Optional[HipSmithWatermanBinding] = None) -> None:
        self.aligner = binding or HipSmithWatermanBinding()
        if binding is None and not self.aligner.load():
            from .pairhmm import HardwareFeatureException
            raise HardwareFeatureException("Machine does not support the gfx950 SmithWaterman.")

    def align(self, reference, alternate, parameters: SWParameters,
              overhang_strategy: SWOverhangStrategy) -> SmithWatermanAlignment:
        ref, alt = bytes(_as_u8(reference)), bytes(_as_u8(alternate))
        if overhang_strategy in (SWOverhangStrategy.SOFTCLIP, SWOverhangStrategy.IGNORE):
            k = last_index_of(ref, alt)
            if k != -1:
                return SmithWatermanAlignment(f"{len(alt)}M", k)
        cigar, off = self.aligner.align(ref, alt, parameters, overhang_strategy)
        return SmithWatermanAlignment(cigar, off)

    def close(self) -> None:
        self.aligner.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
