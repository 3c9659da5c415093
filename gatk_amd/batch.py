"""Packed multi-region batches and the device-resident batched forward engine.

A ``PackedBatch`` holds many active regions at once in the layout the batch C-ABI takes
(include/gatk_pairhmm.h, ``pairhmm_batch_desc``): read byte planes (bases, base quals, insertion
GOP, deletion GOP, GCP) concatenated with per-read offsets/lengths, haplotype bases concatenated
with per-haplotype offsets/lengths, and an explicit pair list.  Pairs are stored region by region
and, inside a region, read-major -- the order ``computeLikelihoods`` writes
(VectorLoglessPairHMM.java:131, 141-152) -- so ``out[region_pair_off[k]:region_pair_off[k+1]]``
reshaped to (reads, haplotypes) is region k's likelihood matrix.
"""
from __future__ import annotations

import atexit
import ctypes
import weakref
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _native
from .pairhmm import HardwareFeatureException, PairHMMNativeArguments, _as_u8, _raise_for


@dataclass
class PackedBatch:
    read_bases: np.ndarray
    read_quals: np.ndarray
    read_ins: np.ndarray
    read_del: np.ndarray
    read_gcp: np.ndarray
    read_off: np.ndarray  # int64 [n_reads]
    read_len: np.ndarray  # int32 [n_reads]
    hap_bases: np.ndarray
    hap_off: np.ndarray  # int64 [n_haps]
    hap_len: np.ndarray  # int32 [n_haps]
    pair_read: np.ndarray  # int32 [n_pairs]
    pair_hap: np.ndarray  # int32 [n_pairs]
    region_pair_off: np.ndarray  # int64 [n_regions + 1]
    region_read_off: np.ndarray  # int64 [n_regions + 1]
    region_hap_off: np.ndarray  # int64 [n_regions + 1]

    @property
    def n_pairs(self) -> int:
        return int(len(self.pair_read))

    @property
    def n_regions(self) -> int:
        return int(len(self.region_pair_off) - 1)

    def cells(self, pairs: Optional[np.ndarray] = None) -> int:
        """Sum over pairs of R * H (real DP cells, LoglessPairHMM.java:47-48)."""
        pr, ph = self.pair_read, self.pair_hap
        if pairs is not None:
            pr, ph = pr[pairs], ph[pairs]
        return int(np.dot(self.read_len[pr].astype(np.int64), self.hap_len[ph].astype(np.int64)))

    def region(self, k: int):
        """(reads, haplotypes) of region k as lists of byte arrays (for the drop-in path)."""
        r0, r1 = int(self.region_read_off[k]), int(self.region_read_off[k + 1])
        h0, h1 = int(self.region_hap_off[k]), int(self.region_hap_off[k + 1])
        reads = []
        for r in range(r0, r1):
            o, n = int(self.read_off[r]), int(self.read_len[r])
            reads.append(tuple(p[o:o + n] for p in (self.read_bases, self.read_quals, self.read_ins,
                                                    self.read_del, self.read_gcp)))
        haps = [self.hap_bases[int(self.hap_off[h]):int(self.hap_off[h]) + int(self.hap_len[h])] for h in range(h0, h1)]
        return reads, haps

    def subset_regions(self, regions: Sequence[int]) -> "PackedBatch":
        """A new batch holding only the listed regions (re-packed, same per-region layout)."""
        regs = []
        for k in regions:
            reads, haps = self.region(int(k))
            regs.append((haps, reads))
        return pack_regions(regs)

    def desc(self) -> _native.BatchDesc:
        arrs = {
            "read_bases": np.ascontiguousarray(self.read_bases, dtype=np.uint8),
            "read_quals": np.ascontiguousarray(self.read_quals, dtype=np.uint8),
            "read_ins": np.ascontiguousarray(self.read_ins, dtype=np.uint8),
            "read_del": np.ascontiguousarray(self.read_del, dtype=np.uint8),
            "read_gcp": np.ascontiguousarray(self.read_gcp, dtype=np.uint8),
            "read_off": np.ascontiguousarray(self.read_off, dtype=np.int64),
            "read_len": np.ascontiguousarray(self.read_len, dtype=np.int32),
            "hap_bases": np.ascontiguousarray(self.hap_bases, dtype=np.uint8),
            "hap_off": np.ascontiguousarray(self.hap_off, dtype=np.int64),
            "hap_len": np.ascontiguousarray(self.hap_len, dtype=np.int32),
            "pair_read": np.ascontiguousarray(self.pair_read, dtype=np.int32),
            "pair_hap": np.ascontiguousarray(self.pair_hap, dtype=np.int32),
        }
        d = _native.BatchDesc()
        d.n_reads = len(arrs["read_len"])
        d.read_bytes = len(arrs["read_bases"])
        d.n_haps = len(arrs["hap_len"])
        d.hap_bytes = len(arrs["hap_bases"])
        d.n_pairs = len(arrs["pair_read"])
        for k, a in arrs.items():
            setattr(d, k, a.ctypes.data)
        d._keep = arrs  # keep the arrays alive while the descriptor is used
        return d


def reads_routed_to_fp64(batch: PackedBatch) -> np.ndarray:
    """Per pair: whether the engine sends it to the FP64 kernel without an FP32 pass.  The host
    predicate of the C-ABI: the read's expected log10 emission along a perfect alignment,
    sum_i (1-e_i) log10(1-e_i) + e_i log10(e_i/3) over its base-quality error rates, each term
    rounded to 2^-40 fixed point and summed exactly, is below -90."""
    import math

    score_of_q = np.array([math.floor(((1 - e) * math.log10(1 - e) + e * math.log10(e / 3) if e < 1 else -1.0)
                                      * 2.0 ** 40 + 0.5) for e in (10.0 ** (-q / 10.0) for q in range(256))],
                          dtype=np.int64)
    per_base = score_of_q[batch.read_quals]
    csum = np.concatenate([[0], np.cumsum(per_base, dtype=np.int64)])
    off = batch.read_off.astype(np.int64)
    read_score = csum[off + batch.read_len] - csum[off]
    return read_score[batch.pair_read] < -90 * 2 ** 40


def pack_regions(regions: Sequence) -> PackedBatch:
    """Pack regions given as (haplotypes, reads) where each read is a 5-tuple of byte arrays
    (bases, quals, ins_gop, del_gop, gcp).  Pairs are region-major, read-major inside a region."""
    hap_chunks: List[np.ndarray] = []
    read_chunks: List[List[np.ndarray]] = [[], [], [], [], []]
    hap_len, read_len = [], []
    pair_read, pair_hap = [], []
    rpo, rro, rho = [0], [0], [0]
    n_r = n_h = 0
    for haps, reads in regions:
        for h in haps:
            a = _as_u8(h)
            hap_chunks.append(a)
            hap_len.append(len(a))
        for rd in reads:
            for k in range(5):
                read_chunks[k].append(_as_u8(rd[k]))
            read_len.append(len(rd[0]))
        A, N = len(haps), len(reads)
        pair_read.append(np.repeat(np.arange(n_r, n_r + N, dtype=np.int32), A))
        pair_hap.append(np.tile(np.arange(n_h, n_h + A, dtype=np.int32), N))
        n_r += N
        n_h += A
        rpo.append(rpo[-1] + N * A)
        rro.append(n_r)
        rho.append(n_h)
    hap_len = np.asarray(hap_len, dtype=np.int32)
    read_len = np.asarray(read_len, dtype=np.int32)
    cat = lambda xs: np.concatenate(xs) if xs else np.zeros(0, np.uint8)  # noqa: E731
    return PackedBatch(
        read_bases=cat(read_chunks[0]), read_quals=cat(read_chunks[1]), read_ins=cat(read_chunks[2]),
        read_del=cat(read_chunks[3]), read_gcp=cat(read_chunks[4]),
        read_off=np.concatenate([[0], np.cumsum(read_len[:-1], dtype=np.int64)]).astype(np.int64) if len(read_len) else np.zeros(0, np.int64),
        read_len=read_len,
        hap_bases=cat(hap_chunks),
        hap_off=np.concatenate([[0], np.cumsum(hap_len[:-1], dtype=np.int64)]).astype(np.int64) if len(hap_len) else np.zeros(0, np.int64),
        hap_len=hap_len,
        pair_read=np.concatenate(pair_read) if pair_read else np.zeros(0, np.int32),
        pair_hap=np.concatenate(pair_hap) if pair_hap else np.zeros(0, np.int32),
        region_pair_off=np.asarray(rpo, dtype=np.int64),
        region_read_off=np.asarray(rro, dtype=np.int64),
        region_hap_off=np.asarray(rho, dtype=np.int64),
    )


@dataclass
class RunStats:
    total_ms: float
    fp32_kernel_ms: float
    rescue_kernel_ms: float
    n_rescued: int           # pairs computed by the FP64 kernel
    cells: int
    n_routed_fp64: int = 0   # of n_rescued: routed by their read's qualities (the rest flagged by FP32)

    @property
    def gcups(self) -> float:
        return self.cells / (self.total_ms * 1e-3) / 1e9 if self.total_ms > 0 else float("nan")


# Engines still open when the interpreter exits (a failed test keeps its locals alive in the
# traceback) are closed by an atexit hook, while the HIP runtime is still up.
_LIVE_ENGINES: "weakref.WeakSet" = weakref.WeakSet()


@atexit.register
def _close_live_engines() -> None:
    for eng in list(_LIVE_ENGINES):
        try:
            eng.close()
        except Exception:
            pass


class DeviceBatch:
    """A PackedBatch resident in HBM (pairhmm_batch_upload) with its work queue."""

    def __init__(self, engine: "BatchEngine", batch: PackedBatch) -> None:
        self.engine = engine
        self.n_pairs = batch.n_pairs
        self._batch = batch  # host copy, for `cells` on demand (RunStats carries the count too)
        self._h = ctypes.c_void_p()
        d = batch.desc()
        _raise_for(engine.lib.pairhmm_batch_upload(engine.ctx, ctypes.byref(d), ctypes.byref(self._h)), "batch_upload")
        engine._batches.add(self)  # a batch is freed before its engine's context (pairhmm_batch_free reads it)

    @property
    def cells(self) -> int:
        """Sum over pairs of R * H (computed on the host when asked)."""
        return self._batch.cells()

    def forward(self, mode: int = _native.RUN_PERSISTENT, out: Optional[np.ndarray] = None) -> RunStats:
        st = _native.RunStats()
        ptr = None
        if out is not None:
            if out.dtype != np.float64 or not out.flags.c_contiguous or out.size < self.n_pairs:
                raise ValueError("out must be a contiguous float64 array with n_pairs entries")
            ptr = out.ctypes.data
        _raise_for(self.engine.lib.pairhmm_batch_forward(self.engine.ctx, self._h, int(mode), ptr, ctypes.byref(st)),
                   "batch_forward")
        return RunStats(st.total_ms, st.fp32_kernel_ms, st.rescue_kernel_ms, int(st.n_rescued), int(st.cells),
                        int(st.n_routed_fp64))

    def free(self) -> None:
        if self._h:
            self.engine.lib.pairhmm_batch_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class BatchEngine:
    """Owns one engine context (HIP stream + tables) for batch work."""

    def __init__(self, args: Optional[PairHMMNativeArguments] = None) -> None:
        self.lib = _native.load_library()
        if not self.lib.pairhmm_load():
            raise HardwareFeatureException("no gfx950 device visible to this process")
        a = args or PairHMMNativeArguments()
        na = _native.NativeArguments(int(a.max_number_of_threads), int(bool(a.use_double_precision)))
        self.ctx = ctypes.c_void_p()
        _raise_for(self.lib.pairhmm_initialize(ctypes.byref(na), ctypes.byref(self.ctx)), "initialize")
        self._batches: "weakref.WeakSet" = weakref.WeakSet()  # released before the context
        self._streams: "weakref.WeakSet" = weakref.WeakSet()
        _LIVE_ENGINES.add(self)

    def upload(self, batch: PackedBatch) -> DeviceBatch:
        return DeviceBatch(self, batch)

    def close(self) -> None:
        if self.ctx:
            for st in list(self._streams):
                st.close()
            for bt in list(self._batches):
                bt.free()
            self.lib.pairhmm_done(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
