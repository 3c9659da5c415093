"""Streaming many active regions through the device (double-buffered host queue).

``RegionStream`` wraps the C-ABI stream (``pairhmm_stream_*``): ``submit`` validates and packs a
chunk of regions into pinned staging and returns while the device still works on the previous
chunk; results land in the caller's arrays when the slot is reused or at ``finish``.

``run_streaming`` drives a synthetic genome-scale run: chunks of regions are generated in worker
processes (seeded per chunk), submitted in order, and the whole run's cells, rescue count and wall
time are returned.
"""
from __future__ import annotations

import ctypes
import time
from collections import deque
from concurrent.futures import ProcessPoolExecutor
from dataclasses import dataclass
from typing import Deque, List, Optional

import numpy as np

from . import _native
from .batch import BatchEngine, PackedBatch
from .pairhmm import _raise_for


class RegionStream:
    def __init__(self, engine: BatchEngine) -> None:
        self.engine = engine
        self._h = ctypes.c_void_p()
        _raise_for(engine.lib.pairhmm_stream_create(engine.ctx, ctypes.byref(self._h)), "stream_create")
        self._pending: Deque[np.ndarray] = deque()

    def submit(self, chunk: PackedBatch, out: np.ndarray) -> None:
        if out.dtype != np.float64 or not out.flags.c_contiguous or out.size < chunk.n_pairs:
            raise ValueError("out must be a contiguous float64 array with n_pairs entries")
        d = chunk.desc()
        _raise_for(self.engine.lib.pairhmm_stream_submit(self._h, ctypes.byref(d), out.ctypes.data), "stream_submit")
        self._pending.append(out)  # filled two submits later or at finish
        while len(self._pending) > 2:
            self._pending.popleft()

    def finish(self) -> _native.RunStats:
        st = _native.RunStats()
        _raise_for(self.engine.lib.pairhmm_stream_finish(self._h, ctypes.byref(st)), "stream_finish")
        self._pending.clear()
        return st

    def close(self) -> None:
        if self._h:
            self.engine.lib.pairhmm_stream_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _make_chunk(args):
    from . import synth

    n, seed, str_fraction = args
    return synth.generate(synth.BATCHED_UNEVEN, n, seed, mix=(synth.PRECISION_STRESS, str_fraction))


@dataclass
class StreamingResult:
    regions: int
    pairs: int
    cells: int
    rescued: int
    wall_s: float
    gpu_busy_s: float

    @property
    def gcups(self) -> float:
        return self.cells / self.wall_s / 1e9


def run_streaming(n_regions: int, chunk_regions: int = 500, workers: int = 8, seed: int = 4_000_000,
                  str_fraction: float = 0.10, keep_outputs: bool = False, log=None):
    """Stream n_regions synthetic regions (90% batched_uneven-shaped, str_fraction STR-rich)."""
    n_chunks = (n_regions + chunk_regions - 1) // chunk_regions
    sizes = [min(chunk_regions, n_regions - k * chunk_regions) for k in range(n_chunks)]
    outs: List[np.ndarray] = []
    pairs = 0
    with BatchEngine() as eng, RegionStream(eng) as rs, ProcessPoolExecutor(max_workers=workers) as pool:
        futs = deque()
        nxt = 0
        while nxt < n_chunks and len(futs) < 2 * workers:
            futs.append(pool.submit(_make_chunk, (sizes[nxt], seed + nxt, str_fraction)))
            nxt += 1
        t0 = time.perf_counter()
        done = 0
        while futs:
            chunk = futs.popleft().result()
            if nxt < n_chunks:
                futs.append(pool.submit(_make_chunk, (sizes[nxt], seed + nxt, str_fraction)))
                nxt += 1
            out = np.empty(chunk.n_pairs)
            rs.submit(chunk, out)
            pairs += chunk.n_pairs
            if keep_outputs:
                outs.append(out)
            done += 1
            if log and done % max(1, n_chunks // 10) == 0:
                log(f"stream: {done}/{n_chunks} chunks submitted, {time.perf_counter() - t0:.1f} s")
        st = rs.finish()
        wall = time.perf_counter() - t0
    res = StreamingResult(regions=n_regions, pairs=pairs, cells=int(st.cells), rescued=int(st.n_rescued),
                          wall_s=wall, gpu_busy_s=st.total_ms / 1e3)
    return (res, outs) if keep_outputs else res
