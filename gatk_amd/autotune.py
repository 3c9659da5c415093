"""Autotuner of the batched forward launch schedule.

The persistent forward kernel has five schedule knobs (``pairhmm_tuning`` in
include/gatk_pairhmm.h): wavefronts per workgroup (which sets how the per-wave LDS slices pack into
a CU's 160 KB), the persistent grid as a fraction of resident capacity, the work-queue order
(descending cost vs submission order), the stripe schedule of the FP32 kernel: packed stripes
(the rows of consecutive pairs share a wave's 64-row stripes) or aligned stripes (one pair per
stripe, the read's last stripe padded to 64 rows), and the packed kernel's loop body (8 or 16
wave-steps per unrolled block, or per LDS class by occupancy).  ``tune`` uploads a sample of the workload, times
every candidate's eager pass (FP32 forward plus FP64 rescue, HIP events, best of ``repeats``),
re-times the fastest few on a larger sample, and returns the fastest.  Results are cached in a JSON
file keyed by device and a coarse workload signature (haplotype / read length classes), so later
runs on the same device and workload shape reuse them.
"""
from __future__ import annotations

import ctypes
import itertools
import json
import os
from dataclasses import asdict, dataclass
from typing import Dict, Iterable, List, Optional

import numpy as np

from . import _native
from .batch import BatchEngine, PackedBatch
from .pairhmm import _raise_for

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "gatk_amd", "autotune.json")


@dataclass(frozen=True)
class Schedule:
    waves_per_block: int = 4
    grid_percent: int = 100
    queue_order: int = 0
    stripe_schedule: int = 0  # 0 packed stripes, 1 aligned stripes
    step_unroll: int = 0  # packed stripes: 8 or 16 wave-steps per loop body, 0 per LDS class

    def apply(self, engine: BatchEngine) -> None:
        t = _native.Tuning(self.waves_per_block, self.grid_percent, self.queue_order, self.stripe_schedule,
                           self.step_unroll)
        _raise_for(engine.lib.pairhmm_set_tuning(engine.ctx, ctypes.byref(t)), "set_tuning")


# The loop body matters only to the packed kernel, so aligned candidates keep the default.
CANDIDATES: List[Schedule] = [Schedule(w, g, o, s, u) for w, g, o, s, u in
                              itertools.product((1, 2, 4, 8), (50, 100, 200), (0, 1), (0, 1), (0, 8, 16))
                              if s == 0 or u == 0]


def workload_signature(batch: PackedBatch) -> str:
    """Coarse shape of a workload: 64-column haplotype class of the 95th percentile and 63-row
    stripe count of the 95th-percentile read."""
    h95 = int(np.percentile(batch.hap_len, 95))
    r95 = int(np.percentile(batch.read_len, 95)) if len(batch.read_len) else 0
    return f"h{(h95 + 63) // 64 * 64}-s{(r95 + 62) // 63}"


def device_name() -> str:
    try:
        import torch

        return torch.cuda.get_device_name(0)
    except Exception:  # torch absent: one key per host
        return "gfx950"


def _load(path: str) -> Dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _time(eng: BatchEngine, sample: PackedBatch, cands: Iterable[Schedule], repeats: int, verbose: bool,
          tag: str):
    results = []
    for cand in cands:
        cand.apply(eng)
        dev = eng.upload(sample)
        best = float("inf")
        for _ in range(repeats):
            st = dev.forward(mode=_native.RUN_PERSISTENT_EAGER)
            best = min(best, st.total_ms)
        dev.free()
        results.append((best, cand))
        if verbose:
            print(f"[autotune] {tag} {cand}: {best:.2f} ms")
    results.sort(key=lambda x: x[0])
    return results


def tune(batch: PackedBatch, candidates: Iterable[Schedule] = CANDIDATES, sample_regions: int = 300,
         final_regions: int = 10_000, finalists: int = 4, repeats: int = 2,
         cache_path: Optional[str] = DEFAULT_CACHE, use_cache: bool = True, verbose: bool = False) -> Schedule:
    """Pick the fastest schedule for this workload shape (see module docstring).

    Two stages: every candidate is timed on the first ``sample_regions`` regions, then the
    ``finalists`` fastest, and the default schedule, are timed again on up to ``final_regions``
    regions, where the queue's tail and the FP64 rescue weigh as they do on a full batch (a small
    sample favours schedules whose advantage is only in the tail).  The time compared is the eager
    pass's FP32 forward plus FP64 rescue (HIP events around both kernels)."""
    key = f"{device_name()}|{workload_signature(batch)}"
    if use_cache and cache_path:
        hit = _load(cache_path).get(key)
        if hit:
            return Schedule(**hit["schedule"])
    n = min(sample_regions, batch.n_regions)
    sample = batch.subset_regions(range(n)) if n < batch.n_regions else batch
    with BatchEngine() as eng:
        results = _time(eng, sample, candidates, repeats, verbose, "stage 1")
        n2 = min(final_regions, batch.n_regions)
        if n2 > n and finalists > 0:
            final = batch.subset_regions(range(n2)) if n2 < batch.n_regions else batch
            final_set = [c for _, c in results[:finalists]]
            if Schedule() not in final_set:
                final_set.append(Schedule())
            results = _time(eng, final, final_set, repeats, verbose, "stage 2")
            n = n2
    best_ms, best = results[0]
    if cache_path:
        os.makedirs(os.path.dirname(cache_path), exist_ok=True)
        db = _load(cache_path)
        db[key] = {"schedule": asdict(best), "pass_ms": best_ms, "sample_regions": n,
                   "ranking": [[ms, asdict(c)] for ms, c in results[:5]]}
        tmp = cache_path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
        os.replace(tmp, cache_path)
    return best
