"""PairHMM forward throughput on one MI355X (GCUPS), batched persistent kernel.

Workload: ``batched_uneven`` -- 10,000 synthetic active regions (haplotypes 2-128 per region,
150-600 bp; reads Poisson(80) per region, 100-250 bp, 1% substitutions, 0.1% indels; base Q 10-40,
GOP 45, GCP 10).  A step is one forward pass over every read x haplotype pair of all regions
(FP32 cells, FP64 rescue of underflowing pairs), replayed from a captured hipGraph, with the
inputs resident in HBM before timing starts; the results stay resident in HBM.

Prints one JSON line (see the keys below).  ``roofline`` is the FP32 forward kernel measured with
HIP events on the engine's stream; ``cpu_baseline`` times the FP64 CPU port on a bounded sample
of the same pairs on this host's cores and doubles as the accuracy check of that sample.
"""
from __future__ import annotations

import argparse
import faulthandler
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_CELL = 12          # M: 6, I: 3, D: 3 (LoglessPairHMM.java:51-55)
FP32_PEAK_TFLOPS = 157.3    # MI355X vector FP32 (published)
# HBM bytes per cell of the forward pass: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE over the packed
# FP32 forward kernel's launches on 1,000 batched_uneven regions, each scaled by its reading for a
# 1 GiB elementwise copy (x2.0 and x1.0); the rocprof summary is in profiles/.
HBM_BYTES_PER_CELL = 0.0055


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(batch, out, seconds: float, threads: int):
    """Time the FP64 oracle on random pairs of the workload until ~`seconds` of wall time."""
    from oracle import oracle

    rng = np.random.default_rng(1234)
    order = rng.permutation(batch.n_pairs)
    done, cells, t_used, worst = 0, 0, 0.0, 0.0
    chunk = 256
    while t_used < seconds and done < len(order):
        idx = np.sort(order[done:done + chunk])
        t0 = time.perf_counter()
        ref = oracle.logless_batch(batch, n_threads=threads, pairs=idx)
        t_used += time.perf_counter() - t0
        cells += batch.cells(idx)
        fin = np.isfinite(ref)
        if out is not None and fin.any():
            worst = max(worst, float(np.max(np.abs(out[idx][fin] - ref[fin]))))
        done += len(idx)
        chunk = min(chunk * 2, 8192)
    return dict(value=cells / t_used / 1e9, unit="GCUPS", cores=threads, kind="port",
                sample=f"{done} random pairs of the timed workload ({cells:.3e} cells), FP64 LoglessPairHMM port "
                       f"(oracle/pairhmm_oracle.c), OpenMP over pairs, {t_used:.1f} s"), worst, done


def host_cpu() -> str:
    """The host CPU model and logical CPU count (from /proc/cpuinfo)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}, {os.cpu_count()} logical CPUs visible"


def cpu_proxy_baseline(batch, out, seconds: float, threads: int):
    """Time the GKL-class CPU proxy (oracle/simd_proxy.c: FP32 SIMD lanes over a haplotype's reads, FP64
    recompute under 1e-28) on whole regions of the workload, in random order, until ~`seconds`."""
    from oracle import oracle

    rng = np.random.default_rng(4321)
    regions = rng.permutation(batch.n_regions)
    done, cells, t_used, worst, rescued = 0, 0, 0.0, 0.0, 0
    step = 8
    while t_used < seconds and done < len(regions):
        sel = np.sort(regions[done:done + step])
        idx = np.concatenate([np.arange(batch.region_pair_off[k], batch.region_pair_off[k + 1]) for k in sel])
        t0 = time.perf_counter()
        got, n = oracle.simd_proxy_batch(batch, n_threads=threads, pairs=idx)
        t_used += time.perf_counter() - t0
        rescued += n
        cells += batch.cells(idx)
        fin = np.isfinite(got)
        if out is not None and fin.any():
            worst = max(worst, float(np.max(np.abs(out[idx][fin] - got[fin]))))
        done += len(sel)
        step = min(step * 2, 256)
    return dict(value=round(cells / t_used / 1e9, 3), unit="GCUPS", cores=threads, kind="port",
                sample=f"{done} random whole regions of the timed workload ({cells:.3e} cells), FP32 SIMD "
                       f"(AVX-512 where the host has it) with FP64 recompute of {rescued} pairs, flush-to-zero, "
                       f"OpenMP over haplotypes, {t_used:.1f} s",
                max_abs_dlog10L_vs_gpu=worst)


def fwd_bwd_leg(args):
    """Forward + backward GCUPS on a batched_uneven batch and forward GCUPS on the same pairs."""
    import torch

    from gatk_amd import _native, synth
    from gatk_amd.autograd import GradEngine, GradProblem
    from gatk_amd.batch import BatchEngine

    gb = synth.workload("batched_uneven", n_regions=args.bwd_regions, seed=3)
    cells = gb.cells()
    prob = GradProblem.from_batch(gb)
    n = len(gb.read_bases)
    qb = torch.from_numpy(gb.read_quals.astype(np.float32)).cuda()
    qi = torch.full((n,), 45.0, dtype=torch.float32, device="cuda")
    qd = torch.full((n,), 45.0, dtype=torch.float32, device="cuda")
    qc = torch.full((n,), 10.0, dtype=torch.float32, device="cuda")
    with GradEngine() as eng:
        out, grads = eng.run(prob, qb, qi, qd, qc)  # warm-up
        t0 = time.perf_counter()
        for _ in range(args.bwd_steps):
            out, grads = eng.run(prob, qb, qi, qd, qc)
        el = time.perf_counter() - t0
        st = eng.last_stats
    gcups = cells * args.bwd_steps / el / 1e9
    with BatchEngine() as beng:
        dev = beng.upload(gb)
        dev.forward()
        t0 = time.perf_counter()
        for _ in range(args.bwd_steps):
            dev.forward()
        fel = time.perf_counter() - t0
        dev.free()
    fwd = cells * args.bwd_steps / fel / 1e9
    log(f"fwd+bwd: {gcups:.1f} GCUPS ({el / args.bwd_steps * 1e3:.0f} ms/step, fp32 kernel {st.fp32_kernel_ms:.0f} ms, "
        f"fp64 {st.rescue_kernel_ms:.0f} ms, {st.n_rescued} pairs in FP64); forward on the same pairs {fwd:.1f} GCUPS")
    res = {"gcups": round(gcups, 2), "fwd_gcups_same_pairs": round(fwd, 2), "ratio_to_fwd": round(gcups / fwd, 3),
           "workload": f"batched_uneven {args.bwd_regions} regions (seed 3), qb = base quals, qi = qd = 45, qc = 10",
           "pairs": gb.n_pairs, "cells": cells, "fp64_pairs": int(st.n_rescued)}
    ctx = dict(batch=gb, out=out.cpu().numpy(), grads={k: v.cpu().numpy() for k, v in grads.items()},
               q=(gb.read_quals.astype(np.float32), np.full(n, 45.0, np.float32), np.full(n, 45.0, np.float32),
                  np.full(n, 10.0, np.float32)))
    return res, ctx


def cpu_grad_baseline(ctx, threads, seconds=10.0):
    """FP64 adjoint port (full matrices) on every pair of randomly sampled reads, read after read until
    about `seconds` of wall time: CPU fwd+bwd GCUPS, and the accuracy of the GPU's per-read gradient
    sums and log10 L for those reads."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle

    b = ctx["batch"]
    rng = np.random.default_rng(7)
    order = rng.permutation(len(b.read_len))

    def one(p):
        r, h = int(b.pair_read[p]), int(b.pair_hap[p])
        o, R = int(b.read_off[r]), int(b.read_len[r])
        ho, H = int(b.hap_off[h]), int(b.hap_len[h])
        qs = [x[o:o + R].astype(np.float64) for x in ctx["q"]]
        return oracle.logless_float_grad(b.hap_bases[ho:ho + H], b.read_bases[o:o + R], *qs), R * H

    pairs_of = {}
    for p, r in enumerate(b.pair_read):
        pairs_of.setdefault(int(r), []).append(p)
    el, cells, n_pairs, worst_l, worst_g, done = 0.0, 0, 0, 0.0, 0.0, 0
    with ThreadPoolExecutor(max_workers=threads) as ex:
        while el < seconds and done < len(order):
            reads = order[done:done + 2 * threads]
            done += len(reads)
            pairs = [p for r in reads for p in pairs_of.get(int(r), [])]
            t0 = time.perf_counter()
            res = list(ex.map(one, pairs))
            el += time.perf_counter() - t0
            cells += sum(c for _, c in res)
            n_pairs += len(pairs)
            worst_l = max([worst_l] + [abs(ctx["out"][p] - g["log10_l"]) for p, (g, _) in zip(pairs, res)])
            for r in reads:
                o, R = int(b.read_off[r]), int(b.read_len[r])
                want = {k: np.zeros(R) for k in ("qb", "qi", "qd", "qc")}
                for p, (g, _) in zip(pairs, res):
                    if b.pair_read[p] == r:
                        for k in want:
                            want[k] += g[k]
                a = np.concatenate([ctx["grads"][k][o:o + R] for k in want])
                w = np.concatenate([want[k] for k in want])
                worst_g = max(worst_g, float(np.linalg.norm(a - w) / max(np.linalg.norm(w), 1e-300)))
    return {"cpu_fwd_bwd_baseline": {"value": cells / el / 1e9, "unit": "GCUPS", "cores": threads, "kind": "port",
                                     "sample": f"{n_pairs} pairs of {done} random reads, FP64 adjoint port "
                                               f"(oracle/pairhmm_oracle.c, full matrices), {el:.1f} s"},
            "max_abs_dlog10L_sample": worst_l, "grad_rel_err_sample": worst_g, "grad_checked_reads": done}


def reference_cpu_paths() -> dict:
    """Whether the reference's own CPU paths can run on this host: Java LoglessPairHMM needs a JVM,
    GKL's AVX_LOGLESS_CACHING(_OMP) needs the GKL jar with its native library.  The image ships
    neither (no network to fetch them), so both are normally reported unavailable."""
    import glob
    import shutil

    java = shutil.which("java")
    jars = []
    for pat in ("/usr/share/java/gkl*.jar", "/opt/gkl/**/gkl*.jar",
                os.path.expanduser("~/.m2/repository/com/intel/gkl/**/gkl*.jar"),
                os.path.expanduser("~/.gradle/caches/**/com.intel.gkl/**/gkl*.jar")):
        jars += glob.glob(pat, recursive=True)
    return {"java_logless_caching": "available" if java else "unavailable (no java on this host)",
            "gkl_avx_logless_caching_omp": ("available: " + jars[0]) if (java and jars) else
            "unavailable (no GKL jar" + ("" if java else ", no JVM") + " on this host)"}


def streaming_leg(args):
    """Regions streamed through the double-buffered stream API in chunks generated by worker processes."""
    from gatk_amd.stream import run_streaming

    workers = min(16, os.cpu_count() or 1)
    res = run_streaming(args.stream_regions, chunk_regions=args.stream_chunk, workers=workers, log=log)
    d = {"regions": res.regions, "pairs": res.pairs, "cells": res.cells, "wall_s": round(res.wall_s, 2),
         "gcups": round(res.gcups, 1), "rescued_pairs": res.rescued, "chunk_regions": args.stream_chunk,
         "generator_workers": workers,
         "workload": "90% batched_uneven-shaped regions, 10% precision_stress (STR-rich) regions, "
                     "double-buffered host queue, wall time includes chunk hand-over, packing and H2D/D2H"}
    log(f"streaming: {d}")
    return d


def readprep_leg(args):
    """The input side on the device over the reads of the batched_uneven workload: DRAGstr GOP / GCP
    imputation, then the PCR indel error model (CONSERVATIVE) and the minimum-quality caps.  Device
    time per step (inputs resident); CPU baselines: the host C++ DRAGstr imputer (one thread) on all
    reads, and the read-prep oracle (PCR + caps, pure Python) on a sample."""
    import torch

    from gatk_amd import synth
    from gatk_amd.dragstr import DragstrPairHMMInputScoreImputator, DragstrParams
    from gatk_amd.readprep import PCRErrorModel, ReadPrep
    from oracle import readprep_oracle as rp

    b = synth.workload("batched_uneven", n_regions=args.readprep_regions, seed=1)
    n_bases, n_reads = len(b.read_bases), len(b.read_len)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    rng = np.random.default_rng(8)
    mapq = rng.integers(20, 61, n_reads).astype(np.int32)
    params = DragstrParams.default()
    d_b, d_o, d_l, d_mq = t(b.read_bases), t(b.read_off), t(b.read_len.astype(np.int32)), t(mapq)
    gop = torch.empty(n_bases, dtype=torch.uint8, device="cuda")
    gcp = torch.empty(n_bases, dtype=torch.uint8, device="cuda")
    q, i, d = t(b.read_quals.copy()), t(b.read_ins.copy()), t(b.read_del.copy())
    with ReadPrep() as prep:
        prep.dragstr_impute(params, d_b, d_o, d_l, gop, gcp)  # warm-up
        prep.modify_read_qualities(d_b, d_o, d_l, q, i, d, PCRErrorModel.CONSERVATIVE, mapping_quality=d_mq)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        prep.dragstr_impute(params, d_b, d_o, d_l, gop, gcp)
        t1 = time.perf_counter()
        prep.modify_read_qualities(d_b, d_o, d_l, q, i, d, PCRErrorModel.CONSERVATIVE, mapping_quality=d_mq)
        t2 = time.perf_counter()
    imp = DragstrPairHMMInputScoreImputator(params)
    c0 = time.perf_counter()
    want_gop, want_gcp = imp.impute_batch(b)
    c1 = time.perf_counter()
    same = bool(np.array_equal(gop.cpu().numpy(), want_gop) and np.array_equal(gcp.cpu().numpy(), want_gcp))
    sample = rng.choice(n_reads, size=min(args.readprep_sample, n_reads), replace=False)
    s_bases = int(b.read_len[sample].sum())
    c2 = time.perf_counter()
    for r in sample:
        o, L = int(b.read_off[r]), int(b.read_len[r])
        wi, wd = rp.apply_pcr_error_model(bytes(b.read_bases[o:o + L]), b.read_ins[o:o + L], b.read_del[o:o + L],
                                          "CONSERVATIVE")
        rp.cap_minimum_read_qualities(b.read_quals[o:o + L], wi, wd, int(mapq[r]), 18, False)
    c3 = time.perf_counter()
    d = {"reads": n_reads, "bases": n_bases,
         "dragstr_gbases_per_s": round(n_bases / (t1 - t0) / 1e9, 3),
         "pcr_and_caps_gbases_per_s": round(n_bases / (t2 - t1) / 1e9, 3),
         "dragstr_matches_host": same,
         "cpu_dragstr_host_cpp_gbases_per_s": round(n_bases / (c1 - c0) / 1e9, 4),
         "cpu_pcr_and_caps_oracle_gbases_per_s": round(s_bases / (c3 - c2) / 1e9, 6),
         "cpu_sample": f"host C++ DRAGstr imputer on all reads (1 thread, {c1 - c0:.2f} s); read-prep oracle "
                       f"(pure Python) on {len(sample)} reads ({c3 - c2:.2f} s)",
         "workload": f"reads of batched_uneven {args.readprep_regions} regions (seed 1), default DRAGstr tables, "
                     f"PCR model CONSERVATIVE, mapping quality 20-60, base-quality threshold 18"}
    log(f"read prep: {d}")
    return d


def pd_regions(n_regions, seed=21):
    """batched_uneven regions whose haplotypes are partially determined: per haplotype 3 SNP columns
    (1-3 alternative bases) and one optional deletion span of 1-8 bases, PD bytes laid out as
    PartiallyDeterminedHaplotype.getPDBytesForHaplotypes does."""
    from gatk_amd import synth
    from gatk_amd.pdhmm import DEL_END, DEL_START, SNP

    b = synth.workload("batched_uneven", n_regions=n_regions, seed=seed)
    rng = np.random.default_rng(seed)
    pd = np.zeros(len(b.hap_bases), np.uint8)
    for h in range(len(b.hap_len)):
        o, H = int(b.hap_off[h]), int(b.hap_len[h])
        for j in rng.choice(H, 3, replace=False):
            pd[o + j] = SNP | (int(rng.integers(1, 16)) << 3)
        s = int(rng.integers(0, H - 8))
        pd[o + s] |= DEL_START
        pd[o + s + int(rng.integers(0, 8))] |= DEL_END
    return b, pd


def pd_leg(args, threads):
    """Partially determined PairHMM through its drop-in (pairhmm_pd_compute_likelihoods), one call per
    active region as VectorLoglessPairPDHMM makes them: wall rate per call including the host->device
    staging, and the FP64 oracle on `threads` host threads over a sample of the same regions."""
    from gatk_amd.pairhmm import ReadDataHolder
    from gatk_amd.pdhmm import HaplotypeDataHolder, HipPDHmm
    from oracle import oracle

    b, pd = pd_regions(args.pd_regions)
    regions = []
    for k in range(len(b.region_read_off) - 1):
        r0, r1 = int(b.region_read_off[k]), int(b.region_read_off[k + 1])
        h0, h1 = int(b.region_hap_off[k]), int(b.region_hap_off[k + 1])
        reads = []
        for r in range(r0, r1):
            o, L = int(b.read_off[r]), int(b.read_len[r])
            reads.append(tuple(np.ascontiguousarray(x[o:o + L]) for x in
                               (b.read_bases, b.read_quals, b.read_ins, b.read_del, b.read_gcp)))
        haps = [(bytes(b.hap_bases[int(b.hap_off[h]):int(b.hap_off[h]) + int(b.hap_len[h])]),
                 bytes(pd[int(b.hap_off[h]):int(b.hap_off[h]) + int(b.hap_len[h])])) for h in range(h0, h1)]
        cells = int(sum(len(r[0]) for r in reads)) * int(sum(len(h) for h, _ in haps))
        regions.append((reads, haps, cells))
    hmm = HipPDHmm()
    hmm.load()
    hmm.initialize()
    holders = [([ReadDataHolder(*r) for r in reads],
                [HaplotypeDataHolder(np.frombuffer(h, np.uint8), np.frombuffer(p, np.uint8)) for h, p in haps],
                np.empty(len(reads) * len(haps))) for reads, haps, _ in regions]
    for rh, hh, out in holders[:4]:
        hmm.compute_likelihoods(rh, hh, out)  # warm-up
    t0 = time.perf_counter()
    for rh, hh, out in holders:
        hmm.compute_likelihoods(rh, hh, out)
    wall = time.perf_counter() - t0
    hmm.done()
    cells = sum(c for _, _, c in regions)
    rng = np.random.default_rng(3)
    sample = rng.choice(len(regions), size=min(args.pd_sample, len(regions)), replace=False)
    s_cells, worst = 0, 0.0
    c0 = time.perf_counter()
    for k in sample:
        reads, haps, c = regions[k]
        want = oracle.pd_logless_batch(reads, haps, n_threads=threads).ravel()
        worst = max(worst, float(np.max(np.abs(want - holders[k][2]))))
        s_cells += c
    c1 = time.perf_counter()
    d = {"regions": len(regions), "pairs": int(sum(len(r) * len(h) for r, h, _ in regions)), "cells": cells,
         "dropin_wall_gcups": round(cells / wall / 1e9, 2), "wall_s": round(wall, 3),
         "max_abs_dlog10L_vs_oracle_sample": worst,
         "cpu_oracle_gcups": round(s_cells / (c1 - c0) / 1e9, 4),
         "cpu_sample": f"FP64 PD oracle on {threads} threads over {len(sample)} regions ({c1 - c0:.1f} s)",
         "dtype": "f64",
         "workload": f"batched_uneven {args.pd_regions} regions (seed 21), each haplotype with 3 SNP columns and "
                     f"one 1-8 base deletion span; one computeLikelihoods call per region"}
    log(f"pd-hmm: {d}")
    return d


def postprocess_leg(args):
    """Likelihood post-processing on the device over one read-major matrix resident in HBM (reads x
    haplotypes, 150-base reads, DRAGEN dynamic threshold on): device time per call and the algorithmic
    bytes it moves; the oracle (pure Python) on a sample of reads as the CPU baseline."""
    import torch

    from gatk_amd.postprocess import LikelihoodPostProcessor, log10_mismapping_cap
    from oracle import postprocess_oracle as pp

    n, A, L = args.post_reads, 8, 150
    rng = np.random.default_rng(5)
    lk = -rng.exponential(15.0, (n, A))
    quals = rng.integers(2, 41, n * L).astype(np.uint8)
    lens = np.full(n, L, np.int32)
    off = (np.arange(n, dtype=np.int64) * L)
    cap = log10_mismapping_cap(45)
    d_lk = torch.from_numpy(lk).cuda()
    d_q, d_len, d_off = (torch.from_numpy(x).cuda() for x in (quals, lens, off))
    keep = torch.empty(n, dtype=torch.uint8, device="cuda")
    th = torch.empty(n, dtype=torch.float64, device="cuda")
    reps = 20
    with LikelihoodPostProcessor() as p:
        p.normalize_and_filter(d_lk, d_len, 0, cap, False, d_q, d_off, 0.02, True, 1.0, keep, th)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):  # idempotent: a second cap leaves every value in place
            p.normalize_and_filter(d_lk, d_len, 0, cap, False, d_q, d_off, 0.02, True, 1.0, keep, th)
        torch.cuda.synchronize()
        per_call = (time.perf_counter() - t0) / reps
    nbytes = n * (2 * 8 * A + L + 4 + 8 + 1 + 8)
    sample = rng.choice(n, size=min(args.post_sample, n), replace=False)
    got_lk, got_keep = d_lk.cpu().numpy(), keep.cpu().numpy().astype(bool)
    c0 = time.perf_counter()
    norm, want_keep, _ = pp.postprocess(lk[sample], [quals[off[r]:off[r] + L] for r in sample], cap, False, 0, True,
                                        0.02, 1.0)
    c1 = time.perf_counter()
    same = bool(np.array_equal(norm, got_lk[sample]) and np.array_equal(want_keep, got_keep[sample]))
    d = {"reads": n, "alleles": A, "read_len": L, "ms_per_call": round(per_call * 1e3, 3),
         "mreads_per_s": round(n / per_call / 1e6, 1),
         "algorithmic_gb_per_s": round(nbytes / per_call / 1e9, 1), "matches_oracle_sample": same,
         "cpu_oracle_mreads_per_s": round(len(sample) / (c1 - c0) / 1e6, 5),
         "cpu_sample": f"oracle (pure Python, 1 thread) on {len(sample)} reads ({c1 - c0:.2f} s)",
         "note": "wall time per call on the context's stream (synchronous entry point), inputs in HBM"}
    log(f"post-processing: {d}")
    return d


def sw_leg(args):
    """Read -> haplotype Smith-Waterman realignment (HaplotypeCaller's step after the PairHMM) through
    the batch API: every read of batched_uneven regions against a haplotype of its region, with
    ALIGNMENT_TO_BEST_HAPLOTYPE_SW_PARAMETERS and SOFTCLIP, one call (host in / host out); the oracle
    (one thread, C) on a sample as the CPU baseline, CIGARs compared on that sample."""
    from gatk_amd import synth
    from gatk_amd.smithwaterman import (ALIGNMENT_TO_BEST_HAPLOTYPE_SW_PARAMETERS as P, HipSmithWatermanBinding,
                                        SWOverhangStrategy)
    from oracle import oracle

    b = synth.workload("batched_uneven", n_regions=args.sw_regions, seed=31)
    rng = np.random.default_rng(31)
    refs, alts = [], []
    for k in range(len(b.region_read_off) - 1):
        h0, h1 = int(b.region_hap_off[k]), int(b.region_hap_off[k + 1])
        for r in range(int(b.region_read_off[k]), int(b.region_read_off[k + 1])):
            h = int(rng.integers(h0, h1))
            refs.append(b.hap_bases[int(b.hap_off[h]):int(b.hap_off[h]) + int(b.hap_len[h])].tobytes())
            alts.append(b.read_bases[int(b.read_off[r]):int(b.read_off[r]) + int(b.read_len[r])].tobytes())
    cells = sum(len(r) * len(a) for r, a in zip(refs, alts))
    sw = HipSmithWatermanBinding()
    sw.load()
    # Warm-up at the batch's own size: the first call of a size grows the context's pinned and
    # device staging (about 0.2 s of allocation), which a steady-state caller pays once.
    sw.align_batch(refs, alts, P, SWOverhangStrategy.SOFTCLIP)
    t0 = time.perf_counter()
    got = sw.align_batch(refs, alts, P, SWOverhangStrategy.SOFTCLIP)
    wall = time.perf_counter() - t0
    sw.close()
    sample = rng.choice(len(refs), size=min(args.sw_sample, len(refs)), replace=False)
    pp = (P.match_value, P.mismatch_penalty, P.gap_open_penalty, P.gap_extend_penalty)
    s_cells, same = 0, True
    c0 = time.perf_counter()
    for k in sample:
        same &= oracle.sw_align(refs[k], alts[k], pp, "SOFTCLIP") == got[k]
        s_cells += len(refs[k]) * len(alts[k])
    c1 = time.perf_counter()
    d = {"alignments": len(refs), "cells": cells, "wall_s": round(wall, 3),
         "alignments_per_s": round(len(refs) / wall, 1), "wall_gcups": round(cells / wall / 1e9, 2),
         "matches_oracle_sample": bool(same),
         "cpu_oracle_gcups": round(s_cells / (c1 - c0) / 1e9, 4),
         "cpu_sample": f"SW oracle (C, 1 thread) on {len(sample)} alignments ({c1 - c0:.2f} s)",
         "workload": f"every read of batched_uneven {args.sw_regions} regions (seed 31) against a random haplotype of "
                     f"its region, parameters (10, -15, -30, -5), SOFTCLIP; one batch call (packed CIGAR text), "
                     f"host in / host out, after one untimed call of the same size"}
    log(f"smith-waterman: {d}")
    return d


def chunked_leg(args, chunk_regions=500):
    """The streaming workload's region mix (90% batched_uneven-shaped, 10% STR-rich precision_stress
    regions) in 500-region chunks through the batch API (upload + persistent graph + download per
    chunk).  Two contexts alternate: a worker thread uploads chunk k+1 into one while chunk k runs on
    the other, so the host-side upload overlaps the device (contexts may be used from different
    threads; each is used by one thread at a time).  Chunks are generated before the timed region.
    GPU-side GCUPS counts the forward launches only; the wall figure runs from the first upload to the
    last download."""
    from concurrent.futures import ThreadPoolExecutor

    from gatk_amd import _native
    from gatk_amd.batch import BatchEngine
    from gatk_amd.stream import _make_chunk

    n = args.chunked_regions
    n_gen = (n + chunk_regions - 1) // chunk_regions
    g0 = time.perf_counter()
    distinct = [_make_chunk((min(chunk_regions, n - k * chunk_regions), 4_000_000 + k, 0.10)) for k in range(n_gen)]
    gen = time.perf_counter() - g0
    # --chunked-passes > 1 runs the generated chunks again (every pass a fresh upload and forward):
    # a larger streamed workload without generating more regions.
    chunks = distinct * max(1, args.chunked_passes)
    n_chunks = len(chunks)
    cells = pairs = rescued = 0
    gpu_ms = 0.0
    engines = [BatchEngine(), BatchEngine()]
    with ThreadPoolExecutor(max_workers=1) as ex:
        t0 = time.perf_counter()
        fut = ex.submit(engines[0].upload, chunks[0])
        for k in range(n_chunks):
            dev = fut.result()
            if k + 1 < n_chunks:  # the other context finished chunk k-1 before this iteration
                fut = ex.submit(engines[(k + 1) % 2].upload, chunks[k + 1])
            out = np.empty(chunks[k].n_pairs)
            st = dev.forward(mode=_native.RUN_PERSISTENT, out=out)
            dev.free()
            if k == 0:
                first_out = out
            gpu_ms += st.total_ms
            cells += st.cells
            pairs += chunks[k].n_pairs
            rescued += st.n_rescued
            if (k + 1) % 5 == 0:
                log(f"chunked: {k + 1} chunks, {pairs} pairs, GPU {gpu_ms / 1e3:.2f} s")
        wall = time.perf_counter() - t0
    # The first chunk again, alone on one context: the pipelined results must match bit for bit.
    again = np.empty(chunks[0].n_pairs)
    dev = engines[0].upload(chunks[0])
    dev.forward(mode=_native.RUN_PERSISTENT_EAGER, out=again)
    dev.free()
    same = bool(np.array_equal(first_out, again, equal_nan=True))
    for e in engines:
        e.close()
    d = {"regions": n * max(1, args.chunked_passes), "distinct_regions": n, "passes": max(1, args.chunked_passes),
         "chunk_regions": chunk_regions, "pairs": pairs, "cells": cells, "rescued_pairs": rescued,
         "gpu_gcups": round(cells / (gpu_ms * 1e-3) / 1e9, 1), "wall_gcups_with_transfers": round(cells / wall / 1e9, 1),
         "generation_s": round(gen, 1), "first_chunk_matches_serial_run": same,
         "workload": "streaming mix: 90% batched_uneven-shaped, 10% precision_stress regions; batch API per "
                     "500-region chunk on two contexts, the next chunk's upload overlapping the current forward"}
    log(f"chunked streaming workload: {d}")
    return d


def precision_leg(args, threads):
    """Precision stress: STR-rich regions with haplotypes of 800-1,500 bp, reads of 250-600 bp, base
    Q 2-15 and 5% substitutions, so that many pairs fall below the FP32 underflow threshold.  Reports
    the rescue fraction, the FP32 kernel's and the FP64 rescue's time and throughput, and |d log10 L|
    against the FP64 oracle on a sample of FP32-kept and of rescued pairs."""
    from gatk_amd import _native, synth
    from gatk_amd.batch import BatchEngine
    from oracle import oracle

    b = synth.workload("precision_stress", n_regions=args.precision_regions, seed=2)
    cells = b.cells()
    log(f"precision stress: {args.precision_regions} regions, {b.n_pairs} pairs, {cells:.4e} cells")
    with BatchEngine() as eng:
        dev = eng.upload(b)
        out = np.empty(b.n_pairs)
        dev.forward(out=out)  # graph capture + warm-up
        t0 = time.perf_counter()
        st = dev.forward(out=out)
        wall = time.perf_counter() - t0
        ste = dev.forward(mode=_native.RUN_PERSISTENT_EAGER, out=out)
        dev.free()
    # FP64 pairs: routed by their read's qualities before any FP32 pass, or flagged by the FP32 pass
    # (scaled sum L * 2^120 below 1e-28).  A routed pair is a false positive of the routing predicate
    # when its FP64 result shows the FP32 pass would have kept it (L * 2^120 >= 1e-28).
    from gatk_amd.batch import reads_routed_to_fp64

    routed = reads_routed_to_fp64(b)
    below = out + 120 * np.log10(2.0) < -28.0
    rescued = routed | below
    r_idx, k_idx = np.nonzero(rescued)[0], np.nonzero(~rescued & np.isfinite(out))[0]
    r_cells = b.cells(r_idx)
    routed_fp = int(np.count_nonzero(routed & ~below))
    rng = np.random.default_rng(5)
    pick = lambda idx, n: np.sort(rng.choice(idx, size=min(n, len(idx)), replace=False))  # noqa: E731
    sk, sr = pick(k_idx, args.precision_sample), pick(r_idx, args.precision_sample)
    t1 = time.perf_counter()
    ref = oracle.logless_batch(b, n_threads=threads, pairs=np.concatenate([sk, sr]))
    el = time.perf_counter() - t1
    err_k = np.abs(out[sk] - ref[:len(sk)])
    err_r = np.abs(out[sr] - ref[len(sk):])
    hist_edges = [0, 1e-9, 1e-8, 1e-7, 1e-6, 1e-5, 1e-4, np.inf]
    hist = lambda e: np.histogram(e, bins=hist_edges)[0].tolist()  # noqa: E731
    d = {"regions": args.precision_regions, "pairs": b.n_pairs, "cells": cells,
         "rescued_pairs": int(ste.n_rescued), "rescue_fraction": round(ste.n_rescued / b.n_pairs, 4),
         "fp64_routed_pairs": int(ste.n_routed_fp64), "fp64_flagged_by_fp32_pairs": int(ste.n_rescued - ste.n_routed_fp64),
         "routed_pairs_fp32_would_keep": routed_fp,
         "routed_predicate_agrees_with_host": bool(int(routed.sum()) == int(ste.n_routed_fp64)),
         "rescued_cells": int(r_cells), "gcups": round(cells / (st.total_ms * 1e-3) / 1e9, 1),
         "wall_gcups": round(cells / wall / 1e9, 1),
         "fp32_kernel_ms": round(ste.fp32_kernel_ms, 1), "fp64_rescue_ms": round(ste.rescue_kernel_ms, 1),
         "fp64_rescue_gcups": round(r_cells / (ste.rescue_kernel_ms * 1e-3) / 1e9, 1) if ste.rescue_kernel_ms > 0 else None,
         "max_abs_dlog10L_fp32_kept": float(err_k.max()) if len(err_k) else None,
         "max_abs_dlog10L_fp64_rescued": float(err_r.max()) if len(err_r) else None,
         "error_histogram_edges": ["0", "1e-9", "1e-8", "1e-7", "1e-6", "1e-5", "1e-4", "inf"],
         "error_histogram_fp32_kept": hist(err_k), "error_histogram_fp64_rescued": hist(err_r),
         "oracle_sample": f"{len(sk)} kept + {len(sr)} rescued random pairs, FP64 oracle on {threads} threads, {el:.1f} s",
         "workload": "precision_stress (seed 2): STR-rich haplotypes 800-1,500 bp, reads 250-600 bp, base Q 2-15, "
                     "5% substitutions, batched_uneven haplotype/read counts"}
    log(f"precision stress: {d}")
    return d


def str_fit_leg(args):
    """Per-(STR period, repeats) GOP / GCP fit with the adjoint kernel against known penalties."""
    from gatk_amd import strfit

    spec = strfit.FitSpec(n_regions=args.fit_regions)
    t0 = time.perf_counter()
    wl, res, rep = strfit.run_gpu_fit(spec, max_iter=args.fit_iters, log=log)
    total = time.perf_counter() - t0
    b = wl.batch
    d = {"regions": spec.n_regions, "reads": int(len(b.read_len)), "cells_per_evaluation": res.cells_per_eval,
         "evaluations": res.evaluations, "fit_seconds": round(res.seconds, 2), "total_seconds": round(total, 2),
         "fwd_bwd_gcups_in_fit": round(res.cells_per_eval * res.evaluations / res.seconds / 1e9, 1),
         "final_neg_mean_log10L": round(res.loss, 6), "recovered": rep.ok, **rep.as_dict()}
    log(f"STR fit: {d}")
    return d


def main() -> int:
    faulthandler.enable()  # a fatal signal prints every thread's Python stack to stderr
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--regions", type=int, default=10_000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--plain-regions", type=int, default=100)
    ap.add_argument("--small-calls", type=int, default=20,
                    help="drop-in calls timed on plain_small (64 reads x 16 haplotypes); 0 skips that leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--bwd-regions", type=int, default=2_000,
                    help="regions of the forward+backward leg (0 skips it)")
    ap.add_argument("--bwd-steps", type=int, default=2)
    ap.add_argument("--fit-regions", type=int, default=1_000,
                    help="regions of the STR penalty fit leg (100 reads each; 0 skips it)")
    ap.add_argument("--fit-iters", type=int, default=200)
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-schedule autotuner")
    ap.add_argument("--stripes", choices=("packed", "aligned"), default=None,
                    help="force the FP32 stripe schedule of the headline leg (with --no-tune)")
    ap.add_argument("--stream-regions", type=int, default=0,
                    help="regions of the double-buffered stream-API leg with a generator process pool "
                         "(off by default; see the streaming notes of the design document)")
    ap.add_argument("--stream-chunk", type=int, default=500)
    ap.add_argument("--chunked-regions", type=int, default=10_000,
                    help="regions of the chunked streaming-workload leg (90%% batched_uneven, 10%% STR-rich; "
                         "0 skips it)")
    ap.add_argument("--chunked-passes", type=int, default=20,
                    help="times the chunked leg runs its generated chunks (each pass uploads and computes every "
                         "chunk again; 20 passes of 10,000 regions stream 200,000)")
    ap.add_argument("--readprep-regions", type=int, default=2_000,
                    help="regions whose reads the input-side leg prepares (0 skips it)")
    ap.add_argument("--readprep-sample", type=int, default=300)
    ap.add_argument("--precision-regions", type=int, default=2_000,
                    help="regions of the precision-stress leg (0 skips it)")
    ap.add_argument("--precision-sample", type=int, default=10_000,
                    help="FP32-kept and rescued pairs each checked against the FP64 oracle")
    ap.add_argument("--pd-regions", type=int, default=1_000,
                    help="regions of the partially determined PairHMM leg (0 skips it)")
    ap.add_argument("--pd-sample", type=int, default=20, help="regions the PD oracle re-scores")
    ap.add_argument("--post-reads", type=int, default=1_000_000,
                    help="reads of the likelihood post-processing leg (0 skips it)")
    ap.add_argument("--post-sample", type=int, default=20_000)
    ap.add_argument("--sw-regions", type=int, default=1_000,
                    help="regions of the Smith-Waterman realignment leg (0 skips it)")
    ap.add_argument("--sw-sample", type=int, default=2_000)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB read + 1 GiB write elementwise kernel (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        x = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": x.numel() * 4, "bytes_written": y.numel() * 4,
                          "launches": 3}))
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (regions shard across GPUs "
              f"with no exchange; multi-GPU runs are out of scope here)", file=sys.stderr)
        return 2

    # Streaming leg first: its generator pool starts before this process touches the GPU.
    stream = None
    if args.stream_regions > 0:
        stream = streaming_leg(args)

    from gatk_amd import _native, synth
    from gatk_amd.batch import BatchEngine
    from gatk_amd.pairhmm import HaplotypeDataHolder, HipPairHmm, PairHMMNativeArguments, ReadDataHolder

    t0 = time.perf_counter()
    batch = synth.workload("batched_uneven", n_regions=args.regions)
    cells = batch.cells()
    log(f"generated {batch.n_regions} regions, {batch.n_pairs} pairs, {cells:.4e} cells in "
        f"{time.perf_counter() - t0:.1f} s")

    eng = BatchEngine()
    schedule = None
    if not args.no_tune:
        from gatk_amd import autotune

        t0 = time.perf_counter()
        schedule = autotune.tune(batch, use_cache=False, cache_path=None)
        schedule.apply(eng)
        log(f"autotuner picked {schedule} in {time.perf_counter() - t0:.1f} s")
    elif args.stripes:
        from gatk_amd import autotune

        schedule = autotune.Schedule(stripe_schedule=0 if args.stripes == "packed" else 1)
        schedule.apply(eng)
    t0 = time.perf_counter()
    dev = eng.upload(batch)
    log(f"uploaded to HBM in {time.perf_counter() - t0:.1f} s")

    out = np.empty(batch.n_pairs)
    for w in range(args.warmup):
        st = dev.forward(mode=_native.RUN_PERSISTENT, out=out if w == 0 else None)
        log(f"warmup {w}: {st.total_ms:.1f} ms, rescued {st.n_rescued}")
    if args.warmup == 0:
        dev.forward(mode=_native.RUN_PERSISTENT, out=out)

    t0 = time.perf_counter()
    rescued = 0
    for k in range(args.steps):
        st = dev.forward(mode=_native.RUN_PERSISTENT)
        rescued = st.n_rescued
        log(f"step {k}: {st.total_ms:.1f} ms ({st.gcups:.1f} GCUPS device time)")
    elapsed = time.perf_counter() - t0
    gcups = cells * args.steps / elapsed / 1e9
    ms_per_step = elapsed / args.steps * 1e3

    # Dominant kernel: FP32 forward, HIP events on the engine stream (eager launch).
    st = dev.forward(mode=_native.RUN_PERSISTENT_EAGER)
    achieved = cells * FLOP_PER_CELL / (st.fp32_kernel_ms * 1e-3) / 1e12
    achieved_gcups = cells / (st.total_ms * 1e-3) / 1e9  # FP32 and FP64 kernels of one eager pass
    log(f"eager: fp32 kernel {st.fp32_kernel_ms:.1f} ms, rescue kernel {st.rescue_kernel_ms:.1f} ms")
    dev.free()

    # Plain per-region drop-in path (one synchronous computeLikelihoods per region).
    plain = plain_dev = None
    if args.plain_regions > 0:
        binding = HipPairHmm()
        binding.load(None)
        binding.initialize(PairHMMNativeArguments())
        regs = range(min(args.plain_regions, batch.n_regions))
        prepared = []
        pcells = 0
        for k in regs:
            reads, haps = batch.region(k)
            prepared.append(([ReadDataHolder(*r) for r in reads], [HaplotypeDataHolder(h) for h in haps],
                             np.empty(len(reads) * len(haps))))
            p0, p1 = batch.region_pair_off[k], batch.region_pair_off[k + 1]
            pcells += batch.cells(np.arange(p0, p1))
        binding.compute_likelihoods(*prepared[0])
        t0 = time.perf_counter()
        for rh, hh, o in prepared:
            binding.compute_likelihoods(rh, hh, o)
        pt = time.perf_counter() - t0
        binding.done()
        plain = pcells / pt / 1e9
        # Device time of the same per-region work: each region as its own batch (same FP32 and FP64
        # kernels at a per-region grid), kernels timed by HIP events with inputs already resident.
        dev_ms = 0.0
        for k in regs:
            rb = batch.subset_regions([k])
            rdev = eng.upload(rb)
            rdev.forward(mode=_native.RUN_PERSISTENT_EAGER)
            dev_ms += rdev.forward(mode=_native.RUN_PERSISTENT_EAGER).total_ms
            rdev.free()
        plain_dev = pcells / (dev_ms * 1e-3) / 1e9
        log(f"plain per-region drop-in: {plain:.1f} GCUPS wall, {plain_dev:.1f} GCUPS device time over "
            f"{len(prepared)} regions")
    # Config 0: the plain forward on 64 reads x 16 haplotypes (one drop-in call), best of the calls.
    plain_small = None
    if args.small_calls > 0:
        small = synth.workload("plain_small")
        s_reads, s_haps = small.region(0)
        binding = HipPairHmm()
        binding.load(None)
        binding.initialize(PairHMMNativeArguments())
        rh0, hh0 = [ReadDataHolder(*r) for r in s_reads], [HaplotypeDataHolder(h) for h in s_haps]
        o0 = np.empty(len(s_reads) * len(s_haps))
        best = float("inf")
        for _ in range(args.small_calls + 1):
            t0 = time.perf_counter()
            binding.compute_likelihoods(rh0, hh0, o0)
            best = min(best, time.perf_counter() - t0)
        binding.done()
        sdev = eng.upload(small)
        sdev.forward(mode=_native.RUN_PERSISTENT_EAGER)
        s_ms = min(sdev.forward(mode=_native.RUN_PERSISTENT_EAGER).total_ms for _ in range(5))
        sdev.free()
        plain_small = {"pairs": small.n_pairs, "cells": small.cells(),
                       "dropin_wall_gcups": round(small.cells() / best / 1e9, 2),
                       "device_gcups": round(small.cells() / (s_ms * 1e-3) / 1e9, 2),
                       "workload": "plain_small: 64 reads of 150 bp x 16 haplotypes of 300 bp, one computeLikelihoods "
                                   f"call (wall: best of {args.small_calls + 1} calls, host in/out); device: the same "
                                   "pairs as one batch, kernels only (best of 5)"}
        log(f"config 0 plain_small: {plain_small}")
    eng.close()

    # Forward + backward (adjoint) leg: d log10 L / d(per-position Phred qb, qi, qd, qc).
    bwd = None
    if args.bwd_regions > 0:
        bwd, bwd_ctx = fwd_bwd_leg(args)
    fit = None
    if args.fit_regions > 0:
        fit = str_fit_leg(args)
    chunked = None
    if args.chunked_regions > 0:
        chunked = chunked_leg(args)
    readprep = None
    if args.readprep_regions > 0:
        readprep = readprep_leg(args)
    pdhmm = None
    if args.pd_regions > 0:
        pdhmm = pd_leg(args, min(16, os.cpu_count() or 1))
    post = None
    if args.post_reads > 0:
        post = postprocess_leg(args)
    swr = None
    if args.sw_regions > 0:
        swr = sw_leg(args)
    precision = None
    if args.precision_regions > 0:
        precision = precision_leg(args, min(16, os.cpu_count() or 1))
    cpu, worst, n_checked, proxy = None, None, 0, None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cpu, worst, n_checked = cpu_baseline(batch, out, args.cpu_seconds, threads)
        log(f"cpu baseline {cpu['value']:.3f} GCUPS on {threads} threads; max |dlog10L| on {n_checked} pairs = {worst:.2e}")
        one, _, _ = cpu_baseline(batch, None, args.cpu_seconds / 3, 1)
        cpu["value_1_thread"] = round(one["value"], 4)
        cpu["host"] = host_cpu()
        proxy = cpu_proxy_baseline(batch, out, args.cpu_seconds, threads)
        log(f"cpu GKL-class proxy: {proxy}; FP64 port on 1 thread {one['value']:.3f} GCUPS; host {cpu['host']}")
        if bwd is not None:
            bwd.update(cpu_grad_baseline(bwd_ctx, threads))
        cpu["reference_paths"] = reference_cpu_paths()

    line = {
        "metric": "PairHMM GCUPS fwd & fwd+bwd on 1 MI355X at |dlog10L|<=1e-4 vs FP64",
        "value": round(gcups, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 cells, f64 rescue + last-row sum",
        "data": "synthetic (seeded PCG64)",
        "config": {"workload": "batched_uneven: 10,000 regions, 2-128 haplotypes of 150-600 bp, "
                               "Poisson(80) reads of 100-250 bp, hipGraph persistent kernel",
                   "regions": batch.n_regions, "pairs": batch.n_pairs, "cells": cells,
                   "rescued_pairs": rescued},
        "fwd_bwd_gcups": bwd["gcups"] if bwd else None,
        "fwd_bwd": bwd,
        "str_fit": fit,
        "streaming": stream,
        "streaming_workload_chunked": chunked,
        "precision_stress": precision,
        "read_prep": readprep,
        "pd_hmm": pdhmm,
        "likelihood_postprocessing": post,
        "smith_waterman": swr,
        "schedule": schedule.__dict__ if schedule else None,
        "plain_per_region_gcups": round(plain, 2) if plain else None,
        "plain_per_region_device_gcups": round(plain_dev, 2) if plain else None,
        "batched_over_plain": round(gcups / plain, 2) if plain else None,
        "batched_over_plain_device": round(achieved_gcups / plain_dev, 2) if plain else None,
        "plain_small": plain_small,
        "max_abs_dlog10L_vs_fp64_sample": worst,
        "checked_pairs": n_checked,
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": int(HBM_BYTES_PER_CELL * cells),
                     "kernel": ("forward_kernel_packed<C>" if (schedule is None or schedule.stripe_schedule == 0)
                               else "forward_kernel<float, C>") + " over its LDS-class launches of one pass",
                     "kernel_ms": round(st.fp32_kernel_ms, 3),
                     "flop_per_cell": FLOP_PER_CELL},
        "cpu_baseline": cpu,
        "cpu_gkl_class_proxy": proxy,
    }
    print(json.dumps(line), flush=True)
    _reap_children()
    return 0


def _reap_children() -> None:
    """Joins multiprocessing children and logs any child process still alive at exit."""
    import multiprocessing

    for p in multiprocessing.active_children():
        p.join(timeout=10)
    kids = []
    try:
        for t in os.listdir(f"/proc/{os.getpid()}/task"):
            with open(f"/proc/{os.getpid()}/task/{t}/children") as f:
                kids += f.read().split()
    except OSError:
        pass
    for k in kids:
        try:
            with open(f"/proc/{k}/cmdline", "rb") as f:
                cmd = f.read().replace(b"\0", b" ").decode(errors="replace")[:200]
            log(f"child process {k} alive at exit: {cmd}")
        except OSError:
            pass


if __name__ == "__main__":
    sys.exit(main())
