"""Seeded synthetic active regions for the PairHMM workloads.

Every workload is a list of active regions; each region has haplotypes mutated from a random
reference window and reads sampled from those haplotypes with sequencing errors.  The generator
writes straight into the packed batch layout (``gatk_amd.batch.PackedBatch``) so that ~10^6 reads
can be produced in seconds.  numpy ``PCG64`` with one fixed seed per workload.

Workloads (sizes as named):
  * ``plain_small``      -- 1 region: reference 1,000 bp; 16 haplotypes = the same 300 bp window
                            with 0-3 SNPs and 0-2 indels of 1-8 bp; 64 reads of 150 bp, 1% subs;
                            base Q 20-40, GOP 45, GCP 10 (1,024 pairs).
  * ``batched_uneven``   -- 10,000 regions: haplotypes 2-128 per region, window 150-600 bp;
                            reads Poisson(80) clamped to 1-1,000, length 100-250, 1% subs,
                            0.1% 1-3 bp indels; base Q 10-40, GOP 45, GCP 10.
  * ``precision_stress`` -- 2,000 regions: STR-rich haplotypes (1-6 bp motifs x 5-40), window
                            800-1,500 bp; reads 250-600 bp, 5% subs; base Q 2-15.  Haplotype and read
                            counts follow ``batched_uneven``.
  * ``streaming``        -- ``batched_uneven`` regions with 10% drawn from ``precision_stress``.
"""
from __future__ import annotations

from dataclasses import dataclass, replace
from typing import List, Optional, Tuple

import numpy as np

from .batch import PackedBatch

BASES = np.frombuffer(b"ACGT", dtype=np.uint8)
_CODE = np.zeros(256, dtype=np.int64)
for _i, _b in enumerate(b"ACGT"):
    _CODE[_b] = _i


@dataclass(frozen=True)
class RegionSpec:
    n_haps: Tuple[int, int] = (2, 128)          # uniform inclusive
    window: Tuple[int, int] = (150, 600)        # reference window length, uniform inclusive
    hap_snps: Tuple[int, int] = (0, 3)
    hap_indels: Tuple[int, int] = (0, 2)
    hap_indel_len: Tuple[int, int] = (1, 8)
    reads_mean: float = 80.0                    # Poisson mean
    reads_clamp: Tuple[int, int] = (1, 1000)
    read_len: Tuple[int, int] = (100, 250)
    sub_rate: float = 0.01
    indel_rate: float = 0.001
    indel_len: Tuple[int, int] = (1, 3)
    base_q: Tuple[int, int] = (10, 40)
    gop: int = 45
    gcp: int = 10
    str_rich: bool = False
    str_motif: Tuple[int, int] = (1, 6)
    str_repeats: Tuple[int, int] = (5, 40)
    fixed_reads: Optional[int] = None
    fixed_haps: Optional[int] = None


BATCHED_UNEVEN = RegionSpec()
PRECISION_STRESS = RegionSpec(window=(800, 1500), read_len=(250, 600), sub_rate=0.05, base_q=(2, 15), str_rich=True)
PLAIN_SMALL = RegionSpec(window=(300, 300), read_len=(150, 150), base_q=(20, 40), indel_rate=0.0,
                         fixed_reads=64, fixed_haps=16)

WORKLOADS = {
    "plain_small": (PLAIN_SMALL, 1, 0),
    "batched_uneven": (BATCHED_UNEVEN, 10_000, 1),
    "precision_stress": (PRECISION_STRESS, 2_000, 2),
    "streaming": (BATCHED_UNEVEN, 200_000, 4),
}


def _ui(rng, lo_hi) -> int:
    return int(rng.integers(lo_hi[0], lo_hi[1] + 1))


def _random_seq(rng, n: int) -> np.ndarray:
    return BASES[rng.integers(0, 4, n)]


def _str_rich_seq(rng, n: int, spec: RegionSpec) -> np.ndarray:
    """Random sequence interleaved with short tandem repeats (motif 1-6 bp x 5-40)."""
    parts, total = [], 0
    while total < n:
        flank = _random_seq(rng, int(rng.integers(10, 60)))
        motif = _random_seq(rng, _ui(rng, spec.str_motif))
        tract = np.tile(motif, _ui(rng, spec.str_repeats))
        parts += [flank, tract]
        total += len(flank) + len(tract)
    return np.concatenate(parts)[:n]


def _mutate(rng, window: np.ndarray, spec: RegionSpec) -> np.ndarray:
    h = window.copy()
    n_snp = _ui(rng, spec.hap_snps)
    if n_snp:
        pos = rng.integers(0, len(h), n_snp)
        h[pos] = BASES[(_CODE[h[pos]] + rng.integers(1, 4, n_snp)) % 4]
    for _ in range(_ui(rng, spec.hap_indels)):
        L = _ui(rng, spec.hap_indel_len)
        p = int(rng.integers(0, len(h)))
        if rng.random() < 0.5:
            h = np.concatenate([h[:p], _random_seq(rng, L), h[p:]])
        elif len(h) - L >= 2:
            h = np.concatenate([h[:p], h[p + L:]])
    return h


def _gen_region(rng, spec: RegionSpec, window: Optional[np.ndarray] = None):
    """One region: (hap list, 5 read planes concatenated, read lengths)."""
    W = _ui(rng, spec.window)
    if window is None:
        window = _str_rich_seq(rng, W, spec) if spec.str_rich else _random_seq(rng, W)
    A = spec.fixed_haps if spec.fixed_haps else _ui(rng, spec.n_haps)
    haps = [_mutate(rng, window, spec) for _ in range(A)]
    if spec.fixed_reads:
        N = spec.fixed_reads
    else:
        N = int(np.clip(rng.poisson(spec.reads_mean), spec.reads_clamp[0], spec.reads_clamp[1]))
    hl = np.array([len(h) for h in haps])
    hoff = np.concatenate([[0], np.cumsum(hl)[:-1]])
    hcat = np.concatenate(haps)
    src = rng.integers(0, A, N)
    R = rng.integers(spec.read_len[0], spec.read_len[1] + 1, N)
    R = np.minimum(R, hl[src])
    start = (rng.random(N) * (hl[src] - R + 1)).astype(np.int64)
    Rmax = int(R.max())
    idx = hoff[src][:, None] + start[:, None] + np.arange(Rmax)[None, :]
    idx = np.minimum(idx, len(hcat) - 1)
    mat = hcat[idx]
    mask = np.arange(Rmax)[None, :] < R[:, None]
    sub = (rng.random(mat.shape) < spec.sub_rate) & mask
    ns = int(sub.sum())
    if ns:
        mat[sub] = BASES[(_CODE[mat[sub]] + rng.integers(1, 4, ns)) % 4]
    reads = [mat[k, : R[k]] for k in range(N)]
    if spec.indel_rate > 0:
        n_ev = rng.binomial(R, spec.indel_rate)
        for k in np.nonzero(n_ev)[0]:
            s = reads[k]
            for _ in range(int(n_ev[k])):
                L = _ui(rng, spec.indel_len)
                p = int(rng.integers(0, len(s)))
                if rng.random() < 0.5:
                    s = np.concatenate([s[:p], _random_seq(rng, L), s[p:]])[: R[k]]
                elif len(s) - L >= 1:
                    s = np.concatenate([s[:p], s[p + L:]])
            reads[k] = s
    rl = np.array([len(s) for s in reads], dtype=np.int32)
    bases = np.concatenate(reads)
    n = len(bases)
    quals = rng.integers(spec.base_q[0], spec.base_q[1] + 1, n).astype(np.uint8)
    ins = np.full(n, spec.gop, dtype=np.uint8)
    dele = np.full(n, spec.gop, dtype=np.uint8)
    gcp = np.full(n, spec.gcp, dtype=np.uint8)
    return haps, (bases, quals, ins, dele, gcp), rl


def generate(spec: RegionSpec, n_regions: int, seed: int, mix: Optional[Tuple[RegionSpec, float]] = None,
             plain_reference: bool = False) -> PackedBatch:
    """Generate n_regions regions.  ``mix=(spec2, frac)`` draws that fraction of regions from spec2.
    ``plain_reference`` draws the 300 bp window from a 1,000 bp random reference (plain_small)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    hap_chunks: List[np.ndarray] = []
    hap_lens: List[np.ndarray] = []
    planes: List[List[np.ndarray]] = [[], [], [], [], []]
    read_lens: List[np.ndarray] = []
    pr_chunks, ph_chunks = [], []
    rpo, rro, rho = [0], [0], [0]
    n_r = n_h = 0
    for _ in range(n_regions):
        sp = spec
        if mix is not None and rng.random() < mix[1]:
            sp = mix[0]
        window = None
        if plain_reference:
            ref = _random_seq(rng, 1000)
            W = _ui(rng, sp.window)
            s0 = int(rng.integers(0, 1000 - W + 1))
            window = ref[s0:s0 + W]
        haps, rp, rl = _gen_region(rng, sp, window)
        hap_chunks += haps
        hap_lens.append(np.array([len(h) for h in haps], dtype=np.int32))
        for k in range(5):
            planes[k].append(rp[k])
        read_lens.append(rl)
        A, N = len(haps), len(rl)
        pr_chunks.append(np.repeat(np.arange(n_r, n_r + N, dtype=np.int32), A))
        ph_chunks.append(np.tile(np.arange(n_h, n_h + A, dtype=np.int32), N))
        n_r += N
        n_h += A
        rpo.append(rpo[-1] + N * A)
        rro.append(n_r)
        rho.append(n_h)
    hl = np.concatenate(hap_lens)
    rl = np.concatenate(read_lens)
    return PackedBatch(
        read_bases=np.concatenate(planes[0]), read_quals=np.concatenate(planes[1]),
        read_ins=np.concatenate(planes[2]), read_del=np.concatenate(planes[3]), read_gcp=np.concatenate(planes[4]),
        read_off=np.concatenate([[0], np.cumsum(rl, dtype=np.int64)[:-1]]).astype(np.int64), read_len=rl,
        hap_bases=np.concatenate(hap_chunks),
        hap_off=np.concatenate([[0], np.cumsum(hl, dtype=np.int64)[:-1]]).astype(np.int64), hap_len=hl,
        pair_read=np.concatenate(pr_chunks), pair_hap=np.concatenate(ph_chunks),
        region_pair_off=np.asarray(rpo, dtype=np.int64), region_read_off=np.asarray(rro, dtype=np.int64),
        region_hap_off=np.asarray(rho, dtype=np.int64),
    )


def workload(name: str, n_regions: Optional[int] = None, seed: Optional[int] = None) -> PackedBatch:
    """Named workload (see module docstring); ``n_regions`` scales the region count down/up."""
    spec, default_n, default_seed = WORKLOADS[name]
    n = default_n if n_regions is None else int(n_regions)
    s = default_seed if seed is None else int(seed)
    if name == "plain_small":
        return generate(spec, n, s, plain_reference=True)
    if name == "streaming":
        return generate(spec, n, s, mix=(PRECISION_STRESS, 0.10))
    return generate(spec, n, s)
