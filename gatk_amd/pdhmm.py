"""Partially determined PairHMM on the gfx950 engine: the drop-in for GKL's IntelPDHMM.

``HipPDHmm`` mirrors the native binding VectorLoglessPairPDHMM calls (IntelPDHMM: load / initialize /
computeLikelihoods / done, VectorLoglessPairPDHMM.java:49-61, 115); ``HipLoglessPairPDHMM`` mirrors
VectorLoglessPairPDHMM.computeLog10Likelihoods (:72-140) around it, including the -inf mask of
pairs whose read does not reach the haplotype's determined alleles (:125-135) and the results file
(PDPairHMM.java:349-370).  Every likelihood is computed by pdhmm.hip through
``pairhmm_pd_compute_likelihoods``; there is no host fallback.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from . import _native
from .pairhmm import (LikelihoodMatrix, ReadDataHolder, _as_u8, _raise_for, java_format_e, marshal_planes,
                      marshal_reads, phred_to_fastq)

PD_HAPLOTYPE_HOLDER_DTYPE = np.dtype([("haplotype_bases", "<u8"), ("haplotype_pd_bases", "<u8"), ("length", "<i4"),
                                      ("_pad", "<i4")])

# PartiallyDeterminedHaplotype byte layout (PartiallyDeterminedHaplotype.java:59-66).
SNP, DEL_START, DEL_END = 1, 2, 4
A, C, G, T = 8, 16, 32, 64

PD_RESULTS_HEADER = ("# hap-bases\thap-pd-bases\tread-bases\tread-qual\tread-ins-qual\tread-del-qual\tgcp\t"
                     "expected-result")  # PDPairHMM.java:352


def pd_bytes_for_alleles(ref: bytes, alt: bytes) -> bytes:
    """PartiallyDeterminedHaplotype.getPDBytesForHaplotypes (:176-203): SNP with the alternative's base
    bit on the first base, or a deletion span DEL_START .. DEL_END over the deleted bases."""
    if len(alt) == len(ref):
        bit = {ord("A"): A, ord("C"): C, ord("T"): T, ord("G"): G}.get(alt[0])
        if bit is None:
            raise RuntimeError("Found unexpected base in alt alleles")
        out = bytearray(len(ref))
        out[0] = (SNP + bit) & 0xFF
        return bytes(out)
    out = bytearray(len(ref) - 1)
    out[0] = (out[0] + DEL_START) & 0xFF
    out[-1] = (out[-1] + DEL_END) & 0xFF
    return bytes(out)


@dataclass
class PDHMMNativeArguments:
    """PDHMMNativeArguments (PDPairHMMNativeArgumentCollection.java:33-48); no device meaning."""
    max_number_of_threads: int = 4
    avx_level: str = "FASTEST_AVAILABLE"
    open_mp_setting: str = "FASTEST_AVAILABLE"


@dataclass
class HaplotypeDataHolder:
    """HaplotypeDataHolder of the PD binding (VectorLoglessPairPDHMM.java:101-106)."""
    haplotype_bases: np.ndarray
    haplotype_pd_bases: np.ndarray


class HipPDHmm:
    """IntelPDHMM's load / initialize / computeLikelihoods / done over the gfx950 engine."""

    def __init__(self) -> None:
        self._ctx = ctypes.c_void_p()
        self._lib = None

    def load(self, tmp_dir=None) -> bool:
        self._lib = _native.load_library()
        return bool(self._lib.pairhmm_load())

    def initialize(self, args: Optional[PDHMMNativeArguments] = None) -> None:
        a = _native.NativeArguments(int((args or PDHMMNativeArguments()).max_number_of_threads), 1)
        _raise_for(self._lib.pairhmm_initialize(ctypes.byref(a), ctypes.byref(self._ctx)), "initialize")

    def compute_likelihoods(self, reads: Sequence[ReadDataHolder], haplotypes: Sequence[HaplotypeDataHolder],
                            likelihoods: np.ndarray) -> None:
        """computeLikelihoods(reads, haplotypes, out): out[r * n_haps + h], read-major."""
        if len(reads) == 0 or len(haplotypes) == 0:
            return
        if (likelihoods.dtype != np.float64 or not likelihoods.flags.c_contiguous
                or likelihoods.size < len(reads) * len(haplotypes)):
            raise ValueError("likelihoods must be a contiguous float64 array of n_reads * n_haplotypes")
        rd, keep = marshal_reads(reads)
        try:
            hd, keep2 = marshal_planes(PD_HAPLOTYPE_HOLDER_DTYPE, [[h.haplotype_bases for h in haplotypes],
                                                                   [h.haplotype_pd_bases for h in haplotypes]])
        except ValueError:
            raise ValueError("haplotype bases and PD bases aren't the same size") from None
        st = self._lib.pairhmm_pd_compute_likelihoods(
            self._ctx, rd.ctypes.data_as(ctypes.POINTER(_native.ReadDataHolder)), len(reads),
            hd.ctypes.data_as(ctypes.POINTER(_native.PDHaplotypeDataHolder)), len(haplotypes),
            likelihoods.ctypes.data)
        _raise_for(st, "computeLikelihoods")

    def done(self) -> None:
        if self._lib is not None and self._ctx:
            self._lib.pairhmm_done(self._ctx)
            self._ctx = ctypes.c_void_p()


def overlaps_with_margin(a: Tuple[int, int], b: Tuple[int, int], margin: int) -> bool:
    """SimpleInterval.overlapsWithMargin (SimpleInterval.java:239-247) on one contig, 1-based closed."""
    if margin < 0:
        raise ValueError(f"given margin is negative: {margin}")
    return a[0] <= b[1] + margin and b[0] - margin <= a[1]


@dataclass
class PDAllele:
    """The PartiallyDeterminedHaplotype fields the HMM reads: bases, PD bytes and the extent of its
    determined alleles (getMaximumExtentOfSiteDeterminedAlleles)."""
    bases: bytes
    pd_bases: bytes
    determined_extent: Optional[Tuple[int, int]] = None


def pd_results_line(hap: bytes, pd: bytes, read: ReadDataHolder, lk: float) -> str:
    """One record of the PD results file (PDPairHMM.java:362-364); PD bytes print as Java's
    Arrays.toString of signed bytes."""
    signed = ", ".join(str(int(x) - 256 if int(x) >= 128 else int(x)) for x in _as_u8(pd))
    return ("\n" + bytes(_as_u8(hap)).decode("ascii") + "\t[" + signed + "]\t" +
            bytes(_as_u8(read.read_bases)).decode("ascii") + "\t" + phred_to_fastq(read.read_quals) + "\t" +
            phred_to_fastq(read.insertion_gop) + "\t" + phred_to_fastq(read.deletion_gop) + "\t" +
            phred_to_fastq(read.overall_gcp) + "\t" + java_format_e(lk))


def parse_pd_results_line(line: str):
    """The reconstitution of PDPairHMMLikelihoodCalculationEngineUnitTest.java:45-54: haplotype, PD
    bytes, a ReadDataHolder and the expected log10 likelihood of one results-file record."""
    f = line.rstrip("\n").split("\t")
    pd = bytes(int(x.strip()) & 0xFF for x in f[1][1:-1].split(","))
    q = [np.frombuffer(s.encode("ascii"), np.uint8) - 33 for s in f[3:7]]
    read = ReadDataHolder(np.frombuffer(f[2].encode("ascii"), np.uint8), *[x.astype(np.uint8) for x in q])
    return f[0].encode("ascii"), pd, read, float(f[7])


class HipLoglessPairPDHMM:
    """VectorLoglessPairPDHMM (utils/pairhmm/VectorLoglessPairPDHMM.java:34-140) over the gfx950 engine."""

    def __init__(self, args: Optional[PDHMMNativeArguments] = None) -> None:
        self._binding = HipPDHmm()
        if not self._binding.load():
            from .pairhmm import HardwareFeatureException
            raise HardwareFeatureException("Machine does not support the gfx950 PD-PairHMM.")
        self._binding.initialize(args)
        self._writer = None
        self.m_log_likelihood_array: Optional[np.ndarray] = None

    def set_and_initialize_debug_output_stream(self, writer) -> None:
        self._writer = writer
        writer.write(PD_RESULTS_HEADER)

    def compute_log10_likelihoods(self, log_likelihoods: LikelihoodMatrix, alleles: Sequence[PDAllele],
                                  processed_reads, input_score_imputator,
                                  range_for_read_overlap_to_determined_bases: int = -1,
                                  read_spans: Optional[Sequence[Tuple[int, int]]] = None) -> None:
        """computeLog10Likelihoods (:72-140).  ``processed_reads`` are ``pairhmm.Read``s; ``read_spans``
        are their unclipped spans, needed when the overlap range is >= 0."""
        if len(processed_reads) == 0:
            return
        holders: List[ReadDataHolder] = []
        for read in processed_reads:
            imp = input_score_imputator.impute(read)
            holders.append(ReadDataHolder(read.bases, read.base_qualities, imp.ins_open_penalties,
                                          imp.del_open_penalties, imp.gap_continuation_penalties))
        haps = [HaplotypeDataHolder(_as_u8(a.bases), _as_u8(a.pd_bases)) for a in alleles]
        n_r, n_a = len(holders), len(haps)
        out = np.empty(n_r * n_a, dtype=np.float64)
        self._binding.compute_likelihoods(holders, haps, out)
        for r in range(n_r):
            for a, allele in enumerate(alleles):
                k = r * n_a + a
                if range_for_read_overlap_to_determined_bases >= 0:
                    if read_spans is None or allele.determined_extent is None:
                        raise ValueError("read spans and determined extents are needed for the overlap range")
                    if not overlaps_with_margin(allele.determined_extent, read_spans[r],
                                                range_for_read_overlap_to_determined_bases + 1):
                        out[k] = -np.inf
                log_likelihoods.set(a, r, out[k])
                if self._writer is not None:
                    self._writer.write(pd_results_line(allele.bases, allele.pd_bases, holders[r], out[k]))
        self.m_log_likelihood_array = out

    def get_log_likelihood_array(self) -> Optional[np.ndarray]:
        return self.m_log_likelihood_array

    def close(self) -> None:
        self._binding.done()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
