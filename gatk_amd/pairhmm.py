"""Host-side mirror of GATK's PairHMM interface over the gfx950 engine.

Names, argument meaning and error behaviour follow the reference (paths relative to
src/main/java/org/broadinstitute/hellbender/):

* ``HipPairHmm`` plays the role of ``PairHMMNativeBinding`` (the GKL ``IntelPairHmm`` object that
  VectorLoglessPairHMM.java:44,62-81 holds): ``load`` / ``initialize`` / ``compute_likelihoods`` /
  ``done`` over the C-ABI.
* ``HipLoglessPairHMM`` plays the role of ``VectorLoglessPairHMM`` (utils/pairhmm/VectorLoglessPairHMM.java):
  ``initialize(haplotypes, per_sample_reads, ...)`` stores haplotype holders in list order (:89-102),
  ``compute_log10_likelihoods(matrix, reads, imputator)`` builds read holders (:118-129), calls the
  engine (:138) and maps results into the matrix in allele order (:140-153); ``close`` -> done (:163-168).
* ``Implementation`` mirrors ``PairHMM.Implementation`` (utils/pairhmm/PairHMM.java:40-109) for the
  native choices; construction raises ``HardwareFeatureException`` when the device is absent, the
  signal FASTEST_AVAILABLE uses to fall back (PairHMM.java:79-98).
* ``HipLoglessPairHMM.set_and_initialize_debug_output_stream`` writes the reference's HMM results file
  (PairHMM.java:364-385): a header line, then one line per (read, haplotype) with the inputs and the
  log10 likelihood in Java's ``%e`` format (``java_format_e``).
"""
from __future__ import annotations

import ctypes
import enum
import math
from dataclasses import dataclass, field
from decimal import ROUND_HALF_UP, Decimal
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _native

DEFAULT_GOP = 45  # ReadUtils default base insertion/deletion quality (ReadUtils.java:44, 838-846)
DEFAULT_GCP = 10  # --pair-hmm-gap-continuation-penalty (LikelihoodEngineArgumentCollection.java:61-63)


class HardwareFeatureException(RuntimeError):
    """UserException.HardwareFeatureException (UserException.java:435)."""


class GATKException(RuntimeError):
    """GATKException: a native call failed after the library loaded."""


def _raise_for(status: int, what: str) -> None:
    if status == _native.PAIRHMM_OK:
        return
    msg = f"{what}: {_native.last_error()}"
    if status == _native.PAIRHMM_ERR_NO_DEVICE:
        raise HardwareFeatureException(msg)
    if status == _native.PAIRHMM_ERR_BAD_ARG:
        raise ValueError(msg)  # IllegalArgumentException
    if status == _native.PAIRHMM_ERR_STATE:
        raise RuntimeError(msg)  # IllegalStateException
    raise GATKException(msg)


def _as_u8(x) -> np.ndarray:
    if isinstance(x, (bytes, bytearray)):
        return np.frombuffer(bytes(x), dtype=np.uint8)
    if isinstance(x, str):
        return np.frombuffer(x.encode("latin-1"), dtype=np.uint8)
    return np.ascontiguousarray(np.asarray(x), dtype=np.uint8)


@dataclass
class PairHMMNativeArguments:
    """PairHMMNativeArguments (PairHMMNativeArgumentCollection.java:11-22; defaults 4 / false)."""
    max_number_of_threads: int = 4
    use_double_precision: bool = False


@dataclass
class ReadDataHolder:
    """ReadDataHolder (VectorLoglessPairHMM.java:122-127): five byte arrays of one read."""
    read_bases: np.ndarray
    read_quals: np.ndarray
    insertion_gop: np.ndarray
    deletion_gop: np.ndarray
    overall_gcp: np.ndarray


@dataclass
class HaplotypeDataHolder:
    """HaplotypeDataHolder (VectorLoglessPairHMM.java:97-98)."""
    haplotype_bases: np.ndarray


# numpy images of the C-ABI holder structs (5 pointers + length; 1 pointer + length), so a whole
# call's holders are filled with a few vectorised assignments instead of one ctypes struct per read.
READ_HOLDER_DTYPE = np.dtype([("read_bases", "<u8"), ("read_quals", "<u8"), ("insertion_gop", "<u8"),
                              ("deletion_gop", "<u8"), ("overall_gcp", "<u8"), ("length", "<i4"), ("_pad", "<i4")])
HAPLOTYPE_HOLDER_DTYPE = np.dtype([("haplotype_bases", "<u8"), ("length", "<i4"), ("_pad", "<i4")])


def _concat_u8(items, n_bytes: Optional[int] = None) -> np.ndarray:
    """The items' bytes as one contiguous uint8 plane.  Byte-sized buffers (uint8 ndarrays, bytes)
    are joined through the buffer protocol in one C loop; the result is kept only if it holds the
    expected n_bytes (the items' lengths summed), so a wider dtype, a list or a string goes through
    the per-item uint8 conversion instead."""
    try:
        joined = b"".join(items)
        if n_bytes is not None and len(joined) == n_bytes:
            return np.frombuffer(joined, dtype=np.uint8)
    except (TypeError, BufferError, ValueError):
        pass
    return np.concatenate([_as_u8(x) for x in items])


def marshal_planes(dtype: np.dtype, columns):
    """Concatenates each column (a list of per-item byte arrays of equal lengths item by item) into
    one contiguous plane and returns a holder array of ``dtype`` whose pointer fields (in field
    order) address each item's bytes, plus the planes to keep alive for the call."""
    n = len(columns[0])
    if any(len(c) != n for c in columns):
        raise ValueError("columns differ in item count")
    items = [x for c in columns for x in c]
    all_lens = np.fromiter(map(len, items), dtype=np.int64, count=len(items)).reshape(len(columns), n)
    lens = all_lens[0]
    if not (all_lens == lens).all():
        raise ValueError("Read bases and read quals / GOP / GCP arrays aren't the same size")
    offs = np.zeros(n, dtype=np.uint64)
    if n > 1:
        offs[1:] = np.cumsum(lens[:-1], dtype=np.uint64)
    holders = np.zeros(n, dtype=dtype)
    total = int(lens.sum())
    # One concatenation for every column (column k's plane starts at k * total): a single call's
    # overhead instead of one per column.
    plane = np.zeros(1, np.uint8) if n == 0 or total == 0 else _concat_u8(items, total * len(columns))
    base = plane.ctypes.data
    for k, name in enumerate(dtype.names[:len(columns)]):
        holders[name] = np.uint64(base + k * total) + offs
    holders["length"] = lens
    return holders, [plane]


def marshal_reads(reads):
    return marshal_planes(READ_HOLDER_DTYPE, [[r.read_bases for r in reads], [r.read_quals for r in reads],
                                              [r.insertion_gop for r in reads], [r.deletion_gop for r in reads],
                                              [r.overall_gcp for r in reads]])


class HipPairHmm:
    """The native binding: PairHMMNativeBinding's load / initialize / computeLikelihoods / done."""

    def __init__(self) -> None:
        self._ctx = ctypes.c_void_p()
        self._lib = None

    def load(self, tmp_dir=None) -> bool:
        """PairHMMNativeBinding.load(File): True when a gfx950 device is usable."""
        self._lib = _native.load_library()
        return bool(self._lib.pairhmm_load())

    def initialize(self, args: PairHMMNativeArguments) -> None:
        a = _native.NativeArguments(int(args.max_number_of_threads), int(bool(args.use_double_precision)))
        _raise_for(self._lib.pairhmm_initialize(ctypes.byref(a), ctypes.byref(self._ctx)), "initialize")

    def compute_likelihoods(self, reads: Sequence[ReadDataHolder], haplotypes: Sequence[HaplotypeDataHolder],
                            likelihoods: np.ndarray) -> None:
        """computeLikelihoods(reads, haplotypes, out): out[r * n_haps + h], read-major."""
        if len(reads) == 0 or len(haplotypes) == 0:
            return
        out = likelihoods
        if out.dtype != np.float64 or not out.flags.c_contiguous or out.size < len(reads) * len(haplotypes):
            raise ValueError("likelihoods must be a contiguous float64 array of n_reads * n_haplotypes")
        rd, keep = marshal_reads(reads)
        hd, keep2 = marshal_planes(HAPLOTYPE_HOLDER_DTYPE, [[h.haplotype_bases for h in haplotypes]])
        st = self._lib.pairhmm_compute_likelihoods(
            self._ctx, rd.ctypes.data_as(ctypes.POINTER(_native.ReadDataHolder)), len(reads),
            hd.ctypes.data_as(ctypes.POINTER(_native.HaplotypeDataHolder)), len(haplotypes), out.ctypes.data)
        _raise_for(st, "computeLikelihoods")

    def done(self) -> None:
        if self._lib is not None and self._ctx:
            self._lib.pairhmm_done(self._ctx)
            self._ctx = ctypes.c_void_p()

    @property
    def context(self) -> ctypes.c_void_p:
        return self._ctx


@dataclass
class Read:
    """The GATKRead fields the PairHMM consumes: bases, base qualities and the BI/BD tags."""
    bases: np.ndarray
    base_qualities: np.ndarray
    insertion_qualities: Optional[np.ndarray] = None
    deletion_qualities: Optional[np.ndarray] = None

    def __post_init__(self):
        self.bases = _as_u8(self.bases)
        self.base_qualities = _as_u8(self.base_qualities)
        n = len(self.bases)
        if self.insertion_qualities is None:
            self.insertion_qualities = np.full(n, DEFAULT_GOP, dtype=np.uint8)
        if self.deletion_qualities is None:
            self.deletion_qualities = np.full(n, DEFAULT_GOP, dtype=np.uint8)
        self.insertion_qualities = _as_u8(self.insertion_qualities)
        self.deletion_qualities = _as_u8(self.deletion_qualities)

    def __len__(self) -> int:
        return len(self.bases)


@dataclass
class PairHMMInputScoreImputation:
    ins_open_penalties: np.ndarray
    del_open_penalties: np.ndarray
    gap_continuation_penalties: np.ndarray


class StandardPairHMMInputScoreImputator:
    """StandardPairHMMInputScoreImputator (tools/walkers/haplotypecaller/StandardPairHMMInputScoreImputator.java:20-49):
    GOPs from the read's BI/BD qualities and a constant GCP."""

    def __init__(self, constant_gcp: int = DEFAULT_GCP) -> None:
        self.constant_gcp = int(constant_gcp)

    @classmethod
    def new_instance(cls, constant_gcp: int) -> "StandardPairHMMInputScoreImputator":
        return cls(constant_gcp)

    def impute(self, read: Read) -> PairHMMInputScoreImputation:
        return PairHMMInputScoreImputation(read.insertion_qualities, read.deletion_qualities,
                                           np.full(len(read), self.constant_gcp, dtype=np.uint8))


class LikelihoodMatrix:
    """A dense allele x read log10-likelihood matrix (the LikelihoodMatrix sink,
    utils/genotyper/LikelihoodMatrix.java:14-82)."""

    def __init__(self, alleles: Sequence[bytes], evidence_count: int = 0) -> None:
        self._alleles = list(alleles)
        self.values = np.zeros((len(self._alleles), evidence_count))

    def alleles(self) -> List[bytes]:
        return self._alleles

    def number_of_alleles(self) -> int:
        return len(self._alleles)

    def set(self, allele_index: int, evidence_index: int, value: float) -> None:
        if evidence_index >= self.values.shape[1]:
            grow = np.zeros((self.values.shape[0], evidence_index + 1))
            grow[:, : self.values.shape[1]] = self.values
            self.values = grow
        self.values[allele_index, evidence_index] = value

    def get(self, allele_index: int, evidence_index: int) -> float:
        return float(self.values[allele_index, evidence_index])


class Implementation(enum.Enum):
    """Native PairHMM choices (PairHMM.Implementation, PairHMM.java:40-109): the gfx950 engine
    added next to AVX_LOGLESS_CACHING / AVX_LOGLESS_CACHING_OMP."""
    HIP_LOGLESS_CACHING = "HIP_LOGLESS_CACHING"

    def make_new_hmm(self, args: Optional[PairHMMNativeArguments] = None) -> "HipLoglessPairHMM":
        return HipLoglessPairHMM(args or PairHMMNativeArguments())


RESULTS_HEADER = "# hap-bases read-bases read-qual read-ins-qual read-del-qual gcp expected-result"


def java_format_e(v: float) -> str:
    """Java's String.format("%e", v) (java.util.Formatter): the double's shortest decimal digits rounded
    HALF_UP to 7 significant digits, d.dddddde+XX with at least two exponent digits; "Infinity",
    "-Infinity" and "NaN" for the non-finite values."""
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "Infinity" if v > 0 else "-Infinity"
    if v == 0.0:
        return ("-" if math.copysign(1.0, v) < 0 else "") + "0.000000e+00"
    d = Decimal(repr(float(v)))
    exp = d.adjusted()
    q = d.scaleb(-exp).quantize(Decimal("1.000000"), rounding=ROUND_HALF_UP)
    if abs(q) >= 10:
        exp += 1
        q = d.scaleb(-exp).quantize(Decimal("1.000000"), rounding=ROUND_HALF_UP)
    return f"{q}e{'+' if exp >= 0 else '-'}{abs(exp):02d}"


def phred_to_fastq(quals) -> str:
    """htsjdk SAMUtils.phredToFastq: Phred q as the character q + 33 (q in 0..93)."""
    q = np.asarray(quals, dtype=np.int64)
    if q.size and (q.min() < 0 or q.max() > 93):
        raise ValueError("Phred scores must lie in 0..93 to print as FASTQ characters")
    return bytes((q + 33).astype(np.uint8)).decode("ascii")


def results_line(hap: bytes, read: "ReadDataHolder", lk: float) -> str:
    """One record of the HMM results file (PairHMM.writeToResultsFileIfApplicable, PairHMM.java:377-385)."""
    return ("\n" + bytes(_as_u8(hap)).decode("ascii") + " " + bytes(_as_u8(read.read_bases)).decode("ascii") + " " +
            phred_to_fastq(read.read_quals) + " " + phred_to_fastq(read.insertion_gop) + " " +
            phred_to_fastq(read.deletion_gop) + " " + phred_to_fastq(read.overall_gcp) + " " + java_format_e(lk))


class HipLoglessPairHMM:
    """VectorLoglessPairHMM (utils/pairhmm/VectorLoglessPairHMM.java:24-169) over the gfx950 engine."""

    def __init__(self, args: Optional[PairHMMNativeArguments] = None) -> None:
        self._binding = HipPairHmm()
        if not self._binding.load(None):
            raise HardwareFeatureException("Machine does not support the gfx950 HIP PairHMM.")
        self._binding.initialize(args or PairHMMNativeArguments())
        self._hap_index: Dict[bytes, int] = {}
        self._hap_holders: List[HaplotypeDataHolder] = []
        self._log_likelihoods: Optional[np.ndarray] = None
        self._debug = None

    def set_and_initialize_debug_output_stream(self, writer) -> None:
        """PairHMM.setAndInitializeDebugOutputStream (PairHMM.java:364-372): attach a text writer and
        write the results-file header; every later likelihood is recorded on it."""
        self._debug = writer
        writer.write(RESULTS_HEADER)

    def initialize(self, haplotypes: Sequence[bytes], per_sample_read_list=None, read_max_length: int = 0,
                   haplotype_max_length: int = 0) -> None:
        """Stores haplotype holders in list order (VectorLoglessPairHMM.java:89-102)."""
        self._hap_holders = []
        self._hap_index = {}
        for idx, h in enumerate(haplotypes):
            hb = bytes(_as_u8(h))
            self._hap_holders.append(HaplotypeDataHolder(_as_u8(hb)))
            self._hap_index[hb] = idx

    def compute_log10_likelihoods(self, log_likelihoods: LikelihoodMatrix, processed_reads: Sequence[Read],
                                  input_score_imputator) -> None:
        """VectorLoglessPairHMM.computeLog10Likelihoods (:108-159)."""
        if len(processed_reads) == 0:
            return
        holders = []
        for read in processed_reads:
            imp = input_score_imputator.impute(read)
            holders.append(ReadDataHolder(read.bases, read.base_qualities, imp.ins_open_penalties,
                                          imp.del_open_penalties, imp.gap_continuation_penalties))
        n_haps = log_likelihoods.number_of_alleles()
        out = np.zeros(len(processed_reads) * n_haps)
        self._binding.compute_likelihoods(holders, self._hap_holders, out)
        self._log_likelihoods = out
        for r in range(len(processed_reads)):
            for a, hap in enumerate(log_likelihoods.alleles()):
                idx = self._hap_index[bytes(_as_u8(hap))]
                log_likelihoods.set(a, r, out[r * n_haps + idx])
                if self._debug is not None:  # VectorLoglessPairHMM.java:149
                    self._debug.write(results_line(hap, holders[r], float(out[r * n_haps + idx])))

    def get_log_likelihood_array(self) -> Optional[np.ndarray]:
        return self._log_likelihoods

    def close(self) -> None:
        self._binding.done()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
