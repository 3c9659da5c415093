"""Likelihood post-processing on the device: the mismapping-rate cap and the poorly-modelled read
filter that PairHMMLikelihoodCalculationEngine.computeReadLikelihoods applies after the PairHMM
(PairHMMLikelihoodCalculationEngine.java:200-201), run by postprocess.hip through
``pairhmm_normalize_and_filter_likelihoods`` on a matrix that stays in device memory."""
from __future__ import annotations

import ctypes
import math
from typing import Optional

from . import _native
from .pairhmm import _raise_for

DEFAULT_EXPECTED_ERROR_RATE_PER_BASE = 0.02        # ReadLikelihoodCalculationEngine.java:188
DEFAULT_DYNAMIC_DISQUALIFICATION_SCALE_FACTOR = 1.0  # PairHMMLikelihoodCalculationEngine.java:41


def log10_mismapping_cap(phred_scaled_global_read_mismapping_rate: int) -> float:
    """AssemblyBasedCallerUtils.getGlobalMismatchingRateFromArgs (AssemblyBasedCallerUtils.java:286-289):
    QualityUtils.qualToErrorProbLog10(phred) = phred * -0.1, or -inf (no cap) for a negative rate."""
    p = int(phred_scaled_global_read_mismapping_rate)
    return -math.inf if p < 0 else p * -0.1


class LikelihoodPostProcessor:
    def __init__(self) -> None:
        self.lib = _native.load_library()
        self.ctx = ctypes.c_void_p()
        a = _native.NativeArguments(1, 0)
        _raise_for(self.lib.pairhmm_initialize(ctypes.byref(a), ctypes.byref(self.ctx)), "initialize")

    def normalize_and_filter(self, likelihoods, read_len, reference_allele_index: int = 0,
                             log10_cap: float = -4.5, symmetric: bool = False, hmm_base_quals=None,
                             read_off=None, expected_error_rate_per_base: float = DEFAULT_EXPECTED_ERROR_RATE_PER_BASE,
                             dynamic_disqualification: bool = False,
                             disqualification_scale: float = DEFAULT_DYNAMIC_DISQUALIFICATION_SCALE_FACTOR,
                             keep=None, threshold=None):
        """``likelihoods``: device float64 tensor [n_reads, n_alleles] (read-major, normalised in place);
        ``read_len`` int32, ``read_off`` int64 and ``hmm_base_quals`` uint8 device tensors describe each
        read's HMM base qualities.  Returns (keep uint8 [n_reads], threshold float64 [n_reads] or None)."""
        import torch

        if likelihoods.dtype != torch.float64 or not likelihoods.is_cuda or likelihoods.dim() != 2:
            raise ValueError("likelihoods must be a 2-D float64 device tensor")
        if not likelihoods.is_contiguous():
            raise ValueError("likelihoods must be contiguous (read-major)")
        n_reads, n_alleles = likelihoods.shape
        if read_len.dtype != torch.int32 or read_len.numel() != n_reads:
            raise ValueError("read_len must be int32 with one entry per read")
        if keep is None:
            keep = torch.empty(n_reads, dtype=torch.uint8, device=likelihoods.device)
        ptr = lambda t: t.data_ptr() if t is not None else None  # noqa: E731
        st = self.lib.pairhmm_normalize_and_filter_likelihoods(
            self.ctx, likelihoods.data_ptr(), n_reads, n_alleles, int(reference_allele_index), float(log10_cap),
            int(bool(symmetric)), ptr(hmm_base_quals), ptr(read_off), read_len.data_ptr(),
            float(expected_error_rate_per_base), int(bool(dynamic_disqualification)), float(disqualification_scale),
            keep.data_ptr(), ptr(threshold))
        _raise_for(st, "normalize_and_filter_likelihoods")
        return keep, threshold

    def close(self) -> None:
        if self.ctx:
            self.lib.pairhmm_done(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
