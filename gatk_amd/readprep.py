"""HaplotypeCaller's read-quality preparation on the device (PairHMMLikelihoodCalculationEngine.
modifyReadQualities, tools/walkers/haplotypecaller/PairHMMLikelihoodCalculationEngine.java:283-305):
the PCR indel error model (applyPCRErrorModel, :360-371), then the minimum-quality caps
(capMinimumReadQualities, :307-320), over the packed read planes of a batch held in torch device
tensors.  Both steps run as HIP kernels through the C-ABI (pairhmm_apply_pcr_error_model,
pairhmm_cap_read_qualities); there is no host fallback.
"""
from __future__ import annotations

import ctypes
import enum
from typing import Optional

import torch

from . import _native
from .pairhmm import HardwareFeatureException, PairHMMNativeArguments, _raise_for

MIN_USABLE_Q_SCORE = 6           # QualityUtils.java:37
DEFAULT_BASE_QUALITY_THRESHOLD = 18


class PCRErrorModel(enum.IntEnum):
    """PairHMMLikelihoodCalculationEngine.PCRErrorModel (:53-61); the value is the rate factor."""
    NONE = 0
    HOSTILE = 1
    AGGRESSIVE = 2
    CONSERVATIVE = 3


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _check(t: torch.Tensor, dtype, name: str) -> None:
    if not t.is_cuda or t.dtype != dtype or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous {dtype} device tensor")


class ReadPrep:
    """One engine context for read-quality preparation."""

    def __init__(self, args: Optional[PairHMMNativeArguments] = None) -> None:
        self.lib = _native.load_library()
        if not self.lib.pairhmm_load():
            raise HardwareFeatureException("no gfx950 device visible to this process")
        a = args or PairHMMNativeArguments()
        na = _native.NativeArguments(int(a.max_number_of_threads), int(bool(a.use_double_precision)))
        self.ctx = ctypes.c_void_p()
        _raise_for(self.lib.pairhmm_initialize(ctypes.byref(na), ctypes.byref(self.ctx)), "initialize")

    def apply_pcr_error_model(self, read_bases, read_off, read_len, ins_quals, del_quals,
                              model: PCRErrorModel) -> None:
        """In place on ins_quals / del_quals (uint8 device planes)."""
        for t, dt, n in ((read_bases, torch.uint8, "read_bases"), (read_off, torch.int64, "read_off"),
                         (read_len, torch.int32, "read_len"), (ins_quals, torch.uint8, "ins_quals"),
                         (del_quals, torch.uint8, "del_quals")):
            _check(t, dt, n)
        _raise_for(self.lib.pairhmm_apply_pcr_error_model(self.ctx, _ptr(read_bases), _ptr(read_off), _ptr(read_len),
                                                          int(read_off.numel()), int(model), _ptr(ins_quals),
                                                          _ptr(del_quals)), "apply_pcr_error_model")

    def cap_read_qualities(self, read_off, read_len, base_quals, ins_quals, del_quals,
                           mapping_quality: Optional[torch.Tensor] = None,
                           base_quality_threshold: int = DEFAULT_BASE_QUALITY_THRESHOLD,
                           disable_cap_to_mapq: bool = False) -> None:
        """In place on the three uint8 quality planes; mapping_quality is int32 per read."""
        for t, dt, n in ((read_off, torch.int64, "read_off"), (read_len, torch.int32, "read_len"),
                         (base_quals, torch.uint8, "base_quals"), (ins_quals, torch.uint8, "ins_quals"),
                         (del_quals, torch.uint8, "del_quals")):
            _check(t, dt, n)
        if not disable_cap_to_mapq:
            if mapping_quality is None:
                raise ValueError("mapping_quality is required unless disable_cap_to_mapq")
            _check(mapping_quality, torch.int32, "mapping_quality")
        _raise_for(self.lib.pairhmm_cap_read_qualities(self.ctx, _ptr(read_off), _ptr(read_len), int(read_off.numel()),
                                                       _ptr(mapping_quality), int(base_quality_threshold),
                                                       int(bool(disable_cap_to_mapq)), _ptr(base_quals),
                                                       _ptr(ins_quals), _ptr(del_quals)), "cap_read_qualities")

    def dragstr_impute(self, params, read_bases, read_off, read_len, out_gop, out_gcp,
                       max_read_len: Optional[int] = None) -> None:
        """DragstrPairHMMInputScoreImputator.impute on the device: per-position GOP / GCP bytes into
        out_gop / out_gcp (uint8 device planes) from a ``gatk_amd.dragstr.DragstrParams``."""
        import numpy as np

        for t, dt, n in ((read_bases, torch.uint8, "read_bases"), (read_off, torch.int64, "read_off"),
                         (read_len, torch.int32, "read_len"), (out_gop, torch.uint8, "out_gop"),
                         (out_gcp, torch.uint8, "out_gcp")):
            _check(t, dt, n)
        gop = np.ascontiguousarray(params.gop_table, dtype=np.float64)
        gcp = np.ascontiguousarray(params.gcp_table, dtype=np.float64)
        if max_read_len is None:
            max_read_len = int(read_len.max().item()) if read_len.numel() else 0
        _raise_for(self.lib.pairhmm_dragstr_impute_device(
            self.ctx, gop.ctypes.data, gcp.ctypes.data, int(gop.shape[0]), int(gop.shape[1]), _ptr(read_bases),
            _ptr(read_off), _ptr(read_len), int(read_off.numel()), int(max_read_len), _ptr(out_gop), _ptr(out_gcp)),
            "dragstr_impute_device")

    def modify_read_qualities(self, read_bases, read_off, read_len, base_quals, ins_quals, del_quals,
                              model: PCRErrorModel, mapping_quality: Optional[torch.Tensor] = None,
                              base_quality_threshold: int = DEFAULT_BASE_QUALITY_THRESHOLD,
                              disable_cap_to_mapq: bool = False) -> None:
        """modifyReadQualities: the PCR model, then the caps (:296-297)."""
        self.apply_pcr_error_model(read_bases, read_off, read_len, ins_quals, del_quals, model)
        self.cap_read_qualities(read_off, read_len, base_quals, ins_quals, del_quals, mapping_quality,
                                base_quality_threshold, disable_cap_to_mapq)

    def close(self) -> None:
        if self.ctx:
            self.lib.pairhmm_done(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
