"""ctypes binding of libgatk_pairhmm.so (the C-ABI declared in include/gatk_pairhmm.h).

The library is built in-tree by ``gatk_amd/csrc/Makefile`` (``__graft_entry__.build()``).  There is
no CPU fallback: if the library is missing, loading raises ``NativeLibraryMissing``.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgatk_pairhmm.so")

PAIRHMM_OK = 0
PAIRHMM_ERR_NO_DEVICE = 1
PAIRHMM_ERR_BAD_ARG = 2
PAIRHMM_ERR_OOM = 3
PAIRHMM_ERR_KERNEL = 4
PAIRHMM_ERR_STATE = 5

RUN_PERSISTENT = 0
RUN_PERSISTENT_EAGER = 1
RUN_FP64 = 2

# Every symbol include/gatk_pairhmm.h declares.
EXPORTED_SYMBOLS = (
    "pairhmm_load",
    "pairhmm_initialize",
    "pairhmm_compute_likelihoods",
    "pairhmm_done",
    "pairhmm_last_error",
    "pairhmm_batch_upload",
    "pairhmm_batch_forward",
    "pairhmm_batch_free",
    "pairhmm_forward_backward",
    "pairhmm_forward_backward_noalloc",
    "pairhmm_set_tuning",
    "pairhmm_get_tuning",
    "pairhmm_stream_create",
    "pairhmm_stream_submit",
    "pairhmm_stream_finish",
    "pairhmm_stream_destroy",
    "pairhmm_dragstr_analyze",
    "pairhmm_dragstr_contexts",
    "pairhmm_dragstr_impute",
    "pairhmm_apply_pcr_error_model",
    "pairhmm_cap_read_qualities",
    "pairhmm_dragstr_impute_device",
    "pairhmm_pd_compute_likelihoods",
    "pairhmm_normalize_and_filter_likelihoods",
    "pairhmm_sw_align",
    "pairhmm_sw_align_batch",
    "pairhmm_sw_align_batch_packed",
)


class NativeLibraryMissing(RuntimeError):
    pass


class ReadDataHolder(ctypes.Structure):
    _fields_ = [
        ("read_bases", ctypes.c_void_p),
        ("read_quals", ctypes.c_void_p),
        ("insertion_gop", ctypes.c_void_p),
        ("deletion_gop", ctypes.c_void_p),
        ("overall_gcp", ctypes.c_void_p),
        ("length", ctypes.c_int32),
    ]


class HaplotypeDataHolder(ctypes.Structure):
    _fields_ = [("haplotype_bases", ctypes.c_void_p), ("length", ctypes.c_int32)]


class SWParameters(ctypes.Structure):
    _fields_ = [("match_value", ctypes.c_int32), ("mismatch_penalty", ctypes.c_int32),
                ("gap_open_penalty", ctypes.c_int32), ("gap_extend_penalty", ctypes.c_int32)]


class PDHaplotypeDataHolder(ctypes.Structure):
    _fields_ = [("haplotype_bases", ctypes.c_void_p), ("haplotype_pd_bases", ctypes.c_void_p),
                ("length", ctypes.c_int32)]


class NativeArguments(ctypes.Structure):
    _fields_ = [("max_number_of_threads", ctypes.c_int32), ("use_double_precision", ctypes.c_int32)]


class BatchDesc(ctypes.Structure):
    _fields_ = [
        ("n_reads", ctypes.c_int64),
        ("read_bases", ctypes.c_void_p),
        ("read_quals", ctypes.c_void_p),
        ("read_ins", ctypes.c_void_p),
        ("read_del", ctypes.c_void_p),
        ("read_gcp", ctypes.c_void_p),
        ("read_off", ctypes.c_void_p),
        ("read_len", ctypes.c_void_p),
        ("read_bytes", ctypes.c_int64),
        ("n_haps", ctypes.c_int64),
        ("hap_bases", ctypes.c_void_p),
        ("hap_off", ctypes.c_void_p),
        ("hap_len", ctypes.c_void_p),
        ("hap_bytes", ctypes.c_int64),
        ("n_pairs", ctypes.c_int64),
        ("pair_read", ctypes.c_void_p),
        ("pair_hap", ctypes.c_void_p),
    ]


class GradDesc(ctypes.Structure):
    _fields_ = [
        ("n_reads", ctypes.c_int64),
        ("n_haps", ctypes.c_int64),
        ("n_pairs", ctypes.c_int64),
        ("read_bytes", ctypes.c_int64),
        ("read_bases", ctypes.c_void_p),
        ("read_qb", ctypes.c_void_p),
        ("read_qi", ctypes.c_void_p),
        ("read_qd", ctypes.c_void_p),
        ("read_qc", ctypes.c_void_p),
        ("read_off", ctypes.c_void_p),
        ("read_len", ctypes.c_void_p),
        ("hap_bases", ctypes.c_void_p),
        ("hap_off", ctypes.c_void_p),
        ("hap_len", ctypes.c_void_p),
        ("pair_read", ctypes.c_void_p),
        ("pair_hap", ctypes.c_void_p),
        ("pair_weight", ctypes.c_void_p),
        ("max_read_len", ctypes.c_int32),
        ("max_hap_len", ctypes.c_int32),
        ("use_double", ctypes.c_int32),
        ("out_log10_likelihood", ctypes.c_void_p),
        ("grad_qb", ctypes.c_void_p),
        ("grad_qi", ctypes.c_void_p),
        ("grad_qd", ctypes.c_void_p),
        ("grad_qc", ctypes.c_void_p),
        ("grad_base_error", ctypes.c_void_p),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [("waves_per_block", ctypes.c_int32), ("grid_percent", ctypes.c_int32),
                ("queue_order", ctypes.c_int32), ("stripe_schedule", ctypes.c_int32),
                ("step_unroll", ctypes.c_int32)]


class RunStats(ctypes.Structure):
    _fields_ = [
        ("total_ms", ctypes.c_double),
        ("fp32_kernel_ms", ctypes.c_double),
        ("rescue_kernel_ms", ctypes.c_double),
        ("n_rescued", ctypes.c_int64),
        ("cells", ctypes.c_int64),
        ("n_routed_fp64", ctypes.c_int64),
    ]


_lock = threading.Lock()
_lib = None


def load_library() -> ctypes.CDLL:
    """Load the in-tree native library (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        # PyTorch's ROCm wheel bundles its own HIP runtime (torch/lib/libamdhip64.so).  If torch is
        # present it is imported first, so that the library binds to that same runtime instead of
        # loading a second HIP runtime into the process (torch then fails to see the GPU).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                f"{LIB_PATH} is missing; build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        L = ctypes.CDLL(LIB_PATH)
        declare(L)
        _lib = L
        return L


def declare(L: ctypes.CDLL) -> ctypes.CDLL:
    """Sets the argument and result types of every C-ABI entry point on a loaded library."""
    vp, i32 = ctypes.c_void_p, ctypes.c_int32
    L.pairhmm_load.argtypes = []
    L.pairhmm_load.restype = i32
    L.pairhmm_initialize.argtypes = [ctypes.POINTER(NativeArguments), ctypes.POINTER(vp)]
    L.pairhmm_initialize.restype = i32
    L.pairhmm_compute_likelihoods.argtypes = [vp, ctypes.POINTER(ReadDataHolder), i32,
                                              ctypes.POINTER(HaplotypeDataHolder), i32, vp]
    L.pairhmm_compute_likelihoods.restype = i32
    L.pairhmm_done.argtypes = [vp]
    L.pairhmm_done.restype = i32
    L.pairhmm_last_error.argtypes = []
    L.pairhmm_last_error.restype = ctypes.c_char_p
    L.pairhmm_batch_upload.argtypes = [vp, ctypes.POINTER(BatchDesc), ctypes.POINTER(vp)]
    L.pairhmm_batch_upload.restype = i32
    L.pairhmm_batch_forward.argtypes = [vp, vp, i32, vp, ctypes.POINTER(RunStats)]
    L.pairhmm_batch_forward.restype = i32
    L.pairhmm_batch_free.argtypes = [vp]
    L.pairhmm_batch_free.restype = i32
    L.pairhmm_forward_backward.argtypes = [vp, ctypes.POINTER(GradDesc), ctypes.POINTER(RunStats)]
    L.pairhmm_forward_backward.restype = i32
    L.pairhmm_forward_backward_noalloc.argtypes = [vp, ctypes.POINTER(GradDesc), ctypes.POINTER(RunStats)]
    L.pairhmm_forward_backward_noalloc.restype = i32
    L.pairhmm_set_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    L.pairhmm_set_tuning.restype = i32
    L.pairhmm_get_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    L.pairhmm_get_tuning.restype = i32
    L.pairhmm_stream_create.argtypes = [vp, ctypes.POINTER(vp)]
    L.pairhmm_stream_create.restype = i32
    L.pairhmm_stream_submit.argtypes = [vp, ctypes.POINTER(BatchDesc), vp]
    L.pairhmm_stream_submit.restype = i32
    L.pairhmm_stream_finish.argtypes = [vp, ctypes.POINTER(RunStats)]
    L.pairhmm_stream_finish.restype = i32
    L.pairhmm_stream_destroy.argtypes = [vp]
    L.pairhmm_stream_destroy.restype = i32
    i64 = ctypes.c_int64
    L.pairhmm_dragstr_analyze.argtypes = [vp, i32, i32, vp, vp]
    L.pairhmm_dragstr_analyze.restype = i32
    L.pairhmm_dragstr_contexts.argtypes = [vp, vp, vp, i64, i64, i32, i32, vp]
    L.pairhmm_dragstr_contexts.restype = i32
    L.pairhmm_dragstr_impute.argtypes = [vp, vp, i32, i32, vp, vp, vp, i64, i64, vp, vp]
    L.pairhmm_dragstr_impute.restype = i32
    L.pairhmm_apply_pcr_error_model.argtypes = [vp, vp, vp, vp, i64, i32, vp, vp]
    L.pairhmm_apply_pcr_error_model.restype = i32
    L.pairhmm_cap_read_qualities.argtypes = [vp, vp, vp, i64, vp, i32, i32, vp, vp, vp]
    L.pairhmm_cap_read_qualities.restype = i32
    L.pairhmm_dragstr_impute_device.argtypes = [vp, vp, vp, i32, i32, vp, vp, vp, i64, i32, vp, vp]
    L.pairhmm_dragstr_impute_device.restype = i32
    L.pairhmm_pd_compute_likelihoods.argtypes = [vp, ctypes.POINTER(ReadDataHolder), i32,
                                                 ctypes.POINTER(PDHaplotypeDataHolder), i32, vp]
    L.pairhmm_pd_compute_likelihoods.restype = i32
    dbl = ctypes.c_double
    L.pairhmm_normalize_and_filter_likelihoods.argtypes = [vp, vp, i64, i32, i32, dbl, i32, vp, vp, vp, dbl,
                                                           i32, dbl, vp, vp]
    L.pairhmm_normalize_and_filter_likelihoods.restype = i32
    L.pairhmm_sw_align.argtypes = [vp, vp, i32, vp, i32, ctypes.POINTER(SWParameters), i32, ctypes.c_char_p,
                                   i32, ctypes.POINTER(i32)]
    L.pairhmm_sw_align.restype = i32
    L.pairhmm_sw_align_batch.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, ctypes.POINTER(SWParameters), i32, vp,
                                         vp, vp, vp]
    L.pairhmm_sw_align_batch.restype = i32
    L.pairhmm_sw_align_batch_packed.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, ctypes.POINTER(SWParameters), i32,
                                                vp, i64, vp, vp]
    L.pairhmm_sw_align_batch_packed.restype = i32
    return L


def last_error() -> str:
    msg = load_library().pairhmm_last_error()
    return msg.decode("utf-8", "replace") if msg else ""
