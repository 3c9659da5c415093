"""torch.autograd wrapper of the gfx950 forward + backward (adjoint) kernel.

``pairhmm_log10_likelihood(problem, qb, qi, qd, qc)`` returns log10 L for every pair of a
``GradProblem`` as a differentiable function of per-read-position Phred parameters (float tensors
of length read_bytes on the engine's device):
    qb -- base quality, qi -- insertion gap-open, qd -- deletion gap-open, qc -- gap continuation.
The model is LoglessPairHMM's recurrence (LoglessPairHMM.java:47-57) with the continuous transition
probabilities MM = 1 - e(qi) - e(qd), MI = e(qi), MD = e(qd), GM = 1 - e(qc), II = DD = e(qc) and
priors 1 - e(qb) / e(qb)/3, e(q) = 10^(-q/10).

The kernel evaluates log10 L and sum_p w_p d log10 L_p / dq in one call.  The forward call runs it
with w = 1 and keeps the gradient; ``backward`` reuses it when grad_output is all ones and runs the
kernel again with w = grad_output otherwise.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _native
from .batch import PackedBatch
from .pairhmm import HardwareFeatureException, PairHMMNativeArguments, _raise_for


@dataclass
class GradProblem:
    """Reads, haplotypes and pairs resident on the device (torch tensors)."""
    read_bases: torch.Tensor  # uint8 [read_bytes]
    read_off: torch.Tensor    # int64 [n_reads]
    read_len: torch.Tensor    # int32 [n_reads]
    hap_bases: torch.Tensor   # uint8
    hap_off: torch.Tensor     # int64 [n_haps]
    hap_len: torch.Tensor     # int32 [n_haps]
    pair_read: torch.Tensor   # int32 [n_pairs]
    pair_hap: torch.Tensor    # int32 [n_pairs]
    max_read_len: int
    max_hap_len: int

    @property
    def n_pairs(self) -> int:
        return int(self.pair_read.numel())

    @property
    def read_bytes(self) -> int:
        return int(self.read_bases.numel())

    @classmethod
    def from_batch(cls, batch: PackedBatch, device="cuda") -> "GradProblem":
        t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(device)  # noqa: E731
        return cls(read_bases=t(batch.read_bases, np.uint8), read_off=t(batch.read_off, np.int64),
                   read_len=t(batch.read_len, np.int32), hap_bases=t(batch.hap_bases, np.uint8),
                   hap_off=t(batch.hap_off, np.int64), hap_len=t(batch.hap_len, np.int32),
                   pair_read=t(batch.pair_read, np.int32), pair_hap=t(batch.pair_hap, np.int32),
                   max_read_len=int(batch.read_len.max()) if len(batch.read_len) else 0,
                   max_hap_len=int(batch.hap_len.max()))


class GradEngine:
    """One engine context for forward + backward calls."""

    def __init__(self, use_double: bool = False, no_alloc: bool = False) -> None:
        self.lib = _native.load_library()
        if not self.lib.pairhmm_load():
            raise HardwareFeatureException("no gfx950 device visible to this process")
        na = _native.NativeArguments(1, 0)
        self.ctx = ctypes.c_void_p()
        _raise_for(self.lib.pairhmm_initialize(ctypes.byref(na), ctypes.byref(self.ctx)), "initialize")
        self.use_double = use_double
        # no_alloc: pairhmm_forward_backward_noalloc (fixed device scratch, no allocation per call)
        self.no_alloc = no_alloc
        self.last_stats = None

    def run(self, p: GradProblem, qb, qi, qd, qc, weight: Optional[torch.Tensor] = None, want_eb: bool = False):
        """Returns (log10 L [n_pairs], dict of per-position gradients) as float64 device tensors."""
        dev = p.read_bases.device
        f32 = lambda x: x.detach().to(device=dev, dtype=torch.float32).contiguous()  # noqa: E731
        qb, qi, qd, qc = f32(qb), f32(qi), f32(qd), f32(qc)
        for name, x in (("qb", qb), ("qi", qi), ("qd", qd), ("qc", qc)):
            if x.numel() != p.read_bytes:
                raise ValueError(f"{name} must have one value per read position ({p.read_bytes})")
        out = torch.empty(p.n_pairs, dtype=torch.float64, device=dev)
        grads = {k: torch.empty(p.read_bytes, dtype=torch.float64, device=dev) for k in ("qb", "qi", "qd", "qc")}
        if want_eb:
            grads["eb"] = torch.empty(p.read_bytes, dtype=torch.float64, device=dev)
        w = None
        if weight is not None:
            w = weight.detach().to(device=dev, dtype=torch.float64).contiguous()
        d = _native.GradDesc()
        d.n_reads, d.n_haps, d.n_pairs = int(p.read_len.numel()), int(p.hap_len.numel()), p.n_pairs
        d.read_bytes = p.read_bytes
        d.read_bases = p.read_bases.data_ptr()
        d.read_qb, d.read_qi, d.read_qd, d.read_qc = qb.data_ptr(), qi.data_ptr(), qd.data_ptr(), qc.data_ptr()
        d.read_off, d.read_len = p.read_off.data_ptr(), p.read_len.data_ptr()
        d.hap_bases, d.hap_off, d.hap_len = p.hap_bases.data_ptr(), p.hap_off.data_ptr(), p.hap_len.data_ptr()
        d.pair_read, d.pair_hap = p.pair_read.data_ptr(), p.pair_hap.data_ptr()
        d.pair_weight = w.data_ptr() if w is not None else None
        d.max_read_len, d.max_hap_len = int(p.max_read_len), int(p.max_hap_len)
        d.use_double = int(self.use_double)
        d.out_log10_likelihood = out.data_ptr()
        d.grad_qb, d.grad_qi, d.grad_qd, d.grad_qc = (grads[k].data_ptr() for k in ("qb", "qi", "qd", "qc"))
        d.grad_base_error = grads["eb"].data_ptr() if want_eb else None
        st = _native.RunStats()
        torch.cuda.synchronize(dev)  # inputs were written on torch's stream; the engine uses its own
        fn = self.lib.pairhmm_forward_backward_noalloc if self.no_alloc else self.lib.pairhmm_forward_backward
        _raise_for(fn(self.ctx, ctypes.byref(d), ctypes.byref(st)), "forward_backward")
        self.last_stats = st
        return out, grads

    def close(self) -> None:
        if self.ctx:
            self.lib.pairhmm_done(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class _PairHMMLog10Likelihood(torch.autograd.Function):
    @staticmethod
    def forward(ctx, qb, qi, qd, qc, problem, engine):
        out, grads = engine.run(problem, qb, qi, qd, qc)
        ctx.problem, ctx.engine = problem, engine
        ctx.grads = grads
        ctx.save_for_backward(qb, qi, qd, qc)
        return out.to(qb.dtype) if qb.dtype == torch.float64 else out

    @staticmethod
    def backward(ctx, grad_out):
        qb, qi, qd, qc = ctx.saved_tensors
        g = grad_out.to(torch.float64)
        if bool(torch.all(g == 1.0)):
            grads = ctx.grads
        else:
            _, grads = ctx.engine.run(ctx.problem, qb, qi, qd, qc, weight=g)
        cast = lambda x, like: x.to(like.dtype)  # noqa: E731
        return cast(grads["qb"], qb), cast(grads["qi"], qi), cast(grads["qd"], qd), cast(grads["qc"], qc), None, None


def pairhmm_log10_likelihood(problem: GradProblem, qb, qi, qd, qc, engine: GradEngine):
    """Differentiable log10 L per pair (see module docstring)."""
    return _PairHMMLog10Likelihood.apply(qb, qi, qd, qc, problem, engine)
