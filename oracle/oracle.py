"""Python handle on the FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY.

This module is the checker for the gfx950 PairHMM engine in ``gatk_amd``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import it.  The
product path never does.

* ``liboracle.so`` (``pairhmm_oracle.c``) restates GATK's FP64 Java path:
  ``LoglessPairHMM`` (LoglessPairHMM.java:20-93), ``Log10PairHMM`` EXACT/ORIGINAL
  (Log10PairHMM.java:58-183), the ``PairHMMModel`` / ``QualityUtils`` / ``JacobianLogTable``
  tables (PairHMMModel.java:71-117, 373-418; QualityUtils.java:43-57; MathUtils.java:406-480),
  plus an FP64 adjoint of the recurrence for continuous Phred parameters, and the partially
  determined ``LoglessPDPairHMM`` (LoglessPDPairHMM.java:34-194); ``sw_oracle.c`` restates
  ``SmithWatermanJavaAligner`` (SmithWatermanJavaAligner.java:104-377); ``simd_proxy.c`` is not a
  restatement but a GKL-class FP32 SIMD CPU PairHMM that only the bench's CPU leg times.
* ``torch_log10_likelihood`` is an independent anti-diagonal restatement of the same continuous
  model in torch FP64, used with ``torch.autograd.gradcheck`` to pin the C adjoint.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lock = threading.Lock()
_lib = None

_u8p = ctypes.POINTER(ctypes.c_uint8)
_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build() -> str:
    """Compile liboracle.so with oracle/Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(_LIB_PATH):
                build()
            L = ctypes.CDLL(_LIB_PATH)
            L.oracle_init.restype = None
            for name in ("oracle_qual_to_error_prob", "oracle_qual_to_prob", "oracle_qual_to_prob_log10"):
                getattr(L, name).argtypes = [ctypes.c_int]
                getattr(L, name).restype = ctypes.c_double
            L.oracle_jacobian_table_get.argtypes = [ctypes.c_double]
            L.oracle_jacobian_table_get.restype = ctypes.c_double
            L.oracle_approx_log10_sum_log10.argtypes = [ctypes.c_double, ctypes.c_double]
            L.oracle_approx_log10_sum_log10.restype = ctypes.c_double
            for name in ("oracle_match_to_match_prob", "oracle_match_to_match_prob_log10"):
                getattr(L, name).argtypes = [ctypes.c_int, ctypes.c_int]
                getattr(L, name).restype = ctypes.c_double
            L.oracle_qual_to_trans_probs.argtypes = [ctypes.c_int8] * 3 + [_dp]
            L.oracle_qual_to_trans_probs.restype = ctypes.c_int
            L.oracle_first_position_haplotypes_differ.argtypes = [_u8p, ctypes.c_int, _u8p, ctypes.c_int]
            L.oracle_first_position_haplotypes_differ.restype = ctypes.c_int
            pair = [_u8p, ctypes.c_int, _u8p, _u8p, _u8p, _u8p, _u8p, ctypes.c_int, ctypes.c_int]
            L.oracle_logless.argtypes = pair + [_dp]
            L.oracle_logless.restype = ctypes.c_int
            L.oracle_pd_logless.argtypes = [_u8p, _u8p, ctypes.c_int, _u8p, _u8p, _u8p, _u8p, _u8p, ctypes.c_int,
                                            ctypes.c_int, _dp]
            L.oracle_pd_logless.restype = ctypes.c_int
            L.oracle_pd_logless_batch.argtypes = [ctypes.c_int32, ctypes.c_int32] + [_u8p] * 5 + [
                _i64p, _i32p, _u8p, _u8p, _i64p, _i32p, ctypes.c_int, ctypes.c_int, _dp]
            L.oracle_pd_logless_batch.restype = ctypes.c_int
            L.oracle_sw_align.argtypes = [_u8p, ctypes.c_int, _u8p, ctypes.c_int] + [ctypes.c_int] * 5 + [
                ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int), _i32p, _i32p]
            L.oracle_sw_align.restype = ctypes.c_int
            L.oracle_simd_proxy_batch.argtypes = [ctypes.c_int64, _i32p, _i32p, _u8p, _u8p, _u8p, _u8p, _u8p,
                                                  _i64p, _i32p, _u8p, _i64p, _i32p, ctypes.c_int32, ctypes.c_int,
                                                  ctypes.c_int, _dp]
            L.oracle_simd_proxy_batch.restype = ctypes.c_long
            L.oracle_log10.argtypes = pair + [ctypes.c_int, _dp]
            L.oracle_log10.restype = ctypes.c_int
            L.oracle_logless_batch.argtypes = [ctypes.c_int64, _i32p, _i32p, _u8p, _u8p, _u8p, _u8p, _u8p,
                                               _i64p, _i32p, _u8p, _i64p, _i32p, ctypes.c_int, ctypes.c_int, _dp]
            L.oracle_logless_batch.restype = ctypes.c_int
            fl = [_u8p, ctypes.c_int, _u8p, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int]
            L.oracle_logless_float.argtypes = fl + [_dp]
            L.oracle_logless_float.restype = ctypes.c_int
            L.oracle_logless_float_grad.argtypes = fl + [_dp] * 6
            L.oracle_logless_float_grad.restype = ctypes.c_int
            L.oracle_init()
            _lib = L
    return _lib


def _bytes(a) -> np.ndarray:
    if isinstance(a, (bytes, bytearray)):
        return np.frombuffer(bytes(a), dtype=np.uint8).copy()
    if isinstance(a, str):
        return np.frombuffer(a.encode("latin-1"), dtype=np.uint8).copy()
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint8))


def _p(a: np.ndarray, t=_u8p):
    return a.ctypes.data_as(t)


class OracleError(ValueError):
    """Raised where the Java path throws IllegalArgumentException."""


def qual_to_error_prob(q: int) -> float:
    return lib().oracle_qual_to_error_prob(int(q))


def qual_to_prob(q: int) -> float:
    return lib().oracle_qual_to_prob(int(q))


def qual_to_prob_log10(q: int) -> float:
    return lib().oracle_qual_to_prob_log10(int(q))


def jacobian_table_get(diff: float) -> float:
    return lib().oracle_jacobian_table_get(float(diff))


def approximate_log10_sum_log10(a: float, b: float) -> float:
    return lib().oracle_approx_log10_sum_log10(float(a), float(b))


def match_to_match_prob(ins_qual: int, del_qual: int) -> float:
    return lib().oracle_match_to_match_prob(int(ins_qual), int(del_qual))


def match_to_match_prob_log10(ins_qual: int, del_qual: int) -> float:
    return lib().oracle_match_to_match_prob_log10(int(ins_qual), int(del_qual))


def qual_to_trans_probs(ins: int, dele: int, gcp: int) -> np.ndarray:
    out = np.zeros(6)
    st = lib().oracle_qual_to_trans_probs(ctypes.c_int8(np.int8(np.uint8(ins & 0xFF))),
                                          ctypes.c_int8(np.int8(np.uint8(dele & 0xFF))),
                                          ctypes.c_int8(np.int8(np.uint8(gcp & 0xFF))), _p(out, _dp))
    if st != 0:
        raise OracleError("quality cannot be negative")
    return out


def first_position_haplotypes_differ(h1, h2) -> int:
    a, b = _bytes(h1), _bytes(h2)
    r = lib().oracle_first_position_haplotypes_differ(_p(a), len(a), _p(b), len(b))
    if r < 0:
        raise OracleError("haplotype is bad")
    return r


def _pair_args(hap, read, quals, ins, dels, gcp):
    arrs = [_bytes(x) for x in (hap, read, quals, ins, dels, gcp)]
    R = len(arrs[1])
    for a in arrs[2:]:
        if len(a) != R:
            raise OracleError("read arrays differ in length")
    return arrs, R


def logless(hap, read, quals, ins, dels, gcp, tristate: bool = True) -> float:
    """LoglessPairHMM log10 likelihood (FP64, Java arithmetic order)."""
    (h, r, q, i, d, g), R = _pair_args(hap, read, quals, ins, dels, gcp)
    out = ctypes.c_double()
    st = lib().oracle_logless(_p(h), len(h), _p(r), _p(q), _p(i), _p(d), _p(g), R, int(tristate),
                              ctypes.byref(out))
    if st != 0:
        raise OracleError(f"oracle_logless status {st}")
    return out.value


def pd_logless(hap, hap_pd, read, quals, ins, dels, gcp, tristate: bool = True) -> float:
    """LoglessPDPairHMM log10 likelihood (FP64, Java arithmetic order)."""
    (h, r, q, i, d, g), R = _pair_args(hap, read, quals, ins, dels, gcp)
    pd = _bytes(hap_pd)
    if len(pd) != len(h):
        raise OracleError("hap_pd must have one byte per haplotype base")
    out = ctypes.c_double()
    st = lib().oracle_pd_logless(_p(h), _p(pd), len(h), _p(r), _p(q), _p(i), _p(d), _p(g), R, int(tristate),
                                 ctypes.byref(out))
    if st != 0:
        raise OracleError(f"oracle_pd_logless status {st}")
    return out.value


def pd_logless_batch(reads, haps, n_threads: int = 1, tristate: bool = True) -> np.ndarray:
    """Every read against every PD haplotype, read-major [n_reads, n_haps].

    ``reads``: sequence of (bases, quals, ins, dels, gcp); ``haps``: sequence of (bases, pd_bytes)."""
    rb = [tuple(_bytes(x) for x in r) for r in reads]
    hb = [(_bytes(h), _bytes(pd)) for h, pd in haps]
    rl = np.array([len(r[0]) for r in rb], dtype=np.int32)
    ro = np.zeros(len(rb), dtype=np.int64)
    if len(rb) > 1:
        ro[1:] = np.cumsum(rl[:-1])
    hl = np.array([len(h) for h, _ in hb], dtype=np.int32)
    ho = np.zeros(len(hb), dtype=np.int64)
    if len(hb) > 1:
        ho[1:] = np.cumsum(hl[:-1])
    planes = [np.concatenate([r[k] for r in rb]) if rb else np.zeros(0, np.uint8) for k in range(5)]
    planes = [np.ascontiguousarray(p, dtype=np.uint8) for p in planes]
    hcat = np.ascontiguousarray(np.concatenate([h for h, _ in hb]), dtype=np.uint8)
    pcat = np.ascontiguousarray(np.concatenate([pd for _, pd in hb]), dtype=np.uint8)
    out = np.empty(len(rb) * len(hb), dtype=np.float64)
    st = lib().oracle_pd_logless_batch(len(rb), len(hb), *[_p(p) for p in planes], _p(ro, _i64p), _p(rl, _i32p),
                                       _p(hcat), _p(pcat), _p(ho, _i64p), _p(hl, _i32p), int(tristate),
                                       int(n_threads), _p(out, _dp))
    if st != 0:
        raise OracleError(f"oracle_pd_logless_batch status {st}")
    return out.reshape(len(rb), len(hb))


SW_STRATEGIES = {"SOFTCLIP": 0, "INDEL": 1, "LEADING_INDEL": 2, "IGNORE": 3}


def sw_align(ref, alt, params, strategy: str, matrices: bool = False):
    """SmithWatermanJavaAligner's full alignment (no substring shortcut): (cigar text, offset), plus the
    score and backtrack matrices ([R+1, A+1] int32) when ``matrices``.  ``params`` is SWParameters'
    (match, mismatch, open, extend)."""
    r, a = _bytes(ref), _bytes(alt)
    cap = 16 * (len(r) + len(a) + 8)
    buf = ctypes.create_string_buffer(cap)
    off = ctypes.c_int()
    sc = bt = None
    if matrices:
        sc = np.zeros((len(r) + 1, len(a) + 1), np.int32)
        bt = np.zeros_like(sc)
    st = lib().oracle_sw_align(_p(r), len(r), _p(a), len(a), *[int(x) for x in params], SW_STRATEGIES[strategy], buf,
                               cap, ctypes.byref(off), _p(sc, _i32p) if matrices else None,
                               _p(bt, _i32p) if matrices else None)
    if st != 0:
        raise OracleError(f"oracle_sw_align status {st}")
    res = (buf.value.decode("ascii"), off.value)
    return res + (sc, bt) if matrices else res


def log10_pairhmm(hap, read, quals, ins, dels, gcp, exact: bool, tristate: bool = True) -> float:
    """Log10PairHMM log10 likelihood: EXACT (exact=True) or ORIGINAL."""
    (h, r, q, i, d, g), R = _pair_args(hap, read, quals, ins, dels, gcp)
    out = ctypes.c_double()
    st = lib().oracle_log10(_p(h), len(h), _p(r), _p(q), _p(i), _p(d), _p(g), R, int(tristate), int(exact),
                            ctypes.byref(out))
    if st != 0:
        raise OracleError(f"oracle_log10 status {st}")
    return out.value


def logless_batch(batch, n_threads: int = 1, tristate: bool = True, pairs=None) -> np.ndarray:
    """LoglessPairHMM over a packed batch (``gatk_amd.batch.PackedBatch`` layout).

    ``pairs`` optionally selects a subset of pair indices; results come back in that order.
    """
    pr = np.ascontiguousarray(batch.pair_read, dtype=np.int32)
    ph = np.ascontiguousarray(batch.pair_hap, dtype=np.int32)
    if pairs is not None:
        pairs = np.asarray(pairs)
        pr, ph = np.ascontiguousarray(pr[pairs]), np.ascontiguousarray(ph[pairs])
    out = np.empty(len(pr), dtype=np.float64)
    ro = np.ascontiguousarray(batch.read_off, dtype=np.int64)
    rl = np.ascontiguousarray(batch.read_len, dtype=np.int32)
    ho = np.ascontiguousarray(batch.hap_off, dtype=np.int64)
    hl = np.ascontiguousarray(batch.hap_len, dtype=np.int32)
    st = lib().oracle_logless_batch(len(pr), _p(pr, _i32p), _p(ph, _i32p), _p(batch.read_bases), _p(batch.read_quals),
                                    _p(batch.read_ins), _p(batch.read_del), _p(batch.read_gcp), _p(ro, _i64p),
                                    _p(rl, _i32p), _p(batch.hap_bases), _p(ho, _i64p), _p(hl, _i32p), int(tristate),
                                    int(n_threads), _p(out, _dp))
    if st != 0:
        raise OracleError(f"oracle_logless_batch status {st}")
    return out


def simd_proxy_batch(batch, n_threads: int = 1, tristate: bool = True, pairs=None):
    """The GKL-class CPU proxy (simd_proxy.c: FP32 SIMD lanes over reads, FP64 recompute under 1e-28)
    over a packed batch; returns (log10 L per selected pair, pairs recomputed in FP64)."""
    pr = np.ascontiguousarray(batch.pair_read, dtype=np.int32)
    ph = np.ascontiguousarray(batch.pair_hap, dtype=np.int32)
    if pairs is not None:
        pairs = np.asarray(pairs)
        pr, ph = np.ascontiguousarray(pr[pairs]), np.ascontiguousarray(ph[pairs])
    out = np.empty(len(pr), dtype=np.float64)
    ro = np.ascontiguousarray(batch.read_off, dtype=np.int64)
    rl = np.ascontiguousarray(batch.read_len, dtype=np.int32)
    ho = np.ascontiguousarray(batch.hap_off, dtype=np.int64)
    hl = np.ascontiguousarray(batch.hap_len, dtype=np.int32)
    n = lib().oracle_simd_proxy_batch(len(pr), _p(pr, _i32p), _p(ph, _i32p), _p(batch.read_bases),
                                      _p(batch.read_quals), _p(batch.read_ins), _p(batch.read_del),
                                      _p(batch.read_gcp), _p(ro, _i64p), _p(rl, _i32p), _p(batch.hap_bases),
                                      _p(ho, _i64p), _p(hl, _i32p), len(hl), int(tristate), int(n_threads),
                                      _p(out, _dp))
    if n < 0:
        raise OracleError("oracle_simd_proxy_batch: out of memory")
    return out, int(n)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def logless_float(hap, read, qb, qi, qd, qc, tristate: bool = True) -> float:
    """Continuous-Phred model (MM = 1 - e_ins - e_del) forward, log10 L."""
    h, r = _bytes(hap), _bytes(read)
    qb, qi, qd, qc = (_f64(x) for x in (qb, qi, qd, qc))
    out = ctypes.c_double()
    st = lib().oracle_logless_float(_p(h), len(h), _p(r), len(r), _p(qb, _dp), _p(qi, _dp), _p(qd, _dp), _p(qc, _dp),
                                    int(tristate), ctypes.byref(out))
    if st != 0:
        raise OracleError(f"oracle_logless_float status {st}")
    return out.value


def logless_float_grad(hap, read, qb, qi, qd, qc, tristate: bool = True) -> dict:
    """FP64 adjoint: log10 L and d log10 L / d(qi, qd, qc Phred; base error prob; qb Phred)."""
    h, r = _bytes(hap), _bytes(read)
    R = len(r)
    qb, qi, qd, qc = (_f64(x) for x in (qb, qi, qd, qc))
    outs = {k: np.zeros(R) for k in ("qi", "qd", "qc", "eb", "qb")}
    ll = ctypes.c_double()
    st = lib().oracle_logless_float_grad(_p(h), len(h), _p(r), R, _p(qb, _dp), _p(qi, _dp), _p(qd, _dp), _p(qc, _dp),
                                         int(tristate), ctypes.byref(ll), *(_p(outs[k], _dp) for k in
                                                                            ("qi", "qd", "qc", "eb", "qb")))
    if st != 0:
        raise OracleError(f"oracle_logless_float_grad status {st}")
    outs["log10_l"] = ll.value
    return outs


def torch_log10_likelihood(hap, read, qb, qi, qd, qc, tristate: bool = True):
    """Torch FP64 restatement of the continuous model, swept by anti-diagonals (autograd-able).

    Cell (i, j) lies on anti-diagonal d = i + j.  M(i,j) reads diagonal d-2, I and D read d-1
    (LoglessPairHMM.java:51-55).  Row 0 holds M = I = 0 and D = 2^1020 / H (:30-36); column 0
    holds zeros.  Returns log10(sum_j M(R,j) + I(R,j)) - log10(2^1020) (:62-67).
    """
    import torch

    h, r = _bytes(hap), _bytes(read)
    H, R = len(h), len(r)
    dt = torch.float64
    ten = 10.0
    e = lambda q: torch.pow(torch.tensor(ten, dtype=dt), -q / 10.0)  # noqa: E731
    eb, ei, ed, ec = e(qb), e(qi), e(qd), e(qc)
    z1 = torch.zeros(1, dtype=dt)
    # parameters indexed by row i = 0..R (row 0 unused)
    MM = torch.cat([z1, 1.0 - ei - ed])
    MI = torch.cat([z1, ei])
    MD = torch.cat([z1, ed])
    GM = torch.cat([z1, 1.0 - ec])
    GG = torch.cat([z1, ec])
    PM = torch.cat([z1, 1.0 - eb])
    PX = torch.cat([z1, eb / (3.0 if tristate else 1.0)])
    match = (r[:, None] == h[None, :]) | (r[:, None] == ord("N")) | (h[None, :] == ord("N"))
    match_t = torch.zeros((R + 1, H + 1), dtype=torch.bool)
    match_t[1:, 1:] = torch.from_numpy(match)
    ii = torch.arange(R + 1)
    init = 2.0 ** 1020 / H
    zeros = torch.zeros(R + 1, dtype=dt)

    def shift(x):  # x[i-1] at position i
        return torch.cat([z1, x[:-1]])

    Mm2 = Im2 = Dm2 = zeros  # diagonal d-2
    # diagonal 0 : only (0,0): D = init
    D0 = zeros.clone()
    D0[0] = init
    Mm1, Im1, Dm1 = zeros, zeros, D0  # diagonal d-1 (= diagonal 0)
    total = torch.zeros((), dtype=dt)
    for d in range(1, R + H + 1):
        j = d - ii
        valid = (ii >= 1) & (j >= 1) & (j <= H)
        jc = j.clamp(0, H)
        p = torch.where(match_t[ii, jc], PM, PX)
        M = p * (MM * shift(Mm2) + GM * (shift(Im2) + shift(Dm2)))
        I = MI * shift(Mm1) + GG * shift(Im1)
        D = MD * Mm1 + GG * Dm1
        vm = valid.to(dt)
        M, I, D = M * vm, I * vm, D * vm
        if d <= H:  # row 0 cell (0, d)
            row0 = torch.zeros(R + 1, dtype=dt)
            row0[0] = init
            D = D + row0
        if 1 <= d - R <= H:
            total = total + M[R] + I[R]
        Mm2, Im2, Dm2 = Mm1, Im1, Dm1
        Mm1, Im1, Dm1 = M, I, D
    return torch.log10(total) - 1020.0 * np.log10(2.0)
