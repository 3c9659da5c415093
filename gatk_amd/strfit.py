"""Fit of per-(STR period, repeat count) gap penalties with the adjoint kernel.

The DRAGstr model gives every read position a context -- the best STR period p (1..6) and its copy
count k (1..20) (DragstrReadSTRAnalyzer.java:69-139) -- and reads its gap-open (GOP, used for
insertions and deletions alike) and gap-continuation (GCP) penalties from (p, k) tables
(DragstrPairHMMInputScoreImputator.java:53-70).  This module recovers those tables from reads by
maximum likelihood: it maximises sum over reads of log10 L(read | source haplotype) under the
PairHMM with continuous per-position penalties, where the gradient with respect to every table
cell is the sum of the adjoint kernel's per-position gradients over the positions in that context.

``generate`` draws a synthetic workload with known tables: STR-rich haplotypes (random flanks
between tandem repeats of 1-6 bp motifs with 2-20 copies) and reads sampled by walking the
PairHMM's own transitions, so that the maximum-likelihood tables are the ones used to draw them:
  * a read starts in the match state at a uniform haplotype offset;
  * from the match state at row r-1 a deletion opens with e(GOP_{r-1}) and extends with e(GCP_{r-1})
    (MD and DD of row r-1, LoglessPairHMM.java:55), an insertion opens into row r with e(GOP_r)
    (MI of row r, :53); otherwise the next haplotype base is matched (MM, :51);
  * from the insertion state at row r-1 the insertion extends with e(GCP_r) (II, :53);
  * a matched base is read wrong with probability e(q_base), as one of the other three bases
    (the prior of LoglessPairHMM.java:89-90); an inserted base is uniform;
  * row r's penalties come from the context of read position r-1 of the error-free read the walk
    starts from (the haplotype window), and the last read position uses 45 / 10 like the imputer.
A read's realised context can differ from the window's where an error breaks or changes a repeat,
and the first row's entry factor GM_1 (LoglessPairHMM.java:51 with D of row 0, :30-36) has no
counterpart in a draw, so the fit leaves read position 0 out of the gradient.

``fit`` runs L-BFGS over the 2 x 120 Phred values (GOP kept above 4, GCP above 0.2 by a softplus
parameterisation).  Cells that no read position visits keep their starting value; ``FitReport``
compares the recovered cells against the truth within the standard error the data allow.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Dict, Optional, Tuple

import numpy as np

from . import dragstr
from .batch import PackedBatch

MAX_PERIOD = 6
MAX_REPEATS = 20
N_CTX = MAX_PERIOD * MAX_REPEATS
END_GOP, END_GCP = 45.0, 10.0
GOP_MIN, GCP_MIN = 4.0, 0.2
PHRED_SE = 10.0 / math.log(10.0)  # d q / d ln e

_ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)


def true_tables() -> Tuple[np.ndarray, np.ndarray]:
    """Truth for the synthetic fit: the DRAGstr default tables for periods 1-6
    (DragstrParams.java:49-58, :114-117) with GOP capped at 40 as the imputer caps it (:31)."""
    d = dragstr.DragstrParams.default()
    return np.minimum(40.0, d.gop_table[:MAX_PERIOD]).copy(), d.gcp_table[:MAX_PERIOD].copy()


def err(q):
    return np.power(10.0, -np.asarray(q, dtype=np.float64) / 10.0)


@dataclass(frozen=True)
class FitSpec:
    n_regions: int = 2000
    reads_per_region: int = 100
    hap_len: Tuple[int, int] = (300, 450)
    read_len: Tuple[int, int] = (100, 150)
    base_q: Tuple[int, int] = (20, 40)
    flank: Tuple[int, int] = (4, 40)
    motif: Tuple[int, int] = (1, 6)
    repeats: Tuple[int, int] = (2, 20)


@dataclass
class FitWorkload:
    batch: PackedBatch            # one haplotype per region, pairs = (read, its source haplotype)
    context: np.ndarray           # int32 [read_bytes]: realised read contexts (-1 at read ends)
    gop_true: np.ndarray          # [6, 20]
    gcp_true: np.ndarray
    opens: np.ndarray             # [N_CTX, 2]: (gap opens drawn, open opportunities) by window context
    extends: np.ndarray           # [N_CTX, 2]: (extensions drawn, extension opportunities)


def _str_haplotype(rng, n: int, spec: FitSpec) -> np.ndarray:
    parts, total = [], 0
    while total < n:
        f = _ACGT[rng.integers(0, 4, int(rng.integers(spec.flank[0], spec.flank[1] + 1)))]
        p = int(rng.integers(spec.motif[0], spec.motif[1] + 1))
        k = int(rng.integers(spec.repeats[0], spec.repeats[1] + 1))
        t = np.tile(_ACGT[rng.integers(0, 4, p)], k)
        parts += [f, t]
        total += len(f) + len(t)
    return np.concatenate(parts)[:n]


def _packed_reads(mat: np.ndarray, lens: np.ndarray):
    """Row-major [n, Lmax] byte matrix -> concatenated bytes + offsets."""
    mask = np.arange(mat.shape[1])[None, :] < lens[:, None]
    off = np.concatenate([[0], np.cumsum(lens, dtype=np.int64)[:-1]]).astype(np.int64)
    return np.ascontiguousarray(mat[mask]), off, mask


def generate(spec: FitSpec = FitSpec(), seed: int = 3_000_003, gop_true: Optional[np.ndarray] = None,
             gcp_true: Optional[np.ndarray] = None) -> FitWorkload:
    """Synthetic reads drawn from the PairHMM with known (period, repeats) penalties (module docstring)."""
    if gop_true is None or gcp_true is None:
        gop_true, gcp_true = true_tables()
    gop_flat = np.append(np.asarray(gop_true, dtype=np.float64).ravel(), END_GOP)  # index -1 -> end values
    gcp_flat = np.append(np.asarray(gcp_true, dtype=np.float64).ravel(), END_GCP)
    e_open, e_ext = err(gop_flat), err(gcp_flat)
    rng = np.random.Generator(np.random.PCG64(seed))
    A = spec.n_regions
    haps = [_str_haplotype(rng, int(rng.integers(spec.hap_len[0], spec.hap_len[1] + 1)), spec) for _ in range(A)]
    hl = np.array([len(h) for h in haps], dtype=np.int32)
    hoff = np.concatenate([[0], np.cumsum(hl, dtype=np.int64)[:-1]]).astype(np.int64)
    hcat = np.concatenate(haps)
    n = A * spec.reads_per_region
    src = np.repeat(np.arange(A, dtype=np.int32), spec.reads_per_region)
    R = rng.integers(spec.read_len[0], spec.read_len[1] + 1, n).astype(np.int32)
    slack = hl[src] - R - 40
    if np.any(slack < 0):
        raise ValueError("haplotypes must be at least 40 bases longer than reads")
    start = (rng.random(n) * (slack + 1)).astype(np.int64)
    Rmax = int(R.max())
    # error-free window reads and their contexts (row r uses window position r-1)
    widx = hoff[src][:, None] + start[:, None] + np.arange(Rmax)[None, :]
    win = hcat[np.minimum(widx, len(hcat) - 1)]
    wbytes, woff, wmask = _packed_reads(win, R)

    class _W:
        read_bases, read_off, read_len = wbytes, woff, R

    wctx_flat = dragstr.contexts(_W, MAX_PERIOD, MAX_REPEATS)
    wctx = np.full((n, Rmax), -1, dtype=np.int32)
    wctx[wmask] = wctx_flat
    # the walk
    quals = rng.integers(spec.base_q[0], spec.base_q[1] + 1, (n, Rmax)).astype(np.uint8)
    out = np.zeros((n, Rmax), dtype=np.uint8)
    hend = hoff[src] + hl[src]  # exclusive end in hcat
    col = hoff[src] + start     # next haplotype base (global index)
    in_ins = np.zeros(n, dtype=bool)
    bad = np.zeros(n, dtype=bool)
    opens = np.zeros((N_CTX + 1, 2), dtype=np.int64)
    extends = np.zeros((N_CTX + 1, 2), dtype=np.int64)

    def tally(tab, ctx, hit, mask):
        c = np.where(ctx[mask] < 0, N_CTX, ctx[mask])
        np.add.at(tab[:, 1], c, 1)
        np.add.at(tab[:, 0], c, hit[mask].astype(np.int64))

    for r in range(1, Rmax + 1):
        act = r <= R
        base = np.zeros(n, dtype=np.uint8)
        if r == 1:
            match = act.copy()
        else:
            cp, cc = wctx[:, r - 2], wctx[:, r - 1]
            from_m = act & ~in_ins
            from_i = act & in_ins
            u = rng.random(n)
            pd, pi = e_open[cp], e_open[cc]
            dele = from_m & (u < pd)
            ins = from_m & ~dele & (u < pd + pi)
            ext_i = from_i & (rng.random(n) < e_ext[cc])
            # bookkeeping by window context (rows 2..R-1 carry fitted cells; 1 and R are excluded
            # from the fit through the gradient mask and the fixed end values)
            tally(opens, cp, dele, from_m)
            tally(opens, cc, ins, from_m)
            tally(extends, cc, ext_i, from_i)
            if np.any(dele):
                L = rng.geometric(1.0 - e_ext[cp[dele]])
                # D->D extensions (L-1 of them) with one stop, at row r-1's continuation
                np.add.at(extends[:, 0], np.where(cp[dele] < 0, N_CTX, cp[dele]), (L - 1).astype(np.int64))
                np.add.at(extends[:, 1], np.where(cp[dele] < 0, N_CTX, cp[dele]), L.astype(np.int64))
                col[dele] += L
            new_ins = ins | ext_i
            match = act & ~new_ins
            in_ins = act & new_ins
            ib = act & new_ins
            base[ib] = _ACGT[rng.integers(0, 4, int(ib.sum()))]
        over = match & (col >= hend)
        bad |= over
        ok = match & ~over
        hb = hcat[np.minimum(col, len(hcat) - 1)]
        e = err(quals[:, r - 1])
        wrong = ok & (rng.random(n) < e)
        sub = _ACGT[(np.searchsorted(_ACGT, hb) + rng.integers(1, 4, n)) % 4]
        base[ok] = np.where(wrong[ok], sub[ok], hb[ok])
        col[ok] += 1
        out[:, r - 1] = np.where(act, base, out[:, r - 1])
    keep = ~bad
    out, quals, R, src = out[keep], quals[keep], R[keep], src[keep]
    rbytes, roff, _ = _packed_reads(out, R)
    qbytes, _, _ = _packed_reads(quals, R)
    nb = len(rbytes)
    z = np.zeros(nb, dtype=np.uint8)
    n_keep = len(R)
    # regions: reads stay grouped by source haplotype (src is sorted)
    rro = np.searchsorted(src, np.arange(A + 1)).astype(np.int64)
    batch = PackedBatch(
        read_bases=rbytes, read_quals=qbytes, read_ins=z + 45, read_del=z + 45, read_gcp=z + 10,
        read_off=roff, read_len=R.astype(np.int32), hap_bases=hcat, hap_off=hoff, hap_len=hl,
        pair_read=np.arange(n_keep, dtype=np.int32), pair_hap=src.astype(np.int32),
        region_pair_off=rro.copy(), region_read_off=rro, region_hap_off=np.arange(A + 1, dtype=np.int64))
    ctx = dragstr.contexts(batch, MAX_PERIOD, MAX_REPEATS)
    return FitWorkload(batch=batch, context=ctx, gop_true=np.asarray(gop_true, dtype=np.float64),
                       gcp_true=np.asarray(gcp_true, dtype=np.float64), opens=opens[:N_CTX],
                       extends=extends[:N_CTX])


# ---------------------------------------------------------------------------------------------
# the fit


def _softplus_inv(y):
    import torch

    return y + torch.log(-torch.expm1(-y))


@dataclass
class FitResult:
    gop: np.ndarray   # [6, 20] fitted Phred
    gcp: np.ndarray
    loss: float       # final -mean log10 L per read
    evaluations: int
    seconds: float
    cells_per_eval: int


LoglikFn = Callable[..., "object"]  # (qb, qi, qd, qc) -> log10 L per pair (differentiable torch tensor)


def objective(loglik: LoglikFn, context, qb, read_off, n_reads: int):
    """f(gop, gcp) = -mean over reads of log10 L, for [N_CTX] Phred tensors (module docstring).

    ``loglik(qb, qi, qd, qc)`` returns log10 L per pair as a differentiable function of per-position
    Phred tensors (``gatk_amd.autograd.pairhmm_log10_likelihood`` bound to a problem and engine);
    ``context`` holds the table cell per read position (-1 = fixed end values); ``read_off`` marks
    each read's first position, whose penalties enter without gradient.
    """
    import torch

    dev = qb.device
    ctx = torch.as_tensor(np.asarray(context), device=dev).long()
    fixed = ctx < 0
    cidx = ctx.clamp(min=0)
    first = torch.zeros(ctx.numel(), dtype=torch.bool, device=dev)
    first[torch.as_tensor(np.asarray(read_off), device=dev).long()] = True

    def f(gop, gcp):
        qo = torch.where(fixed, torch.full_like(gop[cidx], END_GOP), gop[cidx])
        qc = torch.where(fixed, torch.full_like(gcp[cidx], END_GCP), gcp[cidx])
        qo = torch.where(first, qo.detach(), qo)
        qc = torch.where(first, qc.detach(), qc)
        return -loglik(qb, qo, qo, qc).sum() / n_reads

    return f


def fit(loglik: LoglikFn, context, qb, read_off, n_reads: int, init_gop: float = 30.0, init_gcp: float = 6.0,
        max_iter: int = 60, history: int = 20, tolerance: float = 1e-9, log=None) -> FitResult:
    """L-BFGS fit of the 6 x 20 GOP / GCP tables over ``objective`` (same arguments)."""
    import time

    import torch

    dev = qb.device
    f = objective(loglik, context, qb, read_off, n_reads)
    a_gop = _softplus_inv(torch.full((N_CTX,), init_gop - GOP_MIN, dtype=torch.float64, device=dev))
    a_gcp = _softplus_inv(torch.full((N_CTX,), init_gcp - GCP_MIN, dtype=torch.float64, device=dev))
    a_gop.requires_grad_(True)
    a_gcp.requires_grad_(True)
    opt = torch.optim.LBFGS([a_gop, a_gcp], lr=1.0, max_iter=max_iter, max_eval=int(max_iter * 1.5),
                            history_size=history, tolerance_grad=tolerance, tolerance_change=1e-12,
                            line_search_fn="strong_wolfe")
    n_eval = [0]
    last = [float("nan")]

    def tables():
        return GOP_MIN + torch.nn.functional.softplus(a_gop), GCP_MIN + torch.nn.functional.softplus(a_gcp)

    def closure():
        opt.zero_grad()
        loss = f(*tables())
        loss.backward()
        n_eval[0] += 1
        last[0] = float(loss.detach())
        if log and n_eval[0] % 10 == 0:
            log(f"fit: evaluation {n_eval[0]}, -mean log10 L = {last[0]:.6f}")
        return loss

    t0 = time.perf_counter()
    opt.step(closure)
    if dev.type == "cuda":
        torch.cuda.synchronize(dev)
    secs = time.perf_counter() - t0
    g, c = tables()
    return FitResult(gop=g.detach().cpu().numpy().reshape(MAX_PERIOD, MAX_REPEATS),
                     gcp=c.detach().cpu().numpy().reshape(MAX_PERIOD, MAX_REPEATS), loss=last[0],
                     evaluations=n_eval[0], seconds=secs, cells_per_eval=0)


# ---------------------------------------------------------------------------------------------
# recovery report


@dataclass
class FitReport:
    gop_checked: int          # cells with enough expected gap opens to pin GOP
    gop_recovered: int        # of those, within the tolerance
    gop_rmse: float
    gop_max_z: float          # max |error| / standard error over checked cells
    gcp_checked: int
    gcp_recovered: int
    gcp_rmse: float
    gcp_max_z: float
    worst: Dict[str, object]

    @property
    def ok(self) -> bool:
        return self.gop_checked > 0 and self.gop_recovered == self.gop_checked and \
            self.gcp_recovered == self.gcp_checked

    def as_dict(self) -> Dict[str, object]:
        return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in self.__dict__.items()}


def position_counts(context: np.ndarray, read_off: np.ndarray, read_len: np.ndarray) -> np.ndarray:
    """Read positions per table cell that carry gradient (first and last positions excluded)."""
    ctx = np.asarray(context).copy()
    ctx[np.asarray(read_off)] = -1
    c = ctx[ctx >= 0]
    return np.bincount(c, minlength=N_CTX)[:N_CTX]


def evaluate(res: FitResult, wl: FitWorkload, min_opens: float = 30.0, min_extends: float = 20.0,
             z_tol: float = 3.0, slack: float = 0.5) -> FitReport:
    """Recovered cell <=> |fit - truth| <= z_tol * SE + slack, with SE from the expected event counts:
    GOP: SE = 4.34 / sqrt(expected opens), expected opens = 2 e(GOP) x positions;
    GCP: SE = 4.34 sqrt((1 - e) / (e T)), T = expected continuation trials = opens / (1 - e(GCP)).
    Only cells with >= min_opens expected opens (GOP) or >= min_extends expected extensions (GCP) are
    checked; the rest have too little data to pin them."""
    b = wl.batch
    npos = position_counts(wl.context, b.read_off, b.read_len).astype(np.float64)
    gt, ct = wl.gop_true.ravel(), wl.gcp_true.ravel()
    gf, cf = res.gop.ravel(), res.gcp.ravel()
    eo, ec = err(gt), err(ct)
    exp_opens = 2.0 * eo * npos
    trials = exp_opens / (1.0 - ec)
    exp_ext = ec * trials
    se_g = PHRED_SE / np.sqrt(np.maximum(exp_opens, 1e-300))
    se_c = PHRED_SE * np.sqrt((1.0 - ec) / np.maximum(exp_ext, 1e-300))
    mg = exp_opens >= min_opens
    mc = exp_ext >= min_extends
    dg, dc = np.abs(gf - gt), np.abs(cf - ct)
    okg = dg <= z_tol * se_g + slack
    okc = dc <= z_tol * se_c + slack
    worst = {}
    if mg.any():
        k = int(np.argmax(np.where(mg, dg / se_g, -1)))
        worst["gop"] = {"period": k // MAX_REPEATS + 1, "repeats": k % MAX_REPEATS + 1, "fit": float(gf[k]),
                        "truth": float(gt[k]), "se": float(se_g[k])}
    if mc.any():
        k = int(np.argmax(np.where(mc, dc / se_c, -1)))
        worst["gcp"] = {"period": k // MAX_REPEATS + 1, "repeats": k % MAX_REPEATS + 1, "fit": float(cf[k]),
                        "truth": float(ct[k]), "se": float(se_c[k])}
    rm = lambda d, m: float(np.sqrt(np.mean(d[m] ** 2))) if m.any() else float("nan")  # noqa: E731
    mz = lambda d, s, m: float(np.max(d[m] / s[m])) if m.any() else 0.0  # noqa: E731
    return FitReport(gop_checked=int(mg.sum()), gop_recovered=int((okg & mg).sum()), gop_rmse=rm(dg, mg),
                     gop_max_z=mz(dg, se_g, mg), gcp_checked=int(mc.sum()), gcp_recovered=int((okc & mc).sum()),
                     gcp_rmse=rm(dc, mc), gcp_max_z=mz(dc, se_c, mc), worst=worst)


def run_gpu_fit(spec: FitSpec = FitSpec(), seed: int = 3_000_003, max_iter: int = 60, log=None):
    """Generate, upload and fit on the GPU; returns (workload, FitResult, FitReport)."""
    import time

    import torch

    from .autograd import GradEngine, GradProblem, pairhmm_log10_likelihood

    wl = generate(spec, seed)
    b = wl.batch
    prob = GradProblem.from_batch(b)
    qb = torch.from_numpy(b.read_quals.astype(np.float32)).cuda()
    with GradEngine() as eng:
        ll = lambda a, i, d, c: pairhmm_log10_likelihood(prob, a, i, d, c, eng)  # noqa: E731
        t0 = time.perf_counter()
        res = fit(ll, wl.context, qb, b.read_off, len(b.read_len), max_iter=max_iter, log=log)
        res.seconds = time.perf_counter() - t0
        res.cells_per_eval = b.cells()
    return wl, res, evaluate(res, wl)
