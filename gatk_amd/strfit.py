"""Fit of per-(STR period, repeat count) gap penalties with the adjoint kernel.

The DRAGstr model gives every read position a context -- the best STR period p (1..6) and its copy
count k (1..20) (DragstrReadSTRAnalyzer.java:69-139) -- and reads its gap-open (GOP, used for
insertions and deletions alike) and gap-continuation (GCP) penalties from (p, k) tables
(DragstrPairHMMInputScoreImputator.java:53-70).  This module recovers those tables from reads by
maximum likelihood under the PairHMM with continuous per-position penalties.  The gradient with
respect to every table cell is the sum of the adjoint kernel's per-position gradients over the read
positions in that cell.

The PairHMM is not a normalised distribution over reads: an inserted base has weight 1
(LoglessPairHMM.java:53, no prior on the I state) where a matched base's prior sums to 1 over the
four bases (:89-90).  For a read of R rows with per-row penalties theta the total weight of all
reads is
    W(theta) = sum over row paths of (transition weights) x 4^(inserted rows),
a row-level recurrence that needs no haplotype (deletion excursions D(r-1,.) -> M(r,.) sum to
MD_{r-1} GM_r / (1 - DD_{r-1})).  The read distribution is P(read) = L(read) / W, so the fit maximises
    sum over reads of log10 L(read | haplotype) - log10 W(theta of the read's rows),
which is consistent for reads drawn from P.  Maximising log10 L alone is not: it rewards weight
4 e(GCP) per extra inserted base and drives GOP and GCP down.

``generate`` draws reads from P exactly (up to the haplotype ends, which reads stay 40 bases away
from):
  * STR-rich haplotypes: random flanks between tandem repeats of 1-6 bp motifs with 2-20 copies;
  * a read of R rows starts at a uniform haplotype offset; row r's penalties come from the DRAGstr
    cell of position r-1 of the error-free haplotype window under the read (so they do not depend on
    the drawn errors) and the last row uses 45 / 10 like the imputer;
  * the row path (M or I per row, deletion excursions between rows) is drawn from the transition
    weights tilted by backward weights of the row-level recurrence, deletion lengths geometric in
    DD; an M row emits the haplotype base or, with probability e(q_base), one of the other three;
    an I row emits a uniform base.
The truth tables are the DRAGstr defaults (DragstrParams.java:49-58, :114-117) with GOP capped at
40 as the imputer caps it (:31) and GCP floored at 7: below 6.02 Phred the weight 4 e(GCP) of one
more inserted base exceeds 1, and P then puts most of its mass on insertions running to the read's
end.

``fit`` runs projected Adam over the 2 x 120 Phred values.  Cells that no read position visits keep
their starting value; ``FitReport`` compares the recovered cells against the truth within the
standard error the drawn events allow.
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
GCP_FLOOR = 7.0
INS_WEIGHT = 4.0  # total weight of one inserted base over the four bases
PHRED_SE = 10.0 / math.log(10.0)  # d q / d ln e

_ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)


def true_tables() -> Tuple[np.ndarray, np.ndarray]:
    """Truth for the synthetic fit (module docstring)."""
    d = dragstr.DragstrParams.default()
    return np.minimum(40.0, d.gop_table[:MAX_PERIOD]).copy(), np.maximum(GCP_FLOOR, d.gcp_table[:MAX_PERIOD])


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
    context: np.ndarray           # int32 [read_bytes]: cell of each read position (-1 at read ends)
    gop_true: np.ndarray          # [6, 20]
    gcp_true: np.ndarray
    opens: np.ndarray             # [N_CTX]: gap opens drawn, by the cell whose GOP they use
    extends: np.ndarray           # [N_CTX, 2]: (extensions drawn, stops drawn), by the cell whose GCP they use


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
    """Row-major [n, Lmax] matrix -> concatenated values + offsets + mask."""
    mask = np.arange(mat.shape[1])[None, :] < lens[:, None]
    off = np.concatenate([[0], np.cumsum(lens, dtype=np.int64)[:-1]]).astype(np.int64)
    return np.ascontiguousarray(mat[mask]), off, mask


def row_weights(e_open, e_ext):
    """Per-row transition weights of the continuous model from per-row e(GOP), e(GCP) arrays
    [..., R]: (MM, MI, GM, II, DEL) where DEL[r] = MD_{r-1} GM_r / (1 - DD_{r-1}) is the summed
    deletion excursion from M(r-1) to M(r) (DEL[0] = 0).  Works on numpy arrays and torch tensors."""
    MM = 1.0 - 2.0 * e_open
    GM = 1.0 - e_ext
    if isinstance(e_open, np.ndarray):
        DEL = np.zeros_like(e_open)
        DEL[..., 1:] = e_open[..., :-1] * GM[..., 1:] / (1.0 - e_ext[..., :-1])
    else:
        import torch

        DEL = torch.cat([torch.zeros_like(e_open[..., :1]), e_open[..., :-1] * GM[..., 1:] / (1.0 - e_ext[..., :-1])],
                        dim=-1)
    return MM, e_open, GM, e_ext, DEL


def generate(spec: FitSpec = FitSpec(), seed: int = 3_000_003, gop_true: Optional[np.ndarray] = None,
             gcp_true: Optional[np.ndarray] = None) -> FitWorkload:
    """Reads drawn from the normalised PairHMM read distribution with known (period, repeats)
    penalties (module docstring)."""
    if gop_true is None or gcp_true is None:
        gop_true, gcp_true = true_tables()
    gop_flat = np.append(np.asarray(gop_true, dtype=np.float64).ravel(), END_GOP)  # index -1 -> end values
    gcp_flat = np.append(np.asarray(gcp_true, dtype=np.float64).ravel(), END_GCP)
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
    rows = np.arange(Rmax)[None, :]
    active = rows < R[:, None]
    # per-row cells from the error-free window under the read
    widx = hoff[src][:, None] + start[:, None] + rows
    win = hcat[np.minimum(widx, len(hcat) - 1)]
    wbytes, woff, wmask = _packed_reads(win, R)

    class _W:
        read_bases, read_off, read_len = wbytes, woff, R

    ctx = np.full((n, Rmax), -1, dtype=np.int32)
    ctx[wmask] = dragstr.contexts(_W, MAX_PERIOD, MAX_REPEATS)
    eo, ec = err(gop_flat)[ctx], err(gcp_flat)[ctx]
    MM, MI, GM, II, DEL = row_weights(eo, ec)
    # backward weights of the row-level recurrence, normalised per row
    bM = np.ones((n, Rmax))
    bI = np.ones((n, Rmax))
    for k in range(Rmax - 2, -1, -1):
        m = (MM[:, k + 1] + DEL[:, k + 1]) * bM[:, k + 1] + INS_WEIGHT * MI[:, k + 1] * bI[:, k + 1]
        i = GM[:, k + 1] * bM[:, k + 1] + INS_WEIGHT * II[:, k + 1] * bI[:, k + 1]
        s = np.maximum(m, i)
        last = k >= R - 1
        bM[:, k] = np.where(last, 1.0, m / s)
        bI[:, k] = np.where(last, 1.0, i / s)
    # forward draw
    quals = rng.integers(spec.base_q[0], spec.base_q[1] + 1, (n, Rmax)).astype(np.uint8)
    out = np.zeros((n, Rmax), dtype=np.uint8)
    hend = hoff[src] + hl[src]
    col = hoff[src] + start
    in_ins = np.zeros(n, dtype=bool)
    bad = np.zeros(n, dtype=bool)
    opens = np.zeros(N_CTX + 1, dtype=np.int64)
    extends = np.zeros((N_CTX + 1, 2), dtype=np.int64)

    def cell(c):
        return np.where(c < 0, N_CTX, c)

    for k in range(Rmax):
        act = active[:, k]
        base = np.zeros(n, dtype=np.uint8)
        if k == 0:
            match = act.copy()
        else:
            u = rng.random(n)
            wd, wl_, wi = MM[:, k] * bM[:, k], DEL[:, k] * bM[:, k], INS_WEIGHT * MI[:, k] * bI[:, k]
            tot = wd + wl_ + wi
            from_m = act & ~in_ins
            dele = from_m & (u * tot < wl_)
            ins = from_m & ~dele & (u * tot < wl_ + wi)
            wm, wi2 = GM[:, k] * bM[:, k], INS_WEIGHT * II[:, k] * bI[:, k]
            from_i = act & in_ins
            ext = from_i & (rng.random(n) * (wm + wi2) < wi2)
            np.add.at(opens, cell(ctx[dele, k - 1]), 1)
            np.add.at(opens, cell(ctx[ins, k]), 1)
            np.add.at(extends[:, 0], cell(ctx[ext, k]), 1)
            np.add.at(extends[:, 1], cell(ctx[from_i & ~ext, k]), 1)
            if np.any(dele):
                dd = II[dele, k - 1]
                L = rng.geometric(1.0 - dd)
                np.add.at(extends[:, 0], cell(ctx[dele, k - 1]), (L - 1).astype(np.int64))
                np.add.at(extends[:, 1], cell(ctx[dele, k - 1]), 1)
                col[dele] += L
            ins_row = ins | ext
            match = act & ~ins_row
            in_ins = ins_row
            base[ins_row] = _ACGT[rng.integers(0, 4, int(ins_row.sum()))]
        over = match & (col >= hend)
        bad |= over
        ok = match & ~over
        hb = hcat[np.minimum(col, len(hcat) - 1)]
        wrong = ok & (rng.random(n) < err(quals[:, k]))
        sub = _ACGT[(np.searchsorted(_ACGT, hb) + rng.integers(1, 4, n)) % 4]
        base[ok] = np.where(wrong[ok], sub[ok], hb[ok])
        col[ok] += 1
        out[:, k] = np.where(act, base, out[:, k])
    keep = ~bad
    out, quals, R, src, ctx = out[keep], quals[keep], R[keep], src[keep], ctx[keep]
    rbytes, roff, _ = _packed_reads(out, R)
    qbytes, _, _ = _packed_reads(quals, R)
    cflat, _, _ = _packed_reads(ctx, R)
    z = np.zeros(len(rbytes), dtype=np.uint8)
    rro = np.searchsorted(src, np.arange(A + 1)).astype(np.int64)  # reads stay grouped by haplotype
    batch = PackedBatch(
        read_bases=rbytes, read_quals=qbytes, read_ins=z + 45, read_del=z + 45, read_gcp=z + 10,
        read_off=roff, read_len=R.astype(np.int32), hap_bases=hcat, hap_off=hoff, hap_len=hl,
        pair_read=np.arange(len(R), dtype=np.int32), pair_hap=src.astype(np.int32),
        region_pair_off=rro.copy(), region_read_off=rro, region_hap_off=np.arange(A + 1, dtype=np.int64))
    return FitWorkload(batch=batch, context=cflat.astype(np.int32), gop_true=np.asarray(gop_true, dtype=np.float64),
                       gcp_true=np.asarray(gcp_true, dtype=np.float64), opens=opens[:N_CTX],
                       extends=extends[:N_CTX])


# ---------------------------------------------------------------------------------------------
# the fit


@dataclass
class FitResult:
    gop: np.ndarray   # [6, 20] fitted Phred
    gcp: np.ndarray
    loss: float       # final -mean (log10 L - log10 W) per read
    evaluations: int
    seconds: float
    cells_per_eval: int


LoglikFn = Callable[..., "object"]  # (qb, qi, qd, qc) -> log10 L per pair (differentiable torch tensor)


def log10_total_weight(qo, qc, read_off, read_len):
    """log10 W per read (module docstring) from per-position Phred GOP / GCP tensors, FP64 torch."""
    import torch

    dev = qo.device
    off = torch.as_tensor(np.asarray(read_off), device=dev).long()
    lens = torch.as_tensor(np.asarray(read_len), device=dev).long()
    rmax = int(lens.max()) if lens.numel() else 0
    rows = torch.arange(rmax, device=dev)
    idx = (off[:, None] + rows[None, :]).clamp(max=qo.numel() - 1)
    act = rows[None, :] < lens[:, None]
    ten = torch.tensor(10.0, dtype=torch.float64, device=dev)
    zero = torch.zeros((), dtype=torch.float64, device=dev)
    eo = torch.where(act, torch.pow(ten, -qo.double()[idx] / 10.0), zero)
    ec = torch.where(act, torch.pow(ten, -qc.double()[idx] / 10.0), zero)
    MM, MI, GM, II, DEL = row_weights(eo, ec)
    wM, wI = GM[:, 0], torch.zeros_like(GM[:, 0])
    logw = torch.zeros_like(wM)
    for k in range(1, rmax):
        a = act[:, k]
        nM = wM * (MM[:, k] + DEL[:, k]) + wI * GM[:, k]
        nI = INS_WEIGHT * (wM * MI[:, k] + wI * II[:, k])
        s = nM + nI
        wM = torch.where(a, nM / s, wM)
        wI = torch.where(a, nI / s, wI)
        logw = logw + torch.where(a, torch.log10(s), zero)
    return logw + torch.log10(wM + wI)


def objective(loglik: LoglikFn, context, qb, read_off, read_len):
    """f(gop, gcp) = -mean over reads of (log10 L - log10 W), for [N_CTX] Phred tensors.

    ``loglik(qb, qi, qd, qc)`` returns log10 L per pair as a differentiable function of per-position
    Phred tensors (``gatk_amd.autograd.pairhmm_log10_likelihood`` bound to a problem and engine),
    one pair per read in read order; ``context`` holds the table cell per read position (-1 = fixed
    end values).
    """
    import torch

    dev = qb.device
    ctx = torch.as_tensor(np.asarray(context), device=dev).long()
    fixed = ctx < 0
    cidx = ctx.clamp(min=0)
    n_reads = len(read_len)

    def f(gop, gcp):
        qo = torch.where(fixed, torch.full_like(gop[cidx], END_GOP), gop[cidx])
        qc = torch.where(fixed, torch.full_like(gcp[cidx], END_GCP), gcp[cidx])
        ll = loglik(qb, qo, qo, qc)
        return -(ll.sum() - log10_total_weight(qo, qc, read_off, read_len).sum()) / n_reads

    return f


def fit(loglik: LoglikFn, context, qb, read_off, read_len, init_gop: float = 30.0, init_gcp: float = 12.0,
        max_iter: int = 200, lr: float = 1.5, lr_final: float = 0.01, log=None) -> FitResult:
    """Projected Adam over the 6 x 20 GOP / GCP Phred tables on ``objective`` (same arguments).

    Adam's per-parameter normalisation suits the table: cells differ by orders of magnitude in how
    many read positions they see, the FP32 likelihood is noisy at the 1e-7 level that a curvature
    line search would have to resolve, and a step never exceeds the learning rate (in Phred), which
    decays from ``lr`` to ``lr_final`` on a cosine schedule.  Values are kept in [GOP_MIN, 60] and
    [GCP_MIN, 40]; cells with no read position have zero gradient and keep their starting value.
    """
    import time

    import torch

    dev = qb.device
    f = objective(loglik, context, qb, read_off, read_len)
    gop = torch.full((N_CTX,), float(init_gop), dtype=torch.float64, device=dev, requires_grad=True)
    gcp = torch.full((N_CTX,), float(init_gcp), dtype=torch.float64, device=dev, requires_grad=True)
    opt = torch.optim.Adam([gop, gcp], lr=lr, betas=(0.9, 0.99))
    last = float("nan")
    t0 = time.perf_counter()
    for it in range(max_iter):
        for gr in opt.param_groups:
            gr["lr"] = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * it / max(1, max_iter - 1)))
        opt.zero_grad()
        loss = f(gop, gcp)
        loss.backward()
        opt.step()
        with torch.no_grad():
            gop.clamp_(GOP_MIN, 60.0)
            gcp.clamp_(GCP_MIN, 40.0)
        last = float(loss.detach())
        if log and (it + 1) % 25 == 0:
            log(f"fit: iteration {it + 1}, -mean log10(L / W) = {last:.6f}")
    with torch.no_grad():
        last = float(f(gop, gcp))
    if dev.type == "cuda":
        torch.cuda.synchronize(dev)
    return FitResult(gop=gop.detach().cpu().numpy().reshape(MAX_PERIOD, MAX_REPEATS),
                     gcp=gcp.detach().cpu().numpy().reshape(MAX_PERIOD, MAX_REPEATS), loss=last,
                     evaluations=max_iter + 1, seconds=time.perf_counter() - t0, cells_per_eval=0)


# ---------------------------------------------------------------------------------------------
# recovery report


@dataclass
class FitReport:
    gop_checked: int          # cells with enough drawn gap opens to pin GOP
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


def evaluate(res: FitResult, wl: FitWorkload, min_opens: float = 30.0, min_extends: float = 20.0,
             z_tol: float = 3.0, slack: float = 0.5) -> FitReport:
    """Recovered cell <=> |fit - truth| <= z_tol * SE + slack, with SE from the events drawn in the
    cell: GOP: SE = 4.34 / sqrt(opens); GCP: SE = 4.34 sqrt((1 - e) / (e T)), T = extensions + stops.
    Only cells with >= min_opens opens (GOP) or >= min_extends extensions (GCP) are checked; the rest
    have too little data to pin them."""
    gt, ct = wl.gop_true.ravel(), wl.gcp_true.ravel()
    gf, cf = res.gop.ravel(), res.gcp.ravel()
    ec = err(ct)
    opens = wl.opens.astype(np.float64)
    ext, stops = wl.extends[:, 0].astype(np.float64), wl.extends[:, 1].astype(np.float64)
    se_g = PHRED_SE / np.sqrt(np.maximum(opens, 1e-300))
    se_c = PHRED_SE * np.sqrt((1.0 - ec) / np.maximum(ec * (ext + stops), 1e-300))
    mg = opens >= min_opens
    mc = ext >= min_extends
    dg, dc = np.abs(gf - gt), np.abs(cf - ct)
    okg = dg <= z_tol * se_g + slack
    okc = dc <= z_tol * se_c + slack
    worst = {}
    for name, m, dd, se, fv, tv in (("gop", mg, dg, se_g, gf, gt), ("gcp", mc, dc, se_c, cf, ct)):
        if m.any():
            k = int(np.argmax(np.where(m, dd / se, -1)))
            worst[name] = {"period": k // MAX_REPEATS + 1, "repeats": k % MAX_REPEATS + 1, "fit": float(fv[k]),
                           "truth": float(tv[k]), "se": float(se[k])}
    rm = lambda d, m: float(np.sqrt(np.mean(d[m] ** 2))) if m.any() else float("nan")  # noqa: E731
    mz = lambda d, s, m: float(np.max(d[m] / s[m])) if m.any() else 0.0  # noqa: E731
    return FitReport(gop_checked=int(mg.sum()), gop_recovered=int((okg & mg).sum()), gop_rmse=rm(dg, mg),
                     gop_max_z=mz(dg, se_g, mg), gcp_checked=int(mc.sum()), gcp_recovered=int((okc & mc).sum()),
                     gcp_rmse=rm(dc, mc), gcp_max_z=mz(dc, se_c, mc), worst=worst)


def run_gpu_fit(spec: FitSpec = FitSpec(), seed: int = 3_000_003, max_iter: int = 200, log=None):
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
        res = fit(ll, wl.context, qb, b.read_off, b.read_len, max_iter=max_iter, log=log)
        res.seconds = time.perf_counter() - t0
        res.cells_per_eval = b.cells()
    return wl, res, evaluate(res, wl)
