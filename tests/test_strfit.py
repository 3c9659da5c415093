"""CPU checks of the STR penalty fit: the synthetic draw follows the stated penalties, and the fit
objective's gradient with respect to the (period, repeats) tables is right.

The objective is exercised with the torch FP64 restatement of the continuous PairHMM from the
oracle standing in for the GPU kernel (same call signature), against central finite differences
of the objective itself.
"""
import numpy as np
import pytest
import torch

from gatk_amd import dragstr, strfit
from gatk_amd.batch import pack_regions
from oracle import oracle


@pytest.fixture(scope="module")
def small_workload():
    spec = strfit.FitSpec(n_regions=60, reads_per_region=40)
    return strfit.generate(spec, seed=5)


def _brute_force_log10_w(qo, qc):
    """W by enumerating row paths explicitly from the PairHMM transitions (LoglessPairHMM.java:51-57):
    a path is a state (M or I) per row plus a deletion length between consecutive M rows; weight =
    GM_1 x transitions x 4 per I row; deletion excursions of length L weigh MD DD^(L-1) GM."""
    import itertools

    R = len(qo)
    e = lambda q: 10 ** (-q / 10)  # noqa: E731
    MM = [1 - 2 * e(q) for q in qo]
    MI = [e(q) for q in qo]
    MD = MI
    GM = [1 - e(q) for q in qc]
    II = [e(q) for q in qc]
    total = 0.0
    for states in itertools.product("MI", repeat=R):
        if states[0] != "M":
            continue
        w = GM[0]
        for r in range(1, R):
            a, b = states[r - 1], states[r]
            if a == "M" and b == "M":
                dele = sum(MD[r - 1] * II[r - 1] ** (L - 1) * GM[r] for L in range(1, 400))
                w *= MM[r] + dele
            elif a == "M" and b == "I":
                w *= 4 * MI[r]
            elif a == "I" and b == "M":
                w *= GM[r]
            else:
                w *= 4 * II[r]
        total += w
    return np.log10(total)


def test_total_weight_matches_path_enumeration():
    rng = np.random.default_rng(11)
    lens = [1, 2, 4, 7]
    qo = np.concatenate([rng.uniform(8, 40, n) for n in lens])
    qc = np.concatenate([rng.uniform(7, 12, n) for n in lens])
    off = np.concatenate([[0], np.cumsum(lens)[:-1]])
    got = strfit.log10_total_weight(torch.from_numpy(qo), torch.from_numpy(qc), off, np.array(lens)).numpy()
    for k, n in enumerate(lens):
        want = _brute_force_log10_w(qo[off[k]:off[k] + n], qc[off[k]:off[k] + n])
        assert abs(got[k] - want) < 1e-12, (n, got[k], want)


def test_workload_layout(small_workload):
    wl = small_workload
    b = wl.batch
    assert b.n_pairs == len(b.read_len)
    assert np.all(b.pair_read == np.arange(b.n_pairs))
    assert np.all(np.diff(b.pair_hap) >= 0)
    assert wl.context.shape == b.read_bases.shape
    ends = b.read_off + b.read_len - 1
    assert np.all(wl.context[ends] == -1)
    assert set(np.unique(b.read_bases)) <= set(b"ACGT")
    assert int(b.region_read_off[-1]) == len(b.read_len)


def _cpu_loglik(batch, pairs):
    def f(qb, qi, qd, qc):
        out = []
        for p in pairs:
            r, h = int(batch.pair_read[p]), int(batch.pair_hap[p])
            o, n = int(batch.read_off[r]), int(batch.read_len[r])
            ho, hn = int(batch.hap_off[h]), int(batch.hap_len[h])
            out.append(oracle.torch_log10_likelihood(batch.hap_bases[ho:ho + hn], batch.read_bases[o:o + n],
                                                     qb[o:o + n], qi[o:o + n], qd[o:o + n], qc[o:o + n]))
        return torch.stack(out)
    return f


def test_objective_gradient_matches_finite_differences(small_workload):
    wl = small_workload
    b = wl.batch
    # three short reads against their haplotypes, trimmed for the pure-torch restatement
    keep_reads = [0, 40, 80]
    regs = []
    for rr in keep_reads:
        reads_k, haps_k = b.region(int(np.searchsorted(b.region_read_off, rr, side="right") - 1))
        rd = reads_k[0]
        regs.append(([haps_k[0][:160]], [tuple(x[:60] for x in rd)]))
    pb = pack_regions(regs)
    ctx = dragstr.contexts(pb, strfit.MAX_PERIOD, strfit.MAX_REPEATS)
    qb = torch.from_numpy(pb.read_quals.astype(np.float64))
    f = strfit.objective(_cpu_loglik(pb, range(pb.n_pairs)), ctx, qb, pb.read_off, pb.read_len)
    rng = np.random.default_rng(1)
    g0 = torch.tensor(rng.uniform(20, 40, strfit.N_CTX), dtype=torch.float64, requires_grad=True)
    c0 = torch.tensor(rng.uniform(7, 12, strfit.N_CTX), dtype=torch.float64, requires_grad=True)
    loss = f(g0, c0)
    loss.backward()
    used = np.unique(ctx[ctx >= 0])
    checked = 0
    for cell in used[:6]:
        for tab, grad in ((g0, g0.grad), (c0, c0.grad)):
            h = 1e-3
            with torch.no_grad():
                tp = tab.detach().clone(); tp[cell] += h
                tm = tab.detach().clone(); tm[cell] -= h
                fp = f(tp, c0.detach()) if tab is g0 else f(g0.detach(), tp)
                fm = f(tm, c0.detach()) if tab is g0 else f(g0.detach(), tm)
            fd = float((fp - fm) / (2 * h))
            assert abs(float(grad[cell]) - fd) <= 1e-6 + 1e-4 * abs(fd), (cell, float(grad[cell]), fd)
            checked += 1
    assert checked >= 6
    unused = np.setdiff1d(np.arange(strfit.N_CTX), used)
    assert torch.all(g0.grad[unused] == 0) and torch.all(c0.grad[unused] == 0)


def test_evaluate_flags_wrong_cells(small_workload):
    wl = small_workload
    res = strfit.FitResult(gop=wl.gop_true.copy(), gcp=wl.gcp_true.copy(), loss=0.0, evaluations=0, seconds=0.0,
                           cells_per_eval=0)
    rep = strfit.evaluate(res, wl, min_opens=1.0, min_extends=1.0)
    assert rep.gop_checked > 0 and rep.ok
    bad = strfit.FitResult(gop=wl.gop_true + 8.0, gcp=wl.gcp_true.copy(), loss=0.0, evaluations=0, seconds=0.0,
                           cells_per_eval=0)
    rep2 = strfit.evaluate(bad, wl, min_opens=1.0, min_extends=1.0)
    assert rep2.gop_recovered < rep2.gop_checked and not rep2.ok


def test_fit_converges_on_separable_likelihood():
    # Per read position a Bernoulli gap-open model: log10 L_pos = x log10 e(q) + (1 - x) log10(1 - 2 e(q))
    # with x = the expected event rate 2 e(q_true) per position, so each cell's optimum is q_true.
    rng = np.random.default_rng(3)
    n_pos, n_pairs = 6000, 60
    ctx = rng.integers(0, 40, n_pos).astype(np.int32)
    read_off = np.arange(0, n_pos, n_pos // n_pairs)
    truth = rng.uniform(15, 40, strfit.N_CTX)
    t = torch.from_numpy(truth)
    c = torch.from_numpy(ctx).long()
    x = 2 * 10 ** (-t[c] / 10)
    seg = torch.from_numpy(np.repeat(np.arange(n_pairs), n_pos // n_pairs))

    def loglik(qb, qi, qd, qc):
        e = 10 ** (-qi / 10)
        per = x * torch.log10(2 * e) + (1 - x) * torch.log10(1 - 2 * e)
        return torch.zeros(n_pairs, dtype=per.dtype).index_add(0, seg, per)

    qb = torch.zeros(n_pos, dtype=torch.float64)
    read_len = np.full(n_pairs, n_pos // n_pairs)
    # the test likelihood is already normalised: cancel the PairHMM normaliser the objective subtracts
    ll = lambda a, i, d, c: loglik(a, i, d, c) + strfit.log10_total_weight(i, c, read_off, read_len)  # noqa: E731
    res = strfit.fit(ll, ctx, qb, read_off, read_len, max_iter=300)
    got = res.gop.ravel()
    used = np.unique(ctx)
    assert np.max(np.abs(got[used] - truth[used])) < 0.1
    unused = np.setdiff1d(np.arange(strfit.N_CTX), np.unique(ctx))
    assert np.all(got[unused] == 30.0)
