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


def test_draw_rates_follow_truth(small_workload):
    wl = strfit.generate(strfit.FitSpec(n_regions=300, reads_per_region=60), seed=9)
    for tab, truth in ((wl.opens, strfit.err(wl.gop_true.ravel())), (wl.extends, strfit.err(wl.gcp_true.ravel()))):
        hits, trials = tab[:, 0].astype(float), tab[:, 1].astype(float)
        m = trials * truth >= 20
        assert m.sum() >= 20
        z = (hits[m] - truth[m] * trials[m]) / np.sqrt(truth[m] * (1 - truth[m]) * trials[m])
        assert np.max(np.abs(z)) < 4.5, np.max(np.abs(z))


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
    f = strfit.objective(_cpu_loglik(pb, range(pb.n_pairs)), ctx, qb, pb.read_off, len(pb.read_len))
    rng = np.random.default_rng(1)
    g0 = torch.tensor(rng.uniform(20, 40, strfit.N_CTX), dtype=torch.float64, requires_grad=True)
    c0 = torch.tensor(rng.uniform(2, 10, strfit.N_CTX), dtype=torch.float64, requires_grad=True)
    loss = f(g0, c0)
    loss.backward()
    used = np.unique(ctx[ctx >= 0])
    # cells used only at read position 0 carry no gradient
    first_only = set(ctx[pb.read_off].tolist()) - set(np.delete(ctx, pb.read_off)[np.delete(ctx, pb.read_off) >= 0].tolist())
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
            if cell in first_only:
                assert float(grad[cell]) == 0.0
            else:
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
