"""GPU: the adjoint kernel's gradients fit per-(STR period, repeats) GOP / GCP tables that recover
the penalties the synthetic reads were drawn with.

Recovery criterion (``strfit.evaluate``): every table cell with at least 30 expected gap opens
(GOP) or 20 expected continuations (GCP) lies within 3 standard errors + 0.5 Phred of the truth,
where the standard error follows from the expected event counts of that cell.
"""
import sys

import numpy as np
import pytest

from gatk_amd import strfit

pytestmark = pytest.mark.gpu


def test_fit_recovers_str_penalties():
    spec = strfit.FitSpec(n_regions=500, reads_per_region=80)
    print(f"[strfit] spec={spec} seed=77", flush=True)
    wl, res, rep = strfit.run_gpu_fit(spec, seed=77, max_iter=200,
                                      log=lambda m: print(m, file=sys.stderr, flush=True))
    print(f"[strfit] {res.evaluations} evaluations in {res.seconds:.1f} s; report {rep.as_dict()}", flush=True)
    assert np.isfinite(res.loss)
    assert rep.gop_checked >= 40 and rep.gcp_checked >= 10
    assert rep.ok, rep.as_dict()
