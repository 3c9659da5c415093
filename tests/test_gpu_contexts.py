"""Two contexts driven from two threads at once (the C-ABI allows it: contexts are independent, each
used by one thread at a time).  Every batch's results must equal, bit for bit, the same batch run
alone: batches of different shapes upload and run on two contexts concurrently, one thread each,
several rounds, including precision-stress batches whose FP64 pass overlaps the other context's
FP32 pass.  Graph rounds capture on one context while the other uploads: the library keeps its
copies off the legacy null stream, whose use during another thread's capture breaks that capture."""
import threading

import numpy as np
import pytest

from gatk_amd import _native, synth
from gatk_amd.batch import BatchEngine

pytestmark = pytest.mark.gpu


def _serial(batch):
    with BatchEngine() as eng:
        dev = eng.upload(batch)
        out = np.empty(batch.n_pairs)
        dev.forward(mode=_native.RUN_PERSISTENT_EAGER, out=out)
        dev.free()
    return out


def test_two_contexts_from_two_threads_match_serial_runs():
    batches = [synth.workload("batched_uneven", n_regions=40, seed=71),
               synth.workload("precision_stress", n_regions=6, seed=72),
               synth.workload("batched_uneven", n_regions=25, seed=73),
               synth.workload("precision_stress", n_regions=4, seed=74),
               synth.workload("batched_uneven", n_regions=2, seed=75),
               synth.workload("batched_uneven", n_regions=3, seed=76)]
    want = [_serial(b) for b in batches]
    print(f"[contexts] pairs per batch {[b.n_pairs for b in batches]}", flush=True)
    engines = [BatchEngine(), BatchEngine()]
    got = {}
    errors = []

    def work(t):
        try:
            eng = engines[t]
            for rnd in range(6):
                for k in range(t, len(batches), 2):
                    dev = eng.upload(batches[k])
                    out = np.empty(batches[k].n_pairs)
                    mode = _native.RUN_PERSISTENT if rnd % 2 == 0 else _native.RUN_PERSISTENT_EAGER
                    dev.forward(mode=mode, out=out)
                    dev.free()
                    got[(rnd, k)] = out
        except Exception as e:  # reported in the main thread
            errors.append(e)

    threads = [threading.Thread(target=work, args=(t,)) for t in range(2)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=150)
    for e in engines:
        e.close()
    assert not errors, errors
    assert len(got) == 6 * len(batches)
    for (rnd, k), out in got.items():
        assert np.array_equal(out, want[k], equal_nan=True), (rnd, k)
