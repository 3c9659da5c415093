"""Drives the engine's stream path on the CPU build (libstream_harness.so: host C-ABI code with
AddressSanitizer, HIP replaced by host memory) through the real Python wrappers: ``RegionStream``
with empty chunks between real ones, then ``run_streaming`` with its spawn-context generator pool.

Every result must equal the stub kernel's value -(R * 1000 + H) * 1e-6 for its pair, so results that
land in the wrong array, the wrong place or a freed array are caught; AddressSanitizer aborts the
process on any out-of-bounds or freed-memory access in the C++ host code.

Usage: python drive_stream.py LIB REGIONS CHUNK WORKERS
"""
import ctypes
import os
import sys
import time
import types

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def expected(batch):
    import numpy as np
    return -(batch.read_len[batch.pair_read].astype(np.float64) * 1000.0
             + batch.hap_len[batch.pair_hap].astype(np.float64)) * 1e-6


class _Tolerant:
    """Lets the binding declare entry points the harness library does not build."""

    def __init__(self, lib):
        self.lib = lib

    def __getattr__(self, name):
        try:
            return getattr(self.lib, name)
        except AttributeError:
            return types.SimpleNamespace()


def main() -> int:
    lib_path, n_regions, chunk, workers = sys.argv[1], int(sys.argv[2]), int(sys.argv[3]), int(sys.argv[4])
    sys.path.insert(0, ROOT)
    import numpy as np

    from gatk_amd import _native, synth
    from gatk_amd.batch import BatchEngine
    from gatk_amd.stream import RegionStream, run_streaming

    lib = ctypes.CDLL(lib_path)
    _native.declare(_Tolerant(lib))
    _native._lib = lib

    # 1. Empty chunks between real ones, with only the stream's own references to the outputs
    #    (the pattern that released an output array still in flight).
    chunks = [synth.generate(synth.BATCHED_UNEVEN, 3, 8100 + k, mix=(synth.PRECISION_STRESS, 0.3)) for k in range(4)]
    empty = chunks[0].subset_regions([])
    checks = []
    with BatchEngine() as eng, RegionStream(eng) as rs:
        for rnd in range(2):
            for k, c in enumerate(chunks):
                out = np.full(c.n_pairs, np.nan)
                rs.submit(c, out)
                checks.append((c, out.__array_interface__["data"][0], c.n_pairs))
                del out
                if k % 2 == 1:
                    rs.submit(empty, np.empty(0))
                    rs.submit(empty, np.empty(0))
            st = rs.finish()
            assert st.cells == sum(c.cells() for c in chunks), st.cells
    print(f"harness: {len(checks)} chunks with empty chunks between them: clean", flush=True)

    # 2. Kept outputs must hold the stub's value for every pair.
    outs = []
    with BatchEngine() as eng, RegionStream(eng) as rs:
        for c in chunks:
            o = np.full(c.n_pairs, np.nan)
            rs.submit(c, o)
            rs.submit(empty, np.empty(0))
            outs.append(o)
        rs.finish()
    for c, o in zip(chunks, outs):
        np.testing.assert_array_equal(o, expected(c))
    print("harness: kept outputs equal the stub kernel's values", flush=True)

    # 3. run_streaming as the bench drives it: spawn pool, outputs not kept.
    t0 = time.perf_counter()
    res = run_streaming(n_regions, chunk_regions=chunk, workers=workers, seed=4_000_000,
                        log=lambda m: print(f"harness: {m}", flush=True))
    print(f"harness: run_streaming {res.regions} regions, {res.pairs} pairs in "
          f"{time.perf_counter() - t0:.1f} s: clean", flush=True)
    res, outs = run_streaming(min(n_regions, 4 * chunk), chunk_regions=chunk, workers=min(workers, 4),
                              seed=4_100_000, keep_outputs=True)
    assert res.pairs == sum(len(o) for o in outs)
    assert all(np.all(o < 0) for o in outs)
    print("harness: done", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
