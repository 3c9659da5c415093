"""Build-time guard: the gfx950 device code must not write through the scalar data cache.

Compiles every .hip source of this directory to device assembly and fails if any scalar-memory
store, scalar atomic, or scalar-cache write-back/discard instruction appears.  Run by
__graft_entry__.build() in the build container only.
"""
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
FORBIDDEN = re.compile(r"^\s*(s_store_|s_buffer_store|s_scratch_store|s_atomic_|s_buffer_atomic|"
                       r"s_dcache_wb|s_dcache_discard)", re.M)
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "--cuda-device-only", "-S"]


def main() -> int:
    bad = []
    for name in sorted(os.listdir(HERE)):
        if not name.endswith(".hip"):
            continue
        with tempfile.TemporaryDirectory() as td:
            out = os.path.join(td, name + ".s")
            subprocess.run(["/opt/rocm/bin/hipcc", *FLAGS, "-o", out, os.path.join(HERE, name)], check=True,
                           stderr=subprocess.DEVNULL)
            asm = open(out).read()
        hits = FORBIDDEN.findall(asm)
        if hits:
            bad.append((name, sorted(set(hits))))
    for name, hits in bad:
        print(f"check_isa: {name} contains forbidden scalar-memory writes: {hits}", file=sys.stderr)
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main())
