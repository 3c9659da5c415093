"""Build-time guards on the gfx950 device code.

Compiles every .hip source of this directory to device assembly (with the per-file flags the
Makefile gives it) and fails if
  * any scalar-memory store, scalar atomic, or scalar-cache write-back/discard instruction appears:
    the device code must not write through the scalar data cache;
  * a VALU instruction within two instructions after a ds_write_b96/b128 overwrites that store's
    data VGPRs.  The FP64 forward step gave wrong results exactly when its code held that pattern
    (post-RA scheduling on) and correct ones when it did not (post-RA scheduling off), and the
    compiler does not guard against it; 8-byte LDS stores (ds_write2_b64) followed at once by such
    a write compute correctly and are not flagged.
Run by __graft_entry__.build() in the build container only.
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
# Per-file flags of the Makefile's object rules.
FILE_FLAGS = {"pairhmm_fwd.hip": ["-mllvm", "-enable-post-misched=false"]}
WIDE_LDS_STORE = re.compile(r"^ds_(write|store)_b(96|128)\b")


def _vgprs(operand):
    out = set()
    for m in re.finditer(r"v\[(\d+):(\d+)\]", operand):
        out.update(range(int(m.group(1)), int(m.group(2)) + 1))
    out.update(int(m.group(1)) for m in re.finditer(r"\bv(\d+)\b", operand))
    return out


def wide_lds_store_overwrites(asm):
    """(store, overwriting instruction) pairs where a VALU op within two instructions after a
    ds_write_b96/b128 writes one of the store's data VGPRs."""
    ins = [l.split(";")[0].strip() for l in asm.split("\n")
           if l.startswith("\t") and not l.startswith("\t.")]
    hits = []
    for k, line in enumerate(ins):
        if not WIDE_LDS_STORE.match(line):
            continue
        data = set()
        for operand in line.split(None, 1)[1].split(",")[1:]:
            data |= _vgprs(operand)
        for nxt in ins[k + 1:k + 3]:
            if nxt.startswith("v_") and _vgprs(nxt.split(",")[0]) & data:
                hits.append((line, nxt))
    return hits


def main() -> int:
    bad = []
    for name in sorted(os.listdir(HERE)):
        if not name.endswith(".hip"):
            continue
        with tempfile.TemporaryDirectory() as td:
            out = os.path.join(td, name + ".s")
            subprocess.run(["/opt/rocm/bin/hipcc", *FLAGS, *FILE_FLAGS.get(name, []), "-o", out,
                            os.path.join(HERE, name)], check=True, stderr=subprocess.DEVNULL)
            asm = open(out).read()
        hits = FORBIDDEN.findall(asm)
        if hits:
            bad.append(f"{name} contains forbidden scalar-memory writes: {sorted(set(hits))}")
        over = wide_lds_store_overwrites(asm)
        if over:
            bad.append(f"{name}: {len(over)} LDS stores over 8 bytes have their data overwritten within two "
                       f"instructions, e.g. {over[0]}")
    for msg in bad:
        print(f"check_isa: {msg}", file=sys.stderr)
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main())
