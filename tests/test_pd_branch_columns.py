"""The PD-HMM kernel skips the branch update on step blocks where no lane's column is special
(pdhmm.hip: pd_step<kBranch = false>; special columns are set by pd_column_codes in
pdhmm_capi.cpp: DEL_END, either inside bit, or the column before an inside column).

This pins the argument on the CPU with a restatement of the kernel's recurrence in which every cell
of a non-special column takes the skipping path (the most skipping any block schedule can do): its
results equal, bit for bit, the same recurrence with the branch update on every cell, and agree
with the FP64 oracle (LoglessPDPairHMM.java:34-150) to rounding.  Dropping the deletion-start term
of the special rule must change some result, so the check can tell.
"""
import math

import numpy as np

from oracle import oracle

SNP, DEL_START, DEL_END = 1, 2, 4
CLASS = {ord("A"): 8, ord("a"): 8, ord("C"): 16, ord("c"): 16, ord("G"): 32, ord("g"): 32, ord("T"): 64,
         ord("t"): 64}


def _inside_pass(pd, state):
    """LoglessPDPairHMM's state machine over one row (:63-136): inside flags and the end state."""
    inside = []
    for b in pd:
        inside.append(state == "inside")
        if state == "after":
            state = "normal"
        if b & DEL_START:
            state = "inside"
        if b & DEL_END:
            state = "after"
    return inside, state


def _special(pd, in1, in2, with_start=True):
    H = len(pd)
    out = []
    for j in range(H):
        s = in1[j] or in2[j] or bool(pd[j] & DEL_END)
        if with_start and j + 1 < H:
            s = s or in1[j + 1] or in2[j + 1]
        out.append(s)
    return out


def _restated(hap, pd, read, quals, ins, dels, gcp, skip, with_start=True):
    """log10 L by the kernel's restated recurrence; skip: non-special cells keep their branch values
    and hand on E = X."""
    H, R = len(hap), len(read)
    in1, end1 = _inside_pass(pd, "normal")
    in2, _ = _inside_pass(pd, end1)
    special = _special(pd, in1, in2, with_start)
    d0 = 1.0 / H
    EM = [0.0] * (H + 1)
    EI = [0.0] * (H + 1)
    ED = [d0] * (H + 1)  # row 0: D = 1 / H on columns 0..H
    total = 0.0
    for i in range(1, R + 1):
        t = oracle.qual_to_trans_probs(int(ins[i - 1]), int(dels[i - 1]), int(gcp[i - 1]))
        MM, GM, MI, II, MD, DD = (float(x) for x in t)
        err = oracle.qual_to_error_prob(int(quals[i - 1]))
        x = int(read[i - 1])
        pm, px = 1.0 - err, err / 3.0
        if x == ord("N"):
            px = pm
        inside = in1 if i == 1 else in2
        nM, nI, nD = [0.0] * (H + 1), [0.0] * (H + 1), [0.0] * (H + 1)
        bM = bI = bD = 0.0
        for c in range(1, H + 1):
            y, p = int(hap[c - 1]), int(pd[c - 1])
            alt = (p & 0x78) if p & SNP else 0
            if y == ord("N"):
                alt |= 0xF8
            match = y == x or (alt & CLASS.get(x, 0x80)) != 0
            M = (pm if match else px) * (EM[c - 1] * MM + (EI[c - 1] + ED[c - 1]) * GM)
            I = EM[c] * MI + EI[c] * II
            D = nM[c - 1] * MD + nD[c - 1] * DD
            if i == R:
                total += M + I
            if skip and not special[c - 1]:
                nM[c], nI[c], nD[c] = M, I, D
                continue
            if not inside[c - 1]:
                bM, bI, bD = nM[c - 1], nI[c - 1], nD[c - 1]
            f = 1.0 if p & DEL_END else 0.0
            nM[c], nI[c], nD[c] = max(bM * f, M), max(bI * f, I), max(bD * f, D)
        EM, EI, ED = nM, nI, nD
    return math.log10(total) if total > 0 else -math.inf


def _case(rng, H, n_del, n_snp, R):
    hap = bytearray(rng.choice(np.frombuffer(b"ACGT", np.uint8), H))
    pd = bytearray(H)
    for j in rng.choice(H, min(n_snp, H), replace=False):
        pd[j] |= SNP | int(rng.choice([8, 16, 32, 64]))
    for _ in range(n_del):
        s = int(rng.integers(0, H))
        pd[s] |= DEL_START
        pd[min(H - 1, s + int(rng.integers(0, 5)))] |= DEL_END
    if rng.random() < 0.3:
        pd[-1] |= DEL_START
    src = bytes(hap)
    s = int(rng.integers(0, max(1, H - R + 1)))
    read = bytearray(src[s:s + R].ljust(R, b"A"))
    for k in rng.choice(R, max(1, R // 8), replace=False):
        read[k] = int(rng.choice(list(b"ACGTN")))
    quals = rng.integers(6, 41, R).astype(np.uint8)
    ins = rng.integers(20, 46, R).astype(np.uint8)
    dels = rng.integers(20, 46, R).astype(np.uint8)
    gcp = rng.integers(5, 16, R).astype(np.uint8)
    return bytes(hap), bytes(pd), np.frombuffer(bytes(read), np.uint8).copy(), quals, ins, dels, gcp


def test_skipping_non_special_cells_is_exact():
    rng = np.random.default_rng(5)
    n_differ_without_start = 0
    for k in range(60):
        H = int(rng.integers(4, 45))
        hap, pd, read, quals, ins, dels, gcp = _case(rng, H, int(rng.integers(0, 7)), int(rng.integers(0, 3)),
                                                     int(rng.integers(1, 25)))
        full = _restated(hap, pd, read, quals, ins, dels, gcp, skip=False)
        skip = _restated(hap, pd, read, quals, ins, dels, gcp, skip=True)
        assert skip == full, k
        want = oracle.pd_logless(hap, pd, read, quals, ins, dels, gcp)
        assert abs(full - want) <= 1e-12 * max(1.0, abs(want)), (k, full, want)
        if _restated(hap, pd, read, quals, ins, dels, gcp, skip=True, with_start=False) != full:
            n_differ_without_start += 1
    assert n_differ_without_start > 0
