"""gatk_amd -- a gfx950 (MI355X) PairHMM engine for GATK's read-likelihood path.

Modules:
  pairhmm  -- mirror of GATK's PairHMM / VectorLoglessPairHMM interface over the C-ABI
  batch    -- packed multi-region batches, device-resident batched forward (persistent kernel)
  synth    -- seeded synthetic active regions
  _native  -- ctypes binding of libgatk_pairhmm.so (include/gatk_pairhmm.h)
"""
__all__ = ["pairhmm", "batch", "synth"]
