"""gatk_amd -- a gfx950 (MI355X) PairHMM engine for GATK's read-likelihood path.

Modules:
  pairhmm   -- mirror of GATK's PairHMM / VectorLoglessPairHMM interface over the C-ABI
  batch     -- packed multi-region batches, device-resident batched forward (persistent kernel)
  stream    -- double-buffered streaming of region chunks through the batched forward
  autograd  -- forward + adjoint kernel as a torch.autograd.Function
  autotune  -- launch-schedule autotuner of the batched forward
  dragstr   -- DRAGstr STR contexts, parameter tables and GOP / GCP imputation
  strfit    -- per-(STR period, repeats) GOP / GCP fit with the adjoint kernel
  synth     -- seeded synthetic active regions
  _native   -- ctypes binding of libgatk_pairhmm.so (include/gatk_pairhmm.h)
"""
__all__ = ["pairhmm", "batch", "stream", "autograd", "autotune", "dragstr", "strfit", "synth"]
