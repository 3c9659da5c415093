/*
 * gatk_pairhmm.h -- C-ABI of the gfx950 PairHMM engine (libgatk_pairhmm.so).
 *
 * The first block is the drop-in boundary: it is exactly what GATK's native PairHMM binding calls
 * (interface org.broadinstitute.gatk.nativebindings.pairhmm.PairHMMNativeBinding, external jar
 * gatk-native-bindings 1.1.0, build.gradle:303), as evidenced by its only caller
 * src/main/java/org/broadinstitute/hellbender/utils/pairhmm/VectorLoglessPairHMM.java:
 *   load(File)                                     :63, :71   -> pairhmm_load
 *   initialize(PairHMMNativeArguments)             :81        -> pairhmm_initialize
 *   computeLikelihoods(ReadDataHolder[],
 *                      HaplotypeDataHolder[], double[]) :138  -> pairhmm_compute_likelihoods
 *   done()                                         :164       -> pairhmm_done
 * A JNI shim maps those four Java methods onto these functions one to one.
 *
 * Later blocks: new entry points with no counterpart in the reference (a device-resident batch API,
 * one persistent launch over many active regions; streaming; the backward (adjoint) API; tuning;
 * DRAGstr; read-quality preparation and likelihood post-processing on the device), and two more
 * drop-in boundaries of GATK's native libraries: GKL's IntelPDHMM (partially determined PairHMM,
 * VectorLoglessPairPDHMM.java:115) and GKL's SWAlignerNativeBinding (Smith-Waterman,
 * SWNativeAlignerWrapper.java:46).
 *
 * All functions return a pairhmm_status code; pairhmm_last_error() describes the last failure of
 * the calling thread.  Pointers are borrowed for the duration of the call only.
 */
#ifndef GATK_PAIRHMM_H
#define GATK_PAIRHMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The JNI shim turns PAIRHMM_ERR_NO_DEVICE into
 * UserException.HardwareFeatureException (VectorLoglessPairHMM.java:64-66, 72-74) and anything
 * else into GATKException. */
enum pairhmm_status {
    PAIRHMM_OK = 0,
    PAIRHMM_ERR_NO_DEVICE = 1, /* no usable gfx950 device                                   */
    PAIRHMM_ERR_BAD_ARG = 2,   /* invalid lengths / qualities (IllegalArgumentException)     */
    PAIRHMM_ERR_OOM = 3,       /* device or host allocation failed                           */
    PAIRHMM_ERR_KERNEL = 4,    /* a HIP launch or copy failed                                */
    PAIRHMM_ERR_STATE = 5      /* call on a context in the wrong state (IllegalStateException) */
};

/* ReadDataHolder (VectorLoglessPairHMM.java:122-127): five byte arrays of equal length.
 * Bases are ASCII bytes; qualities are raw Phred bytes (not +33).  As in PairHMMModel
 * (PairHMMModel.java:108-111) insertion/deletion/GCP bytes >= 128 are negative Java bytes and
 * are rejected; a base quality byte of 255 is rejected (QualityUtils.java:48 holds 255 entries). */
typedef struct pairhmm_read_data_holder {
    const uint8_t *read_bases;
    const uint8_t *read_quals;
    const uint8_t *insertion_gop;
    const uint8_t *deletion_gop;
    const uint8_t *overall_gcp;
    int32_t length;
} pairhmm_read_data_holder;

/* HaplotypeDataHolder (VectorLoglessPairHMM.java:97-98). */
typedef struct pairhmm_haplotype_data_holder {
    const uint8_t *haplotype_bases;
    int32_t length;
} pairhmm_haplotype_data_holder;

/* PairHMMNativeArguments (PairHMMNativeArgumentCollection.java:18-22). */
typedef struct pairhmm_native_arguments {
    int32_t max_number_of_threads; /* accepted and ignored: the device supplies the parallelism */
    int32_t use_double_precision;  /* nonzero: every pair runs the FP64 kernel                  */
} pairhmm_native_arguments;

typedef struct pairhmm_context pairhmm_context;

/* ---- drop-in boundary ------------------------------------------------------------------- */

/* PairHMMNativeBinding.load(File): 1 if a gfx950 device is usable by this process, else 0. */
int32_t pairhmm_load(void);

/* PairHMMNativeBinding.initialize(PairHMMNativeArguments): creates a context (own HIP stream,
 * device tables, staging buffers).  Contexts are independent and may be used from different
 * threads concurrently; one context must not be used by two threads at once. */
int32_t pairhmm_initialize(const pairhmm_native_arguments *args, pairhmm_context **out_ctx);

/* PairHMMNativeBinding.computeLikelihoods(reads, haplotypes, likelihoods):
 * likelihoods[r * n_haplotypes + h] = log10 P(read r | haplotype h) (read-major, the layout
 * VectorLoglessPairHMM.java:131,141-152 reads).  Synchronous: returns after the results are in
 * host memory.  FP32 cells with FP64 recompute of pairs whose FP32 sum falls under the underflow
 * threshold, or FP64 throughout when use_double_precision was set or a haplotype is longer than
 * 4,832 bases.  Haplotypes longer than 32,608 bases are PAIRHMM_ERR_BAD_ARG.  Values are <= 0 or
 * -inf. */
int32_t pairhmm_compute_likelihoods(pairhmm_context *ctx, const pairhmm_read_data_holder *reads,
                                    int32_t n_reads, const pairhmm_haplotype_data_holder *haplotypes,
                                    int32_t n_haplotypes, double *likelihoods);

/* PairHMMNativeBinding.done(): releases the context and everything it holds. */
int32_t pairhmm_done(pairhmm_context *ctx);

/* Message for the last non-OK status returned on the calling thread. */
const char *pairhmm_last_error(void);

/* ---- batch API (new; no counterpart in the reference) ------------------------------------ */

/* A packed batch of many active regions.  Reads and haplotypes are concatenated byte planes
 * indexed by per-item offsets and lengths; pair p scores read pair_read[p] against haplotype
 * pair_hap[p] and its result lands in out[p].  All arrays are host memory. */
typedef struct pairhmm_batch_desc {
    int64_t n_reads;
    const uint8_t *read_bases, *read_quals, *read_ins, *read_del, *read_gcp; /* read_bytes each */
    const int64_t *read_off;
    const int32_t *read_len;
    int64_t read_bytes;
    int64_t n_haps;
    const uint8_t *hap_bases; /* hap_bytes */
    const int64_t *hap_off;
    const int32_t *hap_len;
    int64_t hap_bytes;
    int64_t n_pairs;
    const int32_t *pair_read;
    const int32_t *pair_hap;
} pairhmm_batch_desc;

typedef struct pairhmm_batch pairhmm_batch;

/* Validates the batch, copies it into device memory (HBM-resident until pairhmm_batch_free) and
 * builds the work queue (pairs ordered by descending cell cost). */
int32_t pairhmm_batch_upload(pairhmm_context *ctx, const pairhmm_batch_desc *desc, pairhmm_batch **out_batch);

/* Run modes for pairhmm_batch_forward. */
enum pairhmm_run_mode {
    PAIRHMM_RUN_PERSISTENT = 0, /* persistent FP32 kernel fed by a device queue + FP64 rescue,
                                   replayed from a captured hipGraph                          */
    PAIRHMM_RUN_PERSISTENT_EAGER = 1, /* same kernels launched without the graph, each timed */
    PAIRHMM_RUN_FP64 = 2              /* every pair through the FP64 kernel                   */
};

typedef struct pairhmm_run_stats {
    double total_ms;        /* wall time of the device work (HIP events), excluding D2H        */
    double fp32_kernel_ms;  /* duration of the FP32 forward kernel (eager mode only)           */
    double rescue_kernel_ms;/* duration of the FP64 rescue kernel (eager mode only)            */
    int64_t n_rescued;      /* pairs computed by the FP64 kernel: flagged by the FP32 pass, or
                               sent there directly because their read's qualities predict an
                               FP32 underflow (expected log10 emission below -90) or their
                               haplotype is too long for the FP32 kernels (over 4,832 bases)    */
    int64_t cells;          /* sum over pairs of read length x haplotype length                */
    int64_t n_routed_fp64;  /* of n_rescued: pairs sent to the FP64 kernel without an FP32 pass
                               (the rest were flagged by it)                                    */
} pairhmm_run_stats;

/* Runs the forward pass over the whole batch.  out_host (n_pairs doubles, may be NULL) receives
 * log10 likelihoods in pair order; stats (may be NULL) receives timings. */
int32_t pairhmm_batch_forward(pairhmm_context *ctx, pairhmm_batch *batch, int32_t mode, double *out_host,
                              pairhmm_run_stats *stats);

/* Releases a batch.  Batches (and streams) must be released before pairhmm_done of their context. */
int32_t pairhmm_batch_free(pairhmm_batch *batch);

/* Launch schedule of the forward (new).  Applied by later pairhmm_batch_upload, stream and
 * computeLikelihoods calls on the context; gatk_amd/autotune.py searches it per device and
 * workload. */
typedef struct pairhmm_tuning {
    int32_t waves_per_block; /* wavefronts per workgroup, 1..8 (0: 4)                          */
    int32_t grid_percent;    /* persistent grid as a percentage of resident capacity (0: 100)  */
    int32_t queue_order;     /* 0: pairs by descending cost (default), 1: submission order     */
    int32_t stripe_schedule; /* FP32 pass: 0 packed stripes (rows of consecutive pairs share a
                                wave's 64-row stripes; default), 1 aligned (one pair per stripe,
                                the read's last stripe padded).  Batches whose haplotypes hold the
                                bytes 0 or 1 always run aligned.                                 */
    int32_t step_unroll;     /* packed FP32 kernel: wave-steps per unrolled loop body, 8 or 16;
                                0 (default): per LDS class, 16 where its extra registers cost no
                                occupancy (long haplotypes), else 8                              */
} pairhmm_tuning;

int32_t pairhmm_set_tuning(pairhmm_context *ctx, const pairhmm_tuning *tuning);
int32_t pairhmm_get_tuning(pairhmm_context *ctx, pairhmm_tuning *tuning);

/* ---- streaming (new; no counterpart in the reference) ------------------------------------- */

/* Double-buffered streaming of region batches: pairhmm_stream_submit validates and packs a chunk
 * (same layout as pairhmm_batch_desc, host memory) into pinned staging and returns while the device
 * still computes the previous chunk; the H2D copy overlaps that compute.  The chunk's arrays may be
 * released when submit returns; out_host (n_pairs doubles) must stay valid until it is filled,
 * which happens when its slot is reused (two submits later) or at pairhmm_stream_finish. */
typedef struct pairhmm_stream pairhmm_stream;

int32_t pairhmm_stream_create(pairhmm_context *ctx, pairhmm_stream **out_stream);
int32_t pairhmm_stream_submit(pairhmm_stream *stream, const pairhmm_batch_desc *chunk, double *out_host);
/* Waits for every submitted chunk; stats: host wall time since the first submit, cells, rescued. */
int32_t pairhmm_stream_finish(pairhmm_stream *stream, pairhmm_run_stats *stats);
int32_t pairhmm_stream_destroy(pairhmm_stream *stream);

/* ---- forward + backward (new; no counterpart in the reference) ------------------------------ */

/* Differentiable forward over continuous per-position Phred parameters.  With e(q) = 10^(-q/10):
 * MM = 1 - e(qi) - e(qd), MI = e(qi), MD = e(qd), GM = 1 - e(qc), II = DD = e(qc), prior =
 * 1 - e(qb) on a match and e(qb)/3 otherwise; the recurrence is LoglessPairHMM's
 * (LoglessPairHMM.java:47-57).  All pointers are DEVICE pointers on the context's device. */
typedef struct pairhmm_grad_desc {
    int64_t n_reads, n_haps, n_pairs;
    int64_t read_bytes;
    const uint8_t *read_bases;
    const float *read_qb, *read_qi, *read_qd, *read_qc; /* Phred per read position (read_bytes) */
    const int64_t *read_off;
    const int32_t *read_len;
    const uint8_t *hap_bases;
    const int64_t *hap_off;
    const int32_t *hap_len;
    const int32_t *pair_read, *pair_hap;
    const double *pair_weight;        /* w_p, may be NULL (all 1)                              */
    int32_t max_read_len, max_hap_len; /* bounds over the batch (size the scratch and LDS)     */
    int32_t use_double;               /* nonzero: every pair through the FP64 kernel            */
    double *out_log10_likelihood;     /* [n_pairs]                                             */
    double *grad_qb, *grad_qi, *grad_qd, *grad_qc; /* [read_bytes]: sum_p w_p d log10 L_p / d q */
    double *grad_base_error;          /* [read_bytes] or NULL: sum_p w_p d log10 L_p / d e(qb)  */
} pairhmm_grad_desc;

/* Runs the forward pass and the adjoint over every pair; overwrites the outputs.  FP32 cells with
 * FP32 adjoint sums and FP64 per-position reductions; pairs whose scaled likelihood is too small for
 * the FP32 adjoint range are recomputed in FP64.  Synchronous. */
int32_t pairhmm_forward_backward(pairhmm_context *ctx, const pairhmm_grad_desc *desc, pairhmm_run_stats *stats);

/* pairhmm_forward_backward without allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments, results and outputs as pairhmm_forward_backward, but the
 * work lists and checkpoints live in one fixed device array of 256 MiB per GPU (allocated when the
 * library's code object is loaded on that GPU and held until the process ends) instead of the
 * context's growable buffers; no HIP allocation, free or event creation happens in the call.  Each
 * persistent grid is held to the waves whose checkpoints fit, so large problems run at lower
 * occupancy.  Fails with PAIRHMM_ERR_BAD_ARG when the pair list (4 bytes x (64 + 2 n_pairs)) or
 * one workgroup's checkpoints (set by max_read_len and max_hap_len) exceed the array.  Every call on
 * a GPU shares that GPU's array and leaves its contents behind: calls that overlap on one GPU (two
 * threads, two contexts) race and must be serialised by the caller.  stats->total_ms and the kernel
 * times are 0 (no events are created). */
int32_t pairhmm_forward_backward_noalloc(pairhmm_context *ctx, const pairhmm_grad_desc *desc,
                                         pairhmm_run_stats *stats);

/* ---- DRAGstr input side (host code; no device needed) -------------------------------------- */

/* Per read position: the period (1..max_period) with the most tandem copies overlapping the
 * position or starting right after it, shorter period on ties, and that copy count.
 * Replaces DragstrReadSTRAnalyzer.of(bases, maxPeriod).mostRepeatedPeriod(i) /
 * numberOfMostRepeats(i) (DragstrReadSTRAnalyzer.java:69-139). */
int32_t pairhmm_dragstr_analyze(const uint8_t *bases, int32_t length, int32_t max_period, int32_t *out_period,
                                int32_t *out_repeats);

/* Context index per read position of a packed read set: (min(period, P) - 1) * R + min(repeats, R) - 1
 * for every position but the last of each read, -1 for the last (whose penalties are fixed).  The
 * index is the cell DragstrParams.gop/gcp(period, repeats) reads (DragstrParams.java:194-200). */
int32_t pairhmm_dragstr_contexts(const uint8_t *read_bases, const int64_t *read_off, const int32_t *read_len,
                                 int64_t n_reads, int64_t read_bytes, int32_t max_period, int32_t max_repeats,
                                 int32_t *out_context);

/* DragstrPairHMMInputScoreImputator.impute for a packed read set
 * (DragstrPairHMMInputScoreImputator.java:53-70): GOP = round(min(40, gop[period][repeats])), used
 * for insertions and deletions alike, GCP = round(gcp[period][repeats]), last position 45 / 10.
 * Tables are row-major [max_period][max_repeats] Phred values. */
int32_t pairhmm_dragstr_impute(const double *gop_table, const double *gcp_table, int32_t max_period,
                               int32_t max_repeats, const uint8_t *read_bases, const int64_t *read_off,
                               const int32_t *read_len, int64_t n_reads, int64_t read_bytes, uint8_t *out_gop,
                               uint8_t *out_gcp);

/* ---- Read-quality preparation on the device ------------------------------------------------- */

/* PCR indel error models (PairHMMLikelihoodCalculationEngine.PCRErrorModel,
 * PairHMMLikelihoodCalculationEngine.java:53-61); the value is the rate factor. */
typedef enum pairhmm_pcr_error_model {
    PAIRHMM_PCR_NONE = 0,
    PAIRHMM_PCR_HOSTILE = 1,
    PAIRHMM_PCR_AGGRESSIVE = 2,
    PAIRHMM_PCR_CONSERVATIVE = 3
} pairhmm_pcr_error_model;

/* PairHMMLikelihoodCalculationEngine.applyPCRErrorModel (PairHMMLikelihoodCalculationEngine.java:360-371)
 * for a packed read set: at positions 0 .. L-2 of every read, the insertion and deletion GOP become
 * min(GOP, getErrorModelAdjustedQual(repeat length, rate factor)) (:356-358), the repeat length being
 * ReadLikelihoodCalculationEngine.findTandemRepeatUnits (ReadLikelihoodCalculationEngine.java:193-253).
 * All array pointers are DEVICE pointers; ins_quals / del_quals are updated in place.  Synchronous on
 * the context's stream. */
int32_t pairhmm_apply_pcr_error_model(pairhmm_context *ctx, const uint8_t *read_bases, const int64_t *read_off,
                                      const int32_t *read_len, int64_t n_reads, int32_t pcr_model, uint8_t *ins_quals,
                                      uint8_t *del_quals);

/* PairHMMLikelihoodCalculationEngine.capMinimumReadQualities (:307-320) for a packed read set: base
 * quality capped by the read's mapping quality (mapping_quality[n_reads], DEVICE, ignored and may be
 * NULL when disable_cap_to_mapq is nonzero), then set to 6 (QualityUtils.MIN_USABLE_Q_SCORE) where it
 * is below base_quality_threshold; insertion / deletion quality set to 6 where below 6.  Comparisons
 * are Java signed-byte comparisons, as in setToFixedValueIfTooLow.  DEVICE pointers, in place,
 * synchronous. */
int32_t pairhmm_cap_read_qualities(pairhmm_context *ctx, const int64_t *read_off, const int32_t *read_len,
                                   int64_t n_reads, const int32_t *mapping_quality, int32_t base_quality_threshold,
                                   int32_t disable_cap_to_mapq, uint8_t *base_quals, uint8_t *ins_quals,
                                   uint8_t *del_quals);

/* DragstrPairHMMInputScoreImputator.impute (DragstrPairHMMInputScoreImputator.java:53-70) on the device:
 * the same outputs as pairhmm_dragstr_impute for reads held in DEVICE memory (read_bases, read_off,
 * read_len, out_gop, out_gcp are device pointers; the tables are host arrays, at most 512 cells).
 * max_read_len bounds the reads (at most 4096).  Synchronous on the context's stream. */
int32_t pairhmm_dragstr_impute_device(pairhmm_context *ctx, const double *gop_table, const double *gcp_table,
                                      int32_t max_period, int32_t max_repeats, const uint8_t *read_bases,
                                      const int64_t *read_off, const int32_t *read_len, int64_t n_reads,
                                      int32_t max_read_len, uint8_t *out_gop, uint8_t *out_gcp);

/* ---- Partially determined PairHMM (drop-in for GKL's IntelPDHMM) -------------------------- */

/* HaplotypeDataHolder of the PD binding (VectorLoglessPairPDHMM.java:101-106): the haplotype bases
 * and, per base, its PartiallyDeterminedHaplotype byte (PartiallyDeterminedHaplotype.java:59-66:
 * SNP = 1, DEL_START = 2, DEL_END = 4, alternative A/C/G/T = 8/16/32/64). */
typedef struct pairhmm_pd_haplotype_data_holder {
    const uint8_t *haplotype_bases;
    const uint8_t *haplotype_pd_bases;
    int32_t length;
} pairhmm_pd_haplotype_data_holder;

/* IntelPDHMM.computeLikelihoods(ReadDataHolder[], HaplotypeDataHolder[], double[])
 * (VectorLoglessPairPDHMM.java:115): likelihoods[r * n_haplotypes + h] = the LoglessPDPairHMM log10
 * likelihood (LoglessPDPairHMM.java:34-150), FP64 throughout.  IntelPDHMM.load / initialize / done
 * (VectorLoglessPairPDHMM.java:55-60) map to pairhmm_load / pairhmm_initialize / pairhmm_done; the
 * PDHMMNativeArguments (thread count, AVX level, OpenMP setting) have no device meaning.  A read
 * base outside ACGTacgtN meeting a SNP column whose base is neither it nor 'N' is
 * PAIRHMM_ERR_BAD_ARG, where LoglessPDPairHMM.isBasePDMatching throws (:195-196), and so is a
 * haplotype of more than 8,060 bases (one wave's LDS holds 20 bytes per column).  Synchronous. */
int32_t pairhmm_pd_compute_likelihoods(pairhmm_context *ctx, const pairhmm_read_data_holder *reads,
                                       int32_t n_reads, const pairhmm_pd_haplotype_data_holder *haplotypes,
                                       int32_t n_haplotypes, double *likelihoods);

/* ---- Likelihood post-processing on the device ---------------------------------------------- */

/* The two steps PairHMMLikelihoodCalculationEngine.computeReadLikelihoods applies to the finished
 * read x haplotype matrix (PairHMMLikelihoodCalculationEngine.java:200-201), on a read-major matrix
 * in DEVICE memory (likelihoods[r * n_alleles + a], the computeLikelihoods layout; samples simply
 * partition the reads):
 *   AlleleLikelihoods.normalizeLikelihoods(log10_mismapping_cap, symmetric) (AlleleLikelihoods.java:
 *   416-458), in place; log10_mismapping_cap must be negative, -INFINITY skips the step; the
 *   reference allele (reference_allele_index, -1 if none) is excluded from the best-allele search
 *   unless symmetric (:505-533);
 *   ReadLikelihoodCalculationEngine.filterPoorlyModeledEvidence(dynamic_disqualification,
 *   expected_error_rate_per_base, disqualification_scale) (ReadLikelihoodCalculationEngine.java:
 *   98-155): keep[r] = 0 where read r's best likelihood is below its threshold, else 1; threshold[r]
 *   (may be NULL) receives the threshold.  hmm_base_quals / read_off / read_len describe each read's
 *   HMM base qualities (the dynamic threshold reads them; their length is the qualified read length).
 * DEVICE pointers; synchronous on the context's stream.  No counterpart in the native binding: the
 * reference runs these steps in Java on the host. */
int32_t pairhmm_normalize_and_filter_likelihoods(pairhmm_context *ctx, double *likelihoods, int64_t n_reads,
                                                 int32_t n_alleles, int32_t reference_allele_index,
                                                 double log10_mismapping_cap, int32_t symmetric,
                                                 const uint8_t *hmm_base_quals, const int64_t *read_off,
                                                 const int32_t *read_len, double expected_error_rate_per_base,
                                                 int32_t dynamic_disqualification, double disqualification_scale,
                                                 uint8_t *keep, double *threshold);

/* ---- Smith-Waterman (drop-in for GKL's SWAlignerNativeBinding) ----------------------------- */

/* SWParameters (gatk-native-bindings; SmithWatermanAlignmentConstants.java:30-68 holds the presets). */
typedef struct pairhmm_sw_parameters {
    int32_t match_value;
    int32_t mismatch_penalty;
    int32_t gap_open_penalty;
    int32_t gap_extend_penalty;
} pairhmm_sw_parameters;

/* SWOverhangStrategy, in the binding's declaration order. */
typedef enum pairhmm_sw_overhang_strategy {
    PAIRHMM_SW_SOFTCLIP = 0,
    PAIRHMM_SW_INDEL = 1,
    PAIRHMM_SW_LEADING_INDEL = 2,
    PAIRHMM_SW_IGNORE = 3
} pairhmm_sw_overhang_strategy;

/* SWAlignerNativeBinding.align(ref, alt, parameters, overhangStrategy) -> SWNativeAlignerResult
 * (SWNativeAlignerWrapper.java:46; SmithWatermanIntelAligner.java:35): the CIGAR as text into cigar
 * (cigar_cap bytes including the terminating 0) and the alignment offset.  The result is that of
 * SmithWatermanJavaAligner's full alignment (SmithWatermanJavaAligner.java:104-377); the substring
 * shortcut of the callers (SWNativeAlignerWrapper.java:35-43) is theirs to take, as with GKL.  Empty
 * sequences are PAIRHMM_ERR_BAD_ARG (:55-57); sequences up to 32767 bases.  Synchronous. */
int32_t pairhmm_sw_align(pairhmm_context *ctx, const uint8_t *ref, int32_t ref_len, const uint8_t *alt,
                         int32_t alt_len, const pairhmm_sw_parameters *parameters, int32_t overhang_strategy,
                         char *cigar, int32_t cigar_cap, int32_t *alignment_offset);

/* Batch form (new; no counterpart in the reference): pair p aligns alt_bases[alt_off[p] ..] to
 * ref_bases[ref_off[p] ..] with one set of parameters; its CIGAR text goes to cigars + cigar_off[p]
 * (cigar_cap[p] bytes) and its offset to alignment_offset[p].  Host memory; one launch. */
int32_t pairhmm_sw_align_batch(pairhmm_context *ctx, int64_t n_pairs, const uint8_t *ref_bases,
                               const int64_t *ref_off, const int32_t *ref_len, const uint8_t *alt_bases,
                               const int64_t *alt_off, const int32_t *alt_len,
                               const pairhmm_sw_parameters *parameters, int32_t overhang_strategy, char *cigars,
                               const int64_t *cigar_off, const int32_t *cigar_cap, int32_t *alignment_offset);

/* Packed batch form (new; no counterpart in the reference): as pairhmm_sw_align_batch, but the CIGARs
 * are written back to back, without separators, into cigar_text (cigar_text_cap bytes; one byte past
 * the last CIGAR receives a 0).  Pair p's CIGAR is cigar_text[cigar_end[p-1] .. cigar_end[p]), with
 * cigar_end[-1] = 0.  A caller decodes the text once instead of one buffer per pair; too small a
 * buffer is PAIRHMM_ERR_BAD_ARG. */
int32_t pairhmm_sw_align_batch_packed(pairhmm_context *ctx, int64_t n_pairs, const uint8_t *ref_bases,
                                      const int64_t *ref_off, const int32_t *ref_len, const uint8_t *alt_bases,
                                      const int64_t *alt_off, const int32_t *alt_len,
                                      const pairhmm_sw_parameters *parameters, int32_t overhang_strategy,
                                      char *cigar_text, int64_t cigar_text_cap, int64_t *cigar_end,
                                      int32_t *alignment_offset);

#ifdef __cplusplus
}
#endif

#endif /* GATK_PAIRHMM_H */
