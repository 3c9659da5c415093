// pairhmm_kernels.hpp -- device-side data views and launch helpers of the PairHMM forward kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include <cstdint>

namespace gatk_pairhmm {

constexpr int kWave = 64;      // CDNA wavefront width: one read row per lane
constexpr int kLdsPad = 64;    // boundary-row / haplotype slots before column 0 (virtual columns)

// Packed batch resident in device memory.  Reads are five byte planes sharing rd_off/rd_len;
// haplotypes one plane.  When pair_read is null the pairs are implicit and read-major:
// pair p = r * n_haps + h (the computeLikelihoods output layout).
struct DeviceBatch {
    const uint8_t *hap_bases;
    const int64_t *hap_off;
    const int32_t *hap_len;
    const uint8_t *rd_bases, *rd_quals, *rd_ins, *rd_del, *rd_gcp;
    const int64_t *rd_off;
    const int32_t *rd_len;
    const int32_t *pair_read;
    const int32_t *pair_hap;
    int32_t n_haps;  // used only for implicit pairs
    double *out;     // log10 likelihood per pair
    // FP64 launches: each wave's two boundary planes (2 x (kCols + 4) doubles, wave w at
    // w * 2 * (kCols + 4)) live in global memory; waves past f64_plane_waves stay idle.
    double *f64_planes;
    uint32_t f64_plane_waves;
};

// Qual -> probability tables in device memory (see model_tables.hpp).
struct DeviceTables {
    const double *qual_err;  // 512 entries: 10^(q/-10) at q, that value / 3 at 256 + q
    const double *mm_prob;   // triangular matchToMatch table
};

// Where a launch takes its pairs from.
struct WorkSource {
    const uint32_t *items;        // item -> pair index; null: item is the pair index
    uint32_t n_items;             // item count when n_items_dev is null
    const uint32_t *n_items_dev;  // item count read on the device (e.g. the rescue count)
    uint32_t *head;               // non-null: dynamic queue (atomic head); null: grid-stride
    uint32_t chunk = 0;           // packed FP32 schedule: items per queue chunk (set at launch)
};

// FP32 launches append pairs whose scaled last-row sum is below `threshold` to `list`.
struct RescueSink {
    uint32_t *list;   // null: no rescue (FP64 launches)
    uint32_t *count;
    double threshold;
};

// Column capacity classes of a wave's LDS slice (haplotype bytes + boundary planes).  The class is
// a template parameter of every kernel so each LDS access is one base register plus an immediate
// offset.  A class holds columns -64..H+63+8 plus slack: H + 160 <= class.  Classes up to
// kMaxFp32Cols run the FP32 kernels (their planes are in LDS); the wider ones only the FP64 kernel,
// whose LDS holds just the haplotype (its planes are in global memory), so haplotypes of up to
// 32,608 bases are taken, their pairs computed in FP64 throughout.
inline constexpr int kColClasses[] = {384, 512, 640, 768, 1024, 1280, 1664, 2560, 4992, 8192, 16384, 32768};
inline constexpr int kNumColClasses = static_cast<int>(sizeof(kColClasses) / sizeof(kColClasses[0]));
inline constexpr int kMaxFp32Cols = 4992;

// Calls f(std::integral_constant<int, C>{}) for the class C == cols; hipErrorInvalidValue if cols
// is not a class.  Every kernel family instantiates through this, so all share one class list.
template <int I = 0, typename F>
hipError_t dispatch_col_class(int cols, F &&f) {
    if constexpr (I == kNumColClasses) {
        return hipErrorInvalidValue;
    } else {
        if (cols == kColClasses[I]) return f(std::integral_constant<int, kColClasses[I]>{});
        return dispatch_col_class<I + 1>(cols, static_cast<F &&>(f));
    }
}

// Elements between the boundary planes of a wave's LDS slice, host and device alike: 16-byte
// aligned (vector reads of lane 0's sums) and not a multiple of 64 dwords (single ds_write_b32 off
// one base register rather than ds_write2st64 pairs).
constexpr int lds_plane_stride(int cols) { return cols + 4; }

// Bytes of LDS one wavefront needs for haplotypes up to max_hap_len.
int lds_columns(int max_hap_len);
// LDS of one wave: FP32 aligned stripes (haplotype + two planes), FP32 packed stripes (two haplotype
// buffers + two planes) or FP64 (haplotype; its planes live in global memory).
size_t lds_bytes_per_wave(int max_hap_len, bool fp64, bool packed = false);
// Items per queue chunk of a packed FP32 launch of `waves` waves over n_items items (1..16).
uint32_t packed_chunk(int64_t n_items, int64_t waves);
// Index of the LDS column class that holds haplotypes of this length (0 = smallest), -1 if none.
// Queues split by class let short haplotypes run at the occupancy their smaller LDS slice allows.
int lds_class_index(int hap_len);

// Dynamic LDS budget per workgroup: the device's limit (queried once), at most 160 KB (gfx950).
size_t max_lds_per_block();

// Launch geometry of the forward kernel for one batch.
struct LaunchShape {
    int waves_per_block = 0;  // 0: the haplotype is too long for one wave's LDS slice
    int blocks_per_cu = 0;    // resident blocks per CU at this LDS size (occupancy API)
    size_t lds_bytes = 0;     // dynamic LDS per block
    bool packed = false;      // FP32 packed-stripe schedule (rows of consecutive pairs share stripes)
    int unroll = 8;           // packed schedule: wave-steps per unrolled loop body (8 or 16)
};

// Largest waves-per-block <= requested whose LDS fits kMaxLdsPerBlock, and its occupancy.  packed
// selects the FP32 packed-stripe kernel (ignored for FP64), which takes its pairs from a queue;
// unroll picks its loop body: 8 or 16 wave-steps, or 0 for 16 where that keeps the occupancy of 8.
LaunchShape forward_shape(bool fp64, int max_hap_len, int requested_waves_per_block, bool packed = false,
                          int unroll = 0);

// Bytes of global boundary planes one FP64 wave needs for haplotypes up to max_hap_len (0 if too long).
size_t fp64_plane_bytes_per_wave(int max_hap_len);

hipError_t launch_forward(bool fp64, const DeviceBatch &batch, const DeviceTables &tables, const WorkSource &work,
                          const RescueSink &rescue, int max_hap_len, const LaunchShape &shape, int grid_blocks,
                          hipStream_t stream);

// Zeroes n counters with a kernel (graph-capturable; no memset node).
hipError_t launch_zero_counters(uint32_t *counters, int n, hipStream_t stream);
// Sets n counters from a device array with a kernel (graph-capturable).
hipError_t launch_copy_counters(uint32_t *counters, const uint32_t *init, int n, hipStream_t stream);

// ---- forward + backward (adjoint) over continuous Phred parameters --------------------------

constexpr int kGradTile = 8;  // wave-steps per recompute tile of the adjoint

// Device pointers of a gradient batch.  Read parameters are float Phred values per read position.
struct GradBatch {
    const uint8_t *hap_bases;
    const int64_t *hap_off;
    const int32_t *hap_len;
    const uint8_t *rd_bases;
    const float *rd_qb, *rd_qi, *rd_qd, *rd_qc;
    const int64_t *rd_off;
    const int32_t *rd_len;
    const int32_t *pair_read;
    const int32_t *pair_hap;
    const double *pair_weight;  // may be null (weight 1)
    double *out_log10;          // [n_pairs]
    double *g_qb, *g_qi, *g_qd, *g_qc;  // [read bytes], accumulated
    double *g_eb;               // [read bytes] or null
    void *scratch;              // per-wave checkpoints and relay rows
    size_t scratch_per_wave;    // elements (float or double) per wave
    int max_stripes, max_tiles;
};

size_t grad_lds_bytes_per_wave(int max_hap_len, bool fp64);
size_t grad_scratch_elems_per_wave(int max_read_len, int max_hap_len, int *max_stripes, int *max_tiles);
LaunchShape grad_shape(bool fp64, int max_hap_len, int requested_waves_per_block);
hipError_t launch_grad(bool fp64, const GradBatch &g, const WorkSource &work, const RescueSink &rescue, int max_hap_len,
                       const LaunchShape &shape, int grid_blocks, hipStream_t stream);

// Device-side binning of the adjoint's pairs by LDS column class, so that each adjoint launch gets
// the LDS slice (and so the occupancy) of its own haplotypes rather than the batch's longest.
// Classes above max_class count as max_class.  Counting adds the per-class pair counts to counts[];
// scattering places each pair at run_begin[run] + an atomic cursor of its run, where a run is one
// launch covering one or more merged classes.
struct ClassRuns {
    int32_t run_of_class[kNumColClasses];
    uint32_t run_begin[kNumColClasses];
};
hipError_t launch_grad_class_count(const int32_t *pair_hap, const int32_t *hap_len, uint32_t n_pairs, int max_class,
                                   uint32_t *counts, hipStream_t stream);
hipError_t launch_grad_class_scatter(const int32_t *pair_hap, const int32_t *hap_len, uint32_t n_pairs, int max_class,
                                     const ClassRuns &runs, uint32_t *cursors, uint32_t *order, hipStream_t stream);

// Read-quality preparation (readprep.hip).  table.q[k] is the PCR-model quality for repeat length k.
struct PcrQualTable {
    uint8_t q[21];
};
hipError_t launch_pcr_error_model(const uint8_t *bases, const int64_t *read_off, const int32_t *read_len,
                                  int64_t n_reads, const PcrQualTable &table, uint8_t *ins, uint8_t *del,
                                  hipStream_t stream);
hipError_t launch_cap_qualities(const int64_t *read_off, const int32_t *read_len, int64_t n_reads, const int32_t *mapq,
                                int32_t bq_threshold, int32_t disable_mapq_cap, uint8_t *quals, uint8_t *ins,
                                uint8_t *del, hipStream_t stream);

// DRAGstr imputation on the device: per-(period, repeats) GOP / GCP bytes, already rounded and
// capped as the imputer does, row-major [max_period][max_repeats].
struct DragstrByteTables {
    int32_t max_period, max_repeats;
    uint8_t gop[512], gcp[512];
};
hipError_t launch_dragstr_impute(const uint8_t *bases, const int64_t *read_off, const int32_t *read_len,
                                 int64_t n_reads, const DragstrByteTables &tables, uint8_t *out_gop, uint8_t *out_gcp,
                                 hipStream_t stream);
int dragstr_max_read_length();

// ---- partially determined PairHMM (pdhmm.hip) ----------------------------------------------

// Column code of a partially determined haplotype (built on the host from the haplotype base and
// its PartiallyDeterminedHaplotype byte): bits 0-7 the base; bits 8-15 the bases the column also
// accepts (the SNP alternatives A/C/G/T at the PD byte's own bit positions, or every class bit for
// an 'N' base); kPdDelEnd when the base ends a deletion span; kPdInsideRow1 / kPdInsideLater when
// row 1 / rows >= 2 process the column in state INSIDE_DEL; kPdSpecial on every column that carries
// kPdDelEnd or an inside bit and on the column before any column with an inside bit (a deletion
// start): the only columns where a row's branch values are read or must be set.
constexpr uint32_t kPdAltA = 8, kPdAltC = 16, kPdAltG = 32, kPdAltT = 64;
constexpr uint32_t kPdDelEnd = 1u << 24, kPdInsideRow1 = 1u << 25, kPdInsideLater = 1u << 26;
constexpr uint32_t kPdSpecial = 1u << 27;

// Every read against every haplotype; out[r * n_haps + h] (read-major).
struct PdBatch {
    const uint32_t *hap_code;  // column codes, haplotype h at hap_off[h]
    const int64_t *hap_off;
    const int32_t *hap_len;
    const uint8_t *rd_bases, *rd_quals, *rd_ins, *rd_del, *rd_gcp;
    const int64_t *rd_off;
    const int32_t *rd_len;
    int32_t n_haps;
    int32_t tristate;
    double *out;
    const uint32_t *order;     // launch order of the pairs (longest first), or null
};

// LDS of one wave (one workgroup) for haplotypes up to max_hap_len: a 16-byte record and a 4-byte
// column code per column.
size_t pd_lds_bytes(int max_hap_len);
hipError_t launch_pd_forward(const PdBatch &batch, const DeviceTables &tables, uint32_t n_pairs, int max_hap_len,
                             int grid_blocks, hipStream_t stream);

// ---- likelihood post-processing (postprocess.hip) ---------------------------------------

struct PostprocessArgs {
    double *likelihoods;        // [n_reads * n_alleles], read-major, normalised in place
    int64_t n_reads;
    int32_t n_alleles;
    int32_t ref_index;          // reference allele, -1 if none
    int32_t normalize;          // 0: cap is -inf, no normalisation
    int32_t symmetric;
    double cap;
    const uint8_t *base_quals;  // HMM base qualities (dynamic threshold only)
    const int64_t *read_off;
    const int32_t *read_len;
    double max_error_per_base;
    int32_t dynamic;
    double scale;
    uint8_t *keep;              // [n_reads]
    double *threshold;          // [n_reads] or null
};
hipError_t launch_postprocess(const PostprocessArgs &args, hipStream_t stream);

// ---- Smith-Waterman (sw.hip) ----------------------------------------------------------------

struct SwBatch {
    const uint8_t *ref_bases;
    const int64_t *ref_off;
    const int32_t *ref_len;
    const uint8_t *alt_bases;
    const int64_t *alt_off;
    const int32_t *alt_len;
    int32_t match, mismatch, open, extend;  // SWParameters
    int32_t strategy;                       // 0 SOFTCLIP, 1 INDEL, 2 LEADING_INDEL, 3 IGNORE
    int32_t max_ref_len;
    uint8_t *scratch;                       // per wave: backtrack codes, last column, bottom row
    size_t scratch_per_wave, bt_bytes_per_wave;
    uint32_t *elems;                        // CIGAR elements (len << 2 | op), reference list order
    const int64_t *elem_off;                // per pair; capacity ref_len + alt_len + 4
    uint32_t *elems_short;                  // the first kSwShortElems elements of pair p at p * kSwShortElems
    int32_t *n_elems;
    int32_t *offset;                        // alignment offset
};
// Elements per pair in the compact copy the host reads back; a pair with more (rare: every element
// but a clip consumes an indel or a run of matches) is read from the full list.
constexpr int kSwShortElems = 16;
size_t sw_lds_bytes(int max_ref_len);
size_t sw_bt_bytes_per_wave(int max_ref_len, int max_alt_len);
size_t sw_scratch_bytes_per_wave(int max_ref_len, int max_alt_len);
hipError_t launch_sw(const SwBatch &batch, uint32_t n_pairs, int grid_blocks, hipStream_t stream);

}  // namespace gatk_pairhmm
