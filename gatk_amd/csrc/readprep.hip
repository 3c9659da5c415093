// readprep.hip -- HaplotypeCaller's read-quality preparation on the device, over packed read planes:
//   * the PCR indel error model: PairHMMLikelihoodCalculationEngine.applyPCRErrorModel
//     (tools/walkers/haplotypecaller/PairHMMLikelihoodCalculationEngine.java:360-371) caps the
//     insertion and deletion GOP at read positions 0 .. L-2 by a quality that falls with the tandem
//     repeat length there (ReadLikelihoodCalculationEngine.findTandemRepeatUnits,
//     ReadLikelihoodCalculationEngine.java:193-253, repeat units up to 8 bases, lengths capped at 20);
//   * the minimum-quality caps of capMinimumReadQualities (:307-320): base quality capped by the
//     read's mapping quality (unless disabled) and set to 6 below the base-quality threshold,
//     insertion / deletion quality set to 6 below 6 -- compared as Java signed bytes, so a byte of
//     128 or more counts as low.
// One wave per read, lanes over its positions; every thread writes only its own position.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pairhmm_kernels.hpp"

namespace gatk_pairhmm {
namespace {

constexpr int kMaxStrUnit = 8;
constexpr int kMaxRepeat = 20;

__device__ __forceinline__ bool same(const uint8_t *a, const uint8_t *b, int n) {
    for (int k = 0; k < n; ++k)
        if (a[k] != b[k]) return false;
    return true;
}

// GATKVariantContextUtils.findNumberOfRepetitions (GATKVariantContextUtils.java:976-1010): how many
// copies of unit[0, ulen) the test string [0, tlen) starts (leading) or ends (trailing) with.
__device__ int repetitions(const uint8_t *unit, int ulen, const uint8_t *test, int tlen, bool leading) {
    if (tlen == 0) return 0;
    const int diff = tlen - ulen;
    int n = 0;
    if (leading) {
        for (int start = 0; start <= diff; start += ulen) {
            if (!same(test + start, unit, ulen)) return n;
            ++n;
        }
    } else {
        for (int start = diff; start >= 0; start -= ulen) {
            if (!same(test + start, unit, ulen)) return n;
            ++n;
        }
    }
    return n;
}

// findTandemRepeatUnits(read, offset).getRight(): the repeat length at offset.
__device__ int tandem_repeat_length(const uint8_t *read, int len, int offset) {
    int max_bw = 0;
    const uint8_t *bw = read + offset;  // best backward unit (initially the base itself)
    int bw_len = 1;
    for (int s = 1; s <= kMaxStrUnit; ++s) {
        if (offset + 1 - s < 0) break;
        max_bw = repetitions(read + offset - s + 1, s, read, offset + 1, false);
        if (max_bw > 1) {
            bw = read + offset - s + 1;
            bw_len = s;
            break;
        }
    }
    int max_rl = max_bw;
    if (offset < len - 1) {
        const uint8_t *fw = read + offset + 1;
        int fw_len = 1;
        int max_fw = 0;
        for (int s = 1; s <= kMaxStrUnit; ++s) {
            if (offset + s + 1 > len) break;
            max_fw = repetitions(read + offset + 1, s, read + offset + 1, len - offset - 1, true);
            if (max_fw > 1) {
                fw_len = s;
                break;
            }
        }
        if (fw_len == bw_len && same(fw, bw, fw_len)) {
            max_rl = max_bw + max_fw;
        } else {
            max_bw = repetitions(fw, fw_len, read, offset + 1, false);
            max_rl = max_fw + max_bw;
        }
    }
    return max_rl > kMaxRepeat ? kMaxRepeat : max_rl;
}

__global__ __launch_bounds__(256) void pcr_error_model_kernel(const uint8_t *__restrict__ bases,
                                                              const int64_t *__restrict__ read_off,
                                                              const int32_t *__restrict__ read_len, int64_t n_reads,
                                                              PcrQualTable table, uint8_t *ins, uint8_t *del) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t waves = static_cast<int64_t>(gridDim.x) * (blockDim.x / kWave);
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave; r < n_reads;
         r += waves) {
        const int64_t o = read_off[r];
        const int len = read_len[r];
        for (int i = 1 + lane; i < len; i += kWave) {  // position i - 1, as in the reference loop
            const uint8_t q = table.q[tandem_repeat_length(bases + o, len, i - 1)];
            if (ins[o + i - 1] > q) ins[o + i - 1] = q;
            if (del[o + i - 1] > q) del[o + i - 1] = q;
        }
    }
}

__device__ __forceinline__ uint8_t fixed_if_low(uint8_t v, int8_t min_q) {
    return static_cast<int8_t>(v) < min_q ? static_cast<uint8_t>(6) : v;  // QualityUtils.MIN_USABLE_Q_SCORE
}

__global__ __launch_bounds__(256) void cap_qualities_kernel(const int64_t *__restrict__ read_off,
                                                            const int32_t *__restrict__ read_len, int64_t n_reads,
                                                            const int32_t *__restrict__ mapq, int32_t bq_threshold,
                                                            int32_t disable_mapq_cap, uint8_t *quals, uint8_t *ins,
                                                            uint8_t *del) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t waves = static_cast<int64_t>(gridDim.x) * (blockDim.x / kWave);
    const int8_t threshold = static_cast<int8_t>(bq_threshold);
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave; r < n_reads;
         r += waves) {
        const int64_t o = read_off[r];
        const int len = read_len[r];
        const int mq = disable_mapq_cap ? 255 : mapq[r];
        for (int i = lane; i < len; i += kWave) {
            int q = quals[o + i];
            if (!disable_mapq_cap && mq < q) q = mq;
            quals[o + i] = fixed_if_low(static_cast<uint8_t>(q), threshold);
            ins[o + i] = fixed_if_low(ins[o + i], 6);
            del[o + i] = fixed_if_low(del[o + i], 6);
        }
    }
}

int read_grid(int64_t n_reads) {
    const int64_t blocks = (n_reads + 3) / 4;  // four waves (reads) per 256-thread block
    return static_cast<int>(blocks < 1 ? 1 : (blocks > 65536 ? 65536 : blocks));
}

}  // namespace

hipError_t launch_pcr_error_model(const uint8_t *bases, const int64_t *read_off, const int32_t *read_len,
                                  int64_t n_reads, const PcrQualTable &table, uint8_t *ins, uint8_t *del,
                                  hipStream_t stream) {
    if (n_reads <= 0) return hipSuccess;
    hipLaunchKernelGGL(pcr_error_model_kernel, dim3(read_grid(n_reads)), dim3(256), 0, stream, bases, read_off,
                       read_len, n_reads, table, ins, del);
    return hipGetLastError();
}

hipError_t launch_cap_qualities(const int64_t *read_off, const int32_t *read_len, int64_t n_reads, const int32_t *mapq,
                                int32_t bq_threshold, int32_t disable_mapq_cap, uint8_t *quals, uint8_t *ins,
                                uint8_t *del, hipStream_t stream) {
    if (n_reads <= 0) return hipSuccess;
    hipLaunchKernelGGL(cap_qualities_kernel, dim3(read_grid(n_reads)), dim3(256), 0, stream, read_off, read_len,
                       n_reads, mapq, bq_threshold, disable_mapq_cap, quals, ins, del);
    return hipGetLastError();
}

}  // namespace gatk_pairhmm

// ---------------------------------------------------------------------------------------------
// DRAGstr imputation on the device (DragstrPairHMMInputScoreImputator.impute,
// DragstrPairHMMInputScoreImputator.java:53-70, over the contexts of DragstrReadSTRAnalyzer,
// DragstrReadSTRAnalyzer.java:69-139).  The copy count of period p at position i is the reference
// test's brute-force definition (DragstrReadSTRAnalyzerTest.java:126-171): the longest run of equal
// adjacent p-mers through any p-mer starting in [i-p+1, i+1], one copy discounted for the p-mer at 0
// when i < p-1; the best period has the most copies, the shorter on ties.  Each lane walks those runs
// directly over the read staged in LDS.
// ---------------------------------------------------------------------------------------------
namespace gatk_pairhmm {
namespace {

constexpr int kDragstrMaxRead = 4096;  // bases of one read staged per wave

__device__ __forceinline__ bool pmer_eq(const uint8_t *s, int a, int b, int p) {
    for (int k = 0; k < p; ++k)
        if (s[a + k] != s[b + k]) return false;
    return true;
}

__device__ int copies_at(const uint8_t *s, int L, int i, int p) {
    if (p > L) return 0;
    const int lo = i - p + 1 > 0 ? i - p + 1 : 0;
    const int hi = i + 1 < L - p ? i + 1 : L - p;
    int best = 0;
    for (int u = lo; u <= hi; ++u) {
        int f = 1;  // equal p-mers from u forward
        while (u + f * p <= L - p && pmer_eq(s, u + (f - 1) * p, u + f * p, p)) ++f;
        int b = 0;  // and backward
        while (u - (b + 1) * p >= 0 && pmer_eq(s, u - (b + 1) * p, u - b * p, p)) ++b;
        int c = f + b;
        if (u == 0 && i < p - 1) c = c - 1 > 1 ? c - 1 : 1;
        best = c > best ? c : best;
    }
    return best;
}

__global__ __launch_bounds__(256) void dragstr_impute_kernel(const uint8_t *__restrict__ bases,
                                                             const int64_t *__restrict__ read_off,
                                                             const int32_t *__restrict__ read_len, int64_t n_reads,
                                                             DragstrByteTables tables, uint8_t *out_gop,
                                                             uint8_t *out_gcp) {
    __shared__ uint8_t stage[4][kDragstrMaxRead];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    uint8_t *s = stage[wave];
    const int64_t waves = static_cast<int64_t>(gridDim.x) * (blockDim.x / kWave);
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + wave; r < n_reads; r += waves) {
        const int64_t o = read_off[r];
        const int L = read_len[r];
        if (L > kDragstrMaxRead) continue;  // beyond the stage (the host rejects such batches up front)
        __builtin_amdgcn_wave_barrier();
        for (int i = lane; i < L; i += kWave) s[i] = bases[o + i];
        __builtin_amdgcn_wave_barrier();
        for (int i = lane; i < L; i += kWave) {
            if (i == L - 1) {  // DragstrPairHMMInputScoreImputator.java:22-23,69-70
                out_gop[o + i] = 45;
                out_gcp[o + i] = 10;
                continue;
            }
            int best_p = 1, best_c = copies_at(s, L, i, 1);
            for (int p = 2; p <= tables.max_period; ++p) {
                const int c = copies_at(s, L, i, p);
                if (c > best_c) {
                    best_c = c;
                    best_p = p;
                }
            }
            const int rep = best_c < 1 ? 1 : (best_c > tables.max_repeats ? tables.max_repeats : best_c);
            const int cell = (best_p - 1) * tables.max_repeats + rep - 1;
            out_gop[o + i] = tables.gop[cell];
            out_gcp[o + i] = tables.gcp[cell];
        }
    }
}

}  // namespace

hipError_t launch_dragstr_impute(const uint8_t *bases, const int64_t *read_off, const int32_t *read_len,
                                 int64_t n_reads, const DragstrByteTables &tables, uint8_t *out_gop, uint8_t *out_gcp,
                                 hipStream_t stream) {
    if (n_reads <= 0) return hipSuccess;
    hipLaunchKernelGGL(dragstr_impute_kernel, dim3(read_grid(n_reads)), dim3(256), 0, stream, bases, read_off,
                       read_len, n_reads, tables, out_gop, out_gcp);
    return hipGetLastError();
}

int dragstr_max_read_length() { return kDragstrMaxRead; }

}  // namespace gatk_pairhmm
