// postprocess.hip -- the likelihood post-processing that follows the PairHMM, on the device, so the
// read x haplotype matrix need not travel to the host for it
// (PairHMMLikelihoodCalculationEngine.computeReadLikelihoods,
// src/main/java/org/broadinstitute/hellbender/tools/walkers/haplotypecaller/PairHMMLikelihoodCalculationEngine.java:200-201):
//   1. the mismapping-rate cap, AlleleLikelihoods.normalizeLikelihoods
//      (src/main/java/org/broadinstitute/hellbender/utils/genotyper/AlleleLikelihoods.java:416-458):
//      per read, every likelihood below best + cap is raised to it, best being searchBestAllele's
//      (:505-533; the reference allele is skipped unless normalisation is symmetric);
//   2. the poorly-modelled read filter, filterPoorlyModeledEvidence
//      (ReadLikelihoodCalculationEngine.java:98-155, AlleleLikelihoods.java:1351-1375): a read is
//      dropped when its best likelihood is below ceil(L * rate) * -4 (at most 2 errors unless the
//      DRAGEN dynamic threshold is on, which takes the lower of that and
//      -0.1 (sum mean_q + K sqrt(sum var_q)) over the read's HMM base qualities).
// One thread per read; a wave's 64 matrix rows and quality strings are staged through LDS with
// coalesced loads.  FP64 with contraction off, so every sum and product rounds where Java's do (the
// quality sums run in read order).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pairhmm_kernels.hpp"

namespace gatk_pairhmm {
namespace {

// ReadLikelihoodCalculationEngine.dynamicReadQualThreshLookupTable (:161-176): mean and variance for
// base qualities 1..40, the reference's values.
__constant__ double kDynMean[40] = {
    5.996842844, 5.870018422, 5.401558531, 4.818940919, 4.218758304, 3.646319832, 3.122346753, 2.654731979,
    2.244479156, 1.88893867,  1.583645342, 1.3233807,   1.102785365, 0.916703025, 0.760361881, 0.629457387,
    0.520175654, 0.42918208,  0.353590663, 0.290923699, 0.23906788,  0.196230431, 0.160897421, 0.131795374,
    0.1078567,   0.088189063, 0.072048567, 0.058816518, 0.047979438, 0.039111985, 0.031862437, 0.025940415,
    0.021106532, 0.017163711, 0.013949904, 0.011332027, 0.009200898, 0.007467036, 0.006057179, 0.004911394};
__constant__ double kDynVar[40] = {
    0.196616587, 1.388545569, 5.641990128, 10.33176216, 14.25799688, 17.02880749, 18.64537883, 19.27521677,
    19.13584613, 18.43922003, 17.36842261, 16.07088712, 14.65952563, 13.21718577, 11.80207947, 10.45304833,
    9.194183767, 8.038657241, 6.991779595, 6.053379213, 5.219610436, 4.484302033, 3.839943445, 3.27839108,
    2.791361596, 2.370765375, 2.008921719, 1.698687797, 1.433525748, 1.207526336, 1.015402928, 0.852465956,
    0.714585285, 0.598145851, 0.500000349, 0.41742159,  0.348056286, 0.289881373, 0.241163527, 0.200422214};

constexpr int kPpWaves = 2;             // waves per workgroup, each over 64 consecutive reads
constexpr int kQualStage = 10240;       // bytes of HBM qualities a wave stages in LDS
constexpr int kRowStage = 512;          // likelihoods a wave stages in LDS (64 reads x 8 haplotypes)
constexpr int kQualBatch = 10;          // quality words per lane in flight while staging

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sums of the table's mean and variance over a read's qualities, in read order (the order the
// reference adds them).  Unrolled so the reads of later bases issue ahead of the ordered FP64 sums.
__device__ __forceinline__ void quality_sums(const uint8_t *src, int L, const double2 *tab, double &s_mean,
                                             double &s_var) {
#pragma clang fp contract(off)
#pragma unroll 8
    for (int i = 0; i < L; ++i) {
        const int bq = src[i];
        const int e = bq <= 1 ? 0 : (bq < 40 ? bq : 40) - 1;
        const double2 mv = tab[e];
        s_mean += mv.x;
        s_var += mv.y;
    }
}

__global__ __launch_bounds__(kPpWaves * 64) void postprocess_kernel(PostprocessArgs p) {
#pragma clang fp contract(off)
    // The table in LDS: per-base lookups differ lane to lane, which LDS serves far better than
    // gathers from global memory.
    __shared__ double2 tab[40];
    // Each wave's reads are adjacent in the matrix and (usually) in the quality plane: both are
    // brought into LDS with coalesced loads, worked on there, and the rows written back coalesced.
    __shared__ uint32_t qstage[kPpWaves][kQualStage / 4];
    __shared__ double rstage[kPpWaves][kRowStage];
    if (threadIdx.x < 40) tab[threadIdx.x] = make_double2(kDynMean[threadIdx.x], kDynVar[threadIdx.x]);
    __syncthreads();
    const int w = static_cast<int>(threadIdx.x) / 64, lane = static_cast<int>(threadIdx.x) % 64;
    const int64_t r0 = (static_cast<int64_t>(blockIdx.x) * kPpWaves + w) * 64;
    if (r0 >= p.n_reads) return;  // the whole wave
    const int nr = p.n_reads - r0 < 64 ? static_cast<int>(p.n_reads - r0) : 64;
    const bool has = lane < nr;
    const int64_t r = r0 + lane;
    const int A = p.n_alleles;

    const bool stage_rows = static_cast<int64_t>(nr) * A <= kRowStage;
    double *const grow = p.likelihoods + r0 * A;
    const int L = has ? p.read_len[r] : 0;
    const uint8_t *q = p.base_quals + (has && p.dynamic ? p.read_off[r] : 0);
    // Every HBM load of the wave is issued before the first wait: the row loads go to registers
    // first (nr * A <= kRowStage = 8 * 64, so eight per lane), then the qualities are staged, then
    // the rows are stored to LDS; one wave barrier covers both.
    double rt[kRowStage / 64];
    if (stage_rows) {
#pragma unroll
        for (int u = 0; u < kRowStage / 64; ++u) {
            const int e = lane + 64 * u;
            rt[u] = e < nr * A ? grow[e] : 0.0;
        }
    }
    // The wave's quality span [lo, hi) in global addresses; staged when it fits, read in place
    // otherwise.
    bool stage_q = false;
    uint64_t blo = 0;
    uint8_t *const qb = reinterpret_cast<uint8_t *>(qstage[w]);
    if (p.dynamic) {
        uint64_t lo = has ? reinterpret_cast<uintptr_t>(q) : ~0ull, hi = has ? lo + L : 0ull;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const uint64_t olo = __shfl_xor(static_cast<unsigned long long>(lo), off, 64);
            const uint64_t ohi = __shfl_xor(static_cast<unsigned long long>(hi), off, 64);
            lo = olo < lo ? olo : lo;
            hi = ohi > hi ? ohi : hi;
        }
        blo = lo & ~3ull;
        stage_q = hi > lo && hi - blo <= static_cast<uint64_t>(kQualStage);
        if (stage_q) {
            const uint64_t wlo = (lo + 3) & ~3ull, whi = hi & ~3ull;
            // Derived from the plane's pointer (not from the integer address) so the loads stay global.
            const uint8_t *gb = p.base_quals + (blo - reinterpret_cast<uintptr_t>(p.base_quals));
            if (wlo < whi) {
                const __attribute__((address_space(1))) uint32_t *gw =
                    (const __attribute__((address_space(1))) uint32_t *)(gb + (wlo - blo));
                const int nw = static_cast<int>((whi - wlo) / 4), w0 = static_cast<int>((wlo - blo) / 4);
                // Batches of kQualBatch loads per lane in flight before their stores.
                for (int k0 = lane; k0 < nw; k0 += 64 * kQualBatch) {
                    uint32_t t[kQualBatch];
#pragma unroll
                    for (int u = 0; u < kQualBatch; ++u) {
                        const int k = k0 + 64 * u;
                        t[u] = k < nw ? gw[k] : 0u;
                    }
#pragma unroll
                    for (int u = 0; u < kQualBatch; ++u) {
                        const int k = k0 + 64 * u;
                        if (k < nw) qstage[w][w0 + k] = t[u];
                    }
                }
                const int head = static_cast<int>(wlo - lo), tail = static_cast<int>(hi - whi);
                if (lane < head) qb[lo - blo + lane] = gb[lo - blo + lane];
                if (lane < tail) qb[whi - blo + lane] = gb[whi - blo + lane];
            } else {
                for (int k = lane; k < static_cast<int>(hi - lo); k += 64) qb[lo - blo + k] = gb[lo - blo + k];
            }
        }
    }
    if (stage_rows) {
#pragma unroll
        for (int u = 0; u < kRowStage / 64; ++u) {
            const int e = lane + 64 * u;
            if (e < nr * A) rstage[w][e] = rt[u];
        }
    }
    if (stage_rows || stage_q) wave_lds_sync();
    double *v = !has ? rstage[w] : stage_rows ? &rstage[w][lane * A] : p.likelihoods + r * A;

    if (has && p.normalize && A >= 2) {
        const bool can_ref = p.symmetric != 0;
        const int b0 = (can_ref || p.ref_index != 0) ? 0 : 1;
        double best = v[b0];
        for (int a = b0 + 1; a < A; ++a) {
            if (!can_ref && a == p.ref_index) continue;
            const double c = v[a];
            if (c > best) best = c;
        }
        const double worst = best + p.cap;
        for (int a = 0; a < A; ++a)
            if (v[a] < worst) v[a] = worst;
    }
    const double errors = ceil(static_cast<double>(L) * p.max_error_per_base);
    double t = (p.dynamic ? errors : fmin(2.0, errors)) * -4.0;
    if (p.dynamic) {
        // Sums in read order, as the reference adds them; two calls so each loop sees one address
        // space (LDS or global) and its reads pipeline.
        double s_mean = 0.0, s_var = 0.0;
        if (stage_q)
            quality_sums(qb + (reinterpret_cast<uintptr_t>(q) - blo), L, tab, s_mean, s_var);
        else
            quality_sums(q, L, tab, s_mean, s_var);
        const double d = (s_mean + p.scale * sqrt(s_var)) * -0.1;
        if (d < t) t = d;
    }
    if (has) {
        double m = -__builtin_inf();
        for (int a = 0; a < A; ++a)
            if (v[a] > m) m = v[a];
        p.keep[r] = (m < t) ? 0 : 1;
        if (p.threshold) p.threshold[r] = t;
    }
    if (stage_rows && p.normalize && A >= 2) {
        wave_lds_sync();
        for (int e = lane; e < nr * A; e += 64) grow[e] = rstage[w][e];
    }
}

}  // namespace

hipError_t launch_postprocess(const PostprocessArgs &args, hipStream_t stream) {
    if (args.n_reads <= 0) return hipSuccess;
    const int64_t per_block = 64 * kPpWaves;
    const int64_t grid = (args.n_reads + per_block - 1) / per_block;
    hipLaunchKernelGGL(postprocess_kernel, dim3(static_cast<unsigned>(grid)), dim3(per_block), 0, stream, args);
    return hipGetLastError();
}

}  // namespace gatk_pairhmm
