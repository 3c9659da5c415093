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
// One thread per read over its row of the read-major matrix; FP64 with contraction off, so every
// sum and product rounds where Java's do (the quality sums run in read order).
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

__global__ void postprocess_kernel(PostprocessArgs p) {
#pragma clang fp contract(off)
    // The table in LDS: per-base lookups differ lane to lane, which LDS serves far better than
    // gathers from global memory.
    __shared__ double2 tab[40];
    if (threadIdx.x < 40) tab[threadIdx.x] = make_double2(kDynMean[threadIdx.x], kDynVar[threadIdx.x]);
    __syncthreads();
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= p.n_reads) return;
    const int A = p.n_alleles;
    double *v = p.likelihoods + r * A;
    if (p.normalize && A >= 2) {
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
    const int L = p.read_len[r];
    const double errors = ceil(static_cast<double>(L) * p.max_error_per_base);
    double t = (p.dynamic ? errors : fmin(2.0, errors)) * -4.0;
    if (p.dynamic) {
        // Sums in read order, as the reference adds them; the qualities are fetched 8 at a time
        // from aligned words wholly inside the read, byte by byte at the two ends.
        const uint8_t *q = p.base_quals + p.read_off[r];
        double s_mean = 0.0, s_var = 0.0;
        auto add = [&](int bq) {
            const int e = bq <= 1 ? 0 : (bq < 40 ? bq : 40) - 1;
            const double2 mv = tab[e];
            s_mean += mv.x;
            s_var += mv.y;
        };
        const int head = static_cast<int>((8u - (reinterpret_cast<uintptr_t>(q) & 7u)) & 7u);
        const int h = head < L ? head : L;
        int i = 0;
        for (; i < h; ++i) add(q[i]);
        for (; i + 8 <= L; i += 8) {
            const uint64_t w = *reinterpret_cast<const uint64_t *>(q + i);
#pragma unroll
            for (int k = 0; k < 8; ++k) add(static_cast<int>((w >> (8 * k)) & 0xFFu));
        }
        for (; i < L; ++i) add(q[i]);
        const double d = (s_mean + p.scale * sqrt(s_var)) * -0.1;
        if (d < t) t = d;
    }
    double m = -__builtin_inf();
    for (int a = 0; a < A; ++a)
        if (v[a] > m) m = v[a];
    p.keep[r] = (m < t) ? 0 : 1;
    if (p.threshold) p.threshold[r] = t;
}

}  // namespace

hipError_t launch_postprocess(const PostprocessArgs &args, hipStream_t stream) {
    if (args.n_reads <= 0) return hipSuccess;
    const int block = 256;
    const int64_t grid = (args.n_reads + block - 1) / block;
    hipLaunchKernelGGL(postprocess_kernel, dim3(static_cast<unsigned>(grid)), dim3(block), 0, stream, args);
    return hipGetLastError();
}

}  // namespace gatk_pairhmm
