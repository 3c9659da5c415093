// model_tables.hpp -- host-side PairHMM model tables uploaded to the device at initialize().
//
// The per-row transition probabilities and emission priors are functions of byte-valued Phred
// scores.  The tables reproduce the FP64 values of GATK's Java caches so that the FP64 kernel
// sees exactly the doubles LoglessPairHMM sees:
//   qual_err[q]  = 10^(q / -10)                         QualityUtils.java:52-57, 138-141
//   mm_prob[t]   = matchToMatchProb (triangular)        PairHMMModel.java:71, 86-94, 373-388
// where matchToMatchProb(i, j) = 10^(log1p(-min(1, 10^s)) / ln 10) with
// s = approximateLog10SumLog10(-0.1 max, -0.1 min), the Jacobian-table log-sum of
// MathUtils.java:406-430, 467-480 (step 1e-4, tolerance 8.0).
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

namespace gatk_pairhmm {

constexpr int kMaxQual = 254;                                       // QualityUtils.java:43
constexpr int kMMTableSize = ((kMaxQual + 1) * (kMaxQual + 2)) >> 1;  // PairHMMModel.java:71

struct ModelTables {
    // [q] = 10^(q/-10) (entry 255 is unused: such bytes are rejected); [256 + q] = that value / 3.0,
    // the mismatch prior with the tristate correction (LoglessPairHMM.java:89-90), divided here in
    // double exactly as Java does, so the kernels need no division.
    double qual_err[512];
    std::vector<double> mm_prob;          // kMMTableSize entries, index max*(max+1)/2 + min

    ModelTables() : mm_prob(kMMTableSize) {
        for (int q = 0; q < 256; ++q) qual_err[q] = std::pow(10.0, static_cast<double>(q) / -10.0);
        for (int q = 0; q < 256; ++q) qual_err[256 + q] = qual_err[q] / 3.0;
        const double tol = 8.0, step = 0.0001;
        const int jac_len = static_cast<int>(tol / step) + 1;
        std::vector<double> jac(jac_len);
        for (int k = 0; k < jac_len; ++k) jac[k] = std::log10(1.0 + std::pow(10.0, static_cast<double>(-k) * step));
        const double inv_step = 1.0 / step;
        auto fast_round = [](double d) { return d > 0.0 ? static_cast<int>(d + 0.5) : static_cast<int>(d - 0.5); };
        auto approx_sum = [&](double a, double b) {  // a = -0.1*max <= b = -0.1*min
            if (a > b) std::swap(a, b);
            const double diff = b - a;
            return b + (diff < tol ? jac[fast_round(diff * inv_step)] : 0.0);
        };
        const double inv_ln10 = 1.0 / std::log(10.0);
        for (int i = 0, off = 0; i <= kMaxQual; off += ++i) {
            for (int j = 0; j <= i; ++j) {
                const double p = std::pow(10.0, approx_sum(-0.1 * i, -0.1 * j));
                const double lg = std::log1p(-(p < 1.0 ? p : 1.0)) * inv_ln10;
                mm_prob[off + j] = std::pow(10.0, lg);
            }
        }
    }
};

// Host reference of one row's transition probabilities (PairHMMModel.qualToTransProbs,
// PairHMMModel.java:107-117) used by validation code; the kernels do the same lookups on device.
inline int mm_index(int ins, int del) {
    const int mn = ins < del ? ins : del, mx = ins < del ? del : ins;
    return ((mx * (mx + 1)) >> 1) + mn;
}

}  // namespace gatk_pairhmm
