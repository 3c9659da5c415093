// dragstr.cpp -- DRAGstr read STR context and per-position gap penalty imputation (host code).
//
// For every read position the DRAGstr model picks the repeat unit (period 1..P) whose tandem run
// overlapping the position, or starting right after it, has the most copies; ties go to the shorter
// period (DragstrReadSTRAnalyzer.java:11-47, :152-178).  The imputer then takes GOP (insertion and
// deletion alike, capped at 40) and GCP from the (period, repeats) tables, rounded to bytes, and
// pins the last read position to 45 / 10 (DragstrPairHMMInputScoreImputator.java:22-31, :53-70;
// table lookup clamps period and repeats to the table size, DragstrParams.java:194-200).
//
// The repeat count of period p at position i is defined by the reference's brute force
// (DragstrReadSTRAnalyzerTest.java:126-171): the largest run of equal adjacent p-mers over the
// p-mers that start in [i-p+1, i+1], counted forward and backward from that p-mer, with one copy
// discounted for p-mers starting at 0 when i < p-1.  This file computes it in O(L*p) per period:
// one backward pass finds where adjacent p-mers are equal (a run of p matching bases at distance
// p), chains give the copy counts up- and downstream of every start, and a window of p+1 starts
// gives the per-position maximum.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "capi_internal.hpp"

using namespace gatk_pairhmm::capi;

namespace {

// repeats[i] for one period; work vectors are reused across periods.
void repeats_for_period(const uint8_t *s, int L, int p, std::vector<int> &run, std::vector<int> &fwd,
                        std::vector<int> &bwd, int32_t *out) {
    if (p > L) {
        std::fill(out, out + L, 0);
        return;
    }
    const int n_units = L - p + 1;  // p-mer starts 0..L-p
    // run[t]: number of consecutive t' >= t with s[t'] == s[t'+p], for t in [0, L-p).
    run.assign(static_cast<size_t>(L) + 1, 0);
    for (int t = L - p - 1; t >= 0; --t) run[t] = (s[t] == s[t + p]) ? run[t + 1] + 1 : 0;
    auto eq_next = [&](int u) { return u + p <= L - p && run[u] >= p; };  // p-mer u == p-mer u+p
    fwd.assign(n_units, 1);
    for (int u = n_units - 1; u >= 0; --u)
        if (eq_next(u)) fwd[u] = fwd[u + p] + 1;
    bwd.assign(n_units, 0);
    for (int u = p; u < n_units; ++u)
        if (eq_next(u - p)) bwd[u] = bwd[u - p] + 1;
    for (int i = 0; i < L; ++i) {
        const int lo = std::max(0, i - p + 1), hi = std::min(i + 1, L - p);
        int best = 0;
        for (int u = lo; u <= hi; ++u) {
            int c = fwd[u] + bwd[u];
            if (u == 0 && i < p - 1) c = std::max(1, c - 1);
            best = std::max(best, c);
        }
        out[i] = best;
    }
}

void analyze(const uint8_t *s, int L, int max_period, int32_t *period, int32_t *repeats, std::vector<int32_t> &tmp,
             std::vector<int> &run, std::vector<int> &fwd, std::vector<int> &bwd) {
    if (L <= 0) return;
    repeats_for_period(s, L, 1, run, fwd, bwd, repeats);
    std::fill(period, period + L, 1);
    tmp.resize(L);
    for (int p = 2; p <= max_period; ++p) {
        repeats_for_period(s, L, p, run, fwd, bwd, tmp.data());
        for (int i = 0; i < L; ++i)
            if (tmp[i] > repeats[i]) {
                repeats[i] = tmp[i];
                period[i] = p;
            }
    }
}

bool valid_reads(const uint8_t *bases, const int64_t *off, const int32_t *len, int64_t n_reads, int64_t bytes) {
    for (int64_t r = 0; r < n_reads; ++r)
        if (len[r] < 0 || off[r] < 0 || off[r] + len[r] > bytes) return false;
    return bases != nullptr || bytes == 0;
}

}  // namespace

extern "C" int32_t pairhmm_dragstr_analyze(const uint8_t *bases, int32_t length, int32_t max_period,
                                           int32_t *out_period, int32_t *out_repeats) {
    if (length < 0 || max_period < 1) return fail(PAIRHMM_ERR_BAD_ARG, "length must be >= 0 and max_period >= 1");
    if (length > 0 && (!bases || !out_period || !out_repeats)) return fail(PAIRHMM_ERR_BAD_ARG, "null argument");
    std::vector<int32_t> tmp;
    std::vector<int> run, fwd, bwd;
    analyze(bases, length, max_period, out_period, out_repeats, tmp, run, fwd, bwd);
    return PAIRHMM_OK;
}

extern "C" int32_t pairhmm_dragstr_contexts(const uint8_t *read_bases, const int64_t *read_off, const int32_t *read_len,
                                            int64_t n_reads, int64_t read_bytes, int32_t max_period,
                                            int32_t max_repeats, int32_t *out_context) {
    if (n_reads < 0 || read_bytes < 0 || max_period < 1 || max_repeats < 1)
        return fail(PAIRHMM_ERR_BAD_ARG, "bad sizes");
    if (n_reads > 0 && (!read_off || !read_len || !out_context)) return fail(PAIRHMM_ERR_BAD_ARG, "null argument");
    if (!valid_reads(read_bases, read_off, read_len, n_reads, read_bytes))
        return fail(PAIRHMM_ERR_BAD_ARG, "read offsets/lengths exceed read_bytes");
    std::vector<int32_t> per, rep, tmp;
    std::vector<int> run, fwd, bwd;
    for (int64_t r = 0; r < n_reads; ++r) {
        const int L = read_len[r];
        if (L == 0) continue;
        per.resize(L);
        rep.resize(L);
        analyze(read_bases + read_off[r], L, max_period, per.data(), rep.data(), tmp, run, fwd, bwd);
        int32_t *o = out_context + read_off[r];
        for (int i = 0; i < L - 1; ++i) {
            const int pi = std::min(per[i], max_period) - 1;
            const int ri = std::min(std::max(rep[i], 1), max_repeats) - 1;
            o[i] = pi * max_repeats + ri;
        }
        o[L - 1] = -1;  // fixed end-of-read penalties
    }
    return PAIRHMM_OK;
}

extern "C" int32_t pairhmm_dragstr_impute(const double *gop_table, const double *gcp_table, int32_t max_period,
                                          int32_t max_repeats, const uint8_t *read_bases, const int64_t *read_off,
                                          const int32_t *read_len, int64_t n_reads, int64_t read_bytes,
                                          uint8_t *out_gop, uint8_t *out_gcp) {
    if (!gop_table || !gcp_table) return fail(PAIRHMM_ERR_BAD_ARG, "null table");
    if (n_reads > 0 && (!out_gop || !out_gcp)) return fail(PAIRHMM_ERR_BAD_ARG, "null argument");
    const int64_t n_ctx = static_cast<int64_t>(max_period) * max_repeats;
    for (int64_t k = 0; k < n_ctx && max_period >= 1 && max_repeats >= 1; ++k)
        if (!(gop_table[k] >= 0.0 && std::isfinite(gop_table[k]) && gcp_table[k] >= 0.0 && std::isfinite(gcp_table[k])))
            return fail(PAIRHMM_ERR_BAD_ARG, "tables must hold finite values >= 0");
    std::vector<int32_t> ctx(static_cast<size_t>(read_bytes));
    const int32_t st = pairhmm_dragstr_contexts(read_bases, read_off, read_len, n_reads, read_bytes, max_period,
                                                max_repeats, ctx.data());
    if (st != PAIRHMM_OK) return st;
    constexpr double kMaxGopInRead = 40.0;
    constexpr uint8_t kGopAtEnd = 45, kGcpAtEnd = 10;
    for (int64_t r = 0; r < n_reads; ++r) {
        for (int64_t i = read_off[r], e = read_off[r] + read_len[r]; i < e; ++i) {
            const int32_t c = ctx[i];
            if (c < 0) {
                out_gop[i] = kGopAtEnd;
                out_gcp[i] = kGcpAtEnd;
            } else {
                // Java's Math.round: floor(x + 0.5); values here are >= 0 and small.
                out_gop[i] = static_cast<uint8_t>(static_cast<int>(std::floor(std::min(kMaxGopInRead, gop_table[c]) + 0.5)));
                out_gcp[i] = static_cast<uint8_t>(static_cast<int>(std::floor(gcp_table[c] + 0.5)));
            }
        }
    }
    return PAIRHMM_OK;
}
