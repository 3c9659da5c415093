// sw_capi.cpp -- C-ABI of the gfx950 Smith-Waterman aligner: the drop-in for GKL's
// SWAlignerNativeBinding (IntelSmithWaterman) as called through
// src/main/java/org/broadinstitute/hellbender/utils/smithwaterman/SWNativeAlignerWrapper.java:33-53
// and SmithWatermanIntelAligner.java:25-38, plus a batch form for many alignments per launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <string>
#include <vector>

#include "capi_internal.hpp"

using namespace gatk_pairhmm;
using namespace gatk_pairhmm::capi;

namespace {

constexpr int kMaxSwLength = 32767;       // backtrack codes are int16
constexpr int32_t kMaxSwPenalty = 1000000;  // keeps every score far from int32 overflow
constexpr size_t kSwScratchBudget = size_t(2) << 30;

// TextCigarCodec.encode of the elements in reverse list order (calculateCigar's Lists.reverse):
// decimal length, then the operator.  Returns the text length (a terminating 0 follows it), -1 if
// the text and its 0 do not fit in cap.
int32_t format_cigar(const uint32_t *e, int n, char *out, int32_t cap) {
    static const char kOp[4] = {'M', 'I', 'D', 'S'};
    int32_t pos = 0;
    for (int k = n - 1; k >= 0; --k) {
        char digits[12];
        int nd = 0;
        uint32_t len = e[k] >> 2;
        do {
            digits[nd++] = static_cast<char>('0' + len % 10u);
            len /= 10u;
        } while (len != 0);
        if (pos + nd + 2 > cap) return -1;
        while (nd > 0) out[pos++] = digits[--nd];
        out[pos++] = kOp[e[k] & 3u];
    }
    if (pos + 1 > cap) return -1;
    out[pos] = 0;
    return pos;
}

// Results of sw_run_batch: the compact element copy, element counts, offsets (in the context's
// pinned staging) and the full element lists of the pairs with more than kSwShortElems elements.
struct SwResults {
    const uint32_t *short_elems = nullptr;
    const int32_t *n_el = nullptr, *offs = nullptr;
    std::vector<int64_t> long_pair;     // ascending pair indices
    std::vector<int64_t> long_start;    // their first element in long_elems
    std::vector<uint32_t> long_elems;
    // Pair p's elements; `cursor` walks long_pair for pairs visited in ascending order.
    const uint32_t *elements(int64_t p, size_t &cursor) const {
        if (n_el[p] <= kSwShortElems) return short_elems + p * kSwShortElems;
        while (long_pair[cursor] != p) ++cursor;
        return long_elems.data() + long_start[cursor];
    }
};

// The device side of both batch forms: validation, staging, one launch, results back in the
// context's pinned staging: the first kSwShortElems elements of every pair, the counts and the
// offsets (a few KB per thousand pairs, where the full lists are sized for the worst case), then the
// full lists of the rare pairs with more elements.
int32_t sw_run_batch(pairhmm_context *ctx, int64_t n_pairs, const uint8_t *ref_bases, const int64_t *ref_off,
                     const int32_t *ref_len, const uint8_t *alt_bases, const int64_t *alt_off, const int32_t *alt_len,
                     const pairhmm_sw_parameters *params, int32_t overhang_strategy, SwResults &res) {
    if (!ctx) return fail(PAIRHMM_ERR_STATE, "context is null");
    if (n_pairs < 0 || n_pairs > 0xFFFFFFFFll) return fail(PAIRHMM_ERR_BAD_ARG, "bad pair count");
    if (!params) return fail(PAIRHMM_ERR_BAD_ARG, "null SW parameters");
    if (overhang_strategy < 0 || overhang_strategy > 3) return fail(PAIRHMM_ERR_BAD_ARG, "unknown overhang strategy");
    for (int32_t v : {params->match_value, params->mismatch_penalty, params->gap_open_penalty,
                      params->gap_extend_penalty})
        if (v > kMaxSwPenalty || v < -kMaxSwPenalty)
            return fail(PAIRHMM_ERR_BAD_ARG, "SW parameters beyond +-1e6");
    if (n_pairs == 0) return PAIRHMM_OK;
    if (!ref_bases || !ref_off || !ref_len || !alt_bases || !alt_off || !alt_len)
        return fail(PAIRHMM_ERR_BAD_ARG, "null array argument");
    int max_r = 1, max_a = 1;
    int64_t ref_end = 0, alt_end = 0, elem_total = 0;
    for (int64_t p = 0; p < n_pairs; ++p) {
        // SmithWatermanJavaAligner.java:55-57 (IllegalArgumentException for empty sequences).
        if (ref_len[p] <= 0 || alt_len[p] <= 0)
            return fail(PAIRHMM_ERR_BAD_ARG,
                        "Non-null, non-empty sequences are required for the Smith-Waterman calculation (pair " +
                            std::to_string(p) + ")");
        if (ref_len[p] > kMaxSwLength || alt_len[p] > kMaxSwLength)
            return fail(PAIRHMM_ERR_BAD_ARG, "sequences longer than 32767 bases");
        if (ref_off[p] < 0 || alt_off[p] < 0) return fail(PAIRHMM_ERR_BAD_ARG, "negative offset");
        max_r = std::max(max_r, ref_len[p]);
        max_a = std::max(max_a, alt_len[p]);
        ref_end = std::max(ref_end, ref_off[p] + ref_len[p]);
        alt_end = std::max(alt_end, alt_off[p] + alt_len[p]);
        elem_total += ref_len[p] + alt_len[p] + 4;
    }
    if (sw_lds_bytes(max_r) > max_lds_per_block())
        return fail(PAIRHMM_ERR_BAD_ARG, "reference of " + std::to_string(max_r) + " bases exceeds the LDS budget");

    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    const size_t o_rb = take(ref_end), o_ab = take(alt_end), o_roff = take(8 * n_pairs), o_rlen = take(4 * n_pairs),
                 o_aoff = take(8 * n_pairs), o_alen = take(4 * n_pairs), o_eoff = take(8 * n_pairs);
    const size_t in_bytes = off;
    const size_t o_short = take(4 * static_cast<size_t>(kSwShortElems) * n_pairs), o_n = take(4 * n_pairs),
                 o_off = take(4 * n_pairs);
    const size_t out_begin = o_short, out_bytes = off - o_short;
    const size_t o_elems = take(4 * static_cast<size_t>(elem_total));  // device only
    const size_t per_wave = sw_scratch_bytes_per_wave(max_r, max_a);
    const int cus = std::max(ctx->num_cus, 1);
    // 20 waves per CU: the 12-byte LDS row records of 600-base references allow it, and the sweep's
    // lane-to-lane dependency chain needs many resident waves to hide its latency.
    int64_t waves = std::min<int64_t>(n_pairs, static_cast<int64_t>(cus) * 20);
    waves = std::max<int64_t>(1, std::min<int64_t>(waves, static_cast<int64_t>(kSwScratchBudget / per_wave)));
    const size_t o_scratch = take(per_wave * static_cast<size_t>(waves));

    HIP_TRY(hipSetDevice(ctx->device), PAIRHMM_ERR_KERNEL);
    if (ctx->stage_host.reserve(std::max(in_bytes, out_bytes)) != hipSuccess)
        return fail(PAIRHMM_ERR_OOM, "pinned staging allocation failed");
    if (ctx->stage_dev.reserve(off) != hipSuccess) return fail(PAIRHMM_ERR_OOM, "device staging allocation failed");
    uint8_t *hs = static_cast<uint8_t *>(ctx->stage_host.ptr);
    std::memcpy(hs + o_rb, ref_bases, static_cast<size_t>(ref_end));
    std::memcpy(hs + o_ab, alt_bases, static_cast<size_t>(alt_end));
    std::memcpy(hs + o_roff, ref_off, 8 * n_pairs);
    std::memcpy(hs + o_rlen, ref_len, 4 * n_pairs);
    std::memcpy(hs + o_aoff, alt_off, 8 * n_pairs);
    std::memcpy(hs + o_alen, alt_len, 4 * n_pairs);
    int64_t *eoff = reinterpret_cast<int64_t *>(hs + o_eoff);
    for (int64_t p = 0, e = 0; p < n_pairs; ++p) {
        eoff[p] = e;
        e += ref_len[p] + alt_len[p] + 4;
    }
    uint8_t *ds = static_cast<uint8_t *>(ctx->stage_dev.ptr);
    HIP_TRY(hipMemcpyAsync(ds, hs, in_bytes, hipMemcpyHostToDevice, ctx->stream), PAIRHMM_ERR_KERNEL);

    SwBatch b{};
    b.ref_bases = ds + o_rb;
    b.ref_off = reinterpret_cast<const int64_t *>(ds + o_roff);
    b.ref_len = reinterpret_cast<const int32_t *>(ds + o_rlen);
    b.alt_bases = ds + o_ab;
    b.alt_off = reinterpret_cast<const int64_t *>(ds + o_aoff);
    b.alt_len = reinterpret_cast<const int32_t *>(ds + o_alen);
    b.match = params->match_value;
    b.mismatch = params->mismatch_penalty;
    b.open = params->gap_open_penalty;
    b.extend = params->gap_extend_penalty;
    b.strategy = overhang_strategy;
    b.max_ref_len = max_r;
    b.scratch = ds + o_scratch;
    b.scratch_per_wave = per_wave;
    b.bt_bytes_per_wave = sw_bt_bytes_per_wave(max_r, max_a);
    b.elems = reinterpret_cast<uint32_t *>(ds + o_elems);
    b.elem_off = reinterpret_cast<const int64_t *>(ds + o_eoff);
    b.elems_short = reinterpret_cast<uint32_t *>(ds + o_short);
    b.n_elems = reinterpret_cast<int32_t *>(ds + o_n);
    b.offset = reinterpret_cast<int32_t *>(ds + o_off);
    HIP_TRY(launch_sw(b, static_cast<uint32_t>(n_pairs), static_cast<int>(waves), ctx->stream), PAIRHMM_ERR_KERNEL);
    // The input staging (in_bytes) has been consumed by the copy above before the kernel ran (same
    // stream), so the outputs can come back into the same pinned buffer.
    HIP_TRY(hipMemcpyAsync(hs, ds + out_begin, out_bytes, hipMemcpyDeviceToHost, ctx->stream), PAIRHMM_ERR_KERNEL);
    HIP_TRY(hipStreamSynchronize(ctx->stream), PAIRHMM_ERR_KERNEL);
    res.short_elems = reinterpret_cast<const uint32_t *>(hs);
    res.n_el = reinterpret_cast<const int32_t *>(hs + (o_n - out_begin));
    res.offs = reinterpret_cast<const int32_t *>(hs + (o_off - out_begin));
    res.long_pair.clear();
    res.long_start.clear();
    res.long_elems.clear();
    int64_t long_total = 0;
    std::vector<int64_t> long_src;  // first element of each long pair in the device's full lists
    for (int64_t p = 0, e = 0; p < n_pairs; ++p) {
        if (res.n_el[p] > kSwShortElems) {
            res.long_pair.push_back(p);
            res.long_start.push_back(long_total);
            long_src.push_back(e);
            long_total += res.n_el[p];
        }
        e += ref_len[p] + alt_len[p] + 4;
    }
    if (!res.long_pair.empty()) {
        res.long_elems.resize(static_cast<size_t>(long_total));
        for (size_t k = 0; k < res.long_pair.size(); ++k)
            HIP_TRY(hipMemcpyAsync(res.long_elems.data() + res.long_start[k], b.elems + long_src[k],
                                   4 * static_cast<size_t>(res.n_el[res.long_pair[k]]), hipMemcpyDeviceToHost,
                                   ctx->stream),
                    PAIRHMM_ERR_KERNEL);
        HIP_TRY(hipStreamSynchronize(ctx->stream), PAIRHMM_ERR_KERNEL);
    }
    return PAIRHMM_OK;
}

}  // namespace

extern "C" int32_t pairhmm_sw_align_batch(pairhmm_context *ctx, int64_t n_pairs, const uint8_t *ref_bases,
                                          const int64_t *ref_off, const int32_t *ref_len, const uint8_t *alt_bases,
                                          const int64_t *alt_off, const int32_t *alt_len,
                                          const pairhmm_sw_parameters *params, int32_t overhang_strategy,
                                          char *cigars, const int64_t *cigar_off, const int32_t *cigar_cap,
                                          int32_t *alignment_offset) {
    if (n_pairs > 0 && (!cigars || !cigar_off || !cigar_cap || !alignment_offset))
        return fail(PAIRHMM_ERR_BAD_ARG, "null array argument");
    SwResults res;
    const int32_t st = sw_run_batch(ctx, n_pairs, ref_bases, ref_off, ref_len, alt_bases, alt_off, alt_len, params,
                                    overhang_strategy, res);
    if (st != PAIRHMM_OK || n_pairs == 0) return st;
    size_t cursor = 0;
    for (int64_t p = 0; p < n_pairs; ++p) {
        if (format_cigar(res.elements(p, cursor), res.n_el[p], cigars + cigar_off[p], cigar_cap[p]) < 0)
            return fail(PAIRHMM_ERR_BAD_ARG, "CIGAR buffer of pair " + std::to_string(p) + " is too small");
        alignment_offset[p] = res.offs[p];
    }
    return PAIRHMM_OK;
}

extern "C" int32_t pairhmm_sw_align_batch_packed(pairhmm_context *ctx, int64_t n_pairs, const uint8_t *ref_bases,
                                                 const int64_t *ref_off, const int32_t *ref_len,
                                                 const uint8_t *alt_bases, const int64_t *alt_off,
                                                 const int32_t *alt_len, const pairhmm_sw_parameters *params,
                                                 int32_t overhang_strategy, char *cigar_text, int64_t cigar_text_cap,
                                                 int64_t *cigar_end, int32_t *alignment_offset) {
    if (n_pairs > 0 && (!cigar_text || !cigar_end || !alignment_offset))
        return fail(PAIRHMM_ERR_BAD_ARG, "null array argument");
    if (cigar_text_cap < 0) return fail(PAIRHMM_ERR_BAD_ARG, "negative CIGAR text capacity");
    SwResults res;
    const int32_t st = sw_run_batch(ctx, n_pairs, ref_bases, ref_off, ref_len, alt_bases, alt_off, alt_len, params,
                                    overhang_strategy, res);
    if (st != PAIRHMM_OK || n_pairs == 0) return st;
    int64_t pos = 0;
    size_t cursor = 0;
    for (int64_t p = 0; p < n_pairs; ++p) {
        // format_cigar also writes a terminating 0, which the next CIGAR overwrites; the capacity
        // check keeps that byte inside the buffer too.
        const int64_t room = cigar_text_cap - pos;
        const int32_t w = format_cigar(res.elements(p, cursor), res.n_el[p], cigar_text + pos,
                                       static_cast<int32_t>(std::min<int64_t>(room, INT32_MAX)));
        if (w < 0)
            return fail(PAIRHMM_ERR_BAD_ARG, "CIGAR text buffer is too small at pair " + std::to_string(p));
        pos += w;
        cigar_end[p] = pos;
        alignment_offset[p] = res.offs[p];
    }
    return PAIRHMM_OK;
}

extern "C" int32_t pairhmm_sw_align(pairhmm_context *ctx, const uint8_t *ref, int32_t ref_len, const uint8_t *alt,
                                    int32_t alt_len, const pairhmm_sw_parameters *params, int32_t overhang_strategy,
                                    char *cigar, int32_t cigar_cap, int32_t *alignment_offset) {
    const int64_t zero = 0;
    return pairhmm_sw_align_batch(ctx, 1, ref, &zero, &ref_len, alt, &zero, &alt_len, params, overhang_strategy, cigar,
                                  &zero, &cigar_cap, alignment_offset);
}
