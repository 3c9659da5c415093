// pdhmm_capi.cpp -- C-ABI of the partially determined PairHMM (the drop-in for GKL's IntelPDHMM)
// and of the likelihood post-processing on the device.
//
// PD-HMM: the drop-in for GKL's IntelPDHMM
// as called by src/main/java/org/broadinstitute/hellbender/utils/pairhmm/VectorLoglessPairPDHMM.java:
// 49-61 (load / initialize) and :72-140 (computeLikelihoods over ReadDataHolder[] and
// HaplotypeDataHolder[] with haplotypeBases + haplotypePDBases, read-major double[] out).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "capi_internal.hpp"

using namespace gatk_pairhmm;
using namespace gatk_pairhmm::capi;

namespace {

// PartiallyDeterminedHaplotype byte layout (PartiallyDeterminedHaplotype.java:59-66).
constexpr uint8_t kPdSnp = 1, kPdDelStart = 2, kPdDelEndByte = 4, kPdAltMask = 0x78;

enum class PdState { Normal, InsideDel, AfterDel };  // LoglessPDPairHMM.java:22-31

// One pass of LoglessPDPairHMM's state machine over a row (:63-136): ORs `inside_bit` into the code
// of every column processed in INSIDE_DEL and returns the state the row ends in.
PdState pd_state_pass(const uint8_t *pd, int H, PdState state, uint32_t inside_bit, uint32_t *code) {
    for (int j = 0; j < H; ++j) {
        if (state == PdState::InsideDel) code[j] |= inside_bit;
        if (state == PdState::AfterDel) state = PdState::Normal;
        if (pd[j] & kPdDelStart) state = PdState::InsideDel;
        if (pd[j] & kPdDelEndByte) state = PdState::AfterDel;
    }
    return state;
}

// Column codes of one haplotype (layout in pairhmm_kernels.hpp).  Row 1 starts NORMAL; every later
// row starts in the state the row before ended in (the state variable lives outside the row loop,
// :63), which is the same state for all rows >= 2.
void pd_column_codes(const uint8_t *bases, const uint8_t *pd, int H, uint32_t *code) {
    for (int j = 0; j < H; ++j) {
        const uint32_t y = bases[j];
        uint32_t alt = (pd[j] & kPdSnp) ? (pd[j] & kPdAltMask) : 0u;
        if (y == 'N') alt |= 0xF8u;  // an 'N' haplotype base matches every read base (:163)
        code[j] = y | (alt << 8) | ((pd[j] & kPdDelEndByte) ? kPdDelEnd : 0u);
    }
    const PdState end1 = pd_state_pass(pd, H, PdState::Normal, kPdInsideRow1, code);
    pd_state_pass(pd, H, end1, kPdInsideLater, code);
    constexpr uint32_t inside = kPdInsideRow1 | kPdInsideLater;
    for (int j = 0; j < H; ++j)
        if ((code[j] & (inside | kPdDelEnd)) || (j + 1 < H && (code[j + 1] & inside))) code[j] |= kPdSpecial;
}

bool is_pd_base(uint8_t x) {
    switch (x) {
    case 'A': case 'a': case 'C': case 'c': case 'G': case 'g': case 'T': case 't': case 'N': return true;
    default: return false;
    }
}

}  // namespace

extern "C" int32_t pairhmm_pd_compute_likelihoods(pairhmm_context *ctx, const pairhmm_read_data_holder *reads,
                                                  int32_t n_reads, const pairhmm_pd_haplotype_data_holder *haps,
                                                  int32_t n_haps, double *likelihoods) {
    if (!ctx) return fail(PAIRHMM_ERR_STATE, "context is null (initialize was not called)");
    if (n_reads < 0 || n_haps < 0) return fail(PAIRHMM_ERR_BAD_ARG, "negative read or haplotype count");
    if (n_reads == 0 || n_haps == 0) return PAIRHMM_OK;
    if (!reads || !haps || !likelihoods) return fail(PAIRHMM_ERR_BAD_ARG, "null array argument");
    int64_t read_bytes = 0, hap_cols = 0;
    int max_h = 1;
    for (int32_t r = 0; r < n_reads; ++r) {
        const pairhmm_read_data_holder &rd = reads[r];
        int32_t st = validate_read(rd.read_bases, rd.read_quals, rd.insertion_gop, rd.deletion_gop, rd.overall_gcp,
                                   rd.length, r);
        if (st != PAIRHMM_OK) return st;
        read_bytes += rd.length;
    }
    // isBasePDMatching throws for a read base outside ACGTacgtN on a SNP column whose haplotype base
    // is neither that base nor 'N' (LoglessPDPairHMM.java:163, 195-196).  Haplotype side: which
    // bases sit on SNP columns.
    std::vector<std::array<bool, 256>> snp_bases(n_haps);
    for (int32_t h = 0; h < n_haps; ++h) {
        const pairhmm_pd_haplotype_data_holder &hp = haps[h];
        if (hp.length <= 0 || hp.haplotype_bases == nullptr || hp.haplotype_pd_bases == nullptr)
            return fail(PAIRHMM_ERR_BAD_ARG, "haplotype " + std::to_string(h) + " is empty or has a null array");
        hap_cols += hp.length;
        max_h = std::max(max_h, hp.length);
        snp_bases[h].fill(false);
        for (int j = 0; j < hp.length; ++j)
            if (hp.haplotype_pd_bases[j] & kPdSnp) snp_bases[h][hp.haplotype_bases[j]] = true;
    }
    for (int32_t r = 0; r < n_reads; ++r) {
        const pairhmm_read_data_holder &rd = reads[r];
        for (int i = 0; i < rd.length; ++i) {
            const uint8_t x = rd.read_bases[i];
            if (is_pd_base(x)) continue;
            for (int32_t h = 0; h < n_haps; ++h)
                for (int y = 0; y < 256; ++y)
                    if (snp_bases[h][y] && y != x && y != 'N')
                        return fail(PAIRHMM_ERR_BAD_ARG, "read " + std::to_string(r) + " base " +
                                                             std::to_string(static_cast<int>(x)) +
                                                             " meets a SNP column: Found unexpected base in alt alleles");
        }
    }
    const int64_t n_pairs = static_cast<int64_t>(n_reads) * n_haps;
    if (n_pairs > 0xFFFFFFFFll) return fail(PAIRHMM_ERR_BAD_ARG, "too many pairs in one call");
    if (pd_lds_bytes(max_h) > max_lds_per_block())
        return fail(PAIRHMM_ERR_BAD_ARG, "haplotype of " + std::to_string(max_h) + " bases exceeds the LDS budget");

    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    const size_t o_rb = take(read_bytes), o_rq = take(read_bytes), o_ri = take(read_bytes), o_rd = take(read_bytes),
                 o_rg = take(read_bytes), o_hc = take(sizeof(uint32_t) * hap_cols),
                 o_roff = take(sizeof(int64_t) * n_reads), o_rlen = take(sizeof(int32_t) * n_reads),
                 o_hoff = take(sizeof(int64_t) * n_haps), o_hlen = take(sizeof(int32_t) * n_haps),
                 o_order = take(sizeof(uint32_t) * n_pairs);
    const size_t in_bytes = off;
    const size_t o_out = take(sizeof(double) * n_pairs);
    const size_t total = off;

    HIP_TRY(hipSetDevice(ctx->device), PAIRHMM_ERR_KERNEL);
    if (ctx->stage_host.reserve(std::max(in_bytes, sizeof(double) * static_cast<size_t>(n_pairs))) != hipSuccess)
        return fail(PAIRHMM_ERR_OOM, "pinned staging allocation failed");
    if (ctx->stage_dev.reserve(total) != hipSuccess) return fail(PAIRHMM_ERR_OOM, "device staging allocation failed");
    uint8_t *hs = static_cast<uint8_t *>(ctx->stage_host.ptr);
    int64_t *roff = reinterpret_cast<int64_t *>(hs + o_roff);
    int32_t *rlen = reinterpret_cast<int32_t *>(hs + o_rlen);
    int64_t *hoff = reinterpret_cast<int64_t *>(hs + o_hoff);
    int32_t *hlen = reinterpret_cast<int32_t *>(hs + o_hlen);
    uint32_t *hcode = reinterpret_cast<uint32_t *>(hs + o_hc);
    int64_t pos = 0;
    for (int32_t r = 0; r < n_reads; ++r) {
        const pairhmm_read_data_holder &rd = reads[r];
        roff[r] = pos;
        rlen[r] = rd.length;
        if (rd.length > 0) {
            std::memcpy(hs + o_rb + pos, rd.read_bases, rd.length);
            std::memcpy(hs + o_rq + pos, rd.read_quals, rd.length);
            std::memcpy(hs + o_ri + pos, rd.insertion_gop, rd.length);
            std::memcpy(hs + o_rd + pos, rd.deletion_gop, rd.length);
            std::memcpy(hs + o_rg + pos, rd.overall_gcp, rd.length);
        }
        pos += rd.length;
    }
    pos = 0;
    for (int32_t h = 0; h < n_haps; ++h) {
        hoff[h] = pos;
        hlen[h] = haps[h].length;
        pd_column_codes(haps[h].haplotype_bases, haps[h].haplotype_pd_bases, haps[h].length, hcode + pos);
        pos += haps[h].length;
    }
    // Workgroups are dispatched in order: longest reads first, and each read's haplotypes longest
    // first, keeps the last wave of a launch short (outputs stay read-major).
    uint32_t *order = reinterpret_cast<uint32_t *>(hs + o_order);
    std::vector<int32_t> rs(n_reads), hsrt(n_haps);
    for (int32_t r = 0; r < n_reads; ++r) rs[r] = r;
    for (int32_t h = 0; h < n_haps; ++h) hsrt[h] = h;
    std::stable_sort(rs.begin(), rs.end(), [&](int32_t a, int32_t c) { return rlen[a] > rlen[c]; });
    std::stable_sort(hsrt.begin(), hsrt.end(), [&](int32_t a, int32_t c) { return hlen[a] > hlen[c]; });
    for (int64_t k = 0, i = 0; i < n_reads; ++i)
        for (int32_t j = 0; j < n_haps; ++j, ++k)
            order[k] = static_cast<uint32_t>(static_cast<int64_t>(rs[i]) * n_haps + hsrt[j]);
    uint8_t *ds = static_cast<uint8_t *>(ctx->stage_dev.ptr);
    HIP_TRY(hipMemcpyAsync(ds, hs, in_bytes, hipMemcpyHostToDevice, ctx->stream), PAIRHMM_ERR_KERNEL);

    PdBatch b{};
    b.order = reinterpret_cast<const uint32_t *>(ds + o_order);
    b.hap_code = reinterpret_cast<const uint32_t *>(ds + o_hc);
    b.hap_off = reinterpret_cast<const int64_t *>(ds + o_hoff);
    b.hap_len = reinterpret_cast<const int32_t *>(ds + o_hlen);
    b.rd_bases = ds + o_rb;
    b.rd_quals = ds + o_rq;
    b.rd_ins = ds + o_ri;
    b.rd_del = ds + o_rd;
    b.rd_gcp = ds + o_rg;
    b.rd_off = reinterpret_cast<const int64_t *>(ds + o_roff);
    b.rd_len = reinterpret_cast<const int32_t *>(ds + o_rlen);
    b.n_haps = n_haps;
    b.tristate = 1;
    b.out = reinterpret_cast<double *>(ds + o_out);
    DeviceTables tab{ctx->d_qual_err, ctx->d_mm_prob};
    // One wave per workgroup and per pair; the hardware dispatcher balances uneven pairs.
    const int grid = static_cast<int>(std::min<int64_t>(n_pairs, 1 << 20));
    HIP_TRY(launch_pd_forward(b, tab, static_cast<uint32_t>(n_pairs), max_h, grid, ctx->stream), PAIRHMM_ERR_KERNEL);
    HIP_TRY(hipMemcpyAsync(hs, ds + o_out, sizeof(double) * n_pairs, hipMemcpyDeviceToHost, ctx->stream),
            PAIRHMM_ERR_KERNEL);
    HIP_TRY(hipStreamSynchronize(ctx->stream), PAIRHMM_ERR_KERNEL);
    std::memcpy(likelihoods, hs, sizeof(double) * n_pairs);
    return PAIRHMM_OK;
}

// PairHMMLikelihoodCalculationEngine.java:200-201 on the device (postprocess.hip).
extern "C" int32_t pairhmm_normalize_and_filter_likelihoods(
    pairhmm_context *ctx, double *likelihoods, int64_t n_reads, int32_t n_alleles, int32_t reference_allele_index,
    double log10_mismapping_cap, int32_t symmetric, const uint8_t *hmm_base_quals, const int64_t *read_off,
    const int32_t *read_len, double expected_error_rate_per_base, int32_t dynamic_disqualification,
    double disqualification_scale, uint8_t *keep, double *threshold) {
    if (!ctx) return fail(PAIRHMM_ERR_STATE, "context is null");
    if (n_reads < 0) return fail(PAIRHMM_ERR_BAD_ARG, "negative read count");
    // AlleleLikelihoods.java:417-418 and :1352.
    if (!(log10_mismapping_cap < 0.0)) return fail(PAIRHMM_ERR_BAD_ARG, "the minimum reference likelihood fall must be negative");
    if (n_alleles <= 0) return fail(PAIRHMM_ERR_BAD_ARG, "unsupported for read-likelihood collections with no alleles");
    if (reference_allele_index < -1 || reference_allele_index >= n_alleles)
        return fail(PAIRHMM_ERR_BAD_ARG, "reference allele index out of range");
    if (!(expected_error_rate_per_base >= 0.0)) return fail(PAIRHMM_ERR_BAD_ARG, "negative expected error rate");
    if (n_reads == 0) return PAIRHMM_OK;
    if (!likelihoods || !read_len || !keep || (dynamic_disqualification && (!hmm_base_quals || !read_off)))
        return fail(PAIRHMM_ERR_BAD_ARG, "null device pointer");
    PostprocessArgs a{};
    a.likelihoods = likelihoods;
    a.n_reads = n_reads;
    a.n_alleles = n_alleles;
    a.ref_index = reference_allele_index;
    a.normalize = log10_mismapping_cap == -HUGE_VAL ? 0 : 1;
    a.symmetric = symmetric ? 1 : 0;
    a.cap = log10_mismapping_cap;
    a.base_quals = hmm_base_quals;
    a.read_off = read_off;
    a.read_len = read_len;
    a.max_error_per_base = expected_error_rate_per_base;
    a.dynamic = dynamic_disqualification ? 1 : 0;
    a.scale = disqualification_scale;
    a.keep = keep;
    a.threshold = threshold;
    HIP_TRY(hipSetDevice(ctx->device), PAIRHMM_ERR_KERNEL);
    HIP_TRY(launch_postprocess(a, ctx->stream), PAIRHMM_ERR_KERNEL);
    HIP_TRY(hipStreamSynchronize(ctx->stream), PAIRHMM_ERR_KERNEL);
    return PAIRHMM_OK;
}
