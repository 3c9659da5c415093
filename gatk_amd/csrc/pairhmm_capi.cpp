// pairhmm_capi.cpp -- C-ABI of the gfx950 PairHMM engine: contexts, validation, staging, batches.
//
// Drop-in boundary: PairHMMNativeBinding as called by
// src/main/java/org/broadinstitute/hellbender/utils/pairhmm/VectorLoglessPairHMM.java:57-168.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "capi_internal.hpp"

using namespace gatk_pairhmm;
using namespace gatk_pairhmm::capi;

struct pairhmm_batch {
    pairhmm_context *ctx = nullptr;
    int64_t n_pairs = 0;
    int64_t cells = 0;
    int max_hap_len = 1;
    void *d_mem = nullptr;  // one allocation holding every device array of the batch
    DeviceBatch view{};
    uint32_t *d_order = nullptr;    // work queue: pair indices by descending cost
    uint32_t *d_rescue = nullptr;   // rescue list (n_pairs capacity)
    uint32_t *d_counters = nullptr; // kCounterSlots: queue heads and rescue counters
    uint32_t *d_counter_init = nullptr; // their values at the start of a pass
    hipGraphExec_t graph = nullptr;
    std::vector<QueueSlice> slices; // FP32 and FP64 launches over the queue, one per LDS class run
    double *d_f64 = nullptr;        // FP64 boundary planes of the batch's FP64 launches
    size_t f64_bytes = 0;
    bool packed = false;            // FP32 pass on packed stripes (fixed at upload with the queue order)
};

extern "C" {

const char *pairhmm_last_error(void) { return g_last_error.c_str(); }

int32_t pairhmm_load(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    return device_is_gfx950(dev) ? 1 : 0;
}

int32_t pairhmm_initialize(const pairhmm_native_arguments *args, pairhmm_context **out_ctx) {
    if (out_ctx == nullptr) return fail(PAIRHMM_ERR_BAD_ARG, "out_ctx is null");
    *out_ctx = nullptr;
    if (!pairhmm_load()) return fail(PAIRHMM_ERR_NO_DEVICE, "no gfx950 device visible to this process");
    pairhmm_context *ctx = new (std::nothrow) pairhmm_context();
    if (!ctx) return fail(PAIRHMM_ERR_OOM, "context allocation failed");
    ctx->use_double = args ? args->use_double_precision != 0 : false;
    hipError_t e = hipGetDevice(&ctx->device);
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, ctx->device);
    if (e == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    static const ModelTables tables;  // built once per process (C++ magic static)
    if (e == hipSuccess) e = hipMalloc(&ctx->d_qual_err, sizeof(tables.qual_err));
    if (e == hipSuccess) e = hipMalloc(&ctx->d_mm_prob, sizeof(double) * tables.mm_prob.size());
    // Copies on the context's own stream, never the legacy null stream: a legacy-stream operation
    // while another thread captures a graph on its context invalidates that capture.
    if (e == hipSuccess)
        e = hipMemcpyAsync(ctx->d_qual_err, tables.qual_err, sizeof(tables.qual_err), hipMemcpyHostToDevice,
                           ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ctx->d_mm_prob, tables.mm_prob.data(), sizeof(double) * tables.mm_prob.size(),
                           hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        pairhmm_done(ctx);
        return fail(e == hipErrorOutOfMemory ? PAIRHMM_ERR_OOM : PAIRHMM_ERR_KERNEL,
                    std::string("initialize: ") + hipGetErrorString(e));
    }
    *out_ctx = ctx;
    return PAIRHMM_OK;
}

int32_t pairhmm_done(pairhmm_context *ctx) {
    if (!ctx) return PAIRHMM_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->stage_dev.release();
    ctx->stage_host.release();
    ctx->f64_planes.release();
    ctx->grad_scratch32.release();
    ctx->grad_scratch64.release();
    ctx->grad_work.release();
    if (ctx->d_qual_err) (void)hipFree(ctx->d_qual_err);
    if (ctx->d_mm_prob) (void)hipFree(ctx->d_mm_prob);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PAIRHMM_OK;
}

}  // extern "C"

extern "C" int32_t pairhmm_compute_likelihoods(pairhmm_context *ctx, const pairhmm_read_data_holder *reads,
                                               int32_t n_reads, const pairhmm_haplotype_data_holder *haps,
                                               int32_t n_haps, double *likelihoods) {
    if (!ctx) return fail(PAIRHMM_ERR_STATE, "context is null (initialize was not called)");
    if (n_reads < 0 || n_haps < 0) return fail(PAIRHMM_ERR_BAD_ARG, "negative read or haplotype count");
    if (n_reads == 0 || n_haps == 0) return PAIRHMM_OK;
    if (!reads || !haps || !likelihoods) return fail(PAIRHMM_ERR_BAD_ARG, "null array argument");
    int64_t read_bytes = 0, hap_bytes = 0;
    int max_h = 1;
    for (int32_t r = 0; r < n_reads; ++r) {
        const pairhmm_read_data_holder &rd = reads[r];
        int32_t st = validate_read(rd.read_bases, rd.read_quals, rd.insertion_gop, rd.deletion_gop, rd.overall_gcp,
                                   rd.length, r);
        if (st != PAIRHMM_OK) return st;
        read_bytes += rd.length;
    }
    for (int32_t h = 0; h < n_haps; ++h) {
        if (haps[h].length <= 0 || haps[h].haplotype_bases == nullptr)
            return fail(PAIRHMM_ERR_BAD_ARG, "haplotype " + std::to_string(h) + " is empty");
        hap_bytes += haps[h].length;
        max_h = std::max(max_h, haps[h].length);
    }
    const int64_t n_pairs = static_cast<int64_t>(n_reads) * n_haps;
    if (n_pairs > 0xFFFFFFFFll) return fail(PAIRHMM_ERR_BAD_ARG, "too many pairs in one call");
    // A haplotype too long for the FP32 kernels' LDS planes puts the whole call in FP64.
    const bool fp64_all = ctx->use_double || lds_columns(max_h) > kMaxFp32Cols;
    // Reads whose qualities predict an FP32 underflow: their pairs start on the rescue list and skip
    // the FP32 pass (the batch and stream APIs route them the same way).
    std::vector<uint8_t> direct(static_cast<size_t>(n_reads), 0);
    int64_t n_direct_reads = 0;
    if (!fp64_all)
        for (int32_t r = 0; r < n_reads; ++r)
            n_direct_reads += direct[r] = read_routes_direct(reads[r].read_quals, reads[r].length);
    const int64_t n_direct = n_direct_reads * n_haps, n32 = n_pairs - n_direct;

    // Staging layout (host pinned and device identical): 5 read planes, hap plane, offsets, lengths.
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    const size_t o_rb = take(read_bytes), o_rq = take(read_bytes), o_ri = take(read_bytes), o_rd = take(read_bytes),
                 o_rg = take(read_bytes), o_hb = take(hap_bytes), o_roff = take(sizeof(int64_t) * n_reads),
                 o_rlen = take(sizeof(int32_t) * n_reads), o_hoff = take(sizeof(int64_t) * n_haps),
                 o_hlen = take(sizeof(int32_t) * n_haps), o_cnt = take(sizeof(uint32_t) * 4),
                 o_order = take(sizeof(uint32_t) * (n_direct ? n32 : 0)),
                 o_rescue = take(sizeof(uint32_t) * n_pairs);
    // The counters and the head of the rescue list travel with the inputs; the copy covers the
    // rescue list up to its direct pairs.
    const size_t in_bytes = o_rescue + sizeof(uint32_t) * n_direct;
    const size_t o_out = take(sizeof(double) * n_pairs);
    const size_t total = off;

    HIP_TRY(hipSetDevice(ctx->device), PAIRHMM_ERR_KERNEL);
    if (ctx->stage_host.reserve(in_bytes) != hipSuccess) return fail(PAIRHMM_ERR_OOM, "pinned staging allocation failed");
    if (ctx->stage_dev.reserve(total) != hipSuccess) return fail(PAIRHMM_ERR_OOM, "device staging allocation failed");
    uint8_t *hs = static_cast<uint8_t *>(ctx->stage_host.ptr);
    int64_t *roff = reinterpret_cast<int64_t *>(hs + o_roff);
    int32_t *rlen = reinterpret_cast<int32_t *>(hs + o_rlen);
    int64_t *hoff = reinterpret_cast<int64_t *>(hs + o_hoff);
    int32_t *hlen = reinterpret_cast<int32_t *>(hs + o_hlen);
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
        std::memcpy(hs + o_hb + pos, haps[h].haplotype_bases, haps[h].length);
        pos += haps[h].length;
    }
    uint32_t *hcnt = reinterpret_cast<uint32_t *>(hs + o_cnt);
    hcnt[0] = hcnt[2] = hcnt[3] = 0;
    hcnt[1] = static_cast<uint32_t>(n_direct);
    if (n_direct) {
        uint32_t *ord = reinterpret_cast<uint32_t *>(hs + o_order), *res = reinterpret_cast<uint32_t *>(hs + o_rescue);
        for (int32_t r = 0; r < n_reads; ++r)
            for (int32_t h = 0; h < n_haps; ++h) {
                const uint32_t p = static_cast<uint32_t>(static_cast<int64_t>(r) * n_haps + h);
                *(direct[r] ? res++ : ord++) = p;
            }
    }
    uint8_t *ds = static_cast<uint8_t *>(ctx->stage_dev.ptr);
    HIP_TRY(hipMemcpyAsync(ds, hs, in_bytes, hipMemcpyHostToDevice, ctx->stream), PAIRHMM_ERR_KERNEL);

    DeviceBatch b{};
    b.hap_bases = ds + o_hb;
    b.hap_off = reinterpret_cast<const int64_t *>(ds + o_hoff);
    b.hap_len = reinterpret_cast<const int32_t *>(ds + o_hlen);
    b.rd_bases = ds + o_rb;
    b.rd_quals = ds + o_rq;
    b.rd_ins = ds + o_ri;
    b.rd_del = ds + o_rd;
    b.rd_gcp = ds + o_rg;
    b.rd_off = reinterpret_cast<const int64_t *>(ds + o_roff);
    b.rd_len = reinterpret_cast<const int32_t *>(ds + o_rlen);
    b.pair_read = nullptr;  // implicit read-major pairs
    b.pair_hap = nullptr;
    b.n_haps = n_haps;
    b.out = reinterpret_cast<double *>(ds + o_out);
    DeviceTables tab{ctx->d_qual_err, ctx->d_mm_prob};
    uint32_t *counters = reinterpret_cast<uint32_t *>(ds + o_cnt);
    uint32_t *rescue = reinterpret_cast<uint32_t *>(ds + o_rescue);

    bool packable = true;
    for (int32_t h = 0; h < n_haps && packable; ++h) packable = haps_packable(haps[h].haplotype_bases, haps[h].length);
    const LaunchShape sh32 = forward_shape(false, max_h, ctx->waves_per_block, use_packed(ctx, packable), ctx->step_unroll);
    const LaunchShape sh64 = forward_shape(true, max_h, ctx->waves_per_block);
    if (sh64.waves_per_block == 0 || (!fp64_all && sh32.waves_per_block == 0))
        return fail(PAIRHMM_ERR_BAD_ARG, "haplotype of " + std::to_string(max_h) + " bases exceeds the LDS budget");
    const int cus = std::max(ctx->num_cus, 1);
    auto grid_for = [&](const LaunchShape &sh) {
        const int64_t need = (n_pairs + sh.waves_per_block - 1) / sh.waves_per_block;
        return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(need, static_cast<int64_t>(sh.blocks_per_cu) * cus)));
    };
    WorkSource all{nullptr, static_cast<uint32_t>(n_pairs), nullptr, nullptr};
    // The FP32 pass takes its pairs from a queue whose head is counters[0] (zero in the staged
    // counters); the packed schedule takes them in chunks from it.
    const WorkSource fp32_items{n_direct ? reinterpret_cast<const uint32_t *>(ds + o_order) : nullptr,
                                static_cast<uint32_t>(n32), nullptr, counters};
    RescueSink none{nullptr, nullptr, 0.0};
    // FP64 boundary planes for every wave of the FP64 grid (a grid-stride launch needs them all).
    const int grid64 = grid_for(sh64);
    const size_t waves64 = static_cast<size_t>(grid64) * sh64.waves_per_block;
    if (ctx->f64_planes.reserve(waves64 * fp64_plane_bytes_per_wave(max_h)) != hipSuccess)
        return fail(PAIRHMM_ERR_OOM, "FP64 plane allocation failed");
    b.f64_planes = static_cast<double *>(ctx->f64_planes.ptr);
    b.f64_plane_waves = static_cast<uint32_t>(waves64);
    if (fp64_all) {
        HIP_TRY(launch_forward(true, b, tab, all, none, max_h, sh64, grid64, ctx->stream), PAIRHMM_ERR_KERNEL);
    } else {
        // counters[1] (the rescue count) starts at the direct pairs, copied in with the inputs.
        RescueSink rs{rescue, counters + 1, kFp32RescueThreshold};
        if (n32 > 0)
            HIP_TRY(launch_forward(false, b, tab, fp32_items, rs, max_h, sh32, grid_for(sh32), ctx->stream),
                    PAIRHMM_ERR_KERNEL);
        // FP64 recompute of the flagged pairs; the count is read on the device.
        WorkSource again{rescue, 0, counters + 1, nullptr};
        HIP_TRY(launch_forward(true, b, tab, again, none, max_h, sh64, grid64, ctx->stream), PAIRHMM_ERR_KERNEL);
    }
    HIP_TRY(hipMemcpyAsync(likelihoods, b.out, sizeof(double) * n_pairs, hipMemcpyDeviceToHost, ctx->stream),
            PAIRHMM_ERR_KERNEL);
    HIP_TRY(hipStreamSynchronize(ctx->stream), PAIRHMM_ERR_KERNEL);
    return PAIRHMM_OK;
}

// ---------------------------------------------------------------------------------------------
// Batch API
// ---------------------------------------------------------------------------------------------

namespace gatk_pairhmm {
namespace capi {

bool read_routes_direct(const uint8_t *quals, int32_t len) {
    // Per-quality terms in 2^-40 fixed point (floor(x * 2^40 + 0.5)): integer sums are exact, so the
    // predicate does not depend on the summation order and the loop carries no FP latency chain.
    static const std::vector<int64_t> score_of_q = [] {
        std::vector<int64_t> t(256);
        for (int q = 0; q < 256; ++q) {
            const double e = std::pow(10.0, -q / 10.0);
            const double v = e < 1.0 ? (1.0 - e) * std::log10(1.0 - e) + e * std::log10(e / 3.0) : -1.0;
            t[q] = static_cast<int64_t>(std::floor(v * 0x1p40 + 0.5));
        }
        return t;
    }();
    int64_t sc = 0;
    for (int32_t i = 0; i < len; ++i) sc += score_of_q[quals[i]];
    return sc < static_cast<int64_t>(kDirectFp64Score * 0x1p40);
}

void direct_rescue_init(const std::vector<QueueSlice> &slices, const std::vector<uint32_t> &order,
                        std::vector<uint32_t> &counter_init, std::vector<uint32_t> &rescue_init) {
    counter_init.assign(kCounterSlots, 0u);
    rescue_init.assign(order.size(), 0u);
    for (size_t k = 0; k < slices.size(); ++k) {
        const QueueSlice &q = slices[k];
        *slice_rescue_count(counter_init.data(), k) = static_cast<uint32_t>(q.n_direct);
        for (int64_t i = 0; i < q.n_direct; ++i)
            rescue_init[q.begin + i] = order[q.begin + q.count - q.n_direct + i];
    }
}

int32_t validate_and_order(const pairhmm_batch_desc *d, int queue_order, std::vector<uint32_t> &order, int64_t &cells,
                           int &max_h, std::vector<QueueSlice> &slices, bool route_direct, bool by_hap_len) {
    if (d->n_reads < 0 || d->n_haps <= 0 || d->n_pairs < 0 || d->read_bytes < 0 || d->hap_bytes <= 0)
        return fail(PAIRHMM_ERR_BAD_ARG, "bad batch sizes");
    if (d->n_pairs > 0xFFFFFFFFll) return fail(PAIRHMM_ERR_BAD_ARG, "batch exceeds 2^32 pairs");
    if (d->n_reads > 0 && (!d->read_off || !d->read_len)) return fail(PAIRHMM_ERR_BAD_ARG, "null read tables");
    if (!d->hap_bases || !d->hap_off || !d->hap_len) return fail(PAIRHMM_ERR_BAD_ARG, "null haplotype tables");
    if (d->n_pairs > 0 && (!d->pair_read || !d->pair_hap)) return fail(PAIRHMM_ERR_BAD_ARG, "null pair tables");
    max_h = 1;
    for (int64_t h = 0; h < d->n_haps; ++h) {
        if (d->hap_len[h] <= 0 || d->hap_off[h] < 0 || d->hap_off[h] + d->hap_len[h] > d->hap_bytes)
            return fail(PAIRHMM_ERR_BAD_ARG, "haplotype " + std::to_string(h) + " is empty or out of range");
        max_h = std::max(max_h, d->hap_len[h]);
    }
    for (int64_t r = 0; r < d->n_reads; ++r) {
        if (d->read_len[r] < 0 || d->read_off[r] < 0 || d->read_off[r] + d->read_len[r] > d->read_bytes)
            return fail(PAIRHMM_ERR_BAD_ARG, "read " + std::to_string(r) + " is out of range");
    }
    if (d->read_bytes > 0) {
        if (!d->read_bases || !d->read_quals || !d->read_ins || !d->read_del || !d->read_gcp)
            return fail(PAIRHMM_ERR_BAD_ARG, "null read plane");
        int32_t st = validate_read(d->read_bases, d->read_quals, d->read_ins, d->read_del, d->read_gcp, d->read_bytes, -1);
        if (st != PAIRHMM_OK) return st;
    }
    const int64_t np = d->n_pairs;
    order.resize(static_cast<size_t>(np));
    slices.clear();
    cells = 0;
    std::vector<uint32_t> cost(static_cast<size_t>(np));
    uint32_t max_cost = 0;
    // LDS class of every haplotype, once (a batch has far fewer haplotypes than pairs), and the pairs
    // and longest haplotype per class, counted in the pair loop.  top < 0: a haplotype too long for
    // any class, reported after the pair checks.
    const int top = lds_class_index(max_h);
    std::vector<uint8_t> hap_cls;
    std::vector<int64_t> n_in(static_cast<size_t>(std::max(top, 0)) + 1, 0);
    std::vector<int> h_in(static_cast<size_t>(std::max(top, 0)) + 1, 1);
    if (top >= 0) {
        hap_cls.resize(static_cast<size_t>(d->n_haps));
        for (int64_t h = 0; h < d->n_haps; ++h) hap_cls[h] = static_cast<uint8_t>(lds_class_index(d->hap_len[h]));
    }
    for (int64_t p = 0; p < np; ++p) {
        const int32_t r = d->pair_read[p], h = d->pair_hap[p];
        if (r < 0 || r >= d->n_reads || h < 0 || h >= d->n_haps)
            return fail(PAIRHMM_ERR_BAD_ARG, "pair " + std::to_string(p) + " indexes outside the batch");
        const int R = d->read_len[r], H = d->hap_len[h];
        cells += static_cast<int64_t>(R) * H;
        if (top >= 0) {
            const int k = hap_cls[h];
            n_in[k]++;
            h_in[k] = std::max(h_in[k], H);
        }
        // Aligned schedule: the pair's wave-steps, ceil(R/64) stripes of H + 64 steps.  Packed: the
        // haplotype length, so the pairs a wave takes one after another share stripe lengths.
        const uint32_t c = by_hap_len ? static_cast<uint32_t>(H)
                                      : static_cast<uint32_t>((R + kWave - 1) / kWave) * static_cast<uint32_t>(H + kWave);
        cost[p] = c;
        max_cost = std::max(max_cost, c);
    }
    if (np == 0) return PAIRHMM_OK;
    if (top < 0)
        return fail(PAIRHMM_ERR_BAD_ARG, "haplotype of " + std::to_string(max_h) + " bases exceeds the LDS budget");
    if (queue_order == 1) {
        // Submission order, one slice; an FP64-only class sends every pair to FP64.
        for (int64_t p = 0; p < np; ++p) order[p] = static_cast<uint32_t>(p);
        slices.push_back(QueueSlice{0, np, max_h, lds_columns(max_h) > kMaxFp32Cols ? np : 0});
        return PAIRHMM_OK;
    }
    // Counting sort by descending cost, then a stable counting sort by descending LDS class.
    std::vector<uint32_t> by_cost(static_cast<size_t>(np));
    {
        std::vector<int64_t> start(static_cast<size_t>(max_cost) + 2, 0);
        for (int64_t p = 0; p < np; ++p) start[max_cost - cost[p] + 1]++;
        for (size_t i = 1; i < start.size(); ++i) start[i] += start[i - 1];
        for (int64_t p = 0; p < np; ++p) by_cost[start[max_cost - cost[p]]++] = static_cast<uint32_t>(p);
    }
    // Classes with few pairs are merged into the next larger one: a launch per class costs a drain
    // of the persistent grid, worth it only for a class with real work.  The last FP32 class always
    // closes a run, so no FP32 pair is merged into an FP64-only class.  At most kMaxSlices runs.
    std::vector<int> run_of(static_cast<size_t>(top) + 1, 0);
    std::vector<int64_t> run_start;
    std::vector<int> run_class;
    {
        int64_t carried = 0;
        for (int k = 0; k <= top; ++k) {
            carried += n_in[k];
            const bool close = (k == top) || kColClasses[k] == kMaxFp32Cols ||
                               carried >= std::max<int64_t>(4096, np / 64);
            run_of[k] = static_cast<int>(run_class.size());
            if (close && carried > 0) {
                run_class.push_back(k);
                carried = 0;
            }
        }
        // classes merged into a later run take that run's index
        for (int k = top; k >= 0; --k)
            if (run_of[k] >= static_cast<int>(run_class.size())) run_of[k] = static_cast<int>(run_class.size()) - 1;
        while (static_cast<int>(run_class.size()) > kMaxSlices) {
            // fold the smallest run into its larger neighbour
            for (int k = 0; k <= top; ++k)
                if (run_of[k] == 0) run_of[k] = 1;
            run_class.erase(run_class.begin());
            for (int k = 0; k <= top; ++k) run_of[k] -= 1;
        }
    }
    const int n_runs = static_cast<int>(run_class.size());
    std::vector<uint8_t> hap_run(static_cast<size_t>(d->n_haps));
    for (int64_t h = 0; h < d->n_haps; ++h) hap_run[h] = static_cast<uint8_t>(run_of[hap_cls[h]]);
    std::vector<int64_t> run_n(static_cast<size_t>(n_runs), 0), run_direct(static_cast<size_t>(n_runs), 0);
    std::vector<int> run_h(static_cast<size_t>(n_runs), 1);
    for (int k = 0; k <= top; ++k) {
        run_n[run_of[k]] += n_in[k];
        if (n_in[k] > 0) run_h[run_of[k]] = std::max(run_h[run_of[k]], h_in[k]);
    }
    // FP64 pairs ahead of any FP32 pass: with route_direct, those whose read predicts an FP32
    // underflow; always, those of a run whose class is FP64-only (haplotypes too long for FP32).
    std::vector<uint8_t> direct;
    const bool wide = kColClasses[run_class.back()] > kMaxFp32Cols;
    if (route_direct || wide) {
        std::vector<uint8_t> read_direct(static_cast<size_t>(d->n_reads), 0);
        if (route_direct)
            for (int64_t r = 0; r < d->n_reads; ++r)
                read_direct[r] = read_routes_direct(d->read_quals + d->read_off[r], d->read_len[r]);
        direct.resize(static_cast<size_t>(np));
        for (int64_t p = 0; p < np; ++p) {
            const int r = hap_run[d->pair_hap[p]];
            direct[p] = read_direct[d->pair_read[p]] || kColClasses[run_class[r]] > kMaxFp32Cols;
            if (direct[p]) run_direct[r]++;
        }
    }
    // Largest class first: runs in descending order of their class; within a run the FP32 queue by
    // descending cost, then the direct FP64 pairs by descending cost.
    std::vector<int64_t> pos(static_cast<size_t>(n_runs), 0), pos_direct(static_cast<size_t>(n_runs), 0);
    int64_t acc = 0;
    for (int r = n_runs - 1; r >= 0; --r) {
        pos[r] = acc;
        pos_direct[r] = acc + run_n[r] - run_direct[r];
        if (run_n[r] > 0) slices.push_back(QueueSlice{acc, run_n[r], run_h[r], run_direct[r]});
        acc += run_n[r];
    }
    for (int64_t i = 0; i < np; ++i) {
        const uint32_t p = by_cost[i];
        const int r = hap_run[d->pair_hap[p]];
        if (!direct.empty() && direct[p])
            order[pos_direct[r]++] = p;
        else
            order[pos[r]++] = p;
    }
    return PAIRHMM_OK;
}

int persistent_grid(const pairhmm_context *ctx, const LaunchShape &shape, int64_t n_items) {
    const int64_t g = static_cast<int64_t>(shape.blocks_per_cu) * std::max(ctx->num_cus, 1) * ctx->grid_percent / 100;
    const int64_t need = (std::max<int64_t>(n_items, 1) + shape.waves_per_block - 1) / std::max(shape.waves_per_block, 1);
    return static_cast<int>(std::max<int64_t>(1, std::min(g, need)));
}

hipError_t launch_fp32_slices(const pairhmm_context *ctx, const DeviceBatch &batch, const DeviceTables &tab,
                              const uint32_t *order, const std::vector<QueueSlice> &slices, uint32_t *counters,
                              uint32_t *rescue_list, hipStream_t s, bool packed) {
    if (slices.size() > static_cast<size_t>(kMaxSlices)) return hipErrorInvalidValue;
    for (size_t k = 0; k < slices.size(); ++k) {
        const QueueSlice &q = slices[k];
        const int64_t n32 = q.count - q.n_direct;
        if (n32 == 0) continue;  // an FP64-only slice, or every pair routed to FP64
        const LaunchShape sh = forward_shape(false, q.max_hap_len, ctx->waves_per_block, packed, ctx->step_unroll);
        if (sh.waves_per_block == 0) return hipErrorInvalidValue;
        RescueSink rs{rescue_list + q.begin, slice_rescue_count(counters, k), kFp32RescueThreshold};
        WorkSource w{order + q.begin, static_cast<uint32_t>(n32), nullptr, slice_head(counters, k)};
        const hipError_t e = launch_forward(false, batch, tab, w, rs, q.max_hap_len, sh,
                                            persistent_grid(ctx, sh, n32), s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

size_t fp64_slices_plane_bytes(const pairhmm_context *ctx, const std::vector<QueueSlice> &slices) {
    size_t need = 0;
    for (const QueueSlice &q : slices) {
        const LaunchShape sh = forward_shape(true, q.max_hap_len, ctx->waves_per_block);
        const size_t waves = static_cast<size_t>(persistent_grid(ctx, sh, q.count)) * sh.waves_per_block;
        need = std::max(need, waves * fp64_plane_bytes_per_wave(q.max_hap_len));
    }
    return need;
}

hipError_t launch_fp64_slices(const pairhmm_context *ctx, const DeviceBatch &batch, const DeviceTables &tab,
                              const uint32_t *list, const std::vector<QueueSlice> &slices, uint32_t *counters,
                              bool rescue_pass, double *planes, size_t plane_bytes, hipStream_t s) {
    if (slices.size() > static_cast<size_t>(kMaxSlices)) return hipErrorInvalidValue;
    RescueSink none{nullptr, nullptr, 0.0};
    for (size_t k = 0; k < slices.size(); ++k) {
        const QueueSlice &q = slices[k];
        const LaunchShape sh = forward_shape(true, q.max_hap_len, ctx->waves_per_block);
        const size_t per_wave = fp64_plane_bytes_per_wave(q.max_hap_len);
        if (sh.waves_per_block == 0 || per_wave == 0 || planes == nullptr) return hipErrorInvalidValue;
        const size_t cap_waves = plane_bytes / per_wave;
        if (cap_waves < static_cast<size_t>(sh.waves_per_block)) return hipErrorInvalidValue;
        // A rescue pass reads its item count on the device; the grid is sized for the slice's
        // capacity and surplus waves leave at their first queue fetch.
        WorkSource w = rescue_pass ? WorkSource{list + q.begin, 0, slice_rescue_count(counters, k), slice_rescue_head(counters, k)}
                                   : WorkSource{list + q.begin, static_cast<uint32_t>(q.count), nullptr, slice_head(counters, k)};
        const int grid = static_cast<int>(std::min<size_t>(static_cast<size_t>(persistent_grid(ctx, sh, q.count)),
                                                           cap_waves / sh.waves_per_block));
        DeviceBatch v = batch;
        v.f64_planes = planes;
        v.f64_plane_waves = static_cast<uint32_t>(std::min<size_t>(cap_waves, 0xFFFFFFFFu));
        const hipError_t e = launch_forward(true, v, tab, w, none, q.max_hap_len, sh, grid, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace capi
}  // namespace gatk_pairhmm

extern "C" int32_t pairhmm_batch_upload(pairhmm_context *ctx, const pairhmm_batch_desc *d, pairhmm_batch **out) {
    if (!ctx) return fail(PAIRHMM_ERR_STATE, "context is null");
    if (!d || !out) return fail(PAIRHMM_ERR_BAD_ARG, "null argument");
    *out = nullptr;
    std::vector<uint32_t> order;
    std::vector<QueueSlice> slices;
    int64_t cells = 0;
    int max_h = 1;
    const bool packed = d->hap_bases && use_packed(ctx, haps_packable(d->hap_bases, d->hap_bytes));
    int32_t vst = validate_and_order(d, ctx->queue_order, order, cells, max_h, slices, true, packed);
    if (vst != PAIRHMM_OK) return vst;
    const int64_t np = d->n_pairs;
    // Counter values at the start of every pass (the rescue list's head is copied once, below).
    std::vector<uint32_t> counter_init, rescue_init;
    direct_rescue_init(slices, order, counter_init, rescue_init);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    const size_t rb = static_cast<size_t>(d->read_bytes), hb = static_cast<size_t>(d->hap_bytes);
    const size_t o_rb = take(rb), o_rq = take(rb), o_ri = take(rb), o_rd = take(rb), o_rg = take(rb), o_hb = take(hb),
                 o_roff = take(8 * d->n_reads), o_rlen = take(4 * d->n_reads), o_hoff = take(8 * d->n_haps),
                 o_hlen = take(4 * d->n_haps), o_pr = take(4 * np), o_ph = take(4 * np), o_ord = take(4 * np),
                 o_out = take(8 * np), o_res = take(4 * np), o_cnt = take(4 * kCounterSlots),
                 o_cinit = take(4 * kCounterSlots);
    HIP_TRY(hipSetDevice(ctx->device), PAIRHMM_ERR_KERNEL);
    pairhmm_batch *bt = new (std::nothrow) pairhmm_batch();
    if (!bt) return fail(PAIRHMM_ERR_OOM, "batch allocation failed");
    bt->ctx = ctx;
    bt->n_pairs = np;
    bt->cells = cells;
    bt->max_hap_len = max_h;
    bt->packed = packed;
    if (hipMalloc(&bt->d_mem, off) != hipSuccess) {
        delete bt;
        return fail(PAIRHMM_ERR_OOM, "device allocation of " + std::to_string(off) + " bytes failed");
    }
    uint8_t *m = static_cast<uint8_t *>(bt->d_mem);
    struct Cp {
        size_t o;
        const void *src;
        size_t n;
    } copies[] = {{o_rb, d->read_bases, rb}, {o_rq, d->read_quals, rb}, {o_ri, d->read_ins, rb},
                  {o_rd, d->read_del, rb},   {o_rg, d->read_gcp, rb},   {o_hb, d->hap_bases, hb},
                  {o_roff, d->read_off, 8 * (size_t)d->n_reads}, {o_rlen, d->read_len, 4 * (size_t)d->n_reads},
                  {o_hoff, d->hap_off, 8 * (size_t)d->n_haps},   {o_hlen, d->hap_len, 4 * (size_t)d->n_haps},
                  {o_pr, d->pair_read, 4 * (size_t)np},          {o_ph, d->pair_hap, 4 * (size_t)np},
                  {o_ord, order.data(), 4 * (size_t)np},         {o_res, rescue_init.data(), 4 * (size_t)np},
                  {o_cinit, counter_init.data(), 4 * (size_t)kCounterSlots}};
    // On the context's stream, not the legacy null stream (see pairhmm_initialize); the sources are
    // the caller's pageable arrays, borrowed for this call only, so the copies complete before return.
    for (const Cp &c : copies) {
        if (c.n == 0) continue;
        hipError_t e = hipMemcpyAsync(m + c.o, c.src, c.n, hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) {
            (void)hipStreamSynchronize(ctx->stream);
            pairhmm_batch_free(bt);
            return fail(PAIRHMM_ERR_KERNEL, std::string("batch upload: ") + hipGetErrorString(e));
        }
    }
    {
        const hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            pairhmm_batch_free(bt);
            return fail(PAIRHMM_ERR_KERNEL, std::string("batch upload: ") + hipGetErrorString(e));
        }
    }
    DeviceBatch &v = bt->view;
    v.rd_bases = m + o_rb;
    v.rd_quals = m + o_rq;
    v.rd_ins = m + o_ri;
    v.rd_del = m + o_rd;
    v.rd_gcp = m + o_rg;
    v.hap_bases = m + o_hb;
    v.rd_off = reinterpret_cast<const int64_t *>(m + o_roff);
    v.rd_len = reinterpret_cast<const int32_t *>(m + o_rlen);
    v.hap_off = reinterpret_cast<const int64_t *>(m + o_hoff);
    v.hap_len = reinterpret_cast<const int32_t *>(m + o_hlen);
    v.pair_read = reinterpret_cast<const int32_t *>(m + o_pr);
    v.pair_hap = reinterpret_cast<const int32_t *>(m + o_ph);
    v.n_haps = static_cast<int32_t>(d->n_haps);
    v.out = reinterpret_cast<double *>(m + o_out);
    bt->d_order = reinterpret_cast<uint32_t *>(m + o_ord);
    bt->d_rescue = reinterpret_cast<uint32_t *>(m + o_res);
    bt->d_counters = reinterpret_cast<uint32_t *>(m + o_cnt);
    bt->d_counter_init = reinterpret_cast<uint32_t *>(m + o_cinit);
    bt->slices = std::move(slices);
    bt->f64_bytes = fp64_slices_plane_bytes(ctx, bt->slices);
    if (bt->f64_bytes > 0 && hipMalloc(reinterpret_cast<void **>(&bt->d_f64), bt->f64_bytes) != hipSuccess) {
        bt->d_f64 = nullptr;
        pairhmm_batch_free(bt);
        return fail(PAIRHMM_ERR_OOM, "FP64 plane allocation failed");
    }
    // FP32 slices hold haplotypes of FP32 classes only (validate_and_order), so the FP64 kernel's
    // shape is the one to check.
    if (forward_shape(true, max_h, ctx->waves_per_block).waves_per_block == 0) {
        pairhmm_batch_free(bt);
        return fail(PAIRHMM_ERR_BAD_ARG, "haplotype of " + std::to_string(max_h) + " bases exceeds the LDS budget");
    }
    *out = bt;
    return PAIRHMM_OK;
}

namespace {

// Enqueues one full forward pass of the batch: counter reset, persistent FP32 kernel over the
// queue, persistent FP64 kernel over the rescue list.
hipError_t enqueue_forward(pairhmm_batch *bt, hipStream_t s, hipEvent_t ev[4]) {
    pairhmm_context *ctx = bt->ctx;
    DeviceTables tab{ctx->d_qual_err, ctx->d_mm_prob};
    hipError_t e = launch_copy_counters(bt->d_counters, bt->d_counter_init, kCounterSlots, s);
    if (e != hipSuccess) return e;
    if (ev) (void)hipEventRecord(ev[0], s);
    e = launch_fp32_slices(ctx, bt->view, tab, bt->d_order, bt->slices, bt->d_counters, bt->d_rescue, s, bt->packed);
    if (e != hipSuccess) return e;
    if (ev) (void)hipEventRecord(ev[1], s);
    e = launch_fp64_slices(ctx, bt->view, tab, bt->d_rescue, bt->slices, bt->d_counters, true, bt->d_f64, bt->f64_bytes,
                           s);
    if (ev) (void)hipEventRecord(ev[2], s);
    return e;
}

}  // namespace

extern "C" int32_t pairhmm_batch_forward(pairhmm_context *ctx, pairhmm_batch *bt, int32_t mode, double *out_host,
                                         pairhmm_run_stats *stats) {
    if (!ctx || !bt || bt->ctx != ctx) return fail(PAIRHMM_ERR_STATE, "batch does not belong to this context");
    HIP_TRY(hipSetDevice(ctx->device), PAIRHMM_ERR_KERNEL);
    hipStream_t s = ctx->stream;
    hipEvent_t ev[4] = {};
    for (auto &e : ev) HIP_TRY(hipEventCreate(&e), PAIRHMM_ERR_KERNEL);
    struct EvGuard {
        hipEvent_t *e;
        ~EvGuard() {
            for (int i = 0; i < 4; ++i) (void)hipEventDestroy(e[i]);
        }
    } guard{ev};
    pairhmm_run_stats st{};
    st.cells = bt->cells;
    if (bt->n_pairs == 0) {
        if (stats) *stats = st;
        return PAIRHMM_OK;
    }
    if (mode == PAIRHMM_RUN_PERSISTENT) {
        if (!bt->graph) {
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), PAIRHMM_ERR_KERNEL);
            hipError_t e = enqueue_forward(bt, s, nullptr);
            hipError_t e2 = hipStreamEndCapture(s, &g);
            if (e != hipSuccess || e2 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return fail(PAIRHMM_ERR_KERNEL, std::string("graph capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
            }
            e = hipGraphInstantiate(&bt->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return fail(PAIRHMM_ERR_KERNEL, std::string("graph instantiate: ") + hipGetErrorString(e));
        }
        HIP_TRY(hipEventRecord(ev[0], s), PAIRHMM_ERR_KERNEL);
        HIP_TRY(hipGraphLaunch(bt->graph, s), PAIRHMM_ERR_KERNEL);
        HIP_TRY(hipEventRecord(ev[3], s), PAIRHMM_ERR_KERNEL);
    } else if (mode == PAIRHMM_RUN_PERSISTENT_EAGER) {
        HIP_TRY(enqueue_forward(bt, s, ev), PAIRHMM_ERR_KERNEL);
        HIP_TRY(hipEventRecord(ev[3], s), PAIRHMM_ERR_KERNEL);
    } else if (mode == PAIRHMM_RUN_FP64) {
        DeviceTables tab{ctx->d_qual_err, ctx->d_mm_prob};
        HIP_TRY(launch_zero_counters(bt->d_counters, kCounterSlots, s), PAIRHMM_ERR_KERNEL);
        HIP_TRY(hipEventRecord(ev[0], s), PAIRHMM_ERR_KERNEL);
        HIP_TRY(launch_fp64_slices(ctx, bt->view, tab, bt->d_order, bt->slices, bt->d_counters, false, bt->d_f64,
                                   bt->f64_bytes, s),
                PAIRHMM_ERR_KERNEL);
        HIP_TRY(hipEventRecord(ev[3], s), PAIRHMM_ERR_KERNEL);
    } else {
        return fail(PAIRHMM_ERR_BAD_ARG, "unknown run mode");
    }
    uint32_t counters[kCounterSlots] = {};
    HIP_TRY(hipMemcpyAsync(counters, bt->d_counters, sizeof(counters), hipMemcpyDeviceToHost, s), PAIRHMM_ERR_KERNEL);
    if (out_host)
        HIP_TRY(hipMemcpyAsync(out_host, bt->view.out, sizeof(double) * bt->n_pairs, hipMemcpyDeviceToHost, s),
                PAIRHMM_ERR_KERNEL);
    HIP_TRY(hipStreamSynchronize(s), PAIRHMM_ERR_KERNEL);
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ev[0], ev[3]) == hipSuccess) st.total_ms = ms;
    if (mode == PAIRHMM_RUN_PERSISTENT_EAGER) {
        if (hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) st.fp32_kernel_ms = ms;
        if (hipEventElapsedTime(&ms, ev[1], ev[2]) == hipSuccess) st.rescue_kernel_ms = ms;
    }
    st.n_rescued = (mode == PAIRHMM_RUN_FP64) ? bt->n_pairs : rescued_total(counters);
    if (mode != PAIRHMM_RUN_FP64)
        for (const QueueSlice &q : bt->slices) st.n_routed_fp64 += q.n_direct;
    if (stats) *stats = st;
    return PAIRHMM_OK;
}

extern "C" int32_t pairhmm_batch_free(pairhmm_batch *bt) {
    if (!bt) return PAIRHMM_OK;
    if (bt->ctx) (void)hipSetDevice(bt->ctx->device);
    if (bt->ctx && bt->ctx->stream) (void)hipStreamSynchronize(bt->ctx->stream);
    if (bt->graph) (void)hipGraphExecDestroy(bt->graph);
    if (bt->d_mem) (void)hipFree(bt->d_mem);
    if (bt->d_f64) (void)hipFree(bt->d_f64);
    delete bt;
    return PAIRHMM_OK;
}

// ---------------------------------------------------------------------------------------------
// Forward + backward
// ---------------------------------------------------------------------------------------------

namespace {

// Fixed device memory a forward + backward pass carves its work lists and checkpoints from instead of
// the context's growable buffers (pairhmm_forward_backward_noalloc).
struct FixedScratch {
    uint8_t *ptr = nullptr;
    size_t bytes = 0;
};

// pairhmm_forward_backward over the context's buffers (fixed == nullptr), or over `fixed`: then
// nothing is allocated, no event is created, and each launch's persistent grid is held to the
// waves whose checkpoints fit.
int32_t forward_backward(pairhmm_context *ctx, const pairhmm_grad_desc *d, pairhmm_run_stats *stats,
                         const FixedScratch *fixed) {
    if (!ctx) return fail(PAIRHMM_ERR_STATE, "context is null");
    if (!d) return fail(PAIRHMM_ERR_BAD_ARG, "null descriptor");
    if (d->n_pairs < 0 || d->n_reads < 0 || d->n_haps < 0 || d->read_bytes < 0)
        return fail(PAIRHMM_ERR_BAD_ARG, "negative sizes");
    if (d->n_pairs > 0xFFFFFFFFll) return fail(PAIRHMM_ERR_BAD_ARG, "more than 2^32 pairs");
    pairhmm_run_stats st{};
    if (d->n_pairs == 0) {
        if (stats) *stats = st;
        return PAIRHMM_OK;
    }
    if (d->max_hap_len <= 0 || d->max_read_len < 0)
        return fail(PAIRHMM_ERR_BAD_ARG, "max_hap_len must be > 0 and max_read_len >= 0");
    if (!d->read_bases || !d->read_qb || !d->read_qi || !d->read_qd || !d->read_qc || !d->read_off || !d->read_len ||
        !d->hap_bases || !d->hap_off || !d->hap_len || !d->pair_read || !d->pair_hap || !d->out_log10_likelihood ||
        !d->grad_qb || !d->grad_qi || !d->grad_qd || !d->grad_qc)
        return fail(PAIRHMM_ERR_BAD_ARG, "null device pointer");
    HIP_TRY(hipSetDevice(ctx->device), PAIRHMM_ERR_KERNEL);
    const LaunchShape sh32max = grad_shape(false, d->max_hap_len, 4);
    const LaunchShape sh64 = grad_shape(true, d->max_hap_len, 4);
    if (sh32max.waves_per_block == 0 || sh64.waves_per_block == 0)
        return fail(PAIRHMM_ERR_BAD_ARG, "haplotype of " + std::to_string(d->max_hap_len) +
                                             " bases exceeds the adjoint kernel's LDS budget");
    const int max_class = lds_class_index(d->max_hap_len);
    const uint32_t np = static_cast<uint32_t>(d->n_pairs);
    // Work area: 64 counter slots ([1] FP64 item count, [2] FP64 head, [16+k] pairs of class k,
    // [32+r] scatter cursor of run r, [48+r] queue head of run r), the FP64 list, the class order.
    const size_t work_bytes = sizeof(uint32_t) * (64 + 2 * static_cast<size_t>(np));
    uint32_t *counters = nullptr;
    if (fixed) {
        if (align_up(work_bytes, 256) >= fixed->bytes)
            return fail(PAIRHMM_ERR_BAD_ARG, std::to_string(np) + " pairs exceed the fixed scratch (" +
                                                 std::to_string(fixed->bytes) + " bytes)");
        counters = reinterpret_cast<uint32_t *>(fixed->ptr);
    } else {
        if (ctx->grad_work.reserve(work_bytes) != hipSuccess)
            return fail(PAIRHMM_ERR_OOM, "adjoint work-list allocation failed");
        counters = static_cast<uint32_t *>(ctx->grad_work.ptr);
    }
    uint32_t *rescue = counters + 64;
    uint32_t *order = rescue + np;
    hipStream_t s = ctx->stream;
    hipEvent_t ev[3] = {};
    if (!fixed)
        for (auto &e : ev) HIP_TRY(hipEventCreate(&e), PAIRHMM_ERR_KERNEL);
    struct EvGuard {
        hipEvent_t *e;
        ~EvGuard() {
            for (int i = 0; i < 3; ++i)
                if (e[i]) (void)hipEventDestroy(e[i]);
        }
    } guard{ev};
    if (!fixed) HIP_TRY(hipEventRecord(ev[0], s), PAIRHMM_ERR_KERNEL);
    HIP_TRY(launch_zero_counters(counters, 64, s), PAIRHMM_ERR_KERNEL);

    // FP32 launches: one per run of LDS classes, largest class first; classes with few pairs merge
    // into the next larger class present (the forward's queue slices, binned on the device here
    // because the pair table is device-resident).
    struct Run {
        int cls;
        uint32_t begin, count;
        LaunchShape shape;
        int grid;
    };
    struct RunList {
        Run item[kNumColClasses];
        size_t n = 0;
        size_t size() const { return n; }
        void push_back(const Run &r) { item[n++] = r; }
        Run &operator[](size_t i) { return item[i]; }
        const Run &operator[](size_t i) const { return item[i]; }
        Run *begin() { return item; }
        Run *end() { return item + n; }
        const Run *begin() const { return item; }
        const Run *end() const { return item + n; }
    } runs;  // at most one run per LDS class; fixed storage
    ClassRuns cr{};
    if (!d->use_double) {
        HIP_TRY(launch_grad_class_count(d->pair_hap, d->hap_len, np, max_class, counters + 16, s), PAIRHMM_ERR_KERNEL);
        uint32_t counts[kNumColClasses] = {};
        HIP_TRY(hipMemcpyAsync(counts, counters + 16, sizeof(counts), hipMemcpyDeviceToHost, s), PAIRHMM_ERR_KERNEL);
        HIP_TRY(hipStreamSynchronize(s), PAIRHMM_ERR_KERNEL);
        const uint32_t merge_below = std::max<uint32_t>(4096, np / 64);
        uint32_t pending = 0;
        int members[kNumColClasses];
        int n_members = 0;
        for (int k = 0; k < kNumColClasses; ++k) cr.run_of_class[k] = -1;
        for (int k = 0; k <= max_class; ++k) {
            if (counts[k] == 0) continue;
            pending += counts[k];
            members[n_members++] = k;
            bool later = false;
            for (int j = k + 1; j <= max_class; ++j) later |= counts[j] > 0;
            if (pending >= merge_below || !later) {
                for (int i = 0; i < n_members; ++i) cr.run_of_class[members[i]] = static_cast<int32_t>(runs.size());
                runs.push_back(Run{k, 0, pending, LaunchShape{}, 0});
                pending = 0;
                n_members = 0;
            }
        }
        uint32_t acc = 0;  // largest class first in the order list
        for (size_t r = runs.size(); r-- > 0;) {
            runs[r].begin = acc;
            cr.run_begin[r] = acc;
            acc += runs[r].count;
            const int cap = kColClasses[runs[r].cls] - 2 * kLdsPad - 32;  // longest haplotype of the class
            runs[r].shape = grad_shape(false, std::min(cap, d->max_hap_len), 4);
            if (runs[r].shape.waves_per_block == 0) return fail(PAIRHMM_ERR_KERNEL, "adjoint class shape");
            runs[r].grid = persistent_grid(ctx, runs[r].shape, runs[r].count);
        }
        HIP_TRY(launch_grad_class_scatter(d->pair_hap, d->hap_len, np, max_class, cr, counters + 32, order, s),
                PAIRHMM_ERR_KERNEL);
    }
    int max_stripes = 0, max_tiles = 0;
    const size_t per_wave = grad_scratch_elems_per_wave(d->max_read_len, d->max_hap_len, &max_stripes, &max_tiles);
    const int cus = std::max(ctx->num_cus, 1);
    int grid64 = sh64.blocks_per_cu * cus;
    void *scratch32 = nullptr, *scratch64 = nullptr;
    if (fixed) {
        // The FP32 runs and the FP64 launch follow each other on one stream, so their checkpoints
        // share the space after the work area; each grid is held to the blocks whose waves fit.
        const size_t ck_off = align_up(work_bytes, 256);
        const size_t ck_bytes = fixed->bytes - ck_off;
        auto blocks_fit = [&](const LaunchShape &sh, size_t elem) {
            return ck_bytes / (static_cast<size_t>(sh.waves_per_block) * per_wave * elem);
        };
        if (blocks_fit(sh64, sizeof(double)) < 1)
            return fail(PAIRHMM_ERR_BAD_ARG, "reads of " + std::to_string(d->max_read_len) + " and haplotypes of " +
                                                 std::to_string(d->max_hap_len) +
                                                 " bases need more checkpoint space than the fixed scratch holds");
        grid64 = static_cast<int>(std::min<size_t>(static_cast<size_t>(grid64), blocks_fit(sh64, sizeof(double))));
        for (Run &r : runs) {
            const size_t fit = blocks_fit(r.shape, sizeof(float));
            if (fit < 1) return fail(PAIRHMM_ERR_BAD_ARG, "adjoint checkpoints exceed the fixed scratch");
            r.grid = static_cast<int>(std::min<size_t>(static_cast<size_t>(r.grid), fit));
        }
        scratch32 = scratch64 = fixed->ptr + ck_off;
    } else {
        size_t waves32 = 0;
        for (const Run &r : runs) waves32 = std::max(waves32, static_cast<size_t>(r.grid) * r.shape.waves_per_block);
        const size_t waves64 = static_cast<size_t>(grid64) * sh64.waves_per_block;
        if (ctx->grad_scratch32.reserve(std::max<size_t>(waves32, 1) * per_wave * sizeof(float)) != hipSuccess ||
            ctx->grad_scratch64.reserve(waves64 * per_wave * sizeof(double)) != hipSuccess)
            return fail(PAIRHMM_ERR_OOM, "adjoint scratch allocation failed");
        scratch32 = ctx->grad_scratch32.ptr;
        scratch64 = ctx->grad_scratch64.ptr;
    }
    GradBatch g{};
    g.hap_bases = d->hap_bases;
    g.hap_off = d->hap_off;
    g.hap_len = d->hap_len;
    g.rd_bases = d->read_bases;
    g.rd_qb = d->read_qb;
    g.rd_qi = d->read_qi;
    g.rd_qd = d->read_qd;
    g.rd_qc = d->read_qc;
    g.rd_off = d->read_off;
    g.rd_len = d->read_len;
    g.pair_read = d->pair_read;
    g.pair_hap = d->pair_hap;
    g.pair_weight = d->pair_weight;
    g.out_log10 = d->out_log10_likelihood;
    g.g_qb = d->grad_qb;
    g.g_qi = d->grad_qi;
    g.g_qd = d->grad_qd;
    g.g_qc = d->grad_qc;
    g.g_eb = d->grad_base_error;
    g.scratch_per_wave = per_wave;
    g.max_stripes = max_stripes;
    g.max_tiles = max_tiles;
    const size_t gbytes = sizeof(double) * static_cast<size_t>(d->read_bytes);
    HIP_TRY(hipMemsetAsync(d->grad_qb, 0, gbytes, s), PAIRHMM_ERR_KERNEL);
    HIP_TRY(hipMemsetAsync(d->grad_qi, 0, gbytes, s), PAIRHMM_ERR_KERNEL);
    HIP_TRY(hipMemsetAsync(d->grad_qd, 0, gbytes, s), PAIRHMM_ERR_KERNEL);
    HIP_TRY(hipMemsetAsync(d->grad_qc, 0, gbytes, s), PAIRHMM_ERR_KERNEL);
    if (d->grad_base_error) HIP_TRY(hipMemsetAsync(d->grad_base_error, 0, gbytes, s), PAIRHMM_ERR_KERNEL);
    RescueSink none{nullptr, nullptr, 0.0};
    if (d->use_double) {
        if (!fixed) HIP_TRY(hipEventRecord(ev[1], s), PAIRHMM_ERR_KERNEL);
        g.scratch = scratch64;
        WorkSource all{nullptr, static_cast<uint32_t>(d->n_pairs), nullptr, counters + 2};
        HIP_TRY(launch_grad(true, g, all, none, d->max_hap_len, sh64, grid64, s), PAIRHMM_ERR_KERNEL);
    } else {
        g.scratch = scratch32;
        RescueSink rs{rescue, counters + 1, kFp32RescueThreshold};  // as the forward, plus the adjoint's range
        for (size_t r = runs.size(); r-- > 0;) {
            const Run &q = runs[r];
            WorkSource w{order + q.begin, q.count, nullptr, counters + 48 + r};
            HIP_TRY(launch_grad(false, g, w, rs, std::min(kColClasses[q.cls] - 2 * kLdsPad - 32, d->max_hap_len),
                                q.shape, q.grid, s),
                    PAIRHMM_ERR_KERNEL);
        }
        if (!fixed) HIP_TRY(hipEventRecord(ev[1], s), PAIRHMM_ERR_KERNEL);
        g.scratch = scratch64;
        WorkSource again{rescue, 0, counters + 1, counters + 2};
        HIP_TRY(launch_grad(true, g, again, none, d->max_hap_len, sh64, grid64, s), PAIRHMM_ERR_KERNEL);
    }
    if (!fixed) HIP_TRY(hipEventRecord(ev[2], s), PAIRHMM_ERR_KERNEL);
    uint32_t host_counters[4] = {};
    HIP_TRY(hipMemcpyAsync(host_counters, counters, sizeof(host_counters), hipMemcpyDeviceToHost, s), PAIRHMM_ERR_KERNEL);
    HIP_TRY(hipStreamSynchronize(s), PAIRHMM_ERR_KERNEL);
    if (!fixed) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev[0], ev[2]) == hipSuccess) st.total_ms = ms;
        if (hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) st.fp32_kernel_ms = ms;
        if (hipEventElapsedTime(&ms, ev[1], ev[2]) == hipSuccess) st.rescue_kernel_ms = ms;
    }
    st.n_rescued = d->use_double ? d->n_pairs : host_counters[1];
    if (stats) *stats = st;
    return PAIRHMM_OK;
}

}  // namespace

extern "C" int32_t pairhmm_forward_backward(pairhmm_context *ctx, const pairhmm_grad_desc *d, pairhmm_run_stats *stats) {
    return forward_backward(ctx, d, stats, nullptr);
}

// ---------------------------------------------------------------------------------------------
// Forward + backward without allocation
// ---------------------------------------------------------------------------------------------

// The fixed scratch of pairhmm_forward_backward_noalloc: one array per device, held until the
// process ends.  In host code `scratch` names a host-side stand-in, never the device memory: the
// device address comes from hipGetSymbolAddress, once per device, and only that address is used.
static constexpr size_t kNoAllocScratchBytes = size_t(256) << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up on the first call there).
int32_t device_scratch(int device, uint8_t **out) {
    if (device < 0 || device >= kMaxDevices) return fail(PAIRHMM_ERR_STATE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return fail(PAIRHMM_ERR_KERNEL, std::string("scratch symbol lookup: ") + hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<uint8_t *>(p);
    }
    *out = g_scratch_ptr[device];
    return PAIRHMM_OK;
}

}  // namespace

extern "C" int32_t pairhmm_forward_backward_noalloc(pairhmm_context *ctx, const pairhmm_grad_desc *d,
                                                    pairhmm_run_stats *stats) {
    if (!ctx) return fail(PAIRHMM_ERR_STATE, "context is null");
    HIP_TRY(hipSetDevice(ctx->device), PAIRHMM_ERR_KERNEL);
    FixedScratch fx;
    const int32_t st = device_scratch(ctx->device, &fx.ptr);
    if (st != PAIRHMM_OK) return st;
    fx.bytes = kNoAllocScratchBytes;
    return forward_backward(ctx, d, stats, &fx);
}

// ---------------------------------------------------------------------------------------------
// Read-quality preparation
// ---------------------------------------------------------------------------------------------

extern "C" int32_t pairhmm_apply_pcr_error_model(pairhmm_context *ctx, const uint8_t *read_bases,
                                                 const int64_t *read_off, const int32_t *read_len, int64_t n_reads,
                                                 int32_t pcr_model, uint8_t *ins_quals, uint8_t *del_quals) {
    if (!ctx) return fail(PAIRHMM_ERR_STATE, "context is null");
    if (n_reads < 0) return fail(PAIRHMM_ERR_BAD_ARG, "negative read count");
    if (pcr_model < PAIRHMM_PCR_NONE || pcr_model > PAIRHMM_PCR_CONSERVATIVE)
        return fail(PAIRHMM_ERR_BAD_ARG, "unknown PCR error model");
    if (pcr_model == PAIRHMM_PCR_NONE || n_reads == 0) return PAIRHMM_OK;  // applyPCRErrorModel returns at once
    if (!read_bases || !read_off || !read_len || !ins_quals || !del_quals)
        return fail(PAIRHMM_ERR_BAD_ARG, "null device pointer");
    // initializePCRErrorModel (:339-353): the quality per repeat length, in Java's double arithmetic and
    // MathUtils.fastRound (MathUtils.java:428-430).
    PcrQualTable table{};
    const double rate = static_cast<double>(pcr_model);
    for (int k = 0; k <= 20; ++k) {
        const double x = 40.0 - std::exp(k / (rate * 3.14159265358979323846)) + 1.0;
        const int r = x > 0.0 ? static_cast<int>(x + 0.5) : static_cast<int>(x - 0.5);
        table.q[k] = static_cast<uint8_t>(std::max(10, r));
    }
    HIP_TRY(hipSetDevice(ctx->device), PAIRHMM_ERR_KERNEL);
    HIP_TRY(launch_pcr_error_model(read_bases, read_off, read_len, n_reads, table, ins_quals, del_quals, ctx->stream),
            PAIRHMM_ERR_KERNEL);
    HIP_TRY(hipStreamSynchronize(ctx->stream), PAIRHMM_ERR_KERNEL);
    return PAIRHMM_OK;
}

extern "C" int32_t pairhmm_cap_read_qualities(pairhmm_context *ctx, const int64_t *read_off, const int32_t *read_len,
                                              int64_t n_reads, const int32_t *mapping_quality,
                                              int32_t base_quality_threshold, int32_t disable_cap_to_mapq,
                                              uint8_t *base_quals, uint8_t *ins_quals, uint8_t *del_quals) {
    if (!ctx) return fail(PAIRHMM_ERR_STATE, "context is null");
    if (n_reads < 0) return fail(PAIRHMM_ERR_BAD_ARG, "negative read count");
    if (base_quality_threshold < 0 || base_quality_threshold > 255)
        return fail(PAIRHMM_ERR_BAD_ARG, "base_quality_threshold must be a byte");
    if (n_reads == 0) return PAIRHMM_OK;
    if (!read_off || !read_len || !base_quals || !ins_quals || !del_quals || (!disable_cap_to_mapq && !mapping_quality))
        return fail(PAIRHMM_ERR_BAD_ARG, "null device pointer");
    HIP_TRY(hipSetDevice(ctx->device), PAIRHMM_ERR_KERNEL);
    HIP_TRY(launch_cap_qualities(read_off, read_len, n_reads, mapping_quality, base_quality_threshold,
                                 disable_cap_to_mapq ? 1 : 0, base_quals, ins_quals, del_quals, ctx->stream),
            PAIRHMM_ERR_KERNEL);
    HIP_TRY(hipStreamSynchronize(ctx->stream), PAIRHMM_ERR_KERNEL);
    return PAIRHMM_OK;
}

extern "C" int32_t pairhmm_dragstr_impute_device(pairhmm_context *ctx, const double *gop_table, const double *gcp_table,
                                                 int32_t max_period, int32_t max_repeats, const uint8_t *read_bases,
                                                 const int64_t *read_off, const int32_t *read_len, int64_t n_reads,
                                                 int32_t max_read_len, uint8_t *out_gop, uint8_t *out_gcp) {
    if (!ctx) return fail(PAIRHMM_ERR_STATE, "context is null");
    if (!gop_table || !gcp_table) return fail(PAIRHMM_ERR_BAD_ARG, "null table");
    if (max_period < 1 || max_repeats < 1 || static_cast<int64_t>(max_period) * max_repeats > 512)
        return fail(PAIRHMM_ERR_BAD_ARG, "tables must have 1..512 cells");
    if (n_reads < 0 || max_read_len < 0) return fail(PAIRHMM_ERR_BAD_ARG, "negative sizes");
    if (max_read_len > dragstr_max_read_length())
        return fail(PAIRHMM_ERR_BAD_ARG, "reads longer than " + std::to_string(dragstr_max_read_length()) + " bases");
    DragstrByteTables t{};
    t.max_period = max_period;
    t.max_repeats = max_repeats;
    for (int k = 0; k < max_period * max_repeats; ++k) {
        if (!(gop_table[k] >= 0.0 && std::isfinite(gop_table[k]) && gcp_table[k] >= 0.0 && std::isfinite(gcp_table[k])))
            return fail(PAIRHMM_ERR_BAD_ARG, "tables must hold finite values >= 0");
        // GOP capped at 40 (DragstrPairHMMInputScoreImputator.java:31), both rounded as Math.round.
        t.gop[k] = static_cast<uint8_t>(static_cast<int>(std::floor(std::min(40.0, gop_table[k]) + 0.5)));
        t.gcp[k] = static_cast<uint8_t>(static_cast<int>(std::floor(gcp_table[k] + 0.5)));
    }
    if (n_reads == 0) return PAIRHMM_OK;
    if (!read_bases || !read_off || !read_len || !out_gop || !out_gcp)
        return fail(PAIRHMM_ERR_BAD_ARG, "null device pointer");
    HIP_TRY(hipSetDevice(ctx->device), PAIRHMM_ERR_KERNEL);
    HIP_TRY(launch_dragstr_impute(read_bases, read_off, read_len, n_reads, t, out_gop, out_gcp, ctx->stream),
            PAIRHMM_ERR_KERNEL);
    HIP_TRY(hipStreamSynchronize(ctx->stream), PAIRHMM_ERR_KERNEL);
    return PAIRHMM_OK;
}

extern "C" int32_t pairhmm_set_tuning(pairhmm_context *ctx, const pairhmm_tuning *t) {
    if (!ctx) return fail(PAIRHMM_ERR_STATE, "context is null");
    if (!t) return fail(PAIRHMM_ERR_BAD_ARG, "null tuning");
    const int wpb = t->waves_per_block == 0 ? 4 : t->waves_per_block;
    const int gp = t->grid_percent == 0 ? 100 : t->grid_percent;
    if (wpb < 1 || wpb > 8) return fail(PAIRHMM_ERR_BAD_ARG, "waves_per_block must be 1..8");
    if (gp < 1 || gp > 400) return fail(PAIRHMM_ERR_BAD_ARG, "grid_percent must be 1..400");
    if (t->queue_order != 0 && t->queue_order != 1) return fail(PAIRHMM_ERR_BAD_ARG, "queue_order must be 0 or 1");
    if (t->stripe_schedule != 0 && t->stripe_schedule != 1)
        return fail(PAIRHMM_ERR_BAD_ARG, "stripe_schedule must be 0 or 1");
    if (t->step_unroll != 0 && t->step_unroll != 8 && t->step_unroll != 16)
        return fail(PAIRHMM_ERR_BAD_ARG, "step_unroll must be 0, 8 or 16");
    ctx->waves_per_block = wpb;
    ctx->grid_percent = gp;
    ctx->queue_order = t->queue_order;
    ctx->stripe_schedule = t->stripe_schedule;
    ctx->step_unroll = t->step_unroll;
    return PAIRHMM_OK;
}

extern "C" int32_t pairhmm_get_tuning(pairhmm_context *ctx, pairhmm_tuning *t) {
    if (!ctx) return fail(PAIRHMM_ERR_STATE, "context is null");
    if (!t) return fail(PAIRHMM_ERR_BAD_ARG, "null tuning");
    t->waves_per_block = ctx->waves_per_block;
    t->grid_percent = ctx->grid_percent;
    t->queue_order = ctx->queue_order;
    t->stripe_schedule = ctx->stripe_schedule;
    t->step_unroll = ctx->step_unroll;
    return PAIRHMM_OK;
}
