// pairhmm_stream.cpp -- double-buffered streaming of region batches through the persistent kernels.
//
// New API (no counterpart in the reference): a caller with more regions than it wants resident at
// once (a whole genome's active regions) submits them in chunks.  Two slots alternate; each owns
// pinned host staging and device buffers.  While the device computes chunk k on the context's
// stream, the host validates and packs chunk k+1 into the other slot's pinned staging and its H2D
// copy runs on a separate copy stream; the compute stream waits on that copy with an event.  Results
// come back to pinned staging on the compute stream and are handed to the caller's array when the
// slot is reused or the stream finishes.
#include <chrono>

#include "capi_internal.hpp"

using namespace gatk_pairhmm;
using namespace gatk_pairhmm::capi;

namespace {

struct Slot {
    HostPinned hin;   // packed chunk (the device layout's input part)
    HostPinned hout;  // results + counters
    DevBuf dmem;      // device copy + outputs + rescue list + counters
    DevBuf f64;       // FP64 boundary planes of the chunk's FP64 launches
    hipEvent_t h2d_done = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    double *user_out = nullptr;
    int64_t n_pairs = 0;
};

}  // namespace

struct pairhmm_stream {
    pairhmm_context *ctx = nullptr;
    hipStream_t copy = nullptr;
    Slot slot[2];
    int next = 0;
    int64_t cells = 0;
    int64_t rescued = 0;
    int64_t routed = 0;  // pairs routed to FP64 by their read's qualities
    bool started = false;
    std::chrono::steady_clock::time_point t0;
};

namespace {

int32_t finalize(pairhmm_stream *s, Slot &sl) {
    if (!sl.busy) return PAIRHMM_OK;
    HIP_TRY(hipEventSynchronize(sl.done), PAIRHMM_ERR_KERNEL);
    const double *res = static_cast<const double *>(sl.hout.ptr);
    std::memcpy(sl.user_out, res, sizeof(double) * static_cast<size_t>(sl.n_pairs));
    const uint32_t *cnt = reinterpret_cast<const uint32_t *>(res + sl.n_pairs);
    s->rescued += rescued_total(cnt);
    sl.busy = false;
    sl.user_out = nullptr;
    return PAIRHMM_OK;
}

}  // namespace

extern "C" int32_t pairhmm_stream_create(pairhmm_context *ctx, pairhmm_stream **out) {
    if (!ctx) return fail(PAIRHMM_ERR_STATE, "context is null");
    if (!out) return fail(PAIRHMM_ERR_BAD_ARG, "null argument");
    *out = nullptr;
    HIP_TRY(hipSetDevice(ctx->device), PAIRHMM_ERR_KERNEL);
    pairhmm_stream *s = new (std::nothrow) pairhmm_stream();
    if (!s) return fail(PAIRHMM_ERR_OOM, "stream allocation failed");
    s->ctx = ctx;
    hipError_t e = hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking);
    for (Slot &sl : s->slot) {
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.h2d_done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        pairhmm_stream_destroy(s);
        return fail(PAIRHMM_ERR_KERNEL, std::string("stream create: ") + hipGetErrorString(e));
    }
    *out = s;
    return PAIRHMM_OK;
}

extern "C" int32_t pairhmm_stream_submit(pairhmm_stream *s, const pairhmm_batch_desc *d, double *out_host) {
    if (!s || !s->ctx) return fail(PAIRHMM_ERR_STATE, "stream is null");
    if (!d || !out_host) return fail(PAIRHMM_ERR_BAD_ARG, "null argument");
    if (d->n_pairs < 0) return fail(PAIRHMM_ERR_BAD_ARG, "bad batch sizes");
    if (d->n_pairs == 0) return PAIRHMM_OK;  // nothing to compute; no slot is used
    pairhmm_context *ctx = s->ctx;
    HIP_TRY(hipSetDevice(ctx->device), PAIRHMM_ERR_KERNEL);
    if (!s->started) {
        s->started = true;
        s->t0 = std::chrono::steady_clock::now();
    }
    Slot &sl = s->slot[s->next];
    s->next ^= 1;
    int32_t st = finalize(s, sl);
    if (st != PAIRHMM_OK) return st;
    std::vector<uint32_t> order;
    std::vector<QueueSlice> slices;
    int64_t cells = 0;
    int max_h = 1;
    const bool packed = d->hap_bases && use_packed(ctx, haps_packable(d->hap_bases, d->hap_bytes));
    st = validate_and_order(d, ctx->queue_order, order, cells, max_h, slices, true, packed);
    if (st != PAIRHMM_OK) return st;
    std::vector<uint32_t> counter_init, rescue_init;
    direct_rescue_init(slices, order, counter_init, rescue_init);
    const int64_t np = d->n_pairs;
    if (np == 0) return PAIRHMM_OK;
    // FP32 slices hold haplotypes of FP32 classes only (validate_and_order): the FP64 shape decides.
    const LaunchShape sh64 = forward_shape(true, max_h, ctx->waves_per_block);
    if (sh64.waves_per_block == 0)
        return fail(PAIRHMM_ERR_BAD_ARG, "haplotype of " + std::to_string(max_h) + " bases exceeds the LDS budget");

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
                 o_res = take(4 * np), o_cinit = take(4 * kCounterSlots);
    const size_t in_bytes = off;
    const size_t o_out = take(8 * np), o_cnt = take(4 * kCounterSlots);
    const size_t f64_bytes = fp64_slices_plane_bytes(ctx, slices);
    if (sl.hin.reserve(in_bytes) != hipSuccess || sl.hout.reserve(8 * np + 4 * kCounterSlots) != hipSuccess ||
        sl.dmem.reserve(off) != hipSuccess || sl.f64.reserve(f64_bytes) != hipSuccess)
        return fail(PAIRHMM_ERR_OOM, "stream slot allocation failed");
    uint8_t *h = static_cast<uint8_t *>(sl.hin.ptr);
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
    for (const Cp &c : copies)
        if (c.n) std::memcpy(h + c.o, c.src, c.n);
    uint8_t *m = static_cast<uint8_t *>(sl.dmem.ptr);
    HIP_TRY(hipMemcpyAsync(m, h, in_bytes, hipMemcpyHostToDevice, s->copy), PAIRHMM_ERR_KERNEL);
    HIP_TRY(hipEventRecord(sl.h2d_done, s->copy), PAIRHMM_ERR_KERNEL);

    DeviceBatch v{};
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
    uint32_t *ord = reinterpret_cast<uint32_t *>(m + o_ord);
    uint32_t *cnt = reinterpret_cast<uint32_t *>(m + o_cnt);
    uint32_t *res = reinterpret_cast<uint32_t *>(m + o_res);
    DeviceTables tab{ctx->d_qual_err, ctx->d_mm_prob};
    hipStream_t cs = ctx->stream;
    HIP_TRY(hipStreamWaitEvent(cs, sl.h2d_done, 0), PAIRHMM_ERR_KERNEL);
    HIP_TRY(launch_copy_counters(cnt, reinterpret_cast<const uint32_t *>(m + o_cinit), kCounterSlots, cs),
            PAIRHMM_ERR_KERNEL);
    HIP_TRY(launch_fp32_slices(ctx, v, tab, ord, slices, cnt, res, cs, packed), PAIRHMM_ERR_KERNEL);
    HIP_TRY(launch_fp64_slices(ctx, v, tab, res, slices, cnt, true, static_cast<double *>(sl.f64.ptr), sl.f64.cap, cs),
            PAIRHMM_ERR_KERNEL);
    double *ho = static_cast<double *>(sl.hout.ptr);
    HIP_TRY(hipMemcpyAsync(ho, v.out, 8 * np, hipMemcpyDeviceToHost, cs), PAIRHMM_ERR_KERNEL);
    HIP_TRY(hipMemcpyAsync(ho + np, cnt, 4 * kCounterSlots, hipMemcpyDeviceToHost, cs), PAIRHMM_ERR_KERNEL);
    HIP_TRY(hipEventRecord(sl.done, cs), PAIRHMM_ERR_KERNEL);
    sl.busy = true;
    sl.user_out = out_host;
    sl.n_pairs = np;
    s->cells += cells;
    for (const QueueSlice &q : slices) s->routed += q.n_direct;
    return PAIRHMM_OK;
}

extern "C" int32_t pairhmm_stream_finish(pairhmm_stream *s, pairhmm_run_stats *stats) {
    if (!s || !s->ctx) return fail(PAIRHMM_ERR_STATE, "stream is null");
    HIP_TRY(hipSetDevice(s->ctx->device), PAIRHMM_ERR_KERNEL);
    // Finalize in submission order: the slot after `next` holds the older chunk.
    for (int k = 0; k < 2; ++k) {
        int32_t st = finalize(s, s->slot[(s->next + k) & 1]);
        if (st != PAIRHMM_OK) return st;
    }
    pairhmm_run_stats r{};
    if (s->started)
        r.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - s->t0).count();
    r.cells = s->cells;
    r.n_rescued = s->rescued;
    r.n_routed_fp64 = s->routed;
    if (stats) *stats = r;
    s->started = false;
    s->cells = 0;
    s->rescued = 0;
    s->routed = 0;
    return PAIRHMM_OK;
}

extern "C" int32_t pairhmm_stream_destroy(pairhmm_stream *s) {
    if (!s) return PAIRHMM_OK;
    if (s->ctx) (void)hipSetDevice(s->ctx->device);
    for (Slot &sl : s->slot) {
        if (sl.busy && sl.done) (void)hipEventSynchronize(sl.done);
        if (sl.h2d_done) (void)hipEventDestroy(sl.h2d_done);
        if (sl.done) (void)hipEventDestroy(sl.done);
        sl.hin.release();
        sl.hout.release();
        sl.dmem.release();
        sl.f64.release();
    }
    if (s->copy) (void)hipStreamDestroy(s->copy);
    delete s;
    return PAIRHMM_OK;
}
