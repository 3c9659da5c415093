// capi_internal.hpp -- internals shared by the C-ABI translation units (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gatk_pairhmm.h"
#include "model_tables.hpp"
#include "pairhmm_kernels.hpp"

namespace gatk_pairhmm {
namespace capi {

inline thread_local std::string g_last_error;

inline int32_t fail(int32_t status, const std::string &msg) {
    g_last_error = msg;
    return status;
}

#define HIP_TRY(expr, status)                                                                           \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return fail((status), std::string(#expr) + ": " + hipGetErrorString(e_));                   \
    } while (0)

// Underflow threshold on the FP32 scaled last-row sum (scale 2^120): a pair whose sum is below it
// is recomputed in FP64.  Below 1e-28 the dominant paths' partial sums approach the FP32 normal
// range (1.2e-38) and lose relative precision.
constexpr double kFp32RescueThreshold = 1e-28;

inline bool device_is_gfx950(int dev) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return false;
    return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

// Capacity for a buffer of `cap` bytes asked to hold `n` > cap: at least n, and at least 1.5x the
// old capacity (that step capped at 256 MiB), so a caller whose calls grow a little at a time (one
// region after another) reallocates a few times rather than at every new maximum.
inline size_t grown_capacity(size_t cap, size_t n) {
    const size_t step = std::min<size_t>(cap / 2, size_t(256) << 20);
    return std::max<size_t>({n, size_t(4096), cap + step});
}

// Device buffer that only grows.
struct DevBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    hipError_t reserve(size_t n) {
        if (n <= cap) return hipSuccess;
        const size_t want = grown_capacity(cap, n);
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(&ptr, want);
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
};

struct HostPinned {
    void *ptr = nullptr;
    size_t cap = 0;
    hipError_t reserve(size_t n) {
        if (n <= cap) return hipSuccess;
        const size_t want = grown_capacity(cap, n);
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        cap = 0;
        hipError_t e = hipHostMalloc(&ptr, want, hipHostMallocDefault);
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release() {
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
};

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// The packed-stripe FP32 kernel reserves the haplotype bytes 0 (padding) and 1 (the row-0 marker);
// haplotype planes holding either run the aligned schedule.
inline bool haps_packable(const uint8_t *hap, int64_t n) {
    return n <= 0 || (std::memchr(hap, 0, static_cast<size_t>(n)) == nullptr &&
                      std::memchr(hap, 1, static_cast<size_t>(n)) == nullptr);
}

// PairHMM.computeReadLikelihoodGivenHaplotypeLog10 argument checks (PairHMM.java:284-292) and
// PairHMMModel.qualToTransProbs sign checks (PairHMMModel.java:108-111).
inline int32_t validate_read(const uint8_t *b, const uint8_t *q, const uint8_t *ins, const uint8_t *del, const uint8_t *gcp,
                      int64_t len, int64_t idx) {
    if (len < 0) return fail(PAIRHMM_ERR_BAD_ARG, "read " + std::to_string(idx) + " has negative length");
    if (len == 0) return PAIRHMM_OK;
    if (!b || !q || !ins || !del || !gcp)
        return fail(PAIRHMM_ERR_BAD_ARG, "read " + std::to_string(idx) + " has a null array");
    for (int64_t i = 0; i < len; ++i) {
        if (q[i] > kMaxQual)
            return fail(PAIRHMM_ERR_BAD_ARG, "read " + std::to_string(idx) + ": base quality 255 is out of range");
        if ((ins[i] | del[i] | gcp[i]) & 0x80)
            return fail(PAIRHMM_ERR_BAD_ARG,
                        "read " + std::to_string(idx) + ": insertion/deletion/GCP quality cannot be negative");
    }
    return PAIRHMM_OK;
}


// A contiguous run of the work queue launched as one persistent FP32 kernel with the LDS slice of
// its longest haplotype.  Queue order 0 groups pairs by LDS column class (largest first) and sorts
// each group by descending cost, so short haplotypes run at the occupancy their smaller slice
// allows; queue order 1 is one run in submission order.
struct QueueSlice {
    int64_t begin = 0, count = 0;
    int max_hap_len = 1;
    // The slice's last n_direct queue entries go to the FP64 kernel without an FP32 pass (pairs whose
    // read predicts an FP32 underflow); they also open the slice's rescue list.
    int64_t n_direct = 0;
};
// Device counters of a batch pass, per queue slice k: [k] FP32 queue head, [16 + k] pairs the slice
// sent to the FP64 rescue (its rescue list sits at the slice's own offset in the rescue array),
// [32 + k] FP64 queue head.
constexpr int kMaxSlices = 16;
constexpr int kCounterSlots = 4 * kMaxSlices;

inline uint32_t *slice_head(uint32_t *counters, size_t k) { return counters + k; }
inline uint32_t *slice_rescue_count(uint32_t *counters, size_t k) { return counters + kMaxSlices + k; }
inline uint32_t *slice_rescue_head(uint32_t *counters, size_t k) { return counters + 2 * kMaxSlices + k; }
inline uint32_t rescued_total(const uint32_t *host_counters) {
    uint32_t n = 0;
    for (int k = 0; k < kMaxSlices; ++k) n += host_counters[kMaxSlices + k];
    return n;
}

// Expected log10 emission of a read under its own base qualities, summed over its bases (per base
// (1-e) log10(1-e) + e log10(e/3)).  A pair whose read scores below kDirectFp64Score is sent to the
// FP64 kernel without an FP32 pass: its FP32 sum would fall under the rescue threshold (log10 L near
// -64) with near certainty, and the FP64 result is the one it would get after the rescue anyway.
constexpr double kDirectFp64Score = -90.0;

// Validates a host batch descriptor and builds its work queue and slices.  Fills cells (sum of
// R*H) and the longest haplotype.  With route_direct, pairs whose read scores below
// kDirectFp64Score are placed at the end of their slice (QueueSlice::n_direct).  Queue order 0
// sorts each slice by descending cost for the aligned schedule, and by descending haplotype length
// for the packed one (by_hap_len), whose stripes last as long as their pairs' longest haplotype.
int32_t validate_and_order(const pairhmm_batch_desc *d, int queue_order, std::vector<uint32_t> &order, int64_t &cells,
                           int &max_hap_len, std::vector<QueueSlice> &slices, bool route_direct = false,
                           bool by_hap_len = false);

// The packed schedule for this context and haplotype plane.
inline bool use_packed(const pairhmm_context *ctx, bool packable);

// True when a read's expected log10 emission along a perfect alignment,
// sum_i (1-e_i) log10(1-e_i) + e_i log10(e_i/3) over its base-quality error rates (each term rounded
// to 2^-40 and summed exactly in integers), is below kDirectFp64Score: its pairs would underflow the
// FP32 pass, so they go to the FP64 kernel directly.
bool read_routes_direct(const uint8_t *quals, int32_t len);

// Counter values at the start of a forward pass, and the initial rescue list, for slices built with
// route_direct: each slice's rescue count starts at its direct FP64 pairs, which sit at the head of
// its rescue list.
void direct_rescue_init(const std::vector<QueueSlice> &slices, const std::vector<uint32_t> &order,
                        std::vector<uint32_t> &counter_init, std::vector<uint32_t> &rescue_init);

}  // namespace capi
}  // namespace gatk_pairhmm

// The opaque context of the C-ABI.
struct pairhmm_context {
    int device = 0;
    int num_cus = 0;
    bool use_double = false;
    hipStream_t stream = nullptr;
    double *d_qual_err = nullptr;
    double *d_mm_prob = nullptr;
    gatk_pairhmm::capi::DevBuf stage_dev;    // packed inputs of one computeLikelihoods call + outputs
    gatk_pairhmm::capi::HostPinned stage_host;
    gatk_pairhmm::capi::DevBuf grad_scratch32, grad_scratch64, grad_work;  // forward_backward scratch
    gatk_pairhmm::capi::DevBuf f64_planes;   // FP64 boundary planes of computeLikelihoods calls
    int waves_per_block = 4;  // launch schedule (pairhmm_set_tuning)
    int grid_percent = 100;
    int queue_order = 0;
    int stripe_schedule = 0;  // 0 packed stripes, 1 aligned stripes (FP32 pass)
    int step_unroll = 0;      // packed kernel's loop body: 8 or 16 wave-steps, 0 per LDS class
};

namespace gatk_pairhmm {
namespace capi {

// Persistent grid for a launch shape: resident blocks x CUs x the tuning percentage, and no more
// blocks than the queue has work for.
int persistent_grid(const pairhmm_context *ctx, const LaunchShape &shape, int64_t n_items);

// Enqueues the FP32 forward kernel once per queue slice on `s` (counters must be zeroed first);
// pairs below the rescue threshold are appended to their slice's part of rescue_list (the slice's
// offset and capacity in the queue).
hipError_t launch_fp32_slices(const pairhmm_context *ctx, const DeviceBatch &batch, const DeviceTables &tab,
                              const uint32_t *order, const std::vector<QueueSlice> &slices, uint32_t *counters,
                              uint32_t *rescue_list, hipStream_t s, bool packed);

inline bool use_packed(const pairhmm_context *ctx, bool packable) { return packable && ctx->stripe_schedule == 0; }

// Enqueues the FP64 forward kernel once per slice, each with its own slice's LDS class: over the
// slice's rescue list (rescue_pass, the count read on the device), or over all of its queue.  The
// waves' boundary planes come from `planes` (plane_bytes); a launch's grid is held to what fits.
hipError_t launch_fp64_slices(const pairhmm_context *ctx, const DeviceBatch &batch, const DeviceTables &tab,
                              const uint32_t *list, const std::vector<QueueSlice> &slices, uint32_t *counters,
                              bool rescue_pass, double *planes, size_t plane_bytes, hipStream_t s);

// Bytes of FP64 boundary planes that launch_fp64_slices uses at its full grids.
size_t fp64_slices_plane_bytes(const pairhmm_context *ctx, const std::vector<QueueSlice> &slices);

}  // namespace capi
}  // namespace gatk_pairhmm
