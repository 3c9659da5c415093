// Host stand-ins for the kernel-side entry points of the engine (pairhmm_kernels.hpp), for the CPU
// stream harness only.  The forward "kernel" walks its work source exactly as the persistent kernel
// does (queue items from the head, item count from the device counter when given) and writes a
// value the harness can recompute for every pair:
//     out[pair] = -(R * 1000 + H) * 1e-6,  R, H = the pair's read and haplotype lengths,
// after touching the last byte of every FP64 wave's boundary-plane slice, so AddressSanitizer checks
// every buffer size the host code computed.  Entry points the stream path never calls fail.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "../../gatk_amd/csrc/pairhmm_kernels.hpp"

namespace gatk_pairhmm {

int lds_columns(int max_hap_len) {
    const int need = max_hap_len + 2 * kLdsPad + 32;
    for (int c : kColClasses)
        if (c >= need) return c;
    return 0;
}

int lds_class_index(int hap_len) {
    const int need = hap_len + 2 * kLdsPad + 32;
    for (int k = 0; k < kNumColClasses; ++k)
        if (kColClasses[k] >= need) return k;
    return -1;
}

size_t lds_bytes_per_wave(int max_hap_len, bool fp64, bool packed) {
    const size_t cols = static_cast<size_t>(lds_columns(max_hap_len));
    return (packed && !fp64 ? 16 + 2 * cols : cols) +
           (fp64 ? 0 : 2 * sizeof(float) * static_cast<size_t>(lds_plane_stride(static_cast<int>(cols))));
}

size_t fp64_plane_bytes_per_wave(int max_hap_len) {
    const int cols = lds_columns(max_hap_len);
    return cols == 0 ? 0 : 2 * sizeof(double) * static_cast<size_t>(lds_plane_stride(cols));
}

size_t max_lds_per_block() { return 163840; }

LaunchShape forward_shape(bool fp64, int max_hap_len, int requested_waves_per_block, bool packed, int unroll) {
    LaunchShape sh;
    if (lds_columns(max_hap_len) == 0) return sh;
    sh.packed = packed && !fp64;
    sh.unroll = unroll == 16 ? 16 : 8;
    const size_t per_wave = lds_bytes_per_wave(max_hap_len, fp64, sh.packed);
    int wpb = requested_waves_per_block < 1 ? 1 : requested_waves_per_block;
    while (wpb > 1 && per_wave * static_cast<size_t>(wpb) > max_lds_per_block()) --wpb;
    sh.waves_per_block = wpb;
    sh.lds_bytes = per_wave * static_cast<size_t>(wpb);
    sh.blocks_per_cu = 2;
    return sh;
}

static void die(const char *what) {
    std::fprintf(stderr, "stream harness stub: %s\n", what);
    std::abort();
}

hipError_t launch_forward(bool fp64, const DeviceBatch &b, const DeviceTables &, const WorkSource &ws,
                          const RescueSink &, int max_hap_len, const LaunchShape &shape, int grid_blocks, hipStream_t) {
    if (shape.waves_per_block < 1 || grid_blocks < 1) return hipErrorInvalidValue;
    if (shape.packed && (fp64 || ws.head == nullptr)) die("packed launch without a queue head");
    if (fp64) {
        const size_t waves = static_cast<size_t>(grid_blocks) * shape.waves_per_block;
        if (!b.f64_planes || b.f64_plane_waves < waves) die("FP64 grid exceeds its boundary planes");
        const size_t per_wave = fp64_plane_bytes_per_wave(max_hap_len);
        volatile uint8_t *p = reinterpret_cast<volatile uint8_t *>(b.f64_planes);
        for (size_t w = 0; w < waves; ++w) p[(w + 1) * per_wave - 1] = 0;
    }
    const uint32_t n = ws.n_items_dev ? *ws.n_items_dev : ws.n_items;
    uint32_t it = ws.head ? *ws.head : 0;
    for (; it < n; ++it) {
        const uint32_t pair = ws.items ? ws.items[it] : it;
        int r, h;
        if (b.pair_read) {
            r = b.pair_read[pair];
            h = b.pair_hap[pair];
        } else {
            r = static_cast<int>(pair / static_cast<uint32_t>(b.n_haps));
            h = static_cast<int>(pair % static_cast<uint32_t>(b.n_haps));
        }
        const int H = b.hap_len[h], R = b.rd_len[r];
        if (H > max_hap_len) die("pair launched in an LDS class too small for its haplotype");
        volatile const uint8_t *hb = b.hap_bases + b.hap_off[h];
        (void)hb[H - 1];
        if (R > 0) {
            volatile const uint8_t *q = b.rd_gcp + b.rd_off[r];
            (void)q[R - 1];
        }
        b.out[pair] = -(static_cast<double>(R) * 1000.0 + H) * 1e-6;
    }
    if (ws.head) *ws.head = it;
    return hipSuccess;
}

hipError_t launch_zero_counters(uint32_t *c, int n, hipStream_t) {
    for (int i = 0; i < n; ++i) c[i] = 0;
    return hipSuccess;
}

hipError_t launch_copy_counters(uint32_t *c, const uint32_t *init, int n, hipStream_t) {
    for (int i = 0; i < n; ++i) c[i] = init[i];
    return hipSuccess;
}

// Not on the stream path.
size_t grad_lds_bytes_per_wave(int, bool) { return 0; }
size_t grad_scratch_elems_per_wave(int, int, int *, int *) { return 0; }
LaunchShape grad_shape(bool, int, int) { return LaunchShape{}; }
hipError_t launch_grad(bool, const GradBatch &, const WorkSource &, const RescueSink &, int, const LaunchShape &, int,
                       hipStream_t) {
    return hipErrorNotSupported;
}
hipError_t launch_grad_class_count(const int32_t *, const int32_t *, uint32_t, int, uint32_t *, hipStream_t) {
    return hipErrorNotSupported;
}
hipError_t launch_grad_class_scatter(const int32_t *, const int32_t *, uint32_t, int, const ClassRuns &, uint32_t *,
                                     uint32_t *, hipStream_t) {
    return hipErrorNotSupported;
}
hipError_t launch_pcr_error_model(const uint8_t *, const int64_t *, const int32_t *, int64_t, const PcrQualTable &,
                                  uint8_t *, uint8_t *, hipStream_t) {
    return hipErrorNotSupported;
}
hipError_t launch_cap_qualities(const int64_t *, const int32_t *, int64_t, const int32_t *, int32_t, int32_t, uint8_t *,
                                uint8_t *, uint8_t *, hipStream_t) {
    return hipErrorNotSupported;
}
hipError_t launch_dragstr_impute(const uint8_t *, const int64_t *, const int32_t *, int64_t, const DragstrByteTables &,
                                 uint8_t *, uint8_t *, hipStream_t) {
    return hipErrorNotSupported;
}
int dragstr_max_read_length() { return 0; }

}  // namespace gatk_pairhmm
