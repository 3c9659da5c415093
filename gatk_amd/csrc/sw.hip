// sw.hip -- gfx950 Smith-Waterman (read -> haplotype and haplotype -> reference realignment).
//
// Model: SmithWatermanJavaAligner.calculateMatrix / calculateCigar
// (src/main/java/org/broadinstitute/hellbender/utils/smithwaterman/SmithWatermanJavaAligner.java:
// 104-377), the algorithm behind GKL's SWAlignerNativeBinding.  Integer scores; rows are reference
// positions i, columns alternate positions j; affine gaps in the reference's "best gap" form:
//   V(i,j) = sw(i-1,j) + open  if that beats V(i-1,j) + extend  (then its length is 1, else +1)
//   H(i,j) = sw(i,j-1) + open  if that beats H(i,j-1) + extend
//   sw(i,j) = max(-1e8, diag if diag >= V and diag >= H, else H if H >= V, else V)
// with the backtrack code 0 (diagonal), -len(H) or +len(V) per cell (:160-229).
//
// Mapping: one pair per wave; stripes of 64 alternate columns, one column per lane; lane k owns
// column 64 s + k + 1 and at step t computes row i = t - k + 1 (anti-diagonal sweep over the
// reference).  A lane keeps its column's vertical gap state and its previous row; the row's
// horizontal gap state, the cell to the left and the reference base move lane to lane by DPP
// wave_shr:1, and the cell to the left one step earlier is the diagonal.  Lane 0 takes them from a
// 12-byte LDS record per row that lane 63 of the previous stripe wrote (stripe 0: column 0).  Backtrack codes
// (int16) go to the wave's global scratch in step-major order, 128 coalesced bytes per step; the
// last column and the bottom row of scores go there too.  The wave then picks the end point and walks
// the backtrack as calculateCigar does, 64 diagonal cells per fetch, writing the CIGAR elements in
// the reference's list order (the host reverses them into text).
#include "pairhmm_device.hpp"
#include "pairhmm_kernels.hpp"

namespace gatk_pairhmm {
namespace {

using namespace dev;

// A row's boundary record, 12 bytes: the score and horizontal gap of the stripe's last column, and
// the gap length with the row's reference base (gsh | base << 16; gap lengths stay below 32768).
struct alignas(8) SwScore {
    int32_t sw, bgh;
};

constexpr int kSwPad = 66;
constexpr int32_t kLowInit = INT32_MIN / 2;          // :125
constexpr int32_t kMinCutoff = -100000000;           // (int) -1.0e8, :123
enum : uint32_t { kOpM = 0, kOpI = 1, kOpD = 2, kOpS = 3 };

__device__ __forceinline__ int32_t shr1_keep_i(int32_t old, int32_t src) {
    return __builtin_amdgcn_update_dpp(old, src, kDppWaveShr1, 0xF, 0xF, false);
}

// Score of row 0 / column 0 at position p >= 1 (INDEL and LEADING_INDEL start with gap penalties,
// :133-149), 0 otherwise.
__device__ __forceinline__ int32_t edge_score(int p, bool gap_edges, int32_t open, int32_t extend) {
    return (gap_edges && p >= 1) ? open + (p - 1) * extend : 0;
}

__global__ __launch_bounds__(kWave) void sw_kernel(SwBatch b, uint32_t n_pairs) {
    extern __shared__ SwScore recs[];
    uint32_t *gbase = nullptr;  // set per pair: (R + kSwPad) words after the score records
    const int lane = static_cast<int>(threadIdx.x);
    const int32_t w_match = b.match, w_mismatch = b.mismatch, w_open = b.open, w_extend = b.extend;
    const bool gap_edges = b.strategy == 1 || b.strategy == 2;
    int16_t *bt = reinterpret_cast<int16_t *>(b.scratch + static_cast<size_t>(blockIdx.x) * b.scratch_per_wave);
    int32_t *lastcol = reinterpret_cast<int32_t *>(b.scratch + static_cast<size_t>(blockIdx.x) * b.scratch_per_wave +
                                                   b.bt_bytes_per_wave);
    int32_t *bottom = lastcol + (b.max_ref_len + 1);
    for (uint32_t p = blockIdx.x; p < n_pairs; p += gridDim.x) {
        const int R = b.ref_len[p], A = b.alt_len[p];
        const uint8_t *ref = b.ref_bases + b.ref_off[p];
        const uint8_t *alt = b.alt_bases + b.alt_off[p];
        const int steps = R + kWave - 1;
        const int n_stripes = (A + kWave - 1) / kWave;
        __builtin_amdgcn_wave_barrier();
        gbase = reinterpret_cast<uint32_t *>(recs + (R + kSwPad));
        for (int i = lane; i < R + kSwPad; i += kWave) {
            SwScore v;
            v.sw = i <= R ? edge_score(i, gap_edges, w_open, w_extend) : 0;
            v.bgh = kLowInit;
            recs[i] = v;
            gbase[i] = (i >= 1 && i <= R) ? static_cast<uint32_t>(ref[i - 1]) << 16 : 0u;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int s = 0; s < n_stripes; ++s) {
            const int j = s * kWave + lane + 1;  // this lane's column
            const bool col_ok = j <= A;
            const uint32_t bj = col_ok ? alt[j - 1] : 0xFFFFu;
            // Column state: previous row's score (row 0's edge value to start) and the vertical gap.
            int32_t up = edge_score(j, gap_edges, w_open, w_extend), bgv = kLowInit, gsv = 0;
            // Row state handed lane to lane: score to the left, horizontal gap, reference base.
            int32_t cur = up, bgh = kLowInit, gsh = 0;
            uint32_t a = 0;
            // Diagonal: lane 0 starts from row 0 of the previous stripe's last column.
            int32_t diag = lane == 0 ? edge_score(s * kWave, gap_edges, w_open, w_extend) : 0;
            SwScore nxt = recs[1];
            uint32_t nxt_g = gbase[1];
            int16_t *bts = bt + static_cast<size_t>(s) * steps * kWave;
            for (int t = 0; t < steps; ++t) {
                const SwScore rc = nxt;
                const uint32_t rg = nxt_g;
                nxt = recs[t + 2];
                nxt_g = gbase[t + 2];
                const int i = t - lane + 1;
                // Values of column j-1 at row i (computed one step earlier by the lane to the left).
                const int32_t left = shr1_keep_i(rc.sw, cur);
                const int32_t l_bgh = shr1_keep_i(rc.bgh, bgh);
                const int32_t l_gsh = shr1_keep_i(static_cast<int32_t>(rg & 0xFFFFu), gsh);
                a = static_cast<uint32_t>(shr1_keep_i(static_cast<int32_t>(rg >> 16), static_cast<int32_t>(a)));
                const bool active = col_ok && i >= 1 && i <= R;
                const int32_t step_diag = diag + (a == bj ? w_match : w_mismatch);
                int32_t prev_gap = up + w_open;
                int32_t n_bgv = bgv + w_extend, n_gsv;
                if (prev_gap > n_bgv) {
                    n_bgv = prev_gap;
                    n_gsv = 1;
                } else {
                    n_gsv = gsv + 1;
                }
                prev_gap = left + w_open;
                int32_t n_bgh = l_bgh + w_extend, n_gsh;
                if (prev_gap > n_bgh) {
                    n_bgh = prev_gap;
                    n_gsh = 1;
                } else {
                    n_gsh = l_gsh + 1;
                }
                int32_t v;
                int16_t code;
                if (step_diag >= n_bgv && step_diag >= n_bgh) {
                    v = step_diag;
                    code = 0;
                } else if (n_bgh >= n_bgv) {
                    v = n_bgh;
                    code = static_cast<int16_t>(-n_gsh);
                } else {
                    v = n_bgv;
                    code = static_cast<int16_t>(n_gsv);
                }
                v = v > kMinCutoff ? v : kMinCutoff;
                diag = left;
                if (active) {
                    up = v;
                    bgv = n_bgv;
                    gsv = n_gsv;
                    cur = v;
                    bgh = n_bgh;
                    gsh = n_gsh;
                    bts[static_cast<size_t>(t) * kWave + lane] = code;
                    if (j == A) lastcol[i] = v;
                    if (i == R) bottom[j] = v;
                } else {
                    // Outside the matrix the lane hands on what it holds: its column's row-0 value
                    // before it starts, its last row after it ends (never read by a real cell).
                    bgh = l_bgh;
                    gsh = l_gsh;
                }
                if (lane == kWave - 1 && col_ok && i >= 1 && i <= R) {
                    recs[i].sw = v;
                    recs[i].bgh = n_bgh;
                    gbase[i] = (gbase[i] & 0xFFFF0000u) | static_cast<uint32_t>(n_gsh);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        // Lane 0 reads what every lane stored to global memory: wait for the stores, and drop the
        // vector L1 lines an earlier pair's traceback may hold for this scratch.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        // calculateCigar (:256-377), run by the whole wave on wave-uniform state: lanes fetch 64 end
        // scores or 64 backtrack codes at a time, and every lane replays the reference's sequential
        // decisions on them; lane 0 stores.
        int p1 = 0, p2 = 0, segment = 0;
        if (b.strategy == 1) {
            p1 = R;
            p2 = A;
        } else {
            int32_t best = INT32_MIN;
            p2 = A;
            for (int i0 = 1; i0 <= R; i0 += kWave) {
                const int32_t mine = i0 + lane <= R ? lastcol[i0 + lane] : INT32_MIN;
                const int cnt = R - i0 + 1 < kWave ? R - i0 + 1 : kWave;
                for (int k = 0; k < cnt; ++k) {
                    const int32_t sc = __shfl(mine, k, kWave);
                    if (sc >= best) {
                        p1 = i0 + k;
                        best = sc;
                    }
                }
            }
            if (b.strategy != 2) {
                for (int j0 = 1; j0 <= A; j0 += kWave) {
                    const int32_t mine = j0 + lane <= A ? bottom[j0 + lane] : INT32_MIN;
                    const int cnt = A - j0 + 1 < kWave ? A - j0 + 1 : kWave;
                    for (int k = 0; k < cnt; ++k) {
                        const int32_t sc = __shfl(mine, k, kWave);
                        const int jj = j0 + k;
                        const int d_new = R - jj < 0 ? jj - R : R - jj;
                        const int d_old = p1 - p2 < 0 ? p2 - p1 : p1 - p2;
                        if (sc > best || (sc == best && d_new < d_old)) {
                            p1 = R;
                            p2 = jj;
                            best = sc;
                            segment = A - jj;
                        }
                    }
                }
            }
        }
        uint32_t *elems = b.elems + b.elem_off[p];
        uint32_t *elems_short = b.elems_short + static_cast<size_t>(p) * kSwShortElems;
        int n = 0;
        auto put = [&](uint32_t op, int32_t len) {
            const uint32_t e = (static_cast<uint32_t>(len) << 2) | op;
            if (lane == 0) {
                elems[n] = e;
                if (n < kSwShortElems) elems_short[n] = e;
            }
            ++n;
        };
        if (segment > 0 && b.strategy == 0) {
            put(kOpS, segment);
            segment = 0;
        }
        uint32_t state = kOpM;
        do {
            // The next 64 cells down the diagonal from (p1, p2): a run of diagonal codes is a run of
            // single MATCH steps, taken at once; the first other code is one gap step.
            const int ci = p1 - lane, cj = p2 - lane;
            int code = 1;  // outside the matrix: stop
            if (ci >= 1 && cj >= 1) {
                const int ss = (cj - 1) / kWave, ll = (cj - 1) % kWave;
                code = bt[(static_cast<size_t>(ss) * steps + (ci - 1 + ll)) * kWave + ll];
            }
            const uint64_t stop = __ballot(code != 0);
            const int f = stop ? __builtin_ctzll(stop) : kWave;
            if (f > 0) {
                if (state == kOpM) {
                    segment += f;
                } else {
                    if (segment > 0) put(state, segment);
                    segment = f;
                    state = kOpM;
                }
                p1 -= f;
                p2 -= f;
            }
            if (f < kWave && p1 > 0 && p2 > 0) {
                const int btr = __shfl(code, f, kWave);
                const uint32_t ns = btr > 0 ? kOpD : kOpI;
                const int step = btr > 0 ? btr : -btr;
                if (ns == kOpI)
                    p2 -= step;
                else
                    p1 -= step;
                if (ns == state) {
                    segment += step;
                } else {
                    if (segment > 0) put(state, segment);
                    segment = step;
                    state = ns;
                }
            }
        } while (p1 > 0 && p2 > 0);
        int offset;
        if (b.strategy == 0) {
            put(state, segment);
            if (p2 > 0) put(kOpS, p2);
            offset = p1;
        } else if (b.strategy == 3) {
            put(state, segment + p2);
            offset = p1 - p2;
        } else {
            put(state, segment);
            if (p1 > 0)
                put(kOpD, p1);
            else if (p2 > 0)
                put(kOpI, p2);
            offset = 0;
        }
        if (lane == 0) {
            b.n_elems[p] = n;
            b.offset[p] = offset;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

}  // namespace

size_t sw_lds_bytes(int max_ref_len) {
    return static_cast<size_t>(max_ref_len + kSwPad) * (sizeof(SwScore) + sizeof(uint32_t));
}

size_t sw_bt_bytes_per_wave(int max_ref_len, int max_alt_len) {
    const size_t stripes = static_cast<size_t>((max_alt_len + kWave - 1) / kWave);
    return (stripes * static_cast<size_t>(max_ref_len + kWave - 1) * kWave * sizeof(int16_t) + 255) / 256 * 256;
}

size_t sw_scratch_bytes_per_wave(int max_ref_len, int max_alt_len) {
    const size_t tail = static_cast<size_t>(max_ref_len + 1 + max_alt_len + 1) * sizeof(int32_t);
    return (sw_bt_bytes_per_wave(max_ref_len, max_alt_len) + tail + 255) / 256 * 256;
}

hipError_t launch_sw(const SwBatch &batch, uint32_t n_pairs, int grid_blocks, hipStream_t stream) {
    if (n_pairs == 0) return hipSuccess;
    const size_t lds = sw_lds_bytes(batch.max_ref_len);
    if (lds > max_lds_per_block()) return hipErrorInvalidValue;
    if (lds > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(sw_kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(lds));
    hipLaunchKernelGGL(sw_kernel, dim3(grid_blocks), dim3(kWave), lds, stream, batch, n_pairs);
    return hipGetLastError();
}

}  // namespace gatk_pairhmm
