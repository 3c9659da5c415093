// pdhmm.hip -- gfx950 FP64 kernel of the partially determined PairHMM.
//
// Model: LoglessPDPairHMM.subComputeReadLikelihoodGivenHaplotypeLog10
// (src/main/java/org/broadinstitute/hellbender/utils/pairhmm/LoglessPDPairHMM.java:34-150).  Next to
// M, I, D each row carries branch values bM, bI, bD (:15-17): the main values just before an
// optional deletion span, held across the span and merged back with max() after it.  A per-row
// state machine (NORMAL / INSIDE_DEL / AFTER_DEL, :22-31, 63, 129-136) picks the branch update; SNP
// columns accept a set of alternative bases (isBasePDMatching, :175-194).
//
// Restated for a lane-per-row sweep.  Let E(r,c) be the value a cell hands to its right and lower
// neighbours: E = max(b, X) for X in {M, I, D} when haplotype column c carries DEL_END, else E = X.
// A column is in state AFTER_DEL exactly when the column before it carries DEL_END, and the branch
// and the I rule read column c's own DEL_END bit, so the reference's three cases collapse to
//   M(r,c) = prior (E_M(r-1,c-1) MM + E_I(r-1,c-1) GM + E_D(r-1,c-1) GM)
//   I(r,c) = E_M(r-1,c) MI + E_I(r-1,c) II
//   D(r,c) = E_M(r,c-1) MD + E_D(r,c-1) DD
//   b(r,c) = INSIDE_DEL(r,c) ? b(r,c-1) : E(r,c-1)
// (in NORMAL and INSIDE_DEL the column before carries no DEL_END, so E = X there).  Only
// INSIDE_DEL depends on the row: the state variable is declared outside the row loop (:63), so row 1
// starts NORMAL and every later row starts in the state the previous row ended in, which is the
// same for all rows >= 2.  The host folds this into each column's code (pd_column_codes).
//
// Mapping: one pair per wave, stripes of 64 rows, lane k owns row 64 s + k + 1 and at step t
// computes column c = t - k + 1.  Each lane forms, from its new E values and the transitions of the
// row below (the next lane's row; for lane 63 the next stripe's first row), the two sums that row
// needs:
//   A(r+1, c+1) = E_M(r,c) MM' + (E_I(r,c) + E_D(r,c)) GM'    (so M(r+1,c+1) = prior' A)
//   B(r+1, c)   = E_M(r,c) MI' + E_I(r,c) II'                 (so I(r+1,c) = B)
// and the row below takes A two steps later and B one step later by DPP wave_shr:1: two doubles
// cross lanes per step instead of the three E values.  Lane 0 reads its A and B from an LDS record
// per column that lane 63 of the previous stripe wrote (stripe 0 starts from row 0: M = I = 0,
// D = 2^1020 / H, so A(1,c) = D GM_1 and B(1,c) = 0); every lane reads its own column's code from an
// LDS array.  Records are 16 bytes (A, B) plus the 4-byte code, so a wave holds 20 bytes of LDS per
// column.  Steps run in blocks of 4; a block in which no lane is on a special column (the host's
// kPdSpecial: a deletion end, a column inside a span, or the column before a span) skips the branch
// update and hands on E = X, which is exact because a row's branch values are only read on special
// columns and the column before each span sets them from E.  The row-R lane sums M + I over
// columns 1..H in column order, as the reference does (:132-139), in the last stripe only.
#include "pairhmm_device.hpp"
#include "pairhmm_kernels.hpp"

namespace gatk_pairhmm {
namespace {

using namespace dev;

// One column of a wave's LDS records: A and B of the stripe's first row (the row below lane 63 of
// the stripe before).  The column codes are a separate uint32 array after the records.
struct alignas(16) PdRecord {
    double a, b;
};
static_assert(sizeof(PdRecord) == 16, "PdRecord is one 16-byte LDS word");

constexpr int kPdPadColumns = 68;  // records past H that lane 0 reads (c <= H + 66 with 4-step blocks)
constexpr int kPdLowPad = 64;      // records below column 0, where lane 63's stores of columns < 0 land
constexpr int kPdU = 4;            // wave-steps per unrolled block

// Read-base class bit (placed like the column code's alternative mask) of isBasePDMatching
// (:179-195): A/a, C/c, G/g, T/t, and 0x80 for any other base, which only an 'N' haplotype
// base's full mask matches.
__device__ __forceinline__ uint32_t read_base_bit(uint32_t x) {
    switch (x) {
    case 'A': case 'a': return kPdAltA << 8;
    case 'C': case 'c': return kPdAltC << 8;
    case 'G': case 'g': return kPdAltG << 8;
    case 'T': case 't': return kPdAltT << 8;
    default: return 0x80u << 8;
    }
}

// Row parameters of read position p (row p + 1); rows past the read are zero.
struct PdRow {
    double MM, GM, MI, II, MD, DD, pm, px;
    uint32_t x, xbit;
};

__device__ __forceinline__ PdRow pd_row(const PdBatch &b, const DeviceTables &tab, int64_t ro, int R, int p) {
    PdRow w{0, 0, 0, 0, 0, 0, 0, 0, 0u, 0u};
    if (p < 0 || p >= R) return w;
    w.x = b.rd_bases[ro + p];
    const int q = b.rd_quals[ro + p], qi = b.rd_ins[ro + p], qd = b.rd_del[ro + p], qc = b.rd_gcp[ro + p];
    const int mn = qi < qd ? qi : qd, mx = qi < qd ? qd : qi;
    w.MM = tab.mm_prob[((mx * (mx + 1)) >> 1) + mn];
    w.MI = tab.qual_err[qi];
    w.MD = tab.qual_err[qd];
    w.II = tab.qual_err[qc];
    w.DD = w.II;
    w.GM = 1.0 - w.II;
    const double eq = tab.qual_err[q];
    w.pm = 1.0 - eq;
    w.px = b.tristate ? eq / 3.0 : eq;
    if (w.x == 'N') w.px = w.pm;
    w.xbit = read_base_bit(w.x);
    return w;
}

// Lane state of a stripe: own E and branch values at column c-1, the row-below sums of the last
// two steps, the last-row sum.  The D values (ED, bD) are carried divided by the row's MD:
// D(c) = E_D(c-1) DD + E_M(c-1) MD becomes D/MD = (E_D/MD) DD + E_M, one FMA, and max() commutes with
// the positive scale (rows past R have MD = DD = 0 and stay 0).
struct PdLane {
    double EM = 0, EI = 0, ED = 0, bM = 0, bI = 0, bD = 0;
    double an1 = 0, an2 = 0, bn1 = 0, acc = 0;
};

// One step of a lane at column c.  kBranch: the full branch update; without it E = X and the branch
// values are left alone, which is exact on a block where no lane's column is special: a row's branch
// values are only read at its special columns, and a deletion start (special) sets them from E.
template <bool kLast, bool kBranch>
__device__ __forceinline__ void pd_step(PdLane &st, const PdRecord &cur, uint32_t code, int c, int H, int R, int row,
                                        const PdRow &w, const PdRow &wn, uint32_t inside_bit, bool writer,
                                        PdRecord *wr) {
    const double a = wave_shr1(cur.a, st.an2);  // lane 0 keeps its record's values
    const double I = wave_shr1(cur.b, st.bn1);
    const bool match = ((code & 0xFFu) == w.x) || (code & w.xbit) != 0u;
    const double M = (match ? w.pm : w.px) * a;
    const double D = fma(st.ED, w.DD, st.EM);  // D / MD
    if constexpr (kBranch) {
        const bool inside = (code & inside_bit) != 0u;
        st.bM = inside ? st.bM : st.EM;
        st.bI = inside ? st.bI : st.EI;
        st.bD = inside ? st.bD : st.ED;
        const double f = (code & kPdDelEnd) ? 1.0 : 0.0;  // max(0, X) == X for the cells here
        st.EM = fmax(st.bM * f, M);
        st.EI = fmax(st.bI * f, I);
        st.ED = fmax(st.bD * f, D);
    } else {
        st.EM = M;
        st.EI = I;
        st.ED = D;
    }
    const double an = fma(fma(st.ED, w.MD, st.EI), wn.GM, st.EM * wn.MM);
    const double bn = fma(st.EI, wn.II, st.EM * wn.MI);
    st.an2 = st.an1;
    st.an1 = an;
    st.bn1 = bn;
    if (writer) {  // (B(c), A(c + 1)) of the row below; c >= -62 (low pad)
        wr->b = bn;
        wr[1].a = an;
    }
    if (kLast && row == R && c >= 1 && c <= H) st.acc += M + I;
}

// Steps of one stripe, in blocks of kPdU; kLast adds the row-R lane's sum of M + I.  Lane 0's
// records of a block are read at its start (lane 63 writes record j at step j + 61 or later); every
// lane reads its own columns' codes (cds[c], zero for c < 1 and c > H).  A block runs the branch
// update only if some lane's column in it is special.
template <bool kLast>
__device__ __forceinline__ double pd_stripe(PdRecord *rec, const uint32_t *cds, int lane, int H, int R, int row,
                                            const PdRow &w, const PdRow &wn, uint32_t inside_bit) {
    PdLane st;
    const bool writer = lane == kWave - 1;
    const int steps = round_up(H + kWave - 1, kPdU);
    const uint32_t *cp = cds + 1 - lane;  // column t - lane + 1 at step t
    for (int t0 = 0; t0 < steps; t0 += kPdU) {
        PdRecord cur[kPdU];
        uint32_t cc[kPdU];
        uint32_t any = 0;
#pragma unroll
        for (int u = 0; u < kPdU; ++u) {
            cur[u] = rec[t0 + u + 1];
            cc[u] = cp[t0 + u];
            any |= cc[u];
        }
        PdRecord *wr = rec + (t0 - (kWave - 2));  // lane 63's column at step t0 is t0 - 62
        const int c0 = t0 - lane + 1;
        if (__builtin_amdgcn_ballot_w64((any & kPdSpecial) != 0u) != 0) {
#pragma unroll
            for (int u = 0; u < kPdU; ++u)
                pd_step<kLast, true>(st, cur[u], cc[u], c0 + u, H, R, row, w, wn, inside_bit, writer, wr + u);
        } else {
#pragma unroll
            for (int u = 0; u < kPdU; ++u)
                pd_step<kLast, false>(st, cur[u], cc[u], c0 + u, H, R, row, w, wn, inside_bit, writer, wr + u);
        }
    }
    return st.acc;
}

__global__ __launch_bounds__(kWave) void pd_forward_kernel(PdBatch b, DeviceTables tab, uint32_t n_pairs,
                                                            int rec_columns) {
    extern __shared__ PdRecord lds_rec[];
    PdRecord *const rec = lds_rec + kPdLowPad;
    uint32_t *const cds = reinterpret_cast<uint32_t *>(rec + rec_columns) + kPdLowPad;
    const int lane = static_cast<int>(threadIdx.x);
    for (int c = lane - kPdLowPad; c < 0; c += kWave) cds[c] = 0u;  // columns below 1 carry no code
    const double log10_init = log10(ldexp(1.0, Scale<double>::kExp));
    for (uint32_t k = blockIdx.x; k < n_pairs; k += gridDim.x) {
        const uint32_t pair = b.order ? b.order[k] : k;
        const int r = static_cast<int>(pair / static_cast<uint32_t>(b.n_haps));
        const int h = static_cast<int>(pair % static_cast<uint32_t>(b.n_haps));
        const int R = b.rd_len[r], H = b.hap_len[h];
        if (R == 0) {  // no rows: the final sum over row 0's M and I is 0
            if (lane == 0) b.out[pair] = -__builtin_inf();
            continue;
        }
        const uint32_t *codes = b.hap_code + b.hap_off[h];
        const int64_t ro = b.rd_off[r];
        const double d0 = ldexp(1.0, Scale<double>::kExp) / static_cast<double>(H);
        // Row 1's sums over row 0 (M = I = 0, D = d0 on columns 0..H): A(c) = d0 GM_1 for 1 <= c <= H + 1.
        const double gm1 = pd_row(b, tab, ro, R, 0).GM;
        __builtin_amdgcn_wave_barrier();
        for (int c = lane; c < H + kPdPadColumns && c < rec_columns; c += kWave) {
            // Two 8-byte stores (volatile, so they are not merged into one 16-byte store).
            volatile double *v = &rec[c].a;
            v[0] = (c >= 1 && c <= H + 1) ? d0 * gm1 : 0.0;
            v[1] = 0.0;
            cds[c] = (c >= 1 && c <= H) ? codes[c - 1] : 0u;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

        const int n_stripes = (R + kWave - 1) / kWave;
        double acc = 0.0;
        for (int s = 0; s < n_stripes; ++s) {
            const int row = s * kWave + lane + 1;
            const PdRow w = pd_row(b, tab, ro, R, row - 1);
            const PdRow wn = pd_row(b, tab, ro, R, row);
            const uint32_t inside_bit = row == 1 ? kPdInsideRow1 : kPdInsideLater;
            if (s == n_stripes - 1)
                acc = pd_stripe<true>(rec, cds, lane, H, R, row, w, wn, inside_bit);
            else
                (void)pd_stripe<false>(rec, cds, lane, H, R, row, w, wn, inside_bit);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        const double total = __shfl(acc, (R - 1) & (kWave - 1), kWave);
        if (lane == 0) b.out[pair] = log10(total) - log10_init;
        __builtin_amdgcn_wave_barrier();
    }
}

}  // namespace

int pd_record_columns(int max_hap_len) { return max_hap_len + kPdPadColumns; }

size_t pd_lds_bytes(int max_hap_len) {
    return static_cast<size_t>(kPdLowPad + pd_record_columns(max_hap_len)) * (sizeof(PdRecord) + sizeof(uint32_t));
}

hipError_t launch_pd_forward(const PdBatch &batch, const DeviceTables &tables, uint32_t n_pairs, int max_hap_len,
                             int grid_blocks, hipStream_t stream) {
    if (n_pairs == 0) return hipSuccess;
    const size_t lds = pd_lds_bytes(max_hap_len);
    if (lds > max_lds_per_block()) return hipErrorInvalidValue;
    if (grid_blocks < 1) grid_blocks = 1;
    if (lds > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(pd_forward_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    hipLaunchKernelGGL(pd_forward_kernel, dim3(grid_blocks), dim3(kWave), lds, stream, batch, tables, n_pairs,
                       pd_record_columns(max_hap_len));
    return hipGetLastError();
}

}  // namespace gatk_pairhmm
