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
// computes column c = t - k + 1.  Lane k-1's E values of the last step arrive by DPP wave_shr:1 (the
// cell above); the ones of the step before are the diagonal.  Lane 0 reads the row above from an
// LDS record per column that lane 63 of the previous stripe wrote (E_M, E_I, E_D and the column's
// code; stripe 0 starts from row 0: M = I = 0, D = 2^1020 / H).  Column codes ride down the wave
// with the values, so only lane 0 reads LDS.  The row-R lane sums M + I over columns 1..H in
// column order, as the reference does (:132-139).
#include "pairhmm_device.hpp"
#include "pairhmm_kernels.hpp"

namespace gatk_pairhmm {
namespace {

using namespace dev;

// One column of a wave's LDS record array: the E values of the row above and the column code.
struct alignas(16) PdRecord {
    double em, ei, ed;
    uint32_t code, pad;
};
static_assert(sizeof(PdRecord) == 32, "PdRecord is two 16-byte LDS words");

constexpr int kPdPadColumns = 66;  // records past H that lane 0 reads ahead (c <= H + 64)

__device__ __forceinline__ uint32_t shr1_keep(uint32_t old, uint32_t src) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(old), static_cast<int>(src),
                                                             kDppWaveShr1, 0xF, 0xF, false));
}

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

__global__ __launch_bounds__(kWave) void pd_forward_kernel(PdBatch b, DeviceTables tab, uint32_t n_pairs,
                                                            int rec_columns) {
    extern __shared__ PdRecord rec[];
    const int lane = static_cast<int>(threadIdx.x);
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
        const double d0 = ldexp(1.0, Scale<double>::kExp) / static_cast<double>(H);
        __builtin_amdgcn_wave_barrier();
        for (int c = lane; c < H + kPdPadColumns && c < rec_columns; c += kWave) {
            PdRecord v;
            v.em = 0.0;
            v.ei = 0.0;
            v.ed = c <= H ? d0 : 0.0;
            v.code = (c >= 1 && c <= H) ? codes[c - 1] : 0u;
            v.pad = 0u;
            rec[c] = v;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

        const int64_t ro = b.rd_off[r];
        const int n_stripes = (R + kWave - 1) / kWave;
        double acc = 0.0;
        for (int s = 0; s < n_stripes; ++s) {
            const int row = s * kWave + lane + 1;
            double MM = 0, GM = 0, MI = 0, II = 0, MD = 0, DD = 0, pm = 0, px = 0;
            uint32_t x = 0, xbit = 0;
            if (row <= R) {
                const int p = row - 1;
                x = b.rd_bases[ro + p];
                const int q = b.rd_quals[ro + p], qi = b.rd_ins[ro + p], qd = b.rd_del[ro + p],
                          qc = b.rd_gcp[ro + p];
                const int mn = qi < qd ? qi : qd, mx = qi < qd ? qd : qi;
                MM = tab.mm_prob[((mx * (mx + 1)) >> 1) + mn];
                MI = tab.qual_err[qi];
                MD = tab.qual_err[qd];
                II = tab.qual_err[qc];
                DD = II;
                GM = 1.0 - II;
                const double eq = tab.qual_err[q];
                pm = 1.0 - eq;
                px = b.tristate ? eq / 3.0 : eq;
                if (x == 'N') px = pm;
                xbit = read_base_bit(x);
            }
            const uint32_t inside_bit = row == 1 ? kPdInsideRow1 : kPdInsideLater;
            // Lane state: own E and branch at column c-1, the E of the row above at c-1 (diagonal),
            // the column code.
            double EM = 0, EI = 0, ED = 0, bM = 0, bI = 0, bD = 0;
            double dM = 0, dI = 0, dD = 0;
            uint32_t code = 0;
            PdRecord nxt = rec[1];
            if (lane == 0) {
                const PdRecord c0 = rec[0];
                dM = c0.em;
                dI = c0.ei;
                dD = c0.ed;
            }
            const bool last = s == n_stripes - 1;
            const int steps = H + kWave - 1;
            for (int t = 0; t < steps; ++t) {
                const PdRecord cur = nxt;
                nxt = rec[t + 2];
                code = shr1_keep(cur.code, code);
                const double uM = wave_shr1(cur.em, EM), uI = wave_shr1(cur.ei, EI), uD = wave_shr1(cur.ed, ED);
                const bool match = ((code & 0xFFu) == x) || (code & xbit) != 0u;
                const double prior = match ? pm : px;
                const double M = prior * fma(dD, GM, fma(dI, GM, dM * MM));
                const double I = fma(uI, II, uM * MI);
                const double D = fma(ED, DD, EM * MD);
                const bool inside = (code & inside_bit) != 0u;
                bM = inside ? bM : EM;
                bI = inside ? bI : EI;
                bD = inside ? bD : ED;
                const double f = (code & kPdDelEnd) ? 1.0 : 0.0;  // max(0, X) == X for the cells here
                EM = fmax(bM * f, M);
                EI = fmax(bI * f, I);
                ED = fmax(bD * f, D);
                dM = uM;
                dI = uI;
                dD = uD;
                const int c = t - lane + 1;
                if (last && row == R && c >= 1 && c <= H) acc += M + I;
                if (lane == kWave - 1 && c >= 0 && c <= H) {
                    rec[c].em = EM;
                    rec[c].ei = EI;
                    rec[c].ed = ED;
                }
            }
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
    return static_cast<size_t>(pd_record_columns(max_hap_len)) * sizeof(PdRecord);
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
