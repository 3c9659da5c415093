// pairhmm_fwd.hip -- gfx950 PairHMM forward kernels (FP32 bulk pass and FP64 pass).
//
// Model: LoglessPairHMM.subComputeReadLikelihoodGivenHaplotypeLog10
// (src/main/java/org/broadinstitute/hellbender/utils/pairhmm/LoglessPairHMM.java:20-68):
//   M(r,c) = prior(r,c) * (M(r-1,c-1) MM_r + I(r-1,c-1) GM_r + D(r-1,c-1) GM_r)
//   I(r,c) = M(r-1,c) MI_r + I(r-1,c) II_r
//   D(r,c) = M(r,c-1) MD_r + D(r,c-1) DD_r
// with M(0,.) = I(0,.) = 0, D(0,.) = 2^S / H, column 0 zero, and
// log10 L = log10(sum_c M(R,c) + I(R,c)) - S log10 2.  Row r uses the qualities of read base r-1
// (PairHMMModel.java:169-171); priors follow LoglessPairHMM.java:84-92 with tristate correction.
//
// Mapping onto a CDNA wavefront (one read/haplotype pair per wave):
//   * Rows are processed in stripes of 64, one row per lane: lane k owns row 64 s + k + 1 of stripe
//     s and keeps its row's six transition probabilities and two priors in registers.
//   * At step t lane k computes column c = t - k (an anti-diagonal sweep).  Its upper neighbours
//     (row above, columns c and c-1) are lane k-1's values from the last two steps; they arrive as
//     DPP wave_shr:1 source operands of v_mul / v_fmac, so the exchange costs no instructions of
//     its own.
//   * Lane 0's row above is the last row of the previous stripe, which lane 63 wrote into LDS
//     boundary planes as it went.  Lane 0 does not read that row cell by cell: it takes two
//     ready-made sums per column, A(c) (its diagonal term) and B(c) (its vertical term), 8 columns
//     per LDS read ahead of use, and the DPP products leave lane 0's preloaded value in place
//     (wave_shr:1 without bound_ctrl), so no step waits on LDS.  FP32 forms A and B in a side pass
//     over the three planes lane 63 left (M, I, S = I + D); FP64, whose planes are twice as large,
//     has lane 63 form them directly into two planes.
//   * Rows past R on the last stripe pass the last row down (M = D = 0, I = M + I of the row
//     above), so the last stripe's planes end up holding M(R,c) + I(R,c).
//   * Cells with c <= 0 or c > H are computed too (they are exact zeros, resp. never read by a real
//     cell), which keeps the inner loop free of masks.
//   * The last-row sum, the unscaling and log10 are done in FP64.  FP32 launches use S = 120 and
//     flag pairs whose scaled sum is below the underflow threshold for an FP64 recompute (S = 1020).
//   * Launches are per LDS column class (kColClasses): a wave's LDS slice holds the haplotype and
//     the planes for the class's longest haplotype, so short haplotypes get more waves per CU.
#include "pairhmm_device.hpp"
#include "pairhmm_kernels.hpp"

namespace gatk_pairhmm {
namespace {

using namespace dev;

constexpr int kUnroll = 8;  // wave-steps per unrolled loop body

template <typename T>
struct WaveLds {
    uint8_t *hap;      // [cols]: haplotype byte of column c (1-based) at kLdsPad + c - 1
    T *bm, *bi;        // [cols]: boundary planes, column c at kLdsPad + c
    T *ring;           // FP64: LDS ring of the rows' (B, A) sums for the next row (kRing64 slots)
};

template <int kCols>
constexpr int plane_stride() { return lds_plane_stride(kCols); }

// Plane slots lane 0 reads over a stripe of `steps` wave-steps (8-column blocks from slot kLdsPad),
// rounded up to whole 64-slot chunks of the side pass; never more than kCols.
template <int kCols>
__device__ __forceinline__ int plane_span(int steps) {
    const int j = round_up(kLdsPad + steps, kWave);
    return j < kCols ? j : kCols;
}

// Two boundary planes per wave: lane 63's row (FP32: M and I in LDS, turned into lane 0's sums A
// and B by the side pass; FP64: A and B formed by lane 63 itself, in global memory).
constexpr int kPlanes = 2;
// FP64 step loop: every lane stores the sums it forms for the row below, (B(c), A(c+1)), at its
// own column into an LDS ring of 128 slots (plus 8 for a block that runs past slot 127), from which
// each 8-step block copies lane 63's finished columns out as lane 0's sums of the next stripe.
constexpr int kRing64 = 136;
constexpr size_t kRing64Bytes = kRing64 * 2 * sizeof(double);

template <typename T, int kCols>
constexpr size_t lds_per_wave() {
    return static_cast<size_t>(kCols) + (sizeof(T) == sizeof(float)
                                             ? kPlanes * static_cast<size_t>(plane_stride<kCols>()) * sizeof(T)
                                             : kRing64Bytes);
}

// Transition probabilities and priors of the row that read position p feeds (PairHMMModel.java:
// 158-172, LoglessPairHMM.java:84-92); positions past the read give the pass-through row
// (MI = II = 1, everything else 0).
template <typename T>
struct RowTrans {
    T MM, GM, MI, II, MD, DD, pm, px;
    uint32_t base;
};

template <typename T>
__device__ __forceinline__ RowTrans<T> row_trans(int p, int R, const uint8_t *__restrict__ rb,
                                                 const uint8_t *__restrict__ rq, const uint8_t *__restrict__ ri,
                                                 const uint8_t *__restrict__ rd, const uint8_t *__restrict__ rg,
                                                 const DeviceTables &tab) {
    RowTrans<T> w{0, 0, 0, 0, 0, 0, 0, 0, 0u};
    if (p < R) {
        w.base = rb[p];
        const int q = rq[p], qi = ri[p], qd = rd[p], qc = rg[p];
        const double eq = tab.qual_err[q], ei = tab.qual_err[qi], ed = tab.qual_err[qd], ec = tab.qual_err[qc];
        const int mn = qi < qd ? qi : qd, mx = qi < qd ? qd : qi;
        w.MM = static_cast<T>(tab.mm_prob[((mx * (mx + 1)) >> 1) + mn]);
        w.MI = static_cast<T>(ei);
        w.MD = static_cast<T>(ed);
        w.GM = static_cast<T>(1.0 - ec);
        w.II = static_cast<T>(ec);
        w.DD = w.II;
        w.pm = static_cast<T>(1.0 - eq);
        w.px = static_cast<T>(tab.qual_err[256 + q]);  // eq / 3.0
        if (w.base == 'N') w.px = w.pm;
    } else {
        // Rows past R pass the last row down unchanged: M = D = 0 and I(r,c) = M(r-1,c) + I(r-1,c).
        w.MI = 1;
        w.II = 1;
    }
    return w;
}

// A wave's haplotype bytes (LDS) and boundary planes: LDS for FP32, the wave's slice of the global
// FP64 planes otherwise.
template <typename T, int kCols>
__device__ __forceinline__ WaveLds<T> carve(uint8_t *base, const DeviceBatch &b, size_t gwave) {
    WaveLds<T> w;
    w.hap = base;
    if constexpr (sizeof(T) == sizeof(float)) {
        w.bm = reinterpret_cast<T *>(base + kCols);
        w.ring = nullptr;
    } else {
        w.bm = b.f64_planes + gwave * kPlanes * static_cast<size_t>(plane_stride<kCols>());
        w.ring = reinterpret_cast<T *>(base + kCols);
    }
    w.bi = w.bm + plane_stride<kCols>();
    return w;
}


// Side pass before stripe s > 0.  Lane 63 left the previous stripe's last row r in the two planes
// (M in P0, I in P1).  The pass rebuilds D of that row from its own recurrence,
//   D(c) = M(c-1) MD' + D(c-1) DD'   (MD', DD' of row r),
// and turns the row into lane 0's sums in place, in ascending column order:
//   A(c) = M(c-1) MM0 + (I(c-1) + D(c-1)) GM0 over M,   B(c) = M(c) MI0 + I(c) II0 over I.
// Per 64-column chunk D is a linear scan (linear_scan64) plus the carry D(j0-1) DD'^(l+1); column
// j0-1's M, I and D come from the previous chunk's lane 63 (the chunk itself is already overwritten).
// The pass covers plane slots [0, jmax) (jmax a multiple of 64, at most kCols): the slots lane 0 of
// the next stripe reads.  cd0 is D left of slot 0 (0 for every row but a constant one).  Lane 63
// wrote slots [1, wend) in the previous stripe (column t - 63 at step t); slot 0 and slots from wend
// on still hold what earlier passes left there and are read as M = I = 0.  Leftovers past a
// shorter stripe's end, turned over again by every later pass, can grow to inf, and a summing row
// that meets inf beyond its haplotype (times its prior 0) turns NaN.
template <int kCols>
__device__ __forceinline__ void side_pass_f32(float *P0, float *P1, int lane, float MM0, float GM0, float MI0,
                                              float II0, float MDr, float DDr, int jmax, float cd0 = 0.f,
                                              int wend = kCols) {
    const float a1 = DDr, a2 = a1 * a1, a4 = a2 * a2, a8 = a4 * a4, a16 = a8 * a8, a32 = a16 * a16;
    // DD'^((l mod 16) + 1) by the bits of l mod 16, then the row's 16-lane steps on top:
    // p_in = DD'^(l+1), p_31 = DD'^(l-31) for l >= 32 (1 below).
    float p_row = a1;
    p_row = (lane & 1) ? p_row * a1 : p_row;
    p_row = (lane & 2) ? p_row * a2 : p_row;
    p_row = (lane & 4) ? p_row * a4 : p_row;
    p_row = (lane & 8) ? p_row * a8 : p_row;
    const float p_16 = (lane & 16) ? p_row * a16 : p_row;
    const float p_in = (lane & 32) ? p_16 * a32 : p_16;
    const float p_31 = (lane & 32) ? p_16 : 1.f;
    float cm = 0.f, ci = 0.f, cd = cd0;  // M, I, D of column j0 - 1
    __builtin_amdgcn_wave_barrier();
    for (int j0 = 0; j0 < jmax; j0 += kWave) {
        const int j = j0 + lane;
        const bool written = j >= kLdsPad - (kWave - 1) && j < wend;
        const float mj = written ? P0[j] : 0.f, ij = written ? P1[j] : 0.f;
        const float mprev = wave_shr1(cm, mj), iprev = wave_shr1(ci, ij);
        const float dj = fmaf(cd, p_in, linear_scan64(mprev * MDr, a1, a2, a4, a8, p_row, p_31));
        const float dprev = wave_shr1(cd, dj);
        __builtin_amdgcn_wave_barrier();
        P0[j] = fmaf(iprev + dprev, GM0, mprev * MM0);
        P1[j] = fmaf(ij, II0, mj * MI0);
        cm = lane63(mj);
        ci = lane63(ij);
        cd = lane63(dj);
        __builtin_amdgcn_wave_barrier();
    }
}

// One pass over a pair, stripe by stripe.  Returns the scaled last-row sum (FP64, every lane).
//
// Stripe s covers rows 64 s + 1 .. 64 s + 64, one per lane.  Per step each lane keeps its own M, I,
// D from the previous step and M, S = I + D from two steps back:
//   M(r,c) = prior * (M(r-1,c-1) MM + S(r-1,c-1) GM)     [lane k-1, two steps back]
//   I(r,c) = M(r-1,c) MI + I(r-1,c) II                    [lane k-1, one step back]
//   D(r,c) = M(r,c-1) MD + D(r,c-1) DD                    [own lane, one step back]
// Lane 0's row above is the last row of the previous stripe, which lane 63 left in the LDS
// boundary planes (M and I; D is rebuilt by the side pass).  Before the stripe the side pass turns
// those planes into lane 0's two sums,
//   A(c) = M(r-1,c-1) MM + S(r-1,c-1) GM   and   B(c) = M(r-1,c) MI + I(r-1,c) II,
// in place (A over M, B over I).  Lane 0 reads them ahead with wave-uniform loads and brings them
// into the DPP products, which leave lane 0 unwritten, so no step waits on LDS.
template <typename T, int kCols, bool kHapHasN>
__device__ double forward_pair_f32(const WaveLds<T> &L, int lane, int H, int R, const uint8_t *__restrict__ rb,
                               const uint8_t *__restrict__ rq, const uint8_t *__restrict__ ri,
                               const uint8_t *__restrict__ rd, const uint8_t *__restrict__ rg,
                               const DeviceTables &tab) {
    constexpr int kStripe = kWave;
    const int n_stripes = (R + kStripe - 1) / kStripe;
    const int steps = round_up(H + kWave, kUnroll);
    const int jmax = plane_span<kCols>(steps);
    for (int s = 0; s < n_stripes; ++s) {
        const RowTrans<T> w = row_trans<T>(s * kStripe + lane, R, rb, rq, ri, rd, rg, tab);
        const T MM = w.MM, GM = w.GM, MI = w.MI, II = w.II, MD = w.MD, DD = w.DD, pm = w.pm, px = w.px;
        const uint32_t base = w.base;
        if (s > 0) {
            // Row 64 s (read position 64 s - 1) is the previous stripe's last row; wave-uniform.
            const RowTrans<T> up = row_trans<T>(s * kStripe - 1, R, rb, rq, ri, rd, rg, tab);
            side_pass_f32<kCols>(L.bm, L.bi, lane, rfl(MM), rfl(GM), rfl(MI), rfl(II), up.MD, up.DD, jmax);
        }
    T m1 = 0, i1 = 0, d1 = 0, s1 = 0;  // own cell one step back (column c-1)
        T m2 = 0, s2 = 0;                  // own cell two steps back (column c-2)
        // Byte loads of the haplotype, eight per block.  volatile keeps them from being merged into
        // unaligned 8-byte loads; the explicit LDS address space keeps them ds_read_u8.
        using lds_u8 = const volatile __attribute__((address_space(3))) uint8_t;
        lds_u8 *hp = (lds_u8 *)(L.hap + kLdsPad - 1 - lane);  // address-space cast: C-style only
        // Every lane stores its M and I at its own column c = t - lane at step t (no one-lane exec
        // mask in the step).  Lane 63 writes column c last, at step c + 63, so the planes end with
        // its row; lane 0 reads A, B of column c by step c, before any lane overwrites it.
        using lds_t = __attribute__((address_space(3))) T;
        int lane_w = lane;
        asm volatile("" : "+v"(lane_w));
        lds_t *wm = (lds_t *)(L.bm + kLdsPad) - lane_w, *wi = (lds_t *)(L.bi + kLdsPad) - lane_w;
        for (int t0 = 0; t0 < steps; t0 += kUnroll) {
            lds_t *wmb = wm + t0, *wib = wi + t0;
            T av[kUnroll], bv[kUnroll];
            __builtin_memcpy(av, __builtin_assume_aligned(L.bm + kLdsPad + t0, 16), sizeof(av));
            __builtin_memcpy(bv, __builtin_assume_aligned(L.bi + kLdsPad + t0, 16), sizeof(bv));
            uint32_t hbs[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) hbs[u] = hp[t0 + u];
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const uint32_t hb = hbs[u];
                const bool match = (hb == base) || (kHapHasN && hb == 'N');
                const T prior = match ? pm : px;
                T a = av[u], i = bv[u];
                dpp_upper2(a, m2, MM, s2, GM, i, m1, MI, i1, II);
                const T m = a * prior;
                // The lane carries D' = D / MD: D'(c) = M(c-1) + D'(c-1) DD and S = I + MD D', one FMA
                // each instead of a multiply, an FMA and an add.  D' <= sum_c M(r,c), so it keeps
                // FP32's range even where DD is close to 1.
                const T d = fma(d1, DD, m1);
                const T sx = fma(MD, d, i);
                wmb[u] = m;
                wib[u] = i;
                m2 = m1;
                s2 = s1;
                m1 = m;
                i1 = i;
                d1 = d;
                s1 = sx;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    double acc = 0.0;
    for (int c = 1 + lane; c <= H; c += kWave)
        acc += static_cast<double>(L.bm[kLdsPad + c]) + static_cast<double>(L.bi[kLdsPad + c]);
    return wave_sum(acc);
}

// FP64 pass.  Each lane forms the sums the row below it needs from its own new M, S and I, with the
// transitions of the row below (the next lane's row; for lane 63 the next stripe's first row):
//   A(c+1) = M(c) MM' + S(c) GM'   and   B(c) = M(c) MI' + I(c) II',
// and the row below takes them one lane over: B at the next step (same column), A two steps later
// (the next column).  So two doubles cross lanes per step (four 32-bit DPP moves) where the
// products over the row above would move four (M and S two steps back, M and I one step back), and
// a lane needs only the next row's MM, GM, MI, II and its own MD, DD and priors.  Lane 63's sums are
// lane 0's A and B of the next stripe; after the last stripe the next row is the pass-through row
// (MM = GM = 0, MI = II = 1), so the B plane ends up holding M(R,c) + I(R,c).  Lane 0 takes A and B
// as the old values of the DPP moves (wave_shr:1 without bound_ctrl leaves lane 0 unwritten).
//
// Every lane stores its (B(c), A(c+1)) at its own column c = t - lane into an LDS ring at step t (no
// one-lane exec mask in the step); lane 63 writes column c last, at step c + 63.  After each block
// of 8 steps the whole wave copies lane 63's 8 columns finished in it out to the plane, one 16-byte
// store per column, with every lane storing (lanes 8-63 repeat lanes 0-7).  A lane's ring slots for
// block t0 start at (t0 - lane) mod 128, so a block never wraps and two live columns (less than 128
// apart) never share a slot.
//
// A and B live in the wave's slice of global memory rather than LDS (an FP64 plane is 8 bytes a
// column; in LDS the long-haplotype classes held a CU to 5-7 waves); the slice stays in L2.  They
// are interleaved, P[2 j] = A and P[2 j + 1] = B at j = kLdsPad + c, so a block's 8 columns are 16
// contiguous doubles.  Lane 0's values are read through a VGPR address (an opaque zero offset keeps
// the compiler from turning the uniform load into a scalar one, whose cache does not see vector
// stores) and rolled in half a block ahead of use: after step 3 a block's first half of the buffer
// is reloaded with the next block's first four columns, after step 7 the second half.  A wait for
// the wave's stores at each stripe start orders the previous stripe's stores before these loads;
// all accesses come from one wave on one CU, so its L1 holds no stale copy.
template <int kCols, bool kHapHasN>
__device__ double forward_pair_f64(const WaveLds<double> &L, int lane, int H, int R, const uint8_t *__restrict__ rb,
                                   const uint8_t *__restrict__ rq, const uint8_t *__restrict__ ri,
                                   const uint8_t *__restrict__ rd, const uint8_t *__restrict__ rg,
                                   const DeviceTables &tab) {
    const int n_stripes = (R + kWave - 1) / kWave;
    const int steps = round_up(H + kWave, kUnroll);
    double *P = L.bm;  // A(c) at 2 (kLdsPad + c), B(c) at 2 (kLdsPad + c) + 1
    using lds_u8 = const volatile __attribute__((address_space(3))) uint8_t;
    lds_u8 *hp = (lds_u8 *)(L.hap + kLdsPad - 1 - lane);  // address-space cast: C-style only
    using lds_f64 = __attribute__((address_space(3))) double;
    lds_f64 *const ring = (lds_f64 *)L.ring;  // slot k: B at 2 k, A at 2 k + 1
    constexpr int kLag = kWave - 1;
    int zoff;
    asm volatile("v_mov_b32 %0, 0" : "=v"(zoff));
    const double *r0 = P + 2 * kLdsPad + zoff;
    // Copy-out: lane handles column offset v of the block's finished columns.
    const int v = lane & (kUnroll - 1);
    for (int s = 0; s < n_stripes; ++s) {
        const RowTrans<double> w = row_trans<double>(s * kWave + lane, R, rb, rq, ri, rd, rg, tab);
        const RowTrans<double> wn = row_trans<double>(s * kWave + lane + 1, R, rb, rq, ri, rd, rg, tab);
        // dd = D / MD: D(c) = D(c-1) DD + M(c-1) MD becomes dd(c) = dd(c-1) DD + M(c-1), and
        // S = I + D = I + MD dd, one FMA each (rows past R have MD = DD = 0, so S = I there).
        double m1 = 0, dd1 = 0;
        double an1 = 0, an2 = 0, bn1 = 0;  // sums for the row below, one and two steps back
        // The previous stripe's (or the initial) stores complete before this stripe reads them.
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        double ab[2 * kUnroll];
        __builtin_memcpy(ab, __builtin_assume_aligned(r0, 16), sizeof(ab));
        for (int t0 = 0; t0 < steps; t0 += kUnroll) {
            lds_f64 *wr = ring + 2 * ((t0 - lane) & 127);
            uint32_t hbs[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) hbs[u] = hp[t0 + u];
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const uint32_t hb = hbs[u];
                const bool match = (hb == w.base) || (kHapHasN && hb == 'N');
                const double prior = match ? w.pm : w.px;
                double a = ab[2 * u], i = ab[2 * u + 1];
                a = wave_shr1(a, an2);  // lane 0 keeps its plane values
                i = wave_shr1(i, bn1);
                const double m = a * prior;
                const double dd = fma(dd1, w.DD, m1);
                const double sx = fma(dd, w.MD, i);
                const double an = fma(sx, wn.GM, m * wn.MM);
                const double bn = fma(i, wn.II, m * wn.MI);
                wr[2 * u] = bn;
                wr[2 * u + 1] = an;
                an2 = an1;
                an1 = an;
                bn1 = bn;
                m1 = m;
                dd1 = dd;
                // The next block's halves, loaded unconditionally (the slice has slots past the last
                // block): with a fixed count of loads in the loop, the wait for a half is a counted
                // one instead of a full drain at the block start.
                if (u == kUnroll / 2 - 1)
                    __builtin_memcpy(ab, __builtin_assume_aligned(r0 + 2 * (t0 + kUnroll), 16), sizeof(ab) / 2);
                if (u == kUnroll - 1)
                    __builtin_memcpy(ab + kUnroll, __builtin_assume_aligned(r0 + 2 * (t0 + kUnroll) + kUnroll, 16),
                                     sizeof(ab) / 2);
            }
            __builtin_amdgcn_wave_barrier();
            // Lane 63 finished columns t0 - 63 .. t0 - 56 in this block, at ring slots
            // ((t0 - 63) mod 128) + v: (B(c), A(c + 1)) as they go to the plane.
            {
                const lds_f64 *rs = ring + 2 * (((t0 - kLag) & 127) + v);
                const double ba[2] = {rs[0], rs[1]};
                __builtin_memcpy(P + 2 * (kLdsPad + t0 - kLag + v) + 1, ba, sizeof(ba));
            }
            __builtin_amdgcn_wave_barrier();
        }
        __builtin_amdgcn_wave_barrier();
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the last stripe's B values are complete
    double acc = 0.0;
    for (int c = 1 + lane; c <= H; c += kWave) acc += P[2 * (kLdsPad + c) + 1];
    return wave_sum(acc);
}

template <typename T, int kCols, bool kHapHasN>
__device__ __forceinline__ double forward_pair(const WaveLds<T> &L, int lane, int H, int R, const uint8_t *rb,
                                               const uint8_t *rq, const uint8_t *ri, const uint8_t *rd,
                                               const uint8_t *rg, const DeviceTables &tab) {
    if constexpr (sizeof(T) == sizeof(double))
        return forward_pair_f64<kCols, kHapHasN>(L, lane, H, R, rb, rq, ri, rd, rg, tab);
    else
        return forward_pair_f32<T, kCols, kHapHasN>(L, lane, H, R, rb, rq, ri, rd, rg, tab);
}

template <typename T, int kCols>
__device__ void process_pair(const DeviceBatch &b, const DeviceTables &tab, const RescueSink &rs, const WaveLds<T> &L,
                             int lane, uint32_t pair) {
    int r, h;
    if (b.pair_read) {
        r = b.pair_read[pair];
        h = b.pair_hap[pair];
    } else {
        r = static_cast<int>(pair / static_cast<uint32_t>(b.n_haps));
        h = static_cast<int>(pair % static_cast<uint32_t>(b.n_haps));
    }
    const int H = b.hap_len[h];
    const int R = b.rd_len[r];
    const uint8_t *hap = b.hap_bases + b.hap_off[h];
    const int64_t ro = b.rd_off[r];

    // Stage the haplotype (zero padded) and the row-0 boundary: M = I = 0, D = 2^S / H
    // (LoglessPairHMM.java:30-36).
    int has_n = 0;
    for (int j = lane; j < kCols; j += kWave) {
        const int c = j - kLdsPad + 1;  // hap byte j belongs to column c
        uint8_t v = 0;
        if (c >= 1 && c <= H) {
            v = hap[c - 1];
            has_n |= (v == 'N');
        }
        L.hap[j] = v;
    }
    const T d0 = static_cast<T>(__builtin_ldexp(1.0, Scale<T>::kExp) / static_cast<double>(H));
    // Row 1's sums over row 0 (M = I = 0, D = 2^S / H from column 0 on): A(c) = d0 GM_1 for c >= 1,
    // B(c) = 0 (FP32: A in the M plane, B in the I plane; FP64: interleaved, A at 2 j, B at 2 j + 1).
    const T gm1 = row_trans<T>(0, R, b.rd_bases + ro, b.rd_quals + ro, b.rd_ins + ro, b.rd_del + ro,
                               b.rd_gcp + ro, tab).GM;
    for (int j = lane; j < kCols; j += kWave) {
        const T a0 = j >= kLdsPad + 1 ? d0 * gm1 : T(0);
        if constexpr (sizeof(T) == sizeof(float)) {
            L.bm[j] = a0;
            L.bi[j] = 0;
        } else {
            L.bm[2 * j] = a0;
            L.bm[2 * j + 1] = 0;
        }
    }
    double sum;
    if (__any(has_n))
        sum = forward_pair<T, kCols, true>(L, lane, H, R, b.rd_bases + ro, b.rd_quals + ro, b.rd_ins + ro, b.rd_del + ro,
                                    b.rd_gcp + ro, tab);
    else
        sum = forward_pair<T, kCols, false>(L, lane, H, R, b.rd_bases + ro, b.rd_quals + ro, b.rd_ins + ro, b.rd_del + ro,
                                     b.rd_gcp + ro, tab);
    // Every lane holds the same sum (butterfly reduction), so the result and the rescue decision are
    // wave-uniform; all lanes store the same value to the same address rather than branching on
    // the lane id (a lane-divergent branch here would sit next to the convergent work fetch below).
    const double scale_log10 = static_cast<double>(Scale<T>::kExp) * 0.30102999566398119521;
    b.out[pair] = log10(sum) - scale_log10;
    if (rs.list != nullptr) {
        const int need = __builtin_amdgcn_readfirstlane(static_cast<int>(!(sum >= rs.threshold)));
        if (need) {
            const uint32_t slot = __builtin_amdgcn_readfirstlane(atomicAdd(rs.count, lane == 0 ? 1u : 0u));
            rs.list[slot] = pair;
        }
    }
}

template <typename T, int kCols>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(sizeof(T) == sizeof(double) ? 4 : 1)))
void forward_kernel(DeviceBatch b, DeviceTables tab, WorkSource ws, RescueSink rs) {
    extern __shared__ __align__(16) uint8_t lds_raw[];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int waves_per_block = blockDim.x / kWave;
    constexpr size_t kPerWave = lds_per_wave<T, kCols>();
    const size_t gwave = static_cast<size_t>(blockIdx.x) * waves_per_block + wave;
    // FP64 waves past the planes' capacity stay idle (queue launches only; grid-stride launches are
    // sized to the capacity by the host).
    if (sizeof(T) == sizeof(double) && gwave >= b.f64_plane_waves) return;
    const WaveLds<T> L = carve<T, kCols>(lds_raw + kPerWave * wave, b, gwave);
    const uint32_t n_items = ws.n_items_dev ? *ws.n_items_dev : ws.n_items;
    if (ws.head) {
        for (;;) {
            // All lanes issue the atomic (lane 0 adds 1, the rest add 0), so the fetch has no
            // lane-divergent branch; lane 0's pre-increment value is the wave's item.
            const uint32_t it = __builtin_amdgcn_readfirstlane(atomicAdd(ws.head, lane == 0 ? 1u : 0u));
            if (it >= n_items) break;
            process_pair<T, kCols>(b, tab, rs, L, lane, ws.items ? ws.items[it] : it);
        }
    } else {
        // The wave's index is wave-uniform; readfirstlane tells the compiler so, which keeps the
        // pair, its lengths and every loop over them scalar (uniform branches, no exec masking).
        const uint32_t stride = gridDim.x * waves_per_block;
        for (uint32_t it = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(gwave)); it < n_items; it += stride)
            process_pair<T, kCols>(b, tab, rs, L, lane, ws.items ? ws.items[it] : it);
    }
}

// ---- packed stripes (FP32) --------------------------------------------------------------------
//
// The aligned schedule above gives every pair whole 64-row stripes, so a read of R rows pays for
// 64 ceil(R/64) rows.  Packed stripes stream the rows of successive pairs through a wave's lanes
// instead: a stripe holds the end of one pair and the start of the next.  Each pair contributes
// R + 3 rows to the stream, in order
//   CONST  a row whose S = I + D is 1 in every column (M = I = 0, D' = 1 from the stripe start);
//   ZERO   the pair's row 0: M is d0 = 2^120/H in column -1 only (its prior is d0 on the marker
//          byte staged at column -1, 0 elsewhere, times the CONST row's S = 1), so its D, the running
//          sum of M (DD = MD = 1), is d0 from column 0 on: D(0, c) = 2^120/H as LoglessPairHMM.java:
//          30-36 sets it;
//   READ   rows 1..R, with the read's own transitions, except that row 1 takes MM = MI = 0 (its
//          row above has M = I = 0 in the model, so the terms vanish; here they hold the spike) and
//          row R takes MD = 0 (D(R, .) adds nothing to the result, and S(R, .) = I(R, .) then feeds
//          the summing row);
//   SUM    M(c) = (M + I)(R, c - 1) on columns 2..H+1 (prior 1 on real haplotype bytes read one
//          column to the left, 0 on padding), I = 0, and D' with DD = 1 is the running sum of M: at
//          column H + 2 it holds sum_c (M + I)(R, c) and stays there (M = 0 on padding).  The sum is
//          FP32 in column order, whatever stripe position and neighbours the pair gets, so results
//          are identical from run to run; its relative error is at most (H - 1) 2^-24 and in
//          practice a few ulps, since a last row's terms fall off fast away from the alignment.
// Every value stays a sub-probability times 2^120, so no product of a zeroed transition with a
// neighbour's value can be inf or NaN.  Rows need no other case: the step loop is the aligned
// schedule's, with per-lane haplotype pointers and transitions.  A stripe holds rows of at most two
// pairs (lanes past the second pair's rows idle), whose haplotypes sit in two LDS buffers; it runs
// max(H) + 66 steps, so the SUM row reaches column H + 2 and lane 63 writes every column lane 0 of
// the next stripe needs.  Haplotype bytes 0 (padding) and 1 (the row-0 marker) are reserved; the
// host uses the aligned schedule for batches whose haplotypes contain them.  Waves take the queue in
// fixed chunks (see packed_fetch).

constexpr uint32_t kPadByte = 0;      // haplotype bytes outside 1..H
constexpr uint32_t kZeroRowMark = 1;  // staged at column -1
constexpr uint32_t kNoByte = 0x1FF;   // compares equal to no byte

enum RowKind : int { kRowIdle = 0, kRowConst = 1, kRowZero = 2, kRowRead = 3, kRowSum = 4 };

// A pair in flight on a wave (wave-uniform).
struct PackedPair {
    uint32_t pair;
    int R, H;
    int64_t ro;  // read offset
    int hap_n;   // the haplotype has an 'N'
};

// LDS of a packed wave: a guard (the SUM row's prior reads one byte left of buffer 0 at lane 63's
// first step), two haplotype buffers, the two boundary planes.
constexpr int kHapGuard = 16;

template <int kCols>
constexpr size_t packed_lds_per_wave() {
    return kHapGuard + 2 * static_cast<size_t>(kCols) +
           kPlanes * static_cast<size_t>(plane_stride<kCols>()) * sizeof(float);
}

// A wave takes the queue in chunks of ws.chunk consecutive items, each chunk by one atomic on the
// queue head, and streams a chunk's pairs through its stripes in queue order.  A chunk always
// starts at lane 0 of a fresh stripe and no chunk shares a stripe with another, because a pair's
// bits depend on where its rows fall in the stripes (a row fed through the side pass rounds
// differently from one fed by DPP): chunks are fixed by the queue and the launch, so every run gives
// the same results while the waves still balance their load dynamically.  The launch sets the chunk
// (packed_chunk): at most kPackedChunk items, and few enough that every wave of the grid gets about
// four chunks, so a small queue (one region's pairs) still spreads over the whole GPU.
constexpr uint32_t kPackedChunk = 16;

struct ChunkCursor {
    uint32_t next = 0, end = 0;  // items [next, end) of the wave's current chunk
};

// Takes the next pair of the wave's current chunk; with new_chunk, first takes a new chunk when the
// current one is used up.  A pair with an empty read has L = 0 (log10 L = -inf): it is written at
// once and, as in the aligned schedule, listed for the FP64 pass.
__device__ bool packed_fetch(const DeviceBatch &b, const WorkSource &ws, uint32_t n_items, const RescueSink &rs,
                             int lane, PackedPair &p, ChunkCursor &cc, bool new_chunk) {
    for (;;) {
        if (cc.next >= cc.end) {
            if (!new_chunk) return false;
            // All lanes issue the atomic (lane 0 adds 1): the fetch has no lane-divergent branch.
            const uint32_t c = __builtin_amdgcn_readfirstlane(atomicAdd(ws.head, lane == 0 ? 1u : 0u));
            const uint64_t first = static_cast<uint64_t>(c) * ws.chunk;
            if (first >= n_items) return false;
            cc.next = static_cast<uint32_t>(first);
            cc.end = static_cast<uint32_t>(first + ws.chunk < n_items ? first + ws.chunk : n_items);
        }
        const uint32_t it = cc.next++;
        const uint32_t pair = __builtin_amdgcn_readfirstlane(ws.items ? ws.items[it] : it);
        int r, h;
        if (b.pair_read) {
            r = b.pair_read[pair];
            h = b.pair_hap[pair];
        } else {
            r = static_cast<int>(pair / static_cast<uint32_t>(b.n_haps));
            h = static_cast<int>(pair % static_cast<uint32_t>(b.n_haps));
        }
        p.pair = pair;
        p.R = __builtin_amdgcn_readfirstlane(b.rd_len[r]);
        p.H = __builtin_amdgcn_readfirstlane(b.hap_len[h]);
        const int64_t ro = b.rd_off[r];
        p.ro = (static_cast<int64_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(ro >> 32))) << 32) |
               static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(ro)));
        p.hap_n = 0;
        if (p.R > 0) return true;
        b.out[pair] = -__builtin_inf();
        if (rs.list != nullptr) {
            const uint32_t slot = __builtin_amdgcn_readfirstlane(atomicAdd(rs.count, lane == 0 ? 1u : 0u));
            rs.list[slot] = pair;
        }
    }
}

// Stages pair p's haplotype in LDS buffer `hb`: byte of column c at kLdsPad + c - 1, the row-0 marker
// at column -1, padding everywhere else (the SUM row's prior needs padding up to the stripe's end).
template <int kCols>
__device__ void packed_stage(const DeviceBatch &b, uint8_t *hbuf, int lane, PackedPair &p) {
    // An explicit LDS pointer: through a generic one these become FLAT stores, which count on both
    // wait counters out of order and turn every LDS wait of the step loop into lgkmcnt(0).
    using lds_u8w = __attribute__((address_space(3))) uint8_t;
    lds_u8w *hb = (lds_u8w *)hbuf;  // address-space cast: C-style only
    const int h = b.pair_read ? b.pair_hap[p.pair] : static_cast<int>(p.pair % static_cast<uint32_t>(b.n_haps));
    const uint8_t *hap = b.hap_bases + b.hap_off[h];
    int has_n = 0;
    for (int j = lane; j < kCols; j += kWave) {
        const int c = j - kLdsPad + 1;
        uint8_t v = kPadByte;
        if (c >= 1 && c <= p.H) {
            v = hap[c - 1];
            has_n |= (v == 'N');
        } else if (c == -1) {
            v = kZeroRowMark;
        }
        hb[j] = v;
    }
    p.hap_n = __any(has_n) ? 1 : 0;
}

// Transitions, priors and starting state of one lane's row.
struct PackedRow {
    RowTrans<float> w;
    uint32_t nval;           // second byte the prior treats as a match ('N' for read rows)
    int shift;               // haplotype column offset of the lane's prior (-1 for SUM)
    float m1, s1, d1, s2;    // state left of the lane's first column (columns -k-1 and -k-2)
    int kind;
};

__device__ __forceinline__ PackedRow packed_row(int kind, int idx, const PackedPair &p, const DeviceBatch &b,
                                                const DeviceTables &tab, int lane) {
    PackedRow x;
    x.w = RowTrans<float>{0, 0, 0, 0, 0, 0, 0, 0, kNoByte};
    x.nval = kNoByte;
    x.shift = 0;
    x.m1 = x.s1 = x.d1 = x.s2 = 0.f;
    x.kind = kind;
    if (kind == kRowRead) {
        const int r = idx - 1;  // 1-based read row
        const int64_t ro = p.ro;
        x.w = row_trans<float>(r - 1, p.R, b.rd_bases + ro, b.rd_quals + ro, b.rd_ins + ro, b.rd_del + ro,
                               b.rd_gcp + ro, tab);
        if (r == 1) {
            x.w.MM = 0.f;
            x.w.MI = 0.f;
        }
        if (r == p.R) x.w.MD = 0.f;
        x.nval = 'N';
    } else if (kind == kRowConst) {
        x.w.MD = 1.f;
        x.w.DD = 1.f;
        x.d1 = 1.f;
        x.s1 = 1.f;
        x.s2 = 1.f;
    } else if (kind == kRowZero) {
        const float d0 = static_cast<float>(__builtin_ldexp(1.0, Scale<float>::kExp) / static_cast<double>(p.H));
        x.w.GM = 1.f;
        x.w.MD = 1.f;
        x.w.DD = 1.f;
        x.w.pm = d0;
        x.w.base = kZeroRowMark;
        if (lane == 0) x.m1 = d0;  // column -1 lies left of lane 0's first column
    } else if (kind == kRowSum) {
        x.w.MM = 1.f;
        x.w.GM = 1.f;
        x.w.DD = 1.f;
        x.w.px = 1.f;
        x.w.base = kPadByte;
        x.shift = -1;
    }
    return x;
}

// One stripe of steps (the aligned schedule's step, per-lane haplotype pointers and 'N' bytes).
// Returns the lane's D' after the last step (a SUM row's last-row sum).
template <int kCols, bool kHapHasN, int kU>
__device__ __forceinline__ float packed_steps(const WaveLds<float> &L, const uint8_t *hap, int lane,
                                              const PackedRow &x, int steps) {
    const float MM = x.w.MM, GM = x.w.GM, MI = x.w.MI, II = x.w.II, MD = x.w.MD, DD = x.w.DD, pm = x.w.pm,
                px = x.w.px;
    const uint32_t base = x.w.base, nval = x.nval;
    float m1 = x.m1, i1 = 0.f, d1 = x.d1, s1 = x.s1;
    float m2 = 0.f, s2 = x.s2;
    using lds_u8 = const volatile __attribute__((address_space(3))) uint8_t;
    lds_u8 *hp = (lds_u8 *)(hap + kLdsPad - 1 - lane + x.shift);  // address-space cast: C-style only
    // Every lane stores its M and I at its own column c = t - lane at step t, so the step has no
    // one-lane exec mask (no exec writes, no branch).  Lane 63 writes column c last, at step c + 63,
    // so the planes end with its row; lane 0 has read A and B of column c by step c, before any lane
    // overwrites it.
    using lds_f32 = __attribute__((address_space(3))) float;
    int lane_w = lane;
    asm volatile("" : "+v"(lane_w));
    lds_f32 *wm = (lds_f32 *)(L.bm + kLdsPad) - lane_w, *wi = (lds_f32 *)(L.bi + kLdsPad) - lane_w;
    // No scalar load may be in flight at the loop head: they complete out of order, and one pending
    // there makes every LDS wait in the loop a full lgkmcnt(0) instead of waiting for the first reads.
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    for (int t0 = 0; t0 < steps; t0 += kU) {
        lds_f32 *wmb = wm + t0, *wib = wi + t0;
        float av[kU], bv[kU];
        __builtin_memcpy(av, __builtin_assume_aligned(L.bm + kLdsPad + t0, 16), sizeof(av));
        __builtin_memcpy(bv, __builtin_assume_aligned(L.bi + kLdsPad + t0, 16), sizeof(bv));
        uint32_t hbs[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) hbs[u] = hp[t0 + u];
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const uint32_t hb = hbs[u];
            const bool match = (hb == base) || (kHapHasN && hb == nval);
            const float prior = match ? pm : px;
            float a = av[u], i = bv[u];
            dpp_upper2(a, m2, MM, s2, GM, i, m1, MI, i1, II);
            const float m = a * prior;
            const float d = fmaf(d1, DD, m1);
            const float sx = fmaf(MD, d, i);
            wmb[u] = m;
            wib[u] = i;
            m2 = m1;
            s2 = s1;
            m1 = m;
            i1 = i;
            d1 = d;
            s1 = sx;
        }
    }
    __builtin_amdgcn_wave_barrier();
    return d1;
}

// The packed-stripe FP32 pass of one wave over its share of a queue, kU wave-steps per unrolled
// loop body.
template <int kCols, int kU>
__device__ void packed_wave(const DeviceBatch &b, const DeviceTables &tab, const WorkSource &ws, uint32_t n_items,
                            const RescueSink &rs, uint8_t *lds, int lane) {
    ChunkCursor cc;
    WaveLds<float> L;
    // Haplotype buffer k at lds + kHapGuard + k kCols (chosen by arithmetic, not by indexing an array
    // of pointers, so the compiler keeps them LDS pointers).
    auto hapbuf = [&](int k) { return lds + kHapGuard + (k ? kCols : 0); };
    L.hap = hapbuf(0);
    L.bm = reinterpret_cast<float *>(lds + kHapGuard + 2 * kCols);
    L.bi = L.bm + plane_stride<kCols>();
    for (int j = lane; j < plane_stride<kCols>(); j += kWave) {
        L.bm[j] = 0.f;
        L.bi[j] = 0.f;
    }
    const double scale_log10 = static_cast<double>(Scale<float>::kExp) * 0.30102999566398119521;
    PackedPair cur{}, nxt{};
    bool have_cur = false;
    int cur_next = 0;  // next row of cur: 0 CONST, 1 ZERO, 2..R+1 READ, R+2 SUM
    int buf_cur = 0;
    bool first = true;
    float md63 = 0.f, dd63 = 0.f, cd63 = 0.f;  // lane 63's row of the previous stripe (side pass)
    int wend = 0;                              // plane slots [1, wend) lane 63 wrote in that stripe
    for (;;) {
        if (!have_cur) {
            if (!packed_fetch(b, ws, n_items, rs, lane, cur, cc, true)) break;
            have_cur = true;
            cur_next = 0;
            buf_cur = 0;
            packed_stage<kCols>(b, hapbuf(buf_cur), lane, cur);
        }
        const int rows_cur = cur.R + 3 - cur_next;
        bool two = false;
        if (rows_cur < kWave) {
            if (packed_fetch(b, ws, n_items, rs, lane, nxt, cc, false)) {
                two = true;
                packed_stage<kCols>(b, hapbuf(buf_cur ^ 1), lane, nxt);
            }
        }
        // This lane's row.
        const bool in_cur = lane < rows_cur;
        const int idx = in_cur ? cur_next + lane : lane - rows_cur;
        const PackedPair &pp = in_cur ? cur : nxt;
        int kind = kRowIdle;
        if (in_cur || (two && idx < nxt.R + 3)) {
            const int R = in_cur ? cur.R : nxt.R;
            kind = idx == 0 ? kRowConst : idx == 1 ? kRowZero : idx == R + 2 ? kRowSum : kRowRead;
        }
        const PackedRow x = packed_row(kind, idx, pp, b, tab, lane);
        const uint8_t *hap = hapbuf(in_cur ? buf_cur : buf_cur ^ 1);
        const int hmax = two && nxt.H > cur.H ? nxt.H : cur.H;
        // Wave-uniform by construction; readfirstlane makes it a scalar, so the step loop stays a
        // uniform loop (a vector trip count turns it into a divergent loop with exec masking).
        const int steps = __builtin_amdgcn_readfirstlane(round_up(hmax + kWave + 2, kU));
        if (!first)
            side_pass_f32<kCols>(L.bm, L.bi, lane, rfl(x.w.MM), rfl(x.w.GM), rfl(x.w.MI), rfl(x.w.II), md63, dd63,
                                 plane_span<kCols>(steps), cd63, wend);
        first = false;
        wend = kLdsPad - (kWave - 1) + steps;
        md63 = lane63(x.w.MD);
        dd63 = lane63(x.w.DD);
        cd63 = lane63(x.kind == kRowConst ? 1.f : 0.f);
        const float dfin = (cur.hap_n || (two && nxt.hap_n)) ? packed_steps<kCols, true, kU>(L, hap, lane, x, steps)
                                                            : packed_steps<kCols, false, kU>(L, hap, lane, x, steps);
        // SUM rows that ended here: every lane takes the SUM lane's total, so the result and the
        // rescue decision are wave-uniform.  Unscaling and log10 are FP64.
        const double total = static_cast<double>(dfin);
        const bool cur_done = rows_cur <= kWave;
        const bool nxt_done = two && nxt.R + 3 <= kWave - rows_cur;
        for (int q = 0; q < 2; ++q) {
            if (q == 0 ? !cur_done : !nxt_done) continue;
            const int sl = q == 0 ? rows_cur - 1 : rows_cur + nxt.R + 2;
            const uint64_t tb = __builtin_bit_cast(uint64_t, total);
            const uint32_t lo = __builtin_amdgcn_readlane(static_cast<int>(tb), sl);
            const uint32_t hi = __builtin_amdgcn_readlane(static_cast<int>(tb >> 32), sl);
            const double sum = __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
            const uint32_t pair = q == 0 ? cur.pair : nxt.pair;
            b.out[pair] = log10(sum) - scale_log10;
            if (rs.list != nullptr && !(sum >= rs.threshold)) {
                const uint32_t slot = __builtin_amdgcn_readfirstlane(atomicAdd(rs.count, lane == 0 ? 1u : 0u));
                rs.list[slot] = pair;
            }
        }
        if (!cur_done) {
            cur_next += kWave;
        } else if (two && !nxt_done) {
            cur = nxt;
            cur_next = kWave - rows_cur;
            buf_cur ^= 1;
        } else {
            have_cur = false;
        }
    }
}

template <int kCols, int kU>
__global__ __launch_bounds__(512) void forward_kernel_packed(DeviceBatch b, DeviceTables tab, WorkSource ws,
                                                             RescueSink rs) {
    extern __shared__ __align__(16) uint8_t lds_raw[];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const uint32_t n_items = ws.n_items_dev ? *ws.n_items_dev : ws.n_items;
    packed_wave<kCols, kU>(b, tab, ws, n_items, rs, lds_raw + packed_lds_per_wave<kCols>() * wave, lane);
}

__global__ void zero_counters_kernel(uint32_t *c, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) c[i] = 0u;
}

__global__ void copy_counters_kernel(uint32_t *c, const uint32_t *init, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) c[i] = init[i];
}

}  // namespace

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
    if (cols == 0) return 0;
    if (packed && !fp64)
        return kHapGuard + 2 * cols + kPlanes * sizeof(float) * static_cast<size_t>(lds_plane_stride(static_cast<int>(cols)));
    return cols + (fp64 ? kRing64Bytes : kPlanes * sizeof(float) * static_cast<size_t>(lds_plane_stride(static_cast<int>(cols))));
}

size_t fp64_plane_bytes_per_wave(int max_hap_len) {
    const int cols = lds_columns(max_hap_len);
    return cols == 0 ? 0 : kPlanes * sizeof(double) * static_cast<size_t>(lds_plane_stride(cols));
}

uint32_t packed_chunk(int64_t n_items, int64_t waves) {
    const int64_t c = n_items / (4 * (waves > 0 ? waves : 1));
    return static_cast<uint32_t>(c < 1 ? 1 : c > kPackedChunk ? kPackedChunk : c);
}

namespace {

template <typename F>
hipError_t with_kernel(bool fp64, bool packed, int cols, int unroll, F &&f) {
    return dispatch_col_class(cols, [&](auto C) -> hipError_t {
        constexpr int kC = decltype(C)::value;
        if (fp64) return f(forward_kernel<double, kC>);
        if constexpr (kC > kMaxFp32Cols) {
            return hipErrorInvalidValue;  // FP64-only class
        } else {
            if (!packed) return f(forward_kernel<float, kC>);
            return unroll == 16 ? f(forward_kernel_packed<kC, 16>) : f(forward_kernel_packed<kC, 8>);
        }
    });
}


}  // namespace

hipError_t launch_forward(bool fp64, const DeviceBatch &batch, const DeviceTables &tables, const WorkSource &work,
                          const RescueSink &rescue, int max_hap_len, const LaunchShape &shape, int grid_blocks,
                          hipStream_t stream) {
    if (shape.waves_per_block < 1 || grid_blocks < 1) return hipErrorInvalidValue;
    // The packed schedule takes chunks from a queue with a host-known item count.
    if (shape.packed && (fp64 || work.head == nullptr || work.n_items_dev != nullptr)) return hipErrorInvalidValue;
    WorkSource ws = work;
    if (shape.packed) ws.chunk = packed_chunk(work.n_items, static_cast<int64_t>(grid_blocks) * shape.waves_per_block);
    const dim3 block(kWave * shape.waves_per_block);
    return with_kernel(fp64, shape.packed, lds_columns(max_hap_len), shape.unroll, [&](auto kernel) {
        if (shape.lds_bytes > 65536)
            (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      static_cast<int>(shape.lds_bytes));
        hipLaunchKernelGGL(kernel, dim3(grid_blocks), block, shape.lds_bytes, stream, batch, tables, ws, rescue);
        return hipGetLastError();
    });
}

hipError_t launch_zero_counters(uint32_t *counters, int n, hipStream_t stream) {
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, stream, counters, n);
    return hipGetLastError();
}

hipError_t launch_copy_counters(uint32_t *counters, const uint32_t *init, int n, hipStream_t stream) {
    hipLaunchKernelGGL(copy_counters_kernel, dim3(1), dim3(64), 0, stream, counters, init, n);
    return hipGetLastError();
}

size_t max_lds_per_block() {
    static const size_t limit = [] {
        int dev = 0, v = 0;
        size_t l = 65536;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
            l = static_cast<size_t>(v);
        return l < 163840 ? l : static_cast<size_t>(163840);
    }();
    return limit;
}

LaunchShape forward_shape(bool fp64, int max_hap_len, int requested_waves_per_block, bool packed, int unroll) {
    LaunchShape sh;
    const int cols = lds_columns(max_hap_len);
    if (cols == 0 || (!fp64 && cols > kMaxFp32Cols)) return sh;
    packed = packed && !fp64;
    const size_t per_wave = lds_bytes_per_wave(max_hap_len, fp64, packed);
    const size_t budget = max_lds_per_block();
    int wpb = requested_waves_per_block < 1 ? 1 : requested_waves_per_block;
    while (wpb > 1 && per_wave * static_cast<size_t>(wpb) > budget) --wpb;
    if (per_wave * static_cast<size_t>(wpb) > budget) return sh;
    sh.waves_per_block = wpb;
    sh.lds_bytes = per_wave * static_cast<size_t>(wpb);
    sh.packed = packed;
    auto occupancy = [&](int u) {
        int blocks = 0;
        const hipError_t e = with_kernel(fp64, packed, cols, u, [&](auto kernel) {
            return hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, kWave * wpb, sh.lds_bytes);
        });
        return (e == hipSuccess && blocks > 0) ? blocks : 1;
    };
    sh.unroll = packed && unroll == 16 ? 16 : 8;
    sh.blocks_per_cu = occupancy(sh.unroll);
    if (packed && unroll == 0) {
        // Automatic: the 16-step body issues fewer instructions per step but needs more VGPRs; it
        // wins where the LDS slice, not the registers, limits the waves per CU (long haplotypes).
        const int b16 = occupancy(16);
        if (b16 >= sh.blocks_per_cu) {
            sh.unroll = 16;
            sh.blocks_per_cu = b16;
        }
    }
    return sh;
}

}  // namespace gatk_pairhmm
