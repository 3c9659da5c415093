// pairhmm_bwd.hip -- gfx950 PairHMM forward + backward (adjoint) kernel.
//
// No counterpart in the reference: LoglessPairHMM (LoglessPairHMM.java:20-68) has no gradient.  This
// kernel differentiates the same recurrence with respect to continuous per-position Phred
// parameters: base quality qb, insertion GOP qi, deletion GOP qd and gap continuation qc, with
//   e(q) = 10^(-q/10),  MM = 1 - e(qi) - e(qd),  MI = e(qi),  MD = e(qd),  GM = 1 - e(qc),
//   II = DD = e(qc),  prior = 1 - e(qb) on a match, e(qb)/3 on a mismatch
// (MM = 1 - e1 - e2 is the closed form PairHMMModelUnitTest.java:237-286 checks the integer table
// against).  Outputs per pair: log10 L; per read position: w_p * d log10 L_p / d(qb, qi, qd, qc) and
// d / d e(qb), accumulated over the pairs that share the read.
//
// Schedule (one pair per wave, 64-row stripes with the forward kernel's anti-diagonal skew):
//   1. Forward over all stripes, top down, with lane 0 fed its sums of the row above (A, B) by a
//      side pass, saving in per-wave global scratch, for every stripe, lane 0's A / B planes and,
//      every kTile wave-steps, five values per lane (M, I, D one step back; M, S two steps back)
//      as a checkpoint.
//   2. Adjoint over the stripes bottom up.  Wave-steps run in reverse, so with the same skew lane k
//      finds lane k+1's adjoints from one and two steps earlier (DPP wave_shl:1); lane 63 takes the
//      row below the stripe (its X and gI) from LDS planes that lane 0 of the stripe underneath
//      wrote.  For each tile (right to left) the kernel restores its checkpoint (read one tile
//      ahead), recomputes kTile forward steps into registers and then runs the kTile adjoint steps
//      over them.
//   Adjoints (X = prior * gM, S = I + D):
//      gS(r,c) = GM_{r+1} X(r+1,c+1)
//      gD(r,c) = gS + DD_r gD(r,c+1)
//      gI(r,c) = gS + II_{r+1} gI(r+1,c)
//      gM(r,c) = MM_{r+1} X(r+1,c+1) + MI_{r+1} gI(r+1,c) + MD_r gD(r,c+1)
//   Row r+1's MM, GM, MI, II adjoints need row r's forward values, so lane k accumulates them for
//   the lane below and they are moved down one lane when a stripe ends.  The last-row seed
//   (d L / d M(R,c) = d L / d I(R,c) = 1) enters as a virtual row R+1 with gI = 1, X = 0 and
//   MM = GM = 0, MI = II = 1, which also makes rows beyond R pass the seed up unchanged.
//   Scaling: forward cells carry 2^120 (FP32) / 2^1020 (FP64); the FP32 adjoint seed is 2^sigma with
//   sigma = 30 - ilogb(L_scaled) so every accumulated product stays inside FP32's range; pairs whose
//   L_scaled is below the forward's rescue threshold (1e-28) go to the FP64 kernel, as in the
//   forward.  Per-lane sums are FP32 (FP64 in the FP64 kernel);
//   per-position results are scaled and accumulated in FP64 with atomics.
#include "pairhmm_device.hpp"
#include "pairhmm_kernels.hpp"

namespace gatk_pairhmm {
namespace {

using namespace dev;

constexpr int kCk = 5;  // checkpoint: M, I, D one step back and M, S two steps back

template <typename T>
struct AdjointScale;
template <>
struct AdjointScale<float> {
    // L_scaled * 2^sigma ~ 2^30.  On the likely path f * b ~ L_scaled, so the adjoint b * 2^sigma
    // is about 2^(30 - 111) where the path starts (f ~ 2^120 / H) and the seed 2^sigma at its end;
    // with the seed held below 2^126 this covers every pair down to L_scaled ~ 2^-96, i.e. every
    // pair the FP32 forward keeps (L_scaled >= 1e-28 ~ 2^-93).
    static constexpr int kTarget = 30;
    static constexpr int kMaxSigma = 126;  // seed must stay below FLT_MAX
};
template <>
struct AdjointScale<double> {
    static constexpr int kTarget = 0;
    static constexpr int kMaxSigma = 0;
};

constexpr float kLog2_10_over_10 = 0.33219280948873623479f;

__device__ __forceinline__ float phred_to_err(float q) { return exp2f(-kLog2_10_over_10 * q); }
__device__ __forceinline__ double phred_to_err(double q) { return exp2(-0.33219280948873623479 * q); }

template <typename T>
struct Row {
    T MM, GM, MI, II, MD, DD, pm, px;
    uint32_t base;
    bool n_read;
};

// Row parameters of read position p (row p + 1).  Active rows use the read; rows past the read
// (p >= R) are the pass-through rows that carry the adjoint seed (MI = II = 1, everything else 0).
template <typename T>
__device__ __forceinline__ Row<T> load_row(const GradBatch &g, int64_t ro, int R, int p, bool is_row) {
    Row<T> w{T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0), 0u, false};
    if (!is_row) return w;
    if (p >= R) {
        w.MI = T(1);
        w.II = T(1);
        return w;
    }
    const T ei = phred_to_err(static_cast<T>(g.rd_qi[ro + p]));
    const T ed = phred_to_err(static_cast<T>(g.rd_qd[ro + p]));
    const T ec = phred_to_err(static_cast<T>(g.rd_qc[ro + p]));
    const T eb = phred_to_err(static_cast<T>(g.rd_qb[ro + p]));
    w.MM = T(1) - ei - ed;
    w.MI = ei;
    w.MD = ed;
    w.GM = T(1) - ec;
    w.II = ec;
    w.DD = ec;
    w.pm = T(1) - eb;
    w.px = eb / T(3);
    w.base = g.rd_bases[ro + p];
    w.n_read = (w.base == 'N');
    return w;
}

template <typename T, int kCols>
struct GradLds {
    uint8_t *hap;
    T *bm, *bi, *bs;  // forward boundary row M, I, S by column; lane 0's sums A (over S), B (over I)
    T *bx, *bg;       // adjoint boundary: X and gI of the row below a stripe, by column.  X shares the
                      // M plane, which the adjoint no longer needs (four planes per wave, not five).
};

template <int kCols>
constexpr int grad_stride() { return lds_plane_stride(kCols); }

constexpr int kGradPlanes = 4;

template <typename T, int kCols>
constexpr size_t grad_lds_per_wave() {
    return static_cast<size_t>(kCols) + kGradPlanes * static_cast<size_t>(grad_stride<kCols>()) * sizeof(T);
}

template <typename T, int kCols>
__device__ __forceinline__ GradLds<T, kCols> grad_carve(uint8_t *base) {
    GradLds<T, kCols> w;
    w.hap = base;
    w.bm = reinterpret_cast<T *>(base + kCols);
    w.bi = w.bm + grad_stride<kCols>();
    w.bs = w.bi + grad_stride<kCols>();
    w.bg = w.bs + grad_stride<kCols>();
    w.bx = w.bm;
    return w;
}

// Lane 0's sums of the row above (A = M(c-1) MM + S(c-1) GM, B = M(c) MI + I(c) II) from the boundary
// planes, in place: A over S, B over I, highest columns first so every read of index j-1 precedes
// its overwrite.  The row above is in bm / bi / bs; MM0..II0 are lane 0's row parameters.
template <typename T, int kCols>
__device__ __forceinline__ void side_pass(const GradLds<T, kCols> &L, int lane, T MM0, T GM0, T MI0, T II0) {
    __builtin_amdgcn_wave_barrier();
    for (int j0 = kCols - kWave; j0 >= 0; j0 -= kWave) {
        const int j = j0 + lane;
        const T mj = L.bm[j], ij = L.bi[j];
        const T mp = j > 0 ? L.bm[j - 1] : T(0), sp = j > 0 ? L.bs[j - 1] : T(0);
        __builtin_amdgcn_wave_barrier();
        L.bs[j] = fma(sp, GM0, mp * MM0);
        L.bi[j] = fma(ij, II0, mj * MI0);
        __builtin_amdgcn_wave_barrier();
    }
}

// One forward step of every lane (rows of a 64-row stripe, lane k at column t - k).  Lane 0 takes
// its sums of the row above from a0 / b0; the other lanes take lane k-1's values by DPP.  Returns
// the pre-prior sum a (for the base-error adjoint); updates the rolling state.
template <typename T, bool kHapHasN>
__device__ __forceinline__ T fwd_step(const Row<T> &w, T px, uint32_t hb, T a0, T b0, bool is_first, T &m1, T &i1,
                                      T &d1, T &s1, T &m2, T &s2, T &m_out, T &i_out, T &d_out, T &s_out, T &p_out) {
    const bool match = (hb == w.base) || (kHapHasN && hb == 'N');
    const T prior = match ? w.pm : px;
    T a, i;
    if constexpr (sizeof(T) == sizeof(float)) {
        a = a0;
        i = b0;
        dpp_upper2(a, m2, w.MM, s2, w.GM, i, m1, w.MI, i1, w.II);
    } else {
        a = fma(from_prev_lane(s2), w.GM, from_prev_lane(m2) * w.MM);
        i = fma(from_prev_lane(i1), w.II, from_prev_lane(m1) * w.MI);
        a = is_first ? a0 : a;
        i = is_first ? b0 : i;
    }
    const T m = a * prior;
    // D carried divided by MD: D(c) = D(c-1) DD + M(c-1) MD becomes d(c) = d(c-1) DD + M(c-1), and
    // S = I + MD d, one FMA each (rows past R have MD = DD = 0, so d = 0 and S = I there).
    const T d = fma(d1, w.DD, m1);
    const T sx = fma(d, w.MD, i);
    m_out = m;
    i_out = i;
    d_out = d;
    s_out = sx;
    p_out = prior;
    m2 = m1;
    s2 = s1;
    m1 = m;
    i1 = i;
    d1 = d;
    s1 = sx;
    return a;
}

// One pair: forward with checkpoints, then the adjoint.  Returns nothing; writes the outputs.
//
// Stripes of 64 rows, lane k = read position 64 s + k; rows past R pass the row above down
// (MI = II = 1), in the forward and in the adjoint.  Scratch per stripe: the state registers every
// kTile steps (checkpoints) and lane 0's A / B planes of the stripe.
template <typename T, int kCols, int kTile, bool kHapHasN>
__device__ void grad_pair(const GradBatch &g, const GradLds<T, kCols> &L, T *__restrict__ scratch, int lane,
                          uint32_t pair, int H, int R, int64_t ro, double weight, const RescueSink &rs) {
    constexpr int kStripe = kWave;
    constexpr int kLag = kWave - 1;
    const int n_stripes = (R + kStripe - 1) / kStripe;
    const int steps = round_up(H + kWave, kTile);
    const int n_tiles = steps / kTile;
    T *ck = scratch;                                                        // [stripe][tile][kCk][64]
    T *abp = scratch + static_cast<size_t>(g.max_stripes) * g.max_tiles * kCk * kWave;  // [stripe][2][kCols]
    const T d0 = static_cast<T>(__builtin_ldexp(1.0, Scale<T>::kExp) / static_cast<double>(H));
    using lds_u8 = const volatile __attribute__((address_space(3))) uint8_t;
    lds_u8 *hp = (lds_u8 *)(L.hap + kLdsPad - 1 - lane);  // address-space cast: C-style only
    int lane_w = lane, lane_r = lane, lane_x = lane;
    asm volatile("" : "+v"(lane_w), "+v"(lane_r), "+v"(lane_x));
    const bool is_first = lane_r == 0;

    // ---- 1. forward, top down --------------------------------------------------------------
    for (int s = 0; s < n_stripes; ++s) {
        const Row<T> w = load_row<T>(g, ro, R, s * kStripe + lane, true);
        const T px = w.n_read ? w.pm : w.px;
        side_pass<T, kCols>(L, lane, rfl(w.MM), rfl(w.GM), rfl(w.MI), rfl(w.II));
        {
            T *dst = abp + static_cast<size_t>(s) * 2 * kCols;
            for (int j = lane; j < kCols; j += kWave) {
                dst[j] = L.bs[j];
                dst[kCols + j] = L.bi[j];
            }
        }
        T m1 = 0, i1 = 0, d1 = 0, s1 = 0, m2 = 0, s2 = 0;
        // Every lane stores its M, I and S at its own column c = t - lane at step t (no one-lane exec
        // mask in the step): lane 63 writes column c last, at step c + 63, so the planes end with its
        // row, and lane 0 has read A, B of column c by step c, before any lane overwrites it.
        using lds_t = __attribute__((address_space(3))) T;
        lds_t *wm = (lds_t *)(L.bm + kLdsPad) - lane_w, *wi = (lds_t *)(L.bi + kLdsPad) - lane_w,
              *wsp = (lds_t *)(L.bs + kLdsPad) - lane_w;
        T *ckp = ck + static_cast<size_t>(s) * g.max_tiles * kCk * kWave + lane;
        for (int t0 = 0; t0 < steps; t0 += kTile) {
            ckp[0 * kWave] = m1;  // S one step back is I + D, recomputed on restore
            ckp[1 * kWave] = i1;
            ckp[2 * kWave] = d1;
            ckp[3 * kWave] = m2;
            ckp[4 * kWave] = s2;
            ckp += kCk * kWave;
            T av[kTile], bv[kTile];
            __builtin_memcpy(av, __builtin_assume_aligned(L.bs + kLdsPad + t0, 16), sizeof(av));
            __builtin_memcpy(bv, __builtin_assume_aligned(L.bi + kLdsPad + t0, 16), sizeof(bv));
            uint32_t hbs[kTile];
#pragma unroll
            for (int u = 0; u < kTile; ++u) hbs[u] = hp[t0 + u];
            __builtin_amdgcn_wave_barrier();
            lds_t *wmb = wm + t0, *wib = wi + t0, *wsb = wsp + t0;
#pragma unroll
            for (int u = 0; u < kTile; ++u) {
                T m, i, d, sx, pr;
                fwd_step<T, kHapHasN>(w, px, hbs[u], av[u], bv[u], is_first, m1, i1, d1, s1, m2, s2, m, i, d, sx, pr);
                wmb[u] = m;
                wib[u] = i;
                wsb[u] = sx;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    double acc = 0.0;
    for (int c = 1 + lane; c <= H; c += kWave)
        acc += static_cast<double>(L.bm[kLdsPad + c]) + static_cast<double>(L.bi[kLdsPad + c]);
    const double l_scaled = wave_sum(acc);
    const double scale_log10 = static_cast<double>(Scale<T>::kExp) * 0.30102999566398119521;

    // ---- scaling of the adjoint ----------------------------------------------------------------
    int sigma = 0;
    if (AdjointScale<T>::kMaxSigma > 0) {
        sigma = AdjointScale<T>::kTarget - ilogb(l_scaled);
        const int bad = __builtin_amdgcn_readfirstlane(
            static_cast<int>(!(l_scaled > 0.0) || sigma > AdjointScale<T>::kMaxSigma || l_scaled < rs.threshold));
        if (bad) {  // hand the pair to the FP64 kernel
            if (rs.list != nullptr) {
                const uint32_t slot = __builtin_amdgcn_readfirstlane(atomicAdd(rs.count, lane == 0 ? 1u : 0u));
                rs.list[slot] = pair;
            }
            return;
        }
    }
    g.out_log10[pair] = log10(l_scaled) - scale_log10;
    if (!(l_scaled > 0.0)) return;  // FP64 underflow: log10 L = -inf, no gradient
    const T seed = static_cast<T>(__builtin_ldexp(1.0, sigma));
    const double norm = weight / (l_scaled * __builtin_ldexp(1.0, sigma) * 2.30258509299404568402);
    const double dq = -0.23025850929940456840;  // d e(q)/dq = dq * e(q)

    // Adjoint boundary below the last stripe: the seed row (gI = seed on columns 1..H, X = 0).  X
    // shares the M plane that the sum above just read.
    __builtin_amdgcn_wave_barrier();
    for (int j = lane; j < kCols; j += kWave) {
        const int c = j - kLdsPad;
        L.bx[j] = 0;
        L.bg[j] = (c >= 1 && c <= H) ? seed : T(0);
    }

    // ---- 2. adjoint, bottom up -------------------------------------------------------------------
    for (int s = n_stripes - 1; s >= 0; --s) {
        const int p = s * kStripe + lane;
        const Row<T> w = load_row<T>(g, ro, R, p, true);
        const T px = w.n_read ? w.pm : w.px;
        // Next row's MM, GM, MI, II: lane k+1's own; lane 63's is the first row of the stripe below
        // (a pass-through row below the last stripe).
        const int p_below = (s + 1) * kStripe;
        const Row<T> below = load_row<T>(g, ro, R, p_below, true);
        const T MMn = wave_shl1(below.MM, w.MM), GMn = wave_shl1(below.GM, w.GM);
        const T MIn = wave_shl1(below.MI, w.MI), IIn = wave_shl1(below.II, w.II);
        // Lane 0's A / B planes of this stripe, for the recompute.
        {
            const T *src = abp + static_cast<size_t>(s) * 2 * kCols;
            __builtin_amdgcn_wave_barrier();
            for (int j = lane; j < kCols; j += kWave) {
                L.bs[j] = src[j];
                L.bi[j] = src[kCols + j];
            }
            __builtin_amdgcn_wave_barrier();
        }
        // Lane 63 reads X(r+1, c+1) and gI(r+1, c) of the row below; lane 0 leaves X and gI of this
        // stripe's first row for the stripe above (column c = t at step t).  Steps run in reverse, and
        // every lane stores at its own column c = t - lane (no one-lane exec mask in the step): a
        // column's stores come after lane 63 has read it, and lane 0's, at step c, come last.  Lane
        // 63's X store alone would land before its own read of that column one step later; it goes
        // one column right, onto the column it has just read.
        // Volatile: single-dword LDS accesses at immediate offsets (paired ds_read2 / ds_write2 forms
        // cannot reach the planes' offsets and cost an address add each).
        using lds_t = volatile __attribute__((address_space(3))) T;
        const lds_t *bxr = (const lds_t *)(L.bx + kLdsPad - (kLag - 1)), *bgr = (const lds_t *)(L.bg + kLdsPad - kLag);
        lds_t *bxw = (lds_t *)(L.bx + kLdsPad) - lane_x + (lane_x == kLag ? 1 : 0);
        lds_t *bgw = (lds_t *)(L.bg + kLdsPad) - lane_x;
        T X1 = 0, X2 = 0, gI1 = 0, gD1 = 0;
        T aMMn = 0, aGMn = 0, aMIn = 0, aIIn = 0, aMD = 0, aDD = 0, aE = 0;
        const T *ckbase = ck + static_cast<size_t>(s) * g.max_tiles * kCk * kWave + lane;
        // Checkpoints are read one tile ahead so the global load overlaps the current tile.
        T cur[kCk], nxt[kCk];
#pragma unroll
        for (int k = 0; k < kCk; ++k) cur[k] = ckbase[(static_cast<size_t>(n_tiles - 1) * kCk + k) * kWave];
        for (int tile = n_tiles - 1; tile >= 0; --tile) {
            const int t0 = tile * kTile;
            const int prev = tile > 0 ? tile - 1 : 0;
#pragma unroll
            for (int k = 0; k < kCk; ++k) nxt[k] = ckbase[(static_cast<size_t>(prev) * kCk + k) * kWave];
            T m1 = cur[0], i1 = cur[1], d1 = cur[2], m2 = cur[3], s2 = cur[4];
            T s1 = fma(d1, w.MD, i1);
            const T m_before = m1, d_before = d1;
            T av[kTile], bv[kTile];
            __builtin_memcpy(av, __builtin_assume_aligned(L.bs + kLdsPad + t0, 16), sizeof(av));
            __builtin_memcpy(bv, __builtin_assume_aligned(L.bi + kLdsPad + t0, 16), sizeof(bv));
            uint32_t hbs[kTile];
#pragma unroll
            for (int u = 0; u < kTile; ++u) hbs[u] = hp[t0 + u];
            // Lane 63's row-below values of the tile, read at its start: column c of the X plane is
            // stored by lane k at step c + k (lane 63 at step c + 62, after its read there), and of
            // the gI plane at step c + k, so no store of the tile reaches a column before the step
            // that reads it (steps run t0 + kTile - 1 down to t0).
            T xv[kTile], gv[kTile];
#pragma unroll
            for (int u = 0; u < kTile; ++u) {
                xv[u] = bxr[t0 + u];
                gv[u] = bgr[t0 + u];
            }
            __builtin_amdgcn_wave_barrier();
            T stM[kTile], stI[kTile], stD[kTile], stS[kTile], stP[kTile], stAW[kTile];
#pragma unroll
            for (int u = 0; u < kTile; ++u) {
                const T a = fwd_step<T, kHapHasN>(w, px, hbs[u], av[u], bv[u], is_first, m1, i1, d1, s1, m2, s2,
                                                  stM[u], stI[u], stD[u], stS[u], stP[u]);
                const bool hit = (hbs[u] == w.base) || (kHapHasN && hbs[u] == 'N');
                const T wgt = (hit || w.n_read) ? T(-1) : T(1) / T(3);  // d prior / d e(qb)
                stAW[u] = a * wgt;
            }
#pragma unroll
            for (int u = kTile - 1; u >= 0; --u) {
                const int t = t0 + u;
                const T Xn2 = wave_shl1(xv[u], X2);   // X(r+1, c+1)
                const T gIn1 = wave_shl1(gv[u], gI1); // gI(r+1, c)
                const T gS = GMn * Xn2;
                const T gD = fma(w.DD, gD1, gS);
                const T gI = fma(IIn, gIn1, gS);
                const T gM = fma(w.MD, gD1, fma(MIn, gIn1, MMn * Xn2));
                const T X = stP[u] * gM;
                const T Mp = (u > 0) ? stM[u - 1] : m_before;
                const T Dp = (u > 0) ? stD[u - 1] : d_before;
                aMMn = fma(Xn2, stM[u], aMMn);
                aGMn = fma(Xn2, stS[u], aGMn);
                aMIn = fma(gIn1, stM[u], aMIn);
                aIIn = fma(gIn1, stI[u], aIIn);
                aMD = fma(gD, Mp, aMD);
                aDD = fma(gD, Dp, aDD);
                aE = fma(gM, stAW[u], aE);
                bxw[t] = X;
                bgw[t] = gI;
                X2 = X1;
                X1 = X;
                gI1 = gI;
                gD1 = gD;
            }
#pragma unroll
            for (int k = 0; k < kCk; ++k) cur[k] = nxt[k];
        }
        __builtin_amdgcn_wave_barrier();
        // The next-row sums belong to the lane below: shift them down one lane.  Lane 63's belong to
        // the first row of the stripe below; lane 0's own next-row terms come from the stripe above
        // (or, for row 1, from row 0: S(0, c) = d0, so d L / d GM_1 = d0 sum_c X(1, c)).
        const double gMM = static_cast<double>(wave_shr1(T(0), aMMn));
        const double gGM = static_cast<double>(wave_shr1(T(0), aGMn));
        const double gMI = static_cast<double>(wave_shr1(T(0), aMIn));
        const double gII = static_cast<double>(wave_shr1(T(0), aIIn));
        if (p < R) {
            const double ei = static_cast<double>(w.MI), ed = static_cast<double>(w.MD), ec = static_cast<double>(w.II);
            const double eb = 1.0 - static_cast<double>(w.pm);
            const double d_ei = (-gMM + gMI) * norm, d_ed = (-gMM + static_cast<double>(aMD)) * norm;
            // aDD sums gD times D / MD (the forward's scaled D): d L / d DD = MD aDD.
            const double d_ec = (-gGM + gII + static_cast<double>(aDD) * ed) * norm;
            const double d_eb = static_cast<double>(aE) * norm;
            atomicAdd(&g.g_qi[ro + p], d_ei * dq * ei);
            atomicAdd(&g.g_qd[ro + p], d_ed * dq * ed);
            atomicAdd(&g.g_qc[ro + p], d_ec * dq * ec);
            atomicAdd(&g.g_qb[ro + p], d_eb * dq * eb);
            if (g.g_eb) atomicAdd(&g.g_eb[ro + p], d_eb);
        }
        if (lane == kLag && p_below < R) {
            const double ei = static_cast<double>(below.MI), ed = static_cast<double>(below.MD);
            const double ec = static_cast<double>(below.II);
            const double mmn = static_cast<double>(aMMn);
            atomicAdd(&g.g_qi[ro + p_below], (-mmn + static_cast<double>(aMIn)) * norm * dq * ei);
            atomicAdd(&g.g_qd[ro + p_below], -mmn * norm * dq * ed);
            atomicAdd(&g.g_qc[ro + p_below], (-static_cast<double>(aGMn) + static_cast<double>(aIIn)) * norm * dq * ec);
        }
        if (s == 0) {
            double sx = 0.0;
            for (int c = 1 + lane; c <= H; c += kWave) sx += static_cast<double>(L.bx[kLdsPad + c]);
            sx = wave_sum(sx);
            if (lane == 0) {
                const double ec = static_cast<double>(w.II);
                atomicAdd(&g.g_qc[ro], -static_cast<double>(d0) * sx * norm * dq * ec);
            }
        }
    }
}

template <typename T, int kCols, int kTile>
__device__ void grad_process(const GradBatch &g, const RescueSink &rs, uint8_t *lds, T *scratch, int lane, uint32_t pair) {
    const int r = g.pair_read[pair], h = g.pair_hap[pair];
    const int H = g.hap_len[h], R = g.rd_len[r];
    const int64_t ro = g.rd_off[r];
    const uint8_t *hap = g.hap_bases + g.hap_off[h];
    const double weight = g.pair_weight ? g.pair_weight[pair] : 1.0;
    GradLds<T, kCols> L = grad_carve<T, kCols>(lds);
    // Guards: the scratch and LDS were sized from the caller's max_read_len / max_hap_len; a pair
    // outside those bounds (or with an empty haplotype) gets NaN instead of touching memory.
    const int n_stripes = (R + kWave - 1) / kWave;
    const int n_tiles = round_up(H + kWave, kTile) / kTile;
    if (H <= 0 || H + 2 * kLdsPad + 32 > kCols || n_stripes > g.max_stripes || n_tiles > g.max_tiles) {
        g.out_log10[pair] = __builtin_nan("");
        return;
    }
    int has_n = 0;
    for (int j = lane; j < kCols; j += kWave) {
        const int c = j - kLdsPad + 1;
        uint8_t v = 0;
        if (c >= 1 && c <= H) {
            v = hap[c - 1];
            has_n |= (v == 'N');
        }
        L.hap[j] = v;
    }
    const T d0 = static_cast<T>(__builtin_ldexp(1.0, Scale<T>::kExp) / static_cast<double>(H));
    for (int j = lane; j < kCols; j += kWave) {
        L.bm[j] = 0;
        L.bi[j] = 0;
        L.bs[j] = j >= kLdsPad ? d0 : T(0);
    }
    if (R == 0) {  // empty read: log10 L = -inf, nothing to differentiate (LoglessPairHMM.java:62-67)
        g.out_log10[pair] = -__builtin_inf();
        return;
    }
    if (__any(has_n))
        grad_pair<T, kCols, kTile, true>(g, L, scratch, lane, pair, H, R, ro, weight, rs);
    else
        grad_pair<T, kCols, kTile, false>(g, L, scratch, lane, pair, H, R, ro, weight, rs);
}

template <typename T, int kCols, int kTile>
__device__ __forceinline__ void grad_loop(const GradBatch &g, const WorkSource &ws, const RescueSink &rs) {
    extern __shared__ __align__(16) uint8_t lds_raw[];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int waves_per_block = blockDim.x / kWave;
    uint8_t *lds = lds_raw + grad_lds_per_wave<T, kCols>() * wave;
    const size_t gwave = static_cast<size_t>(blockIdx.x) * waves_per_block + wave;
    T *scratch = reinterpret_cast<T *>(g.scratch) + gwave * g.scratch_per_wave;
    const uint32_t n_items = ws.n_items_dev ? *ws.n_items_dev : ws.n_items;
    for (;;) {
        const uint32_t it = __builtin_amdgcn_readfirstlane(atomicAdd(ws.head, lane == 0 ? 1u : 0u));
        if (it >= n_items) break;
        grad_process<T, kCols, kTile>(g, rs, lds, scratch, lane, ws.items ? ws.items[it] : it);
    }
}

// FP32: held to 168 VGPRs (three waves per SIMD; the few values the allocator then keeps in scratch
// are reloaded once per stripe, outside the step loops).
template <int kCols, int kTile>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void grad_kernel_f32(GradBatch g, WorkSource ws,
                                                                                              RescueSink rs) {
    grad_loop<float, kCols, kTile>(g, ws, rs);
}

template <int kCols, int kTile>
__global__ __launch_bounds__(256) void grad_kernel_f64(GradBatch g, WorkSource ws, RescueSink rs) {
    grad_loop<double, kCols, kTile>(g, ws, rs);
}

// LDS column class of a haplotype length (the same rule as lds_class_index), capped at max_class.
__device__ __forceinline__ int hap_class(int H, int max_class) {
    const int need = H + 2 * kLdsPad + 32;
    int k = max_class;
#pragma unroll
    for (int c = kNumColClasses - 1; c >= 0; --c)
        if (kColClasses[c] >= need) k = c;
    return k < max_class ? k : max_class;
}

__global__ __launch_bounds__(256) void grad_class_count_kernel(const int32_t *pair_hap, const int32_t *hap_len,
                                                               uint32_t n_pairs, int max_class, uint32_t *counts) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & (kWave - 1);
    const int k = p < n_pairs ? hap_class(hap_len[pair_hap[p]], max_class) : -1;
#pragma unroll
    for (int c = 0; c < kNumColClasses; ++c) {  // one atomic per class present in the wave
        const uint64_t m = __ballot(k == c);
        if (m != 0 && lane == __builtin_ctzll(m)) atomicAdd(&counts[c], static_cast<uint32_t>(__builtin_popcountll(m)));
    }
}

__global__ __launch_bounds__(256) void grad_class_scatter_kernel(const int32_t *pair_hap, const int32_t *hap_len,
                                                                 uint32_t n_pairs, int max_class, ClassRuns runs,
                                                                 uint32_t *cursors, uint32_t *order) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & (kWave - 1);
    const int k = p < n_pairs ? hap_class(hap_len[pair_hap[p]], max_class) : -1;
    int run = -1;
#pragma unroll
    for (int c = 0; c < kNumColClasses; ++c)
        if (k == c) run = runs.run_of_class[c];
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (kWave - lane));  // lanes < this one
#pragma unroll
    for (int r = 0; r < kNumColClasses; ++r) {
        const uint64_t m = __ballot(run == r);
        if (m == 0) continue;
        const int leader = __builtin_ctzll(m);
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&cursors[r], static_cast<uint32_t>(__builtin_popcountll(m)));
        base = __shfl(base, leader, kWave);
        if (run == r) order[runs.run_begin[r] + base + static_cast<uint32_t>(__builtin_popcountll(m & below))] = p;
    }
}

}  // namespace

size_t grad_lds_bytes_per_wave(int max_hap_len, bool fp64) {
    const size_t cols = static_cast<size_t>(lds_columns(max_hap_len));
    return cols + kGradPlanes * static_cast<size_t>(lds_plane_stride(static_cast<int>(cols))) *
                      (fp64 ? sizeof(double) : sizeof(float));
}

size_t grad_scratch_elems_per_wave(int max_read_len, int max_hap_len, int *max_stripes, int *max_tiles) {
    const int cols = lds_columns(max_hap_len);
    *max_stripes = (max_read_len + kWave - 1) / kWave;
    if (*max_stripes < 1) *max_stripes = 1;
    *max_tiles = dev::round_up(max_hap_len + kWave, kGradTile) / kGradTile;
    return static_cast<size_t>(*max_stripes) * (static_cast<size_t>(*max_tiles) * kCk * kWave + 2 * static_cast<size_t>(cols));
}

namespace {

constexpr int kMaxGradCols = 2560;  // longer haplotypes exceed the adjoint's LDS and scratch plan

template <typename F>
hipError_t with_grad_kernel(bool fp64, int cols, F &&f) {
    return dispatch_col_class(cols, [&](auto C) -> hipError_t {
        constexpr int kC = decltype(C)::value;
        if constexpr (kC > kMaxGradCols) {
            return hipErrorInvalidValue;
        } else {
            return fp64 ? f(grad_kernel_f64<kC, kGradTile>) : f(grad_kernel_f32<kC, kGradTile>);
        }
    });
}


}  // namespace

LaunchShape grad_shape(bool fp64, int max_hap_len, int requested_waves_per_block) {
    LaunchShape sh;
    const int cols = lds_columns(max_hap_len);
    if (cols == 0 || cols > kMaxGradCols) return sh;
    const size_t per_wave = grad_lds_bytes_per_wave(max_hap_len, fp64);
    const size_t budget = max_lds_per_block();
    int wpb = requested_waves_per_block < 1 ? 1 : requested_waves_per_block;
    while (wpb > 1 && per_wave * static_cast<size_t>(wpb) > budget) --wpb;
    if (per_wave * static_cast<size_t>(wpb) > budget) return sh;
    sh.waves_per_block = wpb;
    sh.lds_bytes = per_wave * static_cast<size_t>(wpb);
    int blocks = 0;
    hipError_t e = with_grad_kernel(fp64, cols, [&](auto kernel) {
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, kWave * wpb, sh.lds_bytes);
    });
    sh.blocks_per_cu = (e == hipSuccess && blocks > 0) ? blocks : 1;
    return sh;
}

hipError_t launch_grad_class_count(const int32_t *pair_hap, const int32_t *hap_len, uint32_t n_pairs, int max_class,
                                   uint32_t *counts, hipStream_t stream) {
    if (n_pairs == 0) return hipSuccess;
    if (max_class < 0 || max_class >= kNumColClasses) return hipErrorInvalidValue;
    hipLaunchKernelGGL(grad_class_count_kernel, dim3((n_pairs + 255) / 256), dim3(256), 0, stream, pair_hap, hap_len,
                       n_pairs, max_class, counts);
    return hipGetLastError();
}

hipError_t launch_grad_class_scatter(const int32_t *pair_hap, const int32_t *hap_len, uint32_t n_pairs, int max_class,
                                     const ClassRuns &runs, uint32_t *cursors, uint32_t *order, hipStream_t stream) {
    if (n_pairs == 0) return hipSuccess;
    if (max_class < 0 || max_class >= kNumColClasses) return hipErrorInvalidValue;
    hipLaunchKernelGGL(grad_class_scatter_kernel, dim3((n_pairs + 255) / 256), dim3(256), 0, stream, pair_hap, hap_len,
                       n_pairs, max_class, runs, cursors, order);
    return hipGetLastError();
}

hipError_t launch_grad(bool fp64, const GradBatch &g, const WorkSource &work, const RescueSink &rescue, int max_hap_len,
                       const LaunchShape &shape, int grid_blocks, hipStream_t stream) {
    if (shape.waves_per_block < 1 || grid_blocks < 1 || work.head == nullptr) return hipErrorInvalidValue;
    const dim3 block(kWave * shape.waves_per_block);
    return with_grad_kernel(fp64, lds_columns(max_hap_len), [&](auto kernel) {
        if (shape.lds_bytes > 65536)
            (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      static_cast<int>(shape.lds_bytes));
        hipLaunchKernelGGL(kernel, dim3(grid_blocks), block, shape.lds_bytes, stream, g, work, rescue);
        return hipGetLastError();
    });
}

}  // namespace gatk_pairhmm
