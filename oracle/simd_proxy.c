/*
 * simd_proxy.c -- a CPU SIMD PairHMM in the manner of GKL's AVX path, timed as a "GKL-class proxy"
 * beside the FP64 port (TEST / BENCH INFRASTRUCTURE ONLY: bench.py's CPU leg calls it; the product
 * never does).
 *
 * GKL (com.intel.gkl 0.9.0) is not in this image, so its speed cannot be measured here.  This proxy
 * follows the public description of its numerics -- an FP32 pass with initial constant 2^120 and an
 * FP64 recompute of pairs whose FP32 last-row sum is below 1e-28 -- with a vectorisation of its own:
 * one SIMD lane per read, 16 reads (AVX-512) or 8 (AVX2) of the same haplotype at a time, rows in
 * lock step, every dependency lane-local.  Reads of a haplotype are sorted by length so a group's
 * lanes finish close together.  The FP32 pass runs with flush-to-zero / denormals-are-zero.  The
 * FP64 recompute is oracle_logless (pairhmm_oracle.c), with the caller's floating-point mode.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <xmmintrin.h>

int oracle_logless(const uint8_t *hap, int H, const uint8_t *read, const uint8_t *quals, const uint8_t *ins,
                   const uint8_t *del, const uint8_t *gcp, int R, int tristate, double *out);
double oracle_qual_to_error_prob(int q);
void oracle_init(void);
double oracle_match_to_match_prob(int ins, int del);

#define LANES 16
typedef float vf __attribute__((vector_size(LANES * 4)));
typedef int32_t vi __attribute__((vector_size(LANES * 4)));

/* One group: lanes = reads of one haplotype; out[k] receives log10 L of lane k (FP32 pass) and
 * underflow[k] = 1 where the scaled sum is below the rescue threshold. */
__attribute__((target_clones("avx512f", "avx2", "default")))
static void run_group(const uint8_t *hap, int H, int n, const uint8_t *const *rb, const uint8_t *const *rq,
                      const uint8_t *const *ri, const uint8_t *const *rd, const uint8_t *const *rg, const int *R,
                      int tristate, double *out, int *underflow, float *work, const double *err,
                      const float *mm) {
    int Rmax = 0;
    for (int k = 0; k < n; k++) Rmax = R[k] > Rmax ? R[k] : Rmax;
    vf *Mp = (vf *)work, *Ip = Mp + (H + 1), *Dp = Ip + (H + 1), *Mc = Dp + (H + 1), *Ic = Mc + (H + 1),
       *Dc = Ic + (H + 1);
    const float d0 = ldexpf(1.0f, 120) / (float)H;
    for (int j = 0; j <= H; j++) {
        Mp[j] = (vf){0};
        Ip[j] = (vf){0};
        for (int k = 0; k < LANES; k++) Dp[j][k] = d0;
    }
    vf acc_final = (vf){0};
    for (int i = 1; i <= Rmax; i++) {
        vf MM, GM, MI, II, MD, DD, pm, px;
        vi base;
        for (int k = 0; k < LANES; k++) {
            if (k < n && i <= R[k]) {
                const int p = i - 1;
                const int qi = ri[k][p] & 127, qd = rd[k][p] & 127, qc = rg[k][p], q = rq[k][p];
                MM[k] = mm[qi * 128 + qd];
                MI[k] = (float)err[qi];
                MD[k] = (float)err[qd];
                II[k] = (float)err[qc];
                DD[k] = II[k];
                GM[k] = 1.0f - II[k];
                const double e = err[q];
                pm[k] = (float)(1.0 - e);
                px[k] = (float)(tristate ? e / 3.0 : e);
                base[k] = rb[k][p];
                if (base[k] == 'N') px[k] = pm[k];
            } else {
                MM[k] = GM[k] = MI[k] = II[k] = MD[k] = DD[k] = pm[k] = px[k] = 0.0f;
                base[k] = -1;
            }
        }
        Mc[0] = (vf){0};
        Ic[0] = (vf){0};
        Dc[0] = (vf){0};
        vf acc = (vf){0};
        for (int j = 1; j <= H; j++) {
            const int32_t y = hap[j - 1];
            const vi yv = (vi){0} + y;
            const vi eq = (base == yv) | ((vi){0} + (y == 'N' ? -1 : 0));
            const vf prior = (vf)(((vi)pm & eq) | ((vi)px & ~eq));
            const vf m = prior * (Mp[j - 1] * MM + (Ip[j - 1] + Dp[j - 1]) * GM);
            const vf ii = Mp[j] * MI + Ip[j] * II;
            const vf d = Mc[j - 1] * MD + Dc[j - 1] * DD;
            Mc[j] = m;
            Ic[j] = ii;
            Dc[j] = d;
            acc += m + ii;
        }
        for (int k = 0; k < n; k++)
            if (R[k] == i) acc_final[k] = acc[k];
        vf *t;
        t = Mp; Mp = Mc; Mc = t;
        t = Ip; Ip = Ic; Ic = t;
        t = Dp; Dp = Dc; Dc = t;
    }
    for (int k = 0; k < n; k++) {
        const double s = (double)acc_final[k];
        underflow[k] = !(s >= 1e-28);
        out[k] = log10(s) - 120.0 * log10(2.0);
    }
}

static int cmp_len_desc(const void *a, const void *b, void *ctx) {
    const int32_t *len = (const int32_t *)ctx;
    const int ia = *(const int *)a, ib = *(const int *)b;
    return len[ib] - len[ia];
}

/* Every pair of a packed batch; pairs are grouped by haplotype (pair_hap), the reads of a haplotype
 * sorted by length and taken LANES at a time.  out is indexed by pair.  Returns the number of pairs
 * recomputed in FP64 (-1 if out of memory). */
long oracle_simd_proxy_batch(int64_t n_pairs, const int32_t *pair_read, const int32_t *pair_hap,
                             const uint8_t *read_bases, const uint8_t *read_quals, const uint8_t *read_ins,
                             const uint8_t *read_del, const uint8_t *read_gcp, const int64_t *read_off,
                             const int32_t *read_len, const uint8_t *hap_bases, const int64_t *hap_off,
                             const int32_t *hap_len, int32_t n_haps, int tristate, int n_threads, double *out) {
    oracle_init();
    /* Phred tables once per call (insertion / deletion bytes are below 128: larger ones are rejected) */
    double err[256];
    float *mm = (float *)malloc(sizeof(float) * 128 * 128);
    if (!mm) return -1;
    for (int q = 0; q < 256; q++) err[q] = oracle_qual_to_error_prob(q);
    for (int a = 0; a < 128; a++)
        for (int c = 0; c < 128; c++) mm[a * 128 + c] = (float)oracle_match_to_match_prob(a, c);
    /* bucket pairs by haplotype */
    int64_t *start = (int64_t *)calloc((size_t)n_haps + 1, sizeof(int64_t));
    int64_t *pairs = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n_pairs > 0 ? n_pairs : 1));
    if (!start || !pairs) {
        free(start);
        free(pairs);
        free(mm);
        return -1;
    }
    for (int64_t p = 0; p < n_pairs; p++) start[pair_hap[p] + 1]++;
    for (int h = 0; h < n_haps; h++) start[h + 1] += start[h];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n_haps + 1));
    memcpy(fill, start, sizeof(int64_t) * (size_t)(n_haps + 1));
    for (int64_t p = 0; p < n_pairs; p++) pairs[fill[pair_hap[p]]++] = p;
    free(fill);
    long rescued = 0;
#pragma omp parallel num_threads(n_threads > 0 ? n_threads : 1) reduction(+ : rescued)
    {
        /* Flush-to-zero and denormals-are-zero for the FP32 pass, as SIMD PairHMMs run (x86 slows down
         * by orders of magnitude on denormal operands); restored before the thread leaves. */
        const unsigned int csr = _mm_getcsr();
        _mm_setcsr(csr | 0x8040u);
        float *work = NULL;
        size_t work_cap = 0;
#pragma omp for schedule(dynamic, 1)
        for (int h = 0; h < n_haps; h++) {
            const int64_t cnt = start[h + 1] - start[h];
            if (cnt == 0) continue;
            const int H = hap_len[h];
            const uint8_t *hap = hap_bases + hap_off[h];
            const size_t need = (size_t)6 * (size_t)(H + 1) * LANES;
            if (need > work_cap) {
                free(work);
                work = (float *)aligned_alloc(64, need * sizeof(float));
                work_cap = need;
            }
            int *idx = (int *)malloc(sizeof(int) * (size_t)cnt);
            int32_t *lens = (int32_t *)malloc(sizeof(int32_t) * (size_t)cnt);
            for (int64_t k = 0; k < cnt; k++) {
                idx[k] = (int)k;
                lens[k] = read_len[pair_read[pairs[start[h] + k]]];
            }
            qsort_r(idx, (size_t)cnt, sizeof(int), cmp_len_desc, lens);
            for (int64_t g = 0; g < cnt; g += LANES) {
                const int n = (int)(cnt - g < LANES ? cnt - g : LANES);
                const uint8_t *rb[LANES], *rq[LANES], *ri[LANES], *rd[LANES], *rg[LANES];
                int R[LANES];
                int64_t pp[LANES];
                double o[LANES];
                int uf[LANES];
                for (int k = 0; k < n; k++) {
                    pp[k] = pairs[start[h] + idx[g + k]];
                    const int r = pair_read[pp[k]];
                    const int64_t ro = read_off[r];
                    rb[k] = read_bases + ro;
                    rq[k] = read_quals + ro;
                    ri[k] = read_ins + ro;
                    rd[k] = read_del + ro;
                    rg[k] = read_gcp + ro;
                    R[k] = read_len[r];
                }
                run_group(hap, H, n, rb, rq, ri, rd, rg, R, tristate, o, uf, work, err, mm);
                for (int k = 0; k < n; k++) {
                    if (uf[k] && R[k] > 0) {
                        oracle_logless(hap, H, rb[k], rq[k], ri[k], rd[k], rg[k], R[k], tristate, &o[k]);
                        rescued++;
                    }
                    out[pp[k]] = o[k];
                }
            }
            free(idx);
            free(lens);
        }
        free(work);
        _mm_setcsr(csr);
    }
    free(start);
    free(pairs);
    free(mm);
    return rescued;
}
