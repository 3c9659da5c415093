/*
 * pairhmm_oracle.c -- FP64 CPU restatement of GATK's PairHMM model.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the gfx950 PairHMM engine in
 * gatk_amd/.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
 * load it, and only to check or to time the CPU path.  The product never links it.
 *
 * What is restated (paths relative to src/main/java/org/broadinstitute/hellbender/):
 *   - QualityUtils.qualToErrorProb / qualToProb / qualToProbLog10 caches
 *       (utils/QualityUtils.java:43-57, 83-85, 138-141, 156-158)
 *   - MathUtils.JacobianLogTable, fastRound, approximateLog10SumLog10(a,b),
 *       approximateLog10SumLog10(double[]), log10SumLog10(double[]), maxElementIndex
 *       (utils/MathUtils.java:406-430, 444-480, 622-650, 786-797)
 *   - PairHMMModel matchToMatch tables and qualToTransProbs[Log10]
 *       (utils/pairhmm/PairHMMModel.java:71-94, 107-117, 214-225, 373-418)
 *   - LoglessPairHMM.subComputeReadLikelihoodGivenHaplotypeLog10 + initializePriors
 *       (utils/pairhmm/LoglessPairHMM.java:8-12, 20-68, 79-93)
 *   - Log10PairHMM (EXACT and ORIGINAL) (utils/pairhmm/Log10PairHMM.java:44-183)
 *   - PairHMM.computeReadLikelihoodGivenHaplotypeLog10 argument checks
 *       (utils/pairhmm/PairHMM.java:284-306)
 *   - PairHMM.findFirstPositionWhereHaplotypesDiffer (utils/pairhmm/PairHMM.java:341-352)
   - LoglessPDPairHMM (partially determined haplotypes: SNP alternatives and optional deletions)
       (utils/pairhmm/LoglessPDPairHMM.java:13-31, 34-150, 158-194), with hapStartIndex 0 as
       PDPairHMM.java:281-288 fixes it.  The reference's own fixture for it
       (src/test/resources/large/expected.PDHMM.hmmresults.txt) is a git-lfs pointer in this
       checkout, so the PD restatement is pinned by reductions instead (tests/test_pd_oracle.py):
       no PD bits equals LoglessPairHMM bit for bit, SNP alternatives equal a substituted
       haplotype, a deletion span reproduces the likelihood of the haplotype with the span removed,
       and every PD cell dominates its LoglessPairHMM cell.  The I rule at a deletion end
       (:79-84) moves results only through second-order paths; it is parity unpinned.
 * and, new (the reference has no gradient): an FP64 adjoint of the LoglessPairHMM recurrence
 * with continuous per-position Phred parameters.
 *
 * Arithmetic follows Java's strict IEEE-754 double evaluation: build with -ffp-contract=off and
 * without -ffast-math so that every product and sum is rounded where Java rounds it.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_OK 0
#define ORACLE_BAD_ARG (-1)
#define ORACLE_OOM (-3)

#define MAX_QUAL 254                       /* QualityUtils.java:43 */
#define MM_TABLE_SIZE (((MAX_QUAL + 1) * (MAX_QUAL + 2)) >> 1) /* PairHMMModel.java:71 */
#define JACOBIAN_MAX_TOLERANCE 8.0         /* MathUtils.java:409 */
#define JACOBIAN_TABLE_STEP 0.0001         /* MathUtils.java:414 */

static double g_qual_err[MAX_QUAL + 1];        /* QualityUtils.qualToErrorProbCache */
static double g_qual_prob_log10[MAX_QUAL + 1]; /* QualityUtils.qualToProbLog10Cache */
static double *g_jacobian;                     /* MathUtils.JacobianLogTable.cache */
static int g_jacobian_len;
static double g_jacobian_inv_step;
static double g_mm_prob[MM_TABLE_SIZE];        /* PairHMMModel.matchToMatchProb */
static double g_mm_log10[MM_TABLE_SIZE];       /* PairHMMModel.matchToMatchLog10 */
static double g_inv_ln10;
static double g_initial_condition;             /* LoglessPairHMM.java:8 */
static double g_initial_condition_log10;       /* LoglessPairHMM.java:9 */
static double g_log10_3;                       /* Log10PairHMM.java:21 */
static pthread_once_t g_once = PTHREAD_ONCE_INIT;

/* MathUtils.fastRound (MathUtils.java:428-430) */
static int fast_round(double d) { return (d > 0.0) ? (int)(d + 0.5) : (int)(d - 0.5); }

static double jacobian_get(double difference) {
    return g_jacobian[fast_round(difference * g_jacobian_inv_step)];
}

/* MathUtils.approximateLog10SumLog10(double, double) (MathUtils.java:467-480) */
double oracle_approx_log10_sum_log10(double a, double b) {
    if (a > b) {
        double t = a;
        a = b;
        b = t;
    } else if (a == -INFINITY) {
        return b;
    }
    const double diff = b - a;
    return b + (diff < JACOBIAN_MAX_TOLERANCE ? jacobian_get(diff) : 0.0);
}

static void init_tables(void) {
    /* QualityUtils static block (QualityUtils.java:52-57) */
    for (int i = 0; i <= MAX_QUAL; i++) {
        g_qual_err[i] = pow(10.0, (double)i / -10.0);
        g_qual_prob_log10[i] = log10(1.0 - g_qual_err[i]);
    }
    /* JacobianLogTable (MathUtils.java:414-418): k in [0, (int)(MAX_TOL / STEP) + 1) */
    g_jacobian_len = (int)(JACOBIAN_MAX_TOLERANCE / JACOBIAN_TABLE_STEP) + 1;
    g_jacobian = (double *)malloc(sizeof(double) * (size_t)g_jacobian_len);
    for (int k = 0; k < g_jacobian_len; k++)
        g_jacobian[k] = log10(1.0 + pow(10.0, (double)(-k) * JACOBIAN_TABLE_STEP));
    g_jacobian_inv_step = 1.0 / JACOBIAN_TABLE_STEP;
    g_inv_ln10 = 1.0 / log(10.0);
    /* PairHMMModel static block (PairHMMModel.java:86-94): i is the larger qual, j the smaller */
    for (int i = 0, offset = 0; i <= MAX_QUAL; offset += ++i) {
        for (int j = 0; j <= i; j++) {
            const double log10_sum = oracle_approx_log10_sum_log10(-0.1 * (double)i, -0.1 * (double)j);
            const double p = pow(10.0, log10_sum);
            g_mm_log10[offset + j] = log1p(-(p < 1.0 ? p : 1.0)) * g_inv_ln10;
            g_mm_prob[offset + j] = pow(10.0, g_mm_log10[offset + j]);
        }
    }
    g_initial_condition = pow(2.0, 1020.0);
    g_initial_condition_log10 = log10(g_initial_condition);
    g_log10_3 = log10(3.0);
}

void oracle_init(void) { pthread_once(&g_once, init_tables); }

/* QualityUtils.qualToErrorProb(byte): cache[qual & 0xff]; index 255 is out of the 255-entry
 * cache in Java (ArrayIndexOutOfBounds), reported here as NaN. */
double oracle_qual_to_error_prob(int q) {
    oracle_init();
    q &= 0xff;
    return q <= MAX_QUAL ? g_qual_err[q] : NAN;
}
double oracle_qual_to_prob(int q) { return 1.0 - oracle_qual_to_error_prob(q); }
double oracle_qual_to_prob_log10(int q) {
    oracle_init();
    q &= 0xff;
    return q <= MAX_QUAL ? g_qual_prob_log10[q] : NAN;
}
double oracle_jacobian_table_get(double diff) {
    oracle_init();
    return jacobian_get(diff);
}

/* PairHMMModel.matchToMatchProb(int, int) (PairHMMModel.java:373-388) */
double oracle_match_to_match_prob(int ins_qual, int del_qual) {
    oracle_init();
    int min_q = ins_qual <= del_qual ? ins_qual : del_qual;
    int max_q = ins_qual <= del_qual ? del_qual : ins_qual;
    if (min_q < 0) return NAN; /* IllegalArgumentException in Java */
    if (MAX_QUAL < max_q)
        return 1.0 - pow(10.0, oracle_approx_log10_sum_log10(-0.1 * (double)min_q, -0.1 * (double)max_q));
    return g_mm_prob[((max_q * (max_q + 1)) >> 1) + min_q];
}

/* PairHMMModel.matchToMatchProbLog10(int, int) (PairHMMModel.java:404-418) */
double oracle_match_to_match_prob_log10(int ins_qual, int del_qual) {
    oracle_init();
    int min_q = ins_qual <= del_qual ? ins_qual : del_qual;
    int max_q = ins_qual <= del_qual ? del_qual : ins_qual;
    if (MAX_QUAL < max_q) {
        const double p = pow(10.0, oracle_approx_log10_sum_log10(-0.1 * (double)min_q, -0.1 * (double)max_q));
        return log1p(-(p < 1.0 ? p : 1.0)) * g_inv_ln10;
    }
    return g_mm_log10[((max_q * (max_q + 1)) >> 1) + min_q];
}

/* PairHMMModel.qualToTransProbs(double[], byte, byte, byte) (PairHMMModel.java:107-117).
 * dest order: MM, GM(indelToMatch), MI, II, MD, DD (PairHMMModel.java:26-51).
 * Byte arguments are signed in Java: values >= 128 are negative and rejected. */
int oracle_qual_to_trans_probs(int8_t ins, int8_t del, int8_t gcp, double *dest) {
    oracle_init();
    if (ins < 0 || del < 0 || gcp < 0) return ORACLE_BAD_ARG;
    dest[0] = oracle_match_to_match_prob(ins & 0xff, del & 0xff);
    dest[2] = g_qual_err[ins & 0xff];
    dest[4] = g_qual_err[del & 0xff];
    dest[1] = 1.0 - g_qual_err[gcp & 0xff];
    dest[3] = dest[5] = g_qual_err[gcp & 0xff];
    return ORACLE_OK;
}

/* PairHMM.findFirstPositionWhereHaplotypesDiffer (PairHMM.java:341-352) */
int oracle_first_position_haplotypes_differ(const uint8_t *h1, int n1, const uint8_t *h2, int n2) {
    if (h1 == NULL || n1 == 0 || h2 == NULL || n2 == 0) return ORACLE_BAD_ARG;
    for (int i = 0; i < n1 && i < n2; i++)
        if (h1[i] != h2[i]) return i;
    return n1 < n2 ? n1 : n2;
}

static int validate_pair(const uint8_t *hap, int H, const uint8_t *read, const uint8_t *quals,
                         const uint8_t *ins, const uint8_t *del, const uint8_t *gcp, int R) {
    if (hap == NULL || H <= 0 || R < 0) return ORACLE_BAD_ARG;
    if (R > 0 && (read == NULL || quals == NULL || ins == NULL || del == NULL || gcp == NULL))
        return ORACLE_BAD_ARG;
    for (int i = 0; i < R; i++) {
        if ((int8_t)ins[i] < 0 || (int8_t)del[i] < 0 || (int8_t)gcp[i] < 0) return ORACLE_BAD_ARG;
        if (quals[i] > MAX_QUAL) return ORACLE_BAD_ARG;
    }
    return ORACLE_OK;
}

/*
 * LoglessPairHMM.subComputeReadLikelihoodGivenHaplotypeLog10 with hapStartIndex = 0 and
 * recacheReadValues = true, the only combination production uses (PairHMM.java:228, 297).
 * Loop structure LoglessPairHMM.java:47-57; priors :84-92; final sum :62-67.
 * Two rolling rows replace the N^2 matrices: each cell reads only row i-1 and cell (i, j-1), so
 * the arithmetic (operands and order) is unchanged.
 */
int oracle_logless(const uint8_t *hap, int H, const uint8_t *read, const uint8_t *quals,
                   const uint8_t *ins, const uint8_t *del, const uint8_t *gcp, int R,
                   int tristate, double *out) {
    oracle_init();
    int st = validate_pair(hap, H, read, quals, ins, del, gcp, R);
    if (st != ORACLE_OK) return st;
    const int W = H + 1;
    double *buf = (double *)malloc(sizeof(double) * (size_t)W * 6);
    if (!buf) return ORACLE_OOM;
    double *Mp = buf, *Ip = buf + W, *Dp = buf + 2 * W, *Mc = buf + 3 * W, *Ic = buf + 4 * W, *Dc = buf + 5 * W;
    const double initial_value = g_initial_condition / (double)H; /* LoglessPairHMM.java:31 */
    for (int j = 0; j < W; j++) {
        Mp[j] = 0.0;
        Ip[j] = 0.0;
        Dp[j] = initial_value; /* :33-35 */
    }
    const double tri = tristate ? 3.0 : 1.0; /* :12, :90 */
    for (int i = 1; i <= R; i++) {
        double t[6];
        oracle_qual_to_trans_probs((int8_t)ins[i - 1], (int8_t)del[i - 1], (int8_t)gcp[i - 1], t);
        const uint8_t x = read[i - 1];
        const uint8_t q = quals[i - 1];
        const double p_match = 1.0 - g_qual_err[q];
        const double p_mis = g_qual_err[q] / tri;
        Mc[0] = 0.0;
        Ic[0] = 0.0;
        Dc[0] = 0.0;
        for (int j = 1; j < W; j++) {
            const uint8_t y = hap[j - 1];
            const double prior = (x == y || x == (uint8_t)'N' || y == (uint8_t)'N') ? p_match : p_mis;
            Mc[j] = prior * (Mp[j - 1] * t[0] + Ip[j - 1] * t[1] + Dp[j - 1] * t[1]);
            Ic[j] = Mp[j] * t[2] + Ip[j] * t[3];
            Dc[j] = Mc[j - 1] * t[4] + Dc[j - 1] * t[5];
        }
        double *s;
        s = Mp; Mp = Mc; Mc = s;
        s = Ip; Ip = Ic; Ic = s;
        s = Dp; Dp = Dc; Dc = s;
    }
    double final_sum = 0.0;
    for (int j = 1; j < W; j++) final_sum += Mp[j] + Ip[j];
    *out = log10(final_sum) - g_initial_condition_log10;
    free(buf);
    return ORACLE_OK;
}

/* MathUtils.maxElementIndex(double[], start, end) (MathUtils.java:786-797) */
static int max_element_index(const double *a, int n) {
    int m = 0;
    for (int i = 1; i < n; i++)
        if (a[i] > a[m]) m = i;
    return m;
}

/* MathUtils.log10SumLog10(double[], 0, n) (MathUtils.java:622-650) */
static double exact_log10_sum_log10(const double *v, int n) {
    const int mi = max_element_index(v, n);
    const double mx = v[mi];
    if (mx == -INFINITY) return mx;
    double sum = 1.0;
    for (int i = 0; i < n; i++) {
        if (i == mi || v[i] == -INFINITY) continue;
        sum += pow(10.0, v[i] - mx);
    }
    return mx + (sum != 1.0 ? log10(sum) : 0.0);
}

/* MathUtils.approximateLog10SumLog10(double[], endIndex) (MathUtils.java:444-461) */
static double approx_log10_sum_log10_array(const double *v, int n) {
    const int mi = max_element_index(v, n);
    double s = v[mi];
    for (int i = 0; i < n; i++) {
        if (i == mi || v[i] == -INFINITY) continue;
        const double diff = s - v[i];
        s += diff < JACOBIAN_MAX_TOLERANCE ? jacobian_get(diff) : 0.0;
    }
    return s;
}

/* Log10PairHMM: EXACT (exact != 0) or ORIGINAL (Log10PairHMM.java:58-183). */
int oracle_log10(const uint8_t *hap, int H, const uint8_t *read, const uint8_t *quals,
                 const uint8_t *ins, const uint8_t *del, const uint8_t *gcp, int R,
                 int tristate, int exact, double *out) {
    oracle_init();
    int st = validate_pair(hap, H, read, quals, ins, del, gcp, R);
    if (st != ORACLE_OK) return st;
    const int W = H + 1;
    double *buf = (double *)malloc(sizeof(double) * (size_t)W * 6);
    if (!buf) return ORACLE_OOM;
    double *Mp = buf, *Ip = buf + W, *Dp = buf + 2 * W, *Mc = buf + 3 * W, *Ic = buf + 4 * W, *Dc = buf + 5 * W;
    const double initial_value = log10(1.0 / (double)H); /* :92 */
    for (int j = 0; j < W; j++) {
        Mp[j] = -INFINITY; /* :47-51 */
        Ip[j] = -INFINITY;
        Dp[j] = initial_value;
    }
    double (*lsum)(const double *, int) = exact ? exact_log10_sum_log10 : approx_log10_sum_log10_array;
    for (int i = 1; i <= R; i++) {
        const int qi = ins[i - 1] & 0xff, qd = del[i - 1] & 0xff, qc = gcp[i - 1] & 0xff;
        /* PairHMMModel.qualToTransProbsLog10 (PairHMMModel.java:214-225) */
        const double mm = oracle_match_to_match_prob_log10(qi, qd);
        const double mi = (double)qi * -0.1, md = (double)qd * -0.1;
        const double gm = g_qual_prob_log10[qc];
        const double ii = (double)qc * -0.1, dd = ii;
        const uint8_t x = read[i - 1];
        const uint8_t q = quals[i - 1];
        /* Log10PairHMM.initializeLog10Priors (:127-128) */
        const double p_match = g_qual_prob_log10[q];
        const double p_mis = (double)q * -0.1 - (tristate ? g_log10_3 : 0.0);
        Mc[0] = -INFINITY;
        Ic[0] = -INFINITY;
        Dc[0] = -INFINITY;
        for (int j = 1; j < W; j++) {
            const uint8_t y = hap[j - 1];
            const double prior = (x == y || x == (uint8_t)'N' || y == (uint8_t)'N') ? p_match : p_mis;
            double v3[3] = {Mp[j - 1] + mm, Ip[j - 1] + gm, Dp[j - 1] + gm}; /* updateCell :177-182 */
            Mc[j] = prior + lsum(v3, 3);
            double v2[2] = {Mp[j] + mi, Ip[j] + ii};
            Ic[j] = lsum(v2, 2);
            double w2[2] = {Mc[j - 1] + md, Dc[j - 1] + dd};
            Dc[j] = lsum(w2, 2);
        }
        double *s;
        s = Mp; Mp = Mc; Mc = s;
        s = Ip; Ip = Ic; Ic = s;
        s = Dp; Dp = Dc; Dc = s;
    }
    /* finalLog10LikelihoodCalculation (:98-105) */
    double v2[2] = {Mp[1], Ip[1]};
    double acc = lsum(v2, 2);
    for (int j = 2; j < W; j++) {
        double v3[3] = {acc, Mp[j], Ip[j]};
        acc = lsum(v3, 3);
    }
    *out = acc;
    free(buf);
    return ORACLE_OK;
}

/*
 * Batched LoglessPairHMM over packed arrays (the layout of gatk_amd's batch API): reads and
 * haplotypes are concatenated byte planes with per-item offsets and lengths; pair p scores
 * read pair_read[p] against haplotype pair_hap[p] and writes out[p].  OpenMP over pairs, like the
 * native OMP implementation's parallel loop over pairs.  Returns the first error status or OK.
 */
int oracle_logless_batch(int64_t n_pairs, const int32_t *pair_read, const int32_t *pair_hap,
                         const uint8_t *read_bases, const uint8_t *read_quals, const uint8_t *read_ins,
                         const uint8_t *read_del, const uint8_t *read_gcp, const int64_t *read_off,
                         const int32_t *read_len, const uint8_t *hap_bases, const int64_t *hap_off,
                         const int32_t *hap_len, int tristate, int n_threads, double *out) {
    oracle_init();
    int status = ORACLE_OK;
#pragma omp parallel for schedule(dynamic, 16) num_threads(n_threads > 0 ? n_threads : 1)
    for (int64_t p = 0; p < n_pairs; p++) {
        const int r = pair_read[p], h = pair_hap[p];
        const int64_t ro = read_off[r];
        int st = oracle_logless(hap_bases + hap_off[h], hap_len[h], read_bases + ro, read_quals + ro,
                                read_ins + ro, read_del + ro, read_gcp + ro, read_len[r], tristate, &out[p]);
        if (st != ORACLE_OK) {
#pragma omp atomic write
            status = st;
        }
    }
    return status;
}

/* ------------------------------------------------------------------------------------------ */
/* Partially determined PairHMM (LoglessPDPairHMM).                                             */
/* ------------------------------------------------------------------------------------------ */

/* PartiallyDeterminedHaplotype bit layout (PartiallyDeterminedHaplotype.java:59-66) */
#define PD_SNP 1
#define PD_DEL_START 2
#define PD_DEL_END 4
#define PD_A 8
#define PD_C 16
#define PD_G 32
#define PD_T 64

enum { PD_NORMAL = 0, PD_INSIDE_DEL = 1, PD_AFTER_DEL = 2 }; /* LoglessPDPairHMM.java:22-31 */

/* LoglessPDPairHMM.isBasePDMatching (:175-194); -1 where the reference throws. */
static int pd_base_matching(uint8_t x, uint8_t pd) {
    if (!(pd & PD_SNP)) return 0;
    switch (x) {
    case 'A': case 'a': return (pd & PD_A) != 0;
    case 'C': case 'c': return (pd & PD_C) != 0;
    case 'T': case 't': return (pd & PD_T) != 0;
    case 'G': case 'g': return (pd & PD_G) != 0;
    default: return -1;
    }
}

static inline double jmax(double a, double b) { return a >= b ? a : b; } /* Math.max on values >= 0 */

/*
 * LoglessPDPairHMM.subComputeReadLikelihoodGivenHaplotypeLog10 (LoglessPDPairHMM.java:34-150) with
 * hapStartIndex = 0 (PDPairHMM.java:281-298 keeps it there).  The state machine's currentState is
 * declared before the row loop (:63), so the state after a row's last column carries into the next
 * row's first column.  Branch matrices (:15-17) are zero in row 0 and column 0, as the reference's
 * never-written cells are.  Priors :158-171 with isBasePDMatching; two rolling rows per matrix.
 */
int oracle_pd_logless(const uint8_t *hap, const uint8_t *hap_pd, int H, const uint8_t *read, const uint8_t *quals,
                      const uint8_t *ins, const uint8_t *del, const uint8_t *gcp, int R, int tristate, double *out) {
    oracle_init();
    int st = validate_pair(hap, H, read, quals, ins, del, gcp, R);
    if (st != ORACLE_OK) return st;
    if (hap_pd == NULL) return ORACLE_BAD_ARG;
    const int W = H + 1;
    double *buf = (double *)malloc(sizeof(double) * (size_t)W * 12);
    if (!buf) return ORACLE_OOM;
    double *Mp = buf, *Ip = buf + W, *Dp = buf + 2 * W, *BMp = buf + 3 * W, *BIp = buf + 4 * W, *BDp = buf + 5 * W;
    double *Mc = buf + 6 * W, *Ic = buf + 7 * W, *Dc = buf + 8 * W, *BMc = buf + 9 * W, *BIc = buf + 10 * W,
           *BDc = buf + 11 * W;
    const double initial_value = g_initial_condition / (double)H; /* :47-51 */
    for (int j = 0; j < W; j++) {
        Mp[j] = Ip[j] = BMp[j] = BIp[j] = BDp[j] = 0.0;
        Dp[j] = initial_value;
    }
    const double tri = tristate ? 3.0 : 1.0;
    int state = PD_NORMAL;
    for (int i = 1; i <= R; i++) {
        double t[6];
        oracle_qual_to_trans_probs((int8_t)ins[i - 1], (int8_t)del[i - 1], (int8_t)gcp[i - 1], t);
        const uint8_t x = read[i - 1];
        const uint8_t q = quals[i - 1];
        const double p_match = 1.0 - g_qual_err[q];
        const double p_mis = g_qual_err[q] / tri;
        Mc[0] = Ic[0] = Dc[0] = BMc[0] = BIc[0] = BDc[0] = 0.0;
        for (int j = 1; j < W; j++) {
            const uint8_t y = hap[j - 1], pd = hap_pd[j - 1];
            int pdm = 0;
            if (!(x == y || x == (uint8_t)'N' || y == (uint8_t)'N')) {
                pdm = pd_base_matching(x, pd);
                if (pdm < 0) {
                    free(buf);
                    return ORACLE_BAD_ARG;
                }
            }
            const double prior = (x == y || x == (uint8_t)'N' || y == (uint8_t)'N' || pdm) ? p_match : p_mis;
            switch (state) {
            case PD_NORMAL:
            case PD_INSIDE_DEL:
                if (state == PD_NORMAL) {
                    BMc[j] = Mc[j - 1];
                    BDc[j] = Dc[j - 1];
                    BIc[j] = Ic[j - 1];
                } else {
                    BMc[j] = BMc[j - 1];
                    BDc[j] = BDc[j - 1];
                    BIc[j] = BIc[j - 1];
                }
                Mc[j] = prior * (Mp[j - 1] * t[0] + Ip[j - 1] * t[1] + Dp[j - 1] * t[1]);
                Dc[j] = Mc[j - 1] * t[4] + Dc[j - 1] * t[5];
                break;
            default: /* PD_AFTER_DEL */
                BMc[j] = jmax(BMc[j - 1], Mc[j - 1]);
                BDc[j] = jmax(BDc[j - 1], Dc[j - 1]);
                BIc[j] = jmax(BIc[j - 1], Ic[j - 1]);
                Mc[j] = prior * (jmax(BMp[j - 1], Mp[j - 1]) * t[0] + jmax(BIp[j - 1], Ip[j - 1]) * t[1] +
                                 jmax(BDp[j - 1], Dp[j - 1]) * t[1]);
                Dc[j] = jmax(BMc[j - 1], Mc[j - 1]) * t[4] + jmax(BDc[j - 1], Dc[j - 1]) * t[5];
                break;
            }
            if (pd & PD_DEL_END)
                Ic[j] = jmax(BMp[j], Mp[j]) * t[2] + jmax(BIp[j], Ip[j]) * t[3];
            else
                Ic[j] = Mp[j] * t[2] + Ip[j] * t[3];
            if (state == PD_AFTER_DEL) state = PD_NORMAL;
            if (pd & PD_DEL_START) state = PD_INSIDE_DEL;
            if (pd & PD_DEL_END) state = PD_AFTER_DEL;
        }
        double *s;
        s = Mp; Mp = Mc; Mc = s;
        s = Ip; Ip = Ic; Ic = s;
        s = Dp; Dp = Dc; Dc = s;
        s = BMp; BMp = BMc; BMc = s;
        s = BIp; BIp = BIc; BIc = s;
        s = BDp; BDp = BDc; BDc = s;
    }
    double final_sum = 0.0;
    for (int j = 1; j < W; j++) final_sum += Mp[j] + Ip[j]; /* :132-139 */
    *out = log10(final_sum) - g_initial_condition_log10;
    free(buf);
    return ORACLE_OK;
}

/* Every read against every haplotype, read-major out[r * n_haps + h] (the computeLikelihoods layout
 * of VectorLoglessPairPDHMM.java:108-115); hap_pd is a plane beside hap_bases. */
int oracle_pd_logless_batch(int32_t n_reads, int32_t n_haps, const uint8_t *read_bases, const uint8_t *read_quals,
                            const uint8_t *read_ins, const uint8_t *read_del, const uint8_t *read_gcp,
                            const int64_t *read_off, const int32_t *read_len, const uint8_t *hap_bases,
                            const uint8_t *hap_pd, const int64_t *hap_off, const int32_t *hap_len, int tristate,
                            int n_threads, double *out) {
    oracle_init();
    int status = ORACLE_OK;
    const int64_t n_pairs = (int64_t)n_reads * n_haps;
#pragma omp parallel for schedule(dynamic, 4) num_threads(n_threads > 0 ? n_threads : 1)
    for (int64_t p = 0; p < n_pairs; p++) {
        const int r = (int)(p / n_haps), h = (int)(p % n_haps);
        const int64_t ro = read_off[r];
        int st = oracle_pd_logless(hap_bases + hap_off[h], hap_pd + hap_off[h], hap_len[h], read_bases + ro,
                                   read_quals + ro, read_ins + ro, read_del + ro, read_gcp + ro, read_len[r],
                                   tristate, &out[p]);
        if (st != ORACLE_OK) {
#pragma omp atomic write
            status = st;
        }
    }
    return status;
}

/* ------------------------------------------------------------------------------------------ */
/* Continuous-parameter model and its FP64 adjoint (no counterpart in the reference).           */
/*                                                                                              */
/* Per read position r (0-based) the parameters are real Phred values: base quality qb[r],      */
/* insertion GOP qi[r], deletion GOP qd[r], gap continuation qc[r].  With e(q) = 10^(-q/10):    */
/*   MM = 1 - e(qi) - e(qd)  (the closed form PairHMMModelUnitTest.java:237-286 checks the       */
/*                             integer table against), MI = e(qi), MD = e(qd),                  */
/*   GM = 1 - e(qc), II = DD = e(qc), prior = 1 - e(qb) on a match, e(qb)/3 on a mismatch.      */
/* The recurrence is LoglessPairHMM's (LoglessPairHMM.java:47-57) with the same 2^1020 scale.   */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    double mm, gm, mi, ii, md, dd, pm, px, eb;
} row_params_t;

static void float_row_params(double qb, double qi, double qd, double qc, int tristate, row_params_t *p) {
    const double eb = pow(10.0, -qb / 10.0), ei = pow(10.0, -qi / 10.0), ed = pow(10.0, -qd / 10.0),
                 ec = pow(10.0, -qc / 10.0);
    p->mm = 1.0 - ei - ed;
    p->mi = ei;
    p->md = ed;
    p->gm = 1.0 - ec;
    p->ii = ec;
    p->dd = ec;
    p->eb = eb;
    p->pm = 1.0 - eb;
    p->px = eb / (tristate ? 3.0 : 1.0);
}

static int is_match(uint8_t x, uint8_t y) { return x == y || x == (uint8_t)'N' || y == (uint8_t)'N'; }

/* Forward likelihood of the continuous-parameter model; log10 L. */
int oracle_logless_float(const uint8_t *hap, int H, const uint8_t *read, int R, const double *qb,
                         const double *qi, const double *qd, const double *qc, int tristate, double *out) {
    oracle_init();
    if (hap == NULL || H <= 0 || R < 0) return ORACLE_BAD_ARG;
    const int W = H + 1;
    double *buf = (double *)malloc(sizeof(double) * (size_t)W * 6);
    if (!buf) return ORACLE_OOM;
    double *Mp = buf, *Ip = buf + W, *Dp = buf + 2 * W, *Mc = buf + 3 * W, *Ic = buf + 4 * W, *Dc = buf + 5 * W;
    for (int j = 0; j < W; j++) {
        Mp[j] = 0.0;
        Ip[j] = 0.0;
        Dp[j] = g_initial_condition / (double)H;
    }
    for (int i = 1; i <= R; i++) {
        row_params_t t;
        float_row_params(qb[i - 1], qi[i - 1], qd[i - 1], qc[i - 1], tristate, &t);
        Mc[0] = Ic[0] = Dc[0] = 0.0;
        for (int j = 1; j < W; j++) {
            const double prior = is_match(read[i - 1], hap[j - 1]) ? t.pm : t.px;
            Mc[j] = prior * (Mp[j - 1] * t.mm + (Ip[j - 1] + Dp[j - 1]) * t.gm);
            Ic[j] = Mp[j] * t.mi + Ip[j] * t.ii;
            Dc[j] = Mc[j - 1] * t.md + Dc[j - 1] * t.dd;
        }
        double *s;
        s = Mp; Mp = Mc; Mc = s;
        s = Ip; Ip = Ic; Ic = s;
        s = Dp; Dp = Dc; Dc = s;
    }
    double final_sum = 0.0;
    for (int j = 1; j < W; j++) final_sum += Mp[j] + Ip[j];
    *out = log10(final_sum) - g_initial_condition_log10;
    free(buf);
    return ORACLE_OK;
}

/*
 * FP64 adjoint of oracle_logless_float.  Outputs, per read position r (arrays of length R):
 *   g_qi, g_qd, g_qc : d log10 L / d(Phred) of insertion GOP, deletion GOP, GCP
 *   g_eb             : d log10 L / d(base error probability e(qb))
 *   g_qb             : d log10 L / d(base quality Phred)
 * Full (R+1) x (H+1) matrices of alpha (forward) and beta (adjoint) values.
 *   beta_M(r,c) = p(r+1,c+1) MM_{r+1} beta_M(r+1,c+1) + MI_{r+1} beta_I(r+1,c) + MD_r beta_D(r,c+1)
 *   beta_I(r,c) = p(r+1,c+1) GM_{r+1} beta_M(r+1,c+1) + II_{r+1} beta_I(r+1,c)
 *   beta_D(r,c) = p(r+1,c+1) GM_{r+1} beta_M(r+1,c+1) + DD_r beta_D(r,c+1)
 * with beta_M = beta_I = 1, beta_D = 0 seeded on the last row (L = sum_c M(R,c) + I(R,c)).
 */
int oracle_logless_float_grad(const uint8_t *hap, int H, const uint8_t *read, int R, const double *qb,
                              const double *qi, const double *qd, const double *qc, int tristate,
                              double *log10_l, double *g_qi, double *g_qd, double *g_qc, double *g_eb,
                              double *g_qb) {
    oracle_init();
    if (hap == NULL || H <= 0 || R <= 0) return ORACLE_BAD_ARG;
    const int W = H + 1;
    const size_t n = (size_t)(R + 1) * (size_t)W;
    double *A = (double *)calloc(6 * n, sizeof(double));
    row_params_t *P = (row_params_t *)calloc((size_t)R + 2, sizeof(row_params_t));
    if (!A || !P) {
        free(A);
        free(P);
        return ORACLE_OOM;
    }
    double *M = A, *I = A + n, *D = A + 2 * n, *bM = A + 3 * n, *bI = A + 4 * n, *bD = A + 5 * n;
#define IX(i, j) ((size_t)(i) * (size_t)W + (size_t)(j))
    for (int j = 0; j < W; j++) D[IX(0, j)] = g_initial_condition / (double)H;
    for (int i = 1; i <= R; i++) float_row_params(qb[i - 1], qi[i - 1], qd[i - 1], qc[i - 1], tristate, &P[i]);
    for (int i = 1; i <= R; i++) {
        const row_params_t t = P[i];
        for (int j = 1; j < W; j++) {
            const double prior = is_match(read[i - 1], hap[j - 1]) ? t.pm : t.px;
            M[IX(i, j)] = prior * (M[IX(i - 1, j - 1)] * t.mm + (I[IX(i - 1, j - 1)] + D[IX(i - 1, j - 1)]) * t.gm);
            I[IX(i, j)] = M[IX(i - 1, j)] * t.mi + I[IX(i - 1, j)] * t.ii;
            D[IX(i, j)] = M[IX(i, j - 1)] * t.md + D[IX(i, j - 1)] * t.dd;
        }
    }
    double L = 0.0;
    for (int j = 1; j < W; j++) L += M[IX(R, j)] + I[IX(R, j)];
    *log10_l = log10(L) - g_initial_condition_log10;
    for (int i = R; i >= 1; i--) {
        const row_params_t t = P[i];
        const row_params_t *nx = (i < R) ? &P[i + 1] : NULL;
        for (int j = H; j >= 1; j--) {
            double gm_ = (i == R) ? 1.0 : 0.0, gi_ = (i == R) ? 1.0 : 0.0, gd_ = 0.0;
            if (nx) {
                const double pnext = (j < H) ? (is_match(read[i], hap[j]) ? nx->pm : nx->px) : 0.0;
                const double bm_next = (j < H) ? bM[IX(i + 1, j + 1)] : 0.0;
                gm_ += pnext * nx->mm * bm_next + nx->mi * bI[IX(i + 1, j)];
                gi_ += pnext * nx->gm * bm_next + nx->ii * bI[IX(i + 1, j)];
                gd_ += pnext * nx->gm * bm_next;
            }
            if (j < H) {
                gm_ += t.md * bD[IX(i, j + 1)];
                gd_ += t.dd * bD[IX(i, j + 1)];
            }
            bM[IX(i, j)] = gm_;
            bI[IX(i, j)] = gi_;
            bD[IX(i, j)] = gd_;
        }
    }
    const double scale = 1.0 / (L * log(10.0));
    const double dq = -log(10.0) / 10.0; /* d e(q) / dq = dq * e(q) */
    for (int i = 1; i <= R; i++) {
        const row_params_t t = P[i];
        double s_mm = 0, s_gm = 0, s_mi = 0, s_ii = 0, s_md = 0, s_dd = 0, s_eb = 0;
        for (int j = 1; j < W; j++) {
            const int m = is_match(read[i - 1], hap[j - 1]);
            const double prior = m ? t.pm : t.px;
            const double pre = M[IX(i - 1, j - 1)] * t.mm + (I[IX(i - 1, j - 1)] + D[IX(i - 1, j - 1)]) * t.gm;
            const double b = bM[IX(i, j)];
            s_mm += b * prior * M[IX(i - 1, j - 1)];
            s_gm += b * prior * (I[IX(i - 1, j - 1)] + D[IX(i - 1, j - 1)]);
            s_mi += bI[IX(i, j)] * M[IX(i - 1, j)];
            s_ii += bI[IX(i, j)] * I[IX(i - 1, j)];
            s_md += bD[IX(i, j)] * M[IX(i, j - 1)];
            s_dd += bD[IX(i, j)] * D[IX(i, j - 1)];
            s_eb += b * pre * (m ? -1.0 : 1.0 / (tristate ? 3.0 : 1.0));
        }
        const double d_ei = -s_mm + s_mi, d_ed = -s_mm + s_md, d_ec = -s_gm + s_ii + s_dd;
        g_qi[i - 1] = d_ei * scale * dq * t.mi;
        g_qd[i - 1] = d_ed * scale * dq * t.md;
        g_qc[i - 1] = d_ec * scale * dq * t.ii;
        g_eb[i - 1] = s_eb * scale;
        g_qb[i - 1] = s_eb * scale * dq * t.eb;
    }
#undef IX
    free(A);
    free(P);
    return ORACLE_OK;
}
