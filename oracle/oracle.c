/*
 * oracle.c -- CPU restatement of the reference's subsequence-search path.
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker, never the product.
 *
 * Build: gcc -O2 -ffp-contract=off -mfma -fopenmp (oracle/Makefile).  All
 * fused multiply-adds below are explicit; nothing else may be contracted.
 */
#include "oracle.h"

#include <omp.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_INF ((float)INFINITY)

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline float fmin2(float a, float b) { return a < b ? a : b; }

/* Local cost of one cell, lp(euclidean, d) in bounds.cuh:150. */
static inline float lp(float d, int l2) { return l2 ? d * d : fabsf(d); }
/* lp(euclidean, d) + acc as nvcc contracts it (one rounding for L2). */
static inline float lpa(float d, float acc, int l2) {
    return l2 ? fmaf(d, d, acc) : fabsf(d) + acc;
}

/* ------------------------------------------------------------------------ */
/* stats.cuh:8-28 znormalize: mean, then sqrt(E[x^2] - mean^2), both
 * accumulated in value_t in index order; then (x - avg) / std in place. */
void or_znormalize_f32(float* x, int32_t n) {
    float avg = 0.0f, sd = 0.0f;
    for (int32_t i = 0; i < n; ++i) avg += x[i];
    avg /= (float)n;
    for (int32_t i = 0; i < n; ++i) sd += x[i] * x[i];
    sd = sqrtf(sd / (float)n - avg * avg);
    for (int32_t i = 0; i < n; ++i) {
        x[i] -= avg;
        x[i] /= sd;
    }
}

void or_znormalize_f64(double* x, int32_t n) {
    double avg = 0.0, sd = 0.0;
    for (int32_t i = 0; i < n; ++i) avg += x[i];
    avg /= (double)n;
    for (int32_t i = 0; i < n; ++i) sd += x[i] * x[i];
    sd = sqrt(sd / (double)n - avg * avg);
    for (int32_t i = 0; i < n; ++i) {
        x[i] -= avg;
        x[i] /= sd;
    }
}

/* ------------------------------------------------------------------------ */
/* bounds.cuh:13-64: Lemire's streaming max/min with two index deques.  The
 * deques only ever grow at the back, so a flat array with head/tail indices
 * holds them.  w >= m is clamped to m - 1 (the envelope is then the global
 * max/min; the reference indexes out of bounds there). */
void or_envelope(const float* q, int32_t w, float* lo, float* up, int32_t m) {
    if (m <= 0) return;
    if (w > m - 1) w = m - 1;
    if (w < 0) w = 0;
    int* ub = (int*)malloc(sizeof(int) * (size_t)(m + 1));
    int* lb = (int*)malloc(sizeof(int) * (size_t)(m + 1));
    int uh = 0, ut = 0, lh = 0, lt = 0;
    ub[ut++] = 0;
    lb[lt++] = 0;
    for (int i = 1; i < m; ++i) {
        if (i > w) {
            up[i - w - 1] = q[ub[uh]];
            lo[i - w - 1] = q[lb[lh]];
        }
        if (q[i] > q[i - 1]) {
            ut--;
            while (ut > uh && q[i] > q[ub[ut - 1]]) ut--;
        } else {
            lt--;
            while (lt > lh && q[i] < q[lb[lt - 1]]) lt--;
        }
        ub[ut++] = i;
        lb[lt++] = i;
        if (i == 2 * w + 1 + ub[uh])
            uh++;
        else if (i == 2 * w + 1 + lb[lh])
            lh++;
    }
    for (int i = m; i < m + w + 1; ++i) {
        up[i - w - 1] = q[ub[uh]];
        lo[i - w - 1] = q[lb[lh]];
        if (i - ub[uh] >= 2 * w + 1) uh++;
        if (i - lb[lh] >= 2 * w + 1) lh++;
    }
    free(ub);
    free(lb);
}

/* ------------------------------------------------------------------------ */
/* stats.cuh:85-113 for one epoch: FP64 inclusive prefix with P[0] = 0
 * (cub_util.cuh:29-54), FP64 window difference / M stored as FP32
 * (stats.cuh:30-38), FP32 finalize (stats.cuh:40-52).
 * The reference scans in place: stats.cuh:100-101 call
 * prefix_sum(Aprefix, Aprefix, N), whose first step writes 0 into Output[0]
 * (cub_util.cuh:35-36), the same element as Input[0]; the scan that follows
 * (:43-44) then reads that 0 instead of the epoch's first sample.  So
 * P[k + 1] sums x[1..k], and only the epoch's first window, (P[M] - P[0]) / M,
 * is affected: it leaves out x[0] (and x[0]^2 in the squares' prefix). */
static void refcompat_epoch(const float* x, float* avg, float* sd, int32_t m,
                            int64_t len) {
    double* p1 = (double*)malloc(sizeof(double) * (size_t)(len + 1));
    double* p2 = (double*)malloc(sizeof(double) * (size_t)(len + 1));
    p1[0] = 0.0;
    p2[0] = 0.0;
    for (int64_t k = 0; k < len; ++k) {
        double v = k == 0 ? 0.0 : (double)x[k];
        p1[k + 1] = p1[k] + v;
        p2[k + 1] = p2[k] + v * v;
    }
    for (int64_t i = 0; i + m <= len; ++i) {
        float X = (float)((p1[i + m] - p1[i]) / (double)m);
        float X2 = (float)((p2[i + m] - p2[i]) / (double)m);
        float v = fmaf(-X, X, X2);
        avg[i] = X;
        sd[i] = v > 0.0f ? sqrtf(v) : 1.0f;
    }
    free(p1);
    free(p2);
}

/* stats.cuh:115-142 avg_std: epochs of `epoch` samples advancing by
 * epoch - M + 1; a final chunk of at most M + 1 samples is merged into the
 * previous epoch.  (With a single short chunk the reference reads an empty
 * vector; here that chunk is simply processed.) */
void or_stats_refcompat(const float* s, int64_t n, int32_t m, int64_t epoch,
                        float* avg, float* sd) {
    if (epoch <= 0) epoch = 10000000;
    int64_t width = epoch - m + 1;
    if (width < 1) width = 1;
    int64_t cap = n / width + 2;
    int64_t* lowers = (int64_t*)malloc(sizeof(int64_t) * (size_t)cap);
    int64_t* uppers = (int64_t*)malloc(sizeof(int64_t) * (size_t)cap);
    int64_t cnt = 0;
    for (int64_t lower = 0; lower < n; lower += width) {
        int64_t upper = lower + epoch < n ? lower + epoch : n;
        if (upper - lower > m + 1 || cnt == 0) {
            lowers[cnt] = lower;
            uppers[cnt] = upper;
            cnt++;
        } else {
            uppers[cnt - 1] = upper;
        }
    }
    for (int64_t e = 0; e < cnt; ++e)
        refcompat_epoch(s + lowers[e], avg + lowers[e], sd + lowers[e], m,
                        uppers[e] - lowers[e]);
    free(lowers);
    free(uppers);
}

/* Non-finite samples: the library's rule (not the reference's, whose running
 * prefix spreads them), stated here as the oracle's: a window holding a NaN
 * or an infinity has NaN statistics, hence NaN distances and bounds. */
void or_stats_fp64(const float* s, int64_t n, int32_t m, float* avg, float* sd,
                   double* avg64, double* sd64) {
    int64_t L = n - m + 1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < L; ++i) {
        int bad = 0;
        for (int32_t k = 0; k < m; ++k) bad |= !isfinite(s[i + k]);
        if (bad) {
            avg[i] = sd[i] = NAN;
            if (avg64) avg64[i] = NAN;
            if (sd64) sd64[i] = NAN;
            continue;
        }
        double sum = 0.0;
        for (int32_t k = 0; k < m; ++k) sum += (double)s[i + k];
        double mean = sum / (double)m;
        double ss = 0.0;
        for (int32_t k = 0; k < m; ++k) {
            double d = (double)s[i + k] - mean;
            ss += d * d;
        }
        double var = ss / (double)m;
        double sdev = var > 0.0 ? sqrt(var) : 1.0;
        avg[i] = (float)mean;
        sd[i] = (float)sdev;
        if (avg64) avg64[i] = mean;
        if (sd64) sd64[i] = sdev;
    }
}

/* ------------------------------------------------------------------------ */
/* bounds.cuh:283-407 register_lb_kim / :152-281 random_lb_kim. */
static float lb_kim_one(const float* s, float avg, float sdv, const float* zq,
                        int32_t m, int l2) {
    float q0, q1, q2, q3, q4, s0, s1, s2, s3, s4, mem;
    float p0, p1, p2, p3, p4, p5, p6, p7, p8, p9;
    q0 = zq[0]; q1 = zq[1]; q2 = zq[2]; q3 = zq[3]; q4 = zq[4];
    s0 = (s[0] - avg) / sdv;
    s1 = (s[1] - avg) / sdv;
    s2 = (s[2] - avg) / sdv;
    s3 = (s[3] - avg) / sdv;
    s4 = (s[4] - avg) / sdv;

    p0 = lp(q0 - s0, l2);
    p1 = lpa(q0 - s1, p0, l2);
    p2 = lpa(q0 - s2, p1, l2);
    p3 = lpa(q0 - s3, p2, l2);
    p4 = lpa(q0 - s4, p3, l2);
    mem = p4;

    p5 = lpa(q1 - s0, p0, l2);
    p6 = lpa(q1 - s1, fmin2(p5, fmin2(p0, p1)), l2);
    p7 = lpa(q1 - s2, fmin2(p6, fmin2(p1, p2)), l2);
    p8 = lpa(q1 - s3, fmin2(p7, fmin2(p2, p3)), l2);
    p9 = lpa(q1 - s4, fmin2(p8, fmin2(p3, p4)), l2);
    mem = fmin2(mem, p9);

    p0 = lpa(q2 - s0, p5, l2);
    p1 = lpa(q2 - s1, fmin2(p0, fmin2(p5, p6)), l2);
    p2 = lpa(q2 - s2, fmin2(p1, fmin2(p6, p7)), l2);
    p3 = lpa(q2 - s3, fmin2(p2, fmin2(p7, p8)), l2);
    p4 = lpa(q2 - s4, fmin2(p3, fmin2(p8, p9)), l2);
    mem = fmin2(mem, p4);

    p5 = lpa(q3 - s0, p0, l2);
    p6 = lpa(q3 - s1, fmin2(p5, fmin2(p0, p1)), l2);
    p7 = lpa(q3 - s2, fmin2(p6, fmin2(p1, p2)), l2);
    p8 = lpa(q3 - s3, fmin2(p7, fmin2(p2, p3)), l2);
    p9 = lpa(q3 - s4, fmin2(p8, fmin2(p3, p4)), l2);
    mem = fmin2(mem, p9);

    p0 = lpa(q4 - s0, p5, l2);
    p1 = lpa(q4 - s1, fmin2(p0, fmin2(p5, p6)), l2);
    p2 = lpa(q4 - s2, fmin2(p1, fmin2(p6, p7)), l2);
    p3 = lpa(q4 - s3, fmin2(p2, fmin2(p7, p8)), l2);
    p4 = lpa(q4 - s4, fmin2(p3, fmin2(p8, p9)), l2);
    mem = fmin2(mem, fmin2(p0, fmin2(p1, fmin2(p2, fmin2(p3, p4)))));

    q0 = zq[m - 5]; q1 = zq[m - 4]; q2 = zq[m - 3]; q3 = zq[m - 2]; q4 = zq[m - 1];
    s0 = (s[m - 5] - avg) / sdv;
    s1 = (s[m - 4] - avg) / sdv;
    s2 = (s[m - 3] - avg) / sdv;
    s3 = (s[m - 2] - avg) / sdv;
    s4 = (s[m - 1] - avg) / sdv;

    p0 = lp(q0 - s0, l2);
    p1 = lp(q0 - s1, l2);
    p2 = lp(q0 - s2, l2);
    p3 = lp(q0 - s3, l2);
    p4 = lp(q0 - s4, l2);

    p5 = lp(q1 - s0, l2);
    p6 = lpa(q1 - s1, fmin2(p5, fmin2(p0, p1)), l2);
    p7 = lpa(q1 - s2, fmin2(p6, fmin2(p1, p2)), l2);
    p8 = lpa(q1 - s3, fmin2(p7, fmin2(p2, p3)), l2);
    p9 = lpa(q1 - s4, fmin2(p8, fmin2(p3, p4)), l2);

    p0 = lp(q2 - s0, l2);
    p1 = lpa(q2 - s1, fmin2(p0, fmin2(p5, p6)), l2);
    p2 = lpa(q2 - s2, fmin2(p1, fmin2(p6, p7)), l2);
    p3 = lpa(q2 - s3, fmin2(p2, fmin2(p7, p8)), l2);
    p4 = lpa(q2 - s4, fmin2(p3, fmin2(p8, p9)), l2);

    p5 = lp(q3 - s0, l2);
    p6 = lpa(q3 - s1, fmin2(p5, fmin2(p0, p1)), l2);
    p7 = lpa(q3 - s2, fmin2(p6, fmin2(p1, p2)), l2);
    p8 = lpa(q3 - s3, fmin2(p7, fmin2(p2, p3)), l2);
    p9 = lpa(q3 - s4, fmin2(p8, fmin2(p3, p4)), l2);

    p0 = lp(q4 - s0, l2);
    p1 = lpa(q4 - s1, fmin2(p0, fmin2(p5, p6)), l2);
    p2 = lpa(q4 - s2, fmin2(p1, fmin2(p6, p7)), l2);
    p3 = lpa(q4 - s3, fmin2(p2, fmin2(p7, p8)), l2);
    p4 = lpa(q4 - s4, fmin2(p3, fmin2(p8, p9)), l2);

    return mem + p4;
}

void or_lb_kim(const float* s, const float* avg, const float* sd,
               const float* zq, int32_t m, const int32_t* idx, int64_t count,
               int32_t metric, float* out) {
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < count; ++g) {
        int64_t w = idx ? idx[g] : g;
        out[g] = lb_kim_one(s + w, avg[w], sd[w], zq, m, metric);
    }
}

/* bounds.cuh:111-144.  The reference adds u*u*(u>0) + l*l*(l<0); at most one
 * term is non-zero (lo <= up), so the sum is accumulated with one fmaf. */
void or_lb_keogh(const float* s, const float* avg, const float* sd,
                 const float* lo, const float* up, int32_t m,
                 const int32_t* idx, int64_t count, int32_t metric, float* out) {
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < count; ++g) {
        int64_t w = idx ? idx[g] : g;
        float a = avg[w], sdv = sd[w], acc = 0.0f;
        for (int32_t i = 0; i < m; ++i) {
            float v = (s[w + i] - a) / sdv;
            float lower = v - lo[i];
            float upper = v - up[i];
            if (metric) {
                if (upper > 0.0f)
                    acc = fmaf(upper, upper, acc);
                else if (lower < 0.0f)
                    acc = fmaf(lower, lower, acc);
            } else {
                if (upper > 0.0f)
                    acc = acc + upper;
                else if (lower < 0.0f)
                    acc = acc - lower;
            }
        }
        out[g] = acc;
    }
}

/* ------------------------------------------------------------------------ */
/* cdtw.cuh:23-88 sequential_cdtw: two rows of M+1 cells, row i = query
 * index, band |i - j| <= W, halo cells set to INFINITY. */
void or_cdtw(const float* zq, const float* s, const float* avg, const float* sd,
             int32_t m, const int32_t* idx, int64_t count, int32_t w,
             int32_t metric, int32_t fused, float* out) {
    const int lane = m + 1;
#pragma omp parallel
    {
        float* pen = (float*)malloc(sizeof(float) * 2 * (size_t)lane);
#pragma omp for schedule(dynamic, 64)
        for (int64_t g = 0; g < count; ++g) {
            int64_t ix = idx ? idx[g] : g;
            const float* sub = s + ix;
            float a = avg[ix], sdv = sd[ix];
            for (int k = 0; k < 2 * lane; ++k) pen[k] = ORC_INF;
            pen[0] = 0.0f;
            int src = 1, trg = 0;
            for (int i = 1; i < lane; ++i) {
                trg = src;
                src = (trg == 0);
                int lower = imax(0, i - w - 1), upper = imin(lane, i + w + 1);
                pen[trg * lane + lower] = ORC_INF;
                for (int j = lower + 1; j < upper; ++j) {
                    float value = zq[i - 1] - (sub[j - 1] - a) / sdv;
                    float mn = fmin2(pen[src * lane + j - 1],
                                     fmin2(pen[src * lane + j], pen[trg * lane + j - 1]));
                    float wt;
                    if (metric)
                        wt = fused ? fmaf(value, value, mn) : value * value + mn;
                    else
                        wt = fabsf(value) + mn;
                    pen[trg * lane + j] = wt;
                }
                if (upper < lane) pen[trg * lane + upper] = ORC_INF;
            }
            out[g] = pen[trg * lane + m];
        }
        free(pen);
    }
}

void or_cdtw_f64(const double* zq, const float* s, const double* avg,
                 const double* sd, int32_t m, const int32_t* idx, int64_t count,
                 int32_t w, int32_t metric, double* out) {
    const int lane = m + 1;
#pragma omp parallel
    {
        double* pen = (double*)malloc(sizeof(double) * 2 * (size_t)lane);
#pragma omp for schedule(dynamic, 64)
        for (int64_t g = 0; g < count; ++g) {
            int64_t ix = idx ? idx[g] : g;
            const float* sub = s + ix;
            double a = avg[ix], sdv = sd[ix];
            for (int k = 0; k < 2 * lane; ++k) pen[k] = INFINITY;
            pen[0] = 0.0;
            int src = 1, trg = 0;
            for (int i = 1; i < lane; ++i) {
                trg = src;
                src = (trg == 0);
                int lower = imax(0, i - w - 1), upper = imin(lane, i + w + 1);
                pen[trg * lane + lower] = INFINITY;
                for (int j = lower + 1; j < upper; ++j) {
                    double value = zq[i - 1] - ((double)sub[j - 1] - a) / sdv;
                    double mn = pen[src * lane + j - 1];
                    if (pen[src * lane + j] < mn) mn = pen[src * lane + j];
                    if (pen[trg * lane + j - 1] < mn) mn = pen[trg * lane + j - 1];
                    pen[trg * lane + j] = (metric ? value * value : fabs(value)) + mn;
                }
                if (upper < lane) pen[trg * lane + upper] = INFINITY;
            }
            out[g] = pen[trg * lane + m];
        }
        free(pen);
    }
}

/* ------------------------------------------------------------------------ */
/* Query gradient of banded cDTW: the recurrence of cdtw.cuh:23-88 in FP64 on
 * band storage (row i keeps columns i-w .. i+w), argmin predecessor per cell,
 * path walked back from the last cell. */
void or_cdtw_grad_f64(const double* zq, const float* s, const double* avg,
                      const double* sd, int32_t m, const int32_t* idx,
                      int64_t count, int32_t w, int32_t metric,
                      int32_t sqrt_dist, const double* g, const double* q_raw,
                      double* dist, double* grad) {
    or_cdtw_grad(zq, s, avg, sd, m, idx, count, w, metric, sqrt_dist, 0, g, q_raw, dist, grad);
}

/* The optimal path of the FP32 DP of or_cdtw (fused = 1): cells rounded as
 * fmaf(d, d, min) / fabsf(d) + min with d = q - z, z = (s - avg) / sd in FP32. */
static void path32_dp(const double* zq, const float* sub, float a, float sdv, int32_t m,
                      int32_t w, int32_t metric, int bw, float* D, uint8_t* dir, double* z) {
    for (int j = 0; j < m; ++j) z[j] = (double)((sub[j] - a) / sdv);
    for (int i = 0; i < m; ++i) {
        const float qi = (float)zq[i];
        for (int t = 0; t < bw; ++t) {
            const int j = i - w + t;
            float v = INFINITY;
            uint8_t dr = 3;
            if (j >= 0 && j < m) {
                float dg = INFINITY, up = INFINITY, lf = INFINITY;
                if (i == 0 && j == 0) {
                    dg = 0.0f;
                } else {
                    if (i > 0 && j > 0) dg = D[(size_t)(i - 1) * bw + t];
                    if (i > 0 && t + 1 < bw) up = D[(size_t)(i - 1) * bw + t + 1];
                    if (j > 0 && t > 0) lf = D[(size_t)i * bw + t - 1];
                }
                float mn;
                if (dg <= up && dg <= lf) {
                    mn = dg;
                    dr = 0;
                } else if (up <= lf) {
                    mn = up;
                    dr = 1;
                } else {
                    mn = lf;
                    dr = 2;
                }
                const float d = qi - (float)z[j];
                v = metric ? fmaf(d, d, mn) : fabsf(d) + mn;
            }
            D[(size_t)i * bw + t] = v;
            dir[(size_t)i * bw + t] = dr;
        }
    }
}

void or_cdtw_grad(const double* zq, const float* s, const double* avg,
                  const double* sd, int32_t m, const int32_t* idx,
                  int64_t count, int32_t w, int32_t metric,
                  int32_t sqrt_dist, int32_t path32, const double* g,
                  const double* q_raw, double* dist, double* grad) {
    if (w > m - 1) w = m - 1;
    const int bw = 2 * w + 1;
    double* G = (double*)calloc((size_t)m, sizeof(double));
    double* D = (double*)malloc(sizeof(double) * (size_t)m * bw);
    float* D32 = (float*)malloc(sizeof(float) * (size_t)m * bw);
    uint8_t* dir = (uint8_t*)malloc((size_t)m * bw);
    double* z = (double*)malloc(sizeof(double) * (size_t)m);
    double* row = (double*)malloc(sizeof(double) * (size_t)m);
    for (int64_t p = 0; p < count; ++p) {
        const int64_t ix = idx ? idx[p] : p;
        if (path32) {
            path32_dp(zq, s + ix, (float)avg[ix], (float)sd[ix], m, w, metric, bw, D32, dir, z);
            const float Dm32 = D32[(size_t)(m - 1) * bw + w];
            double scale = g[p];
            if (sqrt_dist) scale = Dm32 > 0.0f ? scale * 0.5 / sqrt((double)Dm32) : 0.0;
            if (dist) dist[p] = sqrt_dist ? (double)sqrtf(Dm32) : (double)Dm32;
            for (int i = 0; i < m; ++i) row[i] = 0.0;
            int i = m - 1, j = m - 1;
            for (;;) {
                const double d = (double)(float)zq[i] - z[j];
                row[i] += metric ? 2.0 * d : (d > 0 ? 1.0 : (d < 0 ? -1.0 : 0.0));
                if (i == 0 && j == 0) break;
                const uint8_t dr = dir[(size_t)i * bw + (j - i + w)];
                if (dr == 0) {
                    --i;
                    --j;
                } else if (dr == 1) {
                    --i;
                } else {
                    --j;
                }
            }
            for (int r = 0; r < m; ++r) G[r] += scale * row[r];
            continue;
        }
        for (int j = 0; j < m; ++j) z[j] = ((double)s[ix + j] - avg[ix]) / sd[ix];
        for (int i = 0; i < m; ++i) {
            for (int t = 0; t < bw; ++t) {
                const int j = i - w + t;
                double v = INFINITY;
                uint8_t dr = 3;
                if (j >= 0 && j < m) {
                    double dg = INFINITY, up = INFINITY, lf = INFINITY;
                    if (i == 0 && j == 0) {
                        dg = 0.0;
                    } else {
                        if (i > 0 && j > 0) dg = D[(size_t)(i - 1) * bw + t];      /* (i-1, j-1) */
                        if (i > 0 && t + 1 < bw) up = D[(size_t)(i - 1) * bw + t + 1]; /* (i-1, j) */
                        if (j > 0 && t > 0) lf = D[(size_t)i * bw + t - 1];        /* (i, j-1) */
                    }
                    double mn;
                    if (dg <= up && dg <= lf) {
                        mn = dg;
                        dr = 0;
                    } else if (up <= lf) {
                        mn = up;
                        dr = 1;
                    } else {
                        mn = lf;
                        dr = 2;
                    }
                    const double d = zq[i] - z[j];
                    v = (metric ? d * d : fabs(d)) + mn;
                }
                D[(size_t)i * bw + t] = v;
                dir[(size_t)i * bw + t] = dr;
            }
        }
        const double Dm = D[(size_t)(m - 1) * bw + w];
        double scale = g[p];
        if (sqrt_dist) scale = Dm > 0.0 ? scale * 0.5 / sqrt(Dm) : 0.0;
        if (dist) dist[p] = sqrt_dist ? sqrt(Dm) : Dm;
        for (int i = 0; i < m; ++i) row[i] = 0.0;
        int i = m - 1, j = m - 1;
        for (;;) {
            const double d = zq[i] - z[j];
            row[i] += metric ? 2.0 * d : (d > 0 ? 1.0 : (d < 0 ? -1.0 : 0.0));
            if (i == 0 && j == 0) break;
            const uint8_t dr = dir[(size_t)i * bw + (j - i + w)];
            if (dr == 0) {
                --i;
                --j;
            } else if (dr == 1) {
                --i;
            } else {
                --j;
            }
        }
        for (int r = 0; r < m; ++r) G[r] += scale * row[r];
    }
    if (q_raw) {
        double mu = 0.0, var = 0.0, mg = 0.0, mgz = 0.0;
        for (int i = 0; i < m; ++i) mu += q_raw[i];
        mu /= m;
        for (int i = 0; i < m; ++i) var += (q_raw[i] - mu) * (q_raw[i] - mu);
        const double sig = sqrt(var / m);
        for (int i = 0; i < m; ++i) {
            mg += G[i];
            mgz += G[i] * zq[i];
        }
        mg /= m;
        mgz /= m;
        for (int i = 0; i < m; ++i) grad[i] = (G[i] - mg - zq[i] * mgz) / sig;
    } else {
        for (int i = 0; i < m; ++i) grad[i] = G[i];
    }
    free(G);
    free(D);
    free(D32);
    free(dir);
    free(z);
    free(row);
}

/* ------------------------------------------------------------------------ */
/* ldtw.cuh:74-127 z_serial_ldtw; INFTY = 1E10 (ldtw.cuh:8). */
#define LDTW_INFTY ((float)1E10)

static float ldtw_one(const float* q, const float* sub, float a, float sdv,
                      int32_t M, int32_t W, int l2, uint8_t* moves) {
    int i = 0, j = 0, step = 0;
    /* NaN statistics (a non-finite sample): NaN, and no walk -- with NaN
     * costs every comparison fails and the reference's loop would step past
     * the window (ldtw.cuh:92-123). */
    if (!(a == a) || !(sdv == sdv)) {
        if (moves)
            for (; step < 2 * M - 2; ++step) moves[step] = 3;
        return NAN;
    }
    float d = q[i] - (sub[j] - a) / sdv;
    d = l2 ? d * d : (d < 0 ? -d : d);
    while (i + 1 != M || j + 1 != M) {
        float diag = LDTW_INFTY, rght = LDTW_INFTY, down = LDTW_INFTY;
        if (i + 1 < M && j + 1 < M) {
            diag = q[i + 1] - (sub[j + 1] - a) / sdv;
            diag = l2 ? diag * diag : (diag < 0 ? -diag : diag);
        }
        if (i + 1 < M && abs(i - j + 1) <= W) {
            down = q[i + 1] - (sub[j] - a) / sdv;
            down = l2 ? down * down : (down < 0 ? -down : down);
        }
        if (j + 1 < M && abs(j - i + 1) <= W) {
            rght = q[i] - (sub[j + 1] - a) / sdv;
            rght = l2 ? rght * rght : (rght < 0 ? -rght : rght);
        }
        const int diagSdown = diag < down;
        const int diagSrght = diag < rght;
        const int downSrght = down < rght;
        uint8_t mv;
        if (diagSdown && diagSrght) {
            d += diag;
            i++;
            j++;
            mv = 0;
        } else if (!diagSdown && downSrght) {
            d += down;
            i++;
            mv = 1;
        } else {
            d += rght;
            j++;
            mv = 2;
        }
        if (moves) moves[step] = mv;
        step++;
    }
    if (moves)
        for (; step < 2 * M - 2; ++step) moves[step] = 3;
    return d;
}

void or_ldtw(const float* zq, const float* s, const float* avg, const float* sd,
             int32_t m, const int32_t* idx, int64_t count, int32_t w,
             int32_t metric, float* out, uint8_t* moves) {
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < count; ++g) {
        int64_t ix = idx ? idx[g] : g;
        out[g] = ldtw_one(zq, s + ix, avg[ix], sd[ix], m, w, metric,
                          moves ? moves + g * (int64_t)(2 * m - 2) : NULL);
    }
}

/* ldtw.cuh:18-71 serial_ldtw: the same greedy walk on raw values (no
 * z-normalisation of the window; the query is taken as given). */
static float ldtw_raw_one(const float* q, const float* x, int32_t M, int32_t W, int l2) {
    int i = 0, j = 0;
    for (int32_t k = 0; k < M; ++k)
        if (!isfinite(x[k])) return NAN; /* as ldtw_one: a non-finite sample */
    float d = q[0] - x[0];
    d = l2 ? d * d : (d < 0 ? -d : d);
    while (i + 1 != M || j + 1 != M) {
        float dg = LDTW_INFTY, dn = LDTW_INFTY, rt = LDTW_INFTY;
        const int has_i = i + 1 < M, has_j = j + 1 < M;
        if (has_i && has_j) {
            dg = q[i + 1] - x[j + 1];
            dg = l2 ? dg * dg : (dg < 0 ? -dg : dg);
        }
        if (has_i && abs(i + 1 - j) <= W) {
            dn = q[i + 1] - x[j];
            dn = l2 ? dn * dn : (dn < 0 ? -dn : dn);
        }
        if (has_j && abs(j + 1 - i) <= W) {
            rt = q[i] - x[j + 1];
            rt = l2 ? rt * rt : (rt < 0 ? -rt : rt);
        }
        /* diagonal only if strictly below both; else down if the diagonal is
           not below it and it is strictly below right; else right */
        if (dg < dn && dg < rt) {
            d += dg;
            ++i;
            ++j;
        } else if (!(dg < dn) && dn < rt) {
            d += dn;
            ++i;
        } else {
            d += rt;
            ++j;
        }
    }
    return d;
}

void or_ldtw_raw(const float* q, const float* s, int32_t m, const int32_t* idx,
                 int64_t count, int32_t w, int32_t metric, float* out) {
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < count; ++g) {
        int64_t ix = idx ? idx[g] : g;
        out[g] = ldtw_raw_one(q, s + ix, m, w, metric);
    }
}

/* ------------------------------------------------------------------------ */
void or_ed_f64(const double* zq, const float* s, int64_t n, int32_t m,
               const double* avg64, const double* sd64, double* out) {
    int64_t L = n - m + 1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < L; ++i) {
        double acc = 0.0, a = avg64[i], sdv = sd64[i];
        for (int32_t k = 0; k < m; ++k) {
            double z = ((double)s[i + k] - a) / sdv;
            double d = zq[k] - z;
            acc += d * d;
        }
        out[i] = acc;
    }
}

int64_t or_argmin(const float* x, int64_t n) {
    float bsf = ORC_INF;
    int64_t bi = 0;
    for (int64_t i = 0; i < n; ++i)
        if (x[i] < bsf) {
            bsf = x[i];
            bi = i;
        }
    return bi;
}

/* ------------------------------------------------------------------------ */
static const float* g_sort_d;
static int cmp_dist_idx(const void* a, const void* b) {
    int64_t i = *(const int64_t*)a, j = *(const int64_t*)b;
    float di = g_sort_d[i], dj = g_sort_d[j];
    if (di < dj) return -1;
    if (di > dj) return 1;
    return i < j ? -1 : (i > j ? 1 : 0);
}

int32_t or_topk_greedy(const float* d, int64_t n, int32_t k, int64_t excl,
                       int64_t* out_idx, float* out_dist) {
    int64_t* ord = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int64_t nn = 0;
    for (int64_t i = 0; i < n; ++i)
        if (d[i] == d[i]) ord[nn++] = i; /* NaN distances are never picked */
    n = nn;
    g_sort_d = d;
    qsort(ord, (size_t)n, sizeof(int64_t), cmp_dist_idx);
    int32_t got = 0;
    for (int64_t t = 0; t < n && got < k; ++t) {
        int64_t c = ord[t];
        int ok = 1;
        for (int32_t p = 0; p < got; ++p) {
            int64_t dd = c - out_idx[p];
            if (dd < 0) dd = -dd;
            if (dd < excl) {
                ok = 0;
                break;
            }
        }
        if (ok) {
            out_idx[got] = c;
            out_dist[got] = d[c];
            got++;
        }
    }
    free(ord);
    return got;
}

/* ------------------------------------------------------------------------ */
/* FP64 streams (the reference's ED driver, GPU-ed.cu:20-82): two-pass FP64
 * window statistics (sigma := 1 for a constant window, NaN for a window
 * holding a non-finite sample), and the z-normalised squared ED of every
 * window, sum_k (q_k - (x_{i+k} - mu_i) / sigma_i)^2 (ed.cuh:61-69 computes
 * the same quantity as 2M - 2 corr / sigma through FFTs). */
void or_stats_f64d(const double* s, int64_t n, int32_t m, double* avg, double* sd) {
    int64_t L = n - m + 1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < L; ++i) {
        int bad = 0;
        double sum = 0.0;
        for (int32_t k = 0; k < m; ++k) {
            bad |= !isfinite(s[i + k]);
            sum += s[i + k];
        }
        if (bad) {
            avg[i] = sd[i] = NAN;
            continue;
        }
        const double mean = sum / (double)m;
        double ss = 0.0;
        for (int32_t k = 0; k < m; ++k) {
            const double d = s[i + k] - mean;
            ss += d * d;
        }
        const double var = ss / (double)m;
        avg[i] = mean;
        sd[i] = var > 0.0 ? sqrt(var) : 1.0;
    }
}

void or_ed_f64d(const double* zq, const double* s, int64_t n, int32_t m, const double* avg,
                const double* sd, double* out) {
    int64_t L = n - m + 1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < L; ++i) {
        double acc = 0.0;
        for (int32_t k = 0; k < m; ++k) {
            const double d = zq[k] - (s[i + k] - avg[i]) / sd[i];
            acc += d * d;
        }
        out[i] = acc;
    }
}

/* Host threads of the OpenMP loops above (the one-core timings of the CPU
 * baseline, the reference's sequential_cdtw and z_stream_ldtw). */
void or_set_threads(int32_t n) { omp_set_num_threads(n > 0 ? n : 1); }
int32_t or_max_threads(void) { return omp_get_max_threads(); }
