/*
 * cascade.c -- the UCR-suite subsequence search, restated from its published
 * description (Rakthanmanon, Campana, Mueen, Batista, Westover, Zhu, Zakaria,
 * Keogh: "Searching and Mining Trillions of Time Series Subsequences under
 * Dynamic Time Warping", KDD 2012, section 4), as a CPU baseline and a
 * cross-check.
 *
 * TEST INFRASTRUCTURE ONLY: built into liboracle.so with oracle.c; used by
 * tests/ and bench.py's CPU-baseline leg, never by the product library.
 *
 * The reference ships this search as UCR Suite/UCR_DTW_BIN.cpp and
 * UCR_ED_BIN.cpp, whose README (UCR Suite/README.md:10) restricts reuse; the
 * code here follows the paper, not those files.  Its results are pinned to
 * those programs' printed (location, distance) on seeded inputs
 * (tests/golden/ucr_golden.json, written by running them, compiled from
 * their sources, in the build container).
 *
 * The search, per the paper:
 *   - the query is z-normalised once; its band-r envelope (U, L) is built
 *     with a monotone-deque sliding max/min (Lemire 2009), and its
 *     positions are visited in order of decreasing |q| ("reordering early
 *     abandoning": large |q| terms contribute most, so bounds cross the
 *     best-so-far sooner);
 *   - the data are processed in blocks; within a block, every window's mean
 *     and sigma come from running sums of x and x^2 ("online
 *     z-normalisation"), restarted at each block so their drift stays small;
 *   - each window passes a cascade of lower bounds, each stopped as soon as
 *     its partial sum reaches the best-so-far: LB_Kim over the first and last
 *     three points (the hierarchy of the paper's LB_KimFL), LB_Keogh of the
 *     window against the query envelope (EQ), then of the query against the
 *     window's envelope (EC, the block's data envelope normalised by the
 *     window's statistics);
 *   - survivors run the banded DTW, abandoned when a row minimum plus the
 *     suffix of the larger LB_Keogh's per-position terms reaches the
 *     best-so-far.
 * Distances are squared; ties keep the first window (strict improvement).
 */
#define _GNU_SOURCE /* qsort_r */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define CASCADE_INF 1e20
#define CASCADE_BLOCK 131072 /* windows per block of running sums */

static inline double sq(double a) { return a * a; }

/* Sliding max (up) and min (lo) of x over [i - r, i + r] clamped to [0, n). */
static void band_envelope(const double* x, int64_t n, int r, double* lo, double* up,
                          int64_t* qa, int64_t* qb) {
    int64_t ha = 0, ta = 0, hb = 0, tb = 0; /* deques qa (max) and qb (min) */
    for (int64_t j = 0; j < n + r; ++j) {
        if (j < n) {
            while (ta > ha && x[qa[ta - 1]] <= x[j]) --ta;
            qa[ta++] = j;
            while (tb > hb && x[qb[tb - 1]] >= x[j]) --tb;
            qb[tb++] = j;
        }
        const int64_t c = j - r; /* centre whose window [c - r, c + r] ends at j */
        if (c < 0) continue;
        while (qa[ha] < c - r) ++ha;
        while (qb[hb] < c - r) ++hb;
        up[c] = x[qa[ha]];
        lo[c] = x[qb[hb]];
    }
}

typedef struct {
    int m, r;
    double* q;   /* z-normalised query */
    int* ord;    /* positions by decreasing |q| */
    double* qo;  /* q[ord[k]] */
    double* uo;  /* U[ord[k]] */
    double* lo_; /* L[ord[k]] */
} plan_t;

typedef struct {
    double *z, *tq, *td, *suffix, *row_a, *row_b; /* per-window scratch */
    double *dlo, *dup;                            /* block data envelope */
    int64_t *da, *db;                             /* deques */
} work_t;

static int by_abs_desc(const void* a, const void* b, void* q) {
    const double* v = (const double*)q;
    const int i = *(const int*)a, j = *(const int*)b;
    const double x = fabs(v[i]), y = fabs(v[j]);
    if (x != y) return x > y ? -1 : 1;
    return i - j;
}

/* Query z-normalisation with population sigma (the paper's eq. 2). */
static void znorm_query(const double* query, int m, double* q) {
    double s = 0.0, s2 = 0.0;
    for (int i = 0; i < m; ++i) {
        s += query[i];
        s2 += query[i] * query[i];
    }
    const double mu = s / m;
    const double sd = sqrt(s2 / m - mu * mu);
    for (int i = 0; i < m; ++i) q[i] = (query[i] - mu) / sd;
}

static void plan_make(plan_t* p, const double* query, int m, int r) {
    p->m = m;
    p->r = r;
    p->q = malloc(sizeof(double) * m);
    p->ord = malloc(sizeof(int) * m);
    p->qo = malloc(sizeof(double) * m);
    p->uo = malloc(sizeof(double) * m);
    p->lo_ = malloc(sizeof(double) * m);
    znorm_query(query, m, p->q);
    double* u = malloc(sizeof(double) * m);
    double* l = malloc(sizeof(double) * m);
    int64_t* qa = malloc(sizeof(int64_t) * (m + 1));
    int64_t* qb = malloc(sizeof(int64_t) * (m + 1));
    band_envelope(p->q, m, r, l, u, qa, qb);
    for (int i = 0; i < m; ++i) p->ord[i] = i;
    qsort_r(p->ord, (size_t)m, sizeof(int), by_abs_desc, p->q);
    for (int k = 0; k < m; ++k) {
        p->qo[k] = p->q[p->ord[k]];
        p->uo[k] = u[p->ord[k]];
        p->lo_[k] = l[p->ord[k]];
    }
    free(u);
    free(l);
    free(qa);
    free(qb);
}

static void plan_free(plan_t* p) {
    free(p->q);
    free(p->ord);
    free(p->qo);
    free(p->uo);
    free(p->lo_);
}

static void work_make(work_t* w, int m, int r, int64_t block_samples) {
    w->z = malloc(sizeof(double) * m);
    w->tq = malloc(sizeof(double) * m);
    w->td = malloc(sizeof(double) * m);
    w->suffix = malloc(sizeof(double) * (m + 1));
    w->row_a = malloc(sizeof(double) * m);
    w->row_b = malloc(sizeof(double) * m);
    w->dlo = malloc(sizeof(double) * block_samples);
    w->dup = malloc(sizeof(double) * block_samples);
    w->da = malloc(sizeof(int64_t) * (block_samples + 1));
    w->db = malloc(sizeof(int64_t) * (block_samples + 1));
    (void)r;
}

static void work_free(work_t* w) {
    free(w->z);
    free(w->tq);
    free(w->td);
    free(w->suffix);
    free(w->row_a);
    free(w->row_b);
    free(w->dlo);
    free(w->dup);
    free(w->da);
    free(w->db);
}

/* LB_Kim over the first and last three points of the z-normalised window x
 * (raw samples, normalised here), abandoned at bsf. */
static double lb_kim_fl(const double* x, double mu, double sd, const double* q, int m,
                        double bsf) {
#define ZN(i) ((x[(i)] - mu) / sd)
    const double a0 = ZN(0), b0 = ZN(m - 1);
    double lb = sq(a0 - q[0]) + sq(b0 - q[m - 1]);
    if (lb >= bsf || m < 6) return lb;
    const double a1 = ZN(1), b1 = ZN(m - 2);
    double t = fmin(sq(a1 - q[0]), fmin(sq(a0 - q[1]), sq(a1 - q[1])));
    lb += t;
    if (lb >= bsf) return lb;
    t = fmin(sq(b1 - q[m - 1]), fmin(sq(b0 - q[m - 2]), sq(b1 - q[m - 2])));
    lb += t;
    if (lb >= bsf) return lb;
    const double a2 = ZN(2), b2 = ZN(m - 3);
    t = fmin(fmin(sq(a0 - q[2]), sq(a1 - q[2])),
             fmin(sq(a2 - q[2]), fmin(sq(a2 - q[1]), sq(a2 - q[0]))));
    lb += t;
    if (lb >= bsf) return lb;
    t = fmin(fmin(sq(b0 - q[m - 3]), sq(b1 - q[m - 3])),
             fmin(sq(b2 - q[m - 3]), fmin(sq(b2 - q[m - 2]), sq(b2 - q[m - 1]))));
    return lb + t;
#undef ZN
}

/* Banded DTW of the window z against the query q, abandoned once a row's
 * minimum plus the bound on the rows still to come reaches bsf
 * (suffix[i] bounds rows i.. from below). */
static double dtw_abandon(const double* z, const double* q, const double* suffix, int m, int r,
                          double bsf, double* prev, double* cur) {
    for (int j = 0; j < m; ++j) prev[j] = CASCADE_INF;
    for (int i = 0; i < m; ++i) {
        const int j0 = i - r > 0 ? i - r : 0;
        const int j1 = i + r < m - 1 ? i + r : m - 1;
        double rowmin = CASCADE_INF;
        if (j0 > 0) cur[j0 - 1] = CASCADE_INF;
        for (int j = j0; j <= j1; ++j) {
            double best;
            if (i == 0 && j == 0) {
                best = 0.0;
            } else {
                const double up = prev[j];
                const double left = j > 0 ? cur[j - 1] : CASCADE_INF;
                const double diag = j > 0 ? prev[j - 1] : CASCADE_INF;
                best = fmin(diag, fmin(up, left));
            }
            const double d = best + sq(z[i] - q[j]);
            cur[j] = d;
            if (d < rowmin) rowmin = d;
        }
        if (j1 + 1 < m) cur[j1 + 1] = CASCADE_INF;
        if (i + r + 1 < m && rowmin + suffix[i + r + 1] >= bsf) return CASCADE_INF;
        /* Row i + 1 reads row i at columns >= i - r (in its band, or the
         * +inf just left of it) and at i + r + 1 (the +inf just right). */
        double* t = prev;
        prev = cur;
        cur = t;
    }
    return prev[m - 1];
}

/* Windows [w_begin, w_end) of data[0..n): best (squared distance, window) into
 * *bsf / *loc, which hold the best so far on entry (strict improvement keeps
 * the first).  counts: LB_Kim, LB_Keogh EQ, LB_Keogh EC prunes added. */
static void search_range(const double* data, int64_t n, const plan_t* p, work_t* w,
                         int64_t w_begin, int64_t w_end, volatile double* bsf_shared,
                         double* bsf, int64_t* loc, int64_t* counts) {
    const int m = p->m, r = p->r;
    (void)n; /* windows [w_begin, w_end) read data[w_begin .. w_end + m - 1) < n */
    for (int64_t b0 = w_begin; b0 < w_end; b0 += CASCADE_BLOCK) {
        const int64_t b1 = b0 + CASCADE_BLOCK < w_end ? b0 + CASCADE_BLOCK : w_end;
        const int64_t ns = b1 - b0 + m - 1; /* samples of the block */
        const double* x = data + b0;
        band_envelope(x, ns, r, w->dlo, w->dup, w->da, w->db);
        double s = 0.0, s2 = 0.0;
        for (int k = 0; k < m - 1; ++k) {
            s += x[k];
            s2 += x[k] * x[k];
        }
        for (int64_t i = 0; i < b1 - b0; ++i) {
            const double xin = x[i + m - 1];
            s += xin;
            s2 += xin * xin;
            const double mu = s / m;
            const double sd = sqrt(s2 / m - mu * mu);
            const double* xw = x + i;
            double best = *bsf;
            if (bsf_shared) {
                const double g = *bsf_shared;
                if (g < best) best = g;
            }
            double lb = lb_kim_fl(xw, mu, sd, p->q, m, best);
            if (lb >= best) {
                counts[0]++;
            } else {
                /* LB_Keogh EQ: window against the query envelope */
                double eq = 0.0;
                for (int k = 0; k < m; ++k) w->tq[k] = 0.0;
                for (int k = 0; k < m && eq < best; ++k) {
                    const int pos = p->ord[k];
                    const double z = (xw[pos] - mu) / sd;
                    double c = 0.0;
                    if (z > p->uo[k]) c = sq(z - p->uo[k]);
                    else if (z < p->lo_[k]) c = sq(z - p->lo_[k]);
                    w->tq[pos] = c;
                    eq += c;
                }
                if (eq >= best) {
                    counts[1]++;
                } else {
                    /* LB_Keogh EC: query against the window's own envelope */
                    double ec = 0.0;
                    for (int k = 0; k < m; ++k) w->td[k] = 0.0;
                    for (int k = 0; k < m && ec < best; ++k) {
                        const int pos = p->ord[k];
                        const double u = (w->dup[i + pos] - mu) / sd;
                        const double l = (w->dlo[i + pos] - mu) / sd;
                        const double qv = p->qo[k];
                        double c = 0.0;
                        if (qv > u) c = sq(qv - u);
                        else if (qv < l) c = sq(qv - l);
                        w->td[pos] = c;
                        ec += c;
                    }
                    if (ec >= best) {
                        counts[2]++;
                    } else {
                        const double* terms = eq > ec ? w->tq : w->td;
                        w->suffix[m] = 0.0;
                        for (int k = m - 1; k >= 0; --k) w->suffix[k] = w->suffix[k + 1] + terms[k];
                        for (int k = 0; k < m; ++k) w->z[k] = (xw[k] - mu) / sd;
                        const double d =
                            dtw_abandon(w->z, p->q, w->suffix, m, r, best, w->row_a, w->row_b);
                        if (d < *bsf) {
                            *bsf = d;
                            *loc = b0 + i;
                            if (bsf_shared) {
#pragma omp critical(cascade_bsf)
                                if (d < *bsf_shared) *bsf_shared = d;
                            }
                        }
                    }
                }
            }
            const double xout = x[i];
            s -= xout;
            s2 -= xout * xout;
        }
    }
}

int64_t or_ucr_dtw(const double* data, int64_t n, const double* query, int32_t m, int32_t r,
                   double* dist, int64_t* counts) {
    plan_t p;
    work_t w;
    if (r > m - 1) r = m - 1;
    plan_make(&p, query, m, r);
    work_make(&w, m, r, CASCADE_BLOCK + m);
    double bsf = CASCADE_INF;
    int64_t loc = 0, c[3] = {0, 0, 0};
    const int64_t L = n - m + 1;
    if (L > 0) search_range(data, n, &p, &w, 0, L, NULL, &bsf, &loc, c);
    if (dist) *dist = bsf;
    if (counts) {
        counts[0] = c[0];
        counts[1] = c[1];
        counts[2] = c[2];
        counts[3] = L > 0 ? L : 0;
    }
    work_free(&w);
    plan_free(&p);
    return loc;
}

/* The same search on `threads` host threads: the windows split into that
 * many contiguous ranges (each reading M - 1 samples past its end), one
 * best-so-far shared for pruning; the answer is the smallest distance, the
 * lowest window among equal ones (an exact tie across ranges may be pruned
 * in one of them, so equal distances can report another window than the
 * one-thread search). */
int64_t or_ucr_dtw_par(const double* data, int64_t n, const double* query, int32_t m,
                       int32_t r, int32_t threads, double* dist, int64_t* counts) {
    if (r > m - 1) r = m - 1;
    if (threads < 1) threads = 1;
    plan_t p;
    plan_make(&p, query, m, r);
    const int64_t L = n - m + 1;
    volatile double shared = CASCADE_INF;
    double best_d = CASCADE_INF;
    int64_t best_loc = 0;
    int64_t c[3] = {0, 0, 0};
#pragma omp parallel num_threads(threads)
    {
        work_t w;
        work_make(&w, m, r, CASCADE_BLOCK + m);
#pragma omp for schedule(dynamic, 1)
        for (int t = 0; t < threads * 8; ++t) {
            const int64_t a = L * t / (threads * 8), b = L * (t + 1) / (threads * 8);
            double bsf = CASCADE_INF;
            int64_t loc = 0, cc[3] = {0, 0, 0};
            if (b > a) search_range(data, n, &p, &w, a, b, &shared, &bsf, &loc, cc);
#pragma omp critical(cascade_result)
            {
                if (bsf < best_d || (bsf == best_d && loc < best_loc)) {
                    best_d = bsf;
                    best_loc = loc;
                }
                c[0] += cc[0];
                c[1] += cc[1];
                c[2] += cc[2];
            }
        }
        work_free(&w);
    }
    if (dist) *dist = best_d;
    if (counts) {
        counts[0] = c[0];
        counts[1] = c[1];
        counts[2] = c[2];
        counts[3] = L > 0 ? L : 0;
    }
    plan_free(&p);
    return best_loc;
}

/* z-normalised ED search (the paper's section 4.1 techniques without the DTW
 * bounds): online statistics, query positions by decreasing |q|, early
 * abandoning at the best-so-far. */
int64_t or_ucr_ed(const double* data, int64_t n, const double* query, int32_t m, double* dist) {
    double* q = malloc(sizeof(double) * m);
    int* ord = malloc(sizeof(int) * m);
    znorm_query(query, m, q);
    for (int i = 0; i < m; ++i) ord[i] = i;
    qsort_r(ord, (size_t)m, sizeof(int), by_abs_desc, q);
    double bsf = CASCADE_INF;
    int64_t loc = 0;
    const int64_t L = n - m + 1;
    for (int64_t b0 = 0; b0 < L; b0 += CASCADE_BLOCK) {
        const int64_t b1 = b0 + CASCADE_BLOCK < L ? b0 + CASCADE_BLOCK : L;
        const double* x = data + b0;
        double s = 0.0, s2 = 0.0;
        for (int k = 0; k < m - 1; ++k) {
            s += x[k];
            s2 += x[k] * x[k];
        }
        for (int64_t i = 0; i < b1 - b0; ++i) {
            s += x[i + m - 1];
            s2 += x[i + m - 1] * x[i + m - 1];
            const double mu = s / m;
            const double sd = sqrt(s2 / m - mu * mu);
            double acc = 0.0;
            for (int k = 0; k < m && acc < bsf; ++k) {
                const int pos = ord[k];
                acc += sq((x[i + pos] - mu) / sd - q[pos]);
            }
            if (acc < bsf) {
                bsf = acc;
                loc = b0 + i;
            }
            s -= x[i];
            s2 -= x[i] * x[i];
        }
    }
    if (dist) *dist = bsf;
    free(q);
    free(ord);
    return loc;
}
