/*
 * CPU oracle for the discord engine -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C, the reference's CPU path for range
 * discords and the PALMAD length loop that drives it:
 *
 *   oracle_stats_f64 / _f32         compute_statistics_cpu   src/cpu_kernels/cpu_kernels.cpp:7-31
 *   oracle_find_discords_f64 / _f32 find_discords_cpu        src/cpu_kernels/cpu_kernels.cpp:33-78
 *   oracle_topk                     find_top1_discord        src/sets_processing.hpp:11-46
 *                                   (generalised to k > 1, greedy non-overlapping)
 *   oracle_palmad                   PALMAD r-descent         src/main.cu:35-233
 *
 * Only the test suite, the smoke check and the CPU-baseline leg of the bench
 * load this library, and only as the checker.  The product path (the HIP
 * library) never calls into it.
 *
 * Semantics fixed here and shared with the GPU engine:
 *   - distances are squared z-normalised Euclidean distances (Discord.dist);
 *   - row i and column j form a valid pair iff |i - 1 - j| >= m
 *     (cpu_kernels.cpp:57), i.e. j - i >= m-1 or i - j >= m+1;
 *   - a row is pruned at the first valid pair with ED^2 <= r
 *     (cpu_kernels.cpp:59-72), where r is the SQUARED range;
 *   - a window whose standard deviation is zero (or whose variance comes out
 *     negative / NaN) has correlation 0 with everything, i.e. ED^2 = 2m.  The
 *     reference CPU code divides by zero there; its GPU kernels clamp to 2m
 *     (src/PD3_kernels.cuh:63-64).  The clamp is adopted.
 *
 * The f64 variants evaluate the reference's own formulas in double precision
 * (the parity oracle).  The f32 variants evaluate them literally in float and
 * exist as a fidelity witness only.
 *
 * Parity of this restatement is pinned by tests/test_oracle.py against
 * closed forms, an independent brute-force matrix profile (direct
 * z-normalised vectors, not the QT identity) and the committed fixtures in
 * tests/golden/.  The reference cannot be compiled here (its cpu_kernels.hpp
 * includes PD3_types.hpp, which the reference does not ship).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- sliding statistics: cpu_kernels.cpp:7-31 ------------------------- */

void oracle_stats_f64(const float *T, int64_t N, int32_t m, double *mean, double *std)
{
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; i++) {
        double s = 0.0, s2 = 0.0;
        for (int32_t k = 0; k < m; k++) {
            double x = (double)T[i + k];
            s += x;
            s2 += x * x;
        }
        double mu = s / m;
        double var = s2 / m - mu * mu;
        mean[i] = mu;
        std[i] = (var > 0.0) ? sqrt(var) : 0.0;
    }
}

void oracle_stats_f32(const float *T, int64_t N, int32_t m, float *mean, float *std)
{
    for (int64_t i = 0; i < N; i++) {
        float mu = 0.0f, s2 = 0.0f;
        for (int32_t k = 0; k < m; k++) {
            mu += T[i + k];
            s2 += T[i + k] * T[i + k];
        }
        mu = mu / m;
        s2 = s2 / m;
        float sd = (float)sqrt(s2 - mu * mu);
        mean[i] = mu;
        std[i] = sd;
    }
}

/* ---- range discords at squared range r: cpu_kernels.cpp:33-78 ---------- */

static inline int valid_pair(int64_t i, int64_t j, int32_t m)
{
    /* cpu_kernels.cpp:57: not_overlap = |i - 1 - j| >= m */
    int64_t d = i - 1 - j;
    if (d < 0) d = -d;
    return d >= m;
}

/* The pair value from its dot product (cpu_kernels.cpp:55 in double precision). */
static inline double ed2_from_dot(double dot, const double *mean, const double *std,
                                  int64_t i, int64_t j, int32_t m)
{
    /* Products of the two rows' statistics are formed first, so the value is bitwise
     * symmetric in (i, j): mutual nearest neighbours tie exactly and the tie goes to the
     * lower index, as find_top1_discord's strict '>' intends (sets_processing.hpp:28). */
    double den = (double)m * (std[i] * std[j]);
    if (!(den > 0.0))
        return 2.0 * m;
    return 2.0 * m * (1.0 - (dot - m * (mean[i] * mean[j])) / den);
}

static inline double ed2_f64(const float *T, const double *mean, const double *std,
                             int64_t i, int64_t j, int32_t m)
{
    double dot = 0.0;
    for (int32_t k = 0; k < m; k++)
        dot += (double)T[i + k] * (double)T[j + k];
    return ed2_from_dot(dot, mean, std, i, j, m);
}

/*
 * Row i against columns j0 .. j1-1 in ascending order, as the loop of
 * cpu_kernels.cpp:47-75: valid pairs with ED^2 > r lower *nn, the first valid
 * pair with ED^2 <= r ends the scan.  Returns that column, or -1.  Columns are
 * evaluated eight at a time (eight independent dot products, each summed in
 * ascending k exactly as ed2_f64 does) and then consumed in order, so every
 * pair value and every decision is bitwise the one-column-at-a-time loop's;
 * the block only lets the compiler run the eight sums side by side.
 */
#define OJB 8
static int64_t scan_row(const float *T, const double *mean, const double *std, int64_t i,
                        int64_t j0, int64_t j1, int64_t N, double r, int32_t m, double *nn)
{
    int64_t j = j0;
    while (j < j1) {
        if (j + OJB <= j1 && j + OJB <= N) {
            int any = 0;
            for (int q = 0; q < OJB; q++) any |= valid_pair(i, j + q, m);
            if (!any) { j += OJB; continue; }
            double acc[OJB];
            for (int q = 0; q < OJB; q++) acc[q] = 0.0;
            const float *ti = T + i, *tj = T + j;
            for (int32_t k = 0; k < m; k++) {
                const double xi = (double)ti[k];
                for (int q = 0; q < OJB; q++) acc[q] += xi * (double)tj[k + q];
            }
            for (int q = 0; q < OJB; q++) {
                if (!valid_pair(i, j + q, m)) continue;
                const double ed = ed2_from_dot(acc[q], mean, std, i, j + q, m);
                if (ed > r) {
                    if (ed < *nn) *nn = ed;
                } else {
                    return j + q;
                }
            }
            j += OJB;
        } else {
            if (valid_pair(i, j, m)) {
                const double ed = ed2_f64(T, mean, std, i, j, m);
                if (ed > r) {
                    if (ed < *nn) *nn = ed;
                } else {
                    return j;
                }
            }
            j++;
        }
    }
    return -1;
}

/*
 * cand[] must be initialised to 1 and nnDist[] to a large value by the caller
 * (main-check.cu:143-148).  On return cand[i] == 1 iff every valid pair of row
 * i has ED^2 > r, and then nnDist[i] is the exact nearest-neighbour ED^2.
 * Rows are independent; nthreads <= 0 uses the OpenMP default.
 */
void oracle_find_discords_f64(const float *T, const double *mean, const double *std,
                              int32_t *cand, double *nnDist, int64_t N, double r,
                              int32_t m, int32_t nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < N; i++) {
        if (scan_row(T, mean, std, i, 0, N, N, r, m, &nnDist[i]) >= 0) cand[i] = 0;
    }
}

/*
 * oracle_find_discords_f64 on a sample of the rows: the rows of every stride-th run of
 * `run` consecutive rows (runs 0, stride, 2 stride, ..), each against all columns in the
 * reference's order; the other rows' cand/nnDist are left as they are.  Rows are independent
 * (cpu_kernels.cpp:42-76), so the time of the sample times `stride` estimates the time of
 * the whole range pass (the CPU baseline of long searches).
 */
void oracle_find_discords_runs_f64(const float *T, const double *mean, const double *std,
                                   int32_t *cand, double *nnDist, int64_t N, double r,
                                   int32_t m, int32_t nthreads, int64_t run, int64_t stride)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    const int64_t nruns = (N + run - 1) / run;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int64_t q = 0; q < nruns; q += stride) {
        const int64_t e = (q + 1) * run < N ? (q + 1) * run : N;
        for (int64_t i = q * run; i < e; i++)
            if (scan_row(T, mean, std, i, 0, N, N, r, m, &nnDist[i]) >= 0) cand[i] = 0;
    }
}

/*
 * The same range-discord set, for large series (the certificates of long
 * searches): cand[i] and, where cand[i] == 1, nnDist[i] equal
 * oracle_find_discords_f64's bit for bit, since neither depends on the order
 * in which a row meets its columns (cand[i] == 1 iff no valid pair has
 * ED^2 <= r, and nnDist[i] is then the minimum over all valid pairs).  Each
 * thread walks runs of 256 consecutive rows and lets row i first try column
 * p + 1, where p is the column that pruned row i - 1 (windows i-1 and p are
 * close, so windows i and p + 1, which share m - 1 points with them, usually
 * are too); if that pair does not prune, the row is scanned from column 0 as
 * above.  nnDist of pruned rows is unspecified (partial minima), as in the
 * reference.
 */
void oracle_find_discords_hint_f64(const float *T, const double *mean, const double *std,
                                   int32_t *cand, double *nnDist, int64_t N, double r,
                                   int32_t m, int32_t nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    const int64_t RUN = 256;
    const int64_t nruns = (N + RUN - 1) / RUN;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int64_t q = 0; q < nruns; q++) {
        int64_t p = -1;
        const int64_t e = (q + 1) * RUN < N ? (q + 1) * RUN : N;
        for (int64_t i = q * RUN; i < e; i++) {
            int64_t h = p >= 0 ? p + 1 : -1;
            int64_t hit = -1;
            if (h >= 0 && h < N && valid_pair(i, h, m)) {
                double dummy = DBL_MAX;
                hit = scan_row(T, mean, std, i, h, h + 1, N, r, m, &dummy);
            }
            if (hit < 0) hit = scan_row(T, mean, std, i, 0, N, N, r, m, &nnDist[i]);
            if (hit >= 0) cand[i] = 0;
            p = hit;
        }
    }
}

void oracle_find_discords_f32(const float *T, const float *means, const float *stds,
                              int32_t *cand, float *nnDist, int64_t N, float r, int32_t m)
{
    for (int64_t i = 0; i < N; i++) {
        int64_t j = 0;
        int not_pruned = 1;
        while (not_pruned && j < N) {
            float dot = 0.0f;
            for (int32_t k = 0; k < m; k++)
                dot += T[i + k] * T[j + k];
            float ed = 2 * (uint32_t)m *
                       (1 - (dot - m * means[i] * means[j]) / (m * stds[i] * stds[j]));
            if (valid_pair(i, j, m)) {
                if (ed > r) {
                    if (ed < nnDist[i]) nnDist[i] = ed;
                } else {
                    cand[i] = 0;
                    not_pruned = 0;
                }
            }
            j++;
        }
    }
}

/* Exact nearest-neighbour ED^2 of every row (r = -inf): the matrix profile
 * under the oracle's exclusion rule.  Used for certificates on small inputs. */
void oracle_profile_f64(const float *T, const double *mean, const double *std,
                        double *nnDist, int64_t N, int32_t m, int32_t nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < N; i++) {
        double best = DBL_MAX;
        (void)scan_row(T, mean, std, i, 0, N, N, -INFINITY, m, &best);
        nnDist[i] = best;
    }
}

/* ---- top-k: sets_processing.hpp:11-46, generalised --------------------- */

typedef struct { int64_t idx; double d; } od_item;

static int od_cmp(const void *a, const void *b)
{
    const od_item *x = (const od_item *)a, *y = (const od_item *)b;
    if (x->d > y->d) return -1;
    if (x->d < y->d) return 1;
    return (x->idx < y->idx) ? -1 : (x->idx > y->idx);
}

/*
 * Range discords at r are rows with cand == 1 and nnDist > r
 * (sets_processing.hpp:28,37).  They are ranked by (nnDist desc, index asc) --
 * for k == 1 this is exactly find_top1_discord's strict-greater scan -- and
 * accepted greedily when |i - a| >= m for every accepted a.  Returns the
 * number accepted (<= k) and the number of range discords in *ndiscords.
 */
int32_t oracle_topk(const int32_t *cand, const double *nnDist, int64_t N, double r,
                    int32_t m, int32_t k, int64_t *out_idx, double *out_d,
                    int64_t *ndiscords)
{
    int64_t cnt = 0;
    for (int64_t i = 0; i < N; i++)
        if (cand[i] && nnDist[i] > r) cnt++;
    if (ndiscords) *ndiscords = cnt;
    if (cnt == 0) return 0;
    od_item *it = (od_item *)malloc(sizeof(od_item) * (size_t)cnt);
    int64_t c = 0;
    for (int64_t i = 0; i < N; i++)
        if (cand[i] && nnDist[i] > r) { it[c].idx = i; it[c].d = nnDist[i]; c++; }
    qsort(it, (size_t)cnt, sizeof(od_item), od_cmp);
    int32_t acc = 0;
    for (int64_t q = 0; q < cnt && acc < k; q++) {
        int ok = 1;
        for (int32_t a = 0; a < acc; a++) {
            int64_t dd = it[q].idx - out_idx[a];
            if (dd < 0) dd = -dd;
            if (dd < m) { ok = 0; break; }
        }
        if (ok) { out_idx[acc] = it[q].idx; out_d[acc] = it[q].d; acc++; }
    }
    free(it);
    return acc;
}

/* ---- PALMAD length loop: main.cu:35-233 --------------------------------- */

/*
 * The r-schedule of main.cu, over lengths m_j = minL + j*step:
 *   j = 0      r = 2*sqrt(minL); retry with r *= 0.5            (main.cu:38,155)
 *   j = 1..4   r = 0.99*sqrt(d[j-1]); retry with r *= 0.99      (main.cu:171,187)
 *   j >= 5     M, S = mean, population std of sqrt(d[j-5..j-1]);
 *              r = M - 2S; retry with r -= S                    (main.cu:202-214,230)
 * where d[j] is the squared distance of the k-th accepted discord (the top-1
 * distance when k == 1, as in the reference).  Guards the reference lacks:
 * r is clamped at 0 (kernels take r*r, so a negative r would square to a
 * positive range); when S == 0 the retry step falls back to r *= 0.99; after
 * max_attempts failures the next attempt uses r = 0.  An attempt at r = 0
 * that still finds fewer than k discords ends the length with what it found.
 *
 * Outputs per length j: out_idx[j*k + q], out_d[j*k + q] (q < out_cnt[j]),
 * out_attempts[j], out_r[j] (the squared range of the final attempt).
 */
static double next_r_after_fail(int j, double r, double S)
{
    if (j == 0) return 0.5 * r;
    if (j < 5) return 0.99 * r;
    if (S > 0.0) return r - S;
    return 0.99 * r;
}

int32_t oracle_palmad(const float *T, int64_t n, int32_t minL, int32_t maxL, int32_t step,
                      int32_t k, int32_t max_attempts, int32_t nthreads,
                      int64_t *out_idx, double *out_d, int32_t *out_cnt,
                      int32_t *out_attempts, double *out_r)
{
    if (step <= 0 || minL < 2 || maxL < minL || k <= 0 || n < maxL) return -1;
    int32_t nlen = (maxL - minL) / step + 1;
    double *dk = (double *)calloc((size_t)nlen, sizeof(double));
    int64_t Nmax = n - minL + 1;
    double *mean = (double *)malloc(sizeof(double) * (size_t)Nmax);
    double *sd = (double *)malloc(sizeof(double) * (size_t)Nmax);
    double *nn = (double *)malloc(sizeof(double) * (size_t)Nmax);
    int32_t *cand = (int32_t *)malloc(sizeof(int32_t) * (size_t)Nmax);

    for (int32_t j = 0; j < nlen; j++) {
        int32_t m = minL + j * step;
        int64_t N = n - m + 1;
        oracle_stats_f64(T, N, m, mean, sd);
        double r, S = 0.0;
        if (j == 0) {
            r = 2.0 * sqrt((double)minL);
        } else if (j < 5) {
            r = 0.99 * sqrt(dk[j - 1]);
        } else {
            double M = 0.0;
            for (int32_t q = j - 5; q < j; q++) M += sqrt(dk[q]);
            M *= 0.2;
            for (int32_t q = j - 5; q < j; q++) {
                double e = sqrt(dk[q]) - M;
                S += e * e;
            }
            S = sqrt(S * 0.2);
            r = M - 2.0 * S;
        }
        int32_t att = 0;
        for (;;) {
            if (r < 0.0) r = 0.0;
            if (att >= max_attempts) r = 0.0;
            double r2 = r * r;
            for (int64_t i = 0; i < N; i++) { cand[i] = 1; nn[i] = DBL_MAX; }
            oracle_find_discords_f64(T, mean, sd, cand, nn, N, r2, m, nthreads);
            int32_t got = oracle_topk(cand, nn, N, r2, m, k, out_idx + (int64_t)j * k,
                                      out_d + (int64_t)j * k, NULL);
            att++;
            if (got >= k || r == 0.0) {
                out_cnt[j] = got;
                out_attempts[j] = att;
                out_r[j] = r2;
                dk[j] = got > 0 ? out_d[(int64_t)j * k + got - 1] : 0.0;
                break;
            }
            r = next_r_after_fail(j, r, S);
        }
    }
    free(dk); free(mean); free(sd); free(nn); free(cand);
    return nlen;
}
