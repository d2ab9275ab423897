// Hand-written gfx950 kernels of the discord engine.
//
//   k_prefix        blocked, locally centred FP64 prefix sums of the series (once per series)
//   k_derive        per-length window statistics and recurrence terms from the prefix sums
//                   (replaces gpu_compute_statistics / gpu_update_statistics,
//                    src/preprocessing.cuh:12-65)
//   k_scan_select   PD3 candidate selection over a bounded window right of each 64-row block
//                   (replaces gpu_local_candidate_select + gpu_local_define_candidates,
//                    src/PD3_kernels.cuh:9-274)
//   k_scan_refine   exhaustive scan of surviving blocks over the remaining columns, split in
//                   column chunks, early abandoning (replaces gpu_local_discords_refine,
//                   src/PD3_kernels.cuh:276-528, and closes its left-of-segment gap)
//   k_survivors     compaction of surviving rows (replaces sets_processing.hpp:49-113)
//   k_exact         double-precision nearest-neighbour distance of a few candidates
//                   (replaces gpu_multigpus_discords_refine, src/PD3_kernels.cuh:665-759)
//
// Numerics.  The series is never used raw in single precision: every dot product runs on
// data centred by a local reference, and the diagonal recurrence runs on the mean-centred
// covariance (cov(i,j) = cov(i-1,j-1) + df_i dg_j + dg_i df_j with
// df_i = (x[i+m-1]-x[i-1])/2, dg_i = (x[i+m-1]-mu_i) + (x[i-1]-mu_{i-1})), so the values
// stay of the order of m*sigma^2 whatever the level of the series.  A value reaches a lane
// after at most 63 recurrence steps from a directly computed seed.
//
// Layout.  One workgroup = one wave64 = one block of 64 consecutive rows (lane t owns row
// s+t).  At step j every lane is on the same column c+j, so column terms are wave-uniform
// (one broadcast LDS read per step) and each lane receives its diagonal predecessor
// cov(s+t-1, c+j-1) from lane t-1 through DPP wave_shr:1; lane 0 takes a directly computed
// seed cov(s, c+j).  Row pruning, row maxima and the alive mask are lane-local; column
// pruning is a ballot.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>
#include <math.h>

#include "kernels.h"

#define WAVE 64
#define DPP_WAVE_SHR1 0x138

namespace paddi {

__device__ __forceinline__ int ord_key(float f)
{
    int i = __float_as_int(f);
    return i >= 0 ? i : (i ^ 0x7fffffff);
}
__device__ __forceinline__ float ord_val(int k)
{
    return __int_as_float(k >= 0 ? k : (k ^ 0x7fffffff));
}
__device__ __forceinline__ unsigned long long ord_key64(double d)
{
    unsigned long long u = (unsigned long long)__double_as_longlong(d);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

// ---------------------------------------------------------------------------------------
// Prefix sums.  Block b covers x[b*PB, (b+1)*PB); c_b = x[b*PB].  P[b*(PB+1) + k] holds
// (sum, sum of squares) of (x - c_b) over x[b*PB, b*PB+k), k = 0..PB (k = PB is the total).
// Points at or beyond n count as c_b (contribute 0).

__global__ __launch_bounds__(256) void k_prefix(const float *__restrict__ x, int64_t n,
                                                double2 *__restrict__ P, double *__restrict__ refs)
{
    const int t = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * PB;
    const double c = (base < n) ? (double)x[base] : 0.0;
    constexpr int PER = PB / 256;
    double v1[PER], v2[PER];
    double s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int q = 0; q < PER; q++) {
        int64_t g = base + (int64_t)t * PER + q;
        double y = (g < n) ? ((double)x[g] - c) : 0.0;
        v1[q] = s1;
        v2[q] = s2;
        s1 += y;
        s2 += y * y;
    }
    // inclusive scan of the thread totals across the wave
    const int lane = t & 63, w = t >> 6;
    double i1 = s1, i2 = s2;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        double a = __shfl_up(i1, o, 64);
        double b = __shfl_up(i2, o, 64);
        if (lane >= o) { i1 += a; i2 += b; }
    }
    __shared__ double wt1[4], wt2[4];
    if (lane == 63) { wt1[w] = i1; wt2[w] = i2; }
    __syncthreads();
    double o1 = i1 - s1, o2 = i2 - s2;  // exclusive within the wave
    for (int q = 0; q < w; q++) { o1 += wt1[q]; o2 += wt2[q]; }
    double2 *Pb = P + (int64_t)blockIdx.x * (PB + 1);
#pragma unroll
    for (int q = 0; q < PER; q++)
        Pb[t * PER + q] = make_double2(o1 + v1[q], o2 + v2[q]);
    if (t == 255) Pb[PB] = make_double2(o1 + s1, o2 + s2);
    if (t == 0) refs[blockIdx.x] = c;
}

// (sum, sum of squares) of (x - c_b) over the window [i, i+m), b = i / PB, m <= PB.
__device__ __forceinline__ double2 window_sums(const double2 *__restrict__ P,
                                               const double *__restrict__ refs, int64_t i, int m,
                                               double *cb)
{
    const int64_t b = i / PB;
    const int o = (int)(i - b * PB);
    const int e = o + m;
    const double2 *Pb = P + b * (PB + 1);
    const double2 lo = Pb[o];
    *cb = refs[b];
    if (e <= PB) {
        const double2 hi = Pb[e];
        return make_double2(hi.x - lo.x, hi.y - lo.y);
    }
    const double2 tot = Pb[PB];
    const double2 nx = P[(b + 1) * (PB + 1) + (e - PB)];
    const double D = refs[b + 1] - refs[b];
    const double cnt = (double)(e - PB);
    const double a1 = tot.x - lo.x, a2 = tot.y - lo.y;
    return make_double2(a1 + nx.x + cnt * D, a2 + nx.y + 2.0 * D * nx.x + cnt * D * D);
}

__device__ __forceinline__ void window_stats(const double2 *__restrict__ P,
                                             const double *__restrict__ refs, int64_t i, int m,
                                             double *mu, double *invn)
{
    double cb;
    double2 s = window_sums(P, refs, i, m, &cb);
    const double q = s.x * s.x / m;
    double M2 = s.y - q;
    *mu = cb + s.x / m;
    // zero spread (up to rounding of the sums) = degenerate window
    *invn = (M2 > 1e-13 * (s.y + q) && M2 > 0.0) ? 1.0 / sqrt(M2) : 0.0;
}

// Per-length terms for positions i in [0, Npad): rows/columns beyond N get zeros.
__global__ __launch_bounds__(256) void k_derive(const float *__restrict__ x,
                                                const double2 *__restrict__ P,
                                                const double *__restrict__ refs, int64_t N,
                                                int64_t Npad, int m, float4 *__restrict__ colv,
                                                double2 *__restrict__ st64)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Npad) return;
    if (i >= N) {
        colv[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        st64[i] = make_double2(0.0, 0.0);
        return;
    }
    double mu, invn;
    window_stats(P, refs, i, m, &mu, &invn);
    double df = 0.0, dg = 0.0;
    if (i > 0) {
        double mup, invp;
        window_stats(P, refs, i - 1, m, &mup, &invp);
        const double xin = (double)x[i + m - 1], xout = (double)x[i - 1];
        df = 0.5 * (xin - xout);
        dg = (xin - mu) + (xout - mup);
    }
    colv[i] = make_float4((float)df, (float)dg, (float)invn, (float)mu);
    st64[i] = make_double2(mu, invn);
}

// ---------------------------------------------------------------------------------------
// Scan kernels.

struct ScanLds {
    float *xs;    // [64 + m + 4]  x - ref over the current window
    float *q;     // [m]           uniform centred vector
    float4 *tile; // [64]          (df, dg, invn, seed) of the tile's columns
};

// lane t: sum_k q[k] * (x[base+t+k] - mu_{base+t}), using x - ref staged in LDS.
__device__ __forceinline__ float corr_lanes(const ScanLds &L, const float *__restrict__ x,
                                            int64_t base, int m, float qsum, float mu_lane)
{
    const int t = threadIdx.x;
    const float ref = x[base];
    for (int k = t; k < WAVE + m; k += WAVE) L.xs[k] = x[base + k] - ref;
    __syncthreads();
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
    int k = 0;
    for (; k + 4 <= m; k += 4) {
        const float4 qv = *reinterpret_cast<const float4 *>(L.q + k);
        a0 = fmaf(qv.x, L.xs[t + k], a0);
        a1 = fmaf(qv.y, L.xs[t + k + 1], a1);
        a2 = fmaf(qv.z, L.xs[t + k + 2], a2);
        a3 = fmaf(qv.w, L.xs[t + k + 3], a3);
    }
    for (; k < m; k++) a0 = fmaf(L.q[k], L.xs[t + k], a0);
    const float acc = (a0 + a1) + (a2 + a3);
    __syncthreads();
    return acc - (mu_lane - ref) * qsum;
}

// Stage q[k] = x[p+k] - mu_p and return sum q (wave-reduced, uniform).
__device__ __forceinline__ float stage_centred(const ScanLds &L, const float *__restrict__ x,
                                               int64_t p, int m, float mu_p)
{
    const int t = threadIdx.x;
    float s = 0.f;
    for (int k = t; k < m; k += WAVE) {
        const float v = x[p + k] - mu_p;
        L.q[k] = v;
        s += v;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    __syncthreads();
    return s;
}

struct RowState {
    float cov;     // chain value of this lane (row s+t) at the last processed column
    float rdf, rdg;// row recurrence terms (0 on lane 0, which takes seeds)
    float thr;     // prune when cov*invn_j >= thr
    float mx;      // max over valid columns of cov*invn_j
    uint64_t alive;// wave-uniform alive mask of the 64 rows
};

// climit: first column not to evaluate (min of the scan end and N); N: rows at or beyond N
// are padding.
template <bool MASKED>
__device__ __forceinline__ void tile_steps(const ScanLds &L, RowState &R, int64_t s, int64_t c,
                                           int64_t N, int64_t climit, int m, uint64_t &rowhit,
                                           uint64_t &colhit)
{
    const int t = threadIdx.x;
    float cov = R.cov, mx = R.mx;
    const float thr = R.thr, rdf = R.rdf, rdg = R.rdg;
    const int dbase = (int)(c - s) - t;
    const bool rowin = (s + t) < N;
    const int colmax = (int)((climit - c) < 64 ? (climit - c) : 64);
#pragma unroll
    for (int j = 0; j < WAVE; j++) {
        const float4 cd = L.tile[j];
        const int prev = __builtin_amdgcn_update_dpp(__float_as_int(cd.w), __float_as_int(cov),
                                                     DPP_WAVE_SHR1, 0xF, 0xF, false);
        cov = fmaf(rdf, cd.y, __int_as_float(prev));
        cov = fmaf(rdg, cd.x, cov);
        const float cj = cov * cd.z;
        const bool hit = cj >= thr;
        if (!MASKED) {
            mx = fmaxf(mx, cj);
            const uint64_t h = __builtin_amdgcn_ballot_w64(hit);
            rowhit |= h;
            colhit |= (h != 0) ? (1ull << j) : 0ull;
        } else {
            const int d = dbase + j;
            const bool colin = j < colmax;
            const bool rv = colin && ((d >= m - 1) || (d <= -(m + 1)));
            const bool cv = colin && rowin && ((d >= m + 1) || (d <= -(m - 1)));
            mx = rv ? fmaxf(mx, cj) : mx;
            const uint64_t hr = __builtin_amdgcn_ballot_w64(hit && rv);
            const uint64_t hc = __builtin_amdgcn_ballot_w64(hit && cv);
            rowhit |= hr;
            colhit |= (hc != 0) ? (1ull << j) : 0ull;
        }
    }
    R.cov = cov;
    R.mx = mx;
}

// Scan columns [c0, c1) for the block of rows [s, s+64).  c0 % 64 == 0.  Returns false when
// every row of the block is known dead.
__device__ bool scan_range(const ScanLds &L, RowState &R, const float *__restrict__ x,
                           const float4 *__restrict__ colv, uint64_t *__restrict__ alive,
                           int64_t s, int64_t c0, int64_t c1, int64_t N, int m, float qrow_sum,
                           bool qrow_staged)
{
    const int t = threadIdx.x;
    // Chain initialisation: lane u holds cov(s+u, c0-1), or cov(s+u+1, 0) when c0 == 0
    // (position 0 has zero recurrence terms, so the first step reproduces it).
    {
        const int64_t cI = c0 > 0 ? c0 - 1 : 0;
        const int64_t rb = c0 > 0 ? s : s + 1;
        const float muc = colv[cI].w;
        const float qs = stage_centred(L, x, cI, m, muc);
        R.cov = corr_lanes(L, x, rb, m, qs, colv[rb + t].w);
    }
    // Seeds use the centred first row of the block.
    const float murow = colv[s].w;
    const float qsum = stage_centred(L, x, s, m, murow);
    (void)qrow_sum;
    (void)qrow_staged;
    const uint64_t word = (uint64_t)(s >> 6);
    for (int64_t c = c0; c < c1; c += WAVE) {
        const float4 cv = colv[c + t];
        const float seed = corr_lanes(L, x, c, m, qsum, cv.w);
        L.tile[t] = make_float4(cv.x, cv.y, cv.z, seed);
        __syncthreads();
        const bool masked = (c - s - 63 <= m && c - s + 63 >= -m) || (c + WAVE > N) ||
                            (s + WAVE > N) || (c + WAVE > c1);
        uint64_t rowhit = 0, colhit = 0;
        if (masked) {
            // columns at or beyond c1 are outside this scan
            const int64_t lim = c1 < N ? c1 : N;
            tile_steps<true>(L, R, s, c, N, lim, m, rowhit, colhit);
        } else {
            tile_steps<false>(L, R, s, c, N, N, m, rowhit, colhit);
        }
        __syncthreads();
        if (t == 0 && colhit)
            __hip_atomic_fetch_and(alive + (c >> 6), ~colhit, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        uint64_t cur;
        if (t == 0) {
            cur = rowhit ? __hip_atomic_fetch_and(alive + word, ~rowhit, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT)
                         : __hip_atomic_load(alive + word, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        }
        cur = __shfl(cur, 0, 64);
        R.alive &= cur & ~rowhit;
        if (R.alive == 0) return false;
    }
    return true;
}

__device__ __forceinline__ void load_rows(RowState &R, const float4 *__restrict__ colv,
                                          int64_t s, int m, float thr_rho)
{
    const int t = threadIdx.x;
    const float4 rv = colv[s + t];
    R.rdf = (t == 0) ? 0.f : rv.x;
    R.rdg = (t == 0) ? 0.f : rv.y;
    const float inv = rv.z;
    if (inv > 0.f)
        R.thr = thr_rho / inv;
    else
        R.thr = (thr_rho <= 0.f) ? -INFINITY : INFINITY;
    R.mx = -INFINITY;
    (void)m;
}

__device__ __forceinline__ void flush_best(const RowState &R, const float4 *__restrict__ colv,
                                           int *__restrict__ best, int64_t s, int64_t N)
{
    const int t = threadIdx.x;
    if (((R.alive >> t) & 1ull) && (s + t) < N && R.mx > -INFINITY) {
        const float rho = R.mx * colv[s + t].z;
        atomicMax(best + s + t, ord_key(rho));
    }
}

__device__ __forceinline__ ScanLds carve_lds(int m)
{
    extern __shared__ float4 lds4[];
    ScanLds L;
    L.tile = lds4;
    L.q = reinterpret_cast<float *>(lds4 + WAVE);
    L.xs = L.q + ((m + 3) & ~3);
    return L;
}

__global__ __launch_bounds__(64) void k_scan_select(const float *__restrict__ x,
                                                    const float4 *__restrict__ colv,
                                                    uint64_t *__restrict__ alive,
                                                    int *__restrict__ best, int64_t N, int m,
                                                    float thr_rho, int64_t WA)
{
    const int64_t s = (int64_t)blockIdx.x * WAVE;
    if (s >= N) return;
    RowState R;
    R.alive = __hip_atomic_load(alive + (s >> 6), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (R.alive == 0) return;
    ScanLds L = carve_lds(m);
    load_rows(R, colv, s, m, thr_rho);
    const int64_t c1 = (s + WA < N) ? s + WA : N;
    scan_range(L, R, x, colv, alive, s, s, c1, N, m, 0.f, false);
    flush_best(R, colv, best, s, N);
}

// Work item = (surviving block, column chunk).  The chunk [q*CH, (q+1)*CH) minus the window
// [s, s+WA) already scanned by k_scan_select.
__global__ __launch_bounds__(64) void k_scan_refine(const float *__restrict__ x,
                                                    const float4 *__restrict__ colv,
                                                    uint64_t *__restrict__ alive,
                                                    int *__restrict__ best,
                                                    const int *__restrict__ blist, int64_t N,
                                                    int m, float thr_rho, int64_t WA, int64_t CH,
                                                    int nchunks)
{
    const int item = blockIdx.x;
    const int64_t b = blist[item / nchunks];
    const int64_t q = item % nchunks;
    const int64_t s = b * WAVE;
    RowState R;
    R.alive = __hip_atomic_load(alive + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (R.alive == 0) return;
    const int64_t a0 = q * CH;
    const int64_t a1 = ((q + 1) * CH < N) ? (q + 1) * CH : N;
    const int64_t w0 = s, w1 = (s + WA < N) ? s + WA : N;
    ScanLds L = carve_lds(m);
    load_rows(R, colv, s, m, thr_rho);
    bool live = true;
    // part left of the window
    if (a0 < w0 && live) {
        const int64_t e = a1 < w0 ? a1 : w0;
        live = scan_range(L, R, x, colv, alive, s, a0, e, N, m, 0.f, false);
    }
    // part right of the window
    if (a1 > w1 && live) {
        int64_t st = a0 > w1 ? a0 : w1;
        st &= ~(int64_t)63;  // align; columns re-scanned below w1 are harmless
        live = scan_range(L, R, x, colv, alive, s, st, a1, N, m, 0.f, false);
    }
    if (live) flush_best(R, colv, best, s, N);
}

// ---------------------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_fill_alive(uint64_t *__restrict__ alive, int64_t N,
                                                    int64_t nwords)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwords) return;
    const int64_t r0 = w * 64;
    uint64_t v = 0;
    if (r0 + 64 <= N)
        v = ~0ull;
    else if (r0 < N)
        v = (~0ull) >> (64 - (N - r0));
    alive[w] = v;
}

__global__ __launch_bounds__(256) void k_fill_int(int *__restrict__ p, int64_t n, int v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_list_blocks(const uint64_t *__restrict__ alive,
                                                     int64_t nwords, int *__restrict__ list,
                                                     int *__restrict__ count)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwords) return;
    if (alive[w]) {
        const int p = atomicAdd(count, 1);
        list[p] = (int)w;
    }
}

// Surviving rows -> (index, FP32 squared distance); FLT_MAX when no valid pair was seen.
__global__ __launch_bounds__(256) void k_survivors(const uint64_t *__restrict__ alive,
                                                   const int *__restrict__ best, int64_t N, int m,
                                                   int *__restrict__ count,
                                                   int64_t *__restrict__ idx,
                                                   float *__restrict__ d2, int64_t cap)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (!((alive[i >> 6] >> (i & 63)) & 1ull)) return;
    const int k = best[i];
    float d;
    if (k == ord_key(-INFINITY))
        d = FLT_MAX;
    else
        d = 2.0f * (float)m * (1.0f - ord_val(k));
    const int p = atomicAdd(count, 1);
    if (p < cap) {
        idx[p] = i;
        d2[p] = d;
    }
}

// ---------------------------------------------------------------------------------------
// Exact (FP64) nearest-neighbour squared distance of up to MAXU candidates.  Each block
// handles 256 columns; the candidates' centred windows live in LDS.  The pair value is
// formed as sum_k (x[i+k]-mu_i)(x[j+k]-mu_j) in ascending k, times (invn_i*invn_j), which is
// bitwise symmetric in (i, j): mutual nearest neighbours tie exactly and the tie goes to
// the lower index.  outkey[u] receives the atomic minimum of ord_key64(ED^2).

__global__ __launch_bounds__(256) void k_exact(const float *__restrict__ x,
                                               const double2 *__restrict__ st64, int64_t N, int m,
                                               const int64_t *__restrict__ cand, int nU,
                                               unsigned long long *__restrict__ outkey)
{
    extern __shared__ double dsh[];
    double *xs = dsh;                   // [256 + m]   raw x as double
    double *A = dsh + (256 + m + 8);    // [nU][m]     centred candidate windows
    __shared__ double ainv[MAXU];
    __shared__ unsigned long long wmin[MAXU][4];
    const int t = threadIdx.x;
    const int64_t c = (int64_t)blockIdx.x * 256;
    for (int u = 0; u < nU; u++) {
        const int64_t i = cand[u];
        const double2 st = st64[i];
        if (t == 0) ainv[u] = st.y;
        for (int k = t; k < m; k += 256) A[u * m + k] = (double)x[i + k] - st.x;
    }
    for (int k = t; k < 256 + m; k += 256) xs[k] = (double)x[c + k];
    __syncthreads();
    const int64_t j = c + t;
    const double2 sj = (j < N) ? st64[j] : make_double2(0.0, 0.0);
    double acc[MAXU];
#pragma unroll
    for (int u = 0; u < MAXU; u++) acc[u] = 0.0;
    for (int k = 0; k < m; k++) {
        const double b = xs[t + k] - sj.x;
#pragma unroll
        for (int u = 0; u < MAXU; u++)
            if (u < nU) acc[u] = fma(A[u * m + k], b, acc[u]);
    }
    const int lane = t & 63, w = t >> 6;
#pragma unroll
    for (int u = 0; u < MAXU; u++) {
        if (u >= nU) break;
        const int64_t i = cand[u];
        double ed = DBL_MAX;
        const bool valid = (j < N) && ((j - i >= m - 1) || (i - j >= m + 1));
        if (valid) {
            if (ainv[u] > 0.0 && sj.y > 0.0)
                ed = 2.0 * m * (1.0 - acc[u] * (ainv[u] * sj.y));
            else
                ed = 2.0 * m;
        }
        unsigned long long key = ord_key64(ed);
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            const unsigned long long other = __shfl_xor(key, o, 64);
            key = other < key ? other : key;
        }
        if (lane == 0) wmin[u][w] = key;
    }
    __syncthreads();
    if (t < nU) {
        unsigned long long k = wmin[t][0];
        for (int q = 1; q < 4; q++) k = wmin[t][q] < k ? wmin[t][q] : k;
        atomicMin(outkey + t, k);
    }
}

}  // namespace paddi
