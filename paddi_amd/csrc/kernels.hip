// Hand-written gfx950 kernels of the discord engine.
//
//   k_prefix        blocked, locally centred FP64 prefix sums of every series (once per series)
//   k_derive        per-length window statistics and recurrence terms from the prefix sums
//                   (replaces gpu_compute_statistics / gpu_update_statistics,
//                    src/preprocessing.cuh:12-65)
//   k_scan_select   PD3 candidate selection: each 64-row block scans its witness tiles and a
//                   bounded window to its right (replaces gpu_local_candidate_select +
//                   gpu_local_define_candidates, src/PD3_kernels.cuh:9-274)
//   k_scan_refine   exhaustive scan of surviving blocks over the remaining columns, split in
//                   column chunks, early abandoning (replaces gpu_local_discords_refine,
//                   src/PD3_kernels.cuh:276-528, and closes its left-of-segment gap)
//   k_pass_init     per-pass state: alive bits, row maxima, witness slots, counters
//   k_list_blocks   live blocks before each refine round
//   k_survivors     compaction of surviving rows (replaces sets_processing.hpp:49-113)
//   k_block_best    per block, the row with the largest NN-distance bound (first-length
//                   candidate scoring)
//   k_exact         double-precision nearest-neighbour distance of a few candidates
//                   (replaces gpu_multigpus_discords_refine, src/PD3_kernels.cuh:665-759)
//
// Batching.  A launch covers every series of a batch.  Series s owns the global rows
// [roff_s, roff_s + Rpad_s) (roff_s a multiple of 64); its rows at or beyond N_s = n_s - m + 1
// are padding with zero terms.  bser[] maps a global 64-row block to its series (-1 for a
// padding block).  Per-row arrays (colv, st64, best, wit) and the alive bit words are indexed
// by global row; inside a kernel everything is rebased to the series' local indices.
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
// (k_scan_refine: one broadcast LDS read per step; k_scan_select: one read per 16 steps, the
// terms then taken as DPP row broadcasts) and each lane receives its diagonal predecessor
// cov(s+t-1, c+j-1) from lane t-1 through DPP wave_shr:1; lane 0 takes a directly computed
// seed cov(s, c+j).  The seeds of 256 columns at a time are one product on the matrix cores
// (super_seeds).  Row pruning, row maxima and the alive mask are lane-local; column pruning
// is a ballot.  k_exact takes the candidates of groups of three as DPP row broadcasts too.

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
// Prefix sums.  Prefix block b of series s covers x_s[b*PB, (b+1)*PB); c_b = x_s[b*PB].
// P[(pboff_s + b)*(PB+1) + k] holds (sum, sum of squares) of (x - c_b) over
// x_s[b*PB, b*PB+k), k = 0..PB (k = PB is the total).  Points at or beyond n_s count as c_b.
// refs[pboff_s + b] = (c_b, total of squares Y_b, sqrt(PB Y_b), 0): the block's reference and
// the magnitudes that bound the rounding of its sums (window_sums).

__global__ __launch_bounds__(256) void k_prefix(const float *__restrict__ x,
                                                const SerDev *__restrict__ ser,
                                                const int *__restrict__ pser,
                                                double2 *__restrict__ P, double4 *__restrict__ refs,
                                                int *__restrict__ bad)
{
    const int t = threadIdx.x;
    const int gb = blockIdx.x;
    const int sid = pser[gb];
    const SerDev S = ser[sid];
    const float *xs = x + S.xoff;
    const int64_t n = S.n;
    const int64_t base = (int64_t)(gb - S.pboff) * PB;
    const double c = (base < n) ? (double)xs[base] : 0.0;
    constexpr int PER = PB / 256;
    double v1[PER], v2[PER];
    double s1 = 0.0, s2 = 0.0;
    bool nonfinite = false;
#pragma unroll
    for (int q = 0; q < PER; q++) {
        int64_t g = base + (int64_t)t * PER + q;
        const float xv = (g < n) ? xs[g] : 0.f;
        nonfinite |= !isfinite(xv);
        double y = (g < n) ? ((double)xv - c) : 0.0;
        v1[q] = s1;
        v2[q] = s2;
        s1 += y;
        s2 += y * y;
    }
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
    double o1 = i1 - s1, o2 = i2 - s2;
    for (int q = 0; q < w; q++) { o1 += wt1[q]; o2 += wt2[q]; }
    double2 *Pb = P + (int64_t)gb * (PB + 1);
#pragma unroll
    for (int q = 0; q < PER; q++)
        Pb[t * PER + q] = make_double2(o1 + v1[q], o2 + v2[q]);
    if (t == 255) {
        const double y = o2 + s2;
        Pb[PB] = make_double2(o1 + s1, y);
        refs[gb] = make_double4(c, y, sqrt((double)PB * y), 0.0);
    }
    if (nonfinite) atomicOr(bad, 1);
}

// (sum, sum of squares) of (x - c_b) over the window [i, i+m) of one series; m <= PB.  *err
// receives a bound on the rounding error of M2 = sum of squares - sum^2 / m formed from them:
// each prefix value carries at most 64 roundings of the block's running totals (16 serial
// adds per thread, six shuffle levels, the wave offsets), every partial sum of |x - c_b| is
// at most sqrt(PB * total of squares), and the shift of a window crossing into block b+1
// adds a few roundings of its terms.
__device__ __forceinline__ double2 window_sums(const double2 *__restrict__ P,
                                               const double4 *__restrict__ refs, int64_t i, int m,
                                               double invm, double *cb, double *err)
{
    constexpr double G = 64.0 * 1.1102230246251565e-16;
    const int64_t b = i / PB;
    const int o = (int)(i - b * PB);
    const int e = o + m;
    const double2 *Pb = P + b * (PB + 1);
    const double2 lo = Pb[o];
    const double4 rb = refs[b];
    *cb = rb.x;
    double2 s;
    double Y = rb.y, X = rb.z;
    if (e <= PB) {
        const double2 hi = Pb[e];
        s = make_double2(hi.x - lo.x, hi.y - lo.y);
    } else {
        const double2 tot = Pb[PB];
        const double2 nx = P[(b + 1) * (PB + 1) + (e - PB)];
        const double4 rn = refs[b + 1];
        const double D = rn.x - rb.x;
        const double cnt = (double)(e - PB);
        const double a1 = tot.x - lo.x, a2 = tot.y - lo.y;
        s = make_double2(a1 + nx.x + cnt * D, a2 + nx.y + 2.0 * D * nx.x + cnt * D * D);
        Y += rn.y + fabs(2.0 * D * nx.x) + cnt * D * D;
        X += rn.z + fabs(cnt * D);
    }
    *err = G * (Y + 2.0 * fabs(s.x) * X * invm) + 4.0 * 1.1102230246251565e-16 * (fabs(s.y));
    return s;
}

// Statistics of window [i, i+m): mean, 1/sqrt(M2) (0 for a window of zero spread) and M2.
// From the prefix sums where their rounding is below 1e-8 of M2; otherwise -- near-flat
// windows far from their block's reference point, where the sums cancel -- from the window's
// values in two passes, centred on its first value (differences of two floats are exact in
// double precision).
__device__ __forceinline__ void window_stats(const double2 *__restrict__ P,
                                             const double4 *__restrict__ refs,
                                             const float *__restrict__ xs, int64_t i, int m,
                                             double invm, double *mu, double *invn,
                                             double *m2 = nullptr)
{
    double cb, err;
    const double2 s = window_sums(P, refs, i, m, invm, &cb, &err);
    const double sm = s.x * invm;
    double M2 = s.y - s.x * sm;
    double mean = cb + sm;
    if (!(M2 >= 1e8 * err)) {
        const double x0 = (double)xs[i];
        double d = 0.0;
        for (int k = 0; k < m; k++) d += (double)xs[i + k] - x0;
        d *= invm;
        double v = 0.0;
        for (int k = 0; k < m; k++) {
            const double y = ((double)xs[i + k] - x0) - d;
            v = fma(y, y, v);
        }
        M2 = v;
        mean = x0 + d;
    }
    *mu = mean;
    *invn = M2 > 0.0 ? 1.0 / sqrt(M2) : 0.0;
    if (m2) *m2 = M2 > 0.0 ? M2 : 0.0;
}

// DPP row operation of the wave: lanes of the rows in row_mask read src through ctrl, the
// others (and lanes without a source) keep their own value.
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, ROWS, 0xF, false));
}

// Maximum over the 64 lanes (every lane active): pairs, quads, half rows, rows by quad_perm
// and mirror moves, then rows 1 and 3 take the last lane of rows 0 and 2 (row_bcast:15) and
// rows 2 and 3 that of row 1 (row_bcast:31), which leaves the total in lane 63.
__device__ __forceinline__ float wave_max(float v)
{
    v = fmaxf(v, dpp_f<0xB1, 0xF>(v));   // quad_perm [1,0,3,2]
    v = fmaxf(v, dpp_f<0x4E, 0xF>(v));   // quad_perm [2,3,0,1]
    v = fmaxf(v, dpp_f<0x141, 0xF>(v));  // row_half_mirror
    v = fmaxf(v, dpp_f<0x140, 0xF>(v));  // row_mirror
    v = fmaxf(v, dpp_f<0x142, 0xA>(v));  // row_bcast:15
    v = fmaxf(v, dpp_f<0x143, 0xC>(v));  // row_bcast:31
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// DPP row operation with zero fill: lanes of the rows in ROWS and banks (groups of 4 lanes
// of a row) in BANKS read src through CTRL, all others and lanes without a source get 0.
template <int CTRL, int ROWS, int BANKS>
__device__ __forceinline__ float dpp_z(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, BANKS, true));
}

// Inclusive prefix maximum over lanes 0..t of non-negative values (every lane active; 0 is
// the identity): within each row of 16 lanes the three preceding lanes (row_shr:1..3), then
// 4 back for banks 1-3 and 8 back for banks 2-3, which completes the row; then rows 1 and 3
// take the last lane of rows 0 and 2 (row_bcast:15) and rows 2 and 3 that of row 1
// (row_bcast:31).
__device__ __forceinline__ float wave_prefix_max(float x)
{
    float v = fmaxf(fmaxf(x, dpp_z<0x111, 0xF, 0xF>(x)),
                    fmaxf(dpp_z<0x112, 0xF, 0xF>(x), dpp_z<0x113, 0xF, 0xF>(x)));
    v = fmaxf(v, dpp_z<0x114, 0xF, 0xE>(v));
    v = fmaxf(v, dpp_z<0x118, 0xF, 0xC>(v));
    v = fmaxf(v, dpp_z<0x142, 0xA, 0xF>(v));
    v = fmaxf(v, dpp_z<0x143, 0xC, 0xF>(v));
    return v;
}

// Per-length terms for every global row (padding rows get zeros), and per 64-row block
// (one wave; rows is a multiple of 64) the largest sqrt(M2) and invn of its rows; with
// m2 > 0 also (mu, invn) at length m2 into st64n (k_exact's second length).  st64 = null
// skips (mu, invn) at m (already there from the previous length's st64n).  A thread takes
// DERIVE_RPT rows 256 apart (independent loads in flight); the recurrence term dg needs the
// previous window's mean, mu_{i-1} = mu_i + (x[i-1] - x[i+m-1]) / m, whose double-precision
// rounding vanishes in the single-precision dg.
__global__ __launch_bounds__(256) void k_derive(const float *__restrict__ x,
                                                const SerDev *__restrict__ ser,
                                                const int *__restrict__ bser,
                                                const double2 *__restrict__ P,
                                                const double4 *__restrict__ refs, int64_t rows,
                                                int m, float4 *__restrict__ colv,
                                                double2 *__restrict__ st64,
                                                float2 *__restrict__ tmax, int m2,
                                                double2 *__restrict__ st64n,
                                                const int *__restrict__ dev_m)
{
    // the device length loop gives the lengths on the device (m = 0: the loop has stopped)
    if (dev_m) {
        m = dev_m[0];
        m2 = dev_m[1];
        if (m <= 0) return;
    }
    const double invm = 1.0 / m, invm2 = m2 ? 1.0 / m2 : 0.0;
#pragma unroll
    for (int q = 0; q < DERIVE_RPT; q++) {
        const int64_t g = ((int64_t)blockIdx.x * DERIVE_RPT + q) * blockDim.x + threadIdx.x;
        float4 cv = make_float4(0.f, 0.f, 0.f, 0.f);
        double2 sv = make_double2(0.0, 0.0), sv2 = make_double2(0.0, 0.0);
        float sd = 0.f;
        const int sid = g < rows ? bser[g >> 6] : -1;
        if (sid >= 0) {
            const SerDev S = ser[sid];
            const int64_t i = g - S.roff;
            if (i < S.n - m + 1) {
                const float *xs = x + S.xoff;
                const double2 *Ps = P + S.pboff * (PB + 1);
                const double4 *rs = refs + S.pboff;
                double mu, invn, M2;
                window_stats(Ps, rs, xs, i, m, invm, &mu, &invn, &M2);
                double df = 0.0, dg = 0.0;
                if (i > 0) {
                    const double xin = (double)xs[i + m - 1], xout = (double)xs[i - 1];
                    const double mup = mu + (xout - xin) * invm;
                    df = 0.5 * (xin - xout);
                    dg = (xin - mu) + (xout - mup);
                }
                cv = make_float4((float)df, (float)dg, (float)invn, (float)mu);
                sv = make_double2(mu, invn);
                sd = sqrtf((float)M2);
                if (m2 && i < S.n - m2 + 1) {
                    double mu2, inv2;
                    window_stats(Ps, rs, xs, i, m2, invm2, &mu2, &inv2);
                    sv2 = make_double2(mu2, inv2);
                }
            }
        }
        if (g < rows) {
            colv[g] = cv;
            if (st64) st64[g] = sv;
            if (m2) st64n[g] = sv2;
        }
        const float a = wave_max(sd), b = wave_max(cv.z);
        if ((threadIdx.x & 63) == 0 && g < rows) tmax[g >> 6] = make_float2(a, b);
    }
}


// ---------------------------------------------------------------------------------------
// Scan kernels (all indices below are local to one series).

// LDS of a scan workgroup (sizes in floats; scan_ring_w / scan_k4 / scan_lds_bytes).
struct ScanLds {
    float4 *tile; // [2][64]   (df, dg, invn, seed) of a tile's columns, double-buffered
    float *seed;  // [256]     seeds of the super-tile's columns before the mean correction
    float *qz;    // [K4 + 20] the block's first row, centred, at qz[16 + k]; zeros around it
    float *ring;  // raw x of the super-tile window: column col at (col mod W) and
                  // (col mod W) + W, float f of that range at f + f / 16 (bank padding)
    float *xs;    // [64 + m + 4] chain-initialisation scratch (aliases ring)
    float *qi;    // [m]          chain-initialisation scratch (aliases ring)
    int W, K4;
};

// lane t: sum_k v[k] * (x[base+t+k] - mu_{base+t}), using x - ref staged in LDS.
__device__ __forceinline__ float corr_lanes(const ScanLds &L, const float *v,
                                            const float *__restrict__ x, int64_t base, int m,
                                            float vsum, float mu_lane)
{
    const int t = threadIdx.x;
    const float ref = x[base];
    for (int k = t; k < WAVE + m; k += WAVE) L.xs[k] = x[base + k] - ref;
    __syncthreads();
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
    int k = 0;
#pragma unroll 2
    for (; k + 4 <= m; k += 4) {
        const float4 qv = *reinterpret_cast<const float4 *>(v + k);
        a0 = fmaf(qv.x, L.xs[t + k], a0);
        a1 = fmaf(qv.y, L.xs[t + k + 1], a1);
        a2 = fmaf(qv.z, L.xs[t + k + 2], a2);
        a3 = fmaf(qv.w, L.xs[t + k + 3], a3);
    }
    for (; k < m; k++) a0 = fmaf(v[k], L.xs[t + k], a0);
    const float acc = (a0 + a1) + (a2 + a3);
    __syncthreads();
    return acc - (mu_lane - ref) * vsum;
}

// Sum over the wave, uniform: quad butterflies and the row mirrors give every lane its row's
// sum, row_bcast:15 and :31 carry the rows into lane 63.  DPP needs no lane-address registers
// (a shuffle's do, and the scan kernels spilled them).
__device__ __forceinline__ float wave_sum(float v)
{
    v += dpp_z<0xB1, 0xF, 0xF>(v);   // quad_perm [1,0,3,2]
    v += dpp_z<0x4E, 0xF, 0xF>(v);   // quad_perm [2,3,0,1]
    v += dpp_z<0x141, 0xF, 0xF>(v);  // row_half_mirror
    v += dpp_z<0x140, 0xF, 0xF>(v);  // row_mirror
    v += dpp_z<0x142, 0xA, 0xF>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_z<0x143, 0xC, 0xF>(v);  // row_bcast:31 into rows 2 and 3
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Stage dst[k] = x[p+k] - mu_p and return sum dst (wave-reduced, uniform).
__device__ __forceinline__ float stage_centred(float *dst, const float *__restrict__ x,
                                               int64_t p, int m, float mu_p)
{
    const int t = threadIdx.x;
    float s = 0.f;
    for (int k = t; k < m; k += WAVE) {
        const float v = x[p + k] - mu_p;
        dst[k] = v;
        s += v;
    }
    s = wave_sum(s);
    __syncthreads();
    return s;
}

struct RowState {
    float cov;     // chain value of this lane (row s+t) at the last processed column
    float rdf, rdg;// row recurrence terms (0 on lane 0, which takes seeds)
    float thr0;    // (1 - r2/(2m)) / invn_i
    float thr;     // the current super-tile's threshold: prune when cov*invn_j >= thr
    float kb;      // dk * (largest sqrt(M2) of rows s..s+t, the rows this lane's chains cross)
    float mx;      // max over valid columns of cov*invn_j
    float mt;      // the same over the current super-tile
    float mxl, mxh;// max over super-tiles of (their max of cov*invn_j -+ their bound)
    uint64_t alive;// wave-uniform alive mask of the 64 rows
    uint32_t tiles;// 64x64 tiles evaluated (pair counter)
    uint32_t live; // rows alive at the start of those tiles, summed
    uint32_t dtiles; // 64-column tiles of a single live row evaluated directly (refine_item)
    float wbest;   // largest cov*invn_j over the tiles evaluated so far, and its tile
    int wtile;
    uint64_t mine; // rows this workgroup pruned (their witness: wtile, at the end)
    float qsum;    // sum of the staged, centred first row (L.qz)
};

// One series' view of the batch.
struct View {
    const float *x;
    const float4 *colv;
    const float *rho;
    uint64_t *alive;
    int *best, *besth, *bestl;
    int *wit;
    int64_t N;
    int m;
    int ml;       // row i and column j pair iff j - i >= m - 1 or i - j >= ml (ScanScal::ml)
    float dk;
};

// climit: first column not to evaluate (min of the scan end and N); rows at or beyond N
// are padding.  COLS: also collect the columns pruned by this block (colhit), which lets
// k_scan_select retire other blocks' rows early; k_scan_refine skips that bookkeeping (the
// far columns it scans belong to blocks that are almost all dead by then).
template <bool MASKED, bool COLS>
__device__ __forceinline__ void tile_steps(const float4 *tile, RowState &R, int64_t s, int64_t c,
                                           int64_t N, int64_t climit, int m, int ml,
                                           uint64_t &rowhit, uint64_t &colhit)
{
    const int t = threadIdx.x;
    float cov = R.cov, mx = R.mt;
    const float thr = R.thr, rdf = R.rdf, rdg = R.rdg;
    const int dbase = (int)(c - s) - t;
    const bool rowin = (s + t) < N;
    const int colmax = (int)((climit - c) < 64 ? (climit - c) : 64);
#pragma unroll
    for (int j = 0; j < WAVE; j++) {
        const float4 cd = tile[j];
        const int prev = __builtin_amdgcn_update_dpp(__float_as_int(cd.w), __float_as_int(cov),
                                                     DPP_WAVE_SHR1, 0xF, 0xF, false);
        cov = fmaf(rdf, cd.y, __int_as_float(prev));
        cov = fmaf(rdg, cd.x, cov);
        const float cj = cov * cd.z;
        const bool hit = cj >= thr;
        if (!MASKED) {
            mx = fmaxf(mx, cj);
            if (COLS) {
                const uint64_t h = __builtin_amdgcn_ballot_w64(hit);
                rowhit |= h;
                colhit |= (h != 0) ? (1ull << j) : 0ull;
            }
        } else {
            const int d = dbase + j;
            const bool colin = j < colmax;
            const bool rv = colin && ((d >= m - 1) || (d <= -ml));
            mx = rv ? fmaxf(mx, cj) : mx;
            const uint64_t hr = __builtin_amdgcn_ballot_w64(hit && rv);
            rowhit |= hr;
            if (COLS) {
                const bool cv = colin && rowin && ((d >= ml) || (d <= -(m - 1)));
                const uint64_t hc = __builtin_amdgcn_ballot_w64(hit && cv);
                colhit |= (hc != 0) ? (1ull << j) : 0ull;
            }
        }
    }
    // Without column pruning the steps need not test the rows: thr holds over the super-tile
    // and a live row's maximum over it is below thr when the tile starts (a hit retires the
    // row), so the row is hit in this tile iff that maximum has reached thr by its end.
    if (!MASKED && !COLS) rowhit |= __builtin_amdgcn_ballot_w64(mx >= thr);
    R.cov = cov;
    R.mt = mx;
}

// The unmasked tile of k_scan_select with the column terms as DPP operands.  tile_steps
// reads the 16 bytes of column j from LDS in every step, the same address for all 64 lanes.
// Here lane l reads column 16g + (l mod 16) once per 16 steps, and step j takes column j's
// terms from lane j mod 16 of each row of 16 lanes as the first operand of an FMA or multiply
// (DPP row_newbcast: no instruction of its own).  The step's increment
// a = dg_i df_j + df_i dg_j (+ the seed on lane 0, whose row terms are zero: e0 = 1 there,
// 0 elsewhere) does not depend on the chain; the chain itself is one add per step,
// cov_i = cov_{i-1} (shifted in by DPP wave_shr, zero on lane 0) + a.  The increment is summed
// in another order than tile_steps' two FMAs, a rounding difference of the order of one unit
// in the last place per step.  This costs one VALU instruction per step more than tile_steps
// and saves 60 of its 64 LDS reads per tile: k_scan_select is faster with it, k_scan_refine
// (whose loop is bound by VALU issue) is not and keeps tile_steps.
// Written as VALU assembly because the compiler does not fold DPP broadcasts into FMAs.
// Hazards: a DPP operand needs two wait states after a VALU write of it.  A block starts with
// s_nop 1 (its operands may have been written just before it); within a block the chain
// value read by the next step's shifted add was written at least three instructions earlier.
#define PD_CTL(B) " row_newbcast:" #B " row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
#define PD_CHAIN(A, B)                                                                         \
    "v_mul_f32_dpp " A ", %[cx], %[rdg]" PD_CTL(B)                                               \
    "v_fmac_f32_dpp " A ", %[cy], %[rdf]" PD_CTL(B)                                              \
    "v_fmac_f32_dpp " A ", %[cw], %[e0]" PD_CTL(B)                                               \
    "v_add_f32_dpp %[cov], %[cov], " A " wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
#define PD_STEP(A, J, H, B)                                                                    \
    PD_CHAIN(A, B)                                                                               \
    "v_mul_f32_dpp " J ", %[cz], %[cov]" PD_CTL(B)                                               \
    "v_cmp_le_f32_e64 " H ", %[thr], " J "\n"
// Four steps (columns b0..b0+3 of the current group of 16): chain value and row maximum in
// place, and the four hit masks (rows of the step's column that are within range).
#define PD_QUAD_BODY(B0, B1, B2, B3)                                                           \
    "s_nop 1\n"                                                                                  \
    PD_STEP("%[a0]", "%[j0]", "%[h0]", B0)                                                       \
    PD_STEP("%[a1]", "%[j1]", "%[h1]", B1)                                                       \
    "v_max3_f32 %[mx], %[mx], %[j0], %[j1]\n"                                                    \
    PD_STEP("%[a0]", "%[j0]", "%[h2]", B2)                                                       \
    PD_STEP("%[a1]", "%[j1]", "%[h3]", B3)                                                       \
    "v_max3_f32 %[mx], %[mx], %[j0], %[j1]\n"
#define PD_QUAD_IN                                                                             \
    [cw] "v"(cd.w), [cx] "v"(cd.x), [cy] "v"(cd.y), [cz] "v"(cd.z), [e0] "v"(e0),              \
        [rdf] "v"(rdf), [rdg] "v"(rdg), [thr] "v"(thr)
#define PD_QUAD_COLS(B0, B1, B2, B3)                                                           \
    asm(PD_QUAD_BODY(B0, B1, B2, B3)                                                            \
        : [cov] "+v"(cov), [mx] "+v"(mx), [a0] "=&v"(a0), [a1] "=&v"(a1), [j0] "=&v"(j0),      \
          [j1] "=&v"(j1), [h0] "=&s"(h[0]), [h1] "=&s"(h[1]), [h2] "=&s"(h[2]),                \
          [h3] "=&s"(h[3])                                                                      \
        : PD_QUAD_IN)
// Four warm-up steps: the chain only (the next step's increment separates the adds).
#define PD_WARM4(B0, B1, B2, B3)                                                               \
    asm("s_nop 1\n" PD_CHAIN("%[a0]", B0) PD_CHAIN("%[a1]", B1) PD_CHAIN("%[a0]", B2)          \
            PD_CHAIN("%[a1]", B3)                                                               \
        : [cov] "+v"(cov), [a0] "=&v"(a0), [a1] "=&v"(a1)                                       \
        : [cw] "v"(cd.w), [cx] "v"(cd.x), [cy] "v"(cd.y), [e0] "v"(e0), [rdf] "v"(rdf),        \
          [rdg] "v"(rdg))

__device__ __forceinline__ void tile_steps_bc(const float4 *tile, RowState &R, uint64_t &rowhit,
                                              uint64_t &colhit)
{
    const int t = threadIdx.x;
    float cov = R.cov, mx = R.mt;
    const float thr = R.thr, rdf = R.rdf, rdg = R.rdg;
    const float e0 = t == 0 ? 1.f : 0.f;
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const float4 cd = tile[16 * g + (t & 15)];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            float a0, a1, j0, j1;
            uint64_t h[4];
            if (q == 0) PD_QUAD_COLS(0, 1, 2, 3);
            if (q == 1) PD_QUAD_COLS(4, 5, 6, 7);
            if (q == 2) PD_QUAD_COLS(8, 9, 10, 11);
            if (q == 3) PD_QUAD_COLS(12, 13, 14, 15);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                rowhit |= h[i];
                colhit |= (h[i] != 0) ? (1ull << (16 * g + 4 * q + i)) : 0ull;
            }
        }
    }
    R.cov = cov;
    R.mt = mx;
}

__device__ __forceinline__ void warm_steps_bc(const float4 *tile, RowState &R)
{
    const int t = threadIdx.x;
    float cov = R.cov;
    const float rdf = R.rdf, rdg = R.rdg;
    const float e0 = t == 0 ? 1.f : 0.f;
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const float4 cd = tile[16 * g + (t & 15)];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            float a0, a1;
            if (q == 0) PD_WARM4(0, 1, 2, 3);
            if (q == 1) PD_WARM4(4, 5, 6, 7);
            if (q == 2) PD_WARM4(8, 9, 10, 11);
            if (q == 3) PD_WARM4(12, 13, 14, 15);
        }
    }
    R.cov = cov;
}

__device__ __forceinline__ int ring_at(int f) { return f + (f >> 4); }

// Seeds of the 256 columns c .. c+255 of a super-tile on the matrix cores, as one
// 16 x 16 x K4 product (v_mfma_f32_16x16x4f32): row a of A is the window of column c + 16a,
// A[a][k'] = x[c+16a+k'] - x[c+16a], and B[k'][j] = q[k'-j] (zero outside 0..m-1), so
// D[a][j] = sum_k q[k] (x[c+16a+j+k] - x[c+16a]).  Lane l supplies A[l%16][l/16 + 4i] and
// B[l/16 + 4i][l%16] and receives D[4(l/16)+r][l%16] in register r.  Column n = 16a + j then
// has seed D[a][j] - (mu_{c+n} - x[c+16a]) * sum(q), applied per tile.  r0 = c mod W.
__device__ __forceinline__ void super_seeds(const ScanLds &L, int r0, const float *qz)
{
    typedef float f4 __attribute__((ext_vector_type(4)));
    const int l = threadIdx.x, i = l & 15, kk = l >> 4;
    const float *ap = L.ring + ring_at(r0 + 16 * i) + kk;
    const float *bp = qz + 16 + kk - i;
    const float refi = L.ring[ring_at(r0 + 16 * i)];
    f4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
    int k0 = 0;
    for (; k0 + 16 <= L.K4; k0 += 16, ap += 17, bp += 16) {
        acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(ap[0] - refi, bp[0], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ap[4] - refi, bp[4], acc1, 0, 0, 0);
        acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(ap[8] - refi, bp[8], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ap[12] - refi, bp[12], acc1, 0, 0, 0);
    }
    for (int e = 0; k0 < L.K4; k0 += 4, e += 4)
        acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(ap[e] - refi, bp[e], acc0, 0, 0, 0);
    float *out = L.seed + 64 * kk + i;
#pragma unroll
    for (int r = 0; r < 4; r++) out[16 * r] = acc0[r] + acc1[r];
}

// The recurrence alone over a tile (no pruning, no maxima): brings every lane's chain to
// the tile's last column.
__device__ __forceinline__ void warm_steps(const float4 *tile, RowState &R)
{
    float cov = R.cov;
    const float rdf = R.rdf, rdg = R.rdg;
#pragma unroll
    for (int j = 0; j < WAVE; j++) {
        const float4 cd = tile[j];
        const int prev = __builtin_amdgcn_update_dpp(__float_as_int(cd.w), __float_as_int(cov),
                                                     DPP_WAVE_SHR1, 0xF, 0xF, false);
        cov = fmaf(rdf, cd.y, __int_as_float(prev));
        cov = fmaf(rdg, cd.x, cov);
    }
    R.cov = cov;
}

// Ring columns [c, c + cnt) from x (both copies), lanes striding by 64.
__device__ __forceinline__ void ring_fill(const ScanLds &L, const float *__restrict__ x,
                                          int64_t c, int cnt)
{
    const int r0 = (int)(c % L.W);
    for (int p = threadIdx.x; p < cnt; p += WAVE) {
        int pos = r0 + p;
        pos = pos >= L.W ? pos - L.W : pos;
        const float v = x[c + p];
        L.ring[ring_at(pos)] = v;
        L.ring[ring_at(pos + L.W)] = v;
    }
}

// Scan columns [c0, c1) for the block of rows [s, s+64).  c0 % 64 == 0.  Returns false when
// every row of the block is known dead.  With V.wit set, rows pruned by this block record
// the pruning tile and columns pruned by it record the block.
//
// Columns go by super-tiles of 256 (seeds on the matrix cores, then four tiles of 64 steps on
// the vector units).  The series window of the next super-tile (its 256 new columns, four
// floats per lane) and the next tile's column terms are loaded while the current ones are
// processed, and the block's alive word read back by a tile's atomic is consumed one tile
// later, so no global-memory round trip sits on the critical path.
template <bool COLS>
__device__ bool scan_range(const ScanLds &L, RowState &R, const View &V, int64_t s, int64_t c0,
                           int64_t c1)
{
    const int t = threadIdx.x;
    const int m = V.m;
    const int64_t N = V.N;
    const int W = L.W, K4 = L.K4;
    // Chain start.  From c0 >= 64 a warm-up tile over columns [c0-64, c0) runs the
    // recurrence alone: lane 0 takes the seeds, and after its 64 steps lane u holds
    // cov(s+u, c0-1), derived from the seed of column c0-1-u (lanes u > j still hold
    // placeholders at step j, never read by a lane that keeps its value).  From c0 == 0,
    // lane u holds cov(s+u+1, 0) computed directly (position 0 has zero recurrence terms,
    // so the first step reproduces it).
    const bool warm = c0 >= WAVE;
    const int64_t cs = warm ? c0 - WAVE : c0;
    R.cov = 0.f;
    if (!warm) {
        const float qs = stage_centred(L.qi, V.x, 0, m, V.colv[0].w);
        R.cov = corr_lanes(L, L.qi, V.x, s + 1, m, qs, V.colv[s + 1 + t].w);
    }
    const float qsum = R.qsum;
    const uint64_t word = (uint64_t)(s >> 6);
    // the first window, then the next super-tile's new columns into registers
    ring_fill(L, V.x, cs, 256 + K4);
    float nx[4];
    if (cs + 256 < c1) {
#pragma unroll
        for (int p = 0; p < 4; p++) nx[p] = V.x[cs + 256 + K4 + p * WAVE + t];
    }
    float4 pc = V.colv[cs + t];
    uint64_t pend_lane0 = ~0ull;  // lane 0: the alive word returned by the previous tile
    int par = 0;                  // tile buffer in use
    for (int64_t c = cs; c < c1; c += 256) {
        const int r0 = (int)(c % W);
        __syncthreads();  // ring writes of the previous super-tile are visible
        super_seeds(L, r0, L.qz);
        // The rows' thresholds raised by the super-tile's error bound on cov*invn_j: the
        // lane's row scale times the largest column factor of its tiles (tiles start at
        // multiples of 64, so tile T is block (c + 64 T) / 64's columns; positive floats
        // order as their bits).
        uint32_t rb = 0;
        for (int T = 0; T < 4 && c + 64 * T < c1; T++) {
            const uint32_t b = __float_as_uint(V.rho[(c >> 6) + T]);
            rb = rb > b ? rb : b;
        }
        const float inc = R.kb * __uint_as_float(rb);
        R.thr = R.thr0 + inc;
        R.mt = -INFINITY;
        __syncthreads();
        for (int T = 0; T < 4; T++) {
            const int64_t ct = c + 64 * T;
            if (ct >= c1) break;
            const float4 cv = pc;
            if (ct + WAVE < c1) pc = V.colv[ct + WAVE + t];
            const float refa = L.ring[ring_at(r0 + 64 * T + 16 * (t >> 4))];
            const float seed = L.seed[64 * T + t] - (cv.w - refa) * qsum;
            float4 *tile = L.tile + WAVE * par;
            par ^= 1;
            tile[t] = make_float4(cv.x, cv.y, cv.z, seed);
            __syncthreads();
            if (ct < c0) {  // the warm-up tile
                if (COLS)
                    warm_steps_bc(tile, R);
                else
                    warm_steps(tile, R);
                continue;
            }
            R.tiles++;
            R.live += (uint32_t)__builtin_popcountll(R.alive);
            const bool masked = (ct - s - 63 <= m && ct - s + 63 >= -m) || (ct + WAVE > N) ||
                                (s + WAVE > N) || (ct + WAVE > c1);
            uint64_t rowhit = 0, colhit = 0;
            if (masked) {
                const int64_t lim = c1 < N ? c1 : N;
                tile_steps<true, COLS>(tile, R, s, ct, N, lim, m, V.ml, rowhit, colhit);
            } else if (COLS) {
                tile_steps_bc(tile, R, rowhit, colhit);
            } else {
                tile_steps<false, COLS>(tile, R, s, ct, N, N, m, V.ml, rowhit, colhit);
            }
            if (R.mt > R.wbest) {
                R.wbest = R.mt;
                R.wtile = (int)ct;
            }
            R.mine |= rowhit & R.alive;
            if (V.wit) {
                if (((rowhit & R.alive) >> t) & 1ull) V.wit[s + t] = (int)ct;
                if (COLS && ((colhit >> t) & 1ull)) V.wit[ct + t] = (int)s;
            }
            if (COLS && t == 0 && colhit)
                __hip_atomic_fetch_and(V.alive + (ct >> 6), ~colhit, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            // Rows pruned elsewhere, as read back one tile ago, and rows pruned here now.
            const uint64_t pend =
                ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(pend_lane0 >> 32), 0) << 32) |
                (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)pend_lane0, 0);
            const uint64_t newdead = rowhit & R.alive;
            R.alive &= pend & ~rowhit;
            if (t == 0) {
                pend_lane0 = newdead ? __hip_atomic_fetch_and(V.alive + word, ~newdead,
                                                              __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT)
                                     : __hip_atomic_load(V.alive + word, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
            }
            if (R.alive == 0) return false;
        }
        R.mx = fmaxf(R.mx, R.mt);
        R.mxl = fmaxf(R.mxl, R.mt - inc);
        R.mxh = fmaxf(R.mxh, R.mt + inc);
        // The next window: its new columns [c + 256 + K4, c + 512 + K4) replace columns
        // before c + 256 (W >= 256 + K4), then the one after is requested.
        if (c + 256 < c1) {
            int rn = (int)((c + 256 + K4) % W);
#pragma unroll
            for (int p = 0; p < 4; p++) {
                int pos = rn + p * WAVE + t;
                pos = pos >= W ? pos - W : pos;
                L.ring[ring_at(pos)] = nx[p];
                L.ring[ring_at(pos + W)] = nx[p];
            }
            if (c + 512 < c1) {
#pragma unroll
                for (int p = 0; p < 4; p++) nx[p] = V.x[c + 512 + K4 + p * WAVE + t];
            }
        }
    }
    return true;
}

// Columns [c0, c1) for a block with up to DIRECT_ROWS live rows (lanes D.tr[0..nr)), without
// the recurrence: per super-tile a live row's covariances with the 256 columns are the seed
// product itself (super_seeds with the row's centred window: L.qz for the first, the tile
// buffer, which this path does not use, for the others), and lane t takes column
// c + 64 T + t of tile T.  The values are what a chain start holds, so the pruning threshold and the bounds
// are the recurrence's (kb of the row's lane covers its scale).  No lane evaluates a dead
// row: per row and 256 columns one product of K4 / 4 matrix-core steps and about 40 vector
// instructions, against some 1200 for the recurrence's four tiles.  Returns false once every
// row is pruned (here or by another work item).
struct DirectRows {
    int nr;
    int tr[DIRECT_ROWS];
    float qsum[DIRECT_ROWS];
    const float *qz[DIRECT_ROWS];
};

__device__ bool scan_range_direct(const ScanLds &L, RowState &R, const View &V, int64_t s,
                                  const DirectRows &D, int64_t c0, int64_t c1)
{
    const int t = threadIdx.x;
    const int m = V.m;
    const int64_t N = V.N;
    const int W = L.W, K4 = L.K4;
    const uint64_t word = (uint64_t)(s >> 6);
    const int64_t lim = c1 < N ? c1 : N;
    float thr0[DIRECT_ROWS], kb[DIRECT_ROWS], mx[DIRECT_ROWS], mxl[DIRECT_ROWS],
        mxh[DIRECT_ROWS], wbest[DIRECT_ROWS];
    int wtile[DIRECT_ROWS];
    bool on[DIRECT_ROWS];
    uint64_t alive = 0;
#pragma unroll
    for (int u = 0; u < DIRECT_ROWS; u++) {
        const int tu = u < D.nr ? D.tr[u] : D.tr[0];
        thr0[u] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(R.thr0), tu));
        kb[u] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(R.kb), tu));
        mx[u] = mxl[u] = mxh[u] = wbest[u] = -INFINITY;
        wtile[u] = -1;
        on[u] = u < D.nr;
        if (on[u]) alive |= 1ull << tu;
    }
    ring_fill(L, V.x, c0, 256 + K4);
    float nx[4];
    if (c0 + 256 < c1) {
#pragma unroll
        for (int p = 0; p < 4; p++) nx[p] = V.x[c0 + 256 + K4 + p * WAVE + t];
    }
    float4 cvs[4];
#pragma unroll
    for (int T = 0; T < 4; T++) cvs[T] = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int64_t c = c0; c < c1 && alive; c += 256) {
        const int r0 = (int)(c % W);
        uint32_t rb = 0;
        for (int T = 0; T < 4 && c + 64 * T < c1; T++) {
            const uint32_t b = __float_as_uint(V.rho[(c >> 6) + T]);
            rb = rb > b ? rb : b;
            cvs[T] = V.colv[c + 64 * T + t];
        }
        uint64_t newdead = 0;
#pragma unroll
        for (int u = 0; u < DIRECT_ROWS; u++) {
            if (!on[u]) continue;
            __syncthreads();  // ring writes and the previous product's readers are done
            super_seeds(L, r0, D.qz[u]);
            __syncthreads();  // seeds visible
            const int64_t i = s + D.tr[u];
            const float inc = kb[u] * __uint_as_float(rb);
            const float thr = thr0[u] + inc;
            float lmax = -INFINITY;
            int ltile = -1;
            bool hit = false;
#pragma unroll
            for (int T = 0; T < 4; T++) {
                const int64_t ct = c + 64 * T;
                const float refa = L.ring[ring_at(r0 + 64 * T + 16 * (t >> 4))];
                const float seed = L.seed[64 * T + t] - (cvs[T].w - refa) * D.qsum[u];
                const int64_t j = ct + t;
                const bool valid =
                    ct < c1 && j < lim && ((j - i >= m - 1) || (i - j >= V.ml));
                const float cj = valid ? seed * cvs[T].z : -INFINITY;
                hit |= cj >= thr;
                if (cj > lmax) {
                    lmax = cj;
                    ltile = (int)ct;
                }
                R.dtiles += ct < c1;
            }
            const float mt = wave_max(lmax);
            mx[u] = fmaxf(mx[u], mt);
            mxl[u] = fmaxf(mxl[u], mt - inc);
            mxh[u] = fmaxf(mxh[u], mt + inc);
            if (mt > wbest[u]) {
                wbest[u] = mt;
                const uint64_t at = __builtin_amdgcn_ballot_w64(lmax == mt);
                wtile[u] = __builtin_amdgcn_readlane(ltile, at ? __builtin_ctzll(at) : 0);
            }
            if (__builtin_amdgcn_ballot_w64(hit)) newdead |= 1ull << D.tr[u];
        }
        // rows pruned here, then those pruned by other work items
        int lo = 0, hi = 0;
        if (t == 0) {
            const uint64_t a =
                newdead ? __hip_atomic_fetch_and(V.alive + word, ~newdead, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)
                        : __hip_atomic_load(V.alive + word, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
            lo = (int)(uint32_t)a;
            hi = (int)(a >> 32);
        }
        const uint64_t now = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(hi, 0) << 32) |
                             (uint32_t)__builtin_amdgcn_readlane(lo, 0);
        R.mine |= newdead & alive;
        alive &= now & ~newdead;
#pragma unroll
        for (int u = 0; u < DIRECT_ROWS; u++) on[u] = on[u] && ((alive >> D.tr[u]) & 1ull);
        // the next window's new columns, as in scan_range
        if (alive && c + 256 < c1) {
            __syncthreads();  // the last product's readers of the ring are done
            int rn = (int)((c + 256 + K4) % W);
#pragma unroll
            for (int p = 0; p < 4; p++) {
                int pos = rn + p * WAVE + t;
                pos = pos >= W ? pos - W : pos;
                L.ring[ring_at(pos)] = nx[p];
                L.ring[ring_at(pos + W)] = nx[p];
            }
            if (c + 512 < c1) {
#pragma unroll
                for (int p = 0; p < 4; p++) nx[p] = V.x[c + 512 + K4 + p * WAVE + t];
            }
        }
    }
#pragma unroll
    for (int u = 0; u < DIRECT_ROWS; u++)
        if (u < D.nr && t == D.tr[u]) {
            R.mx = fmaxf(R.mx, mx[u]);
            R.mxl = fmaxf(R.mxl, mxl[u]);
            R.mxh = fmaxf(R.mxh, mxh[u]);
            if (wbest[u] > R.wbest) {
                R.wbest = wbest[u];
                R.wtile = wtile[u];
            }
        }
    R.alive = alive;
    return alive != 0;
}

// Columns at cyclic offsets [u0, u1) from the block start s: column (s + u) mod N, with the
// offsets clamped to [U0, N + U0) -- every column exactly once.  Scanning starts at
// U0 = round_up(m + 64, 64), past the exclusion band right of the block, so the early tiles
// that kill most blocks take the unmasked path; the band comes last, wrapped.  u0 is a
// multiple of 64: an unwrapped piece starts aligned, a wrapped one at the aligned-down
// column (re-scanning at most 63 columns is harmless: same values, same decisions).
template <bool COLS, bool DIRECT = false>
__device__ bool scan_cyclic(const ScanLds &L, RowState &R, const View &V, int64_t s, int64_t u0,
                            int64_t u1, int64_t U0, const DirectRows *D = nullptr)
{
    const int64_t N = V.N;
    u0 = u0 > U0 ? u0 : U0;
    u1 = u1 < N + U0 ? u1 : N + U0;
    if (u0 >= u1) return true;
    int64_t c = s + u0;
    bool wrapped = false;
    while (c >= N) {
        c -= N;
        wrapped = true;
    }
    const int64_t e = c + (u1 - u0);
    if (wrapped) c &= ~(int64_t)63;
    if (DIRECT) {
        bool live = scan_range_direct(L, R, V, s, *D, c, e < N ? e : N);
        if (live && e > N) live = scan_range_direct(L, R, V, s, *D, 0, e - N);
        return live;
    }
    bool live = scan_range<COLS>(L, R, V, s, c, e < N ? e : N);
    if (live && e > N) live = scan_range<COLS>(L, R, V, s, 0, e - N);
    return live;
}

__device__ __forceinline__ void load_rows(RowState &R, const View &V, int64_t s, float thr_rho)
{
    const int t = threadIdx.x;
    const float4 rv = V.colv[s + t];
    R.rdf = (t == 0) ? 0.f : rv.x;
    R.rdg = (t == 0) ? 0.f : rv.y;
    const float inv = rv.z;
    // a degenerate row is at 2m from every column: pruned iff 2m <= r2, exactly
    if (inv > 0.f)
        R.thr0 = thr_rho / inv;
    else
        R.thr0 = (thr_rho <= 0.f) ? -INFINITY : INFINITY;
    R.thr = R.thr0;
    R.mx = -INFINITY;
    R.mxl = -INFINITY;
    R.mxh = -INFINITY;
    R.tiles = 0;
    R.live = 0;
    R.dtiles = 0;
    R.wbest = -INFINITY;
    R.wtile = -1;
    R.mine = 0;
}

// Row terms plus the block's centred first row in L.qz (the seeds' right operand).
__device__ __forceinline__ void load_block(const ScanLds &L, RowState &R, const View &V,
                                           int64_t s, float thr_rho)
{
    load_rows(R, V, s, thr_rho);
    const int t = threadIdx.x;
    // sqrt(M2) of row s+t is 1/invn (0 for degenerate and padding rows); prefix maximum
    // over lanes 0..t
    const float inv = V.colv[s + t].z;
    R.kb = V.dk * wave_prefix_max(inv > 0.f ? 1.0f / inv : 0.f);
    if (t < 16) L.qz[t] = 0.f;
    for (int k = 16 + V.m + t; k < L.K4 + 20; k += WAVE) L.qz[k] = 0.f;
    R.qsum = stage_centred(L.qz + 16, V.x, s, V.m, V.colv[s].w);
}

__device__ __forceinline__ void flush_best(const RowState &R, const View &V, int64_t s)
{
    const int t = threadIdx.x;
    if (((R.alive >> t) & 1ull) && (s + t) < V.N && R.mx > -INFINITY) {
        const float inv = V.colv[s + t].z;
        atomicMax(V.best + s + t, ord_key(R.mx * inv));
        atomicMax(V.besth + s + t, ord_key(R.mxh * inv));
        atomicMax(V.bestl + s + t, ord_key(R.mxl * inv));
    }
}

// The witness of each row this workgroup pruned: the tile of its largest correlation among
// the tiles the workgroup evaluated (at least as close a neighbour as the pruning tile's).
__device__ __forceinline__ void record_witness(const RowState &R, const View &V, int64_t s)
{
    const int t = threadIdx.x;
    if (V.wit && ((R.mine >> t) & 1ull) && R.wtile >= 0) V.wit[s + t] = R.wtile;
}

__device__ __forceinline__ ScanLds carve_lds(int m)
{
    extern __shared__ float4 lds4[];
    ScanLds L;
    L.K4 = scan_k4(m);
    L.W = scan_ring_w(m);
    L.tile = lds4;
    L.seed = reinterpret_cast<float *>(lds4 + 2 * WAVE);
    L.qz = L.seed + 256;
    L.ring = L.qz + ((L.K4 + 20 + 3) & ~3);
    L.xs = L.ring;
    L.qi = L.xs + ((64 + m + 4 + 3) & ~3);
    return L;
}

// pairs: this kernel's PassCtl::pairs row; its PassCtl::live row is 2 PAIR_SLOTS further.
__device__ __forceinline__ void count_pairs(const RowState &R,
                                            unsigned long long *__restrict__ pairs)
{
    if (threadIdx.x == 0 && (R.tiles || R.dtiles)) {
        const int slot = blockIdx.x & (PAIR_SLOTS - 1);
        atomicAdd(pairs + slot,
                  (unsigned long long)R.tiles * 4096ull + (unsigned long long)R.dtiles * 64ull);
        atomicAdd(pairs + 2 * PAIR_SLOTS + slot,
                  (unsigned long long)R.live * 64ull + (unsigned long long)R.dtiles * 64ull);
    }
}

// Rebase the batch onto global block gb's series.  Returns the series id (-1: padding).
__device__ __forceinline__ int make_view(const ScanArgs &A, int64_t gb, View &V, int64_t &s,
                                         float &thr)
{
    const int sid = A.bser[gb];
    if (sid < 0) return -1;
    const SerDev S = A.ser[sid];
    V.x = A.x + S.xoff;
    V.colv = A.colv + S.roff;
    V.rho = A.rho + (S.roff >> 6);
    V.alive = A.alive + (S.roff >> 6);
    V.best = A.best + S.roff;
    V.besth = A.besth + S.roff;
    V.bestl = A.bestl + S.roff;
    V.dk = A.dk;
    V.wit = A.wit ? A.wit + S.roff : nullptr;
    V.N = S.N;
    V.m = A.m;
    V.ml = A.ml;
    s = gb * WAVE - S.roff;
    thr = S.thr;
    return sid;
}

// The scan arguments of a pass: buffers from the kernel argument, scalars from the pass
// parameters in device memory (uniform scalar loads).
__device__ __forceinline__ ScanArgs load_args(const ScanPtrs &Q, const PassDev *__restrict__ P)
{
    const ScanScal S = P->S;
    ScanArgs A;
    A.x = Q.x;
    A.ser = Q.ser;
    A.bser = Q.bser;
    A.colv = Q.colv;
    A.tmax = Q.tmax;
    A.rho = Q.rho;
    A.alive = Q.alive;
    A.best = Q.best;
    A.besth = Q.besth;
    A.bestl = Q.bestl;
    A.wit = S.wsel == 0 ? Q.wit0 : (S.wsel == 1 ? Q.wit1 : nullptr);
    A.wit_read = S.rsel == 0 ? Q.wit0 : (S.rsel == 1 ? Q.wit1 : nullptr);
    A.pairs = Q.pairs;
    A.WA = S.WA;
    A.CH = 0;
    A.U0 = S.U0;
    A.m = S.m;
    A.ml = S.ml;
    A.dbudget = S.dbudget;
    A.drows = S.drows;
    A.witness_max = S.witness_max;
    A.dk = S.dk;
    return A;
}

// Five waves per SIMD (a 96-register budget; a few bytes spill to scratch) measured faster
// than the four the unconstrained allocation leaves room for; so for k_scan_refine.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) void k_scan_select(ScanPtrs Q, const PassDev *__restrict__ P)
{
    const ScanArgs A = load_args(Q, P);
    const int64_t gb = blockIdx.x;
    View V;
    int64_t s;
    float thr_rho;
    if (make_view(A, gb, V, s, thr_rho) < 0 || s >= V.N) return;
    RowState R;
    R.alive = __hip_atomic_load(V.alive + (s >> 6), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (R.alive == 0) return;
    ScanLds L = carve_lds(V.m);
    load_block(L, R, V, s, thr_rho);
    bool live = true;
    if (A.wit_read) {
        // Witness tiles: the tiles that pruned these rows at the previous length.  Scan up
        // to witness_max distinct ones first; a run of adjacent witness tiles is one range
        // (one chain initialisation).
        int w = A.wit_read[gb * WAVE + threadIdx.x];
        const int64_t lim = V.N;
        for (int rep = 0; rep < A.witness_max && live; rep++) {
            const uint64_t has = __builtin_amdgcn_ballot_w64(w >= 0 && w < lim && !(w & 63));
            if (!has) break;
            const int lane = __builtin_ctzll(has);
            const int wt = __builtin_amdgcn_readlane(w, lane);
            int we = wt + WAVE;
            if (w == wt) w = -1;
            for (int ext = 0; ext < 3; ext++) {
                if (!__builtin_amdgcn_ballot_w64(w == we)) break;
                if (w == we) w = -1;
                we += WAVE;
            }
            const int64_t e = (we < lim) ? we : lim;
            live = scan_range<true>(L, R, V, s, wt, e);
        }
    }
    if (live) live = scan_cyclic<true>(L, R, V, s, A.U0, A.U0 + A.WA, A.U0);
    if (live) flush_best(R, V, s);
    record_witness(R, V, s);
    count_pairs(R, A.pairs);
}

// One refine work item: global block gb over cyclic offsets [u0, u1) (see scan_cyclic).
__device__ __forceinline__ void refine_item(const ScanArgs &A, int64_t gb, int64_t u0,
                                            int64_t u1)
{
    View V;
    int64_t s;
    float thr_rho;
    if (make_view(A, gb, V, s, thr_rho) < 0) return;
    if (u0 >= u1 || u0 >= V.N + A.U0) return;
    RowState R;
    R.alive = __hip_atomic_load(V.alive + (s >> 6), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (R.alive == 0) return;
    ScanLds L = carve_lds(V.m);
    load_block(L, R, V, s, thr_rho);
    bool live;
    const int nlive = __builtin_popcountll(R.alive);
    const int qs = (L.K4 + 20 + 3) & ~3;  // floats of one staged row window
    if (nlive <= A.drows && nlive * (V.m + 15) <= A.dbudget && (nlive - 1) * qs <= 2 * WAVE * 4) {
        // a few live rows: their covariances straight from the seed product
        // (scan_range_direct); the rows after the first stage their windows in the tile buffer
        DirectRows D;
        D.nr = nlive;
        uint64_t a = R.alive;
        float *tb = reinterpret_cast<float *>(L.tile);
#pragma unroll
        for (int u = 0; u < DIRECT_ROWS; u++) {
            D.tr[u] = a ? __builtin_ctzll(a) : D.tr[0];
            a &= a - 1;
            float *q = u == 0 ? L.qz : tb + (u - 1) * qs;
            D.qz[u] = q;
            D.qsum[u] = 0.f;
            if (u < nlive) {
                if (u > 0) {
                    for (int k = threadIdx.x; k < 16; k += WAVE) q[k] = 0.f;
                    for (int k = 16 + V.m + threadIdx.x; k < L.K4 + 20; k += WAVE) q[k] = 0.f;
                }
                D.qsum[u] = stage_centred(q + 16, V.x, s + D.tr[u], V.m, V.colv[s + D.tr[u]].w);
            }
        }
        live = scan_cyclic<false, true>(L, R, V, s, u0, u1, A.U0, &D);
    } else {
        live = scan_cyclic<false>(L, R, V, s, u0, u1, A.U0);
    }
    if (live) flush_best(R, V, s);
    record_witness(R, V, s);
    count_pairs(R, A.pairs + PAIR_SLOTS);
}

// Refine round: work item = (listed block, chunk q of the round's offset range).  The round's
// block count comes from k_list_blocks, so the host never reads it: rounds of growing offset
// let nearby columns kill most blocks before the far ones are scanned, and the chunk length
// keeps about R.items work items per round whatever that count is.  The grid is fixed, so a
// captured graph can replay it: at least 1.5 R.items workgroups and more than the number of
// blocks; workgroup b takes item b, and the workgroups beyond the round's items exit at once.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) void k_scan_refine(ScanPtrs Q, const PassDev *__restrict__ P,
                                                    const int *__restrict__ blist,
                                                    PassCtl *__restrict__ ctl, int round)
{
    const RefineArgs &RA = P->R;
    const int64_t ub = RA.bound[round], ue = RA.bound[round + 1];
    const int64_t nblk = ctl->list[round];
    if (nblk == 0 || ub >= ue) return;
    const int64_t len = ue - ub;
    int64_t ch = ((len * nblk + RA.items - 1) / RA.items + 63) & ~(int64_t)63;
    ch = ch < RA.min_chunk ? RA.min_chunk : (ch > RA.max_chunk ? RA.max_chunk : ch);
    // Chunks long enough that every item has a workgroup: with G = gridDim.x > nblk,
    // ch >= len * nblk / (G - nblk) gives nblk * ceil(len / ch) <= G.
    const int64_t G = gridDim.x;
    const int64_t chfit = ((len * nblk + (G - nblk) - 1) / (G - nblk) + 63) & ~(int64_t)63;
    ch = ch > chfit ? ch : chfit;
    const int64_t nchunks = (len + ch - 1) / ch;
    const int64_t item = blockIdx.x;
    if (item >= nblk * nchunks) return;
    ScanArgs A = load_args(Q, P);
    A.CH = ch;
    // block-major (a block's chunks dispatched together), or chunk-major (every block's
    // chunk 0, then every block's chunk 1 ...: workgroups are dispatched in order, so a
    // block's nearer chunks have mostly finished and pruned their rows before its farther
    // ones start)
    const int64_t bi = RA.order ? item % nblk : item / nchunks;
    const int64_t gb = blist[bi];
    const int64_t u0 = ub + (RA.order ? item / nblk : item % nchunks) * ch;
    const int64_t u1 = (u0 + ch < ue) ? u0 + ch : ue;
    refine_item(A, gb, u0, u1);
}

void launch_scan_select(const ScanPtrs &Q, const PassDev *P, unsigned grid, size_t lds,
                        hipStream_t st)
{
    hipLaunchKernelGGL(k_scan_select, dim3(grid), dim3(64), lds, st, Q, P);
}

void launch_scan_refine(const ScanPtrs &Q, const PassDev *P, const int *blist, PassCtl *ctl,
                        int round, unsigned grid, size_t lds, hipStream_t st)
{
    hipLaunchKernelGGL(k_scan_refine, dim3(grid), dim3(64), lds, st, Q, P, blist, ctl, round);
}

// ---------------------------------------------------------------------------------------

// Start of a pass: alive bits of every global row (1 for real rows, < N of their series;
// 0 for padding and for series switched off), the column factor of every block's tile,
// row maxima to -inf, this length's witness
// tiles of the pass's series to none (a series switched off keeps those of its last pass at
// this length, which the next length reads), counters to zero.  Fill kernel rather than
// memset nodes, so that a captured graph replays it faithfully.
__global__ __launch_bounds__(256) void k_pass_init(ScanPtrs Q, const PassDev *__restrict__ P,
                                                   int64_t rows, int64_t nwords,
                                                   PassCtl *__restrict__ ctl)
{
    const ScanArgs A = load_args(Q, P);
    uint64_t *alive = A.alive;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (P->S.off) {  // no pass: nothing alive, counters cleared, the rest untouched
        if (g < nwords) alive[g] = 0;
        if (g < (int64_t)(sizeof(PassCtl) / sizeof(int))) reinterpret_cast<int *>(ctl)[g] = 0;
        return;
    }
    if (g < rows) {
        A.best[g] = (int)0x807fffff;  // ord_key(-inf)
        A.besth[g] = (int)0x807fffff;
        A.bestl[g] = (int)0x807fffff;
        if (A.wit) {
            const int sid = P->S.some_off ? A.bser[g >> 6] : -1;
            if (sid < 0 || A.ser[sid].N > 0) A.wit[g] = -1;
        }
    }
    if (g < nwords) {
        const int sid = A.bser[g];
        uint64_t v = 0;
        if (sid >= 0) {
            const SerDev S = A.ser[sid];
            const int64_t r0 = g * 64 - S.roff;
            if (r0 + 64 <= S.N)
                v = ~0ull;
            else if (r0 < S.N)
                v = (~0ull) >> (64 - (S.N - r0));
        }
        alive[g] = v;
        // A scan tile over the 64 columns of block g holds chains started at most 63 columns
        // earlier, so the centred products it sums are bounded through sqrt(M2) over blocks
        // g-1 and g; its correlations are those products times invn of g's columns.
        float r = 0.f;
        if (sid >= 0) {
            const float2 cur = Q.tmax[g];
            float sm = cur.x;
            if (g > 0 && A.bser[g - 1] == sid) sm = fmaxf(sm, Q.tmax[g - 1].x);
            r = sm * cur.y;
        }
        Q.rho[g] = r;
    }
    if (g < (int64_t)(sizeof(PassCtl) / sizeof(int))) reinterpret_cast<int *>(ctl)[g] = 0;
}

__global__ __launch_bounds__(256) void k_list_blocks(const uint64_t *__restrict__ alive,
                                                     int64_t nwords, int *__restrict__ list,
                                                     int *__restrict__ count)
{
    // A workgroup lists LIST_WPT x 256 words and reserves its places with one atomic (an
    // atomic per live word, or per wave, serialised on the counter: 155 us per launch over
    // 2.5M words); within it, a thread's places follow a prefix sum over the lanes and waves.
    __shared__ int s_wave[4];
    __shared__ int s_base;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int64_t c0 = (int64_t)blockIdx.x * (256 * LIST_WPT);
    uint32_t mask = 0;
#pragma unroll
    for (int i = 0; i < LIST_WPT; i++) {
        const int64_t w = c0 + i * 256 + t;
        if (w < nwords && alive[w]) mask |= 1u << i;
    }
    const int n = __builtin_popcount(mask);
    int incl = n;  // inclusive prefix over the wave's lanes
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
    }
    if (lane == 63) s_wave[wv] = incl;
    __syncthreads();
    if (t == 0) {
        int tot = 0;
        for (int q = 0; q < 4; q++) {
            const int c = s_wave[q];
            s_wave[q] = tot;  // the waves' exclusive offsets
            tot += c;
        }
        s_base = tot ? atomicAdd(count, tot) : 0;
    }
    __syncthreads();
    int p = s_base + s_wave[wv] + incl - n;
    for (; mask; mask &= mask - 1) list[p++] = (int)(c0 + __builtin_ctz(mask) * 256 + t);
}

__global__ __launch_bounds__(256) void k_mask_blocks(uint64_t *__restrict__ alive,
                                                     const uint8_t *__restrict__ keep,
                                                     int64_t nwords)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w < nwords && !keep[w]) alive[w] = 0;
}

__global__ __launch_bounds__(256) void k_block_best(ScanPtrs Q, const PassDev *__restrict__ P,
                                                    int64_t nblocks,
                                                    unsigned long long *__restrict__ out)
{
    const int64_t b = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (b >= nblocks) return;
    const ScanArgs A = load_args(Q, P);
    const int sid = A.bser[b];
    unsigned long long key = ~0ull;
    if (sid >= 0) {
        const SerDev S = A.ser[sid];
        const int64_t g = b * 64 + lane;
        if (g - S.roff < S.N)
            key = ((unsigned long long)((unsigned)A.best[g] ^ 0x80000000u) << 32) |
                  (unsigned long long)(unsigned)g;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const unsigned long long v = __shfl_xor(key, o, 64);
        key = v < key ? v : key;
    }
    if (lane == 0) out[b] = key;
}

// Surviving rows -> (global row, the single-precision squared NN distance and its bounds);
// FLT_MAX when no valid pair was seen.
__global__ __launch_bounds__(256) void k_survivors(ScanPtrs Q, const PassDev *__restrict__ P,
                                                   int64_t rows, PassCtl *__restrict__ ctl,
                                                   Surv *__restrict__ out, int64_t cap)
{
    // one thread per 64-row block (a thread per row cost 170 us per pass on 1.6e8 rows, almost
    // all of them dead by now): its live rows, their places reserved by one atomic
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= (rows + 63) >> 6) return;
    const ScanArgs A = load_args(Q, P);
    uint64_t a = A.alive[w];
    if (w * 64 + 64 > rows) a &= (rows - w * 64 >= 64) ? ~0ull : ((1ull << (rows - w * 64)) - 1);
    if (!a) return;
    int p = atomicAdd(&ctl->survivors, __builtin_popcountll(a));
    const float m2 = 2.0f * (float)A.m;
    for (; a; a &= a - 1) {
        const int64_t g = w * 64 + __builtin_ctzll(a);
        const int k = A.best[g];
        float d = FLT_MAX, dlo = FLT_MAX, dhi = FLT_MAX;
        if (k != ord_key(-INFINITY)) {
            d = m2 * (1.0f - ord_val(k));
            dlo = m2 * (1.0f - ord_val(A.besth[g]));
            dhi = m2 * (1.0f - ord_val(A.bestl[g]));
        }
        if (p < cap) out[p] = Surv{g, d, dlo, dhi};
        p++;
    }
}

// ---------------------------------------------------------------------------------------
// Exact (FP64) nearest-neighbour squared distance of candidate groups.  A group is up to
// MAXU candidates of one series; block = (group, EXACT_COLS-column chunk of that series),
// found by binary search in goff (prefix sums of chunk counts).  The pair value is the
// reference's z-normalised distance (src/cpu_kernels/cpu_kernels.cpp:55) in double
// precision, ED^2 = 2m (1 - cov (invn_i invn_j)), with the covariance formed without the
// cancellation of the raw form (dot - m mu_i mu_j): the candidate's values enter centred on
// its mean, a_k = x[i+k] - mu_i (exact: x and mu_i are within a window's spread), the
// column's offset by the chunk's first value c (exact: a difference of two floats), and since
// sum_k a_k (x[j+k] - mu_j) is the centred product,
//     cov = sum_k a_k (x[j+k] - c) - (mu_j - c) sum_k a_k,
// whose rounding is relative to the chunk's local spread, not to the series' level.  The value
// is not bitwise symmetric in (i, j); the top-k treats distances within TIE_REL of each other
// as equal and ranks them by index, as the oracle's exact ties rank.
// outkey[u][b mod EXACT_SLOTS] receives the atomic minimum of ord_key64(ED^2) of candidate u
// over blocks b (the host takes the minimum over the slots).

// NU = the group size of the launch (a template parameter, so the inner loop is straight
// code); a smaller group fills its spare rows with copies of its first candidate and ignores
// their results.
//
// Thread t owns the QC = exact_qc(NU) adjacent columns c + QC t + q of a sub-tile of 256 QC
// columns and keeps a window of QC + 4 doubles of x - c in a register ring: per 4 steps of k it
// reads 4 new doubles (two 16-byte LDS reads; the staged window holds QC values per QC + 2
// doubles, so the lanes' reads, QC + 2 doubles apart, fall in distinct banks) and the
// candidates' next 4 values (broadcast), for 4 QC NU FP64 FMAs.  Value x[c + QC t + j] sits in
// ring slot j mod (QC + 4); the k loop is unrolled over one turn of the ring, so every slot
// index is a constant and the window never moves between registers.  Per column the sum runs
// in ascending k with one FMA per term; candidate rows are zero-padded to a multiple of 4,
// which adds exact zeros.
// The pair distance from its centred covariance (both rows' 1/sqrt(M2) at length m).
__device__ __forceinline__ double pair_ed2(double cov, int m, double invi, double invj)
{
    if (invi > 0.0 && invj > 0.0) return 2.0 * m * (1.0 - cov * (invi * invj));
    return 2.0 * m;
}

template <int QC>
__device__ __forceinline__ int xs_at(int p)
{
    return (QC + 2) * (int)((unsigned)p / QC) + (int)((unsigned)p % QC);
}

// One step of 4 k: xt = xs + (QC + 2) t, thread t's group of the staged window; the offset of
// value QC t + k0 + QC from there is wave-uniform.
template <int NU, int QC, int SLOT>
__device__ __forceinline__ void exact_step(const double *xt, const double *A, int m4, int k0,
                                           double (&w)[QC + 4], double (&acc)[NU][QC])
{
    constexpr int WR = QC + 4;
    {
        const unsigned p = (unsigned)k0 + QC;
        const double2 *g = reinterpret_cast<const double2 *>(xt + (QC + 2) * (p / QC) + p % QC);
        const double2 lo2 = g[0], hi2 = g[1];
        w[(SLOT + QC) % WR] = lo2.x;
        w[(SLOT + QC + 1) % WR] = lo2.y;
        w[(SLOT + QC + 2) % WR] = hi2.x;
        w[(SLOT + QC + 3) % WR] = hi2.y;
    }
    double a[NU][4];
#pragma unroll
    for (int u = 0; u < NU; u++) {
        const double2 *ap = reinterpret_cast<const double2 *>(A + u * m4 + k0);
        const double2 p0 = ap[0], p1 = ap[1];
        a[u][0] = p0.x; a[u][1] = p0.y; a[u][2] = p1.x; a[u][3] = p1.y;
    }
#pragma unroll
    for (int kk = 0; kk < 4; kk++)
#pragma unroll
        for (int u = 0; u < NU; u++)
#pragma unroll
            for (int q = 0; q < QC; q++)
                acc[u][q] = fma(a[u][kk], w[(SLOT + kk + q) % WR], acc[u][q]);
}

// Groups of three: per 16 steps of k each lane holds value k16 + (lane mod 16) of every
// candidate, and step k takes it from lane k mod 16 of each 16-lane row as the first operand
// of the FMA (DPP row_newbcast, no instruction of its own), so the candidates cost one LDS
// read per 16 steps instead of a broadcast read per 2 steps.  Written as VALU assembly (the
// compiler does not fold DPP broadcasts into FMAs); the block starts with s_nop 1, as the
// broadcast operands need two wait states after a VALU write of them.  Per accumulator the
// FMAs still run in ascending k.
#define XF(U, Q, W, K)                                                                         \
    "v_fmac_f64_dpp %[c" #U #Q "], %[a" #U "], %[w" #W "] row_newbcast:" #K                     \
    " row_mask:0xf bank_mask:0xf\n"
#define XK0(U, K) XF(U, 0, 0, K) XF(U, 1, 1, K) XF(U, 2, 2, K) XF(U, 3, 3, K)
#define XK1(U, K) XF(U, 0, 1, K) XF(U, 1, 2, K) XF(U, 2, 3, K) XF(U, 3, 4, K)
#define XK2(U, K) XF(U, 0, 2, K) XF(U, 1, 3, K) XF(U, 2, 4, K) XF(U, 3, 5, K)
#define XK3(U, K) XF(U, 0, 3, K) XF(U, 1, 4, K) XF(U, 2, 5, K) XF(U, 3, 6, K)
#define XQ3(K0, K1, K2, K3)                                                                    \
    asm("s_nop 1\n" XK0(0, K0) XK0(1, K0) XK0(2, K0) XK1(0, K1) XK1(1, K1) XK1(2, K1)           \
            XK2(0, K2) XK2(1, K2) XK2(2, K2) XK3(0, K3) XK3(1, K3) XK3(2, K3)                   \
        : [c00] "+v"(acc[0][0]), [c01] "+v"(acc[0][1]), [c02] "+v"(acc[0][2]),                 \
          [c03] "+v"(acc[0][3]), [c10] "+v"(acc[1][0]), [c11] "+v"(acc[1][1]),                 \
          [c12] "+v"(acc[1][2]), [c13] "+v"(acc[1][3]), [c20] "+v"(acc[2][0]),                 \
          [c21] "+v"(acc[2][1]), [c22] "+v"(acc[2][2]), [c23] "+v"(acc[2][3])                  \
        : [a0] "v"(a0), [a1] "v"(a1), [a2] "v"(a2), [w0] "v"(w[(SLOT + 0) % 8]),               \
          [w1] "v"(w[(SLOT + 1) % 8]), [w2] "v"(w[(SLOT + 2) % 8]), [w3] "v"(w[(SLOT + 3) % 8]), \
          [w4] "v"(w[(SLOT + 4) % 8]), [w5] "v"(w[(SLOT + 5) % 8]), [w6] "v"(w[(SLOT + 6) % 8]))

// Four steps of k (k0 = k16 + B) for a group of three: the window's 4 new values, then the
// 48 FMAs with the candidates' values of lanes B .. B + 3.
template <int SLOT, int B>
__device__ __forceinline__ void exact_quad3(const double *xt, int k0, double a0, double a1,
                                            double a2, double (&w)[8], double (&acc)[3][4])
{
    {
        const unsigned p = (unsigned)k0 + 4;
        const double2 *g = reinterpret_cast<const double2 *>(xt + 6 * (p / 4));
        const double2 lo2 = g[0], hi2 = g[1];
        w[(SLOT + 4) % 8] = lo2.x;
        w[(SLOT + 5) % 8] = lo2.y;
        w[(SLOT + 6) % 8] = hi2.x;
        w[(SLOT + 7) % 8] = hi2.y;
    }
    if (B == 0) XQ3(0, 1, 2, 3);
    if (B == 4) XQ3(4, 5, 6, 7);
    if (B == 8) XQ3(8, 9, 10, 11);
    if (B == 12) XQ3(12, 13, 14, 15);
}

// Scalars of one k_exact work item, per candidate of its group.
struct ExactShared {
    double ainv[MAXU], ainv2[MAXU], aext[MAXU][EXACT_EXT];
    double asa[MAXU], asb[MAXU];   // sum of a_k over k < m, and over k < m2
    double apart[MAXU][4];         // per-wave partial sums of a_k
    int64_t cidx[MAXU];
    unsigned long long wmin[MAXU][4];
};

// One work item: the candidates [u0, u0 + nU) of one series (sid) against its columns
// [chunk EXACT_COLS, (chunk + 1) EXACT_COLS); results into candidate slot `slot` mod
// EXACT_SLOTS.  m and (DUAL) m2 = the lengths.
// DUAL: also the distances at the second length m2 (a separate instantiation: its
// epilogue needs more registers, which would cost the one-length launches occupancy).
template <int NU, bool DUAL>
__device__ __forceinline__ void exact_item(const ExactArgs &E, const int m, const int m2_arg,
                                           const int sid, const int u0, const int nU,
                                           const int64_t chunk, const int64_t slot, double *dsh,
                                           ExactShared &X)
{
    constexpr int QC = exact_qc(NU);
    constexpr int SUBC = 256 * QC;   // columns per sub-tile
    constexpr int WR = QC + 4;       // window ring
    constexpr int UR = WR / 4;       // k steps of 4 per turn of the ring
    const int m4 = exact_mk(m, NU);                     // a multiple of 4 (of 16 for NU = 3)
    double *xs = dsh;                                   // exact_xs_doubles(m, NU)
    double *A = dsh + exact_xs_doubles(m, NU);          // [NU][m4] candidate windows
    double(&ainv)[MAXU] = X.ainv;
    double(&ainv2)[MAXU] = X.ainv2;
    double(&aext)[MAXU][EXACT_EXT] = X.aext;
    double(&asa)[MAXU] = X.asa;
    double(&asb)[MAXU] = X.asb;
    double(&apart)[MAXU][4] = X.apart;
    int64_t(&cidx)[MAXU] = X.cidx;
    unsigned long long(&wmin)[MAXU][4] = X.wmin;
    const int t = threadIdx.x;
    const SerDev S = E.ser[sid];
    const float *x = E.x + S.xoff;
    const double2 *st64 = E.st64 + S.roff;
    const int64_t N = S.n - m + 1;
    const int m2 = DUAL ? m2_arg : 0;
    const int ext = DUAL ? m2 - m : 0;
    const int64_t N2 = DUAL ? S.n - m2 + 1 : 0;
    // pairs iff j - i >= m - 1 or i - j >= ml (ExactArgs::sym)
    const int ml = E.sym ? m - 1 : m + 1, ml2 = E.sym ? m2 - 1 : m2 + 1;
    const double2 *st64n = E.st64n + S.roff;
    const int64_t cblk = chunk * EXACT_COLS;
    const double cref = (double)x[cblk];   // the columns' offset (cblk < N)
    for (int u = 0; u < NU; u++) {
        const int64_t i = E.cand[u0 + (u < nU ? u : 0)];
        const double2 st = st64[i];
        if (t == 0) {
            ainv[u] = st.y;
            const double2 s2 = i < N2 ? st64n[i] : make_double2(0.0, 0.0);
            ainv2[u] = s2.y;
            cidx[u] = i;
            for (int e = 0; e < EXACT_EXT; e++)
                aext[u][e] = (e < ext && i < N2) ? (double)x[i + m + e] - st.x : 0.0;
        }
        double part = 0.0;
        for (int k = t; k < m4; k += 256) {
            const double a = k < m ? (double)x[i + k] - st.x : 0.0;
            A[u * m4 + k] = a;
            part += a;
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) part += __shfl_xor(part, o, 64);
        if ((t & 63) == 0) apart[u][t >> 6] = part;
    }
    __syncthreads();
    if (t < NU) {
        const double sa = (apart[t][0] + apart[t][1]) + (apart[t][2] + apart[t][3]);
        double sb = sa;
        for (int e = 0; e < EXACT_EXT; e++) sb += aext[t][e];
        asa[t] = sa;
        asb[t] = sb;
    }
    unsigned long long kmin[NU], kmin2[NU];
#pragma unroll
    for (int u = 0; u < NU; u++) kmin[u] = kmin2[u] = ~0ull;
    const int lim = SUBC + m4 + 4;  // staged values per sub-tile
    for (int sub = 0; sub < EXACT_COLS / SUBC; sub++) {
        const int64_t c = cblk + sub * SUBC;
        if (c >= N) break;
        // Values at or past n only meet columns >= N (ignored) and are staged as zeros.
        const int avail = (int)(S.n - c < lim ? S.n - c : lim);
        __syncthreads();  // previous sub-tile's readers are done with xs
        // four loads in flight per thread, then the stores
        for (int k = t; k < lim; k += 1024) {
            float v[4];
#pragma unroll
            for (int p = 0; p < 4; p++) v[p] = (k + 256 * p < avail) ? x[c + k + 256 * p] : 0.f;
#pragma unroll
            for (int p = 0; p < 4; p++)
                if (k + 256 * p < lim)
                    xs[xs_at<QC>(k + 256 * p)] = (k + 256 * p < avail) ? (double)v[p] - cref : 0.0;
        }
        __syncthreads();
        double acc[NU][QC];
#pragma unroll
        for (int u = 0; u < NU; u++)
#pragma unroll
            for (int q = 0; q < QC; q++) acc[u][q] = 0.0;
        double w[WR];
#pragma unroll
        for (int q = 0; q < QC; q += 2) {
            const double2 v = *reinterpret_cast<const double2 *>(xs + xs_at<QC>(QC * t + q));
            w[q] = v.x;
            w[q + 1] = v.y;
        }
        const double *xt = xs + (QC + 2) * t;
        if constexpr (NU == 3) {
            for (int k16 = 0; k16 < m4; k16 += 16) {
                const double a0 = A[k16 + (t & 15)], a1 = A[m4 + k16 + (t & 15)],
                             a2 = A[2 * m4 + k16 + (t & 15)];
                exact_quad3<0, 0>(xt, k16, a0, a1, a2, w, acc);
                exact_quad3<4, 4>(xt, k16 + 4, a0, a1, a2, w, acc);
                exact_quad3<0, 8>(xt, k16 + 8, a0, a1, a2, w, acc);
                exact_quad3<4, 12>(xt, k16 + 12, a0, a1, a2, w, acc);
            }
        } else {
            int k0 = 0;
            for (; k0 + 4 * UR <= m4; k0 += 4 * UR) {
                exact_step<NU, QC, 0>(xt, A, m4, k0, w, acc);
                exact_step<NU, QC, 4>(xt, A, m4, k0 + 4, w, acc);
                if (UR == 3) exact_step<NU, QC, 8 % WR>(xt, A, m4, k0 + 8, w, acc);
            }
            if (k0 < m4) exact_step<NU, QC, 0>(xt, A, m4, k0, w, acc);
            if (UR == 3 && k0 + 4 < m4) exact_step<NU, QC, 4>(xt, A, m4, k0 + 4, w, acc);
        }
        // Per column: its statistics, then every candidate's distance at m and, with m2, at m2:
        // the FMA chain continued over k = m .. m2-1 with the candidate's values still centred
        // on its mean at m (the zero-padded terms between add exact zeros), which gives the
        // centred product at m2 after the same correction with the column's mean at m2 and
        // the sum of those values over k < m2 (the product is the same for any centring of
        // one side).  Staged values reach column offset 256 QC + m4 + 3 >=
        // 256 QC - 1 + m2 - 1.
#pragma unroll
        for (int q = 0; q < QC; q++) {
            // keep each column's loads in its own iteration (hoisted, the second length's
            // terms of every column would take registers from the main loop's occupancy)
            if (DUAL) __builtin_amdgcn_sched_barrier(0);
            const int64_t j = c + QC * t + q;
            const double2 sj = (j < N) ? st64[j] : make_double2(0.0, 0.0);
            double2 sj2 = make_double2(0.0, 0.0);
            double xe[EXACT_EXT];
            if (DUAL) {
                if (j < N2) sj2 = st64n[j];
#pragma unroll
                for (int e = 0; e < EXACT_EXT; e++)
                    xe[e] = e < ext ? xs[xs_at<QC>(QC * t + q + m + e)] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < NU; u++) {
                const int64_t i = cidx[u];
                const bool valid = (j < N) && ((j - i >= m - 1) || (i - j >= ml));
                const double cov = fma(-(sj.x - cref), asa[u], acc[u][q]);
                const double ed = valid ? pair_ed2(cov, m, ainv[u], sj.y) : DBL_MAX;
                const unsigned long long key = ord_key64(ed);
                kmin[u] = key < kmin[u] ? key : kmin[u];
                if (DUAL) {
                    double a2 = acc[u][q];
#pragma unroll
                    for (int e = 0; e < EXACT_EXT; e++)
                        if (e < ext) a2 = fma(aext[u][e], xe[e], a2);
                    const bool valid2 =
                        (i < N2) && (j < N2) && ((j - i >= m2 - 1) || (i - j >= ml2));
                    const double cov2 = fma(-(sj2.x - cref), asb[u], a2);
                    const double ed2 = valid2 ? pair_ed2(cov2, m2, ainv2[u], sj2.y) : DBL_MAX;
                    const unsigned long long key2 = ord_key64(ed2);
                    kmin2[u] = key2 < kmin2[u] ? key2 : kmin2[u];
                }
            }
        }
    }
    const int lane = t & 63, w = t >> 6;
#pragma unroll
    for (int u = 0; u < NU; u++) {
        unsigned long long key = kmin[u];
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
        atomicMin(E.outkey + (size_t)(u0 + t) * EXACT_SLOTS + (slot & (EXACT_SLOTS - 1)), k);
    }
    if (DUAL) {
        __syncthreads();  // wmin is read above
#pragma unroll
        for (int u = 0; u < NU; u++) {
            unsigned long long key = kmin2[u];
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
            atomicMin(E.outkey2 + (size_t)(u0 + t) * EXACT_SLOTS + (slot & (EXACT_SLOTS - 1)),
                      k);
        }
    }
}

// Host-planned launch: block = (group, chunk of that group's series), the group found by
// binary search in goff.
template <int NU, bool DUAL>
__global__ __launch_bounds__(256) void k_exact(ExactArgs E)
{
    extern __shared__ double dsh[];
    __shared__ ExactShared X;
    const int64_t item = blockIdx.x;
    int lo = 0, hi = E.ngroups - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (E.goff[mid] <= item) lo = mid; else hi = mid - 1;
    }
    const int u0 = E.gfirst[lo];
    exact_item<NU, DUAL>(E, E.m, E.m2, E.gser[lo], u0, E.gfirst[lo + 1] - u0, item - E.goff[lo],
                         item, dsh, X);
}

// Device-planned launch of the length loop on the device: series s has `gslots` groups of up
// to NU candidates (k_chain_topk or k_chain_rreq writes group g's candidates at
// cand[(s gslots + g) NU ..] and their number at dev_count[s gslots + g]); block = (group slot,
// series, chunk of its columns at the loop's first, longest windows), the series found by
// binary search in goff (fixed per batch; E.ngroups = the number of series), so blocks of an
// empty group, and chunks past its series' windows at the current length, end at once.  The
// lengths come from the device too (dev_m[0] = m, dev_m[1] = m2).
template <int NU, bool DUAL>
__global__ __launch_bounds__(256) void k_exact_chain(ExactArgs E)
{
    extern __shared__ double dsh[];
    __shared__ ExactShared X;
    const int64_t total = E.goff[E.ngroups];
    const int64_t slot = blockIdx.x / total, item = blockIdx.x - slot * total;
    int lo = 0, hi = E.ngroups - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (E.goff[mid] <= item) lo = mid; else hi = mid - 1;
    }
    const int gi = lo * E.gslots + (int)slot;
    const int nU = E.dev_count[gi];
    const int m = E.dev_m[0], m2 = E.dev_m[1];
    if (nU <= 0 || m <= 0) return;
    const int64_t chunk = item - E.goff[lo];
    if (chunk * EXACT_COLS >= E.ser[lo].n - m + 1) return;
    exact_item<NU, DUAL>(E, m, m2, lo, gi * NU, nU, chunk, blockIdx.x, dsh, X);
}

// ---------------------------------------------------------------------------------------
// k_nn: nearest-neighbour distance of a few candidates on the matrix cores, exact where it
// matters (replaces k_exact; same arguments and results).
//
// Filter pass (k_nn / k_nn_chain).  A work item is the candidates [u0, u0 + nU) of one series
// against the EXACT_COLS columns of a chunk, whose series window is staged in LDS once.  Each
// wave takes four of the chunk's 256-column super-tiles and, per candidate, forms the 256
// single-precision covariances of a super-tile as one 16 x 16 x K4 product
// (v_mfma_f32_16x16x4f32 in the Toeplitz arrangement of super_seeds):
// A[a][k'] = x[c+16a+k'] - x[c+16a], B[k'][j] = q[k'-j] with q the candidate's window centred
// on its mean, so D[a][j] = sum_k q[k] (x[c+16a+j+k] - x[c+16a]) and
// cov = D - (mu_n - x[c+16a]) sum(q) at column n = 16a + j.  The rounding of a column's
// distance is bounded per column: |cov32 - cov| <= (K4 + 4) u |q| |y| (K4 accumulations and
// the roundings of q and of x - ref; y the column's window less ref), which in distance units
// is 2m (K4 + 4) u f_n, f_n = |y| invn_n <= 1 + sqrt(m) |(mu_n - ref) invn_n|, plus the
// conversion's rounding (conv_ed2 at 4m); E_t is the largest over the super-tile's columns.
// With d_t the super-tile's smallest single-precision distance, every column's exact distance
// is at least lo_t = d_t - E_t, and the nearest neighbour is no farther than d_t + E_t.  The
// filter writes lo_t per super-tile (tilelo) and the smallest d_t + E_t per workgroup (runub,
// one of EXACT_SLOTS addresses).  k_nn_fix then evaluates in double precision exactly the
// super-tiles with lo_t <= min(the scans' bound for the candidate, every d_t + E_t); the
// minimum over those is the exact nearest-neighbour distance, every other column being
// farther.  With E.lb only the lower bound min_t lo_t is formed (the shifted range, which
// needs a bound below the value).  DUAL: also a lower bound at the second length m2, from the
// same product extended by the window's m2 - m extra terms (still centred on its mean at m)
// and the columns' statistics at m2 (st64n): the next length's range only needs a value at
// most the exact one.
struct NnShared {
    float invn[NN_MAXU], invn2[NN_MAXU], sq[NN_MAXU], sq2[NN_MAXU];
    float qe[NN_MAXU][EXACT_EXT];
    double dinv[NN_MAXU], sa[NN_MAXU], ub[NN_MAXU];
    double part[NN_MAXU][4], fpart[NN_MAXU][4];
    int64_t cidx[NN_MAXU];
    unsigned long long wk[NN_MAXU][4], wk2[NN_MAXU][4];
    int hk[NN_MAXU][4];
    int rel[NN_MAXU];   // candidate row - chunk start (the pair tests in int)
    int in2[NN_MAXU];   // the candidate has a window at m2
};

__device__ __forceinline__ int nn_pad(int p) { return p + (p >> 4); }

__device__ __forceinline__ float nn_wave_min(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double nn_wave_min(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

template <int NU, bool DUAL, bool LB>
__device__ __forceinline__ void nn_item(const ExactArgs &E, const int m, const int m2_arg,
                                        const int sid, const int u0, const int nU,
                                        const int64_t chunk, const int64_t slot, char *dsh,
                                        NnShared &X)
{
    // No contraction beyond the fma calls written out: the values must not depend on the
    // instantiation (group size) that computed them.
#pragma clang fp contract(off)
    typedef float f4 __attribute__((ext_vector_type(4)));
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int K4 = (m + 15 + 3) & ~3;
    const int XS = EXACT_COLS + K4 + 16;
    const int QZ = K4 + 32;
    float *xs = (float *)dsh;
    float *qz = xs + ((XS + (XS >> 4) + 2 + 1) & ~1);  // [NU][QZ]
    const SerDev S = E.ser[sid];
    const float *x = E.x + S.xoff;
    const double2 *st64 = E.st64 + S.roff;
    const int64_t N = S.n - m + 1;
    const int m2 = DUAL ? m2_arg : 0;
    const int ext = DUAL ? m2 - m : 0;
    const int64_t N2 = DUAL ? S.n - m2 + 1 : 0;
    // pairs iff j - i >= m - 1 or i - j >= ml (ExactArgs::sym)
    const int ml = E.sym ? m - 1 : m + 1, ml2 = E.sym ? m2 - 1 : m2 + 1;
    const int64_t cblk = chunk * EXACT_COLS;
    // the chunk's window (values past the series only meet columns >= N: zeros)
    for (int p = t; p < XS; p += 256) {
        const int64_t g = cblk + p;
        xs[nn_pad(p)] = g < S.n ? x[g] : 0.f;
    }
    // the candidates' windows, centred on their means
#pragma unroll
    for (int u = 0; u < NU; u++) {
        const int64_t i = E.cand[u0 + (u < nU ? u : 0)];
        const double mu = st64[i].x;
        double pf = 0.0, pd = 0.0;
        for (int k = t; k < QZ; k += 256) {
            const int kq = k - 16;
            const double a = (kq >= 0 && kq < m) ? (double)x[i + kq] - mu : 0.0;
            const float af = (float)a;
            qz[u * QZ + k] = af;
            pf += (double)af;
            pd += a;
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            pf += __shfl_xor(pf, o, 64);
            pd += __shfl_xor(pd, o, 64);
        }
        if (lane == 0) {
            X.fpart[u][w] = pf;
            X.part[u][w] = pd;
        }
        if (t == 0) {
            const double2 s1 = st64[i];
            X.cidx[u] = i;
            X.rel[u] = (int)(i - cblk);
            X.in2[u] = DUAL && i < N2;
            X.dinv[u] = s1.y;
            X.invn[u] = (float)s1.y;
            X.ub[u] = E.ub ? E.ub[u0 + (u < nU ? u : 0)] : INFINITY;
            const bool in2 = DUAL && i < N2;
            X.invn2[u] = in2 ? (float)E.st64n[S.roff + i].y : 0.f;
            for (int e = 0; e < EXACT_EXT; e++)
                X.qe[u][e] = (e < ext && in2) ? (float)((double)x[i + m + e] - mu) : 0.f;
        }
    }
    __syncthreads();
    if (t < NU) {
        const double f = (X.fpart[t][0] + X.fpart[t][1]) + (X.fpart[t][2] + X.fpart[t][3]);
        X.sq[t] = (float)f;
        double f2 = f;
        for (int e = 0; e < EXACT_EXT; e++) f2 += (double)X.qe[t][e];
        X.sq2[t] = (float)f2;
        X.sa[t] = (X.part[t][0] + X.part[t][1]) + (X.part[t][2] + X.part[t][3]);
    }
    __syncthreads();
    const int i16 = lane & 15, kk = lane >> 4;
    const float m2f = 2.0f * (float)m, m2f2 = 2.0f * (float)m2;
    // (K + 4) u per term count K of the single-precision products (u = 2^-24, with a factor
    // 1.25 of room)
    const float ek = m2f * 1.25f * (float)(K4 + 4) * 5.9604645e-8f;
    const float ek2 = m2f2 * 1.25f * (float)(K4 + 4 + EXACT_EXT) * 5.9604645e-8f;
    const float sqm = sqrtf((float)m), sqm2 = sqrtf((float)m2);
    // the conversion's rounding at the largest distance, 4m (conv_ed2)
    const float cv1 = (float)conv_ed2(m, 4.0 * m), cv2 = (float)conv_ed2(m2, 4.0 * m2);
    const int64_t cbase = cblk;  // columns relative to the chunk, candidates too (X.rel)
    float lbm[NU], lb2[NU], hib[NU];
#pragma unroll
    for (int u = 0; u < NU; u++) {
        lbm[u] = INFINITY;
        lb2[u] = INFINITY;
        hib[u] = INFINITY;
    }
    for (int sct = 0; sct < 4; sct++) {
        const int cl = 256 * (w + 4 * sct);
        const int64_t c = cblk + cl;
        if (c >= N) break;
        f4 acc0[NU], acc1[NU];
#pragma unroll
        for (int u = 0; u < NU; u++) {
            acc0[u] = f4{0.f, 0.f, 0.f, 0.f};
            acc1[u] = f4{0.f, 0.f, 0.f, 0.f};
        }
        const float *ap = xs + nn_pad(cl + 16 * i16) + kk;
        const float refi = xs[nn_pad(cl + 16 * i16)];
        const float *bp = qz + 16 + kk - i16;
        int k0 = 0;
        for (; k0 + 16 <= K4; k0 += 16, ap += 17, bp += 16) {
            const float a0 = ap[0] - refi, a1 = ap[4] - refi, a2 = ap[8] - refi,
                        a3 = ap[12] - refi;
#pragma unroll
            for (int u = 0; u < NU; u++) {
                const float *b = bp + u * QZ;
                acc0[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b[0], acc0[u], 0, 0, 0);
                acc1[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b[4], acc1[u], 0, 0, 0);
                acc0[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a2, b[8], acc0[u], 0, 0, 0);
                acc1[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a3, b[12], acc1[u], 0, 0, 0);
            }
        }
        for (int e = 0; k0 < K4; k0 += 4, e += 4) {
            const float a0 = ap[e] - refi;
#pragma unroll
            for (int u = 0; u < NU; u++)
                acc0[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, bp[u * QZ + e], acc0[u], 0, 0, 0);
        }
        float dmin[NU];
#pragma unroll
        for (int u = 0; u < NU; u++) dmin[u] = INFINITY;
        float emax = 0.f;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int n = 16 * (4 * kk + r) + i16;
            const int64_t j = c + n;
            if (j >= N) continue;
            const float ref = xs[nn_pad(cl + 16 * (4 * kk + r))];
            const float4 cv = E.colv[S.roff + j];
            // |cov32 - cov| <= (K4 + 4) u |q| |y| with y = x - ref over the column's window
            // (K4 accumulations, the rounding of q and of x - ref), |q| = sqrt(M2_i) and
            // |y|^2 = M2_n + m (mu_n - ref)^2; in distance units 2m (K4 + 4) u f_n with
            // f_n = |y| invn_n = sqrt(1 + m ((mu_n - ref) invn_n)^2) <= 1 + sqrt(m) |..|
            const float tn = (cv.w - ref) * cv.z;
            emax = fmaxf(emax, fmaf(sqm, fabsf(tn), 1.0f));
            float mu2 = 0.f, inv2 = 0.f, xe[EXACT_EXT], en2 = 0.f;
            if (DUAL && j < N2) {
                const double2 s2 = E.st64n[S.roff + j];
                mu2 = (float)s2.x;
                inv2 = (float)s2.y;
#pragma unroll
                for (int e = 0; e < EXACT_EXT; e++)
                    xe[e] = e < ext ? xs[nn_pad(cl + n + m + e)] - ref : 0.f;
                en2 = fmaf(ek2, fmaf(sqm2, fabsf((mu2 - ref) * inv2), 1.0f), cv2);
            }
            const int dj = (int)(j - cbase);  // j - i = dj - (i - cbase), both within int
#pragma unroll
            for (int u = 0; u < NU; u++) {
                const int di = dj - X.rel[u];  // j - i
                const float D = acc0[u][r] + acc1[u][r];
                if (di >= m - 1 || -di >= ml) {
                    const float cov = fmaf(-(cv.w - ref), X.sq[u], D);
                    const float d = (X.invn[u] > 0.f && cv.z > 0.f)
                                        ? m2f * fmaf(-cov, X.invn[u] * cv.z, 1.0f) : m2f;
                    dmin[u] = fminf(dmin[u], d);
                }
                if (DUAL && j < N2 && X.in2[u] && (di >= m2 - 1 || -di >= ml2)) {
                    float c2 = D;
#pragma unroll
                    for (int e = 0; e < EXACT_EXT; e++) c2 = fmaf(X.qe[u][e], xe[e], c2);
                    c2 = fmaf(-(mu2 - ref), X.sq2[u], c2);
                    const float d2 = (X.invn2[u] > 0.f && inv2 > 0.f)
                                         ? m2f2 * fmaf(-c2, X.invn2[u] * inv2, 1.0f) : m2f2;
                    lb2[u] = fminf(lb2[u], d2 - en2);
                }
            }
        }
        // the super-tile's bound: lo_t = min d - E_t <= every column's exact distance, and
        // min d + E_t >= the exact distance of the column attaining min d
        const float Et = fmaf(ek, -nn_wave_min(-emax), cv1);
#pragma unroll
        for (int u = 0; u < NU; u++) {
            const float dt = nn_wave_min(dmin[u]);
            const float lo = dt - Et;  // infinite: no valid column
            if (LB) {
                lbm[u] = fminf(lbm[u], lo);
                continue;
            }
            // the super-tile's lower bound for k_nn_fix, and the candidate's upper bound
            hib[u] = fminf(hib[u], dt + Et);
            if (lane == 0 && u < nU) E.tilelo[(size_t)(u0 + u) * E.tiles + (c >> 8)] = lo;
        }
    }
    // per candidate: minimum over the waves, one atomic per workgroup and result (the lower
    // bound; in the filter pass only the second length's)
#pragma unroll
    for (int u = 0; u < NU; u++) {
        const double v = (double)lbm[u];
        const double v2 = DUAL ? (double)nn_wave_min(lb2[u]) : INFINITY;
        if (lane == 0) {
            X.wk[u][w] = (LB && v < INFINITY) ? ord_key64(v) : ~0ull;
            X.wk2[u][w] = v2 < INFINITY ? ord_key64(v2) : ~0ull;
            X.hk[u][w] = ord_key(hib[u]);
        }
    }
    __syncthreads();
    if (t < nU) {
        unsigned long long k = X.wk[t][0], k2 = X.wk2[t][0];
        int h = X.hk[t][0];
        for (int q = 1; q < 4; q++) {
            k = X.wk[t][q] < k ? X.wk[t][q] : k;
            k2 = X.wk2[t][q] < k2 ? X.wk2[t][q] : k2;
            h = X.hk[t][q] < h ? X.hk[t][q] : h;
        }
        // one atomic per workgroup and result, spread over EXACT_SLOTS addresses (the
        // workgroups' atomics on one address serialise)
        const size_t o = (size_t)(u0 + t) * EXACT_SLOTS + (slot & (EXACT_SLOTS - 1));
        if (LB && k != ~0ull) atomicMin(E.outkey + o, k);
        if (DUAL && k2 != ~0ull) atomicMin(E.outkey2 + o, k2);
        if (!LB && h < 0x7f800000) atomicMin(E.runub + o, h);
    }
}

// Double-precision pass of k_nn (after its filter pass): wave = (candidate c, super-tiles
// 64 tb .. 64 tb + 63 of its series).  The super-tiles whose lower bound reaches the
// candidate's upper bound (the scans' bound, and the smallest single-precision upper bound over
// all its columns from the filter pass) are evaluated in double precision, centred as k_exact
// (candidate on its mean, columns offset by the super-tile's first value); the minimum over
// them is the exact nearest-neighbour distance, every other column being farther.  Host-
// planned: candidate c of the launch, series csid[c]; device-planned: c = s nu + u, valid
// when u < dev_count[s].
template <bool DEV>
__global__ __launch_bounds__(64) void k_nn_fix(ExactArgs E)
{
    __shared__ double aw[MAXM];              // the candidate's window, centred on its mean
    __shared__ float xw[256 + MAXM + 8];     // a super-tile's series window
    const int lane = threadIdx.x;
    const int c = (int)(blockIdx.x / E.tb), tb = (int)(blockIdx.x % E.tb);
    int sid, m;
    if (DEV) {
        const int q = c / E.nu_dev, u = c - q * E.nu_dev;
        m = E.dev_m[0];
        if (m <= 0 || u >= E.dev_count[q]) return;
        sid = q;
    } else {
        sid = E.csid[c];
        m = E.m;
    }
    const SerDev S = E.ser[sid];
    const float *x = E.x + S.xoff;
    const double2 *st64 = E.st64 + S.roff;
    const int64_t N = S.n - m + 1;
    const int T = (int)((N + 255) >> 8);
    const int t0 = tb * 64 + lane;
    const float lo = t0 < T ? E.tilelo[(size_t)c * E.tiles + t0] : INFINITY;
    int hk = E.runub[(size_t)c * EXACT_SLOTS + lane];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const int h2 = __shfl_xor(hk, o, 64);
        hk = h2 < hk ? h2 : hk;
    }
    double ub = (double)ord_val(hk);
    if (E.ub) ub = fmin(ub, E.ub[c]);
    // lo is infinite for a super-tile without a valid column (and past the series): never
    // evaluated, also when the bound is infinite (a candidate without any valid column)
    unsigned long long mask = __ballot(lo < INFINITY && (double)lo <= ub);
    if (!mask) return;
    const int64_t i = E.cand[c];
    const double2 si = st64[i];
    double part = 0.0;
    for (int k = lane; k < m; k += 64) {
        const double a = (double)x[i + k] - si.x;
        aw[k] = a;
        part += a;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) part += __shfl_xor(part, o, 64);
    const double sa = part;
    const int ml = E.sym ? m - 1 : m + 1;
    double best = INFINITY;
    while (mask) {
        const int tt = __builtin_ctzll(mask);
        mask &= mask - 1;
        const int64_t c0 = (int64_t)(tb * 64 + tt) * 256;
        __syncthreads();  // the previous super-tile's readers of xw are done
        for (int p = lane; p < 256 + m; p += 64) xw[p] = c0 + p < S.n ? x[c0 + p] : 0.f;
        __syncthreads();
        const double cref = (double)xw[0];
        double cv4[4] = {0.0, 0.0, 0.0, 0.0};
        for (int k = 0; k < m; k++) {
            const double a = aw[k];
#pragma unroll
            for (int h = 0; h < 4; h++) cv4[h] = fma(a, (double)xw[lane + 64 * h + k] - cref, cv4[h]);
        }
#pragma unroll
        for (int h = 0; h < 4; h++) {
            const int64_t j = c0 + lane + 64 * h;
            if (j >= N || !((j - i >= m - 1) || (i - j >= ml))) continue;
            const double2 sj = st64[j];
            const double cov = fma(-(sj.x - cref), sa, cv4[h]);
            const double dd = (si.y > 0.0 && sj.y > 0.0) ? (2.0 * m) * fma(-cov, si.y * sj.y, 1.0)
                                                         : 2.0 * m;
            best = fmin(best, dd);
        }
    }
    best = nn_wave_min(best);
    if (lane == 0 && best < INFINITY)
        atomicMin(E.outkey + (size_t)c * EXACT_SLOTS + (tb & (EXACT_SLOTS - 1)), ord_key64(best));
}

// Host-planned launch (as k_exact): block = (group, chunk of that group's series).
template <int NU, bool DUAL, bool LB>
__global__ __launch_bounds__(256) void k_nn(ExactArgs E)
{
    extern __shared__ char nsh[];
    __shared__ NnShared X;
    const int64_t item = blockIdx.x;
    int lo = 0, hi = E.ngroups - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (E.goff[mid] <= item) lo = mid; else hi = mid - 1;
    }
    const int u0 = E.gfirst[lo];
    nn_item<NU, DUAL, LB>(E, E.m, E.m2, E.gser[lo], u0, E.gfirst[lo + 1] - u0,
                          item - E.goff[lo], item, nsh, X);
}

// Device-planned launch of the length loop (as k_exact_chain).
template <int NU, bool DUAL, bool LB>
__global__ __launch_bounds__(256) void k_nn_chain(ExactArgs E)
{
    extern __shared__ char nsh[];
    __shared__ NnShared X;
    const int64_t total = E.goff[E.ngroups];
    const int64_t slot = blockIdx.x / total, item = blockIdx.x - slot * total;
    int lo = 0, hi = E.ngroups - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (E.goff[mid] <= item) lo = mid; else hi = mid - 1;
    }
    const int gi = lo * E.gslots + (int)slot;
    const int nU = E.dev_count[gi];
    const int m = E.dev_m[0], m2 = E.dev_m[1];
    if (nU <= 0 || m <= 0) return;
    const int64_t chunk = item - E.goff[lo];
    if (chunk * EXACT_COLS >= E.ser[lo].n - m + 1) return;
    nn_item<NU, DUAL, LB>(E, m, m2, lo, gi * NU, nU, chunk, blockIdx.x, nsh, X);
}

template <int NU>
static void launch_nn_u(const ExactArgs &E, bool dual, bool lb, bool dev, unsigned grid,
                        size_t lds, hipStream_t st)
{
    if (dev) {
        if (lb)
            hipLaunchKernelGGL((k_nn_chain<NU, false, true>), dim3(grid), dim3(256), lds, st, E);
        else if (dual)
            hipLaunchKernelGGL((k_nn_chain<NU, true, false>), dim3(grid), dim3(256), lds, st, E);
        else
            hipLaunchKernelGGL((k_nn_chain<NU, false, false>), dim3(grid), dim3(256), lds, st, E);
    } else {
        if (lb)
            hipLaunchKernelGGL((k_nn<NU, false, true>), dim3(grid), dim3(256), lds, st, E);
        else if (dual)
            hipLaunchKernelGGL((k_nn<NU, true, false>), dim3(grid), dim3(256), lds, st, E);
        else
            hipLaunchKernelGGL((k_nn<NU, false, false>), dim3(grid), dim3(256), lds, st, E);
    }
}

void launch_nn_fix(const ExactArgs &E, bool dev, unsigned grid, hipStream_t st)
{
    if (dev)
        hipLaunchKernelGGL((k_nn_fix<true>), dim3(grid), dim3(64), 0, st, E);
    else
        hipLaunchKernelGGL((k_nn_fix<false>), dim3(grid), dim3(64), 0, st, E);
}

void launch_nn(const ExactArgs &E, int nu, bool dual, bool lb, bool dev, unsigned grid,
               size_t lds, hipStream_t st)
{
    switch (nu) {
    case 1: launch_nn_u<1>(E, dual, lb, dev, grid, lds, st); break;
    case 2: launch_nn_u<2>(E, dual, lb, dev, grid, lds, st); break;
    case 3: launch_nn_u<3>(E, dual, lb, dev, grid, lds, st); break;
    default: launch_nn_u<4>(E, dual, lb, dev, grid, lds, st); break;
    }
}

// ---------------------------------------------------------------------------------------
// The length loop on the device (see ChainArgs).

__device__ __forceinline__ double ord_val64(unsigned long long k)
{
    const unsigned long long bits = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)bits);
}

// Start of length j: per series its squared range from the previous length's discords
// evaluated exactly at this length, as Engine::fast takes it (the minimum over the discords
// still inside the series, times 1 - 1e-6; every discord must have its value), the pass's
// series table (N, thr) and scalars (Engine::fill_pass), and the lengths k_derive and
// k_exact_chain read.  A series without k such values stops the loop at this length.  With a
// shifted range (C.shift > 0) the range and the series' pass table entries come later
// (k_chain_range); here only the number of discords is checked.  Run by k_chain_next (one
// workgroup).
__device__ __forceinline__ void chain_begin(const ChainArgs &C, const int j, int *s_bad,
                            unsigned long long *s_nmax)
{
    ChainCtl *ctl = C.ctl;
    const int t = threadIdx.x, nt = blockDim.x;
    const int m = C.minL + j * C.step;
    const bool halted0 = ctl->halt != 0;
    if (t == 0) {
        *s_bad = 0;
        *s_nmax = 0;
    }
    __syncthreads();
    for (int q = t; q < C.nser && !halted0; q += nt) {
        ChainSer &cs = C.cs[q];
        const int64_t Nq = C.ser[q].n - m + 1;
        // the previous length's discords: the loop's initial state, or that length's results
        // (finished by every series, or the loop would have stopped)
        const ChainRes *R = j > ctl->j0 ? &C.res[(size_t)(j - 1 - ctl->j0) * C.nser + q] : nullptr;
        const int64_t *prev = R ? R->i : cs.prev;
        const double *pv = R ? R->v2 : cs.pv;
        int np = R ? R->n : cs.nprev;
        np = np < 0 ? 0 : (np > CHAIN_KMAX ? CHAIN_KMAX : np);
        bool ok = np >= C.k;
        if (C.shift == 0) {
            double mn = INFINITY;
            int cnt = 0;
            for (int u = 0; u < np; u++) {
                ok &= pv[u] >= 0.0;
                if (prev[u] >= Nq) continue;
                cnt++;
                if (pv[u] < 3.0e38) mn = fmin(mn, pv[u]);
            }
            ok &= cnt >= C.k && mn < INFINITY && mn > 0.0;
            cs.r2 = ok ? mn * (1.0 - 1e-6) : 0.0;
        }  // shifted: the range comes from the length's own launch (k_chain_range)
        cs.status = 0;
        cs.nreq = 0;
        if (!ok) atomicOr(s_bad, 1);
        if (Nq > 0) atomicMax(s_nmax, (unsigned long long)Nq);
    }
    __syncthreads();
    const bool halted = halted0 || *s_bad != 0;
    for (int q = t; q < C.nser && (C.shift == 0 || halted); q += nt) {
        SerDev S = C.ser[q];
        S.N = halted ? 0 : S.n - m + 1;
        S.thr = (float)(1.0 - C.cs[q].r2 / (2.0 * m));
        C.pser[q] = S;
    }
    if (t == 0) {
        if (!halted0 && *s_bad) ctl->halt = j + 1;
        ScanScal &A = C.par->S;
        A.WA = C.WA;
        A.U0 = (int64_t)(m + 64 + 63) / 64 * 64;
        A.m = m;
        A.ml = C.sym ? m - 1 : m + 1;
        A.dbudget = C.dbudget;
        A.drows = C.drows;
        A.witness_max = C.witness_max;
        const int w = (ctl->wit0 + (j - ctl->j0)) & 1;
        A.wsel = w;
        A.rsel = w ^ 1;
        A.some_off = halted ? 1 : 0;
        A.dk = (float)(delta_ed2(m) / (2.0 * m));
        A.off = halted ? 1 : 0;
        RefineArgs &R = C.par->R;
        const int64_t end = (halted ? 0 : (int64_t)*s_nmax) + A.U0;
        refine_bounds(A.U0 + A.WA, C.max_chunk, end, C.rrounds, R.bound);
        R.min_chunk = C.min_chunk;
        R.order = C.rorder;
        R.max_chunk = C.max_chunk;
        R.items = C.items;
        ctl->j = j;
        ctl->m = halted ? 0 : m;
        ctl->m2 = (halted || C.shift) ? 0 : m + C.step;
    }
}

// Shifted range, rows: each previous discord i of series q (one wave) at rows i, i-1 ..
// i-shift inside the series at this length, as Engine::fast asks for them, in group slots of
// nu rows (spare rows of a group copy its first), their result slots cleared.
__global__ __launch_bounds__(64) void k_chain_rreq(ChainArgs C)
{
    __shared__ int64_t s_row[CHAIN_KMAX * (EXACT_EXT + 1)];
    __shared__ int s_nrq;
    const int q = blockIdx.x, t = threadIdx.x;
    ChainCtl *ctl = C.ctl;
    ChainSer &cs = C.cs[q];
    const int m = ctl->m;  // 0: the loop has stopped
    const int G = C.gr, nu = C.nur;
    if (t == 0) {
        int nrq = 0;
        if (m > 0) {
            const int j = ctl->j;
            const ChainRes *R =
                j > ctl->j0 ? &C.res[(size_t)(j - 1 - ctl->j0) * C.nser + q] : nullptr;
            const int64_t *prev = R ? R->i : cs.prev;
            int np = R ? R->n : cs.nprev;
            np = np < 0 ? 0 : (np > CHAIN_KMAX ? CHAIN_KMAX : np);
            const int64_t Nq = C.ser[q].n - m + 1;
            for (int u = 0; u < np; u++)
                for (int d = 0; d <= C.shift && nrq < G * nu; d++) {
                    const int64_t row = prev[u] - d;
                    if (row < 0 || row >= Nq) continue;
                    s_row[nrq] = row;
                    cs.rq_u[nrq] = (unsigned char)u;
                    nrq++;
                }
        }
        cs.nrq = nrq;
        s_nrq = nrq;
    }
    __syncthreads();
    const int nrq = s_nrq;
    for (int g = 0; g < G; g++) {
        const int c = nrq - g * nu;
        const int cg = c < 0 ? 0 : (c > nu ? nu : c);
        if (t == 0) C.cntr[q * G + g] = cg;
        if (t < nu && cg > 0) C.candr[((size_t)q * G + g) * nu + t] = s_row[g * nu + (t < cg ? t : 0)];
        for (int e = t; e < cg * EXACT_SLOTS; e += 64)
            C.keyr[((size_t)q * G + g) * nu * EXACT_SLOTS + e] = ~0ull;
    }
}

// Shifted range, value (one wave per series): per previous discord the largest of its rows'
// exact distances, the minimum of those over the discords, times 1 - 1e-6 (Engine::fast's
// rule); the series' pass table entry.  A series without k values switches itself off and
// stops the loop at this length.
__global__ __launch_bounds__(64) void k_chain_range(ChainArgs C)
{
    __shared__ double s_best[CHAIN_KMAX];
    __shared__ int s_any[CHAIN_KMAX];
    const int q = blockIdx.x, t = threadIdx.x;
    ChainCtl *ctl = C.ctl;
    const int m = ctl->m;
    if (m <= 0) return;  // stopped: chain_begin switched the pass table off
    ChainSer &cs = C.cs[q];
    const int nrq = cs.nrq, G = C.gr, nu = C.nur;
    if (t < CHAIN_KMAX) {
        s_best[t] = -INFINITY;
        s_any[t] = 0;
    }
    __syncthreads();
    for (int r = 0; r < nrq; r++) {
        const size_t idx = ((size_t)q * G + r / nu) * nu + r % nu;
        unsigned long long k = C.keyr[idx * EXACT_SLOTS + t];
        for (int w = 32; w >= 1; w >>= 1) {
            const unsigned long long o = __shfl_xor(k, w, 64);
            k = o < k ? o : k;
        }
        if (t == 0) {
            const double v = ord_val64(k);
            const int u = cs.rq_u[r];
            s_any[u] = 1;
            if (v < 3.0e38) s_best[u] = fmax(s_best[u], v);
        }
    }
    __syncthreads();
    if (t == 0) {
        double mn = INFINITY;
        int cnt = 0;
        for (int u = 0; u < CHAIN_KMAX; u++)
            if (s_any[u]) {
                mn = fmin(mn, s_best[u] > -INFINITY ? s_best[u] : INFINITY);
                cnt++;
            }
        const bool ok = cnt >= C.k && mn < INFINITY && mn > 0.0;
        const double r2 = ok ? mn * (1.0 - 1e-6) : 0.0;
        cs.r2 = r2;
        SerDev S = C.ser[q];
        S.N = ok ? S.n - m + 1 : 0;
        S.thr = (float)(1.0 - r2 / (2.0 * m));
        C.pser[q] = S;
        if (!ok) atomicCAS(&ctl->halt, 0, ctl->j + 1);
    }
}

// Close of length j (m = cm, 0 if the loop had stopped) and start of the next, by one
// workgroup after the length's final k_chain_topk: the pass's statistics, when it ran, then
// chain_begin.
__device__ __forceinline__ void chain_close(const ChainArgs &C, const int j, const int cm,
                                            int *s_bad, unsigned long long *s_nmax)
{
    ChainCtl *ctl = C.ctl;
    const int t = threadIdx.x;
    if (cm > 0 && t < 64) {
        const PassCtl *P = C.pctl;
        unsigned long long a0 = P->pairs[0][t], a1 = P->pairs[1][t], a2 = P->live[0][t],
                           a3 = P->live[1][t];
        for (int w = 32; w >= 1; w >>= 1) {
            a0 += __shfl_xor(a0, w, 64);
            a1 += __shfl_xor(a1, w, 64);
            a2 += __shfl_xor(a2, w, 64);
            a3 += __shfl_xor(a3, w, 64);
        }
        if (t == 0) {
            ctl->stats[0] += 1;
            ctl->stats[1] += (unsigned long long)P->survivors;
            ctl->stats[3] += a0;
            ctl->stats[4] += a1;
            ctl->stats[5] += a2;
            ctl->stats[6] += a3;
        }
    }
    __syncthreads();  // every thread has read ctl->j and ctl->m; the top-k's writes are seen
    chain_begin(C, j + 1, s_bad, s_nmax);
}

// chain_close as a kernel of its own (several series: their final k_chain_topk workgroups
// must all have finished).
__global__ __launch_bounds__(1024) void k_chain_next(ChainArgs C)
{
    __shared__ int s_bad;
    __shared__ unsigned long long s_nmax;
    const int j = C.ctl->j, cm = C.ctl->m;
    chain_close(C, j, cm, &s_bad, &s_nmax);
}

// A candidate of the device top-k: its upper bound hi, local row i and position p in the
// workgroup's list; the larger hi ranks first, ties by lower row (TopK::advance's T).
struct TkKey {
    double h;
    int i;
    int p;  // -1: none
};
__device__ __forceinline__ bool tk_better(const TkKey &a, const TkKey &b)
{
    if (a.p < 0) return false;
    if (b.p < 0) return true;
    return a.h > b.h || (a.h == b.h && a.i < b.i);
}
// Best key over the wave (k_chain_topk is one wave); every lane gets it.
__device__ __forceinline__ TkKey tk_reduce(TkKey k)
{
    for (int o = 32; o >= 1; o >>= 1) {
        TkKey b;
        b.h = __shfl_xor(k.h, o, 64);
        b.i = __shfl_xor(k.i, o, 64);
        b.p = __shfl_xor(k.p, o, 64);
        if (tk_better(b, k)) k = b;
    }
    return k;
}
__device__ __forceinline__ int wave_sum_i(int v)
{
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Exact greedy top-k of series q (one wave per series) over its survivors, as
// TopK::advance: every survivor holds [lo, hi] (its single-precision bounds widened by
// conv_ed2, or its exact value once refined); T = argmax hi; the exact maximum lies in
// U = {hi >= lo_T (less the tie tolerance)}; once U is exact its best (ranks_before: TIE_REL,
// then lower row) is accepted if above r2 and its trivial matches dropped.  The greedy is
// replayed from the exact values in every round (it depends on nothing else), so a round only
// adds values: round 0 marks the survivors unrefined, rounds > 0 first fold the previous
// round's k_exact_chain results.  A round that needs values asks for up to nu of U's unrefined
// members (highest bound first), filled up to min(spec, nu) with the candidates most likely to
// be accepted next (descending bound, apart by m from T and from each other), as
// TopK::advance.  Round `rounds` only finishes: a series that still needs values is left
// unfinished (the host runs the length).
__global__ __launch_bounds__(64) void k_chain_topk(ChainArgs C, int round)
{
    // the series' survivors, staged once per round: survivor index, local row, bounds (or the
    // exact value in both), the exact value (NaN: not refined), dropped flag
    __shared__ int s_p[CHAIN_SCAP];
    __shared__ int s_row[CHAIN_SCAP];  // local rows (a call holds fewer than 2^31 rows)
    __shared__ double s_lo[CHAIN_SCAP], s_hi[CHAIN_SCAP], s_v[CHAIN_SCAP];
    __shared__ unsigned char s_gone[CHAIN_SCAP];
    __shared__ int s_n, s_sel[CHAIN_NU], s_anchor[CHAIN_NU + 1], s_acc[CHAIN_KMAX];
    __shared__ int s_cb_bad;  // chain_close's
    __shared__ unsigned long long s_cb_nmax;
    const int q = blockIdx.x, t = threadIdx.x;
    ChainCtl *ctl = C.ctl;
    ChainSer &cs = C.cs[q];
    const int nu = C.nu;
    // Everything the round reads from global memory that depends on nothing else, issued
    // together (a few dependent round trips otherwise make this small kernel slow): the
    // loop's and the series' state, the previous round's result slots (lane t: slot t of
    // every candidate), and the first 64 survivors with their exact values.
    const int j = ctl->j, j0 = ctl->j0, cm = ctl->m, halt = ctl->halt;
    const int status = cs.status, pnreq = cs.nreq;
    const double r2 = cs.r2;
    const SerDev S = C.ser[q];
    const int64_t nsv = C.pctl->survivors;
    unsigned long long ka[CHAIN_NU], kb[CHAIN_NU];
    int rq[CHAIN_NU];
#pragma unroll
    for (int u = 0; u < CHAIN_NU; u++) {
        const size_t o = ((size_t)q * nu + (u < nu ? u : 0)) * EXACT_SLOTS + t;
        ka[u] = C.outkey[o];
        kb[u] = C.outkey2[o];
        rq[u] = cs.req_p[u];
    }
    Surv sv0{};
    double xv0 = 0.0;
    if (t < C.scap) {  // scap <= the buffers' rows
        sv0 = C.surv[t];
        xv0 = C.xv[t];
    }
    const int m = C.minL + j * C.step;
    // lane 0's reading for the whole wave: another series may stop the loop meanwhile
    const int run = __shfl(cm > 0 && halt == 0 && status == 0 ? 1 : 0, 0, 64);
    int nreq = 0;
    if (run) {
        const int64_t roff = S.roff, rend = roff + series_rows(S.n, C.minL);
        // fold the previous round's results: the minimum over the slots
        const int pr = round > 0 ? pnreq : 0;
        double fold[CHAIN_NU];
#pragma unroll
        for (int u = 0; u < CHAIN_NU; u++) {
            for (int w = 32; w >= 1; w >>= 1) {
                const unsigned long long a2 = __shfl_xor(ka[u], w, 64);
                const unsigned long long b2 = __shfl_xor(kb[u], w, 64);
                ka[u] = a2 < ka[u] ? a2 : ka[u];
                kb[u] = b2 < kb[u] ? b2 : kb[u];
            }
            fold[u] = ord_val64(ka[u]);
            if (t == 0 && u < pr) {
                C.xv[rq[u]] = fold[u];
                C.xv2[rq[u]] = ord_val64(kb[u]);
            }
        }
        if (t == 0) s_n = 0;
        __syncthreads();
        const int64_t ns = nsv < C.scap ? nsv : C.scap;
        for (int64_t p = t; p < ns; p += 64) {
            const Surv sv = p == t ? sv0 : C.surv[p];
            if (sv.g < roff || sv.g >= rend) continue;
            const int e = atomicAdd(&s_n, 1);
            if (e >= CHAIN_SCAP) continue;
            double v = __longlong_as_double(0x7ff8000000000000ll);  // NaN: not refined
            if (round == 0) {
                C.xv[p] = v;
            } else {
                v = p == t ? xv0 : C.xv[p];
#pragma unroll
                for (int u = 0; u < CHAIN_NU; u++)  // refined by the previous round: folded here
                    if (u < pr && rq[u] == (int)p) v = fold[u];
            }
            s_p[e] = (int)p;
            s_row[e] = (int)(sv.g - roff);
            s_v[e] = v;
            s_gone[e] = 0;
            if (!isnan(v)) {
                s_lo[e] = v;
                s_hi[e] = v;
            } else if (sv.d2 >= 3.0e38f) {
                s_lo[e] = INFINITY;
                s_hi[e] = INFINITY;
            } else {
                s_lo[e] = (double)sv.d2lo - conv_ed2(m, sv.d2lo);
                s_hi[e] = (double)sv.d2hi + conv_ed2(m, sv.d2hi);
            }
        }
        __syncthreads();
        const int n = s_n < CHAIN_SCAP ? s_n : CHAIN_SCAP;
        // status: 0 searching, -1 no discord above r2 left, 3 unfinished
        int status = (nsv > C.scap || s_n > CHAIN_SCAP) ? 3 : 0;
        int nacc = 0;
        while (status == 0 && nacc < C.k) {
            // T = argmax hi over the candidates left
            TkKey best{-INFINITY, 0, -1};
            for (int e = t; e < n; e += 64)
                if (!s_gone[e]) {
                    const TkKey c{s_hi[e], s_row[e], e};
                    if (tk_better(c, best)) best = c;
                }
            const TkKey T = tk_reduce(best);
            if (T.p < 0) break;
            const double loT = s_lo[T.p];
            const double loTt = loT - 2.0 * TIE_REL * fabs(loT);
            // U and its unrefined members
            int cu = 0, cn = 0;
            for (int e = t; e < n; e += 64)
                if (!s_gone[e] && s_hi[e] >= loTt) {
                    cu++;
                    cn += isnan(s_v[e]) ? 1 : 0;
                }
            cu = wave_sum_i(cu);
            cn = wave_sum_i(cn);
            const bool tonly = cu == 0;  // only when distances are not ordered (NaN)
            const int nneed = tonly ? (isnan(s_v[T.p]) ? 1 : 0) : cn;
            if (nneed > 0) {
                if (round == C.rounds) {
                    status = 3;
                    break;
                }
                // the requests: unrefined members of U by descending bound (T first), then
                // speculative ones (apart by m from T and from each other), up to
                // min(spec, nu) in all
                int nanchor = 1;
                if (t == 0) s_anchor[0] = T.i;
                const int cap = C.spec < nu ? C.spec : nu;
                const int take = nneed < nu ? nneed : nu;
                const int want = take < cap ? cap : take;
                __syncthreads();
                for (int r = 0; r < want; r++) {
                    const bool spec = r >= take;
                    // TopK::advance speculates below spec only, and not for the last acceptance
                    if (spec && (nneed >= C.spec || nacc + 1 >= C.k)) break;
                    TkKey b{-INFINITY, 0, -1};
                    for (int e = t; e < n; e += 64) {
                        if (s_gone[e] || !isnan(s_v[e])) continue;
                        bool used = false;
                        for (int u = 0; u < r; u++) used |= s_sel[u] == e;
                        if (used) continue;
                        if (!spec) {
                            if (tonly ? e != T.p : !(s_hi[e] >= loTt)) continue;
                        } else {
                            bool apart = true;
                            for (int a2 = 0; a2 < nanchor; a2++) {
                                const int d = s_row[e] - s_anchor[a2];
                                apart &= (d < 0 ? -d : d) >= m;
                            }
                            if (!apart) continue;
                        }
                        const TkKey c{s_hi[e], s_row[e], e};
                        if (tk_better(c, b)) b = c;
                    }
                    const TkKey w = tk_reduce(b);
                    if (w.p < 0) break;
                    if (t == 0) {
                        s_sel[r] = w.p;
                        if (spec) s_anchor[nanchor] = w.i;
                    }
                    if (spec) nanchor++;
                    nreq = r + 1;
                    __syncthreads();
                }
                break;
            }
            // U is exact: accept its best if above r2 (ranks_before), drop its trivial matches.
            // Lane-wise best under ranks_before, then the lanes' bests in lane order: the
            // tolerance makes the order matter only between values within TIE_REL.
            int B = -1;
            double vb = 0.0;
            int ib = 0;
            for (int e = t; e < n; e += 64) {
                if (s_gone[e]) continue;
                if (tonly ? e != T.p : !(s_hi[e] >= loTt)) continue;
                const double v = s_v[e];
                const int i = s_row[e];
                bool before = B < 0;
                if (!before) {
                    const double tol = TIE_REL * fmax(fabs(v), fabs(vb));
                    before = v > vb + tol ? true : (vb > v + tol ? false : i < ib);
                }
                if (before) {
                    B = e;
                    vb = v;
                    ib = i;
                }
            }
            // only the lanes holding a member of U take part (usually a few): their bits
            unsigned long long act = __ballot(B >= 0) & ~1ull;
            while (act) {
                const int l = __ffsll((long long)act) - 1;
                act &= act - 1;
                const int Bl = __shfl(B, l, 64);
                const double vl = __shfl(vb, l, 64);
                const int il = __shfl(ib, l, 64);
                if (t != 0) continue;
                bool before = B < 0;
                if (!before) {
                    const double tol = TIE_REL * fmax(fabs(vl), fabs(vb));
                    before = vl > vb + tol ? true : (vb > vl + tol ? false : il < ib);
                }
                if (before) {
                    B = Bl;
                    vb = vl;
                    ib = il;
                }
            }
            B = __shfl(B, 0, 64);
            vb = __shfl(vb, 0, 64);
            ib = __shfl(ib, 0, 64);
            if (!(vb > r2)) {
                status = -1;  // no more discords above r2
                break;
            }
            if (t == 0) s_acc[nacc] = B;
            nacc++;
            for (int e = t; e < n; e += 64) {
                const int d = s_row[e] - ib;
                if ((d < 0 ? -d : d) < m) s_gone[e] = 1;
            }
            __syncthreads();
        }
        __syncthreads();
        // waiting for values (nreq > 0), finished (k discords, or fewer with none above r2
        // left), short of k (2) or unfinished (3)
        int st = status == 3 ? 3 : (nreq > 0 ? 0 : ((nacc >= C.k || r2 <= 0.0) ? 1 : 2));
        if (t == 0) {
            if (st == 1) {
                ChainRes &R = C.res[(size_t)(j - j0) * C.nser + q];
                R.n = nacc;
                R.r2 = r2;
                for (int a = 0; a < nacc; a++) {
                    const int e = s_acc[a];
                    R.i[a] = s_row[e];
                    R.v[a] = s_v[e];
                    R.v2[a] = C.xv2[s_p[e]];
                }
            }
            if (st >= 2) atomicCAS(&ctl->halt, 0, j + 1);
            cs.status = st;
            cs.nreq = st == 0 ? nreq : 0;
            for (int u = 0; u < nreq; u++) {
                cs.req_p[u] = s_p[s_sel[u]];
                C.cand[(size_t)q * nu + u] = s_row[s_sel[u]];
                C.ub[(size_t)q * nu + u] = s_hi[s_sel[u]];
            }
            for (int u = nreq; u < nu && nreq > 0; u++) {  // spare rows of the group: copies
                C.cand[(size_t)q * nu + u] = s_row[s_sel[0]];
                C.ub[(size_t)q * nu + u] = s_hi[s_sel[0]];
            }
            if (nreq > 0) atomicAdd(&ctl->stats[2], (unsigned long long)nreq);
        }
        if (st != 0) nreq = 0;
        for (int e = t; e < nreq * EXACT_SLOTS; e += 64) {
            C.outkey[(size_t)q * nu * EXACT_SLOTS + e] = ~0ull;
            C.outkey2[(size_t)q * nu * EXACT_SLOTS + e] = ~0ull;
            C.runub[(size_t)q * nu * EXACT_SLOTS + e] = 0x7fffffff;
        }
    }
    if (t == 0) C.cnt[q] = nreq;
    // one series: its final round is the length's last work, so it starts the next length
    if (round == C.rounds && C.nser == 1) chain_close(C, j, cm, &s_cb_bad, &s_cb_nmax);
}

template <bool DUAL>
static void launch_exact_chain_t(const ExactArgs &E, int nu, unsigned grid, size_t lds,
                                 hipStream_t st)
{
    switch (nu) {
    case 1: hipLaunchKernelGGL((k_exact_chain<1, DUAL>), dim3(grid), dim3(256), lds, st, E); break;
    case 2: hipLaunchKernelGGL((k_exact_chain<2, DUAL>), dim3(grid), dim3(256), lds, st, E); break;
    case 3: hipLaunchKernelGGL((k_exact_chain<3, DUAL>), dim3(grid), dim3(256), lds, st, E); break;
    default: hipLaunchKernelGGL((k_exact_chain<4, DUAL>), dim3(grid), dim3(256), lds, st, E); break;
    }
}

void launch_exact_chain(const ExactArgs &E, int nu, bool dual, unsigned grid, size_t lds,
                        hipStream_t st)
{
    if (dual)
        launch_exact_chain_t<true>(E, nu, grid, lds, st);
    else
        launch_exact_chain_t<false>(E, nu, grid, lds, st);
}

template <bool DUAL>
static void launch_exact_t(const ExactArgs &E, int nu, unsigned grid, size_t lds, hipStream_t st)
{
    switch (nu) {
    case 1: hipLaunchKernelGGL((k_exact<1, DUAL>), dim3(grid), dim3(256), lds, st, E); break;
    case 2: hipLaunchKernelGGL((k_exact<2, DUAL>), dim3(grid), dim3(256), lds, st, E); break;
    case 3: hipLaunchKernelGGL((k_exact<3, DUAL>), dim3(grid), dim3(256), lds, st, E); break;
    case 4: hipLaunchKernelGGL((k_exact<4, DUAL>), dim3(grid), dim3(256), lds, st, E); break;
    case 5: hipLaunchKernelGGL((k_exact<5, DUAL>), dim3(grid), dim3(256), lds, st, E); break;
    case 6: hipLaunchKernelGGL((k_exact<6, DUAL>), dim3(grid), dim3(256), lds, st, E); break;
    case 7: hipLaunchKernelGGL((k_exact<7, DUAL>), dim3(grid), dim3(256), lds, st, E); break;
    default: hipLaunchKernelGGL((k_exact<8, DUAL>), dim3(grid), dim3(256), lds, st, E); break;
    }
}

void launch_exact(const ExactArgs &E, int nu, unsigned grid, size_t lds, hipStream_t st)
{
    // One candidate at two lengths compiles to 256 registers (one wave per SIMD); the group of
    // two (the spare row a copy of the first, its result ignored) runs three waves per SIMD.
    if (E.m2 && nu == 1 && exact_lds_bytes(E.m, 2) <= 65536 - 1024) {
        nu = 2;
        lds = exact_lds_bytes(E.m, 2);
    }
    if (E.m2)
        launch_exact_t<true>(E, nu, grid, lds, st);
    else
        launch_exact_t<false>(E, nu, grid, lds, st);
}

}  // namespace paddi
