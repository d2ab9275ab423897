// Host side of the discord engine and its C ABI (include/paddi.h).
//
// Per-(m, r) pipeline (the single-GPU slice of the reference's missing
// do_PALMAD_one_length, called at src/main.cu:144,176,219):
//   k_derive        window statistics of length m from the series' prefix sums
//   k_scan_select   64-row blocks scan a bounded window to their right (PD3 phase 1)
//   k_list_blocks   blocks with live rows
//   k_scan_refine   live blocks x column chunks over all remaining columns (PD3 phase 2)
//   k_survivors     surviving rows with their single-precision NN distance
//   host            exact top-k: interval ordering + k_exact (double) where it matters
//
// Exactness.  The scans prune at ED^2 <= r2 - delta, where delta bounds the error of the
// single-precision distance, so the surviving set contains every row whose exact NN distance
// exceeds r2, and each survivor's single-precision distance a_i is within delta of the exact
// one.  select_topk() ranks survivors by exact distance, computing the exact (double)
// distance only for candidates whose rank the +-delta intervals leave undecided, and for
// every accepted discord (whose reported distance is then the double-precision value).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/paddi.h"
#include "engine.h"
#include "kernels.h"

using namespace paddi;

namespace paddi {

const char *status_str(paddi_status s)
{
    switch (s) {
    case PADDI_OK: return "ok";
    case PADDI_ERR_INVALID: return "invalid argument";
    case PADDI_ERR_DEVICE: return "HIP runtime error";
    case PADDI_ERR_NOMEM: return "device allocation failed";
    case PADDI_ERR_CAPACITY: return "output capacity too small";
    case PADDI_ERR_UNSUPPORTED: return "unsupported request";
    default: return "unknown status";
    }
}

struct HipError {
    paddi_status st;
    std::string what;
};

#define HIPC(expr)                                                                     \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            throw HipError{e_ == hipErrorOutOfMemory ? PADDI_ERR_NOMEM : PADDI_ERR_DEVICE, \
                           std::string(#expr) + ": " + hipGetErrorString(e_)};         \
    } while (0)

void DevBuf::reserve(size_t nbytes)
{
    if (nbytes <= bytes) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    HIPC(hipMalloc(&p, nbytes));
    bytes = nbytes;
}
void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

double delta_ed2(int m)
{
    // Bound on |single - exact| squared distance: correlation error of a seed (m-term FP32
    // dot product on locally centred data) plus up to 63 recurrence steps, with a 4x
    // safety factor over the worst case measured on the test workloads.
    const double rho_err = 4.0e-5 + 2.0e-7 * m;
    return 2.0 * m * rho_err;
}

Engine::Engine(int device) : device_(device)
{
    HIPC(hipSetDevice(device));
    HIPC(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    for (auto &e : ev_) HIPC(hipEventCreate(&e));
    HIPC(hipHostMalloc(&pin_, 64, hipHostMallocDefault));
}

Engine::~Engine()
{
    (void)hipSetDevice(device_);
    for (DevBuf *b : {&colv_, &st64_, &alive_, &best_, &blist_, &cnt_, &sidx_, &sd2_, &cand_,
                      &okey_})
        b->release();
    for (auto &e : ev_) (void)hipEventDestroy(e);
    if (pin_) (void)hipHostFree(pin_);
    (void)hipStreamDestroy(stream_);
}

Series *Engine::upload(const float *T, int64_t n)
{
    HIPC(hipSetDevice(device_));
    Series *s = new Series();
    s->n = n;
    s->nb = (n + PB - 1) / PB + 1;
    try {
        s->x.reserve((size_t)(n + XPAD) * sizeof(float));
        s->P.reserve((size_t)s->nb * (PB + 1) * sizeof(double2));
        s->refs.reserve((size_t)s->nb * sizeof(double));
        HIPC(hipMemsetAsync(s->x.p, 0, (size_t)(n + XPAD) * sizeof(float), stream_));
        HIPC(hipMemcpyAsync(s->x.p, T, (size_t)n * sizeof(float), hipMemcpyHostToDevice,
                            stream_));
        HIPC(hipStreamSynchronize(stream_));
    } catch (...) {
        delete s;
        throw;
    }
    return s;
}

void Engine::prefix(Series &s)
{
    if (s.prefix_ready) return;
    hipLaunchKernelGGL(k_prefix, dim3((unsigned)s.nb), dim3(256), 0, stream_, s.xp(), s.n,
                       (double2 *)s.P.p, (double *)s.refs.p);
    HIPC(hipGetLastError());
    s.prefix_ready = true;
}

void Engine::derive(Series &s, int m)
{
    const int64_t N = s.n - m + 1;
    const int64_t Npad = ((N + 63) / 64) * 64 + 128;
    colv_.reserve((size_t)Npad * sizeof(float4));
    st64_.reserve((size_t)Npad * sizeof(double2));
    hipLaunchKernelGGL(k_derive, dim3((unsigned)((Npad + 255) / 256)), dim3(256), 0, stream_,
                       s.xp(), (const double2 *)s.P.p, (const double *)s.refs.p, N, Npad, m,
                       (float4 *)colv_.p, (double2 *)st64_.p);
    HIPC(hipGetLastError());
}

static inline unsigned grid1(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// One (m, r2) pass: returns the survivors (index, single-precision squared NN distance).
void Engine::pass(Series &s, int m, double r2, std::vector<Cand> &out, paddi_stats *st)
{
    const int64_t N = s.n - m + 1;
    const int64_t nwords = (N + 63) / 64;
    alive_.reserve((size_t)nwords * 8);
    best_.reserve((size_t)N * 4);
    blist_.reserve((size_t)nwords * 4);
    cnt_.reserve(16);
    sidx_.reserve((size_t)N * 8);
    sd2_.reserve((size_t)N * 4);
    const double delta = delta_ed2(m);
    const float thr_rho = (float)(1.0 - (r2 - delta) / (2.0 * m));
    const int64_t WA = select_window_;
    const int64_t CH = refine_chunk_;
    const size_t lds = scan_lds_bytes(m);

    HIPC(hipEventRecord(ev_[0], stream_));
    hipLaunchKernelGGL(k_fill_alive, dim3(grid1(nwords, 256)), dim3(256), 0, stream_,
                       (uint64_t *)alive_.p, N, nwords);
    hipLaunchKernelGGL(k_fill_int, dim3(grid1(N, 256)), dim3(256), 0, stream_, (int *)best_.p, N,
                       (int)0x807fffff /* ord_key(-inf) */);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, stream_, (int *)cnt_.p, 4, 0);
    hipLaunchKernelGGL(k_scan_select, dim3((unsigned)nwords), dim3(64), lds, stream_, s.xp(),
                       (const float4 *)colv_.p, (uint64_t *)alive_.p, (int *)best_.p, N, m,
                       thr_rho, WA);
    HIPC(hipGetLastError());
    hipLaunchKernelGGL(k_list_blocks, dim3(grid1(nwords, 256)), dim3(256), 0, stream_,
                       (const uint64_t *)alive_.p, nwords, (int *)blist_.p, (int *)cnt_.p);
    HIPC(hipMemcpyAsync(pin_, cnt_.p, 4, hipMemcpyDeviceToHost, stream_));
    HIPC(hipEventRecord(ev_[1], stream_));
    HIPC(hipStreamSynchronize(stream_));
    const int nblk = ((int *)pin_)[0];
    const int nchunks = (int)((N + CH - 1) / CH);
    if (nblk > 0) {
        hipLaunchKernelGGL(k_scan_refine, dim3((unsigned)((int64_t)nblk * nchunks)), dim3(64), lds,
                           stream_, s.xp(), (const float4 *)colv_.p, (uint64_t *)alive_.p,
                           (int *)best_.p, (const int *)blist_.p, N, m, thr_rho, WA, CH, nchunks);
        HIPC(hipGetLastError());
    }
    hipLaunchKernelGGL(k_survivors, dim3(grid1(N, 256)), dim3(256), 0, stream_,
                       (const uint64_t *)alive_.p, (const int *)best_.p, N, m, (int *)cnt_.p + 1,
                       (int64_t *)sidx_.p, (float *)sd2_.p, N);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync((int *)pin_ + 1, (int *)cnt_.p + 1, 4, hipMemcpyDeviceToHost, stream_));
    HIPC(hipEventRecord(ev_[2], stream_));
    HIPC(hipStreamSynchronize(stream_));
    const int ns = ((int *)pin_)[1];
    out.resize((size_t)ns);
    if (ns > 0) {
        std::vector<int64_t> ix((size_t)ns);
        std::vector<float> dd((size_t)ns);
        HIPC(hipMemcpyAsync(ix.data(), sidx_.p, (size_t)ns * 8, hipMemcpyDeviceToHost, stream_));
        HIPC(hipMemcpyAsync(dd.data(), sd2_.p, (size_t)ns * 4, hipMemcpyDeviceToHost, stream_));
        HIPC(hipStreamSynchronize(stream_));
        for (int q = 0; q < ns; q++) out[(size_t)q] = Cand{ix[(size_t)q], dd[(size_t)q], 0.0, false};
    }
    if (st) {
        float a = 0.f, b = 0.f;
        HIPC(hipEventElapsedTime(&a, ev_[0], ev_[1]));
        HIPC(hipEventElapsedTime(&b, ev_[1], ev_[2]));
        st->kernel_ms[1] += a;
        st->kernel_ms[2] += b;
        st->attempts += 1;
        st->survivors += ns;
    }
}

// Exact (double) NN distance of the given candidates.
void Engine::exact(Series &s, int m, const std::vector<int64_t> &idx, std::vector<double> &out,
                   paddi_stats *st)
{
    const int64_t N = s.n - m + 1;
    out.assign(idx.size(), 0.0);
    const int umax = exact_max_u(m);
    cand_.reserve(MAXU * 8);
    okey_.reserve(MAXU * 8);
    for (size_t b0 = 0; b0 < idx.size(); b0 += (size_t)umax) {
        const int nU = (int)std::min<size_t>((size_t)umax, idx.size() - b0);
        HIPC(hipMemcpyAsync(cand_.p, idx.data() + b0, (size_t)nU * 8, hipMemcpyHostToDevice,
                            stream_));
        HIPC(hipMemsetAsync(okey_.p, 0xff, (size_t)nU * 8, stream_));
        HIPC(hipEventRecord(ev_[0], stream_));
        hipLaunchKernelGGL(k_exact, dim3(grid1(N, 256)), dim3(256), exact_lds_bytes(m, nU),
                           stream_, s.xp(), (const double2 *)st64_.p, N, m,
                           (const int64_t *)cand_.p, nU, (unsigned long long *)okey_.p);
        HIPC(hipGetLastError());
        HIPC(hipEventRecord(ev_[1], stream_));
        unsigned long long keys[MAXU];
        HIPC(hipMemcpyAsync(keys, okey_.p, (size_t)nU * 8, hipMemcpyDeviceToHost, stream_));
        HIPC(hipStreamSynchronize(stream_));
        for (int u = 0; u < nU; u++) {
            unsigned long long k = keys[u];
            unsigned long long bits = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
            double d;
            std::memcpy(&d, &bits, 8);
            out[b0 + (size_t)u] = d;
        }
        if (st) {
            float a = 0.f;
            HIPC(hipEventElapsedTime(&a, ev_[0], ev_[1]));
            st->kernel_ms[3] += a;
            st->exact_refines += nU;
        }
    }
}

// Exact greedy top-k over the survivors (see file header).  Returns the accepted discords in
// rank order; stops early when the next exact distance is <= r2.
void Engine::select_topk(Series &s, int m, double r2, int k, std::vector<Cand> &cands,
                         std::vector<Cand> &acc, paddi_stats *st)
{
    const double delta = delta_ed2(m);
    acc.clear();
    std::vector<char> gone(cands.size(), 0);
    auto lo = [&](const Cand &c) { return c.exact ? c.v : (double)c.a - delta; };
    auto hi = [&](const Cand &c) { return c.exact ? c.v : (double)c.a + delta; };
    auto is_inf = [&](const Cand &c) { return !c.exact && c.a >= 3.0e38f; };
    while ((int)acc.size() < k) {
        // top by upper bound (ties: smaller index)
        int64_t T = -1;
        for (size_t q = 0; q < cands.size(); q++) {
            if (gone[q]) continue;
            if (T < 0) { T = (int64_t)q; continue; }
            const Cand &a = cands[q], &b = cands[(size_t)T];
            const double ha = is_inf(a) ? INFINITY : hi(a), hb = is_inf(b) ? INFINITY : hi(b);
            if (ha > hb || (ha == hb && a.i < b.i)) T = (int64_t)q;
        }
        if (T < 0) break;
        const double loT = is_inf(cands[(size_t)T]) ? INFINITY : lo(cands[(size_t)T]);
        std::vector<size_t> U;
        for (size_t q = 0; q < cands.size(); q++) {
            if (gone[q]) continue;
            const double h = is_inf(cands[q]) ? INFINITY : hi(cands[q]);
            if (h >= loT) U.push_back(q);
        }
        std::vector<int64_t> need;
        std::vector<size_t> needq;
        for (size_t q : U)
            if (!cands[q].exact) { need.push_back(cands[q].i); needq.push_back(q); }
        if (!need.empty()) {
            std::vector<double> v;
            exact(s, m, need, v, st);
            for (size_t u = 0; u < need.size(); u++) {
                cands[needq[u]].exact = true;
                cands[needq[u]].v = v[u];
            }
            continue;
        }
        size_t B = U[0];
        for (size_t q : U) {
            const Cand &a = cands[q], &b = cands[B];
            if (a.v > b.v || (a.v == b.v && a.i < b.i)) B = q;
        }
        if (!(cands[B].v > r2)) break;
        acc.push_back(cands[B]);
        const int64_t bi = cands[B].i;
        for (size_t q = 0; q < cands.size(); q++)
            if (!gone[q] && std::llabs(cands[q].i - bi) < m) gone[q] = 1;
    }
}

static double next_r_after_fail(int j, double r, double S)
{
    if (j == 0) return 0.5 * r;
    if (j < 5) return 0.99 * r;
    if (S > 0.0) return r - S;
    return 0.99 * r;
}

// PALMAD length loop with the r-schedule of src/main.cu:136-233 (guards as in the oracle).
void Engine::palmad(Series &s, int minL, int maxL, int step, int k, const paddi_opts &o,
                    std::vector<paddi_discord> &out, paddi_stats *st, int series_id)
{
    prefix(s);
    const int nlen = (maxL - minL) / step + 1;
    std::vector<double> dk((size_t)nlen, 0.0);
    std::vector<Cand> cands, acc;
    const int max_att = o.max_attempts > 0 ? o.max_attempts : 64;
    for (int j = 0; j < nlen; j++) {
        const int m = minL + j * step;
        derive(s, m);
        double r, S = 0.0;
        if (j == 0) {
            r = 2.0 * std::sqrt((double)minL);
        } else if (j < 5) {
            r = 0.99 * std::sqrt(dk[(size_t)j - 1]);
        } else {
            double M = 0.0;
            for (int q = j - 5; q < j; q++) M += std::sqrt(dk[(size_t)q]);
            M *= 0.2;
            for (int q = j - 5; q < j; q++) {
                const double e = std::sqrt(dk[(size_t)q]) - M;
                S += e * e;
            }
            S = std::sqrt(S * 0.2);
            r = M - 2.0 * S;
        }
        int att = 0;
        for (;;) {
            if (r < 0.0) r = 0.0;
            if (att >= max_att) r = 0.0;
            const double r2 = r * r;
            pass(s, m, r2, cands, st);
            select_topk(s, m, r2, k, cands, acc, st);
            att++;
            if ((int)acc.size() >= k || r == 0.0) {
                for (const Cand &c : acc)
                    out.push_back(paddi_discord{c.i, m, series_id,
                                                c.v >= 3.0e38 ? 3.4028235e38f : (float)c.v,
                                                (float)r2});
                dk[(size_t)j] = acc.empty() ? 0.0 : acc.back().v;
                break;
            }
            r = next_r_after_fail(j, r, S);
        }
    }
}

}  // namespace paddi

// ------------------------------------------------------------------------------------------
// C ABI

struct paddi_ctx {
    paddi::Engine *eng;
    std::string last_error;
};
struct paddi_series {
    paddi_ctx *ctx;
    paddi::Series *s;
};

template <class F>
static paddi_status guarded(paddi_ctx *ctx, F &&f)
{
    try {
        return f();
    } catch (const paddi::HipError &e) {
        if (ctx) ctx->last_error = e.what;
        return e.st;
    } catch (const std::bad_alloc &) {
        return PADDI_ERR_NOMEM;
    }
}

static bool valid_lengths(int64_t n, int32_t minL, int32_t maxL, int32_t step, int32_t k)
{
    return minL >= 4 && maxL >= minL && step >= 1 && k >= 1 && n >= (int64_t)maxL + 1;
}

extern "C" {

const char *paddi_status_str(paddi_status s) { return paddi::status_str(s); }
const char *paddi_version(void) { return "paddi-mi355x 0.1 (gfx950)"; }

paddi_status paddi_create(int32_t device, paddi_ctx **out)
{
    if (!out) return PADDI_ERR_INVALID;
    *out = nullptr;
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || device < 0 || device >= nd)
        return PADDI_ERR_DEVICE;
    paddi_ctx *c = new paddi_ctx{nullptr, {}};
    paddi_status st = guarded(c, [&] {
        c->eng = new paddi::Engine(device);
        return PADDI_OK;
    });
    if (st != PADDI_OK) {
        delete c;
        return st;
    }
    *out = c;
    return PADDI_OK;
}

void paddi_destroy(paddi_ctx *ctx)
{
    if (!ctx) return;
    delete ctx->eng;
    delete ctx;
}

paddi_status paddi_series_upload(paddi_ctx *ctx, const float *T, int64_t n, paddi_series **out)
{
    if (!ctx || !T || !out || n < 8) return PADDI_ERR_INVALID;
    return guarded(ctx, [&] {
        paddi::Series *s = ctx->eng->upload(T, n);
        *out = new paddi_series{ctx, s};
        return PADDI_OK;
    });
}

void paddi_series_free(paddi_series *s)
{
    if (!s) return;
    delete s->s;
    delete s;
}

paddi_status paddi_compute_statistics(paddi_ctx *ctx, const float *T, int64_t n, int32_t m,
                                      float *mean, float *std)
{
    if (!ctx || !T || !mean || !std || m < 2 || n < m) return PADDI_ERR_INVALID;
    if (m > paddi::PB) return PADDI_ERR_UNSUPPORTED;
    return guarded(ctx, [&] {
        paddi::Engine &E = *ctx->eng;
        std::unique_ptr<paddi::Series> s(E.upload(T, n));
        E.prefix(*s);
        E.derive(*s, m);
        const int64_t N = n - m + 1;
        std::vector<double2> st((size_t)N);
        E.copy_st64(st.data(), N);
        const double sm = std::sqrt((double)m);
        for (int64_t i = 0; i < N; i++) {
            mean[i] = (float)st[(size_t)i].x;
            std[i] = st[(size_t)i].y > 0.0 ? (float)(1.0 / (st[(size_t)i].y * sm)) : 0.0f;
        }
        return PADDI_OK;
    });
}

paddi_status paddi_range_discords(paddi_ctx *ctx, const float *T, int64_t n, int32_t m, float r2,
                                  int32_t *is_discord, float *nn_dist2)
{
    if (!ctx || !T || !is_discord || !nn_dist2 || m < 4 || n < (int64_t)m + 1 || !(r2 >= 0.f))
        return PADDI_ERR_INVALID;
    if (m > paddi::MAXM) return PADDI_ERR_UNSUPPORTED;
    return guarded(ctx, [&] {
        paddi::Engine &E = *ctx->eng;
        std::unique_ptr<paddi::Series> s(E.upload(T, n));
        E.prefix(*s);
        E.derive(*s, m);
        const int64_t N = n - m + 1;
        std::vector<paddi::Cand> cands;
        E.pass(*s, m, (double)r2, cands, nullptr);
        std::fill(is_discord, is_discord + N, 0);
        std::fill(nn_dist2, nn_dist2 + N, 3.4028235e38f);
        const double delta = paddi::delta_ed2(m);
        std::vector<int64_t> amb;
        for (const paddi::Cand &c : cands) {
            if (c.a >= 3.0e38f || (double)c.a > (double)r2 + delta) {
                is_discord[c.i] = 1;
                nn_dist2[c.i] = c.a;
            } else if ((double)c.a > (double)r2 - delta) {
                amb.push_back(c.i);
            }
        }
        if (!amb.empty()) {
            std::vector<double> v;
            E.exact(*s, m, amb, v, nullptr);
            for (size_t q = 0; q < amb.size(); q++)
                if (v[q] > (double)r2) {
                    is_discord[amb[q]] = 1;
                    nn_dist2[amb[q]] = (float)v[q];
                }
        }
        return PADDI_OK;
    });
}

paddi_status paddi_discords_series(paddi_ctx *ctx, paddi_series *const *series, int32_t nseries,
                                   int32_t minL, int32_t maxL, int32_t step, int32_t top_k,
                                   const paddi_opts *opts, paddi_discord *out, int64_t cap,
                                   int64_t *count, paddi_stats *stats)
{
    if (!ctx || !series || nseries <= 0 || !count || (!out && cap > 0)) return PADDI_ERR_INVALID;
    for (int32_t q = 0; q < nseries; q++)
        if (!series[q] || series[q]->ctx != ctx ||
            !valid_lengths(series[q]->s->n, minL, maxL, step, top_k))
            return PADDI_ERR_INVALID;
    if (maxL > paddi::MAXM) return PADDI_ERR_UNSUPPORTED;
    paddi_opts o{};
    if (opts) o = *opts;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    const auto t0 = std::chrono::steady_clock::now();
    paddi_status st = guarded(ctx, [&] {
        std::vector<paddi_discord> res;
        for (int32_t q = 0; q < nseries; q++)
            ctx->eng->run(*series[q]->s, minL, maxL, step, top_k, o, res, stats, q);
        *count = (int64_t)res.size();
        if ((int64_t)res.size() > cap) return PADDI_ERR_CAPACITY;
        std::copy(res.begin(), res.end(), out);
        return PADDI_OK;
    });
    if (stats)
        stats->seconds =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return st;
}

paddi_status paddi_discords_batch(paddi_ctx *ctx, const float *const *T, const int64_t *n,
                                  int32_t nseries, int32_t minL, int32_t maxL, int32_t step,
                                  int32_t top_k, const paddi_opts *opts, paddi_discord *out,
                                  int64_t cap, int64_t *count, paddi_stats *stats)
{
    if (!ctx || !T || !n || nseries <= 0 || !count) return PADDI_ERR_INVALID;
    for (int32_t q = 0; q < nseries; q++)
        if (!T[q] || !valid_lengths(n[q], minL, maxL, step, top_k)) return PADDI_ERR_INVALID;
    std::vector<paddi_series *> ss;
    paddi_status st = PADDI_OK;
    for (int32_t q = 0; q < nseries && st == PADDI_OK; q++) {
        paddi_series *p = nullptr;
        st = paddi_series_upload(ctx, T[q], n[q], &p);
        if (st == PADDI_OK) ss.push_back(p);
    }
    if (st == PADDI_OK)
        st = paddi_discords_series(ctx, ss.data(), nseries, minL, maxL, step, top_k, opts, out,
                                   cap, count, stats);
    for (paddi_series *p : ss) paddi_series_free(p);
    return st;
}

paddi_status paddi_discords(paddi_ctx *ctx, const float *T, int64_t n, int32_t minL, int32_t maxL,
                            int32_t step, int32_t top_k, const paddi_opts *opts, paddi_discord *out,
                            int64_t cap, int64_t *count, paddi_stats *stats)
{
    const float *const Ts[1] = {T};
    const int64_t ns[1] = {n};
    return paddi_discords_batch(ctx, Ts, ns, 1, minL, maxL, step, top_k, opts, out, cap, count,
                                stats);
}

}  // extern "C"
