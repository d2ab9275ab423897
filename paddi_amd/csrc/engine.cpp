// Host side of the discord engine and its C ABI (include/paddi.h).
//
// Per length: k_derive (window statistics of length m from the series' prefix sums), then
// per (m, r) pass over a batch of series (the single-GPU slice of the reference's missing
// do_PALMAD_one_length, called at src/main.cu:144,176,219), one host round trip:
//   k_pass_init     alive bits, row maxima, witness slots, counters
//   k_scan_select   64-row blocks scan their witness tiles and a bounded window to their right
//   k_list_blocks   blocks with live rows (before each refine round)
//   k_scan_refine   live blocks x column chunks, three rounds of growing distance
//   k_survivors     surviving rows with their single-precision NN distance
// then TopK: exact top-k by interval ordering + k_exact (double) where it matters.
//
// Exactness.  The scans prune at ED^2 <= r2 - delta_c, where delta_c bounds the error of the
// cell's single-precision distance: delta (the bound on data of uniform scale) times the
// ratio of the largest window scale its chain crossed to the cell's own (per 64-row block
// and 64-column tile, k_pass_init), so the surviving set contains every row whose exact NN
// distance exceeds r2, and each survivor carries bounds alo <= exact <= ahi.  TopK ranks
// survivors by exact distance, computing the exact (double) distance only for candidates
// whose rank those intervals leave undecided and for every accepted discord (whose reported
// distance is then the double-precision value).
//
// Forms.
//   plain   each series alone, each length with the r-descent of src/main.cu:136-233.
//   fast    all series of the call in every launch; per length, r comes from the exact
//           double-precision NN distances (at the new length) of the previous length's
//           discords and their backward shifts -- a lower bound of the new k-th distance
//           whenever they stay mutually non-overlapping -- or, for the first length, from
//           the learned proposal raised to the exact bound of the best-scored candidate
//           rows (learned form) or a geometric probe descent (batched form); blocks first
//           rescan the tiles that pruned their rows at the previous length.  The result is
//           the same discord set: the top-k does not depend on r once r is below the k-th
//           distance.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/paddi.h"
#include "engine.h"
#include "kernels.h"
#include "learned.h"

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

// Nonzero while a no-allocation call runs on this thread (ScratchScope): every buffer it needs
// is bound beforehand, so a buffer that would have to grow is an error, not an allocation.
static thread_local int t_noalloc = 0;

static void refuse_growth(const char *what)
{
    throw HipError(PADDI_ERR_UNSUPPORTED,
                   std::string("no-allocation call: ") + what + " does not fit its reserved space");
}

void DevBuf::reserve(size_t nbytes)
{
    if (nbytes <= bytes) return;
    if (borrowed || t_noalloc) refuse_growth("a device buffer");
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    HIPC(hipMalloc(&p, nbytes));
    bytes = nbytes;
}
void DevBuf::release()
{
    if (p && !borrowed) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    borrowed = false;
}

// Binds device buffers to consecutive 256-byte aligned slices of the scratch array for one
// no-allocation call and gives each back its own memory when the call ends.  Binding checks
// the slices against SCRATCH_BYTES before any device work is queued.
class ScratchScope {
public:
    explicit ScratchScope(Engine &E) : E_(E), base_((char *)E.scratch_)
    {
        if (!base_) throw HipError(PADDI_ERR_DEVICE, "no scratch array on this device");
        t_noalloc++;
        saved_graph_ = E.use_graph_;
        E.use_graph_ = false;   // a graph would bake these slices and allocate to instantiate
    }
    ~ScratchScope()
    {
        for (size_t q = saved_.size(); q-- > 0;) {
            Saved &v = saved_[q];
            v.b->p = v.p;
            v.b->bytes = v.bytes;
            v.b->borrowed = v.borrowed;
        }
        E_.use_graph_ = saved_graph_;
        t_noalloc--;
    }
    ScratchScope(const ScratchScope &) = delete;
    ScratchScope &operator=(const ScratchScope &) = delete;
    void bind(DevBuf &b, size_t bytes)
    {
        off_ = (off_ + 255) & ~(size_t)255;
        if (off_ + bytes > SCRATCH_BYTES)
            throw HipError(PADDI_ERR_UNSUPPORTED,
                           "no-allocation call: the series needs more than the " +
                               std::to_string(SCRATCH_BYTES >> 20) + " MiB scratch array");
        place(b, base_ + off_, bytes);
        off_ += bytes;
    }
    // b shares a's slice (buffers a call never uses together)
    void share(DevBuf &b, const DevBuf &a) { place(b, a.p, a.bytes); }

private:
    struct Saved {
        DevBuf *b;
        void *p;
        size_t bytes;
        bool borrowed;
    };
    void place(DevBuf &b, void *p, size_t bytes)
    {
        saved_.push_back(Saved{&b, b.p, b.bytes, b.borrowed});
        b.p = p;
        b.bytes = bytes;
        b.borrowed = true;
    }
    Engine &E_;
    char *base_;
    size_t off_ = 0;
    bool saved_graph_ = false;
    std::vector<Saved> saved_;
};

static inline unsigned grid1(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// Bytes of one exact launch block (Engine::exact) for nc candidates: EXACT_SLOTS result keys
// per candidate (twice with a second length), the candidates' rows, and per group (at most
// one per candidate, plus one closing entry per launch of one group size) its first block,
// first candidate and series.
static size_t exact_arg_bytes(size_t nc, bool two_lengths)
{
    const size_t ngx = nc + MAXU;
    return nc * EXACT_SLOTS * 8 * (two_lengths ? 2 : 1) + nc * 8 + ngx * 16;
}
static inline int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

// ------------------------------------------------------------------------------------------
// Exact greedy top-k over the survivors of one series.
//
// Every survivor c has an interval [lo_c, hi_c] = [alo_c - conv, ahi_c + conv] (conv_ed2)
// holding its
// exact distance (a point once refined): the scans' bounds, widened by the rounding of
// their conversion to distances.  T = argmax hi; the exact maximum lies in
// U = {c : hi_c >= lo_T}.  Once U is all exact, its maximum (ties: lower index) is the next
// candidate in exact order; it is accepted when its distance exceeds r2, and every
// candidate overlapping it (|i - a| < m) is dropped, since all of them rank below it.

static bool ranks_before(double va, int64_t ia, double vb, int64_t ib)
{
    const double tol = TIE_REL * std::max(std::fabs(va), std::fabs(vb));
    if (va > vb + tol) return true;
    if (vb > va + tol) return false;
    return ia < ib;
}

void TopK::reset(std::vector<Cand> &&c, double r2_, int m_, int k_)
{
    cands = std::move(c);
    gone.assign(cands.size(), 0);
    acc.clear();
    r2 = r2_;
    m = m_;
    k = k_;
    done = false;
}

bool TopK::advance(std::vector<size_t> &need)
{
    need.clear();
    auto lo = [&](const Cand &c) {
        return c.exact ? c.v
                       : (c.a >= 3.0e38f ? INFINITY : (double)c.alo - conv_ed2(m, c.alo));
    };
    auto hi = [&](const Cand &c) {
        return c.exact ? c.v
                       : (c.a >= 3.0e38f ? INFINITY : (double)c.ahi + conv_ed2(m, c.ahi));
    };
    while (!done) {
        if ((int)acc.size() >= k) { done = true; break; }
        int64_t T = -1;
        double hT = -INFINITY;
        for (size_t q = 0; q < cands.size(); q++) {
            if (gone[q]) continue;
            const double h = hi(cands[q]);
            if (T < 0 || h > hT || (h == hT && cands[q].i < cands[(size_t)T].i)) {
                T = (int64_t)q;
                hT = h;
            }
        }
        if (T < 0) { done = true; break; }
        // every candidate that can rank at or above T's lower bound (ties included)
        const double loT = lo(cands[(size_t)T]);
        const double loTt = loT - 2.0 * TIE_REL * std::fabs(loT);
        std::vector<size_t> U;
        for (size_t q = 0; q < cands.size(); q++)
            if (!gone[q] && hi(cands[q]) >= loTt) U.push_back(q);
        if (U.empty()) U.push_back((size_t)T);  // only when distances are not ordered (NaN)
        for (size_t q : U)
            if (!cands[q].exact) need.push_back(q);
        if (!need.empty()) {
            if (need.size() > 64) {
                std::partial_sort(need.begin(), need.begin() + 64, need.end(),
                                  [&](size_t a, size_t b) { return hi(cands[a]) > hi(cands[b]); });
                need.resize(64);
            }
            // Speculate: refine in the same launch the candidates likely to be accepted next,
            // up to `spec` per round: by descending upper bound, skipping any within m of T or
            // of one already taken (those would be dropped as trivial matches).  Once U is
            // exact its best is accepted, so with one acceptance left there is no next.
            if ((int)need.size() < spec && (int)acc.size() + 1 < k) {
                std::vector<char> in(cands.size(), 0);
                for (size_t q : need) in[q] = 1;
                std::vector<size_t> extra;
                for (size_t q = 0; q < cands.size(); q++)
                    if (!gone[q] && !in[q] && !cands[q].exact) extra.push_back(q);
                std::sort(extra.begin(), extra.end(), [&](size_t a, size_t b) {
                    return hi(cands[a]) > hi(cands[b]) ||
                           (hi(cands[a]) == hi(cands[b]) && cands[a].i < cands[b].i);
                });
                std::vector<int64_t> anchor{cands[(size_t)T].i};
                for (size_t q : extra) {
                    if ((int)need.size() >= spec) break;
                    bool apart = true;
                    for (int64_t a : anchor) apart &= std::llabs(cands[q].i - a) >= m;
                    if (!apart) continue;
                    need.push_back(q);
                    anchor.push_back(cands[q].i);
                }
            }
            return true;
        }
        size_t B = U[0];
        for (size_t q : U)
            if (ranks_before(cands[q].v, cands[q].i, cands[B].v, cands[B].i)) B = q;
        if (!(cands[B].v > r2)) { done = true; break; }
        acc.push_back(cands[B]);
        const int64_t bi = cands[B].i;
        for (size_t q = 0; q < cands.size(); q++)
            if (!gone[q] && std::llabs(cands[q].i - bi) < m) gone[q] = 1;
    }
    return false;
}

// ------------------------------------------------------------------------------------------

Engine::Engine(int device) : device_(device)
{
    HIPC(hipSetDevice(device));
    HIPC(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    for (auto &e : ev_) HIPC(hipEventCreate(&e));
    HIPC(hipHostMalloc(&pin_, 64, hipHostMallocDefault));  // k_prefix's non-finite flag
    HIPC(hipHostMalloc(&pin_out_, PASS_OUT_HEAD + (size_t)SCAP * sizeof(Surv),
                       hipHostMallocDefault));
    HIPC(hipEventCreateWithFlags(&ser_ev_, hipEventDisableTiming));
    HIPC(hipEventCreateWithFlags(&sync_ev_, hipEventDisableTiming));
    // Staging the no-allocation call uses, sized here: a one-series pass's parameters and
    // series table, and the arguments and results of one exact launch of
    // EXACT_NOALLOC_CHUNK candidates (exact_arg_bytes).
    ensure_pinned(4096, 4096);
    pin_x_bytes_ = exact_arg_bytes(EXACT_NOALLOC_CHUNK, false);
    HIPC(hipHostMalloc(&pin_x_, pin_x_bytes_, hipHostMallocDefault));
    scratch_ = scratch_setup(device);
    const char *dbg = std::getenv("PADDI_DEBUG");
    debug_ = dbg && dbg[0] == '1';
    const char *enn = std::getenv("PADDI_EXACT_NN");
    if (enn && enn[0] == '0') exact_nn_ = false;
    const char *lim = std::getenv("PADDI_WAIT_LIMIT_S");
    if (lim && *lim) wait_limit_s_ = std::atof(lim);
}

Engine::~Engine()
{
    (void)hipSetDevice(device_);
    (void)hipStreamSynchronize(stream_);
    drop_graphs();
    for (auto &e : ev_) (void)hipEventDestroy(e);
    (void)hipEventDestroy(ser_ev_);
    (void)hipEventDestroy(sync_ev_);
    for (void *p : {pin_, pin_par_, pin_out_, pin_ser_, pin_x_, pin_ch_})
        if (p) (void)hipHostFree(p);
    (void)hipStreamDestroy(stream_);
}

void Engine::drop_graphs()
{
    for (auto &g : graphs_)
        if (g.second.exec) (void)hipGraphExecDestroy(g.second.exec);
    graphs_.clear();
    for (auto &g : chain_graphs_)
        if (g.second.exec) (void)hipGraphExecDestroy(g.second.exec);
    chain_graphs_.clear();
}

// Pinned staging grows with the batch; a graph that copies from the old buffer is dropped.
void Engine::ensure_pinned(size_t par_bytes, size_t ser_bytes)
{
    if ((par_bytes > pin_par_bytes_ || ser_bytes > pin_ser_bytes_) && t_noalloc)
        refuse_growth("the pinned parameter staging");
    if (par_bytes > pin_par_bytes_) {
        HIPC(hipStreamSynchronize(stream_));
        drop_graphs();
        if (pin_par_) HIPC(hipHostFree(pin_par_));
        pin_par_ = nullptr;
        pin_par_bytes_ = 0;
        HIPC(hipHostMalloc(&pin_par_, par_bytes, hipHostMallocDefault));
        pin_par_bytes_ = par_bytes;
    }
    if (ser_bytes > pin_ser_bytes_) {
        HIPC(hipEventSynchronize(ser_ev_));
        if (pin_ser_) HIPC(hipHostFree(pin_ser_));
        pin_ser_ = nullptr;
        pin_ser_bytes_ = 0;
        HIPC(hipHostMalloc(&pin_ser_, ser_bytes, hipHostMallocDefault));
        pin_ser_bytes_ = ser_bytes;
    }
}

Series *Engine::upload(const float *T, int64_t n)
{
    HIPC(hipSetDevice(device_));
    std::unique_ptr<Series> s(new Series());
    s->n = n;
    s->x.reserve((size_t)n * sizeof(float));
    HIPC(hipMemcpyAsync(s->x.p, T, (size_t)n * sizeof(float), hipMemcpyHostToDevice, stream_));
    HIPC(hipStreamSynchronize(stream_));
    return s.release();
}

void Engine::build_batch(Batch &B, const std::vector<const Series *> &ss, int minL,
                         const float *const *host)
{
    derive_m2_ = 0;  // statistics computed ahead belong to the previous row layout
    B.nser = (int)ss.size();
    B.minL = minL;
    B.n.assign(ss.size(), 0);
    B.xoff.assign(ss.size(), 0);
    B.roff.assign(ss.size(), 0);
    B.pboff.assign(ss.size(), 0);
    int64_t xo = 0, ro = 0, po = 0;
    for (size_t q = 0; q < ss.size(); q++) {
        const int64_t n = ss[q]->n;
        B.n[q] = n;
        B.xoff[q] = xo;
        B.roff[q] = ro;
        B.pboff[q] = po;
        xo += round_up(n + XPAD, 64);
        ro += series_rows(n, minL);
        po += (n + PB - 1) / PB + 1;
    }
    B.xtotal = xo;
    B.rows = ro;
    B.nblocks = ro / 64;
    B.npb = po;
    B.x.reserve((size_t)xo * sizeof(float));
    B.P.reserve((size_t)po * (PB + 1) * sizeof(double2));
    B.refs.reserve((size_t)po * sizeof(double4));
    B.ser.reserve(ss.size() * sizeof(SerDev) * 4);
    B.bser.reserve((size_t)B.nblocks * sizeof(int));
    B.pser.reserve((size_t)po * sizeof(int));
    HIPC(hipMemsetAsync(B.x.p, 0, (size_t)xo * sizeof(float), stream_));
    for (size_t q = 0; q < ss.size(); q++)
        HIPC(hipMemcpyAsync(B.x.as<float>() + B.xoff[q], host ? (const void *)host[q] : ss[q]->x.p,
                            (size_t)B.n[q] * 4,
                            host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, stream_));
    B.bser_host.assign((size_t)B.nblocks, -1);
    std::vector<int> pser((size_t)po, 0);
    for (size_t q = 0; q < ss.size(); q++) {
        const int64_t nb = (B.n[q] - minL + 1 + 63) / 64;
        for (int64_t b = 0; b < nb; b++) B.bser_host[(size_t)(B.roff[q] / 64 + b)] = (int)q;
        const int64_t np = (B.n[q] + PB - 1) / PB + 1;
        for (int64_t b = 0; b < np; b++) pser[(size_t)(B.pboff[q] + b)] = (int)q;
    }
    HIPC(hipMemcpyAsync(B.bser.p, B.bser_host.data(), (size_t)B.nblocks * 4,
                        hipMemcpyHostToDevice, stream_));
    HIPC(hipMemcpyAsync(B.pser.p, pser.data(), (size_t)po * 4, hipMemcpyHostToDevice, stream_));
    B.host.assign(ss.size(), SerDev{});
    for (size_t q = 0; q < ss.size(); q++) {
        SerDev &S = B.host[q];
        S.xoff = B.xoff[q];
        S.n = B.n[q];
        S.roff = B.roff[q];
        S.pboff = B.pboff[q];
        S.N = B.n[q] - minL + 1;
        S.thr = 1.0f;
    }
    upload_ser(B);
    HIPC(hipStreamSynchronize(stream_));  // host vectors above must outlive the copies
}

// Wait for everything queued on stream_: the engine's round trips are short, so the host
// spins on an event instead of sleeping in hipStreamSynchronize (whose wake-up costs tens of
// microseconds per round trip); past 2 ms it yields the core between queries and past
// 100 ms sleeps 100 us between them.  Work that has not completed after wait_limit_s_
// (PADDI_WAIT_LIMIT_S, default 600 s) raises PADDI_ERR_DEVICE; the work itself is not
// cancelled, so the context's later calls queue behind it.
void Engine::wait()
{
    HIPC(hipEventRecord(sync_ev_, stream_));
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        const hipError_t e = hipEventQuery(sync_ev_);
        if (e == hipSuccess) return;
        if (e != hipErrorNotReady) HIPC(e);
        const double el =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (el >= wait_limit_s_)
            throw HipError(PADDI_ERR_DEVICE, "device work did not complete within " +
                                                 std::to_string(wait_limit_s_) + " s");
        if (el > 0.1)
            std::this_thread::sleep_for(std::chrono::microseconds(100));
        else if (el > 2e-3)
            std::this_thread::yield();
    }
}

// Series table (N at the current length) to the device, ordered on stream_ before the
// kernels that follow; the host array may change right after the call.
void Engine::upload_ser(Batch &B)
{
    const size_t bytes = B.host.size() * sizeof(SerDev);
    ensure_pinned(0, bytes);
    HIPC(hipEventSynchronize(ser_ev_));  // the previous copy out of pin_ser_ has finished
    std::memcpy(pin_ser_, B.host.data(), bytes);
    HIPC(hipMemcpyAsync(B.ser.p, pin_ser_, bytes, hipMemcpyHostToDevice, stream_));
    HIPC(hipEventRecord(ser_ev_, stream_));
}

void Engine::prefix(Batch &B)
{
    flag_.reserve(64);
    HIPC(hipMemsetAsync(flag_.p, 0, sizeof(int), stream_));
    hipLaunchKernelGGL(k_prefix, dim3((unsigned)B.npb), dim3(256), 0, stream_,
                       B.x.as<float>(), B.ser.as<SerDev>(), B.pser.as<int>(), B.P.as<double2>(),
                       B.refs.as<double4>(), flag_.as<int>());
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(pin_, flag_.p, sizeof(int), hipMemcpyDeviceToHost, stream_));
    wait();
    if (*(volatile int *)pin_)
        throw HipError{PADDI_ERR_INVALID, "the series holds a NaN or infinite value"};
}

void Engine::derive(Batch &B, int m, paddi_stats *st, int m2)
{
    for (int q = 0; q < B.nser; q++) B.host[(size_t)q].N = B.n[(size_t)q] - m + 1;
    colv_.reserve((size_t)B.rows * sizeof(float4));
    st64_.reserve((size_t)B.rows * sizeof(double2));
    tmax_.reserve((size_t)B.nblocks * sizeof(float2));
    // the previous derive's second length is this one: its statistics are already there
    const bool have =
        derive_m2_ == m && derive_m2_ > 0 && st64n_.bytes >= (size_t)B.rows * sizeof(double2);
    if (have) st64_.swap(st64n_);
    if (m2) st64n_.reserve((size_t)B.rows * sizeof(double2));
    derive_m2_ = m2;
    rho_.reserve((size_t)B.nblocks * sizeof(float));
    HIPC(hipEventRecord(ev_[4], stream_));
    hipLaunchKernelGGL(k_derive, dim3(grid1(B.rows, 256 * DERIVE_RPT)), dim3(256), 0, stream_,
                       B.x.as<float>(), B.ser.as<SerDev>(), B.bser.as<int>(), B.P.as<double2>(),
                       B.refs.as<double4>(), B.rows, m, colv_.as<float4>(),
                       have ? nullptr : st64_.as<double2>(), tmax_.as<float2>(), m2,
                       st64n_.as<double2>(), nullptr);
    HIPC(hipGetLastError());
    HIPC(hipEventRecord(ev_[5], stream_));
    derive_st_ = st;  // timed at the next synchronisation (collect_derive)
}

void Engine::collect_derive()
{
    if (!derive_st_) return;
    HIPC(hipEventSynchronize(ev_[5]));
    float a = 0.f;
    HIPC(hipEventElapsedTime(&a, ev_[4], ev_[5]));
    derive_st_->kernel_ms[0] += a;
    derive_st_->launches[0] += 1;
    derive_st_ = nullptr;
}

void Engine::copy_st64(double2 *dst, int64_t N)
{
    HIPC(hipMemcpyAsync(dst, st64_.p, (size_t)N * sizeof(double2), hipMemcpyDeviceToHost,
                        stream_));
    HIPC(hipStreamSynchronize(stream_));
}

// The device work of one pass, in stream order: parameter upload, k_pass_init, k_scan_select,
// then refine_rounds k_scan_refine rounds each preceded by a listing of the live
// blocks, k_survivors, and the download of the counters and the first scap survivors.  No
// step depends on a value read back by the host, so the sequence is captured once per
// LDS class and replayed.  Eager launches are bracketed by events for the statistics;
// a captured sequence has none (event nodes of a graph do not time the kernels between them).
void Engine::enqueue_pass(const ScanPtrs &Q, size_t lds, size_t parb, int64_t rows,
                          int64_t nwords, int64_t scap, bool timed, const uint8_t *keep,
                          bool host_io)
{
    const PassDev *P = par_.as<PassDev>();
    PassCtl *ctl = pout_.as<PassCtl>();
    Surv *surv = (Surv *)((char *)pout_.p + PASS_OUT_HEAD);
    // k_scan_refine needs more workgroups than listed blocks (see there)
    const unsigned grid =
        (unsigned)std::max<int64_t>(refine_waves_ + refine_waves_ / 2, nwords + 64);
    if (host_io) HIPC(hipMemcpyAsync(par_.p, pin_par_, parb, hipMemcpyHostToDevice, stream_));
    hipLaunchKernelGGL(k_pass_init, dim3(grid1(std::max<int64_t>(rows, 1024), 256)), dim3(256), 0,
                       stream_, Q, P, rows, nwords, ctl);
    if (keep)
        hipLaunchKernelGGL(k_mask_blocks, dim3(grid1(nwords, 256)), dim3(256), 0, stream_,
                           alive_.as<uint64_t>(), keep, nwords);
    if (timed) HIPC(hipEventRecord(ev_[0], stream_));
    launch_scan_select(Q, P, (unsigned)nwords, lds, stream_);
    if (timed) HIPC(hipEventRecord(ev_[1], stream_));
    hipLaunchKernelGGL(k_list_blocks, dim3(grid1(nwords, 256 * LIST_WPT)), dim3(256), 0, stream_,
                       (const uint64_t *)alive_.p, nwords, blist_.as<int>(), &ctl->list[0]);
    if (timed) HIPC(hipEventRecord(ev_[2], stream_));
    for (int r = 0; r < refine_rounds_; r++) {
        if (r > 0)
            hipLaunchKernelGGL(k_list_blocks, dim3(grid1(nwords, 256 * LIST_WPT)), dim3(256), 0, stream_,
                               (const uint64_t *)alive_.p, nwords, blist_.as<int>(),
                               &ctl->list[r]);
        launch_scan_refine(Q, P, blist_.as<int>(), ctl, r, grid, lds, stream_);
    }
    if (timed) HIPC(hipEventRecord(ev_[3], stream_));
    hipLaunchKernelGGL(k_survivors, dim3(grid1(nwords, 256)), dim3(256), 0, stream_, Q, P,
                       rows, ctl, surv, rows);
    if (host_io)
        HIPC(hipMemcpyAsync(pin_out_, pout_.p, PASS_OUT_HEAD + (size_t)scap * sizeof(Surv),
                            hipMemcpyDeviceToHost, stream_));
    HIPC(hipGetLastError());
}

// Buffers and the parameter block of a pass (pinned, copied by enqueue_pass): the pass's
// series table (r2 < 0 switches a series off) and scan arguments.  WA = select window;
// refine = false leaves every refine round empty.
ScanPtrs Engine::fill_pass(Batch &B, int m, const std::vector<double> &r2, bool witness,
                           int64_t WA, bool refine, size_t &parb)
{
    const double delta = delta_ed2(m);
    const int64_t rows = B.rows, nwords = B.nblocks;
    parb = sizeof(PassDev) + (size_t)B.nser * sizeof(SerDev);
    alive_.reserve((size_t)nwords * 8);
    best_.reserve((size_t)rows * 4);
    besth_.reserve((size_t)rows * 4);
    bestl_.reserve((size_t)rows * 4);
    blist_.reserve((size_t)nwords * 4);
    pout_.reserve(PASS_OUT_HEAD + (size_t)rows * sizeof(Surv));
    wit_[0].reserve((size_t)rows * 4);
    wit_[1].reserve((size_t)rows * 4);
    par_.reserve(parb);
    ensure_pinned(parb, 0);

    PassDev &P = *(PassDev *)pin_par_;
    SerDev *S = (SerDev *)((char *)pin_par_ + sizeof(PassDev));
    int64_t nmax = 0;
    for (int q = 0; q < B.nser; q++) {
        S[q] = B.host[(size_t)q];
        S[q].N = r2[(size_t)q] < 0.0 ? 0 : B.n[(size_t)q] - m + 1;
        S[q].thr = (float)(1.0 - r2[(size_t)q] / (2.0 * m));
        nmax = std::max(nmax, S[q].N);
    }
    ScanPtrs Q{};
    Q.x = B.x.as<float>();
    Q.ser = (const SerDev *)((char *)par_.p + sizeof(PassDev));
    Q.bser = B.bser.as<int>();
    Q.colv = colv_.as<float4>();
    Q.tmax = tmax_.as<float2>();
    Q.rho = rho_.as<float>();
    Q.alive = alive_.as<uint64_t>();
    Q.best = best_.as<int>();
    Q.besth = besth_.as<int>();
    Q.bestl = bestl_.as<int>();
    Q.wit0 = wit_[0].as<int>();
    Q.wit1 = wit_[1].as<int>();
    Q.pairs = &pout_.as<PassCtl>()->pairs[0][0];
    ScanScal &A = P.S;
    A = ScanScal{};
    A.WA = WA;
    A.U0 = round_up(m + 64, 64);
    A.m = m;
    A.ml = excl_sym_ ? m - 1 : m + 1;
    A.dbudget = direct_budget_;
    A.drows = direct_rows_;
    A.witness_max = witness_tiles_;
    A.wsel = witness ? witcur_ : -1;
    A.rsel = (witness && wread_valid_) ? (witcur_ ^ 1) : -1;
    A.dk = (float)(delta / (2.0 * m));
    A.some_off = 0;
    for (int q = 0; q < B.nser; q++) A.some_off |= r2[(size_t)q] < 0.0;
    // Refine rounds over cyclic offsets [U0 + WA, N + U0) from each block, nearest first;
    // between rounds the live blocks are listed again, so far columns are scanned only for
    // the few blocks the near ones did not kill.
    RefineArgs &R = P.R;
    const int64_t end = refine ? nmax + A.U0 : A.U0 + A.WA;
    refine_bounds(A.U0 + A.WA, refine_chunk_, end, refine_rounds_, R.bound);
    R.min_chunk = min_chunk_;
    R.order = refine_order_ == 2 ? (first_len_ ? 1 : 0) : refine_order_;
    R.max_chunk = refine_chunk_;
    R.items = refine_waves_;
    return Q;
}

// Candidates for a first-length range: a select-only pass at r2 = 0 (nothing is pruned)
// over a window of BOUND_WINDOW columns gives every row of the series in `need` an upper
// bound on its NN distance; each 64-row block offers its row with the largest bound, and the
// `count` best blocks of each series are returned (local rows, best first).
void Engine::bound_candidates(Batch &B, int m, const std::vector<char> &need, int count,
                              std::vector<std::vector<int64_t>> &cand, paddi_stats *st)
{
    std::vector<double> r2((size_t)B.nser, -1.0);
    for (int q = 0; q < B.nser; q++)
        if (need[(size_t)q]) r2[(size_t)q] = 0.0;
    size_t parb = 0;
    const ScanPtrs Q = fill_pass(B, m, r2, false, BOUND_WINDOW, false, parb);
    const int64_t rows = B.rows, nwords = B.nblocks;
    const size_t lds = scan_lds_bytes(m);
    bkey_.reserve((size_t)nwords * 8);
    const PassDev *P = par_.as<PassDev>();
    PassCtl *ctl = pout_.as<PassCtl>();
    HIPC(hipMemcpyAsync(par_.p, pin_par_, parb, hipMemcpyHostToDevice, stream_));
    hipLaunchKernelGGL(k_pass_init, dim3(grid1(std::max<int64_t>(rows, 1024), 256)), dim3(256), 0,
                       stream_, Q, P, rows, nwords, ctl);
    HIPC(hipEventRecord(ev_[0], stream_));
    launch_scan_select(Q, P, (unsigned)nwords, lds, stream_);
    HIPC(hipEventRecord(ev_[1], stream_));
    hipLaunchKernelGGL(k_block_best, dim3(grid1(nwords * 64, 256)), dim3(256), 0, stream_, Q, P,
                       nwords, bkey_.as<unsigned long long>());
    HIPC(hipGetLastError());
    std::vector<unsigned long long> keys((size_t)nwords);
    HIPC(hipMemcpyAsync(keys.data(), bkey_.p, (size_t)nwords * 8, hipMemcpyDeviceToHost, stream_));
    wait();
    collect_derive();
    if (st) {
        float a = 0.f;
        HIPC(hipEventElapsedTime(&a, ev_[0], ev_[1]));
        st->kernel_ms[1] += a;
        st->launches[1] += 1;
    }
    cand.assign((size_t)B.nser, {});
    for (int q = 0; q < B.nser; q++) {
        if (!need[(size_t)q]) continue;
        const int64_t b0 = B.roff[(size_t)q] / 64;
        const int64_t nb = (B.n[(size_t)q] - m + 1 + 63) / 64;
        std::vector<unsigned long long> kq(keys.begin() + b0, keys.begin() + b0 + nb);
        const size_t c = std::min<size_t>((size_t)count, kq.size());
        std::partial_sort(kq.begin(), kq.begin() + (std::ptrdiff_t)c, kq.end());
        for (size_t u = 0; u < c; u++)
            if (kq[u] != ~0ull)
                cand[(size_t)q].push_back((int64_t)(kq[u] & 0xffffffffull) - B.roff[(size_t)q]);
    }
}

// Lower bound of d_k per series from its best-scored candidate rows: their exact NN
// distances, greedily accepted like the top-k (descending, ties to the lower index, no two
// within m); the k-th accepted value is a lower bound of d_1 when k = 1 (any row's NN
// distance is) and the usual estimate otherwise (a failed pass steps below it).
void Engine::first_bound(Batch &B, int m, int k, const std::vector<char> &need,
                         std::vector<double> &lb, paddi_stats *st,
                         std::vector<std::vector<int64_t>> *cand_out)
{
    lb.assign((size_t)B.nser, 0.0);
    std::vector<std::vector<int64_t>> cand;
    bound_candidates(B, m, need, bound_cands_, cand, st);
    if (cand_out) *cand_out = cand;
    std::vector<std::pair<int, int64_t>> req;
    for (int q = 0; q < B.nser; q++)
        for (int64_t i : cand[(size_t)q]) req.emplace_back(q, i);
    if (req.empty()) return;
    std::vector<double> v;
    exact(B, m, req, v, st);
    std::vector<std::vector<std::pair<double, int64_t>>> per((size_t)B.nser);
    for (size_t u = 0; u < req.size(); u++)
        if (v[u] < 3.0e38) per[(size_t)req[u].first].emplace_back(v[u], req[u].second);
    for (int q = 0; q < B.nser; q++) {
        auto &c = per[(size_t)q];
        std::sort(c.begin(), c.end(), [](const std::pair<double, int64_t> &a,
                                         const std::pair<double, int64_t> &b) {
            return a.first > b.first || (a.first == b.first && a.second < b.second);
        });
        std::vector<int64_t> acc;
        double last = 0.0;
        for (const auto &e : c) {
            bool ok = true;
            for (int64_t a : acc) ok &= std::llabs(e.second - a) >= m;
            if (!ok) continue;
            acc.push_back(e.second);
            last = e.first;
            if ((int)acc.size() == k) break;
        }
        if ((int)acc.size() == k && last > 0.0) lb[(size_t)q] = last * (1.0 - 1e-6);
    }
}

// The first-length range of the learned form.  first_bound's candidates are each block's
// best-scored row, usually a neighbour of the block's discord rather than the discord itself,
// so their k-th greedy distance f falls short of d_k.  The candidate blocks are therefore
// scanned in full at r2 = f (a pass restricted to those blocks), and the exact top-k of their
// rows whose NN distance exceeds f gives f' >= f: d_k itself whenever the top-k lie in those
// blocks (for k = 1 any row's NN distance is at most d_1, so f' <= d_1 always; for k > 1 a
// range above d_k only makes the first pass fail and the descent step down).
void Engine::bootstrap(Batch &B, int m, int k, const std::vector<char> &need,
                       std::vector<double> &lb, std::vector<char> &found, paddi_stats *st)
{
    found.assign((size_t)B.nser, 0);
    std::vector<std::vector<int64_t>> cand;
    first_bound(B, m, k, need, lb, st, &cand);
    std::vector<double> r2((size_t)B.nser, -1.0);
    std::vector<uint8_t> keep((size_t)B.nblocks, 0);
    bool any = false;
    for (int q = 0; q < B.nser; q++) {
        if (!need[(size_t)q] || !(lb[(size_t)q] > 0.0)) continue;
        r2[(size_t)q] = lb[(size_t)q];
        for (int64_t i : cand[(size_t)q]) keep[(size_t)((B.roff[(size_t)q] + i) >> 6)] = 1;
        any = true;
    }
    if (!any) return;
    std::vector<std::vector<Cand>> surv;
    pass(B, m, r2, false, surv, st, &keep);
    std::vector<TopK> tk((size_t)B.nser);
    for (int q = 0; q < B.nser; q++) {
        if (r2[(size_t)q] < 0.0) {
            tk[(size_t)q].done = true;
            continue;
        }
        tk[(size_t)q].reset(std::move(surv[(size_t)q]), r2[(size_t)q], m, k);
        tk[(size_t)q].spec = spec_;
    }
    topk_all(B, m, tk, st);
    for (int q = 0; q < B.nser; q++) {
        const std::vector<Cand> &acc = tk[(size_t)q].acc;
        if (r2[(size_t)q] >= 0.0 && (int)acc.size() >= k) {
            found[(size_t)q] = 1;
            if (acc.back().v * (1.0 - 1e-6) > lb[(size_t)q]) lb[(size_t)q] = acc.back().v * (1.0 - 1e-6);
        }
    }
}

// One (m, r2_s) pass: survivors per series.  Series with r2 < 0 are skipped (no rows).
void Engine::pass(Batch &B, int m, const std::vector<double> &r2, bool witness,
                  std::vector<std::vector<Cand>> &out, paddi_stats *st,
                  const std::vector<uint8_t> *keep)
{
    const int64_t rows = B.rows, nwords = B.nblocks;
    // a restricted pass runs eagerly (its block flags are not part of the captured graph)
    const bool graph = use_graph_ && !keep;
    const uint8_t *keep_dev = nullptr;
    if (keep) {
        keep_.reserve((size_t)nwords);
        HIPC(hipMemcpyAsync(keep_.p, keep->data(), (size_t)nwords, hipMemcpyHostToDevice,
                            stream_));
        keep_dev = keep_.as<uint8_t>();
    }
    // a graph is shared by the lengths of a 64-wide class, so it carries the class's LDS
    const size_t lds = graph ? scan_lds_bytes((int)std::min<int64_t>(round_up(m, 64), MAXM))
                                  : scan_lds_bytes(m);
    // The survivors come back with the counters in one download, sized from the previous
    // pass's count (a graph keeps the largest size it was captured with).
    const int64_t scap = std::min<int64_t>(graph ? SCAP : scap_next_, rows);
    size_t parb = 0;
    const ScanPtrs Q = fill_pass(B, m, r2, witness, select_window_, true, parb);

    if (graph) {
        const std::vector<uintptr_t> sig = {
            (uintptr_t)par_.p,   (uintptr_t)pin_par_, (uintptr_t)pin_out_, (uintptr_t)pout_.p,
            (uintptr_t)alive_.p, (uintptr_t)blist_.p,
            (uintptr_t)Q.x,      (uintptr_t)Q.bser,   (uintptr_t)Q.colv,   (uintptr_t)Q.best,
            (uintptr_t)Q.wit0,   (uintptr_t)Q.wit1,   (uintptr_t)Q.besth,  (uintptr_t)Q.bestl,
            (uintptr_t)Q.tmax,   (uintptr_t)Q.rho,    (uintptr_t)rows,     (uintptr_t)nwords,
            (uintptr_t)parb,     (uintptr_t)scap,     (uintptr_t)refine_waves_,
            (uintptr_t)refine_rounds_};
        GraphEnt &G = graphs_[{0, lds}];
        if (!G.exec || G.sig != sig) {
            if (G.exec) HIPC(hipGraphExecDestroy(G.exec));
            G.exec = nullptr;
            hipGraph_t g = nullptr;
            HIPC(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed));
            try {
                enqueue_pass(Q, lds, parb, rows, nwords, scap, false);
            } catch (...) {
                (void)hipStreamEndCapture(stream_, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIPC(hipStreamEndCapture(stream_, &g));
            const hipError_t e = hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIPC(e);
            G.sig = sig;
        }
        HIPC(hipGraphLaunch(G.exec, stream_));
    } else {
        enqueue_pass(Q, lds, parb, rows, nwords, scap, true, keep_dev);
    }
    wait();
    collect_derive();
    const PassCtl C = *(const PassCtl *)pin_out_;
    double pairs_sel = 0.0, pairs_ref = 0.0, live_sel = 0.0, live_ref = 0.0;
    for (int q = 0; q < PAIR_SLOTS; q++) {
        pairs_sel += (double)C.pairs[0][q];
        pairs_ref += (double)C.pairs[1][q];
        live_sel += (double)C.live[0][q];
        live_ref += (double)C.live[1][q];
    }
    const int ns = C.survivors;
    scap_next_ = std::min<int64_t>(SCAP, std::max<int64_t>(SCAP_MIN, 2 * (int64_t)ns + 64));
    if (debug_) {
        float a = 0.f, b = 0.f;
        if (!graph) {
            HIPC(hipEventElapsedTime(&a, ev_[0], ev_[1]));
            HIPC(hipEventElapsedTime(&b, ev_[2], ev_[3]));
        }
        std::fprintf(stderr,
                     "[paddi] m=%d select %.2f ms %.3g pairs | refine blocks %d/%d/%d/%d %.2f ms "
                     "%.3g pairs | survivors %d\n",
                     m, a, pairs_sel, C.list[0], C.list[1], C.list[2],
                     refine_rounds_ > 3 ? C.list[3] : 0, b, pairs_ref, ns);
    }
    out.assign((size_t)B.nser, {});
    if (ns > 0) {
        std::vector<Surv> sv((size_t)ns);
        const int64_t head = std::min<int64_t>(ns, scap);
        std::memcpy(sv.data(), (const char *)pin_out_ + PASS_OUT_HEAD, (size_t)head * sizeof(Surv));
        if (ns > head) {
            HIPC(hipMemcpyAsync(sv.data() + head,
                                (const char *)pout_.p + PASS_OUT_HEAD + (size_t)head * sizeof(Surv),
                                (size_t)(ns - head) * sizeof(Surv), hipMemcpyDeviceToHost,
                                stream_));
            HIPC(hipStreamSynchronize(stream_));
        }
        for (int q = 0; q < ns; q++) {
            const int64_t g = sv[(size_t)q].g;
            const int sid = B.bser_host[(size_t)(g >> 6)];
            out[(size_t)sid].push_back(
                Cand{g - B.roff[(size_t)sid], sv[(size_t)q].d2, 0.0, false, sv[(size_t)q].d2lo,
                     sv[(size_t)q].d2hi});
        }
    }
    if (st) {
        if (!graph) {  // a replayed graph is timed as a whole by the caller only
            float a = 0.f, b = 0.f;
            HIPC(hipEventElapsedTime(&a, ev_[0], ev_[1]));
            HIPC(hipEventElapsedTime(&b, ev_[2], ev_[3]));
            st->kernel_ms[1] += a;
            st->launches[1] += 1;
            st->kernel_ms[2] += b;
            st->launches[2] += refine_rounds_;
        }
        if (!keep) {  // a candidate scan is not an (m, r) attempt of the search
            st->attempts += 1;
            st->survivors += ns;
        }
        st->pairs += (int64_t)(pairs_sel + pairs_ref);
        st->scan_pairs[0] += (int64_t)pairs_sel;
        st->scan_pairs[1] += (int64_t)pairs_ref;
        st->live_pairs[0] += (int64_t)live_sel;
        st->live_pairs[1] += (int64_t)live_ref;
    }
}

// Exact (double) NN distance of (series, local row) requests.
void Engine::exact(Batch &B, int m, const std::vector<std::pair<int, int64_t>> &req,
                   std::vector<double> &out, paddi_stats *st, int m2,
                   std::vector<double> *out2, const std::vector<double> *ub, bool lb)
{
    out.assign(req.size(), 0.0);
    if (out2) out2->assign(req.size(), DBL_MAX);
    if (req.empty()) return;
    // a no-allocation call refines at most EXACT_NOALLOC_CHUNK candidates per launch block
    if (t_noalloc && req.size() > EXACT_NOALLOC_CHUNK) {
        for (size_t a = 0; a < req.size(); a += EXACT_NOALLOC_CHUNK) {
            const size_t e = std::min(req.size(), a + EXACT_NOALLOC_CHUNK);
            const std::vector<std::pair<int, int64_t>> sub(req.begin() + (std::ptrdiff_t)a,
                                                           req.begin() + (std::ptrdiff_t)e);
            std::vector<double> o, o2, subub;
            if (ub) subub.assign(ub->begin() + (std::ptrdiff_t)a, ub->begin() + (std::ptrdiff_t)e);
            exact(B, m, sub, o, st, m2, out2 ? &o2 : nullptr, ub ? &subub : nullptr, lb);
            std::copy(o.begin(), o.end(), out.begin() + (std::ptrdiff_t)a);
            if (out2) std::copy(o2.begin(), o2.end(), out2->begin() + (std::ptrdiff_t)a);
        }
        return;
    }
    // the second length needs its statistics from the last derive
    if (!out2 || m2 <= m || m2 > m + EXACT_EXT || m2 != derive_m2_) m2 = 0;
    // the series table on the device is derive(m)'s: every series at N = n - m + 1
    // k_nn (matrix cores, double precision where it matters) unless turned off; the
    // no-allocation call keeps k_exact, whose staging it sizes when the context is created
    const bool nn = exact_nn_ && !t_noalloc;
    const int umax = nn ? nn_max_u(m) : exact_max_u(m);
    std::vector<size_t> order(req.size());
    for (size_t q = 0; q < req.size(); q++) order[q] = q;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return req[a].first < req[b].first; });
    // Groups: a series' candidates split evenly into groups of at most umax; groups of equal
    // size form one launch (k_exact's work per column is proportional to its group size).
    struct Grp {
        int sid;
        size_t first, n;  // into `order`
    };
    std::vector<std::vector<Grp>> bucket((size_t)MAXU + 1);
    for (size_t q = 0; q < order.size();) {
        const int sid = req[order[q]].first;
        size_t e = q;
        while (e < order.size() && req[order[e]].first == sid) e++;
        const size_t cnt = e - q, ng = (cnt + (size_t)umax - 1) / (size_t)umax;
        const size_t base = cnt / ng, extra = cnt % ng;
        for (size_t g = 0, f = q; g < ng; g++) {
            const size_t n = base + (g < extra ? 1 : 0);
            bucket[n].push_back(Grp{sid, f, n});
            f += n;
        }
        q = e;
    }
    // per launch: gfirst/gser/goff hold ngroups + 1 entries (the last closes gfirst)
    std::vector<int64_t> cand;
    std::vector<size_t> slot;  // cand index -> index into req
    std::vector<int> gfirst, gser;
    std::vector<int64_t> goff;
    struct Launch {
        int nu, ngroups;
        size_t off;
        int64_t blocks;
    };
    std::vector<Launch> launches;
    for (int nu = 1; nu <= MAXU; nu++) {
        if (bucket[(size_t)nu].empty()) continue;
        Launch L{nu, (int)bucket[(size_t)nu].size(), gfirst.size(), 0};
        for (const Grp &g : bucket[(size_t)nu]) {
            gfirst.push_back((int)cand.size());
            gser.push_back(g.sid);
            goff.push_back(L.blocks);
            L.blocks += (B.n[(size_t)g.sid] - m + 1 + EXACT_COLS - 1) / EXACT_COLS;
            for (size_t u = 0; u < g.n; u++) {
                cand.push_back(req[order[g.first + u]].second);
                slot.push_back(order[g.first + u]);
            }
        }
        gfirst.push_back((int)cand.size());
        gser.push_back(0);
        goff.push_back(L.blocks);
        launches.push_back(L);
    }
    // one block: outkey (EXACT_SLOTS per candidate, set to the maximum key) | outkey2 (with
    // m2) | cand | goff | gfirst | gser, staged in pinned memory and copied in one piece
    const size_t nc = cand.size(), ngx = gfirst.size();
    const size_t nk = nc * EXACT_SLOTS, nkeys = m2 ? 2 * nk : nk;
    const size_t o_cand = nkeys * 8, o_goff = o_cand + nc * 8, o_gfirst = o_goff + ngx * 8,
                 o_gser = o_gfirst + ngx * 4, o_ub = (o_gser + ngx * 4 + 7) & ~(size_t)7,
                 o_run = o_ub + (nn ? nc * 8 : 0), o_csid = o_run + (nn ? nk * 4 : 0),
                 xbytes = o_csid + (nn ? nc * 4 : 0);
    // k_nn's filter pass: a lower bound per candidate and 256-column super-tile
    int64_t tiles = 0;
    if (nn) {
        for (size_t q = 0; q < req.size(); q++)
            tiles = std::max<int64_t>(tiles, (B.n[(size_t)req[q].first] - m + 1 + 255) / 256);
        if (!lb) nnlo_.reserve((size_t)nc * (size_t)std::max<int64_t>(tiles, 1) * 4);
    }
    if (xbytes > pin_x_bytes_) {
        if (t_noalloc) refuse_growth("the pinned exact-refinement staging");
        if (pin_x_) HIPC(hipHostFree(pin_x_));
        pin_x_ = nullptr;
        pin_x_bytes_ = 0;
        HIPC(hipHostMalloc(&pin_x_, xbytes * 2, hipHostMallocDefault));
        pin_x_bytes_ = xbytes * 2;
    }
    char *hx = (char *)pin_x_;
    std::memset(hx, 0xff, nkeys * 8);
    std::memcpy(hx + o_cand, cand.data(), nc * 8);
    std::memcpy(hx + o_goff, goff.data(), ngx * 8);
    std::memcpy(hx + o_gfirst, gfirst.data(), ngx * 4);
    std::memcpy(hx + o_gser, gser.data(), ngx * 4);
    if (nn) {
        double *hu = (double *)(hx + o_ub);
        int *hr = (int *)(hx + o_run);
        int *hs = (int *)(hx + o_csid);
        for (size_t q = 0; q < nc; q++) {
            hu[q] = ub ? (*ub)[slot[q]] : INFINITY;
            hs[q] = req[slot[q]].first;
        }
        for (size_t q = 0; q < nk; q++) hr[q] = 0x7fffffff;  // ord_key above every float
    }
    xarg_.reserve(xbytes);
    char *dx = (char *)xarg_.p;
    HIPC(hipMemcpyAsync(dx, hx, xbytes, hipMemcpyHostToDevice, stream_));
    HIPC(hipEventRecord(ev_[6], stream_));
    for (const Launch &L : launches) {
        ExactArgs E{};
        E.x = B.x.as<float>();
        E.ser = B.ser.as<SerDev>();
        E.st64 = st64_.as<double2>();
        E.cand = (const int64_t *)(dx + o_cand);
        E.gfirst = (const int *)(dx + o_gfirst) + L.off;
        E.gser = (const int *)(dx + o_gser) + L.off;
        E.goff = (const int64_t *)(dx + o_goff) + L.off;
        E.outkey = (unsigned long long *)dx;
        E.st64n = st64n_.as<double2>();
        E.outkey2 = (unsigned long long *)dx + nk;
        E.ngroups = L.ngroups;
        E.m = m;
        E.m2 = m2;
        E.sym = excl_sym_ ? 1 : 0;
        E.colv = colv_.as<float4>();
        E.tmax = tmax_.as<float2>();
        E.ub = (const double *)(dx + o_ub);
        E.runub = (int *)(dx + o_run);
        E.lb = lb ? 1 : 0;
        E.tilelo = nnlo_.as<float>();
        E.tiles = (int)tiles;
        if (nn)
            launch_nn(E, L.nu, m2 != 0, lb, false, (unsigned)L.blocks, nn_lds_bytes(m, L.nu),
                      stream_);
        else
            launch_exact(E, L.nu, (unsigned)L.blocks, exact_lds_bytes(m, L.nu), stream_);
        HIPC(hipGetLastError());
    }
    if (nn && !lb && nc > 0 && tiles > 0) {
        // the double-precision pass over the candidates' qualifying super-tiles
        ExactArgs F{};
        F.x = B.x.as<float>();
        F.ser = B.ser.as<SerDev>();
        F.st64 = st64_.as<double2>();
        F.cand = (const int64_t *)(dx + o_cand);
        F.outkey = (unsigned long long *)dx;
        F.m = m;
        F.sym = excl_sym_ ? 1 : 0;
        F.ub = (const double *)(dx + o_ub);
        F.runub = (int *)(dx + o_run);
        F.tilelo = nnlo_.as<float>();
        F.tiles = (int)tiles;
        F.tb = (int)((tiles + 63) / 64);
        F.csid = (const int *)(dx + o_csid);
        launch_nn_fix(F, false, (unsigned)(nc * (size_t)F.tb), stream_);
        HIPC(hipGetLastError());
    }
    HIPC(hipEventRecord(ev_[7], stream_));
    HIPC(hipMemcpyAsync(hx, dx, nkeys * 8, hipMemcpyDeviceToHost, stream_));
    wait();
    collect_derive();
    // minimum over the slots, decoded from ord_key64
    auto value = [&](size_t base, size_t q) {
        unsigned long long k = ~0ull;
        for (int b = 0; b < EXACT_SLOTS; b++)
            k = std::min(k, ((const unsigned long long *)hx)[base + q * EXACT_SLOTS + b]);
        const unsigned long long bits = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
        double d;
        std::memcpy(&d, &bits, 8);
        return d;
    };
    for (size_t q = 0; q < slot.size(); q++) {
        out[slot[q]] = value(0, q);
        if (m2) (*out2)[slot[q]] = value(nk, q);
    }
    if (st) {
        float a = 0.f;
        HIPC(hipEventElapsedTime(&a, ev_[6], ev_[7]));
        st->kernel_ms[3] += a;
        st->launches[3] += (int64_t)launches.size();
        st->exact_refines += (int64_t)cand.size();
    }
    // A value k_nn did not produce (no super-tile qualified, or no valid column: never when
    // its error bound holds) is computed by k_exact over every column.
    if (nn) {
        std::vector<std::pair<int, int64_t>> again;
        std::vector<size_t> where;
        for (size_t q = 0; q < req.size(); q++)
            if (!std::isfinite(out[q]) || (out2 && m2 && !std::isfinite((*out2)[q]))) {
                again.push_back(req[q]);
                where.push_back(q);
            }
        if (!again.empty()) {
            std::vector<double> o, o2;
            exact_nn_ = false;
            try {
                exact(B, m, again, o, st, m2, out2 ? &o2 : nullptr);
            } catch (...) {
                exact_nn_ = true;
                throw;
            }
            exact_nn_ = true;
            for (size_t q = 0; q < where.size(); q++) {
                out[where[q]] = o[q];
                if (out2) (*out2)[where[q]] = o2[q];
            }
        }
    }
}

void Engine::topk_all(Batch &B, int m, std::vector<TopK> &tk, paddi_stats *st, int m2)
{
    for (;;) {
        std::vector<std::pair<int, int64_t>> req;
        std::vector<std::pair<int, size_t>> who;
        std::vector<double> ub;  // the scans' upper bound of each candidate's distance
        std::vector<size_t> need;
        for (int q = 0; q < B.nser; q++) {
            TopK &T = tk[(size_t)q];
            if (T.done) continue;
            if (T.advance(need))
                for (size_t c : need) {
                    const Cand &d = T.cands[c];
                    req.emplace_back(q, d.i);
                    who.emplace_back(q, c);
                    ub.push_back(d.a >= 3.0e38f ? INFINITY
                                                : (double)d.ahi + conv_ed2(m, d.ahi));
                }
        }
        if (req.empty()) break;
        std::vector<double> v, v2;
        exact(B, m, req, v, st, m2, &v2, &ub);
        for (size_t u = 0; u < who.size(); u++) {
            Cand &c = tk[(size_t)who[u].first].cands[who[u].second];
            c.exact = true;
            c.v = v[u];
            c.v2 = v2[u];
        }
    }
}

static double next_r_after_fail(int j, double r, double S)
{
    if (j == 0) return 0.5 * r;
    if (j < 5) return 0.99 * r;
    if (S > 0.0) return r - S;
    return 0.99 * r;
}

static void emit(std::vector<paddi_discord> &out, const std::vector<Cand> &acc, int m, int sid,
                 double r2)
{
    for (const Cand &c : acc)
        out.push_back(paddi_discord{c.i, m, sid, c.v >= 3.0e38 ? 3.4028235e38f : (float)c.v,
                                    (float)r2});
}

static double now_ms()
{
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

void Engine::att_mark(paddi_stats *st)
{
    if (!att_log_ || !st) return;
    collect_derive();
    att_snap_ = *st;
    att_t0_ = now_ms();
}

void Engine::att_end(int m, int attempt, double r2, int discords, paddi_stats *st)
{
    if (!att_log_ || !st) return;
    collect_derive();
    paddi_attempt_stats r{};
    r.m = m;
    r.attempt = attempt;
    r.r2 = (float)r2;
    r.discords = discords;
    r.survivors = st->survivors - att_snap_.survivors;
    r.ms_derive = st->kernel_ms[0] - att_snap_.kernel_ms[0];
    r.ms_scan = (st->kernel_ms[1] - att_snap_.kernel_ms[1]) +
                (st->kernel_ms[2] - att_snap_.kernel_ms[2]);
    r.ms_exact = st->kernel_ms[3] - att_snap_.kernel_ms[3];
    const double t = now_ms();
    r.ms_total = t - att_t0_;
    att_log_->push_back(r);
    att_snap_ = *st;
    att_t0_ = t;
}

// The plain form for series sid: src/main.cu:136-233 (guards as in the oracle).
void Engine::plain(Batch &B, int sid, int minL, int maxL, int step, int k, const paddi_opts &o,
                   std::vector<paddi_discord> &out, paddi_stats *st)
{
    const int nlen = (maxL - minL) / step + 1;
    std::vector<double> dk((size_t)nlen, 0.0);
    std::vector<std::vector<Cand>> surv;
    std::vector<TopK> tk((size_t)B.nser);
    std::vector<double> r2v((size_t)B.nser, -1.0);
    const int max_att = o.max_attempts > 0 ? o.max_attempts : 64;
    wread_valid_ = false;
    for (int j = 0; j < nlen; j++) {
        const int m = minL + j * step;
        att_mark(st);
        derive(B, m, st);
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
            r2v[(size_t)sid] = r2;
            pass(B, m, r2v, false, surv, st);
            tk[(size_t)sid].reset(std::move(surv[(size_t)sid]), r2, m, k);
            tk[(size_t)sid].spec = spec_;
            topk_all(B, m, tk, st);
            att++;
            const std::vector<Cand> &acc = tk[(size_t)sid].acc;
            att_end(m, att, r2, (int)acc.size(), st);
            if ((int)acc.size() >= k || r == 0.0) {
                emit(out, acc, m, sid, r2);
                dk[(size_t)j] = acc.empty() ? 0.0 : acc.back().v;
                break;
            }
            r = next_r_after_fail(j, r, S);
        }
    }
}

// The fast form: every series in every launch (see file header).
void Engine::fast(Batch &B, int minL, int maxL, int step, int k, const paddi_opts &o,
                  std::vector<paddi_discord> &out, paddi_stats *st)
{
    const int nlen = (maxL - minL) / step + 1;
    const int ns = B.nser;
    const int max_att = o.max_attempts > 0 ? o.max_attempts : 64;
    const bool learned = o.mode == PADDI_MODE_LEARNED;
    std::vector<std::vector<int64_t>> prev((size_t)ns);
    // With the previous discords evaluated at i only, their exact distances at the next
    // length come with the exact refinements of this one (k_exact's second length).
    const bool ahead = prev_shift_ == 0 && step <= EXACT_EXT;
    std::vector<std::vector<double>> prev_v((size_t)ns);  // prev's distances at this length
    std::vector<std::vector<paddi_discord>> res((size_t)ns);
    std::vector<std::vector<Cand>> surv;
    std::vector<TopK> tk((size_t)ns);
    wread_valid_ = false;
    // lengths on the device (from the second, up to the last but one, which has a next
    // length for its DUAL refinements); twice without progress turns it off for the call
    bool on_device = device_lengths_ > 0 && !att_log_ && k <= CHAIN_KMAX;
    int stalls = 0;
    for (int j = 0; j < nlen; j++) {
        if (on_device && j > 0 && j + 1 < nlen) {
            const int js = chain(B, minL, step, k, j, nlen - 1, ahead, prev, prev_v, res, st);
            stalls = js == j ? stalls + 1 : 0;
            if (stalls >= 2) on_device = false;
            j = js;
        }
        const int m = minL + j * step;
        const int m_next = (ahead && j + 1 < nlen) ? m + step : 0;
        att_mark(st);
        derive(B, m, st, m_next);
        first_len_ = j == 0;
        if (!wread_valid_ && seed_wit_ && m >= 8 && B.rows >= SAX_MIN_ROWS) {
            seed_witnesses(B, m);
            wread_valid_ = true;
        }
        std::vector<double> r2((size_t)ns, 0.0);
        std::vector<char> from_prev((size_t)ns, 0);
        // r from the previous length's discords, evaluated exactly at this length.  A
        // discord's window re-centres on its anomaly as m grows, so its start stays or moves
        // back: each one is evaluated at i, i-1 .. i-prev_shift_ and the best kept (any row's
        // exact NN distance is a lower bound of d_1); the minimum over the k discords is the
        // estimate of d_k (a pass that finds fewer than k retries lower).
        {
            std::vector<std::pair<int, int64_t>> req;
            std::vector<int> grp;  // request -> (series, discord) group
            int ngrp = 0;
            std::vector<int> gser;
            std::vector<double> mn((size_t)ns, INFINITY);
            std::vector<int> cnt((size_t)ns, 0);
            std::vector<char> done((size_t)ns, 0);
            for (int q = 0; q < ns; q++) {
                const std::vector<double> &pv = prev_v[(size_t)q];
                if ((int)prev[(size_t)q].size() < k || pv.size() != prev[(size_t)q].size())
                    continue;
                bool ok = true;
                for (double v : pv) ok &= v >= 0.0;
                if (!ok) continue;
                // the same rule as below, from the values computed ahead
                const int64_t Nq = B.n[(size_t)q] - m + 1;
                for (size_t u = 0; u < pv.size(); u++) {
                    if (prev[(size_t)q][u] >= Nq) continue;
                    cnt[(size_t)q]++;
                    if (pv[u] < 3.0e38) mn[(size_t)q] = std::min(mn[(size_t)q], pv[u]);
                }
                done[(size_t)q] = 1;
            }
            for (int q = 0; q < ns; q++)
                if ((int)prev[(size_t)q].size() >= k && !done[(size_t)q])
                    for (int64_t i : prev[(size_t)q]) {
                        const int64_t Nq = B.n[(size_t)q] - m + 1;
                        bool any = false;
                        for (int64_t d = -prev_shift_; d <= 0; d++) {
                            const int64_t j = i + d;
                            if (j < 0 || j >= Nq) continue;
                            req.emplace_back(q, j);
                            grp.push_back(ngrp);
                            any = true;
                        }
                        if (any) {
                            gser.push_back(q);
                            ngrp++;
                        }
                    }
            // lower bounds suffice for the range (k_nn in single precision only)
            std::vector<double> v;
            exact(B, m, req, v, st, 0, nullptr, nullptr, true);
            std::vector<double> best((size_t)ngrp, -INFINITY);
            for (size_t u = 0; u < req.size(); u++)
                if (v[u] < 3.0e38) best[(size_t)grp[u]] = std::max(best[(size_t)grp[u]], v[u]);
            for (int g = 0; g < ngrp; g++) {
                const int q = gser[(size_t)g];
                mn[(size_t)q] = std::min(mn[(size_t)q], best[(size_t)g] > -INFINITY
                                                            ? best[(size_t)g] : INFINITY);
                cnt[(size_t)q]++;
            }
            for (int q = 0; q < ns; q++)
                if (cnt[(size_t)q] >= k && std::isfinite(mn[(size_t)q]) && mn[(size_t)q] > 0.0) {
                    r2[(size_t)q] = mn[(size_t)q] * (1.0 - 1e-6);
                    from_prev[(size_t)q] = 1;
                }
        }
        // first length (or no usable previous result): learned proposal, raised to the
        // candidates' lower bound when that is higher; else the probe descent from 2m
        std::vector<double> prop((size_t)ns, 0.0), floor2((size_t)ns, 0.0);
        std::vector<char> need((size_t)ns, 0), found((size_t)ns, 0);
        bool need_prop = false;
        for (int q = 0; q < ns; q++) {
            need[(size_t)q] = !from_prev[(size_t)q];
            need_prop |= !from_prev[(size_t)q];
        }
        if (need_prop && learned) {
            propose_r2(B, m, k, prop);
            if (bootstrap_) bootstrap(B, m, k, need, floor2, found, st);
        }
        const std::vector<double> floor0 = floor2;  // for the trace
        for (int q = 0; q < ns; q++)
            if (!from_prev[(size_t)q]) {
                double r = 2.0 * m;
                if (learned) {
                    // k discords found among the candidate blocks: their k-th distance; else
                    // the proposal, raised to the candidates' bound
                    const double p = prop[(size_t)q], f = floor2[(size_t)q];
                    r = found[(size_t)q] ? f : (p > 0.0 ? std::max(p, f) : (f > 0.0 ? f : r));
                }
                r2[(size_t)q] = r;
            }
        // A series whose windows cannot hold k discords (k of them lie at least m apart)
        // ends every descent at r = 0, after up to max_attempts passes: go there at once.
        for (int q = 0; q < ns; q++)
            if (B.n[(size_t)q] - m < (int64_t)(k - 1) * m) r2[(size_t)q] = 0.0;
        std::vector<char> active((size_t)ns, 1);
        std::vector<int> att((size_t)ns, 0);
        for (int npass = 1;; npass++) {
            std::vector<double> r2p((size_t)ns, -1.0);
            bool any = false;
            double rmin = INFINITY;
            for (int q = 0; q < ns; q++)
                if (active[(size_t)q]) {
                    if (att[(size_t)q] >= max_att) r2[(size_t)q] = 0.0;
                    r2p[(size_t)q] = r2[(size_t)q];
                    rmin = std::min(rmin, r2[(size_t)q]);
                    any = true;
                }
            if (!any) break;
            int finished = 0;
            pass(B, m, r2p, true, surv, st);
            for (int q = 0; q < ns; q++)
                if (active[(size_t)q]) {
                    tk[(size_t)q].reset(std::move(surv[(size_t)q]), r2[(size_t)q], m, k);
                    tk[(size_t)q].spec = spec_;
                } else {
                    tk[(size_t)q].done = true;
                }
            topk_all(B, m, tk, st, m_next);
            for (int q = 0; q < ns; q++) {
                if (!active[(size_t)q]) continue;
                att[(size_t)q]++;
                const std::vector<Cand> &acc = tk[(size_t)q].acc;
                if ((int)acc.size() >= k || r2[(size_t)q] <= 0.0) {
                    if (debug_ && !from_prev[(size_t)q] && q < 8)
                        std::fprintf(stderr,
                                     "[paddi] first length m=%d series %d: proposal %.4g bound "
                                     "%.4g final r2 %.4g d_k %.4g attempts %d\n",
                                     m, q, prop.empty() ? 0.0 : prop[(size_t)q],
                                     floor0[(size_t)q], r2[(size_t)q],
                                     acc.empty() ? 0.0 : acc.back().v, att[(size_t)q]);
                    emit(res[(size_t)q], acc, m, q, r2[(size_t)q]);
                    finished += (int)acc.size();
                    prev[(size_t)q].clear();
                    prev_v[(size_t)q].clear();
                    for (const Cand &c : acc) {
                        prev[(size_t)q].push_back(c.i);
                        if (m_next) prev_v[(size_t)q].push_back(c.v2);
                    }
                    active[(size_t)q] = 0;
                } else {
                    // below the k-th distance the answer is the same: step r down, not
                    // below the candidates' bound unless that bound itself has failed
                    const double f = floor2.empty() ? 0.0 : floor2[(size_t)q];
                    const double nx = r2[(size_t)q] * (from_prev[(size_t)q] ? 0.9 : 0.85);
                    if (f > 0.0 && r2[(size_t)q] > f) {
                        r2[(size_t)q] = std::max(nx, f);
                    } else {
                        if (f > 0.0) floor2[(size_t)q] = 0.0;
                        r2[(size_t)q] = nx;
                    }
                    if (r2[(size_t)q] < 1e-9) r2[(size_t)q] = 0.0;
                }
            }
            att_end(m, npass, rmin, finished, st);
        }
        // this length's witnesses become the next length's read side
        wread_valid_ = true;
        witcur_ ^= 1;
    }
    first_len_ = false;
    for (int q = 0; q < ns; q++) out.insert(out.end(), res[(size_t)q].begin(), res[(size_t)q].end());
}

void Engine::seed_witnesses(Batch &B, int m)
{
    const int64_t rows = B.rows;
    const size_t tb = sax_sort_bytes(rows);
    if (tb == 0) throw HipError{PADDI_ERR_DEVICE, "radix sort sizing failed"};
    wit_[0].reserve((size_t)rows * 4);
    wit_[1].reserve((size_t)rows * 4);
    sax_k_.reserve((size_t)rows * 8);
    sax_k2_.reserve((size_t)rows * 8);
    sax_v_.reserve((size_t)rows * 4);
    sax_v2_.reserve((size_t)rows * 4);
    sax_tmp_.reserve(tb);
    HIPC(launch_sax_witnesses(B.x.as<float>(), B.ser.as<SerDev>(), B.bser.as<int>(),
                              colv_.as<float4>(), rows, m, B.nser,
                              sax_k_.as<unsigned long long>(), sax_k2_.as<unsigned long long>(),
                              sax_v_.as<int>(), sax_v2_.as<int>(), sax_tmp_.p, sax_tmp_.bytes,
                              wit_[witcur_ ^ 1].as<int>(), stream_));
}

// ------------------------------------------------------------------------------------------
// The length loop on the device (ChainArgs in kernels.h).  On entry every series holds the k
// discords of length j0 - 1 in prev and their exact distances at length j0 in prev_v, and the
// last derive was derive(m_{j0-1}, m_{j0}) (its second buffer, st64n_, holds m_{j0}'s
// statistics).  Length j reads the statistics buffer S[(j - j0) & 1] (S0 = st64n_, S1 = st64_
// at entry) and k_derive writes the next length's into the other, so two graphs per length
// class cover the loop.  On return the host state is that of the host path about to run the
// returned length: witness parity, statistics buffers, prev / prev_v.
int Engine::chain(Batch &B, int minL, int step, int k, int j0, int jend, bool ahead,
                  std::vector<std::vector<int64_t>> &prev, std::vector<std::vector<double>> &prev_v,
                  std::vector<std::vector<paddi_discord>> &res, paddi_stats *st)
{
    const int ns = B.nser;
    const int m0 = minL + j0 * step;
    const int shift = ahead ? 0 : prev_shift_;
    if (jend <= j0 || k < 1 || k > CHAIN_KMAX || !wread_valid_ ||
        (ahead && (derive_m2_ != m0 || step > EXACT_EXT ||
                   st64n_.bytes < (size_t)B.rows * sizeof(double2))))
        return j0;
    for (int q = 0; q < ns; q++)
        if ((int)prev[(size_t)q].size() != k ||
            (ahead && prev_v[(size_t)q].size() != (size_t)k))
            return j0;
    const int64_t rows = B.rows, nwords = B.nblocks;
    const int mlast = minL + (jend - 1) * step;
    const int cls_max = (int)std::min<int64_t>(round_up(mlast, 64), MAXM);
    // candidates per round: the speculation width, at least two where they fit (one candidate
    // at two lengths runs one wave per SIMD, see launch_exact)
    const int umax = std::min(CHAIN_NU, exact_nn_ ? nn_max_u(cls_max) : exact_max_u(cls_max));
    const int nu = std::max(std::min(2, umax), std::min(spec_, umax));
    const int rounds = std::max(1, std::min(device_rounds_, CHAIN_ROUNDS));
    // range launch: gr group slots of nur rows per series, the rows split evenly
    const int gr = (k * (shift + 1) + nu - 1) / nu;
    const int nur = (k * (shift + 1) + gr - 1) / gr;
    const int64_t scap = std::min<int64_t>(rows, int64_t(1) << 20);
    const int nl = jend - j0;

    // the pass's buffers (as fill_pass), the loop's own
    alive_.reserve((size_t)nwords * 8);
    best_.reserve((size_t)rows * 4);
    besth_.reserve((size_t)rows * 4);
    bestl_.reserve((size_t)rows * 4);
    blist_.reserve((size_t)nwords * 4);
    pout_.reserve(PASS_OUT_HEAD + (size_t)rows * sizeof(Surv));
    wit_[0].reserve((size_t)rows * 4);
    wit_[1].reserve((size_t)rows * 4);
    par_.reserve(sizeof(PassDev) + (size_t)ns * sizeof(SerDev));
    colv_.reserve((size_t)rows * sizeof(float4));
    tmax_.reserve((size_t)nwords * sizeof(float2));
    rho_.reserve((size_t)nwords * sizeof(float));
    ch_ctl_.reserve(sizeof(ChainCtl));
    ch_ser_.reserve((size_t)ns * sizeof(ChainSer));
    ch_res_.reserve((size_t)nl * ns * sizeof(ChainRes));
    ch_xv_.reserve((size_t)scap * 8);
    ch_xv2_.reserve((size_t)scap * 8);
    ch_cand_.reserve((size_t)ns * CHAIN_NU * 8);
    ch_cnt_.reserve((size_t)ns * 4);
    ch_ub_.reserve((size_t)ns * CHAIN_NU * 8);
    int64_t ttiles = 1;  // super-tiles per series at the first (longest) length
    for (int q = 0; q < ns; q++)
        ttiles = std::max<int64_t>(ttiles, (B.n[(size_t)q] - m0 + 1 + 255) / 256);
    ch_tilelo_.reserve((size_t)ns * CHAIN_NU * (size_t)ttiles * 4);
    ch_runub_.reserve((size_t)ns * CHAIN_NU * EXACT_SLOTS * 4);
    ch_key_.reserve((size_t)ns * CHAIN_NU * EXACT_SLOTS * 8);
    ch_key2_.reserve((size_t)ns * CHAIN_NU * EXACT_SLOTS * 8);
    ch_goff_.reserve((size_t)(ns + 1) * 8);
    ch_candr_.reserve((size_t)ns * gr * nur * 8);
    ch_cntr_.reserve((size_t)ns * gr * 4);
    ch_keyr_.reserve((size_t)ns * gr * nur * EXACT_SLOTS * 8);
    st64_.reserve((size_t)rows * sizeof(double2));
    // pinned: control block (the graphs' download target) | series state | results / goff
    const size_t o_ser = 256, o_res = o_ser + round_up((int64_t)ns * sizeof(ChainSer), 256);
    const size_t pbytes =
        o_res + std::max((size_t)nl * ns * sizeof(ChainRes), (size_t)(ns + 1) * 8);
    if (pbytes > pin_ch_bytes_) {
        if (t_noalloc) refuse_growth("the pinned device-loop staging");
        HIPC(hipStreamSynchronize(stream_));
        drop_graphs();
        if (pin_ch_) HIPC(hipHostFree(pin_ch_));
        pin_ch_ = nullptr;
        pin_ch_bytes_ = 0;
        HIPC(hipHostMalloc(&pin_ch_, pbytes, hipHostMallocDefault));
        pin_ch_bytes_ = pbytes;
    }
    char *hp = (char *)pin_ch_;
    ChainCtl *hc = (ChainCtl *)hp;
    ChainSer *hs = (ChainSer *)(hp + o_ser);
    ChainRes *hr = (ChainRes *)(hp + o_res);
    int64_t *hg = (int64_t *)(hp + o_res);

    ChainCtl *dctl = ch_ctl_.as<ChainCtl>();
    ChainArgs C{};
    C.ctl = dctl;
    C.cs = ch_ser_.as<ChainSer>();
    C.res = ch_res_.as<ChainRes>();
    C.ser = B.ser.as<SerDev>();
    C.pser = (SerDev *)((char *)par_.p + sizeof(PassDev));
    C.par = par_.as<PassDev>();
    C.pctl = pout_.as<PassCtl>();
    C.surv = (const Surv *)((char *)pout_.p + PASS_OUT_HEAD);
    C.xv = ch_xv_.as<double>();
    C.xv2 = ch_xv2_.as<double>();
    C.cand = ch_cand_.as<int64_t>();
    C.cnt = ch_cnt_.as<int>();
    C.ub = ch_ub_.as<double>();
    C.runub = ch_runub_.as<int>();
    C.outkey = ch_key_.as<unsigned long long>();
    C.outkey2 = ch_key2_.as<unsigned long long>();
    C.scap = scap;
    C.WA = select_window_;
    C.max_chunk = refine_chunk_;
    C.min_chunk = min_chunk_;
    C.items = refine_waves_;
    C.witness_max = witness_tiles_;
    C.dbudget = direct_budget_;
    C.drows = direct_rows_;
    C.sym = excl_sym_ ? 1 : 0;
    C.nser = ns;
    C.k = k;
    C.nu = nu;
    C.spec = spec_;
    C.rounds = rounds;
    C.minL = minL;
    C.step = step;
    C.rrounds = refine_rounds_;
    C.rorder = refine_order_ == 1 ? 1 : 0;  // the loop never runs a call's first length
    C.shift = shift;
    C.gr = gr;
    C.nur = nur;
    C.candr = ch_candr_.as<int64_t>();
    C.cntr = ch_cntr_.as<int>();
    C.keyr = ch_keyr_.as<unsigned long long>();
    ScanPtrs Q{};
    Q.x = B.x.as<float>();
    Q.ser = C.pser;
    Q.bser = B.bser.as<int>();
    Q.colv = colv_.as<float4>();
    Q.tmax = tmax_.as<float2>();
    Q.rho = rho_.as<float>();
    Q.alive = alive_.as<uint64_t>();
    Q.best = best_.as<int>();
    Q.besth = besth_.as<int>();
    Q.bestl = bestl_.as<int>();
    Q.wit0 = wit_[0].as<int>();
    Q.wit1 = wit_[1].as<int>();
    Q.pairs = &pout_.as<PassCtl>()->pairs[0][0];
    // k_exact_chain's blocks: series q's column chunks at the loop's first (longest) windows
    hg[0] = 0;
    for (int q = 0; q < ns; q++) {
        const int64_t N = std::max<int64_t>(0, B.n[(size_t)q] - m0 + 1);
        hg[q + 1] = hg[q] + (N + EXACT_COLS - 1) / EXACT_COLS;
    }
    const int64_t xblocks = hg[ns];
    // statistics buffers: with the range from the previous exact rounds, length j reads
    // S[(j - j0) & 1] and k_derive writes the next length's into the other; with a shifted
    // range every length's k_derive writes st64_
    void *const S[2] = {ahead ? st64n_.p : st64_.p, ahead ? st64_.p : st64_.p};

    // graphs: one per (statistics buffer read, length class)
    auto graph_for = [&](int par, int cls) -> hipGraphExec_t {
        const size_t lds_scan = scan_lds_bytes(cls), lds_ex = exact_lds_bytes(cls, nu);
        ExactArgs E{};
        E.x = B.x.as<float>();
        E.ser = B.ser.as<SerDev>();
        E.st64 = (const double2 *)S[par];
        E.st64n = (const double2 *)(ahead ? S[par ^ 1] : st64n_.p);
        E.cand = C.cand;
        E.goff = ch_goff_.as<int64_t>();
        E.outkey = C.outkey;
        E.outkey2 = C.outkey2;
        E.ngroups = ns;
        E.sym = C.sym;
        E.dev_m = &dctl->m;
        E.dev_count = C.cnt;
        E.gslots = 1;
        E.colv = colv_.as<float4>();
        E.tmax = tmax_.as<float2>();
        E.ub = C.ub;
        E.runub = C.runub;
        E.tilelo = ch_tilelo_.as<float>();
        E.tiles = (int)ttiles;
        E.tb = (int)((ttiles + 63) / 64);
        E.nu_dev = nu;
        ExactArgs Er = E;  // the shifted range's launch: lower bounds
        Er.cand = C.candr;
        Er.outkey = C.keyr;
        Er.outkey2 = C.keyr;
        Er.dev_count = C.cntr;
        Er.gslots = gr;
        Er.ub = nullptr;
        Er.runub = nullptr;
        Er.lb = 1;
        const bool nn = exact_nn_;
        const int64_t refine_grid = std::max<int64_t>(refine_waves_ + refine_waves_ / 2, nwords + 64);
        std::string sig;
        auto put = [&](const void *p, size_t n) { sig.append((const char *)p, n); };
        put(&C, sizeof C);
        put(&E, sizeof E);
        put(&Er, sizeof Er);
        put(&Q, sizeof Q);
        const int64_t v[] = {rows, nwords, (int64_t)lds_scan, (int64_t)lds_ex, (int64_t)exact_nn_,
                             xblocks,  refine_grid,     (int64_t)(uintptr_t)pin_ch_,
                             (int64_t)(uintptr_t)B.P.p, (int64_t)(uintptr_t)B.refs.p,
                             (int64_t)(uintptr_t)par_.p};
        put(v, sizeof v);
        ChainGraph &G = chain_graphs_[{(uintptr_t)S[par], cls}];
        if (G.exec && G.sig == sig) return G.exec;
        if (G.exec) HIPC(hipGraphExecDestroy(G.exec));
        G.exec = nullptr;
        hipGraph_t g = nullptr;
        HIPC(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed));
        try {
            if (!ahead) hipLaunchKernelGGL(k_chain_rreq, dim3(ns), dim3(64), 0, stream_, C);
            hipLaunchKernelGGL(k_derive, dim3(grid1(rows, 256 * DERIVE_RPT)), dim3(256), 0,
                               stream_, B.x.as<float>(), B.ser.as<SerDev>(), B.bser.as<int>(),
                               B.P.as<double2>(), B.refs.as<double4>(), rows, 0, colv_.as<float4>(),
                               ahead ? (double2 *)nullptr : (double2 *)S[par],
                               tmax_.as<float2>(), 0,
                               ahead ? (double2 *)S[par ^ 1] : st64n_.as<double2>(), &dctl->m);
            if (!ahead) {
                if (xblocks > 0 && nn)
                    launch_nn(Er, nur, false, true, true, (unsigned)(xblocks * gr),
                              nn_lds_bytes(cls, nur), stream_);
                else if (xblocks > 0)
                    launch_exact_chain(Er, nur, false, (unsigned)(xblocks * gr),
                                       exact_lds_bytes(cls, nur), stream_);
                hipLaunchKernelGGL(k_chain_range, dim3(ns), dim3(64), 0, stream_, C);
            }
            enqueue_pass(Q, lds_scan, 0, rows, nwords, 0, false, nullptr, false);
            for (int r = 0; r < rounds; r++) {
                hipLaunchKernelGGL(k_chain_topk, dim3(ns), dim3(64), 0, stream_, C, r);
                if (xblocks > 0 && nn) {
                    launch_nn(E, nu, ahead, false, true, (unsigned)xblocks,
                              nn_lds_bytes(cls, nu), stream_);
                    launch_nn_fix(E, true, (unsigned)(ns * nu * E.tb), stream_);
                }
                else if (xblocks > 0)
                    launch_exact_chain(E, nu, ahead, (unsigned)xblocks, lds_ex, stream_);
            }
            hipLaunchKernelGGL(k_chain_topk, dim3(ns), dim3(64), 0, stream_, C, rounds);
            if (ns > 1)  // one series: the final k_chain_topk starts the next length itself
                hipLaunchKernelGGL(k_chain_next, dim3(1), dim3(1024), 0, stream_, C);
            HIPC(hipGetLastError());
        } catch (...) {
            (void)hipStreamEndCapture(stream_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIPC(hipStreamEndCapture(stream_, &g));
        const hipError_t e = hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPC(e);
        G.sig = sig;
        return G.exec;
    };
    std::vector<hipGraphExec_t> gl((size_t)nl);
    for (int j = j0; j < jend; j++) {
        const int m = minL + j * step;
        gl[(size_t)(j - j0)] = graph_for(ahead ? (j - j0) & 1 : 0,
                                         (int)std::min<int64_t>(round_up(m, 64), MAXM));
    }

    // initial state, then the lengths
    std::memset(hc, 0, sizeof(ChainCtl));
    hc->j = j0 - 1;  // k_chain_next below starts length j0 (m = 0: no length to close)
    hc->j0 = j0;
    hc->wit0 = witcur_;
    for (int q = 0; q < ns; q++) {
        ChainSer &c = hs[q];
        std::memset(&c, 0, sizeof(ChainSer));
        c.nprev = k;
        for (int u = 0; u < k; u++) {
            c.prev[u] = prev[(size_t)q][(size_t)u];
            c.pv[u] = ahead ? prev_v[(size_t)q][(size_t)u] : 0.0;
        }
    }
    HIPC(hipMemcpyAsync(dctl, hc, sizeof(ChainCtl), hipMemcpyHostToDevice, stream_));
    HIPC(hipMemcpyAsync(ch_ser_.p, hs, (size_t)ns * sizeof(ChainSer), hipMemcpyHostToDevice, stream_));
    HIPC(hipMemcpyAsync(ch_goff_.p, hg, (size_t)(ns + 1) * 8, hipMemcpyHostToDevice, stream_));
    // the first length's start; every later one is started by its predecessor's last kernel
    hipLaunchKernelGGL(k_chain_next, dim3(1), dim3(1024), 0, stream_, C);
    HIPC(hipGetLastError());
    wait();  // the staging is reused for the results
    const int every = std::max(1, device_lengths_);
    auto poll = [&] {  // the control block, once per synchronisation
        HIPC(hipMemcpyAsync(pin_ch_, dctl, sizeof(ChainCtl), hipMemcpyDeviceToHost, stream_));
        wait();
    };
    for (int j = j0; j < jend; j++) {
        HIPC(hipGraphLaunch(gl[(size_t)(j - j0)], stream_));
        if ((j - j0 + 1) % every == 0 && j + 1 < jend) {
            poll();
            if (hc->halt) break;
        }
    }
    poll();
    const ChainCtl cc = *hc;
    const int jstop = cc.halt ? cc.halt - 1 : jend;
    if (debug_)
        std::fprintf(stderr, "[paddi] device lengths %d..%d: finished %d (halt %d)\n", j0,
                     jend - 1, jstop - j0, cc.halt);
    if (jstop > j0) {
        HIPC(hipMemcpyAsync(hr, ch_res_.p, (size_t)(jstop - j0) * ns * sizeof(ChainRes),
                            hipMemcpyDeviceToHost, stream_));
        wait();
        for (int q = 0; q < ns; q++) {
            for (int j = j0; j < jstop; j++) {
                const ChainRes &R = hr[(size_t)(j - j0) * ns + q];
                const int m = minL + j * step;
                for (int a = 0; a < R.n; a++)
                    res[(size_t)q].push_back(paddi_discord{
                        R.i[a], m, q, R.v[a] >= 3.0e38 ? 3.4028235e38f : (float)R.v[a],
                        (float)R.r2});
            }
            const ChainRes &L = hr[(size_t)(jstop - 1 - j0) * ns + q];
            prev[(size_t)q].assign(L.i, L.i + L.n);
            if (ahead)
                prev_v[(size_t)q].assign(L.v2, L.v2 + L.n);
            else
                prev_v[(size_t)q].clear();
        }
    }
    if (st) {
        st->attempts += (int64_t)cc.stats[0];
        st->survivors += (int64_t)cc.stats[1];
        st->exact_refines += (int64_t)cc.stats[2];
        st->pairs += (int64_t)(cc.stats[3] + cc.stats[4]);
        st->scan_pairs[0] += (int64_t)cc.stats[3];
        st->scan_pairs[1] += (int64_t)cc.stats[4];
        st->live_pairs[0] += (int64_t)cc.stats[5];
        st->live_pairs[1] += (int64_t)cc.stats[6];
    }
    // the host path resumes at jstop: its witness parity, m_jstop's statistics in st64n_
    witcur_ = (cc.wit0 + (jstop - j0)) & 1;
    wread_valid_ = true;
    if (ahead) {
        if ((jstop - j0) & 1) st64_.swap(st64n_);
        derive_m2_ = minL + jstop * step;
    } else {
        derive_m2_ = 0;  // st64_ holds the last length's statistics; the host derives again
    }
    return jstop;
}

void Engine::run(const std::vector<const Series *> &ss, int minL, int maxL, int step, int k,
                 const paddi_opts &o, std::vector<paddi_discord> &out, paddi_stats *st)
{
    derive_st_ = nullptr;
    if (o.mode == PADDI_MODE_PLAIN) {
        // the reference's method looped over series, each alone
        for (size_t q = 0; q < ss.size(); q++) {
            build_batch(batch_, {ss[q]}, minL);
            prefix(batch_);
            std::vector<paddi_discord> r;
            plain(batch_, 0, minL, maxL, step, k, o, r, st);
            for (paddi_discord &d : r) d.series = (int32_t)q;
            out.insert(out.end(), r.begin(), r.end());
        }
        collect_derive();
        return;
    }
    if (o.mode == PADDI_MODE_LEARNED) set_model_path(o.model_path);
    build_batch(batch_, ss, minL);
    prefix(batch_);
    fast(batch_, minL, maxL, step, k, o, out, st);
    collect_derive();
}

void Engine::range(Batch &B, int m, float r2, int32_t *is_discord, float *nn_dist2)
{
    const int64_t N = B.n[0] - m + 1;
    std::vector<std::vector<Cand>> surv;
    pass(B, m, {(double)r2}, false, surv, nullptr);
    std::fill(is_discord, is_discord + N, 0);
    std::fill(nn_dist2, nn_dist2 + N, 3.4028235e38f);
    // Rows within the scans' error bound of r2 are decided in double precision.  So are the
    // small distances of range discords: the single-precision value's error is a few 1e-6 of
    // 2m whatever the distance, more than 1e-4 relative below about m/5 (up to
    // RANGE_EXACT_CAP such rows per call; beyond that they keep the single-precision value).
    std::vector<std::pair<int, int64_t>> amb, low;
    for (const Cand &c : surv[0]) {
        if (c.a >= 3.0e38f || (double)c.alo - conv_ed2(m, c.alo) > (double)r2) {
            is_discord[c.i] = 1;
            nn_dist2[c.i] = c.a;
            if ((double)c.a < 0.2 * m) low.emplace_back(0, c.i);
        } else if ((double)c.ahi + conv_ed2(m, c.ahi) > (double)r2) {
            amb.emplace_back(0, c.i);
        }
    }
    const size_t namb = amb.size();
    if (low.size() <= RANGE_EXACT_CAP) amb.insert(amb.end(), low.begin(), low.end());
    if (!amb.empty()) {
        std::vector<double> v;
        exact(B, m, amb, v, nullptr);
        for (size_t q = 0; q < amb.size(); q++)
            if (q >= namb || v[q] > (double)r2) {
                is_discord[amb[q].second] = 1;
                nn_dist2[amb[q].second] = v[q] >= 3.0e38 ? 3.4028235e38f : (float)v[q];
            }
    }
}

// Every device buffer of the call is a slice of the scratch array, bound before any work is
// queued (ScratchScope), the series is copied from the caller's memory straight into its
// slice, and the pinned staging was sized when the context was created.
void Engine::range_noalloc(const float *T, int64_t n, int m, float r2, int32_t *is_discord,
                           float *nn_dist2)
{
    HIPC(hipSetDevice(device_));
    Series s;   // describes the series (its n); the data come from T
    s.n = n;
    Batch B;    // outlives the scope, which gives B's buffers back on exit
    ScratchScope S(*this);
    const int64_t xo = round_up(n + XPAD, 64), npb = (n + PB - 1) / PB + 1;
    const int64_t rows = series_rows(n, m), nblocks = rows / 64;
    S.bind(B.x, (size_t)xo * 4);
    S.bind(B.P, (size_t)npb * (PB + 1) * sizeof(double2));
    S.bind(B.refs, (size_t)npb * sizeof(double4));
    S.bind(B.ser, sizeof(SerDev) * 4);
    S.bind(B.bser, (size_t)nblocks * 4);
    S.bind(B.pser, (size_t)npb * 4);
    S.bind(flag_, 64);
    S.bind(colv_, (size_t)rows * sizeof(float4));
    S.bind(st64_, (size_t)rows * sizeof(double2));
    S.bind(tmax_, (size_t)nblocks * sizeof(float2));
    S.bind(rho_, (size_t)nblocks * sizeof(float));
    S.bind(alive_, (size_t)nblocks * 8);
    S.bind(best_, (size_t)rows * 4);
    S.bind(besth_, (size_t)rows * 4);
    S.bind(bestl_, (size_t)rows * 4);
    S.bind(blist_, (size_t)nblocks * 4);
    S.bind(pout_, PASS_OUT_HEAD + (size_t)rows * sizeof(Surv));
    S.bind(wit_[0], (size_t)rows * 4);   // a pass without witnesses never touches them
    S.share(wit_[1], wit_[0]);
    S.bind(par_, sizeof(PassDev) + sizeof(SerDev));
    S.bind(xarg_, exact_arg_bytes(EXACT_NOALLOC_CHUNK, false));
    const float *const src[1] = {T};
    build_batch(B, {&s}, m, src);
    prefix(B);
    derive(B, m, nullptr);
    range(B, m, r2, is_discord, nn_dist2);
}

}  // namespace paddi

// ------------------------------------------------------------------------------------------
// C ABI

static bool valid_lengths(int64_t n, int32_t minL, int32_t maxL, int32_t step, int32_t k)
{
    return minL >= 4 && maxL >= minL && step >= 1 && k >= 1 && n >= (int64_t)maxL + 1;
}

// The row space of a call fits the engine's 32-bit row indices (MAX_ROWS).
static bool rows_fit(const int64_t *n, int32_t ns, int32_t minL)
{
    int64_t rows = 0;
    for (int32_t q = 0; q < ns; q++) {
        if (n[q] >= paddi::MAX_ROWS) return false;
        rows += paddi::series_rows(n[q], minL);
        if (rows >= paddi::MAX_ROWS) return false;
    }
    return true;
}

extern "C" {

const char *paddi_status_str(paddi_status s) { return paddi::status_str(s); }
const char *paddi_version(void) { return "paddi-mi355x 0.2 (gfx950)"; }

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

paddi_status paddi_get_tunables(paddi_ctx *ctx, paddi_tunables *out)
{
    if (!ctx || !out) return PADDI_ERR_INVALID;
    const paddi::Engine &E = *ctx->eng;
    *out = paddi_tunables{E.select_window_, E.refine_chunk_, E.min_chunk_, E.refine_waves_,
                          E.witness_tiles_, E.spec_, E.use_graph_ ? 1 : 0, E.prev_shift_,
                          E.bootstrap_ ? 1 : 0, E.bound_cands_, E.direct_budget_,
                          E.direct_rows_, E.device_lengths_, E.device_rounds_,
                          E.refine_rounds_, E.seed_wit_ ? 1 : 0, E.refine_order_};
    return PADDI_OK;
}

paddi_status paddi_set_exclusion(paddi_ctx *ctx, int32_t rule)
{
    if (!ctx || (rule != PADDI_EXCL_CPU && rule != PADDI_EXCL_SYMMETRIC))
        return PADDI_ERR_INVALID;
    ctx->eng->excl_sym_ = rule == PADDI_EXCL_SYMMETRIC;
    return PADDI_OK;
}

paddi_status paddi_get_exclusion(paddi_ctx *ctx, int32_t *rule)
{
    if (!ctx || !rule) return PADDI_ERR_INVALID;
    *rule = ctx->eng->excl_sym_ ? PADDI_EXCL_SYMMETRIC : PADDI_EXCL_CPU;
    return PADDI_OK;
}

paddi_status paddi_set_tunables(paddi_ctx *ctx, const paddi_tunables *t)
{
    if (!ctx || !t) return PADDI_ERR_INVALID;
    if (t->select_window < 64 || t->refine_chunk < 64 || t->min_chunk < 64 ||
        t->min_chunk > t->refine_chunk || t->refine_items < 1 || t->witness_tiles < 0 ||
        t->witness_tiles > 64 || t->speculate < 1 || t->speculate > 64 || t->prev_shift < 0 ||
        t->prev_shift > 4 || t->bootstrap < 0 || t->bootstrap > 1 ||
        t->bound_candidates < 0 || t->bound_candidates > 256 || t->direct_budget < 0 ||
        t->direct_rows < 0 || t->direct_rows > paddi::DIRECT_ROWS || t->device_lengths < 0 ||
        t->device_rounds < 1 || t->device_rounds > paddi::CHAIN_ROUNDS ||
        t->refine_rounds < 3 || t->refine_rounds > paddi::REFINE_ROUNDS_MAX ||
        t->seed_witnesses < 0 || t->seed_witnesses > 1 || t->refine_order < 0 ||
        t->refine_order > 2)
        return PADDI_ERR_INVALID;
    paddi::Engine &E = *ctx->eng;
    E.select_window_ = (t->select_window + 63) / 64 * 64;
    E.refine_chunk_ = (t->refine_chunk + 63) / 64 * 64;
    E.min_chunk_ = (t->min_chunk + 63) / 64 * 64;
    E.refine_waves_ = t->refine_items;
    E.witness_tiles_ = t->witness_tiles;
    E.spec_ = t->speculate;
    E.use_graph_ = t->use_graph != 0;
    E.prev_shift_ = t->prev_shift;
    E.bootstrap_ = t->bootstrap != 0;
    E.bound_cands_ = t->bound_candidates > 0 ? t->bound_candidates : paddi::BOUND_CANDIDATES;
    E.direct_budget_ = t->direct_budget;
    E.direct_rows_ = t->direct_rows;
    E.device_lengths_ = t->device_lengths;
    E.device_rounds_ = t->device_rounds;
    E.refine_rounds_ = t->refine_rounds;
    E.seed_wit_ = t->seed_witnesses != 0;
    E.refine_order_ = t->refine_order;
    return PADDI_OK;
}

paddi_status paddi_series_upload(paddi_ctx *ctx, const float *T, int64_t n, paddi_series **out)
{
    if (!ctx || !T || !out || n < 8) return PADDI_ERR_INVALID;
    if (n >= paddi::MAX_ROWS) return PADDI_ERR_UNSUPPORTED;
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
    if (m > paddi::PB || !rows_fit(&n, 1, m)) return PADDI_ERR_UNSUPPORTED;
    return guarded(ctx, [&] {
        paddi::Engine &E = *ctx->eng;
        std::unique_ptr<paddi::Series> s(E.upload(T, n));
        paddi::Batch B;
        E.build_batch(B, {s.get()}, m);
        E.prefix(B);
        E.derive(B, m, nullptr);
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
    if (m > paddi::MAXM || !rows_fit(&n, 1, m)) return PADDI_ERR_UNSUPPORTED;
    return guarded(ctx, [&] {
        paddi::Engine &E = *ctx->eng;
        std::unique_ptr<paddi::Series> s(E.upload(T, n));
        paddi::Batch B;
        E.build_batch(B, {s.get()}, m);
        E.prefix(B);
        E.derive(B, m, nullptr);
        E.range(B, m, r2, is_discord, nn_dist2);
        return PADDI_OK;
    });
}

paddi_status paddi_range_discords_noalloc(paddi_ctx *ctx, const float *T, int64_t n, int32_t m,
                                          float r2, int32_t *is_discord, float *nn_dist2)
{
    if (!ctx || !T || !is_discord || !nn_dist2 || m < 4 || n < (int64_t)m + 1 || !(r2 >= 0.f))
        return PADDI_ERR_INVALID;
    if (m > paddi::MAXM || n > PADDI_NOALLOC_MAX_POINTS) return PADDI_ERR_UNSUPPORTED;
    return guarded(ctx, [&] {
        ctx->eng->range_noalloc(T, n, m, r2, is_discord, nn_dist2);
        return PADDI_OK;
    });
}

paddi_status paddi_discords_series(paddi_ctx *ctx, paddi_series *const *series, int32_t nseries,
                                   int32_t minL, int32_t maxL, int32_t step, int32_t top_k,
                                   const paddi_opts *opts, paddi_discord *out, int64_t cap,
                                   int64_t *count, paddi_stats *stats)
{
    if (!ctx || !series || nseries <= 0 || !count || (!out && cap > 0)) return PADDI_ERR_INVALID;
    std::vector<const paddi::Series *> ss;
    std::vector<int64_t> ns;
    for (int32_t q = 0; q < nseries; q++) {
        if (!series[q] || series[q]->ctx != ctx ||
            !valid_lengths(series[q]->s->n, minL, maxL, step, top_k))
            return PADDI_ERR_INVALID;
        ss.push_back(series[q]->s);
        ns.push_back(series[q]->s->n);
    }
    if (maxL > paddi::MAXM || !rows_fit(ns.data(), nseries, minL)) return PADDI_ERR_UNSUPPORTED;
    paddi_opts o{};
    if (opts) o = *opts;
    if (o.mode < PADDI_MODE_PLAIN || o.mode > PADDI_MODE_LEARNED) return PADDI_ERR_INVALID;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    const auto t0 = std::chrono::steady_clock::now();
    paddi_status st = guarded(ctx, [&] {
        std::vector<paddi_discord> res;
        ctx->eng->run(ss, minL, maxL, step, top_k, o, res, stats);
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
    // refused before any series is read or uploaded
    if (maxL > paddi::MAXM || !rows_fit(n, nseries, minL)) return PADDI_ERR_UNSUPPORTED;
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

paddi_status paddi_discords_timed(paddi_ctx *ctx, const float *T, int64_t n, int32_t minL,
                                  int32_t maxL, int32_t step, int32_t top_k,
                                  const paddi_opts *opts, paddi_discord *out, int64_t cap,
                                  int64_t *count, paddi_attempt_stats *per_attempt,
                                  int64_t cap_attempts, int64_t *n_attempts)
{
    if (!ctx || !n_attempts || cap_attempts < 0 || (!per_attempt && cap_attempts > 0) ||
        !valid_lengths(n, minL, maxL, step, top_k))
        return PADDI_ERR_INVALID;
    std::vector<paddi_attempt_stats> log;
    paddi_stats stats;
    ctx->eng->att_log_ = &log;
    const paddi_status st = paddi_discords(ctx, T, n, minL, maxL, step, top_k, opts, out, cap,
                                           count, &stats);
    ctx->eng->att_log_ = nullptr;
    *n_attempts = (int64_t)log.size();
    const size_t w = std::min<size_t>(log.size(), (size_t)cap_attempts);
    if (w) std::memcpy(per_attempt, log.data(), w * sizeof(paddi_attempt_stats));
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

int32_t paddi_has_mode(int32_t mode)
{
    if (mode == PADDI_MODE_PLAIN || mode == PADDI_MODE_BATCHED) return 1;
    if (mode == PADDI_MODE_LEARNED) return paddi::learned_available() ? 1 : 0;
    return 0;
}

}  // extern "C"
