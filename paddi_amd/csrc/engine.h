// Host engine of the discord search (one HIP device, one stream).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <stdexcept>
#include <stdint.h>
#include <string>
#include <utility>
#include <vector>

#include "../../include/paddi.h"
#include "kernels.h"
#include "mlp.h"

namespace paddi {

// An error inside the engine: the status the C ABI returns for it and a message.  Every
// failure the engine raises is one of these (or std::bad_alloc from the host allocator).
struct HipError : std::runtime_error {
    paddi_status st;
    HipError(paddi_status s, const std::string &what) : std::runtime_error(what), st(s) {}
};

#define HIPC(expr)                                                                       \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw ::paddi::HipError(e_ == hipErrorOutOfMemory ? PADDI_ERR_NOMEM          \
                                                              : PADDI_ERR_DEVICE,        \
                                    std::string(#expr) + ": " + hipGetErrorString(e_));  \
    } while (0)

const char *status_str(paddi_status s);

// Survivors returned with a pass's counters (more: a second copy).  The download is sized
// from the previous pass's count, between SCAP_MIN and SCAP: on MI355X a device-to-pinned
// copy of the 200 KB an 8192-survivor head needs started about 50 us later than one of a few
// KB, every pass.
constexpr int64_t SCAP = 2048;
// Range discords whose small distance is refined in double precision, per call (see range()).
constexpr size_t RANGE_EXACT_CAP = 65536;
constexpr int64_t SCAP_MIN = 256;

// No-allocation range call (paddi_range_discords_noalloc): its device buffers are slices of a
// file-scope device array of SCRATCH_BYTES (scratch.hip), its exact refinements run in
// launches of at most EXACT_NOALLOC_CHUNK candidates (their arguments fit a slice and the
// pinned staging sized when the context is created).
constexpr size_t SCRATCH_BYTES = PADDI_NOALLOC_SCRATCH_BYTES;
constexpr size_t EXACT_NOALLOC_CHUNK = 4096;
// Address of this device's scratch array (hipGetSymbolAddress, taken once per device; the
// device must be current).
void *scratch_setup(int device);
constexpr int64_t BOUND_WINDOW = 1024;  // columns behind the first-length candidates' bounds
constexpr int BOUND_CANDIDATES = 32;    // default candidates per series refined for the first r
// Calls with fewer rows skip the first length's SAX witnesses: their first pass is shorter
// than the sort (c1, 65,536 points: 1.13 -> 1.20 ms with them)
constexpr int64_t SAX_MIN_ROWS = 1 << 18;

// Grow-only device buffer.  A borrowed buffer is a slice of memory the engine does not own
// (the scratch array, ScratchScope): it is never freed and never grows.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    bool borrowed = false;
    void reserve(size_t nbytes);
    void release();
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    template <class T> T *as() const { return (T *)p; }
    void swap(DevBuf &o)
    {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
        std::swap(borrowed, o.borrowed);
    }
};

// A series uploaded to HBM (paddi_series_upload).
struct Series {
    int64_t n = 0;
    DevBuf x;  // n floats
};

// Surviving row: single-precision NN distance a with bounds alo <= exact <= ahi (up to
// conv_ed2), exact distance v once computed (and v2 at the next length, when asked for).
struct Cand {
    int64_t i;
    float a;
    double v;
    bool exact;
    float alo, ahi;
    double v2 = -1.0;
};

// Exact greedy top-k of one series at one (m, r2) (see engine.cpp).
struct TopK {
    std::vector<Cand> cands;
    std::vector<char> gone;
    std::vector<Cand> acc;
    double r2 = 0.0;
    int m = 0, k = 0;
    int spec = 4;      // candidates refined per round (speculative beyond the undecided set)
    bool done = false;
    void reset(std::vector<Cand> &&c, double r2_, int m_, int k_);
    // Advances the greedy; returns true with `need` (indices into cands) when exact values
    // are required before it can continue.
    bool advance(std::vector<size_t> &need);
};

// Series of one call laid out in a shared global row space.
struct Batch {
    int nser = 0;
    int minL = 0;
    std::vector<int64_t> n, xoff, roff, pboff;
    int64_t xtotal = 0, rows = 0, nblocks = 0, npb = 0;
    std::vector<SerDev> host;   // per series, N/thr updated per length
    DevBuf x, P, refs, ser, bser, pser;
    std::vector<int> bser_host;
};

class Engine {
public:
    explicit Engine(int device);
    ~Engine();
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    Series *upload(const float *T, int64_t n);
    // host: a host copy of each series to read instead of the device one (null: device)
    void build_batch(Batch &B, const std::vector<const Series *> &ss, int minL,
                     const float *const *host = nullptr);
    void prefix(Batch &B);
    // Window statistics at m (and at m2 > 0 for k_exact's second length).
    void derive(Batch &B, int m, paddi_stats *st, int m2 = 0);
    // One (m, r2_s) pass over every series of the batch; survivors per series.
    // keep: restrict the pass to these global blocks (a candidate scan, not an attempt).
    void pass(Batch &B, int m, const std::vector<double> &r2, bool witness,
              std::vector<std::vector<Cand>> &out, paddi_stats *st,
              const std::vector<uint8_t> *keep = nullptr);
    // First-length lower bound of d_k per series in `need` (0 = none; see engine.cpp).
    void first_bound(Batch &B, int m, int k, const std::vector<char> &need,
                     std::vector<double> &lb, paddi_stats *st,
                     std::vector<std::vector<int64_t>> *cand_out = nullptr);
    // first_bound, then the candidate blocks scanned in full at that bound (see engine.cpp)
    // found[s]: k discords were found among series s's candidate blocks (lb[s] their k-th)
    void bootstrap(Batch &B, int m, int k, const std::vector<char> &need,
                   std::vector<double> &lb, std::vector<char> &found, paddi_stats *st);
    // First-length candidates of the series in `need` (see engine.cpp).
    void bound_candidates(Batch &B, int m, const std::vector<char> &need, int count,
                          std::vector<std::vector<int64_t>> &cand, paddi_stats *st);
    // Exact NN ED^2 of (series, local row) requests at length m (derive(m) must be current);
    // with m2 in (m, m + EXACT_EXT] also at m2 into out2 (DBL_MAX: none).
    // ub: per request an upper bound of its distance (the scans' bound; null: none); lb:
    // lower bounds suffice (k_nn without double precision).  With k_nn the second length's
    // values are lower bounds.
    void exact(Batch &B, int m, const std::vector<std::pair<int, int64_t>> &req,
               std::vector<double> &out, paddi_stats *st, int m2 = 0,
               std::vector<double> *out2 = nullptr, const std::vector<double> *ub = nullptr,
               bool lb = false);
    // Run the exact top-k of every series to completion (batched exact refines); with m2,
    // every refined candidate also gets its exact distance at m2.
    void topk_all(Batch &B, int m, std::vector<TopK> &tk, paddi_stats *st, int m2 = 0);

    void run(const std::vector<const Series *> &ss, int minL, int maxL, int step, int k,
             const paddi_opts &o, std::vector<paddi_discord> &out, paddi_stats *st);
    // Range discords of a one-series batch at (m, r2) (derive(m) current): the flags and NN
    // distances of paddi_range_discords.
    void range(Batch &B, int m, float r2, int32_t *is_discord, float *nn_dist2);
    // paddi_range_discords on the scratch array only: no device or pinned allocation.
    void range_noalloc(const float *T, int64_t n, int m, float r2, int32_t *is_discord,
                       float *nn_dist2);
    void copy_st64(double2 *dst, int64_t N);

    int64_t select_window_ = 8192;   // columns right of a block scanned by k_scan_select
    int64_t refine_chunk_ = 65536;   // longest k_scan_refine work item (columns)
    int64_t min_chunk_ = 2048;       // shortest k_scan_refine work item (columns)
    int64_t refine_waves_ = 16384;   // target number of k_scan_refine work items
    int witness_tiles_ = 4;          // witness ranges scanned first by k_scan_select
    int spec_ = 4;                   // candidates refined per exact round
    bool use_graph_ = false;         // per-length device pipeline as a hipGraph
    int prev_shift_ = 1;             // next r from previous discords at i, i-1 .. i-prev_shift
    bool bootstrap_ = true;          // learned first length: candidate lower bound on r
    int bound_cands_ = BOUND_CANDIDATES;  // candidate rows per series behind that bound
    int64_t scap_next_ = SCAP;       // survivors downloaded with the next pass's counters
    int direct_budget_ = DIRECT_BUDGET;  // refine: L live rows on the matrix cores while
    int direct_rows_ = DIRECT_ROWS;      // L <= direct_rows_ and L (m + 15) <= the budget
    int device_lengths_ = 0;         // lengths on the device between host synchronisations
                                     // (0: the host runs every length)
    int device_rounds_ = 2;          // exact rounds per length on the device
    int refine_rounds_ = 3;          // refine rounds per pass (refine_bounds)
    bool seed_wit_ = true;           // fast forms, first length: witnesses from SAX words
    int refine_order_ = 0;           // refine work items block-major (0), chunk-major (1), or
                                     // chunk-major at a call's first length only (2)
    bool first_len_ = false;         // fast(): the passes of a call's first length
    bool excl_sym_ = false;          // exclusion rule: |i - j| >= m - 1 instead of |i - 1 - j| >= m
    bool exact_nn_ = true;           // k_nn for exact refinement (PADDI_EXACT_NN=0: k_exact)
    bool debug_ = false;             // per-pass trace on stderr (PADDI_DEBUG=1)
    double wait_limit_s_ = 600.0;    // longest wait for queued device work (PADDI_WAIT_LIMIT_S)
    std::vector<paddi_attempt_stats> *att_log_ = nullptr;  // per-attempt records, when kept

private:
    // per-attempt bookkeeping (att_log_): att_mark snapshots the call's statistics (at a
    // length's start), att_end records what changed since as one attempt and snapshots again
    void att_mark(paddi_stats *st);
    void att_end(int m, int attempt, double r2, int discords, paddi_stats *st);
    paddi_stats att_snap_{};
    double att_t0_ = 0.0;
    void plain(Batch &B, int sid, int minL, int maxL, int step, int k, const paddi_opts &o,
               std::vector<paddi_discord> &out, paddi_stats *st);
    void fast(Batch &B, int minL, int maxL, int step, int k, const paddi_opts &o,
              std::vector<paddi_discord> &out, paddi_stats *st);
    // The read side of the witnesses (wit_[witcur_ ^ 1]) from SAX words at m (derive(m)
    // current), for a call's first length (sax.hip).
    void seed_witnesses(Batch &B, int m);
    // Lengths j0 .. jend-1 of fast() on the device (ChainArgs); returns the first one the
    // device did not finish (jend when all), whose discords and next-length values are
    // then in prev / prev_v, and its results appended to res.
    // ahead: the range from the previous exact rounds' second length (prev_v), else shifted
    // by prev_shift_ with a launch of its own per length.
    int chain(Batch &B, int minL, int step, int k, int j0, int jend, bool ahead,
              std::vector<std::vector<int64_t>> &prev, std::vector<std::vector<double>> &prev_v,
              std::vector<std::vector<paddi_discord>> &res, paddi_stats *st);
public:
    void upload_ser(Batch &B);
    // Learned proposal of the squared range per series at length m (0 = no proposal).
    void propose_r2(Batch &B, int m, int k, std::vector<double> &out);
    // Features of every series at length m for a top-k search (derive(m) must be current);
    // NFEAT per series.
    void features(Batch &B, int m, int k, std::vector<float> &X);
    void mlp_forward(const float *prm, const float *fmu, const float *fsd, const float *X,
                     int Bn, float *Y);
    void mlp_train(float *prm, const float *fmu, const float *fsd, const float *X,
                   const float *T, int Bn, int steps, float lr, float wd, float *losses,
                   float *grad_out);
    void set_model_path(const char *p);

private:
    // One pass's device work (parameter upload .. survivor download) on stream_, eagerly or
    // as a replay of the captured graph of its LDS class.
    // host_io = false: without the parameter upload and the result download (the device
    // length loop writes the parameters and reads the results on the device)
    void enqueue_pass(const ScanPtrs &Q, size_t lds, size_t parb, int64_t rows,
                      int64_t nwords, int64_t scap, bool timed, const uint8_t *keep = nullptr,
                      bool host_io = true);
    void ensure_pinned(size_t par_bytes, size_t ser_bytes);
    ScanPtrs fill_pass(Batch &B, int m, const std::vector<double> &r2, bool witness, int64_t WA,
                       bool refine, size_t &parb);
    void collect_derive();
    void wait();
    paddi_stats *derive_st_ = nullptr;   // derive events not yet read
    struct GraphEnt {
        std::vector<uintptr_t> sig;   // everything baked into the graph
        hipGraphExec_t exec = nullptr;
    };
    std::map<std::pair<int, size_t>, GraphEnt> graphs_;
    // device length loop: one graph per length class and statistics buffer, keyed (the
    // buffer holding the length's (mu, invn), the class's largest m); its signature is every
    // argument it bakes in
    struct ChainGraph {
        std::string sig;
        hipGraphExec_t exec = nullptr;
    };
    std::map<std::pair<uintptr_t, int>, ChainGraph> chain_graphs_;
    DevBuf ch_ctl_, ch_ser_, ch_res_, ch_xv_, ch_xv2_, ch_cand_, ch_cnt_, ch_key_, ch_key2_,
        ch_goff_, ch_candr_, ch_cntr_, ch_keyr_, ch_ub_, ch_runub_, ch_tilelo_;
    void *pin_ch_ = nullptr;       // pinned: the loop's control block, then its downloads
    size_t pin_ch_bytes_ = 0;
    void drop_graphs();

    Model model_;
    std::string model_path_;
    DevBuf feat_, mlpbuf_;
    Batch batch_;
    bool wread_valid_ = false;   // the read-side witness buffer holds the previous length
    int device_;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev_[8] = {};
    void *pin_ = nullptr;
    DevBuf flag_;                  // k_prefix's non-finite flag
    // pinned staging: pass parameters (PassDev + series table), pass results (PassCtl +
    // the first SCAP survivors), series table of derive/exact; ser_ev_ marks the end of
    // the last copy out of pin_ser_
    void *pin_par_ = nullptr, *pin_out_ = nullptr, *pin_ser_ = nullptr;
    size_t pin_par_bytes_ = 0, pin_ser_bytes_ = 0;
    hipEvent_t ser_ev_ = nullptr;
    hipEvent_t sync_ev_ = nullptr;  // wait(): completion of the work queued so far
    DevBuf colv_, st64_, alive_, best_, besth_, bestl_, wit_[2], blist_, par_, pout_;
    DevBuf sax_k_, sax_k2_, sax_v_, sax_v2_, sax_tmp_;  // seed_witnesses' sort
    DevBuf tmax_, rho_;            // per global block: row scales, tile column factors
    DevBuf st64n_;                 // (mu, invn) at the last derive's second length
    int derive_m2_ = 0;            // that length (0: none)
    DevBuf xarg_;                  // exact-refine arguments and results (one block)
    DevBuf nnlo_;                  // k_nn's super-tile lower bounds (its filter pass)
    DevBuf bkey_;                  // per-block candidate keys of bound_candidates
    DevBuf keep_;                  // per-block flags of a restricted pass
    void *pin_x_ = nullptr;        // its pinned staging
    size_t pin_x_bytes_ = 0;
    int witcur_ = 0;
    void *scratch_ = nullptr;      // this device's scratch array (scratch_setup)
    friend class ScratchScope;
};

}  // namespace paddi

// Opaque handles of the C ABI.
struct paddi_ctx {
    paddi::Engine *eng;
    std::string last_error;
};
struct paddi_series {
    paddi_ctx *ctx;
    paddi::Series *s;
};

// Body of a C entry point: its status, or the status of the error it raised (the message is
// kept in ctx->last_error).  No exception leaves the library.
template <class F>
paddi_status guarded(paddi_ctx *ctx, F &&f)
{
    try {
        return f();
    } catch (const paddi::HipError &e) {
        if (ctx) ctx->last_error = e.what();
        return e.st;
    } catch (const std::bad_alloc &) {
        if (ctx) ctx->last_error = "host allocation failed";
        return PADDI_ERR_NOMEM;
    } catch (const std::exception &e) {
        if (ctx) ctx->last_error = e.what();
        return PADDI_ERR_DEVICE;
    } catch (...) {
        if (ctx) ctx->last_error = "unknown error";
        return PADDI_ERR_DEVICE;
    }
}
