// Kernel declarations, argument blocks and engine-wide constants (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace paddi {

constexpr int PB = 4096;       // prefix-sum block (points); also the largest supported m
constexpr int MAXM = 2048;     // largest subsequence length the scan kernels accept
constexpr int MAXU = 8;        // candidates per exact-refine group
constexpr int EXACT_COLS = 4096; // columns per k_exact block (sub-tiles of 256 exact_qc)
constexpr int EXACT_SLOTS = 64;  // result slots per candidate (blocks spread their minima)
constexpr int64_t XPAD = 2048; // zero padding after each series (floats); the scans read
                               // up to n + 63
// Rows of a call (every series' windows at the shortest length, block-padded) must stay below
// 2^31: the candidate keys of k_block_best pack a global row in 32 bits, and the witness
// tables, the scans' tile offsets and the pass counters hold rows and columns as int.
constexpr int64_t MAX_ROWS = int64_t(1) << 31;
// Row-space rows of a series of n points at shortest length minL (Engine::build_batch).
__host__ __device__ inline int64_t series_rows(int64_t n, int minL)
{
    return (n - minL + 1 + 63) / 64 * 64 + 128;
}

// One series of a batch, as the kernels see it.  The batch's table is uploaded once (k_derive,
// k_exact and the feature kernels take N = n - m + 1 from n and m); a pass carries its own
// copy with N and thr of that pass (N = 0 switches a series off).
struct SerDev {
    int64_t xoff;   // offset of x_s in the concatenated series buffer
    int64_t n;      // points
    int64_t roff;   // first global row (multiple of 64)
    int64_t pboff;  // first prefix block
    int64_t N;      // windows at this length, n - m + 1
    float thr;      // 1 - r2/(2m): prune row i at column j when cov invn_i invn_j reaches it
                    // plus the cell's error bound (per tile, see ScanScal::dk)
    int32_t pad;
};

struct ScanArgs {
    const float *x;
    const SerDev *ser;
    const int *bser;          // global block -> series (-1 = padding block)
    const float4 *colv;       // global rows: (df, dg, invn, mu) in single precision
    const float2 *tmax;       // global blocks: (max sqrt(M2), max invn) over the block's rows
    const float *rho;         // global blocks: column factor of a tile (k_pass_init)
    uint64_t *alive;          // one word per global block
    int *best;                // global rows: ord_key of the max correlation
    int *besth, *bestl;       // global rows: ord_key of (max correlation +- its error bound)
    int *wit;                 // global rows: pruning tile recorded at this length (or null)
    const int *wit_read;      // global rows: tiles recorded at the previous length (or null)
    unsigned long long *pairs;  // PassCtl::pairs then PassCtl::live (k_scan_select, k_scan_refine)
    int64_t WA;               // columns right of a block scanned by k_scan_select
    int64_t CH;               // columns per k_scan_refine work item (set per round on device)
    int64_t U0;               // first cyclic offset scanned: round_up(m + 64, 64)
    int m;
    int ml;                   // ScanScal::ml
    int dbudget, drows;       // ScanScal::dbudget, drows
    int witness_max;          // witness ranges scanned first
    float dk;                 // delta(m) / (2m): error bound per unit of block and tile scale
};

// Refine rounds of one pass: round r scans cyclic offsets [bound[r], bound[r+1]) of the
// blocks still alive when it starts, in work items of CH columns, where CH puts about
// `items` work items in the round and is clamped to [min_chunk, max_chunk] (multiple of 64).
constexpr int REFINE_ROUNDS_MAX = 4;
struct RefineArgs {
    int64_t bound[REFINE_ROUNDS_MAX + 1];
    int64_t min_chunk, max_chunk, items;
    int64_t order;  // work item order: 0 block-major, 1 chunk-major (k_scan_refine)
};
// Bounds of `rounds` refine rounds (3 or 4): offsets from b0 = U0 + select window up to 2, 8
// (and 32) times the largest chunk, the last round to the end of the longest series (offset
// end); the bounds past the last round are end (empty rounds).
__host__ __device__ inline void refine_bounds(int64_t b0, int64_t ch, int64_t end, int rounds,
                                              int64_t *bound)
{
    int64_t span = 2 * ch;
    bound[0] = b0 < end ? b0 : end;
    for (int r = 1; r <= REFINE_ROUNDS_MAX; r++, span *= 4) {
        const int64_t b = b0 + span;
        bound[r] = (r < rounds && b < end) ? b : end;
    }
}

// Buffers of a pass, passed by value as a kernel argument: they stay fixed while a batch
// runs, so a captured graph can hold them, and the compiler knows they are global memory.
struct ScanPtrs {
    const float *x;
    const SerDev *ser;        // the pass's series table (N = 0 switches a series off)
    const int *bser;
    const float4 *colv;
    const float2 *tmax;
    float *rho;
    uint64_t *alive;
    int *best, *besth, *bestl;
    int *wit0, *wit1;         // the two witness buffers (one written, one read per length)
    unsigned long long *pairs;
};

// Scalars of one pass.
struct ScanScal {
    int64_t WA, U0;
    int m;
    // Exclusion: row i and column j form a pair iff j - i >= m - 1 or i - j >= ml; ml = m + 1
    // is the reference CPU rule |i - 1 - j| >= m (src/cpu_kernels/cpu_kernels.cpp:57), ml = m - 1
    // its GPU kernels' |i - j| >= m - 1 (src/PD3_kernels.cuh:70).  The scans' masked tiles cover
    // offsets within m + 63 either side of the diagonal, which holds both bands.
    int ml;
    // A refine work item whose block has L <= drows live rows computes them on the matrix
    // cores alone when L (m + 15) <= dbudget (see scan_range_direct).
    int dbudget, drows;
    int witness_max;
    int wsel, rsel;           // witness buffer written / read (0, 1; -1 = none)
    int some_off;             // 1 when the pass switches some series off
    float dk;                 // delta(m) / (2m) (ScanArgs::dk)
    int off;                  // 1: no pass (a stopped device length loop): k_pass_init clears
                              // the alive words and counters only
};

// Everything one pass reads that changes between passes: one host-to-device copy (or one
// memcpy node of a captured graph) brings it in.  The series table follows it.
struct PassDev {
    ScanScal S;
    RefineArgs R;
};

// Device counters of one pass, cleared by k_pass_init.  The pair counts (statistics) are
// spread over PAIR_SLOTS addresses by workgroup, so the workgroups' final adds do not all
// meet on one address.
constexpr int PAIR_SLOTS = 64;
struct PassCtl {
    int list[REFINE_ROUNDS_MAX];  // live blocks listed before refine round r
    int survivors;
    int pad[3];
    unsigned long long pairs[2][PAIR_SLOTS];  // [0] k_scan_select, [1] k_scan_refine
    // the same pair evaluations restricted to rows still alive when their tile started (the
    // useful share of the lane-steps), right after pairs
    unsigned long long live[2][PAIR_SLOTS];
};
// k_scan_refine evaluates a block with L live rows on the matrix cores alone
// (scan_range_direct) when L <= DIRECT_ROWS and L (m + 15) <= the budget: each row costs one
// product of K4 / 4 ~ (m + 15) / 4 matrix-core steps per 256 columns, the recurrence's four
// tiles about 1200 vector instructions whatever m.  Rows after the first stage their windows
// in the tile buffer (512 floats), which bounds L at the longer lengths.
constexpr int DIRECT_ROWS = 4;
constexpr int DIRECT_BUDGET = 700;

struct ExactArgs {
    const float *x;
    const SerDev *ser;
    const double2 *st64;      // global rows: (mu, invn) in double precision
    const int64_t *cand;      // local row index of each candidate
    const int *gfirst;        // group g holds candidates [gfirst[g], gfirst[g+1])
    const int *gser;          // series of group g
    const int64_t *goff;      // first block of group g (prefix sums of 256-column chunks)
    unsigned long long *outkey;  // [candidate][EXACT_SLOTS]: minima by block slot
    // Optional second length m2 in (m, m + 4]: the same candidates' distances at m2 from the
    // same dot products extended by m2 - m terms
    const double2 *st64n;     // global rows: (mu, invn) at m2 (k_derive)
    unsigned long long *outkey2; // as outkey, at m2
    int ngroups;
    int m;
    int m2;                   // 0: one length
    int sym;                  // exclusion rule: 0 the CPU rule, 1 the symmetric one (ScanScal::ml)
    // k_exact_chain only (the device-planned launch): the lengths (dev_m[0] = m, dev_m[1] =
    // m2) and the candidate count of each of a series' gslots groups, read on the device;
    // ngroups is the number of series there, goff their first blocks, m and m2 above are
    // unused
    const int *dev_m;
    const int *dev_count;
    int gslots;
    // k_nn only: the columns' single-precision terms at m (ScanArgs::colv) and block scales
    // (ScanArgs::tmax), per candidate an upper bound of its nearest-neighbour distance (null:
    // none) and EXACT_SLOTS single-precision upper bounds (ord_key of a float, set to
    // 0x7fffffff before the launch), and lb = 1 for lower bounds only (no double precision)
    const float4 *colv;
    const float2 *tmax;
    const double *ub;
    int *runub;
    int lb;
    // k_nn's filter pass and k_nn_fix: per candidate `tiles` super-tile lower bounds, k_nn_fix
    // blocks of 64 super-tiles per candidate (tb), the candidates' series (host-planned) or
    // the group size (device-planned)
    float *tilelo;
    int tiles, tb;
    const int *csid;
    int nu_dev;
};
constexpr int EXACT_EXT = 4;  // largest m2 - m

constexpr int DERIVE_RPT = 4;  // rows per k_derive thread (blocks of 256 DERIVE_RPT rows)
// Blocked prefix sums of every series; *bad becomes nonzero if a value is NaN or infinite.
__global__ void k_prefix(const float *__restrict__ x, const SerDev *__restrict__ ser,
                         const int *__restrict__ pser, double2 *__restrict__ P,
                         double4 *__restrict__ refs, int *__restrict__ bad);
__global__ void k_derive(const float *__restrict__ x, const SerDev *__restrict__ ser,
                         const int *__restrict__ bser, const double2 *__restrict__ P,
                         const double4 *__restrict__ refs, int64_t rows, int m,
                         float4 *__restrict__ colv, double2 *__restrict__ st64,
                         float2 *__restrict__ tmax, int m2, double2 *__restrict__ st64n,
                         const int *__restrict__ dev_m);
// The scan kernels: one 64-row block per workgroup (one wave); LDS from scan_lds_bytes(m).
void launch_scan_select(const ScanPtrs &Q, const PassDev *P, unsigned grid, size_t lds,
                        hipStream_t st);
void launch_scan_refine(const ScanPtrs &Q, const PassDev *P, const int *blist, PassCtl *ctl,
                        int round, unsigned grid, size_t lds, hipStream_t st);
__global__ void k_pass_init(ScanPtrs Q, const PassDev *__restrict__ P, int64_t rows,
                            int64_t nwords, PassCtl *__restrict__ ctl);
constexpr int LIST_WPT = 16;  // alive words per k_list_blocks thread
__global__ void k_list_blocks(const uint64_t *__restrict__ alive, int64_t nwords,
                              int *__restrict__ list, int *__restrict__ count);
// Restrict a pass to the global blocks flagged in keep (the others' alive words cleared).
__global__ void k_mask_blocks(uint64_t *__restrict__ alive, const uint8_t *__restrict__ keep,
                              int64_t nwords);
// A surviving row: global row, single-precision squared NN distance d2 and bounds
// d2lo <= exact <= d2hi (up to the rounding of their conversion).
struct Surv {
    int64_t g;
    float d2;
    float d2lo, d2hi;
};
// Pass results as one device block: the counters, then the survivors (one copy to the host).
constexpr size_t PASS_OUT_HEAD = 2560;
static_assert(sizeof(PassCtl) <= PASS_OUT_HEAD, "PassCtl fits the head of the pass results");
// Per 64-row block, its row with the lowest maximum correlation found (the largest upper
// bound on the NN distance): ((key ^ 0x80000000) << 32 | global row), ~0 for blocks without
// rows; one wave per block.
__global__ void k_block_best(ScanPtrs Q, const PassDev *__restrict__ P, int64_t nblocks,
                             unsigned long long *__restrict__ out);
__global__ void k_survivors(ScanPtrs Q, const PassDev *__restrict__ P, int64_t rows,
                            PassCtl *__restrict__ ctl, Surv *__restrict__ out, int64_t cap);
// k_exact is a template on the largest group size nu of the launch (1..MAXU).
void launch_exact(const ExactArgs &E, int nu, unsigned grid, size_t lds, hipStream_t st);
// k_exact_chain (the device-planned launch of the length loop) for groups of nu <= CHAIN_NU;
// dual: also at the second length.
void launch_exact_chain(const ExactArgs &E, int nu, bool dual, unsigned grid, size_t lds,
                        hipStream_t st);

// k_nn (the matrix-core nearest-neighbour kernel, replacing k_exact): groups of nu <= NN_MAXU
// candidates; dual: also a lower bound at the second length; lb: lower bounds only; dev: the
// device-planned launch of the length loop (as k_exact_chain).
constexpr int NN_MAXU = 4;
void launch_nn(const ExactArgs &E, int nu, bool dual, bool lb, bool dev, unsigned grid,
               size_t lds, hipStream_t st);
// k_nn_fix: the double-precision pass after k_nn's filter pass (grid: candidates x tb).
void launch_nn_fix(const ExactArgs &E, bool dev, unsigned grid, hipStream_t st);
// Dynamic LDS of k_nn: the chunk's series window (padded every 16 floats) and per candidate
// its single-precision window (16 zeros either side of K4 + 16).
inline size_t nn_lds_bytes(int m, int nu)
{
    const int K4 = (m + 15 + 3) & ~3;
    const int xs = 4096 + K4 + 16;
    const size_t xsf = (size_t)(xs + (xs >> 4) + 2);
    return ((xsf * 4 + 7) & ~(size_t)7) + (size_t)nu * (K4 + 32) * 4;
}
// Largest group of k_nn at length m within 64 KiB of LDS.
inline int nn_max_u(int m)
{
    int u = 1;
    while (u < NN_MAXU && nn_lds_bytes(m, u + 1) <= 65536 - 2048) u++;
    return u;
}

// Squared-distance bounds shared by the host and the device.
// delta_ed2: bound on |single - exact| squared distance where the windows a cell's chain
// crossed are of the scale of the cell's own: correlation error of a seed (m-term FP32 dot
// product on locally centred data) plus up to 63 recurrence steps, with a wide safety factor
// over the largest error measured on the test workloads.  The scans scale it by the ratio of
// the largest scale crossed to the cell's (k_pass_init).
__host__ __device__ inline double delta_ed2(int m) { return 2.0 * m * (4.0e-5 + 2.0e-7 * m); }
// conv_ed2: rounding of a scan bound's conversion to a squared distance d (the bound's last
// addition and scaling by invn_i in the scan, then 2m (1 - c) in k_survivors, all in single
// precision): at most 10m + 2|d| units of 2^-24 (|c| <= 1 contributes the 10m), with a factor
// 4 of room.
__host__ __device__ inline double conv_ed2(int m, double d)
{
    return 4.0 * (10.0 * m + 2.0 * (d < 0.0 ? -d : d)) * 6.0e-8;
}
// Exact distances within TIE_REL of each other rank as equal, and then by index.  k_exact's
// value of a pair is not bitwise symmetric (it centres the candidate's side), while a mutual
// nearest-neighbour pair shares one distance, which the reference's strict '>' ranks lower
// index first (sets_processing.hpp:28); the evaluation errors are orders of magnitude below
// TIE_REL, and TIE_REL far below the 1e-4 parity bar.
constexpr double TIE_REL = 1e-9;

// ---------------------------------------------------------------------------------------
// The length loop on the device (Engine::device_lengths).  Every length after the first runs
// as one captured graph: k_derive, the pass (k_pass_init .. k_survivors), then `rounds` x
// (k_chain_topk, k_exact_chain), a final k_chain_topk, which records the length's discords
// and their distances at the next length, and k_chain_next, which starts the next length (its
// range from those distances, the pass parameters, the lengths k_derive and k_exact_chain
// read).  With a shifted range (ChainArgs::shift) the graph starts with the range's own exact
// launch instead: k_chain_rreq, k_derive, k_exact_chain, k_chain_range.  The host only
// replays graphs.  A length the device cannot finish (a pass with fewer than k discords above
// its range, or more exact rounds needed than the graph holds, or more survivors in one series
// than one workgroup takes) stops the loop: `halt` is set, every later length's kernels find
// their series switched off, and the host runs that length itself.
constexpr int CHAIN_KMAX = 8;      // largest top-k of the device loop
constexpr int CHAIN_NU = 4;        // largest exact group per series and round
constexpr int CHAIN_ROUNDS = 4;    // largest number of exact rounds per length
constexpr int CHAIN_SCAP = 4096;   // survivors one series may bring to the device top-k

// Per series: the current length's state, and the discords of the length before the loop's
// first (later lengths take them from the results of the length before).
struct ChainSer {
    int64_t prev[CHAIN_KMAX];  // discords of the length before the first (local rows)
    double pv[CHAIN_KMAX];     // their exact distances at the first length (DBL_MAX: none)
    double r2;                 // the current length's squared range
    int nprev;
    int status;                // current length: 0 searching, 1 done, 2 fewer than k above r2,
                               // 3 unfinished (rounds or survivors exhausted)
    int nreq;                  // exact candidates asked for by the last round
    int nrq;                   // shifted range: rows evaluated for the range (k_chain_rreq)
    int req_p[CHAIN_NU];       // their survivor indices
    unsigned char rq_u[CHAIN_KMAX * (EXACT_EXT + 1)];  // the discord each range row shifts
};
// One series at one length, as the host reads it back.
struct ChainRes {
    int64_t i[CHAIN_KMAX];     // discords (local rows), best first
    double v[CHAIN_KMAX];      // their exact squared NN distances
    double v2[CHAIN_KMAX];     // and at the next length (DBL_MAX: none), its range's inputs
    double r2;                 // range of the pass that found them
    int n;                     // discords (0..k); written once the series finished the length
    int pad;
};
// The loop's scalars.
struct ChainCtl {
    int j;            // length running (index from the call's first length)
    int j0;           // the loop's first length
    int wit0;         // witness buffer written at length j0
    int halt;         // 0, or 1 + the index of the first length the device did not finish
    int m, m2;        // k_derive's and k_exact_chain's lengths (0 once halted)
    // statistics of the finished lengths: attempts, survivors, exact refinements, pair
    // evaluations of select / refine, and of those on live rows
    unsigned long long stats[8];
};
// Arguments of the chain kernels (fixed while a batch runs, so graphs can hold them).
struct ChainArgs {
    ChainCtl *ctl;
    ChainSer *cs;
    ChainRes *res;             // [length index - ChainCtl::j0][series]
    const SerDev *ser;         // the batch's table (n and offsets)
    SerDev *pser;              // the pass's table (N, thr per series)
    PassDev *par;              // the pass's scalars
    const PassCtl *pctl;       // the pass's counters
    const Surv *surv;          // the pass's survivors
    double *xv, *xv2;          // per survivor: exact distance at m and at m2 (NaN: not refined)
    int64_t *cand;             // [series][nu] exact candidates of a round (k_exact_chain)
    int *cnt;                  // [series] their number
    double *ub;                // [series][nu] their upper bounds (k_nn)
    int *runub;                // [series][nu][EXACT_SLOTS] k_nn's single-precision bounds
    unsigned long long *outkey, *outkey2;  // [series][nu][EXACT_SLOTS] (k_exact_chain)
    int64_t scap;              // survivors xv holds
    int64_t WA, max_chunk, min_chunk, items;   // pass scalars that stay fixed (Engine::fill_pass)
    int witness_max, dbudget, drows, sym, rrounds;  // rrounds: refine rounds per pass
    int rorder;                // RefineArgs::order
    int nser, k, nu, spec, rounds;
    int minL, step;            // length j is minL + j step
    // The range of a length: shift = 0, from the previous discords' distances at this length
    // computed by the previous length's exact rounds (m2); shift > 0, from a launch of its own
    // on rows i, i-1 .. i-shift of each previous discord i (Engine::prev_shift_), in gr group
    // slots of nur rows per series
    int shift, gr, nur;
    int64_t *candr;            // [series][gr][nur] rows of the range launch
    int *cntr;                 // [series][gr] their number
    unsigned long long *keyr;  // [series][gr][nur][EXACT_SLOTS] its results
};
// Exact top-k of one series per workgroup of one wave; round = 0 .. rounds (the last only
// finishes).
__global__ void k_chain_topk(ChainArgs C, int round);
// Statistics of the length just run and the start of the next (also of the loop's first,
// launched once with ChainCtl::j = j0 - 1 and m = 0): one workgroup.
__global__ void k_chain_next(ChainArgs C);
// Shifted range (shift > 0), one wave per series: the rows of the range launch, after
// k_chain_next; then, after that launch, the range and the series' pass table entry.
__global__ void k_chain_rreq(ChainArgs C);
__global__ void k_chain_range(ChainArgs C);

// Scan seeds: the MFMA reduction length (m + 15 rounded up to 4) and the series ring
// capacity (a multiple of 16 holding a 256-column super-tile window of 256 + K4 columns).
__host__ __device__ inline int scan_k4(int m) { return (m + 15 + 3) & ~3; }
__host__ __device__ inline int scan_ring_w(int m) { return (256 + scan_k4(m) + 15) & ~15; }
// LDS bytes of the scan kernels for length m: tiles (2 x 64 float4), raw seeds (256),
// the zero-padded first row (K4 + 20), the ring (both copies, padded; it also serves as the
// chain-initialisation scratch).
inline size_t scan_lds_bytes(int m)
{
    const int K4 = scan_k4(m), W = scan_ring_w(m);
    const int ring = 2 * W + (2 * W >> 4) + 4;
    const int scratch = ((64 + m + 4 + 3) & ~3) + ((m + 3) & ~3);
    return (size_t)(2 * 64 * 4 + 256 + ((K4 + 20 + 3) & ~3) + (ring > scratch ? ring : scratch)) * 4;
}
// Columns per thread of k_exact for groups of nU candidates: 8 for one or two candidates, else
// 4 (the accumulators of 8 columns for more would cut the waves a SIMD holds).
__host__ __device__ constexpr int exact_qc(int nU) { return nU <= 2 ? 8 : 4; }
// Candidate window length of k_exact, zero-padded: to a multiple of 16 for groups of three
// (their values enter as DPP broadcasts of 16 lanes), else of 4.
__host__ __device__ constexpr int exact_mk(int m, int nU) { return nU == 3 ? (m + 15) & ~15 : (m + 3) & ~3; }
// Doubles of k_exact's staged series window: 256 qc + mk + 4 values, qc values per qc + 2
// doubles.
__host__ __device__ inline int exact_xs_doubles(int m, int nU)
{
    const int qc = exact_qc(nU);
    return (qc + 2) * ((256 * qc + exact_mk(m, nU) + 4) / qc + 1);
}
// LDS bytes of k_exact for groups of nU candidates of length m.
inline size_t exact_lds_bytes(int m, int nU)
{
    return ((size_t)exact_xs_doubles(m, nU) + (size_t)nU * exact_mk(m, nU)) * 8;
}
// Candidates per exact group at length m: the largest u such that groups of every size up to u
// fit 63 KiB of dynamic LDS (64 KiB less the kernel's static arrays; the footprint is not
// monotone in the group size: groups of three pad the candidates to 16 values); >= 1 for
// m <= MAXM.
inline int exact_max_u(int m)
{
    int u = 1;
    while (u < MAXU && exact_lds_bytes(m, u + 1) <= 65536 - 1024) u++;
    return u;
}

// First-length witnesses from SAX words (sax.hip): per global row the tile of a row of its
// series with the same SAX word at least m + 1 away (-1: none), into wit.  Scratch: keys and
// keys2 (rows x 8 B), vals and vals2 (rows x 4 B), tmp (sax_sort_bytes).
size_t sax_sort_bytes(int64_t rows);
hipError_t launch_sax_witnesses(const float *x, const SerDev *ser, const int *bser,
                                const float4 *colv, int64_t rows, int m, int nser,
                                unsigned long long *keys, unsigned long long *keys2, int *vals,
                                int *vals2, void *tmp, size_t tmp_bytes, int *wit,
                                hipStream_t st);

}  // namespace paddi
