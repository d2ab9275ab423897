// Host engine of the discord search (one HIP device, one stream).
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <stdint.h>
#include <vector>

#include "../../include/paddi.h"

namespace paddi {

const char *status_str(paddi_status s);
double delta_ed2(int m);

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void reserve(size_t nbytes);
    void release();
    ~DevBuf() { release(); }
};

// A series resident in HBM with its prefix sums.
struct Series {
    int64_t n = 0;
    int64_t nb = 0;            // prefix blocks
    DevBuf x;                  // n + XPAD floats, zero padded
    DevBuf P;                  // nb * (PB+1) double2
    DevBuf refs;               // nb doubles
    bool prefix_ready = false;
    const float *xp() const { return (const float *)x.p; }
};

// Surviving row: single-precision distance a, exact distance v once computed.
struct Cand {
    int64_t i;
    float a;
    double v;
    bool exact;
};

class Engine {
public:
    explicit Engine(int device);
    ~Engine();
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    Series *upload(const float *T, int64_t n);
    void prefix(Series &s);
    void derive(Series &s, int m);
    void pass(Series &s, int m, double r2, std::vector<Cand> &out, paddi_stats *st);
    void exact(Series &s, int m, const std::vector<int64_t> &idx, std::vector<double> &out,
               paddi_stats *st);
    void select_topk(Series &s, int m, double r2, int k, std::vector<Cand> &cands,
                     std::vector<Cand> &acc, paddi_stats *st);
    void palmad(Series &s, int minL, int maxL, int step, int k, const paddi_opts &o,
                std::vector<paddi_discord> &out, paddi_stats *st, int series_id);
    void run(Series &s, int minL, int maxL, int step, int k, const paddi_opts &o,
             std::vector<paddi_discord> &out, paddi_stats *st, int series_id)
    {
        palmad(s, minL, maxL, step, k, o, out, st, series_id);
    }
    void copy_st64(double2 *dst, int64_t N)
    {
        (void)hipMemcpy(dst, st64_.p, (size_t)N * sizeof(double2), hipMemcpyDeviceToHost);
    }

    int64_t select_window_ = 8192;   // columns right of a block scanned by k_scan_select
    int64_t refine_chunk_ = 65536;   // columns per k_scan_refine work item

private:
    int device_;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev_[4] = {};
    void *pin_ = nullptr;
    DevBuf colv_, st64_, alive_, best_, blist_, cnt_, sidx_, sd2_, cand_, okey_;
};

}  // namespace paddi
