// Command-line drop-in for the reference's runner binary (src/main.cu:28-33):
//
//   paddi_main <series.txt> <n> <minL> <maxL> <discords.csv> <times.csv>
//              [--top-k K] [--step S] [--mode plain|batched|learned] [--device D]
//              [--range-profile FILE --profile-m M --profile-r2 R2]
//
// The six positional arguments mean what they mean for `main`, so the reference's runner
// scripts can call this binary in its place (without mpirun: one process, one GPU).
//
// Input (src/IOdata.hpp:40-55): whitespace-separated values; the first n are used.  A file
// with fewer values gives a shorter series (the reference pads with FLT_MAX instead; the
// library works with explicit bounds).
//
// Output, appended like the reference's writers:
//   discords.csv (src/IOdata.hpp:111-137): "m;top-1 discord idx;top-1 discord nnDist", then
//     one row per length (top_k rows with --top-k, in rank order); nnDist is the squared
//     z-normalised distance, as the reference writes it.
//   times.csv (src/IOdata.hpp:140-177): the reference's 14 columns, one row per (m, r)
//     attempt as src/main.cu:144,176,219 add them; one GPU, so the MPI and global-phase
//     columns are 0.  Phase 1 = statistics + scans, phase 2 = the double-precision
//     refinement, "# Candidates" = rows left by the scans, "# Discords" = discords the
//     attempt accepted.
//   --range-profile (src/IOdata.hpp:90-108): "m;v_0;..;v_{N-1}" at squared range R2, v_i the
//     NN distance of range discord i and -FLT_MAX for pruned rows (the reference's nnDist
//     after refinement, src/PD3_kernels.cuh:375).
// Exit status: 0 on success, 2 on bad arguments or input, 1 on a library error.
#include <cfloat>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/paddi.h"

static int usage(const char *msg)
{
    std::fprintf(stderr,
                 "paddi_main: %s\nusage: paddi_main <series.txt> <n> <minL> <maxL> "
                 "<discords.csv> <times.csv> [--top-k K] [--step S] "
                 "[--mode plain|batched|learned] [--device D] "
                 "[--range-profile FILE --profile-m M --profile-r2 R2]\n",
                 msg);
    return 2;
}

static bool read_series(const char *path, long long n, std::vector<float> &x)
{
    std::ifstream in(path);
    if (!in.is_open()) return false;
    float v;
    while ((long long)x.size() < n && (in >> v)) x.push_back(v);
    return true;
}

static int fail(paddi_ctx *ctx, paddi_status st, const char *what)
{
    std::fprintf(stderr, "paddi_main: %s: %s\n", what, paddi_status_str(st));
    if (ctx) paddi_destroy(ctx);
    return 1;
}

int main(int argc, char **argv)
{
    if (argc < 7) return usage("expected six positional arguments");
    const char *series_path = argv[1];
    const long long n = std::atoll(argv[2]);
    const int minL = std::atoi(argv[3]), maxL = std::atoi(argv[4]);
    const char *disc_path = argv[5], *times_path = argv[6];
    int top_k = 1, step = 1, device = 0, mode = PADDI_MODE_LEARNED, prof_m = 0;
    const char *prof_path = nullptr;
    double prof_r2 = -1.0;
    for (int a = 7; a < argc; a++) {
        const std::string o = argv[a];
        const bool has = a + 1 < argc;
        if (o == "--top-k" && has) top_k = std::atoi(argv[++a]);
        else if (o == "--step" && has) step = std::atoi(argv[++a]);
        else if (o == "--device" && has) device = std::atoi(argv[++a]);
        else if (o == "--range-profile" && has) prof_path = argv[++a];
        else if (o == "--profile-m" && has) prof_m = std::atoi(argv[++a]);
        else if (o == "--profile-r2" && has) prof_r2 = std::atof(argv[++a]);
        else if (o == "--mode" && has) {
            const std::string v = argv[++a];
            if (v == "plain") mode = PADDI_MODE_PLAIN;
            else if (v == "batched") mode = PADDI_MODE_BATCHED;
            else if (v == "learned") mode = PADDI_MODE_LEARNED;
            else return usage("unknown mode");
        } else {
            return usage(("unknown option " + o).c_str());
        }
    }
    if (n <= 0 || minL < 4 || maxL < minL || top_k < 1 || step < 1)
        return usage("need n > 0, 4 <= minL <= maxL, top-k >= 1, step >= 1");
    if (prof_path && (prof_m < 4 || prof_r2 < 0.0))
        return usage("--range-profile needs --profile-m M (>= 4) and --profile-r2 R2 (>= 0)");
    std::vector<float> x;
    if (!read_series(series_path, n, x)) return usage("cannot read the series file");
    if ((long long)x.size() < (long long)maxL + 1) return usage("series shorter than maxL + 1");
    if (mode == PADDI_MODE_LEARNED && !paddi_has_mode(mode)) mode = PADDI_MODE_BATCHED;

    paddi_ctx *ctx = nullptr;
    paddi_status st = paddi_create(device, &ctx);
    if (st != PADDI_OK) return fail(nullptr, st, "paddi_create");
    const int nlen = (maxL - minL) / step + 1;
    std::vector<paddi_discord> out((size_t)nlen * (size_t)top_k);
    // at most max_attempts (64) + 1 attempts per length
    std::vector<paddi_attempt_stats> pa((size_t)nlen * 65);
    int64_t count = 0, natt = 0;
    paddi_opts opts{};
    opts.mode = mode;
    st = paddi_discords_timed(ctx, x.data(), (int64_t)x.size(), minL, maxL, step, top_k, &opts,
                              out.data(), (int64_t)out.size(), &count, pa.data(),
                              (int64_t)pa.size(), &natt);
    if (st != PADDI_OK) return fail(ctx, st, "paddi_discords_timed");
    pa.resize(std::min<size_t>(pa.size(), (size_t)natt));

    {
        std::ofstream f(disc_path, std::ios::app);
        f << "m;top-1 discord idx;top-1 discord nnDist\n";
        for (int64_t u = 0; u < count; u++) f << out[u].length << ";" << out[u].index << ";"
                                              << out[u].nn_dist2 << "\n";
    }
    {
        std::ofstream f(times_path, std::ios::app);
        static const char *cols[] = {
            "m", "Phase1: find of segment's discords (rank = 0), ms",
            "max statistics among GPUs (kernel), ms", "max PD3 among GPUs (kernel), ms",
            "Phase2: find of fragment's discords (rank = 0), ms",
            "max MatrixMultiply among GPUs (kernel), ms", "# Candidates", "malloc memory, ms",
            "MPI exchange C, ms", "Phase3: find of global discords (rank = 0), ms",
            "max MatrixMultiply among GPUs (kernel), ms", "MPI exchange D, ms", "# Discords",
            "Total time PALMAD (rank = 0), ms"};
        for (int c = 0; c < 14; c++) f << cols[c] << (c < 13 ? ";" : "\n");
        for (const paddi_attempt_stats &r : pa)
            f << r.m << ";" << r.ms_derive + r.ms_scan << ";" << r.ms_derive << ";" << r.ms_scan
              << ";" << r.ms_exact << ";" << r.ms_exact << ";" << r.survivors << ";0;0;0;0;0;"
              << r.discords << ";" << r.ms_total << "\n";
    }
    if (prof_path) {
        if ((long long)x.size() < (long long)prof_m + 1) {
            paddi_destroy(ctx);
            return usage("series shorter than profile-m + 1");
        }
        const int64_t N = (int64_t)x.size() - prof_m + 1;
        std::vector<int32_t> flag((size_t)N);
        std::vector<float> nn((size_t)N);
        st = paddi_range_discords(ctx, x.data(), (int64_t)x.size(), prof_m, (float)prof_r2,
                                  flag.data(), nn.data());
        if (st != PADDI_OK) return fail(ctx, st, "paddi_range_discords");
        std::ofstream f(prof_path, std::ios::app);
        f << prof_m << ";";
        for (int64_t i = 0; i < N; i++)
            f << (flag[(size_t)i] ? nn[(size_t)i] : -FLT_MAX) << (i + 1 < N ? ";" : "\n");
    }
    paddi_destroy(ctx);
    return 0;
}
