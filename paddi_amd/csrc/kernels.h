// Kernel declarations and engine-wide constants (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace paddi {

constexpr int PB = 4096;       // prefix-sum block (points); also the largest supported m
constexpr int MAXM = 2048;     // largest subsequence length the scan kernels accept
constexpr int MAXU = 8;        // candidates per exact-refine launch
constexpr int64_t XPAD = 1024; // zero padding after the series (floats)

__global__ void k_prefix(const float *__restrict__ x, int64_t n, double2 *__restrict__ P,
                         double *__restrict__ refs);
__global__ void k_derive(const float *__restrict__ x, const double2 *__restrict__ P,
                         const double *__restrict__ refs, int64_t N, int64_t Npad, int m,
                         float4 *__restrict__ colv, double2 *__restrict__ st64);
__global__ void k_scan_select(const float *__restrict__ x, const float4 *__restrict__ colv,
                              uint64_t *__restrict__ alive, int *__restrict__ best, int64_t N,
                              int m, float thr_rho, int64_t WA);
__global__ void k_scan_refine(const float *__restrict__ x, const float4 *__restrict__ colv,
                              uint64_t *__restrict__ alive, int *__restrict__ best,
                              const int *__restrict__ blist, int64_t N, int m, float thr_rho,
                              int64_t WA, int64_t CH, int nchunks);
__global__ void k_fill_alive(uint64_t *__restrict__ alive, int64_t N, int64_t nwords);
__global__ void k_fill_int(int *__restrict__ p, int64_t n, int v);
__global__ void k_list_blocks(const uint64_t *__restrict__ alive, int64_t nwords,
                              int *__restrict__ list, int *__restrict__ count);
__global__ void k_survivors(const uint64_t *__restrict__ alive, const int *__restrict__ best,
                            int64_t N, int m, int *__restrict__ count, int64_t *__restrict__ idx,
                            float *__restrict__ d2, int64_t cap);
__global__ void k_exact(const float *__restrict__ x, const double2 *__restrict__ st64, int64_t N,
                        int m, const int64_t *__restrict__ cand, int nU,
                        unsigned long long *__restrict__ outkey);

// LDS bytes of the scan kernels for length m.
inline size_t scan_lds_bytes(int m)
{
    return 64 * 16 + (size_t)((m + 3) & ~3) * 4 + (size_t)(64 + m + 8) * 4;
}
// LDS bytes of k_exact for nU candidates of length m.
inline size_t exact_lds_bytes(int m, int nU)
{
    return ((size_t)nU * m + 256 + m + 8) * 8;
}
// Candidates per exact launch that fit 64 KiB of LDS at length m.
inline int exact_max_u(int m)
{
    int u = (int)((65536 / 8 - 264 - m) / m);
    return u < 1 ? 1 : (u > MAXU ? MAXU : u);
}

}  // namespace paddi
