// First-length witnesses from SAX words.
//
// At a call's first length no witness tiles exist (they are recorded by the scans of the
// previous length), so every block scans its select window and then its refine rounds until
// each row meets a neighbour within the range; rows whose nearest neighbours lie far away
// scan most of the series.  Rows with the same SAX word (the z-normalised window's PAA over
// 8 segments, each quantised by the standard normal's quartiles: 2 bits per segment) are
// likely near neighbours (HOTSAX's inner-loop ordering, Keogh et al. 2005): here each row
// gets as witness the tile of the next row of its series with the same word at least m + 1
// away (or the previous one), and the first length's scans try those tiles first.  Witnesses
// only order the scans: the discords do not depend on them.
#include "kernels.h"

#include <hipcub/hipcub.hpp>

namespace paddi {

// Per global row g: its key (series << 16 | word; all ones for padding rows, windows of zero
// spread and lengths below 8), its row as the sort's value, and no witness yet.
__global__ __launch_bounds__(256) void k_sax_words(const float *__restrict__ x,
                                                   const SerDev *__restrict__ ser,
                                                   const int *__restrict__ bser,
                                                   const float4 *__restrict__ colv,
                                                   int64_t rows, int m,
                                                   unsigned long long *__restrict__ keys,
                                                   int *__restrict__ vals, int *__restrict__ wit)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= rows) return;
    unsigned long long key = ~0ull;
    const int sid = bser[g >> 6];
    if (sid >= 0 && m >= 8) {
        const SerDev S = ser[sid];
        const int64_t i = g - S.roff;
        const float4 cv = colv[g];
        if (i < S.n - m + 1 && cv.z > 0.f) {
            const float *xs = x + S.xoff + i;
            const float ref = xs[0];
            const float scale = cv.z * sqrtf((float)m);  // 1 / sigma
            unsigned word = 0;
            for (int s = 0; s < 8; s++) {
                const int b0 = (s * m) >> 3, b1 = ((s + 1) * m) >> 3;
                float sum = 0.f;
                for (int k = b0; k < b1; k++) sum += xs[k] - ref;
                const float z = (sum / (float)(b1 - b0) + (ref - cv.w)) * scale;
                const unsigned code = (z > -0.6745f) + (z > 0.f) + (z > 0.6745f);
                word = (word << 2) | code;
            }
            key = ((unsigned long long)sid << 16) | word;
        }
    }
    keys[g] = key;
    vals[g] = (int)g;
    wit[g] = -1;
}

// Sorted position p (keys by series and word, rows ascending within a word): the first row of
// the same word at least m + 1 after this one, else the last one at least m + 1 before it,
// each found by a binary search over the next (previous) SAX_REACH positions; its tile
// (local column, a multiple of 64) becomes the row's witness.
constexpr int SAX_REACH = 4096;

__global__ __launch_bounds__(256) void k_sax_witness(const unsigned long long *__restrict__ keys,
                                                     const int *__restrict__ vals, int64_t rows,
                                                     const SerDev *__restrict__ ser, int m,
                                                     int *__restrict__ wit)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= rows) return;
    const unsigned long long key = keys[p];
    if (key == ~0ull) return;
    const int64_t g = vals[p];
    const int64_t roff = ser[key >> 16].roff;
    const int64_t d = m + 1;
    int64_t j = -1;
    {
        // first q in (p, hi] with keys[q] != key or vals[q] >= g + d (true from there on)
        int64_t lo = p + 1, hi = p + SAX_REACH < rows ? p + SAX_REACH : rows;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (keys[mid] != key || vals[mid] >= g + d) hi = mid; else lo = mid + 1;
        }
        if (lo < rows && keys[lo] == key && vals[lo] >= g + d) j = vals[lo];
    }
    if (j < 0) {
        // last q in [lo, p) with keys[q] != key or vals[q] <= g - d (true up to there)
        int64_t lo = p - SAX_REACH > 0 ? p - SAX_REACH : 0, hi = p;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (keys[mid] != key || vals[mid] <= g - d) lo = mid + 1; else hi = mid;
        }
        const int64_t q = lo - 1;
        if (q >= 0 && keys[q] == key && vals[q] <= g - d) j = vals[q];
    }
    if (j >= 0) wit[g] = (int)((j - roff) & ~(int64_t)63);
}

size_t sax_sort_bytes(int64_t rows)
{
    size_t bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const unsigned long long *)nullptr,
                                           (unsigned long long *)nullptr, (const int *)nullptr,
                                           (int *)nullptr, (int)rows, 0, 64) != hipSuccess)
        return 0;
    return bytes;
}

hipError_t launch_sax_witnesses(const float *x, const SerDev *ser, const int *bser,
                                const float4 *colv, int64_t rows, int m, int nser,
                                unsigned long long *keys, unsigned long long *keys2, int *vals,
                                int *vals2, void *tmp, size_t tmp_bytes, int *wit,
                                hipStream_t st)
{
    const unsigned grid = (unsigned)((rows + 255) / 256);
    hipLaunchKernelGGL(k_sax_words, dim3(grid), dim3(256), 0, st, x, ser, bser, colv, rows, m,
                       keys, vals, wit);
    int bits = 16;
    while (bits < 64 && (1ll << (bits - 16)) <= nser) bits++;
    size_t b = tmp_bytes;
    const hipError_t e = hipcub::DeviceRadixSort::SortPairs(tmp, b, keys, keys2, vals, vals2,
                                                            (int)rows, 0, bits, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sax_witness, dim3(grid), dim3(256), 0, st, keys2, vals2, rows, ser, m,
                       wit);
    return hipGetLastError();
}

}  // namespace paddi
