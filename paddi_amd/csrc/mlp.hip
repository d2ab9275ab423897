// Learned range proposal: feature extraction and a small MLP (forward, backward, Adam),
// hand-written for gfx950.  No counterpart in the reference, whose range schedule is the
// blind r-descent of src/main.cu:136-233.
//
// Features of (series s, length m), NFEAT = 34:
//   f0       m / 512
//   f1       log10 of the series' global standard deviation, / 4
//   f2..f33  histogram (fractions) of ED^2 / (2m) in 32 bins over [0, 2) of NPAIR seeded
//            pairs of non-overlapping windows (i uniform, j = i + d mod N, d in [m, N-m))
// Network: NFEAT -> 64 (ReLU) -> 64 (ReLU) -> 1, single precision; inputs are standardised
// with the training set's mean and standard deviation.  Parameters are one flat float array:
//   W1[NFEAT][H] b1[H] W2[H][H] b2[H] W3[H] b3.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include "kernels.h"
#include "mlp.h"

namespace paddi {

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    // splitmix64 finaliser
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// One wave per (series, pair): centred dot product over m, histogram bin by atomics.
__global__ __launch_bounds__(256) void k_feat_pairs(const float *__restrict__ x,
                                                    const SerDev *__restrict__ ser,
                                                    const double2 *__restrict__ st64, int nser,
                                                    int m, uint64_t seed,
                                                    float *__restrict__ hist /*[nser][32]*/)
{
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    const int sid = wave / NPAIR;
    const int p = wave % NPAIR;
    if (sid >= nser) return;
    const SerDev S = ser[sid];
    const int64_t N = S.n - m + 1;
    if (N <= 2 * (int64_t)m + 2) return;
    const uint64_t h0 = mix64(seed ^ ((uint64_t)sid << 32) ^ (uint64_t)p);
    const uint64_t h1 = mix64(h0);
    const int64_t i = (int64_t)(h0 % (uint64_t)N);
    const int64_t d = m + (int64_t)(h1 % (uint64_t)(N - 2 * m));
    const int64_t j = (i + d) % N;
    const float *xs = x + S.xoff;
    const double2 si = st64[S.roff + i], sj = st64[S.roff + j];
    const float mi = (float)si.x, mj = (float)sj.x;
    float acc = 0.f;
    for (int k = lane; k < m; k += 64) acc = fmaf(xs[i + k] - mi, xs[j + k] - mj, acc);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) {
        float q = 1.0f;  // ED^2/(2m) = 1 - rho; degenerate windows count as rho = 0
        if (si.y > 0.0 && sj.y > 0.0) q = 1.0f - acc * (float)(si.y * sj.y);
        int bin = (int)(q * 16.0f);
        bin = bin < 0 ? 0 : (bin > 31 ? 31 : bin);
        atomicAdd(hist + sid * 32 + bin, 1.0f / NPAIR);
    }
}

// Global standard deviation of each series from its prefix blocks (one thread per series).
__global__ __launch_bounds__(64) void k_feat_std(const SerDev *__restrict__ ser,
                                                 const double2 *__restrict__ P,
                                                 const double4 *__restrict__ refs, int nser,
                                                 float *__restrict__ gstd)
{
    // one wave per series, its lanes over the series' prefix blocks (one thread per series
    // read c3's 2,048 blocks one after another: 830 us)
    const int sid = blockIdx.x;
    if (sid >= nser) return;
    const SerDev S = ser[sid];
    const int64_t nb = (S.n + PB - 1) / PB;
    double s1 = 0.0, s2 = 0.0;
    const double c0 = refs[S.pboff].x;
    for (int64_t b = threadIdx.x; b < nb; b += 64) {
        const double2 t = P[(S.pboff + b) * (PB + 1) + PB];
        const double cnt = (double)((b + 1) * PB <= S.n ? PB : S.n - b * PB);
        const double D = refs[S.pboff + b].x - c0;  // shift block sums to the reference c0
        s1 += t.x + cnt * D;
        s2 += t.y + 2.0 * D * t.x + cnt * D * D;
    }
    for (int o = 32; o >= 1; o >>= 1) {
        s1 += __shfl_xor(s1, o, 64);
        s2 += __shfl_xor(s2, o, 64);
    }
    if (threadIdx.x != 0) return;
    const double mu = s1 / S.n;
    const double var = s2 / S.n - mu * mu;
    gstd[sid] = (float)sqrt(var > 0.0 ? var : 0.0);
}

// ---------------------------------------------------------------------------------------
// MLP.  One block of H = 64 threads per sample.

__global__ __launch_bounds__(64) void k_mlp_forward(const float *__restrict__ X, int B,
                                                    const float *__restrict__ prm,
                                                    const float *__restrict__ fmu,
                                                    const float *__restrict__ fsd,
                                                    float *__restrict__ H1, float *__restrict__ H2,
                                                    float *__restrict__ Y)
{
    const int b = blockIdx.x;
    const int u = threadIdx.x;
    if (b >= B) return;
    const float *W1 = prm, *b1 = W1 + NFEAT * HID, *W2 = b1 + HID, *b2 = W2 + HID * HID,
                *W3 = b2 + HID, *b3 = W3 + HID;
    __shared__ float xin[NFEAT], h1s[HID], red[HID];
    if (u < NFEAT) xin[u] = (X[b * NFEAT + u] - fmu[u]) / fsd[u];
    __syncthreads();
    float a = b1[u];
    for (int i = 0; i < NFEAT; i++) a = fmaf(xin[i], W1[i * HID + u], a);
    a = fmaxf(a, 0.f);
    h1s[u] = a;
    if (H1) H1[b * HID + u] = a;
    __syncthreads();
    float c = b2[u];
    for (int v = 0; v < HID; v++) c = fmaf(h1s[v], W2[v * HID + u], c);
    c = fmaxf(c, 0.f);
    if (H2) H2[b * HID + u] = c;
    red[u] = c * W3[u];
    __syncthreads();
    for (int o = HID / 2; o >= 1; o >>= 1) {
        if (u < o) red[u] += red[u + o];
        __syncthreads();
    }
    if (u == 0) Y[b] = red[0] + b3[0];
}

// Per-sample output gradients for the loss mean_b (y_b - t_b)^2.
__global__ __launch_bounds__(64) void k_mlp_back_sample(const float *__restrict__ prm,
                                                        const float *__restrict__ H1,
                                                        const float *__restrict__ H2,
                                                        const float *__restrict__ Y,
                                                        const float *__restrict__ Tg, int B,
                                                        float *__restrict__ DY,
                                                        float *__restrict__ D1,
                                                        float *__restrict__ D2)
{
    const int b = blockIdx.x;
    const int u = threadIdx.x;
    if (b >= B) return;
    const float *W2 = prm + NFEAT * HID + HID, *W3 = W2 + HID * HID + HID;
    __shared__ float d2s[HID];
    const float dy = 2.0f * (Y[b] - Tg[b]) / B;
    if (u == 0) DY[b] = dy;
    const float d2 = (H2[b * HID + u] > 0.f) ? dy * W3[u] : 0.f;
    d2s[u] = d2;
    D2[b * HID + u] = d2;
    __syncthreads();
    float d1 = 0.f;
    for (int v = 0; v < HID; v++) d1 = fmaf(d2s[v], W2[u * HID + v], d1);
    D1[b * HID + u] = (H1[b * HID + u] > 0.f) ? d1 : 0.f;
}

// Parameter gradients: one thread per parameter, summing over the batch.
__global__ __launch_bounds__(256) void k_mlp_back_params(const float *__restrict__ X, int B,
                                                         const float *__restrict__ fmu,
                                                         const float *__restrict__ fsd,
                                                         const float *__restrict__ H1,
                                                         const float *__restrict__ H2,
                                                         const float *__restrict__ DY,
                                                         const float *__restrict__ D1,
                                                         const float *__restrict__ D2,
                                                         float *__restrict__ G)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    float g = 0.f;
    if (p < NFEAT * HID) {
        const int i = p / HID, u = p % HID;
        for (int b = 0; b < B; b++)
            g = fmaf((X[b * NFEAT + i] - fmu[i]) / fsd[i], D1[b * HID + u], g);
    } else if (p < NFEAT * HID + HID) {
        const int u = p - NFEAT * HID;
        for (int b = 0; b < B; b++) g += D1[b * HID + u];
    } else if (p < NFEAT * HID + HID + HID * HID) {
        const int q = p - NFEAT * HID - HID, v = q / HID, u = q % HID;
        for (int b = 0; b < B; b++) g = fmaf(H1[b * HID + v], D2[b * HID + u], g);
    } else if (p < NFEAT * HID + 2 * HID + HID * HID) {
        const int u = p - NFEAT * HID - HID - HID * HID;
        for (int b = 0; b < B; b++) g += D2[b * HID + u];
    } else if (p < NPARAM - 1) {
        const int u = p - NFEAT * HID - 2 * HID - HID * HID;
        for (int b = 0; b < B; b++) g = fmaf(H2[b * HID + u], DY[b], g);
    } else {
        for (int b = 0; b < B; b++) g += DY[b];
    }
    G[p] = g;
}

__global__ __launch_bounds__(256) void k_adam(float *__restrict__ prm, const float *__restrict__ G,
                                              float *__restrict__ M1, float *__restrict__ M2,
                                              float lr, float b1, float b2, float eps, float c1,
                                              float c2, float wd)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    const float g = G[p];
    const float m1 = b1 * M1[p] + (1.f - b1) * g;
    const float m2 = b2 * M2[p] + (1.f - b2) * g * g;
    M1[p] = m1;
    M2[p] = m2;
    prm[p] -= lr * ((m1 / c1) / (sqrtf(m2 / c2) + eps) + wd * prm[p]);
}

__global__ __launch_bounds__(256) void k_sq_err(const float *__restrict__ Y,
                                                const float *__restrict__ Tg, int B,
                                                float *__restrict__ loss)
{
    __shared__ float red[256];
    float s = 0.f;
    for (int b = threadIdx.x; b < B; b += 256) {
        const float e = Y[b] - Tg[b];
        s += e * e;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o >= 1; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = red[0] / B;
}

}  // namespace paddi
