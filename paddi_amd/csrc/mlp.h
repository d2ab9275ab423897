// Learned range proposal: sizes and kernels (mlp.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

namespace paddi {

constexpr int NFEAT = 36;  // m, global std, 32-bin pair histogram, series length, k
constexpr int HID = 64;                        // hidden width of both layers
constexpr int NPAIR = 1024;                    // seeded random pairs per (series, length)
constexpr int NPARAM = NFEAT * HID + HID + HID * HID + HID + HID + 1;
constexpr uint64_t FEAT_SEED = 0x5eed5eed1234ull;

__global__ void k_feat_pairs(const float *__restrict__ x, const SerDev *__restrict__ ser,
                             const double2 *__restrict__ st64, int nser, int m, uint64_t seed,
                             float *__restrict__ hist);
// one wave (64 threads) per series
__global__ void k_feat_std(const SerDev *__restrict__ ser, const double2 *__restrict__ P,
                           const double4 *__restrict__ refs, int nser, float *__restrict__ gstd);
__global__ void k_mlp_forward(const float *__restrict__ X, int B, const float *__restrict__ prm,
                              const float *__restrict__ fmu, const float *__restrict__ fsd,
                              float *__restrict__ H1, float *__restrict__ H2,
                              float *__restrict__ Y);
__global__ void k_mlp_back_sample(const float *__restrict__ prm, const float *__restrict__ H1,
                                  const float *__restrict__ H2, const float *__restrict__ Y,
                                  const float *__restrict__ Tg, int B, float *__restrict__ DY,
                                  float *__restrict__ D1, float *__restrict__ D2);
__global__ void k_mlp_back_params(const float *__restrict__ X, int B,
                                  const float *__restrict__ fmu, const float *__restrict__ fsd,
                                  const float *__restrict__ H1, const float *__restrict__ H2,
                                  const float *__restrict__ DY, const float *__restrict__ D1,
                                  const float *__restrict__ D2, float *__restrict__ G);
__global__ void k_adam(float *__restrict__ prm, const float *__restrict__ G,
                       float *__restrict__ M1, float *__restrict__ M2, float lr, float b1,
                       float b2, float eps, float c1, float c2, float wd);
__global__ void k_sq_err(const float *__restrict__ Y, const float *__restrict__ Tg, int B,
                         float *__restrict__ loss);

// Model: feature standardisation, parameters, and the log-margin subtracted from the
// prediction so that the proposal usually stays below the k-th discord distance.
struct Model {
    float fmu[NFEAT];
    float fsd[NFEAT];
    float prm[NPARAM];
    float margin;   // subtracted from the predicted log(d_k / 2m)
    bool loaded = false;
};

bool model_load(Model &M, const char *path);
bool model_save(const Model &M, const char *path);
const char *default_model_path();

}  // namespace paddi
