// Learned range proposal: host side (features, model file, proposal, training loop) and its
// C ABI.  The network predicts t = log(d_k / 2m), d_k the k-th discord's squared distance at
// length m; the proposal is r2 = 2m * exp(t - margin).  The plain geometric descent remains
// the fallback: a proposal above d_k only costs another pass, never a different result.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <dlfcn.h>
#include <string>
#include <vector>

#include "../../include/paddi.h"
#include "engine.h"
#include "learned.h"
#include "mlp.h"

namespace paddi {

namespace {
const char MAGIC[8] = {'P', 'D', 'R', 'M', 'L', 'P', '0', '2'};
}  // namespace

const char *default_model_path()
{
    static std::string path;
    if (path.empty()) {
        Dl_info info;
        if (dladdr((void *)&default_model_path, &info) && info.dli_fname) {
            std::string so(info.dli_fname);
            const size_t p = so.find_last_of('/');
            path = (p == std::string::npos ? std::string(".") : so.substr(0, p)) +
                   "/model/r_mlp.bin";
        } else {
            path = "model/r_mlp.bin";
        }
    }
    return path.c_str();
}

bool model_load(Model &M, const char *path)
{
    M.loaded = false;
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    char mg[8];
    int32_t hdr[4];
    bool ok = std::fread(mg, 1, 8, f) == 8 && std::memcmp(mg, MAGIC, 8) == 0 &&
              std::fread(hdr, 4, 4, f) == 4 && hdr[0] == NFEAT && hdr[1] == HID &&
              hdr[2] == NPARAM && std::fread(M.fmu, 4, NFEAT, f) == (size_t)NFEAT &&
              std::fread(M.fsd, 4, NFEAT, f) == (size_t)NFEAT &&
              std::fread(M.prm, 4, NPARAM, f) == (size_t)NPARAM &&
              std::fread(&M.margin, 4, 1, f) == 1;
    std::fclose(f);
    M.loaded = ok;
    return ok;
}

bool model_save(const Model &M, const char *path)
{
    FILE *f = std::fopen(path, "wb");
    if (!f) return false;
    const int32_t hdr[4] = {NFEAT, HID, NPARAM, 0};
    bool ok = std::fwrite(MAGIC, 1, 8, f) == 8 && std::fwrite(hdr, 4, 4, f) == 4 &&
              std::fwrite(M.fmu, 4, NFEAT, f) == (size_t)NFEAT &&
              std::fwrite(M.fsd, 4, NFEAT, f) == (size_t)NFEAT &&
              std::fwrite(M.prm, 4, NPARAM, f) == (size_t)NPARAM &&
              std::fwrite(&M.margin, 4, 1, f) == 1;
    std::fclose(f);
    return ok;
}

bool learned_available()
{
    static int state = -1;
    if (state < 0) {
        Model M;
        state = model_load(M, default_model_path()) ? 1 : 0;
    }
    return state == 1;
}

// Features of every series of the batch at length m (derive(m) must be current).
void Engine::features(Batch &B, int m, int k, std::vector<float> &X)
{
    const int ns = B.nser;
    feat_.reserve((size_t)ns * 33 * 4);
    float *hist = feat_.as<float>();
    float *gstd = hist + (size_t)ns * 32;
    HIPC(hipMemsetAsync(hist, 0, (size_t)ns * 33 * 4, stream_));
    const int64_t waves = (int64_t)ns * NPAIR;
    hipLaunchKernelGGL(k_feat_pairs, dim3((unsigned)((waves * 64 + 255) / 256)), dim3(256), 0,
                       stream_, B.x.as<float>(), B.ser.as<SerDev>(), st64_.as<double2>(), ns, m,
                       FEAT_SEED, hist);
    hipLaunchKernelGGL(k_feat_std, dim3((unsigned)ns), dim3(64), 0, stream_,
                       B.ser.as<SerDev>(), B.P.as<double2>(), B.refs.as<double4>(), ns, gstd);
    std::vector<float> h((size_t)ns * 33);
    HIPC(hipMemcpyAsync(h.data(), hist, h.size() * 4, hipMemcpyDeviceToHost, stream_));
    HIPC(hipStreamSynchronize(stream_));
    X.assign((size_t)ns * NFEAT, 0.f);
    for (int q = 0; q < ns; q++) {
        float *r = X.data() + (size_t)q * NFEAT;
        r[0] = m / 512.0f;
        r[1] = std::log10(std::max(h[(size_t)ns * 32 + q], 1e-30f)) / 4.0f;
        for (int b = 0; b < 32; b++) r[2 + b] = h[(size_t)q * 32 + b];
        r[34] = (float)(std::log10((double)B.n[(size_t)q]) / 6.0);
        r[35] = k / 4.0f;
    }
}

void Engine::mlp_forward(const float *prm, const float *fmu, const float *fsd, const float *X,
                         int Bn, float *Y)
{
    mlpbuf_.reserve((size_t)(NPARAM + 2 * NFEAT + (size_t)Bn * (NFEAT + 1)) * 4);
    float *d = mlpbuf_.as<float>();
    float *dp = d, *dmu = dp + NPARAM, *dsd = dmu + NFEAT, *dX = dsd + NFEAT, *dY = dX + (size_t)Bn * NFEAT;
    HIPC(hipMemcpyAsync(dp, prm, NPARAM * 4, hipMemcpyHostToDevice, stream_));
    HIPC(hipMemcpyAsync(dmu, fmu, NFEAT * 4, hipMemcpyHostToDevice, stream_));
    HIPC(hipMemcpyAsync(dsd, fsd, NFEAT * 4, hipMemcpyHostToDevice, stream_));
    HIPC(hipMemcpyAsync(dX, X, (size_t)Bn * NFEAT * 4, hipMemcpyHostToDevice, stream_));
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)Bn), dim3(HID), 0, stream_, dX, Bn, dp, dmu,
                       dsd, (float *)nullptr, (float *)nullptr, dY);
    HIPC(hipMemcpyAsync(Y, dY, (size_t)Bn * 4, hipMemcpyDeviceToHost, stream_));
    HIPC(hipStreamSynchronize(stream_));
}

// Train (Adam, full batch) or, with steps == 0, only return the gradient and loss.
void Engine::mlp_train(float *prm, const float *fmu, const float *fsd, const float *X,
                       const float *T, int Bn, int steps, float lr, float wd, float *losses,
                       float *grad_out)
{
    const size_t nb = (size_t)Bn;
    // params, grad, Adam moments | fmu, fsd | X, T | H1, H2, D1, D2 | Y, DY | loss
    const size_t total = 4 * (size_t)NPARAM + 2 * NFEAT + nb * (NFEAT + 1) + nb * 4 * HID +
                         nb * 2 + 4;
    mlpbuf_.reserve(total * 4);
    float *d = mlpbuf_.as<float>();
    float *dp = d, *G = dp + NPARAM, *M1 = G + NPARAM, *M2 = M1 + NPARAM;
    float *dmu = M2 + NPARAM, *dsd = dmu + NFEAT, *dX = dsd + NFEAT, *dT = dX + nb * NFEAT;
    float *H1 = dT + nb, *H2 = H1 + nb * HID, *D1 = H2 + nb * HID, *D2 = D1 + nb * HID;
    float *Y = D2 + nb * HID, *DY = Y + nb, *loss = DY + nb;
    HIPC(hipMemcpyAsync(dp, prm, NPARAM * 4, hipMemcpyHostToDevice, stream_));
    HIPC(hipMemsetAsync(M1, 0, 2 * NPARAM * 4, stream_));
    HIPC(hipMemcpyAsync(dmu, fmu, NFEAT * 4, hipMemcpyHostToDevice, stream_));
    HIPC(hipMemcpyAsync(dsd, fsd, NFEAT * 4, hipMemcpyHostToDevice, stream_));
    HIPC(hipMemcpyAsync(dX, X, nb * NFEAT * 4, hipMemcpyHostToDevice, stream_));
    HIPC(hipMemcpyAsync(dT, T, nb * 4, hipMemcpyHostToDevice, stream_));
    const int S = steps > 0 ? steps : 1;
    const float b1 = 0.9f, b2 = 0.999f;
    float p1 = 1.f, p2 = 1.f;
    std::vector<float> lh((size_t)S);
    for (int it = 0; it < S; it++) {
        hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)Bn), dim3(HID), 0, stream_, dX, Bn, dp,
                           dmu, dsd, H1, H2, Y);
        hipLaunchKernelGGL(k_sq_err, dim3(1), dim3(256), 0, stream_, Y, dT, Bn, loss + 0);
        hipLaunchKernelGGL(k_mlp_back_sample, dim3((unsigned)Bn), dim3(HID), 0, stream_, dp, H1,
                           H2, Y, dT, Bn, DY, D1, D2);
        hipLaunchKernelGGL(k_mlp_back_params, dim3((NPARAM + 255) / 256), dim3(256), 0, stream_,
                           dX, Bn, dmu, dsd, H1, H2, DY, D1, D2, G);
        if (steps > 0) {
            p1 *= b1;
            p2 *= b2;
            hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, stream_, dp, G,
                               M1, M2, lr, b1, b2, 1e-8f, 1.f - p1, 1.f - p2, wd);
        }
        if (losses) HIPC(hipMemcpyAsync(&lh[(size_t)it], loss, 4, hipMemcpyDeviceToHost, stream_));
    }
    HIPC(hipGetLastError());
    if (grad_out) HIPC(hipMemcpyAsync(grad_out, G, NPARAM * 4, hipMemcpyDeviceToHost, stream_));
    if (steps > 0) HIPC(hipMemcpyAsync(prm, dp, NPARAM * 4, hipMemcpyDeviceToHost, stream_));
    HIPC(hipStreamSynchronize(stream_));
    if (losses) std::memcpy(losses, lh.data(), (size_t)S * 4);
}

void Engine::propose_r2(Batch &B, int m, int k, std::vector<double> &out)
{
    out.assign((size_t)B.nser, 0.0);
    if (!model_.loaded) {
        const char *p = model_path_.empty() ? default_model_path() : model_path_.c_str();
        if (!model_load(model_, p)) return;
    }
    std::vector<float> X, Y((size_t)B.nser);
    features(B, m, k, X);
    mlp_forward(model_.prm, model_.fmu, model_.fsd, X.data(), B.nser, Y.data());
    for (int q = 0; q < B.nser; q++)
        out[(size_t)q] = 2.0 * m * std::exp((double)Y[(size_t)q] - model_.margin);
}

void Engine::set_model_path(const char *p)
{
    const std::string s = p ? p : "";
    if (s != model_path_) {
        model_path_ = s;
        model_.loaded = false;
    }
}

}  // namespace paddi

// ------------------------------------------------------------------------------------------
// C ABI of the learned part (no counterpart in the reference).

extern "C" {

paddi_status paddi_mlp_sizes(int32_t *nfeat, int32_t *hidden, int32_t *nparam)
{
    if (!nfeat || !hidden || !nparam) return PADDI_ERR_INVALID;
    *nfeat = paddi::NFEAT;
    *hidden = paddi::HID;
    *nparam = paddi::NPARAM;
    return PADDI_OK;
}

paddi_status paddi_features(paddi_ctx *ctx, paddi_series *const *series, int32_t nseries,
                            int32_t m, int32_t top_k, float *out)
{
    if (!ctx || !series || nseries <= 0 || !out || m < 4 || m > paddi::MAXM || top_k < 1)
        return PADDI_ERR_INVALID;
    std::vector<const paddi::Series *> ss;
    int64_t rows = 0;
    for (int32_t q = 0; q < nseries; q++) {
        if (!series[q] || series[q]->ctx != ctx || series[q]->s->n < 3 * (int64_t)m + 8)
            return PADDI_ERR_INVALID;
        ss.push_back(series[q]->s);
        rows += paddi::series_rows(series[q]->s->n, m);
    }
    if (rows >= paddi::MAX_ROWS) return PADDI_ERR_UNSUPPORTED;
    return guarded(ctx, [&] {
        paddi::Engine &E = *ctx->eng;
        paddi::Batch B;
        E.build_batch(B, ss, m);
        E.prefix(B);
        E.derive(B, m, nullptr);
        std::vector<float> X;
        E.features(B, m, top_k, X);
        std::memcpy(out, X.data(), X.size() * 4);
        return PADDI_OK;
    });
}

paddi_status paddi_mlp_forward(paddi_ctx *ctx, const float *params, const float *fmu,
                               const float *fsd, const float *X, int32_t B, float *Y)
{
    if (!ctx || !params || !fmu || !fsd || !X || !Y || B <= 0) return PADDI_ERR_INVALID;
    return guarded(ctx, [&] {
        ctx->eng->mlp_forward(params, fmu, fsd, X, B, Y);
        return PADDI_OK;
    });
}

paddi_status paddi_mlp_gradient(paddi_ctx *ctx, const float *params, const float *fmu,
                                const float *fsd, const float *X, const float *T, int32_t B,
                                float *grad, float *loss)
{
    if (!ctx || !params || !fmu || !fsd || !X || !T || !grad || !loss || B <= 0)
        return PADDI_ERR_INVALID;
    return guarded(ctx, [&] {
        std::vector<float> p(params, params + paddi::NPARAM);
        ctx->eng->mlp_train(p.data(), fmu, fsd, X, T, B, 0, 0.f, 0.f, loss, grad);
        return PADDI_OK;
    });
}

paddi_status paddi_mlp_train(paddi_ctx *ctx, float *params, const float *fmu, const float *fsd,
                             const float *X, const float *T, int32_t B, int32_t steps, float lr,
                             float weight_decay, float *losses)
{
    if (!ctx || !params || !fmu || !fsd || !X || !T || B <= 0 || steps <= 0 || !(lr > 0.f))
        return PADDI_ERR_INVALID;
    return guarded(ctx, [&] {
        ctx->eng->mlp_train(params, fmu, fsd, X, T, B, steps, lr, weight_decay, losses, nullptr);
        return PADDI_OK;
    });
}

}  // extern "C"
