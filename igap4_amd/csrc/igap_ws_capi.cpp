// igap_ws_capi.cpp -- C ABI of the warm-start network (include/igap_ws.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/igap_ws.h"
#include "igap_ctx.h"
#include "igap_ws_impl.h"

using igap::ws::Net;
namespace W = igap::ws;

struct igap_ws_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    Net n;
};

static uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static void free_net(Net &n)
{
    void *bufs[] = {n.P, n.G, n.M, n.S, n.Wb, n.WTb, n.y, n.tg, n.x, n.part, n.bpart, n.lpart};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (int l = 0; l < W::NL; l++) {
        if (n.a[l]) (void)hipFree(n.a[l]);
        if (n.dz[l]) (void)hipFree(n.dz[l]);
    }
}

extern "C" int igap_ws_create(int device, const int grid[3], int cin, int cout, const int dil[IGAP_WS_LAYERS], uint64_t seed,
                              igap_ws_ctx **out)
{
    if (!grid || !dil || !out) return igap::fail(IGAP_ERR_ARG, "null argument");
    *out = nullptr;
    if (cin < 1 || cin > W::CH || cout < 1 || cout > W::CH) return igap::fail(IGAP_ERR_ARG, "cin and cout must be in 1..16");
    for (int d = 0; d < 3; d++)
        if (grid[d] < 1) return igap::fail(IGAP_ERR_ARG, "grid sizes must be positive");
    const int64_t V64 = (int64_t)grid[0] * grid[1] * grid[2];
    if (V64 * 16 > 0x7fffffffLL) return igap::fail(IGAP_ERR_ARG, "grid too large (%lld points)", (long long)V64);
    for (int l = 0; l < W::NL; l++)
        if (dil[l] < 1) return igap::fail(IGAP_ERR_ARG, "dilations must be positive");
    igap_ws_ctx *c = new igap_ws_ctx();
    c->device = device;
    Net &n = c->n;
    n.nx = grid[0];
    n.ny = grid[1];
    n.nz = grid[2];
    n.cin = cin;
    n.cout = cout;
    for (int l = 0; l < W::NL; l++) n.dil[l] = dil[l];
    const int V = (int)V64;
    n.nchunk = (V + W::CHUNK - 1) / W::CHUNK;
    const size_t np = (size_t)W::NL * W::LAYER_P;
    auto fail_free = [&](int rc) {
        free_net(n);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    };
#define WSCHK(x)                                                                                              \
    do {                                                                                                      \
        hipError_t e_ = (x);                                                                                  \
        if (e_ != hipSuccess) return fail_free(igap::fail(IGAP_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_))); \
    } while (0)
    WSCHK(hipSetDevice(device));
    WSCHK(hipStreamCreate(&c->stream));
    WSCHK(hipMalloc(&n.P, sizeof(float) * np));
    WSCHK(hipMalloc(&n.G, sizeof(float) * np));
    WSCHK(hipMalloc(&n.M, sizeof(float) * np));
    WSCHK(hipMalloc(&n.S, sizeof(float) * np));
    WSCHK(hipMalloc(&n.Wb, sizeof(uint16_t) * W::NL * W::LAYER_W));
    WSCHK(hipMalloc(&n.WTb, sizeof(uint16_t) * W::NL * W::LAYER_W));
    for (int l = 0; l < W::NL; l++) {
        WSCHK(hipMalloc(&n.a[l], sizeof(uint16_t) * (size_t)V * 16));
        WSCHK(hipMalloc(&n.dz[l], sizeof(uint16_t) * (size_t)V * 16));
    }
    WSCHK(hipMalloc(&n.y, sizeof(float) * (size_t)V * cout));
    WSCHK(hipMalloc(&n.tg, sizeof(float) * (size_t)V * cout));
    WSCHK(hipMalloc(&n.x, sizeof(float) * (size_t)V * cin));
    WSCHK(hipMalloc(&n.part, sizeof(float) * 27 * (size_t)n.nchunk * 256));
    WSCHK(hipMalloc(&n.bpart, sizeof(float) * (size_t)n.nchunk * 16));
    WSCHK(hipMalloc(&n.lpart, sizeof(float) * (size_t)((V * 16 + 255) / 256)));
    // parameters: He-uniform hidden layers over the real input channels, zero last layer
    std::vector<float> P(np, 0.f);
    uint64_t st = seed;
    for (int l = 0; l < W::NL - 1; l++) {
        const int ci_n = (l == 0) ? cin : W::CH;
        const float a = std::sqrt(6.f / (27.f * (float)ci_n));
        for (int t = 0; t < 27; t++)
            for (int co = 0; co < W::CH; co++)
                for (int ci = 0; ci < ci_n; ci++) {
                    const double u = (double)(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
                    P[(size_t)l * W::LAYER_P + (t * 16 + co) * 16 + ci] = (float)((2.0 * u - 1.0) * a);
                }
    }
    WSCHK(hipMemcpy(n.P, P.data(), sizeof(float) * np, hipMemcpyHostToDevice));
    WSCHK(hipMemset(n.M, 0, sizeof(float) * np));
    WSCHK(hipMemset(n.S, 0, sizeof(float) * np));
    WSCHK(W::pack_weights(n, c->stream));
    WSCHK(hipStreamSynchronize(c->stream));
#undef WSCHK
    *out = c;
    return IGAP_OK;
}

extern "C" int igap_ws_destroy(igap_ws_ctx *c)
{
    if (!c) return IGAP_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_net(c->n);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return IGAP_OK;
}

extern "C" int igap_ws_nparams(igap_ws_ctx *c, int64_t *n)
{
    if (!c || !n) return igap::fail(IGAP_ERR_ARG, "null argument");
    *n = (int64_t)W::NL * W::LAYER_P;
    return IGAP_OK;
}

extern "C" int igap_ws_get_params(igap_ws_ctx *c, float *params)
{
    if (!c || !params) return igap::fail(IGAP_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(params, c->n.P, sizeof(float) * W::NL * W::LAYER_P, hipMemcpyDeviceToHost));
    return IGAP_OK;
}

extern "C" int igap_ws_set_params(igap_ws_ctx *c, const float *params)
{
    if (!c || !params) return igap::fail(IGAP_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(c->n.P, params, sizeof(float) * W::NL * W::LAYER_P, hipMemcpyHostToDevice));
    HIPCHK(W::pack_weights(c->n, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return IGAP_OK;
}

static int forward(igap_ws_ctx *c, const float *x)
{
    Net &n = c->n;
    const size_t V = (size_t)n.nx * n.ny * n.nz;
    HIPCHK(hipMemcpyAsync(n.x, x, sizeof(float) * V * n.cin, hipMemcpyHostToDevice, c->stream));
    HIPCHK(W::pack_input(n, n.x, n.a[0], c->stream));
    for (int l = 0; l < W::NL - 1; l++) HIPCHK(W::conv_forward(n, l, n.a[l], n.a[l + 1], nullptr, c->stream));
    HIPCHK(W::conv_forward(n, W::NL - 1, n.a[W::NL - 1], nullptr, n.y, c->stream));
    return IGAP_OK;
}

// loss and parameter gradients into n.G (after forward)
static int backward(igap_ws_ctx *c, const float *target, double *loss)
{
    Net &n = c->n;
    const size_t V = (size_t)n.nx * n.ny * n.nz;
    HIPCHK(hipMemcpyAsync(n.tg, target, sizeof(float) * V * n.cout, hipMemcpyHostToDevice, c->stream));
    HIPCHK(W::loss_grad(n, n.y, n.tg, n.dz[W::NL - 1], c->stream));
    for (int l = W::NL - 1; l >= 0; l--) {
        HIPCHK(W::weight_grad(n, l, n.a[l], n.dz[l], c->stream));
        if (l > 0) HIPCHK(W::conv_input_grad(n, l, n.dz[l], n.a[l], n.dz[l - 1], c->stream));
    }
    if (loss) {
        const int nb = (int)((V * 16 + 255) / 256);
        std::vector<float> lp(nb);
        HIPCHK(hipMemcpyAsync(lp.data(), n.lpart, sizeof(float) * nb, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        double s = 0.0;
        for (int i = 0; i < nb; i++) s += lp[i];
        *loss = s / ((double)V * n.cout);
    }
    return IGAP_OK;
}

extern "C" int igap_ws_forward(igap_ws_ctx *c, const float *x, float *y)
{
    if (!c || !x || !y) return igap::fail(IGAP_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    int rc = forward(c, x);
    if (rc) return rc;
    const size_t V = (size_t)c->n.nx * c->n.ny * c->n.nz;
    HIPCHK(hipMemcpyAsync(y, c->n.y, sizeof(float) * V * c->n.cout, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return IGAP_OK;
}

extern "C" int igap_ws_gradients(igap_ws_ctx *c, const float *x, const float *target, double *loss, float *grad)
{
    if (!c || !x || !target || !grad) return igap::fail(IGAP_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    int rc = forward(c, x);
    if (!rc) rc = backward(c, target, loss);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(grad, c->n.G, sizeof(float) * W::NL * W::LAYER_P, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return IGAP_OK;
}

extern "C" int igap_ws_train_step(igap_ws_ctx *c, const float *x, const float *target, float lr, double *loss)
{
    if (!c || !x || !target) return igap::fail(IGAP_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    int rc = forward(c, x);
    if (!rc) rc = backward(c, target, loss);
    if (rc) return rc;
    Net &n = c->n;
    n.step++;
    const float c1 = 1.f - std::pow(W::ADAM_B1, (float)n.step), c2 = 1.f - std::pow(W::ADAM_B2, (float)n.step);
    HIPCHK(W::adam(n, lr, c1, c2, c->stream));
    HIPCHK(W::pack_weights(n, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return IGAP_OK;
}

extern "C" int igap_ws_selftest_mfma_bf16(float *out)
{
    if (!out) return igap::fail(IGAP_ERR_ARG, "null argument");
    float *d = nullptr;
    HIPCHK(hipMalloc(&d, sizeof(float) * 256));
    hipError_t e = W::selftest_bf16(d, 0);
    if (e == hipSuccess) e = hipMemcpy(out, d, sizeof(float) * 256, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return igap::fail(IGAP_ERR_HIP, "bf16 selftest: %s", hipGetErrorString(e));
    return IGAP_OK;
}
