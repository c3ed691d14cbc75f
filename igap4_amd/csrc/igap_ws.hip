// igap_ws.hip -- learned warm start for the Newton load/time steps: a small
// 3-D convolutional network on the control-point grid, trained and evaluated
// with BF16 matrix-core products and FP32 accumulation (gfx950).
//
// Network: NL = 4 layers of 3x3x3 convolutions (zero padding, per-layer
// dilation), 16 channels between layers, ReLU after layers 0..2, linear last
// layer.  Input channels (<= 16) and output channels (<= 16) are zero-padded
// to 16 so that one 16x16x16 BF16 product covers one tap.
//
// Data layout (all device, channels-last):
//   activations  a[v][16]   BF16 bits, v = (x*ny + y)*nz + z  (the control-
//                point numbering of the assembly, z fastest)
//   weights      W[t][co][ci] BF16 (forward B operand), WT[t][ci][co] (input-
//                gradient B operand), t = 9*(dx+1) + 3*(dy+1) + (dz+1)
//   master copy  FP32 params, per layer W[27][16][16] then b[16]; Adam moments
//
// Convolution as an implicit GEMM on V_MFMA_F32_16X16X16_BF16: M = 16 voxels,
// N = 16 output channels, K = 16 input channels of one tap, 27 taps
// accumulated in one FP32 tile.  Lane l supplies A[l&15][4(l>>4)+j] and
// B[4(l>>4)+j][l&15] (j = 0..3); result register r of lane l holds
// D[4(l>>4)+r][l&15] (pinned by igap_ws_selftest_mfma_bf16).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "igap_ws_impl.h"

namespace igap {
namespace ws {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
    // round to nearest even (inputs are finite)
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ float4_t mma(short4_t a, short4_t b, float4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

enum { FWD_RELU = 0, FWD_LAST = 1, BWD_MASK = 2 };

// One wave per tile of 16 consecutive voxels, all 16 output channels.
//   FWD_RELU  aout = bf16(relu(sum + bias))
//   FWD_LAST  yout[v][c] = sum + bias  (c < cout, FP32)
//   BWD_MASK  aout = bf16(sum * (mask > 0))   (input gradient through ReLU)
// sgn = +1 gathers in[v + off_t] (forward), -1 gathers in[v - off_t]
// (gradient with respect to the layer input).
template <int MODE>
__global__ __launch_bounds__(256) void k_conv(int nx, int ny, int nz, int dil, int sgn, const uint16_t *__restrict__ ain,
                                              const uint16_t *__restrict__ Wk, const float *__restrict__ bias,
                                              const uint16_t *__restrict__ mask, uint16_t *__restrict__ aout,
                                              float *__restrict__ yout, int cout)
{
    const int V = nx * ny * nz;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int v0 = (blockIdx.x * 4 + wv) * 16;
    if (v0 >= V) return;
    const int i = lane & 15, kg = lane >> 4;
    const int v = v0 + i;
    const bool vin = v < V;
    int x = 0, y = 0, z = 0;
    if (vin) {
        x = v / (ny * nz);
        y = (v / nz) % ny;
        z = v % nz;
    }
    float4_t acc = {0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < 27; t++) {
        const int xx = x + sgn * dil * (t / 9 - 1), yy = y + sgn * dil * ((t / 3) % 3 - 1), zz = z + sgn * dil * (t % 3 - 1);
        short4_t a = {0, 0, 0, 0};
        if (vin && xx >= 0 && xx < nx && yy >= 0 && yy < ny && zz >= 0 && zz < nz)
            a = *reinterpret_cast<const short4_t *>(ain + (size_t)((xx * ny + yy) * nz + zz) * 16 + 4 * kg);
        const short4_t b = *reinterpret_cast<const short4_t *>(Wk + (t * 16 + i) * 16 + 4 * kg);
        acc = mma(a, b, acc);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int vo = v0 + 4 * kg + r;
        if (vo >= V) continue;
        float val = acc[r];
        if (MODE == FWD_RELU) {
            val += bias[i];
            aout[(size_t)vo * 16 + i] = f2bf(val > 0.f ? val : 0.f);
        } else if (MODE == FWD_LAST) {
            if (i < cout) yout[(size_t)vo * cout + i] = val + bias[i];
        } else {
            aout[(size_t)vo * 16 + i] = f2bf(bf2f(mask[(size_t)vo * 16 + i]) > 0.f ? val : 0.f);
        }
    }
}

// Weight-gradient partials of one tap over one voxel chunk:
//   part[t][c][ci][co] = sum_{v in chunk c} a[v + off_t][ci] dz[v][co]
// M = ci, N = co, K = voxels (16 per product).  One wave per (chunk, tap).
__global__ __launch_bounds__(64) void k_dw(int nx, int ny, int nz, int dil, const uint16_t *__restrict__ ain,
                                           const uint16_t *__restrict__ dz, int chunk, int nchunk, float *__restrict__ part)
{
    const int V = nx * ny * nz;
    const int c = blockIdx.x, t = blockIdx.y;
    const int lane = threadIdx.x, i = lane & 15, kg = lane >> 4;
    const int vbeg = c * chunk, vend = (vbeg + chunk < V) ? vbeg + chunk : V;
    const int dx = dil * (t / 9 - 1), dy = dil * ((t / 3) % 3 - 1), dzz = dil * (t % 3 - 1);
    float4_t acc = {0.f, 0.f, 0.f, 0.f};
    for (int vb = vbeg; vb < vend; vb += 16) {
        short4_t a, b;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int v = vb + 4 * kg + j;
            short av = 0, bv = 0;
            if (v < vend) {
                const int x = v / (ny * nz), y = (v / nz) % ny, z = v % nz;
                const int xx = x + dx, yy = y + dy, zz = z + dzz;
                if (xx >= 0 && xx < nx && yy >= 0 && yy < ny && zz >= 0 && zz < nz)
                    av = (short)ain[(size_t)((xx * ny + yy) * nz + zz) * 16 + i];
                bv = (short)dz[(size_t)v * 16 + i];
            }
            a[j] = av;
            b[j] = bv;
        }
        acc = mma(a, b, acc);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) part[(((size_t)t * nchunk + c) * 16 + 4 * kg + r) * 16 + i] = acc[r];
}

// grad W[t][co][ci] = sum over chunks (in order) of part[t][c][ci][co]
__global__ void k_dw_reduce(int nchunk, const float *__restrict__ part, float *__restrict__ gW)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;   // t*256 + co*16 + ci
    if (id >= 27 * 256) return;
    const int t = id >> 8, co = (id >> 4) & 15, ci = id & 15;
    float s = 0.f;
    for (int c = 0; c < nchunk; c++) s += part[(((size_t)t * nchunk + c) * 16 + ci) * 16 + co];
    gW[id] = s;
}

// bias-gradient partials: bpart[c][co] = sum_{v in chunk c} dz[v][co]
__global__ void k_db(int V, const uint16_t *__restrict__ dz, int chunk, float *__restrict__ bpart)
{
    const int c = blockIdx.x, co = threadIdx.x;
    const int vbeg = c * chunk, vend = (vbeg + chunk < V) ? vbeg + chunk : V;
    float s = 0.f;
    for (int v = vbeg; v < vend; v++) s += bf2f(dz[(size_t)v * 16 + co]);
    bpart[c * 16 + co] = s;
}

__global__ void k_db_reduce(int nchunk, const float *__restrict__ bpart, float *__restrict__ gb)
{
    const int co = threadIdx.x;
    float s = 0.f;
    for (int c = 0; c < nchunk; c++) s += bpart[c * 16 + co];
    gb[co] = s;
}

// mean-square loss over V x cout: dz = bf16(2 (y - target) / (V cout)) on the
// first cout channels (0 on the rest); lpart[block] = sum of squared errors
__global__ __launch_bounds__(256) void k_loss(int V, int cout, const float *__restrict__ y, const float *__restrict__ tg,
                                              uint16_t *__restrict__ dz, float *__restrict__ lpart)
{
    __shared__ float red[256];
    const int id = blockIdx.x * 256 + threadIdx.x;   // v*16 + c
    float e2 = 0.f;
    if (id < V * 16) {
        const int v = id >> 4, c = id & 15;
        float g = 0.f;
        if (c < cout) {
            const float e = y[(size_t)v * cout + c] - tg[(size_t)v * cout + c];
            e2 = e * e;
            g = 2.f * e / ((float)V * (float)cout);
        }
        dz[id] = f2bf(g);
    }
    red[threadIdx.x] = e2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = red[0];
}

// FP32 input [V][cin] -> BF16 [V][16], zero padded
__global__ void k_pack_input(int V, int cin, const float *__restrict__ x, uint16_t *__restrict__ a)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= V * 16) return;
    const int v = id >> 4, c = id & 15;
    a[id] = c < cin ? f2bf(x[(size_t)v * cin + c]) : (uint16_t)0;
}

// BF16 copies of the master weights: W[l][t][co][ci] and WT[l][t][ci][co]
__global__ void k_pack_weights(int nl, const float *__restrict__ P, uint16_t *__restrict__ Wb, uint16_t *__restrict__ WTb)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;   // l*6912 + t*256 + co*16 + ci
    if (id >= nl * LAYER_W) return;
    const int l = id / LAYER_W, r = id - l * LAYER_W, t = r >> 8, co = (r >> 4) & 15, ci = r & 15;
    const uint16_t b = f2bf(P[(size_t)l * LAYER_P + r]);
    Wb[id] = b;
    WTb[(size_t)l * LAYER_W + t * 256 + ci * 16 + co] = b;
}

// Adam (bias-corrected), FP32 master parameters
__global__ void k_adam(int64_t n, float lr, float b1, float b2, float eps, float c1, float c2, float *__restrict__ P,
                       const float *__restrict__ G, float *__restrict__ M, float *__restrict__ S)
{
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const float g = G[id];
    const float m = b1 * M[id] + (1.f - b1) * g;
    const float s = b2 * S[id] + (1.f - b2) * g * g;
    M[id] = m;
    S[id] = s;
    P[id] -= lr * (m / c1) / (sqrtf(s / c2) + eps);
}

__global__ void k_selftest_bf16(float *out)
{
    const int l = threadIdx.x, i = l & 15, kg = l >> 4;
    short4_t a, b;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int k = 4 * kg + j;
        a[j] = (short)f2bf(1.f + (float)i + 0.25f * (float)k);   // A[i][k]
        b[j] = (short)f2bf(2.f - 0.5f * (float)i + (float)k);    // B[k][n = i]
    }
    const float4_t d = mma(a, b, (float4_t){0.f, 0.f, 0.f, 0.f});
#pragma unroll
    for (int r = 0; r < 4; r++) out[r * 64 + l] = d[r];
}

// ---------------------------------------------------------------- launchers
hipError_t conv_forward(const Net &n, int layer, const uint16_t *ain, uint16_t *aout, float *yout, hipStream_t s)
{
    const int V = n.nx * n.ny * n.nz;
    const dim3 g((unsigned)((V + 63) / 64)), b(256);
    const uint16_t *W = n.Wb + (size_t)layer * LAYER_W;
    const float *bias = n.P + (size_t)layer * LAYER_P + LAYER_W;
    if (layer < NL - 1)
        hipLaunchKernelGGL(k_conv<FWD_RELU>, g, b, 0, s, n.nx, n.ny, n.nz, n.dil[layer], 1, ain, W, bias, nullptr, aout, nullptr, 16);
    else
        hipLaunchKernelGGL(k_conv<FWD_LAST>, g, b, 0, s, n.nx, n.ny, n.nz, n.dil[layer], 1, ain, W, bias, nullptr, nullptr, yout, n.cout);
    return hipGetLastError();
}

hipError_t conv_input_grad(const Net &n, int layer, const uint16_t *dz, const uint16_t *mask, uint16_t *dzprev, hipStream_t s)
{
    const int V = n.nx * n.ny * n.nz;
    const dim3 g((unsigned)((V + 63) / 64)), b(256);
    hipLaunchKernelGGL(k_conv<BWD_MASK>, g, b, 0, s, n.nx, n.ny, n.nz, n.dil[layer], -1, dz, n.WTb + (size_t)layer * LAYER_W,
                       nullptr, mask, dzprev, nullptr, 16);
    return hipGetLastError();
}

hipError_t weight_grad(const Net &n, int layer, const uint16_t *ain, const uint16_t *dz, hipStream_t s)
{
    const int V = n.nx * n.ny * n.nz;
    hipLaunchKernelGGL(k_dw, dim3((unsigned)n.nchunk, 27), dim3(64), 0, s, n.nx, n.ny, n.nz, n.dil[layer], ain, dz, CHUNK, n.nchunk,
                       n.part);
    float *gW = n.G + (size_t)layer * LAYER_P;
    hipLaunchKernelGGL(k_dw_reduce, dim3(27), dim3(256), 0, s, n.nchunk, n.part, gW);
    hipLaunchKernelGGL(k_db, dim3((unsigned)n.nchunk), dim3(16), 0, s, V, dz, CHUNK, n.bpart);
    hipLaunchKernelGGL(k_db_reduce, dim3(1), dim3(16), 0, s, n.nchunk, n.bpart, gW + LAYER_W);
    return hipGetLastError();
}

hipError_t loss_grad(const Net &n, const float *y, const float *tg, uint16_t *dz, hipStream_t s)
{
    const int V = n.nx * n.ny * n.nz;
    hipLaunchKernelGGL(k_loss, dim3((unsigned)((V * 16 + 255) / 256)), dim3(256), 0, s, V, n.cout, y, tg, dz, n.lpart);
    return hipGetLastError();
}

hipError_t pack_input(const Net &n, const float *x, uint16_t *a, hipStream_t s)
{
    const int V = n.nx * n.ny * n.nz;
    hipLaunchKernelGGL(k_pack_input, dim3((unsigned)((V * 16 + 255) / 256)), dim3(256), 0, s, V, n.cin, x, a);
    return hipGetLastError();
}

hipError_t pack_weights(const Net &n, hipStream_t s)
{
    hipLaunchKernelGGL(k_pack_weights, dim3((unsigned)((NL * LAYER_W + 255) / 256)), dim3(256), 0, s, NL, n.P, n.Wb, n.WTb);
    return hipGetLastError();
}

hipError_t adam(const Net &n, float lr, float c1, float c2, hipStream_t s)
{
    const int64_t np = (int64_t)NL * LAYER_P;
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, np, lr, ADAM_B1, ADAM_B2, ADAM_EPS, c1, c2, n.P,
                       n.G, n.M, n.S);
    return hipGetLastError();
}

hipError_t selftest_bf16(float *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_selftest_bf16, dim3(1), dim3(64), 0, s, out);
    return hipGetLastError();
}

}  // namespace ws
}  // namespace igap
