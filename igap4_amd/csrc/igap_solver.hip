// igap_solver.hip -- device-resident Newton-Krylov solve of the assembled system.
//
// The reference drives PETSc SNES NEWTONLS with a backtracking line search,
// KSP GMRES (right preconditioned, modified Gram-Schmidt, restart 10000) and
// PCASM/ILU(0) (src/interface.c:207-250), one SNESSolve per outer sub-
// iteration until ||F|| < 1e-10 (src/interface.c:302-328).  On one GPU the
// same algorithm runs with every vector in HBM:
//   * Newton with the BT line search (Armijo on |F|^2/2, quadratic then cubic
//     backtracking, lambda in [0.1, 0.5] of the previous step),
//   * restarted GMRES(m), right preconditioned, classical Gram-Schmidt with
//     one reorthogonalisation (two batched GEMVs per step instead of j
//     dependent dot products; m is bounded because 10000 Krylov vectors of
//     the large patches do not fit),
//   * ILU(0) of the assembled tangent in the tangent's precision (or FP32
//     with pc_fp32), FP64 or FP32, in the natural dof order (the
//     reference's; rocSPARSE csrilu0 + csrsv) or in a 27-colour order of the
//     control points (colour 9 (i mod 3) + 3 (j mod 3) + (k mod 3), points of
//     one colour share no stencil entry, the three dofs of a point stay
//     together), factored and solved by this file's point-block kernels: the
//     natural order's triangular solves have ~9 n_x levels of dependent rows,
//     the coloured order's 27 colour sweeps,
//   * or a geometric multigrid V-cycle on the patch's B-spline hierarchy
//     (knot-insertion transfers, coarse tangents re-assembled by the same
//     kernels, Chebyshev / point-block Jacobi smoothing, explicit inverse on
//     the coarsest level),
//   * a point-block SpMV (one wave per control point, a lane per neighbour
//     point, the nine values of the 3x3 block per lane).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsparse/rocsparse.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "igap_ctx.h"
#include "igap_launch.h"

namespace igap {

struct MGHier;
void mg_destroy(MGHier *g);

struct Solver {
    MGHier *mg = nullptr;               // multigrid hierarchy (IGAP_PC_MG), built on first use
    int m = 0;                      // GMRES restart
    int64_t n = 0;
    rocsparse_handle sp = nullptr;
    rocsparse_mat_descr dA = nullptr, dL = nullptr, dU = nullptr;
    rocsparse_mat_info info[2] = {nullptr, nullptr};   // per precision
    void *ilu[2] = {nullptr, nullptr};                 // factor values (double / float)
    void *buf[2] = {nullptr, nullptr};
    int analysed[2] = {0, 0};
    int32_t *rowptr32 = nullptr;
    rocblas_handle bl = nullptr;
    double *V = nullptr;            // (m+1) x n, column j = Krylov vector j
    double *w = nullptr, *z = nullptr, *t = nullptr, *d = nullptr, *x = nullptr, *xt = nullptr, *F = nullptr, *Ft = nullptr;
    double *h = nullptr;            // m+1 device
    float *fa = nullptr, *fb = nullptr;
    double *diag = nullptr;         // Jacobi
    double *t2 = nullptr;           // second permuted work vector
    // coloured ordering: perm[new] = old dof, permuted CSR and, per permuted
    // nonzero, its position in the natural CSR
    int32_t *mc_perm = nullptr, *mc_rowptr = nullptr, *mc_col = nullptr, *mc_src = nullptr, *mc_dpos = nullptr;
    int mc_cstart[28] = {};         // first row of each colour (rows of a colour are contiguous)
    int mc_sweep_ok = 0;            // no two points of one colour are coupled (false for a periodic
                                    // direction whose size is not a multiple of 3): colour sweeps valid
    int mc_block_ok = 0;            // point-block pattern (k_mcb_* applicable)
    int mc_ready = 0;               // the coloured ordering is complete
    // the preconditioner of the last refresh (igap_gpu_pc_apply)
    int pc_ready = 0, pc_kind = 0, pc_prec = 0, pc_ord = 0;
    int32_t *mc_pcoord = nullptr;   // per permuted point: owned-grid (i, j, k) packed 10 bits each
    int *zpiv = nullptr;            // first row with a zero pivot of the last own factorisation
    double *part = nullptr;         // partial dot products, (m+2) x ceil(n / DOT_CHUNK)
    rocsparse_mat_info mc_info[2] = {nullptr, nullptr};
    void *mc_ilu[2] = {nullptr, nullptr};
    void *mc_buf[2] = {nullptr, nullptr};
    int mc_analysed[2] = {0, 0};
};

void solver_destroy(Solver *s)
{
    if (!s) return;
    if (s->mg) mg_destroy(s->mg);
    if (s->sp) {
        for (int p = 0; p < 2; p++) {
            if (s->info[p]) rocsparse_destroy_mat_info(s->info[p]);
            if (s->mc_info[p]) rocsparse_destroy_mat_info(s->mc_info[p]);
        }
        if (s->dA) rocsparse_destroy_mat_descr(s->dA);
        if (s->dL) rocsparse_destroy_mat_descr(s->dL);
        if (s->dU) rocsparse_destroy_mat_descr(s->dU);
        rocsparse_destroy_handle(s->sp);
    }
    if (s->bl) rocblas_destroy_handle(s->bl);
    void *bufs[] = {s->ilu[0], s->ilu[1], s->buf[0], s->buf[1], s->rowptr32, s->V, s->w, s->z, s->t, s->d, s->x,
                    s->xt, s->F, s->Ft, s->h, s->fa, s->fb, s->diag, s->t2, s->mc_perm, s->mc_rowptr,
                    s->mc_col, s->mc_src, s->mc_dpos, s->part, s->mc_pcoord, s->zpiv, s->mc_ilu[0], s->mc_ilu[1], s->mc_buf[0], s->mc_buf[1]};
    // teardown: nothing to report to (the context's destroy carries on)
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete s;
}

// ---------------------------------------------------------------- kernels
template <typename T>
__global__ __launch_bounds__(256) void k_spmv(int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ col,
                                              const T *__restrict__ val, const double *__restrict__ x, double *__restrict__ y)
{
    int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (row >= n) return;
    double s = 0.0;
    for (int64_t j = rp[row] + lane; j < rp[row + 1]; j += 64) s += (double)val[j] * x[col[j]];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) y[row] = s;
}

// The assembled pattern is point-blocked: the three rows of control point p
// list the same neighbour points q, each as the columns 3q, 3q+1, 3q+2, and
// rowptr[3p] = 9 (number of blocks before p).  pcol[rowptr[3p]/9 + k] = q_k.
__global__ void k_point_cols(int64_t ncp, const int64_t *__restrict__ rp, const int32_t *__restrict__ col, int32_t *__restrict__ pcol)
{
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < ncp; p += st) {
        const int64_t b0 = rp[3 * p], nn = (rp[3 * p + 1] - b0) / 3;
        for (int64_t k = 0; k < nn; k++) pcol[b0 / 9 + k] = col[b0 + 3 * k] / 3;
    }
}

// y = A x on the point-block pattern: one wave per control point, lane k
// takes neighbour points k and k+64 (one index load, three x loads and the
// nine values of the block for the three rows)
// MODE 0: y = A x;  1: y = b - A x;  2: y -= A x
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_spmv_pb(int64_t ncp, const int64_t *__restrict__ rp, const int32_t *__restrict__ pcol,
                                                 const T *__restrict__ val, const double *__restrict__ x, const double *__restrict__ b,
                                                 double *__restrict__ y)
{
    const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (p >= ncp) return;
    const int64_t b0 = rp[3 * p], b1 = rp[3 * p + 1], b2 = rp[3 * p + 2];
    const int nn = (int)((b1 - b0) / 3);
    const int32_t *pc = pcol + b0 / 9;
    double s0 = 0, s1 = 0, s2 = 0;
    for (int k = lane; k < nn; k += 64) {
        const int64_t q = pc[k];
        const double x0 = x[3 * q], x1 = x[3 * q + 1], x2 = x[3 * q + 2];
        const int64_t o = 3 * k;
        s0 += (double)val[b0 + o] * x0 + (double)val[b0 + o + 1] * x1 + (double)val[b0 + o + 2] * x2;
        s1 += (double)val[b1 + o] * x0 + (double)val[b1 + o + 1] * x1 + (double)val[b1 + o + 2] * x2;
        s2 += (double)val[b2 + o] * x0 + (double)val[b2 + o + 1] * x1 + (double)val[b2 + o + 2] * x2;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o, 64);
        s1 += __shfl_xor(s1, o, 64);
        s2 += __shfl_xor(s2, o, 64);
    }
    if (lane == 0) {
        if (MODE == 0) {
            y[3 * p] = s0;
            y[3 * p + 1] = s1;
            y[3 * p + 2] = s2;
        } else if (MODE == 1) {
            y[3 * p] = b[3 * p] - s0;
            y[3 * p + 1] = b[3 * p + 1] - s1;
            y[3 * p + 2] = b[3 * p + 2] - s2;
        } else {
            y[3 * p] -= s0;
            y[3 * p + 1] -= s1;
            y[3 * p + 2] -= s2;
        }
    }
}

// 16-bit smoothing matrix (multigrid levels with pc_fp32 = 2): block k of
// control point p (its neighbour pcol[blk(p) + k], blk(p) = rowptr[3p] / 9,
// nn neighbours) is the 3x3 block row-major as binary16, the tenth half zero,
// in 5 words, word w at h16[5 blk(p) + w nn + k] (word-major per point, so
// that a wave's load of one word of 64 blocks is 256 contiguous bytes); row r
// of the block holds a / rs[3p + r], with rs a power of two per row that puts
// the row's largest magnitude in [2^14, 2^15) (no overflow; entries down to
// 2^-28 of it stay normal).  20 bytes per block against 36 in FP32.
__device__ __forceinline__ uint32_t pack_h2(float a, float b)
{
    const _Float16 ha = (_Float16)a, hb = (_Float16)b;
    return (uint32_t)__builtin_bit_cast(uint16_t, ha) | ((uint32_t)__builtin_bit_cast(uint16_t, hb) << 16);
}
__device__ __forceinline__ double h_lo(uint32_t w) { return (double)(float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu)); }
__device__ __forceinline__ double h_hi(uint32_t w) { return (double)(float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16)); }

// one wave per control point: the row scales, then the blocks
__global__ __launch_bounds__(256) void k_pack_h16(int64_t ncp, const int64_t *__restrict__ rp, const float *__restrict__ val,
                                                  uint32_t *__restrict__ h16, float *__restrict__ rs)
{
    const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (p >= ncp) return;
    const int64_t b[3] = {rp[3 * p], rp[3 * p + 1], rp[3 * p + 2]};
    const int nn = (int)((b[1] - b[0]) / 3);
    float inv[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        float m = 0.0f;
        for (int j = lane; j < 3 * nn; j += 64) m = fmaxf(m, fabsf(val[b[r] + j]));
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
        const int e = m > 0.0f ? ilogbf(m) - 14 : 0;
        inv[r] = ldexpf(1.0f, -e);
        if (lane == 0) rs[3 * p + r] = ldexpf(1.0f, e);
    }
    uint32_t *hb = h16 + 5 * (b[0] / 9);
    for (int k = lane; k < nn; k += 64) {
        float a[10];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int j = 0; j < 3; j++) a[3 * r + j] = val[b[r] + 3 * k + j] * inv[r];
        a[9] = 0.0f;
#pragma unroll
        for (int w = 0; w < 5; w++) hb[w * nn + k] = pack_h2(a[2 * w], a[2 * w + 1]);
    }
}

// Point grid of a patch without periodic directions, for products that
// compute their CSR positions: the rows of control point p = (px, py, pz)
// list the points q with |q_d - p_d| <= 2 (clipped to the grid) in
// lexicographic order, so its neighbour count, its first block
// (rowptr[3p] / 9) and its k-th neighbour are arithmetic.
struct PointGrid {
    int nb[3];
    int64_t L12;   // stencil_prefix(nb1, nb1) * stencil_prefix(nb2, nb2)
    int64_t L2;    // stencil_prefix(nb2, nb2)
};

// sum over the points a' < a of one direction (n points) of their clipped
// stencil lengths min(a' + 2, n - 1) - max(a' - 2, 0) + 1
__host__ __device__ __forceinline__ int64_t stencil_prefix(int64_t a, int64_t n)
{
    const int64_t m = a < (n - 2 > 0 ? n - 2 : 0) ? a : (n - 2 > 0 ? n - 2 : 0);
    const int64_t smin = m * (m - 1) / 2 + 2 * m + (a - m) * (n - 1);
    const int64_t smax = a >= 3 ? (a - 3) * (a - 2) / 2 : 0;
    return smin - smax + a;
}

// y = A x with the 16-bit blocks (the layout of k_spmv_pb, FP64 sums, the
// row scale applied once); MODE as k_spmv_pb.  Multigrid levels have no
// periodic direction, so the point's first block, its neighbour count and
// its neighbours come from the point grid G instead of rowptr and the
// neighbour index (the blocks and x are the only loads).  PPW control points
// per wave, 64 / PPW lanes each: a point's 20-byte blocks are few (at most
// 125), so a wave per point spends its time waiting on its loads rather
// than on the blocks; every load of a lane is issued before the sums.
template <int MODE, int PPW>
__global__ __launch_bounds__(256) void k_spmv_h16(int64_t ncp, PointGrid G, const uint32_t *__restrict__ h16, const float *__restrict__ rs,
                                                  const double *__restrict__ x, const double *__restrict__ b, double *__restrict__ y)
{
    constexpr int LW = 64 / PPW;   // lanes per control point
    const int lane = threadIdx.x & 63, sl = lane & (LW - 1);
    const int64_t p = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * PPW + lane / LW;
    const bool ok = p < ncp;
    double s0 = 0, s1 = 0, s2 = 0;
    if (ok) {
        const int n1 = G.nb[1], n2 = G.nb[2];
        const int pz = (int)(p % n2), py = (int)((p / n2) % n1), px = (int)(p / ((int64_t)n1 * n2));
        const int lox = px - 2 > 0 ? px - 2 : 0, loy = py - 2 > 0 ? py - 2 : 0, loz = pz - 2 > 0 ? pz - 2 : 0;
        const int lx = (px + 2 < G.nb[0] - 1 ? px + 2 : G.nb[0] - 1) - lox + 1;
        const int ly = (py + 2 < n1 - 1 ? py + 2 : n1 - 1) - loy + 1;
        const int lz = (pz + 2 < n2 - 1 ? pz + 2 : n2 - 1) - loz + 1;
        const int nn = lx * ly * lz, lyz = ly * lz;
        const int64_t blk = stencil_prefix(px, G.nb[0]) * G.L12 + lx * (stencil_prefix(py, n1) * G.L2 + ly * stencil_prefix(pz, n2));
        const uint32_t *hb = h16 + 5 * blk;
        // at most 125 neighbours: a fixed number of rounds
        constexpr int NIT = (125 + LW - 1) / LW;
        uint32_t w[NIT][5];
        double xv[NIT][3];
#pragma unroll
        for (int it = 0; it < NIT; it++) {
            const int k = sl + it * LW;
            if (k < nn) {
                const int kx = k / lyz, r = k - kx * lyz, ky = r / lz, kz = r - ky * lz;
                const int64_t q = ((int64_t)(lox + kx) * n1 + (loy + ky)) * n2 + (loz + kz);
#pragma unroll
                for (int j = 0; j < 3; j++) xv[it][j] = x[3 * q + j];
#pragma unroll
                for (int j = 0; j < 5; j++) w[it][j] = hb[j * nn + k];
            } else {
#pragma unroll
                for (int j = 0; j < 3; j++) xv[it][j] = 0.0;
#pragma unroll
                for (int j = 0; j < 5; j++) w[it][j] = 0u;
            }
        }
#pragma unroll
        for (int it = 0; it < NIT; it++) {
            s0 += h_lo(w[it][0]) * xv[it][0] + h_hi(w[it][0]) * xv[it][1] + h_lo(w[it][1]) * xv[it][2];
            s1 += h_hi(w[it][1]) * xv[it][0] + h_lo(w[it][2]) * xv[it][1] + h_hi(w[it][2]) * xv[it][2];
            s2 += h_lo(w[it][3]) * xv[it][0] + h_hi(w[it][3]) * xv[it][1] + h_lo(w[it][4]) * xv[it][2];
        }
    }
#pragma unroll
    for (int o = LW / 2; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o, 64);
        s1 += __shfl_xor(s1, o, 64);
        s2 += __shfl_xor(s2, o, 64);
    }
    if (ok && sl == 0) {
        s0 *= (double)rs[3 * p];
        s1 *= (double)rs[3 * p + 1];
        s2 *= (double)rs[3 * p + 2];
        if (MODE == 0) {
            y[3 * p] = s0;
            y[3 * p + 1] = s1;
            y[3 * p + 2] = s2;
        } else if (MODE == 1) {
            y[3 * p] = b[3 * p] - s0;
            y[3 * p + 1] = b[3 * p + 1] - s1;
            y[3 * p + 2] = b[3 * p + 2] - s2;
        } else {
            y[3 * p] -= s0;
            y[3 * p + 1] -= s1;
            y[3 * p + 2] -= s2;
        }
    }
}
constexpr int H16_PPW = 2;   // control points per wave of k_spmv_h16

// y[i] = x[idx[i]]  (values or vectors into the coloured order)
// (grid-stride: launched with grid1)
template <typename T, typename U>
__global__ void k_gather(int64_t n, const int32_t *__restrict__ idx, const T *__restrict__ x, U *__restrict__ y)
{
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st) y[i] = (U)x[idx[i]];
}

// y[idx[i]] = x[i]  (back to the natural order; grid-stride)
__global__ void k_scatter(int64_t n, const int32_t *__restrict__ idx, const double *__restrict__ x, double *__restrict__ y)
{
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st) y[idx[i]] = x[i];
}

// Triangular solves of the ILU(0) factor in the coloured order, one colour
// per launch: points of one colour share no stencil entry, so every point's
// three rows depend only on earlier colours (lower) or later ones (upper) and
// on the point's own 3x3 diagonal block, solved exactly by lane 0.  One wave
// per control point; rows r = r0 + 3 qi + d, d = 0..2.  The factor holds L
// (unit diagonal, strictly lower) and U (upper with diagonal) in place;
// dpos[r] is the position of row r's diagonal and the point's own columns
// sit next to it.
template <typename T>
__global__ __launch_bounds__(256) void k_mc_lower(int r0, int npts, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                  const int32_t *__restrict__ dpos, const T *__restrict__ f,
                                                  const double *__restrict__ b, double *__restrict__ y)
{
    const int qi = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (qi >= npts) return;
    const int r = r0 + 3 * qi;
    double s[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double a = 0.0;
        for (int j = rp[r + d] + lane; j < dpos[r + d] - d; j += 64) a += (double)f[j] * y[col[j]];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
        s[d] = a;
    }
    if (lane == 0) {
        const double y0 = b[r] - s[0];
        const double y1 = b[r + 1] - s[1] - (double)f[dpos[r + 1] - 1] * y0;
        const double y2 = b[r + 2] - s[2] - (double)f[dpos[r + 2] - 2] * y0 - (double)f[dpos[r + 2] - 1] * y1;
        y[r] = y0;
        y[r + 1] = y1;
        y[r + 2] = y2;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_mc_upper(int r0, int npts, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                  const int32_t *__restrict__ dpos, const T *__restrict__ f,
                                                  const double *__restrict__ y, double *__restrict__ z)
{
    const int qi = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (qi >= npts) return;
    const int r = r0 + 3 * qi;
    double s[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double a = 0.0;
        for (int j = dpos[r + d] + 3 - d + lane; j < rp[r + d + 1]; j += 64) a += (double)f[j] * z[col[j]];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
        s[d] = a;
    }
    if (lane == 0) {
        const double z2 = (y[r + 2] - s[2]) / (double)f[dpos[r + 2]];
        const double z1 = (y[r + 1] - s[1] - (double)f[dpos[r + 1] + 1] * z2) / (double)f[dpos[r + 1]];
        const double z0 = (y[r] - s[0] - (double)f[dpos[r] + 1] * z1 - (double)f[dpos[r] + 2] * z2) / (double)f[dpos[r]];
        z[r] = z0;
        z[r + 1] = z1;
        z[r + 2] = z2;
    }
}

// The same solves on the point-block structure of the pattern: the three rows
// of a point list the same neighbour points, each as three consecutive
// columns 3Q..3Q+2, lower neighbours first, then the point itself, then the
// upper ones (checked on the host when the ordering is built).  One wave per
// point, lane k takes neighbour points k and k+64: one column load, three
// vector loads and nine factor entries per neighbour, so the dependent load
// chain is col -> y once instead of once per 64 entries of each row.
template <typename T>
__global__ __launch_bounds__(64) void k_mcb_lower(int r0, int npts, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                  const int32_t *__restrict__ dpos, const T *__restrict__ f,
                                                  const double *__restrict__ b, double *__restrict__ y)
{
    const int qi = blockIdx.x, lane = threadIdx.x;
    if (qi >= npts) return;
    const int r = r0 + 3 * qi;
    const int b0 = rp[r], b1 = rp[r + 1], b2 = rp[r + 2];
    const int nl = (dpos[r] - b0) / 3;
    double l10 = 0, l20 = 0, l21 = 0;
    if (lane == 0) {
        l10 = (double)f[b1 + 3 * nl];
        l20 = (double)f[b2 + 3 * nl];
        l21 = (double)f[b2 + 3 * nl + 1];
    }
    double s0 = 0, s1 = 0, s2 = 0;
    for (int k = lane; k < nl; k += 64) {
        const int q = col[b0 + 3 * k];
        const double y0 = y[q], y1 = y[q + 1], y2 = y[q + 2];
        const int o = 3 * k;
        s0 += (double)f[b0 + o] * y0 + (double)f[b0 + o + 1] * y1 + (double)f[b0 + o + 2] * y2;
        s1 += (double)f[b1 + o] * y0 + (double)f[b1 + o + 1] * y1 + (double)f[b1 + o + 2] * y2;
        s2 += (double)f[b2 + o] * y0 + (double)f[b2 + o + 1] * y1 + (double)f[b2 + o + 2] * y2;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o, 64);
        s1 += __shfl_xor(s1, o, 64);
        s2 += __shfl_xor(s2, o, 64);
    }
    if (lane == 0) {
        const double y0 = b[r] - s0;
        const double y1 = b[r + 1] - s1 - l10 * y0;
        const double y2 = b[r + 2] - s2 - l20 * y0 - l21 * y1;
        y[r] = y0;
        y[r + 1] = y1;
        y[r + 2] = y2;
    }
}

template <typename T>
__global__ __launch_bounds__(64) void k_mcb_upper(int r0, int npts, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                  const int32_t *__restrict__ dpos, const T *__restrict__ f,
                                                  const double *__restrict__ y, double *__restrict__ z)
{
    const int qi = blockIdx.x, lane = threadIdx.x;
    if (qi >= npts) return;
    const int r = r0 + 3 * qi;
    const int b0 = rp[r], b1 = rp[r + 1], b2 = rp[r + 2], e0 = rp[r + 3];
    const int nl = (dpos[r] - b0) / 3;
    const int nu = (e0 - b2) / 3 - nl - 1;
    const int u0 = b0 + 3 * nl + 3, u1 = b1 + 3 * nl + 3, u2 = b2 + 3 * nl + 3;   // first upper entry per row
    double d00 = 1, d01 = 0, d02 = 0, d11 = 1, d12 = 0, d22 = 1;
    if (lane == 0) {
        d00 = (double)f[u0 - 3];
        d01 = (double)f[u0 - 2];
        d02 = (double)f[u0 - 1];
        d11 = (double)f[u1 - 2];
        d12 = (double)f[u1 - 1];
        d22 = (double)f[u2 - 1];
    }
    double s0 = 0, s1 = 0, s2 = 0;
    for (int k = lane; k < nu; k += 64) {
        const int q = col[u0 + 3 * k];
        const double z0 = z[q], z1 = z[q + 1], z2 = z[q + 2];
        const int o = 3 * k;
        s0 += (double)f[u0 + o] * z0 + (double)f[u0 + o + 1] * z1 + (double)f[u0 + o + 2] * z2;
        s1 += (double)f[u1 + o] * z0 + (double)f[u1 + o + 1] * z1 + (double)f[u1 + o + 2] * z2;
        s2 += (double)f[u2 + o] * z0 + (double)f[u2 + o + 1] * z1 + (double)f[u2 + o + 2] * z2;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o, 64);
        s1 += __shfl_xor(s1, o, 64);
        s2 += __shfl_xor(s2, o, 64);
    }
    if (lane == 0) {
        const double z2 = (y[r + 2] - s2) / d22;
        const double z1 = (y[r + 1] - s1 - d12 * z2) / d11;
        const double z0 = (y[r] - s0 - d01 * z1 - d02 * z2) / d00;
        z[r] = z0;
        z[r + 1] = z1;
        z[r + 2] = z2;
    }
}

// ILU(0) of the coloured-order matrix on its point-block pattern: the
// scalar ILU(0) (unit lower L, upper U with diagonal, no fill outside the
// pattern) computed a point at a time.  Colours are factored in order (one
// launch each); a point's three rows depend only on rows of earlier colours,
// already final.  One wave per point, its rows in LDS: for each lower
// neighbour Q in increasing order, the 3x3 block L_PQ = A_PQ U_QQ^-1 (the
// scalar eliminations of Q's three columns), then A_PR -= L_PQ U_QR for every
// upper neighbour R of Q that is also a neighbour of P (found through the
// grid offset of R from P, periodic directions taken modulo their size);
// finally the scalar elimination inside the point's own diagonal block, which
// also updates its upper neighbours.  The summation order differs from a
// row-by-row scalar factorisation; the factor is the same up to roundoff.
constexpr int MCB_MAXNB = 125;   // 5^3 stencil points
__device__ __forceinline__ int mcb_offcode(int a, int b, int nb0, int nb1, int nb2, int per0, int per1, int per2)
{
    int di = (b & 1023) - (a & 1023), dj = ((b >> 10) & 1023) - ((a >> 10) & 1023), dk = (b >> 20) - (a >> 20);
    if (per0) { if (di > nb0 / 2) di -= nb0; else if (di < -(nb0 / 2)) di += nb0; }
    if (per1) { if (dj > nb1 / 2) dj -= nb1; else if (dj < -(nb1 / 2)) dj += nb1; }
    if (per2) { if (dk > nb2 / 2) dk -= nb2; else if (dk < -(nb2 / 2)) dk += nb2; }
    if (di < -2 || di > 2 || dj < -2 || dj > 2 || dk < -2 || dk > 2) return -1;
    return (di + 2) * 25 + (dj + 2) * 5 + (dk + 2);
}

template <typename T>
__global__ __launch_bounds__(64) void k_mcb_ilu(int r0, int npts, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                const int32_t *__restrict__ dpos, const int32_t *__restrict__ pcoord, int nb0,
                                                int nb1, int nb2, int per0, int per1, int per2, T *__restrict__ f, int *__restrict__ zpiv)
{
    __shared__ T A[3][3 * MCB_MAXNB];
    __shared__ int nbr[MCB_MAXNB];
    __shared__ int map[MCB_MAXNB];
    __shared__ T L[3][3];
    const int qi = blockIdx.x, lane = threadIdx.x;
    if (qi >= npts) return;
    const int r = r0 + 3 * qi;
    const int b0 = rp[r], b1 = rp[r + 1], b2 = rp[r + 2];
    const int len = b1 - b0, nnb = len / 3, nl = (dpos[r] - b0) / 3;
    for (int i = lane; i < len; i += 64) {
        A[0][i] = f[b0 + i];
        A[1][i] = f[b1 + i];
        A[2][i] = f[b2 + i];
    }
    for (int i = lane; i < MCB_MAXNB; i += 64) map[i] = -1;
    __syncthreads();
    const int cp = pcoord[r / 3];
    for (int t = lane; t < nnb; t += 64) {
        const int q = col[b0 + 3 * t] / 3;
        nbr[t] = q;
        const int code = mcb_offcode(cp, pcoord[q], nb0, nb1, nb2, per0, per1, per2);
        if (code >= 0) map[code] = t;
    }
    __syncthreads();
    for (int s = 0; s < nl; s++) {
        const int Q = nbr[s];
        const int qb0 = rp[3 * Q], qb1 = rp[3 * Q + 1], qb2 = rp[3 * Q + 2];
        const int qd = dpos[3 * Q] - qb0, qnnb = (qb1 - qb0) / 3;
        if (lane < 3) {
            const T u00 = f[qb0 + qd], u01 = f[qb0 + qd + 1], u02 = f[qb0 + qd + 2];
            const T u11 = f[qb1 + qd + 1], u12 = f[qb1 + qd + 2], u22 = f[qb2 + qd + 2];
            const T a0 = A[lane][3 * s] / u00;
            const T a1 = (A[lane][3 * s + 1] - a0 * u01) / u11;
            const T a2 = (A[lane][3 * s + 2] - a0 * u02 - a1 * u12) / u22;
            A[lane][3 * s] = a0;
            A[lane][3 * s + 1] = a1;
            A[lane][3 * s + 2] = a2;
            L[lane][0] = a0;
            L[lane][1] = a1;
            L[lane][2] = a2;
        }
        __syncthreads();
        for (int t = qd / 3 + 1 + lane; t < qnnb; t += 64) {
            const int code = mcb_offcode(cp, pcoord[col[qb0 + 3 * t] / 3], nb0, nb1, nb2, per0, per1, per2);
            const int tp = code >= 0 ? map[code] : -1;
            if (tp < 0) continue;
            T u[3][3];
#pragma unroll
            for (int e = 0; e < 3; e++) {
                u[0][e] = f[qb0 + 3 * t + e];
                u[1][e] = f[qb1 + 3 * t + e];
                u[2][e] = f[qb2 + 3 * t + e];
            }
#pragma unroll
            for (int d = 0; d < 3; d++)
#pragma unroll
                for (int e = 0; e < 3; e++) A[d][3 * tp + e] -= L[d][0] * u[0][e] + L[d][1] * u[1][e] + L[d][2] * u[2][e];
        }
        __syncthreads();
    }
    const int o = 3 * nl;
    const T p0 = A[0][o];
    const T l10 = A[1][o] / p0, l20 = A[2][o] / p0;
    __syncthreads();
    for (int i = o + 1 + lane; i < len; i += 64) {
        A[1][i] -= l10 * A[0][i];
        A[2][i] -= l20 * A[0][i];
    }
    if (lane == 0) {
        A[1][o] = l10;
        A[2][o] = l20;
    }
    __syncthreads();
    const T p1 = A[1][o + 1];
    const T l21 = A[2][o + 1] / p1;
    __syncthreads();
    for (int i = o + 2 + lane; i < len; i += 64) A[2][i] -= l21 * A[1][i];
    if (lane == 0) A[2][o + 1] = l21;
    __syncthreads();
    if (lane == 0) {
        const int zr = p0 == (T)0 ? r : p1 == (T)0 ? r + 1 : A[2][o + 2] == (T)0 ? r + 2 : -1;
        if (zr >= 0) atomicMin(zpiv, zr);
    }
    for (int i = lane; i < len; i += 64) {
        f[b0 + i] = A[0][i];
        f[b1 + i] = A[1][i];
        f[b2 + i] = A[2][i];
    }
}

__global__ void k_set_i32(int *p, int v) { *p = v; }

// One direction of a tensor-product transfer between control-point grids:
// out(i0,i1,i2)[c] = sum_e w[e] in(i0,i1,i2 with the coordinate along dir
// replaced by ci[e])[c] over the entries e of row (coordinate along dir) of a
// 1-D operator in CSR form; three components per point, z fastest.
// Last pass of a transfer: mask (if given) zeroes the Dirichlet rows and
// accumulate adds into out instead of overwriting it.
__global__ void k_apply1d(int dir, int o0, int o1, int o2, int min, const int *__restrict__ rp, const int *__restrict__ ci,
                          const double *__restrict__ w, const double *__restrict__ in, double *__restrict__ out,
                          const int32_t *__restrict__ mask, int accumulate)
{
    const int64_t n = 3LL * o0 * o1 * o2;
    const int d1 = dir == 1 ? min : o1, d2 = dir == 2 ? min : o2;
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += st) {
        const int c = (int)(t % 3);
        const int64_t pt = t / 3;
        const int i2 = (int)(pt % o2), i1 = (int)((pt / o2) % o1), i0 = (int)(pt / ((int64_t)o1 * o2));
        const int row = dir == 0 ? i0 : dir == 1 ? i1 : i2;
        double acc = 0.0;
        for (int e = rp[row]; e < rp[row + 1]; e++) {
            const int j = ci[e];
            const int64_t idx = dir == 0 ? ((int64_t)j * d1 + i1) * d2 + i2
                              : dir == 1 ? ((int64_t)i0 * d1 + j) * d2 + i2
                                         : ((int64_t)i0 * d1 + i1) * d2 + j;
            acc += w[e] * in[3 * idx + c];
        }
        if (mask && mask[t]) acc = 0.0;
        if (accumulate) out[t] += acc;
        else out[t] = acc;
    }
}

// x = s D^-1 b, d = x, r = b   (first Chebyshev step from x = 0)
__global__ void k_cheb_first(int64_t ncp, const double *__restrict__ Dinv, const double *__restrict__ b, double *__restrict__ x,
                             double *__restrict__ d, double *__restrict__ r, double sc)
{
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < ncp; p += st) {
        const double *m = Dinv + 9 * p;
        const double b0 = b[3 * p], b1 = b[3 * p + 1], b2 = b[3 * p + 2];
        const double z0 = sc * (m[0] * b0 + m[1] * b1 + m[2] * b2);
        const double z1 = sc * (m[3] * b0 + m[4] * b1 + m[5] * b2);
        const double z2 = sc * (m[6] * b0 + m[7] * b1 + m[8] * b2);
        x[3 * p] = z0; x[3 * p + 1] = z1; x[3 * p + 2] = z2;
        d[3 * p] = z0; d[3 * p + 1] = z1; d[3 * p + 2] = z2;
        r[3 * p] = b0; r[3 * p + 1] = b1; r[3 * p + 2] = b2;
    }
}

// Inverses of the 3x3 diagonal blocks (one per control point) of a CSR
// matrix whose rows 3p..3p+2 hold columns 3p..3p+2 (a singular block gets
// the identity)
template <typename T>
__global__ void k_block_diag_inv(int64_t ncp, const int64_t *__restrict__ rp, const int32_t *__restrict__ col,
                                 const T *__restrict__ val, double *__restrict__ Dinv)
{
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < ncp; p += st) {
        double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int d = 0; d < 3; d++) {
            const int64_t r = 3 * p + d;
            int64_t lo = rp[r], hi = rp[r + 1] - 1;   // first column >= 3p (columns sorted)
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if (col[mid] < 3 * p) lo = mid + 1;
                else hi = mid;
            }
            for (int64_t j = lo; j < rp[r + 1] && col[j] <= 3 * p + 2; j++)
                if (col[j] >= 3 * p) a[3 * d + (col[j] - 3 * p)] = (double)val[j];
        }
        const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
        const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
        double *o = Dinv + 9 * p;
        if (det == 0.0 || !(det == det)) {
            for (int q = 0; q < 9; q++) o[q] = (q % 4 == 0) ? 1.0 : 0.0;
            continue;
        }
        const double id = 1.0 / det;
        o[0] = c00 * id;
        o[1] = (a[2] * a[7] - a[1] * a[8]) * id;
        o[2] = (a[1] * a[5] - a[2] * a[4]) * id;
        o[3] = c01 * id;
        o[4] = (a[0] * a[8] - a[2] * a[6]) * id;
        o[5] = (a[2] * a[3] - a[0] * a[5]) * id;
        o[6] = c02 * id;
        o[7] = (a[1] * a[6] - a[0] * a[7]) * id;
        o[8] = (a[0] * a[4] - a[1] * a[3]) * id;
    }
}

// d = alpha d + beta Dinv r; x += d   (one Chebyshev step, point-block Jacobi)
__global__ void k_cheb_step(int64_t ncp, const double *__restrict__ Dinv, const double *__restrict__ r, double *__restrict__ d,
                            double *__restrict__ x, double alpha, double beta, int add)
{
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < ncp; p += st) {
        const double *m = Dinv + 9 * p;
        const double r0 = r[3 * p], r1 = r[3 * p + 1], r2 = r[3 * p + 2];
        const double z0 = m[0] * r0 + m[1] * r1 + m[2] * r2;
        const double z1 = m[3] * r0 + m[4] * r1 + m[5] * r2;
        const double z2 = m[6] * r0 + m[7] * r1 + m[8] * r2;
        double d0 = beta * z0, d1 = beta * z1, d2 = beta * z2;
        if (alpha != 0.0) {   // (d is not read on the first step)
            d0 += alpha * d[3 * p];
            d1 += alpha * d[3 * p + 1];
            d2 += alpha * d[3 * p + 2];
        }
        d[3 * p] = d0;
        d[3 * p + 1] = d1;
        d[3 * p + 2] = d2;
        if (add) {
            x[3 * p] += d0;
            x[3 * p + 1] += d1;
            x[3 * p + 2] += d2;
        }
    }
}

// v[i] in [-1, 1): a fixed integer hash of i (start vector of the power iteration)
__global__ void k_fill_hash(int64_t n, double *__restrict__ v)
{
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st) {
        uint64_t h = (uint64_t)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
        h ^= h >> 31;
        h *= 0xBF58476D1CE4E5B9ull;
        h ^= h >> 29;
        v[i] = (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
}

// dense column-major copy of a CSR matrix (n <= a few thousand)
template <typename T>
__global__ void k_csr_to_dense(int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ col, const T *__restrict__ val,
                               double *__restrict__ D)
{
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += st)
        for (int64_t j = rp[r]; j < rp[r + 1]; j++) D[(int64_t)col[j] * n + r] = (double)val[j];
}

// Partial dot products of the Krylov vectors with w: part[k][ch] =
// sum over rows of chunk ch of V[k ld + i] w[i]; one workgroup per (chunk,
// vector), fixed summation order (deterministic).
constexpr int DOT_CHUNK = 4096;
__global__ __launch_bounds__(256) void k_dot_part(int64_t n, int64_t ld, const double *__restrict__ V, const double *__restrict__ w,
                                                  double *__restrict__ part)
{
    __shared__ double red[4];
    const int ch = blockIdx.x, k = blockIdx.y, tid = threadIdx.x;
    const int64_t i0 = (int64_t)ch * DOT_CHUNK;
    const int64_t i1 = i0 + DOT_CHUNK < n ? i0 + DOT_CHUNK : n;
    const double *v = V + (int64_t)k * ld;
    double s = 0;
    for (int64_t i = i0 + tid; i < i1; i += 256) s += v[i] * w[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) part[(int64_t)k * gridDim.x + ch] = (red[0] + red[1]) + (red[2] + red[3]);
}

// h[k] = sum_ch part[k][ch] (one wave per vector; fixed order)
__global__ __launch_bounds__(64) void k_dot_final(int nch, const double *__restrict__ part, double *__restrict__ h)
{
    const int k = blockIdx.x, lane = threadIdx.x;
    double s = 0;
    for (int c = lane; c < nch; c += 64) s += part[(int64_t)k * nch + c];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) h[k] = s;
}

// x *= 1/sqrt(*nrm2) when the squared norm is positive
__global__ void k_scale_rsqrt(int64_t n, const double *__restrict__ nrm2, double *__restrict__ x)
{
    const double q = *nrm2;
    if (!(q > 0.0)) return;
    const double a = 1.0 / sqrt(q);
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st) x[i] *= a;
}

template <typename T>
__global__ void k_diag_inv(int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ col, const T *__restrict__ val,
                           double *__restrict__ dinv)
{
    const int64_t st = (int64_t)gridDim.x * blockDim.x;   // grid-stride: launched with grid1
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += st) {
        // the diagonal by bisection (columns sorted; 1 where the row has none)
        int64_t lo = rp[r], hi = rp[r + 1] - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (col[mid] < r) lo = mid + 1;
            else hi = mid;
        }
        const double dv = (lo < rp[r + 1] && col[lo] == r) ? (double)val[lo] : 1.0;
        dinv[r] = dv != 0.0 ? 1.0 / dv : 1.0;
    }
}

__global__ void k_axpby(int64_t n, double a, const double *__restrict__ x, double b, double *__restrict__ y)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) y[i] = a * x[i] + b * y[i];
}

__global__ void k_waxpy(int64_t n, double a, const double *__restrict__ x, const double *__restrict__ y, double *__restrict__ w)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) w[i] = y[i] + a * x[i];
}

__global__ void k_mul(int64_t n, const double *__restrict__ a, const double *__restrict__ x, double *__restrict__ y)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) y[i] = a[i] * x[i];
}

__global__ void k_d2f(int64_t n, const double *__restrict__ x, float *__restrict__ y)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) y[i] = (float)x[i];
}

__global__ void k_f2d(int64_t n, const float *__restrict__ x, double *__restrict__ y)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) y[i] = (double)x[i];
}

__global__ void k_i64_to_i32(int64_t n, const int64_t *__restrict__ a, int32_t *__restrict__ b)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (int32_t)a[i];
}

static inline dim3 grid1(int64_t n)
{
    int64_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

#define SPCHK(x)                                                                     \
    do {                                                                             \
        rocsparse_status s_ = (x);                                                   \
        if (s_ != rocsparse_status_success) return fail(IGAP_ERR_HIP, "%s: rocsparse status %d", #x, (int)s_); \
    } while (0)
// device allocation through the library's allocator; failure is IGAP_ERR_NOMEM
#define DALLOC(ptr, bytes)                                                                            \
    do {                                                                                              \
        hipError_t e_ = igap::dmalloc((ptr), (bytes));                                                \
        if (e_ != hipSuccess)                                                                         \
            return fail(IGAP_ERR_NOMEM, "device allocation of %zu bytes: %s", (size_t)(bytes),       \
                        hipGetErrorString(e_));                                                       \
    } while (0)
#define BLCHK(x)                                                                     \
    do {                                                                             \
        rocblas_status s_ = (x);                                                     \
        if (s_ != rocblas_status_success) return fail(IGAP_ERR_HIP, "%s: rocblas status %d", #x, (int)s_); \
    } while (0)

}  // namespace igap

using namespace igap;

// ---------------------------------------------------------------- set-up
// a new solver's buffers and library handles (the caller frees s on failure)
static int solver_init(igap_gpu_ctx *c, Solver *s, int m)
{
    s->m = m;
    s->n = c->ndof;
    const int64_t n = c->ndof;
    if (c->nnz > 0x7fffffffLL || n > 0x7fffffffLL)
        return fail(IGAP_ERR_ARG, "the solver's 32-bit sparse indices need nnz < 2^31 (nnz = %lld)", (long long)c->nnz);
    DALLOC(&s->V, sizeof(double) * (size_t)(m + 1) * n);
    double **vecs[] = {&s->w, &s->z, &s->t, &s->t2, &s->d, &s->x, &s->xt, &s->F, &s->Ft, &s->diag};
    for (double **v : vecs) DALLOC(v, sizeof(double) * n);
    DALLOC(&s->h, sizeof(double) * 3 * (m + 2));
    DALLOC(&s->part, sizeof(double) * (size_t)(m + 2) * (size_t)((n + DOT_CHUNK - 1) / DOT_CHUNK));
    DALLOC(&s->fa, sizeof(float) * n);
    DALLOC(&s->fb, sizeof(float) * n);
    DALLOC(&s->rowptr32, sizeof(int32_t) * (n + 1));
    hipLaunchKernelGGL(k_i64_to_i32, dim3((unsigned)((n + 1 + 255) / 256)), dim3(256), 0, c->stream, n + 1, c->d_rowptr,
                       s->rowptr32);
    HIPCHK(hipGetLastError());
    SPCHK(rocsparse_create_handle(&s->sp));
    SPCHK(rocsparse_set_stream(s->sp, c->stream));
    SPCHK(rocsparse_create_mat_descr(&s->dA));
    SPCHK(rocsparse_create_mat_descr(&s->dL));
    SPCHK(rocsparse_set_mat_fill_mode(s->dL, rocsparse_fill_mode_lower));
    SPCHK(rocsparse_set_mat_diag_type(s->dL, rocsparse_diag_type_unit));
    SPCHK(rocsparse_create_mat_descr(&s->dU));
    SPCHK(rocsparse_set_mat_fill_mode(s->dU, rocsparse_fill_mode_upper));
    SPCHK(rocsparse_set_mat_diag_type(s->dU, rocsparse_diag_type_non_unit));
    BLCHK(rocblas_create_handle(&s->bl));
    BLCHK(rocblas_set_stream(s->bl, c->stream));
    BLCHK(rocblas_set_pointer_mode(s->bl, rocblas_pointer_mode_host));
    return IGAP_OK;
}

// the context's solver with GMRES restart m, built on first use (or when m
// changes); published only once complete, so a failed call leaves none
static int solver_get(igap_gpu_ctx *c, int m, Solver **out)
{
    Solver *s = c->solver;
    if (s && s->m != m) {
        solver_destroy(s);
        c->solver = s = nullptr;
    }
    if (!s) {
        s = new Solver();
        const int rc = solver_init(c, s, m);
        if (rc) {
            solver_destroy(s);
            return rc;
        }
        c->solver = s;
    }
    *out = s;
    return IGAP_OK;
}

// the coloured ordering and the permuted CSR pattern (host, once per context)
static int mc_build_body(igap_gpu_ctx *c, Solver *s);

// the coloured ordering, built once; a failed build frees what it allocated
// so the next call starts over
static int mc_build(igap_gpu_ctx *c, Solver *s)
{
    if (s->mc_ready) return IGAP_OK;
    const int rc = mc_build_body(c, s);
    if (rc) {
        int32_t **bufs[] = {&s->mc_pcoord, &s->mc_dpos, &s->mc_perm, &s->mc_rowptr, &s->mc_col, &s->mc_src};
        for (int32_t **b : bufs) {
            if (*b) (void)hipFree(*b);
            *b = nullptr;
        }
        return rc;
    }
    s->mc_ready = 1;
    return IGAP_OK;
}

static int mc_build_body(igap_gpu_ctx *c, Solver *s)
{
    const int64_t n = c->ndof, nnz = c->nnz, ncp = n / 3;
    const int nb1 = c->nb[1], nb2 = c->nb[2];
    std::vector<int64_t> rp(n + 1);
    std::vector<int32_t> col(nnz);
    HIPCHK(hipMemcpyAsync(rp.data(), c->d_rowptr, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(col.data(), c->d_col, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    // control points by (colour, index); dofs 3 cp + i keep their point's place
    std::vector<int32_t> cps(ncp);
    std::iota(cps.begin(), cps.end(), 0);
    auto colour = [&](int64_t a) {
        const int64_t i = a / ((int64_t)nb1 * nb2), j = (a / nb2) % nb1, k = a % nb2;
        return (int)(9 * (i % 3) + 3 * (j % 3) + (k % 3));
    };
    std::stable_sort(cps.begin(), cps.end(), [&](int32_t a, int32_t b) { return colour(a) < colour(b); });
    std::vector<int32_t> perm(n), iperm(n);
    for (int64_t q = 0; q < ncp; q++)
        for (int d = 0; d < 3; d++) {
            perm[3 * q + d] = 3 * cps[q] + d;
            iperm[3 * cps[q] + d] = (int32_t)(3 * q + d);
        }
    std::vector<int32_t> prp(n + 1), pcol(nnz), psrc(nnz);
    std::vector<std::pair<int32_t, int32_t>> row;
    prp[0] = 0;
    for (int64_t r = 0; r < n; r++) {
        const int64_t o = perm[r];
        row.clear();
        for (int64_t j = rp[o]; j < rp[o + 1]; j++) row.emplace_back(iperm[col[j]], (int32_t)j);
        std::sort(row.begin(), row.end());
        const int64_t b = prp[r];
        for (size_t t = 0; t < row.size(); t++) {
            pcol[b + t] = row[t].first;
            psrc[b + t] = row[t].second;
        }
        prp[r + 1] = (int32_t)(b + (int64_t)row.size());
    }
    std::vector<int32_t> pdpos(n);
    for (int64_t r = 0; r < n; r++)
        for (int32_t j = prp[r]; j < prp[r + 1]; j++)
            if (pcol[j] == r) pdpos[r] = j;
    for (int cc = 0; cc <= 27; cc++) s->mc_cstart[cc] = 0;
    {
        int64_t q = 0;
        for (int cc = 0; cc < 27; cc++) {
            s->mc_cstart[cc] = (int)(3 * q);
            while (q < ncp && colour(cps[q]) == cc) q++;
        }
        s->mc_cstart[27] = (int)(3 * q);
    }
    s->mc_sweep_ok = 1;
    for (int64_t r = 0; r < n && s->mc_sweep_ok; r++) {
        const int64_t q0 = 3 * (r / 3);
        for (int32_t j = prp[r]; j < prp[r + 1]; j++) {
            const int32_t cj = pcol[j];
            if ((cj < q0 || cj > q0 + 2) && colour(perm[cj] / 3) == colour(perm[r] / 3)) { s->mc_sweep_ok = 0; break; }
        }
    }
    // point-block structure: the three rows of a point have equal length and
    // list each neighbour point as the columns 3Q, 3Q+1, 3Q+2, the point's own
    // block after its lower neighbours
    s->mc_block_ok = s->mc_sweep_ok;
    for (int64_t r = 0; r < n && s->mc_block_ok; r += 3) {
        const int32_t b0 = prp[r], len = prp[r + 1] - b0;
        if (len % 3 || (pdpos[r] - b0) % 3) { s->mc_block_ok = 0; break; }
        for (int d = 0; d < 3 && s->mc_block_ok; d++) {
            const int32_t bd = prp[r + d];
            if (prp[r + d + 1] - bd != len || pdpos[r + d] - bd != pdpos[r] - b0 + d) { s->mc_block_ok = 0; break; }
            for (int32_t k = 0; k < len; k += 3) {
                const int32_t q = pcol[b0 + k];
                if (q % 3 || pcol[bd + k] != q || pcol[bd + k + 1] != q + 1 || pcol[bd + k + 2] != q + 2) { s->mc_block_ok = 0; break; }
            }
        }
    }
    for (int64_t r = 0; r < n && s->mc_block_ok; r += 3)
        if (prp[r + 1] - prp[r] > 3 * MCB_MAXNB) s->mc_block_ok = 0;
    if (s->mc_block_ok) {
        if (c->nb[0] > 1023 || c->nb[1] > 1023 || c->nb[2] > 1023) s->mc_block_ok = 0;
    }
    {
        std::vector<int32_t> pc(ncp);
        for (int64_t q = 0; q < ncp; q++) {
            const int64_t a = cps[q];
            const int64_t i = a / ((int64_t)nb1 * nb2), j = (a / nb2) % nb1, k = a % nb2;
            pc[q] = (int32_t)(i | (j << 10) | (k << 20));
        }
        DALLOC(&s->mc_pcoord, sizeof(int32_t) * ncp);
        HIPCHK(hipMemcpyAsync(s->mc_pcoord, pc.data(), sizeof(int32_t) * ncp, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    DALLOC(&s->mc_dpos, sizeof(int32_t) * n);
    HIPCHK(hipMemcpyAsync(s->mc_dpos, pdpos.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
    DALLOC(&s->mc_perm, sizeof(int32_t) * n);
    DALLOC(&s->mc_rowptr, sizeof(int32_t) * (n + 1));
    DALLOC(&s->mc_col, sizeof(int32_t) * nnz);
    DALLOC(&s->mc_src, sizeof(int32_t) * nnz);
    HIPCHK(hipMemcpyAsync(s->mc_perm, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(s->mc_rowptr, prp.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(s->mc_col, pcol.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(s->mc_src, psrc.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return IGAP_OK;
}

// ILU(0) of the tangent in the coloured order, precision p
// own factorisation (k_mcb_ilu), colour by colour, of the values in mc_ilu[p]
static int mc_factor_own(igap_gpu_ctx *c, Solver *s, int p)
{
    if (!s->zpiv) DALLOC(&s->zpiv, sizeof(int));
    hipLaunchKernelGGL(k_set_i32, dim3(1), dim3(1), 0, c->stream, s->zpiv, 0x7fffffff);
    for (int cc = 0; cc < 27; cc++) {
        const int r0 = s->mc_cstart[cc], np = (s->mc_cstart[cc + 1] - r0) / 3;
        if (np == 0) continue;
        if (p == IGAP_F64)
            hipLaunchKernelGGL(k_mcb_ilu<double>, dim3((unsigned)np), dim3(64), 0, c->stream, r0, np, s->mc_rowptr, s->mc_col, s->mc_dpos,
                               s->mc_pcoord, c->nb[0], c->nb[1], c->nb[2], c->periodic[0], c->periodic[1], c->periodic[2],
                               (double *)s->mc_ilu[p], s->zpiv);
        else
            hipLaunchKernelGGL(k_mcb_ilu<float>, dim3((unsigned)np), dim3(64), 0, c->stream, r0, np, s->mc_rowptr, s->mc_col, s->mc_dpos,
                               s->mc_pcoord, c->nb[0], c->nb[1], c->nb[2], c->periodic[0], c->periodic[1], c->periodic[2],
                               (float *)s->mc_ilu[p], s->zpiv);
    }
    HIPCHK(hipGetLastError());
    int zp = 0;
    HIPCHK(hipMemcpyAsync(&zp, s->zpiv, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (zp != 0x7fffffff) return fail(IGAP_ERR_STATE, "ILU(0), coloured order: zero pivot in row %d", zp);
    return IGAP_OK;
}

// ILU(0) of the tangent in the coloured order, precision p; impl 0: own
// point-block factorisation where the pattern allows it, 1: rocSPARSE
static int pc_ilu_setup_mc(igap_gpu_ctx *c, Solver *s, int p, int impl = 0)
{
    int rc = mc_build(c, s);
    if (rc) return rc;
    const int n = (int)c->ndof, nnz = (int)c->nnz;
    if (!s->mc_ilu[p]) DALLOC(&s->mc_ilu[p], (p == IGAP_F64 ? sizeof(double) : sizeof(float)) * (size_t)nnz);
    if (p == IGAP_F64)
        hipLaunchKernelGGL((k_gather<double, double>), grid1(nnz), dim3(256), 0, c->stream, (int64_t)nnz, s->mc_src, c->d_val64,
                           (double *)s->mc_ilu[p]);
    else
        hipLaunchKernelGGL((k_gather<float, float>), grid1(nnz), dim3(256), 0, c->stream, (int64_t)nnz, s->mc_src, c->d_val32,
                           (float *)s->mc_ilu[p]);
    HIPCHK(hipGetLastError());
    if (impl == 0 && s->mc_block_ok) return mc_factor_own(c, s, p);
    if (!s->mc_analysed[p]) {
        // leftovers of an analysis that failed part way
        if (s->mc_info[p]) { rocsparse_destroy_mat_info(s->mc_info[p]); s->mc_info[p] = nullptr; }
        if (s->mc_buf[p]) { (void)hipFree(s->mc_buf[p]); s->mc_buf[p] = nullptr; }
        SPCHK(rocsparse_create_mat_info(&s->mc_info[p]));
        size_t b0 = 0, b1 = 0, b2 = 0;
        if (p == IGAP_F64) {
            double *v = (double *)s->mc_ilu[p];
            SPCHK(rocsparse_dcsrilu0_buffer_size(s->sp, n, nnz, s->dA, v, s->mc_rowptr, s->mc_col, s->mc_info[p], &b0));
            SPCHK(rocsparse_dcsrsv_buffer_size(s->sp, rocsparse_operation_none, n, nnz, s->dL, v, s->mc_rowptr, s->mc_col, s->mc_info[p], &b1));
            SPCHK(rocsparse_dcsrsv_buffer_size(s->sp, rocsparse_operation_none, n, nnz, s->dU, v, s->mc_rowptr, s->mc_col, s->mc_info[p], &b2));
        } else {
            float *v = (float *)s->mc_ilu[p];
            SPCHK(rocsparse_scsrilu0_buffer_size(s->sp, n, nnz, s->dA, v, s->mc_rowptr, s->mc_col, s->mc_info[p], &b0));
            SPCHK(rocsparse_scsrsv_buffer_size(s->sp, rocsparse_operation_none, n, nnz, s->dL, v, s->mc_rowptr, s->mc_col, s->mc_info[p], &b1));
            SPCHK(rocsparse_scsrsv_buffer_size(s->sp, rocsparse_operation_none, n, nnz, s->dU, v, s->mc_rowptr, s->mc_col, s->mc_info[p], &b2));
        }
        size_t bs = std::max(b0, std::max(b1, b2));
        DALLOC(&s->mc_buf[p], bs > 0 ? bs : 1);
        if (p == IGAP_F64) {
            double *v = (double *)s->mc_ilu[p];
            SPCHK(rocsparse_dcsrilu0_analysis(s->sp, n, nnz, s->dA, v, s->mc_rowptr, s->mc_col, s->mc_info[p],
                                              rocsparse_analysis_policy_reuse, rocsparse_solve_policy_auto, s->mc_buf[p]));
            SPCHK(rocsparse_dcsrsv_analysis(s->sp, rocsparse_operation_none, n, nnz, s->dL, v, s->mc_rowptr, s->mc_col, s->mc_info[p],
                                            rocsparse_analysis_policy_reuse, rocsparse_solve_policy_auto, s->mc_buf[p]));
            SPCHK(rocsparse_dcsrsv_analysis(s->sp, rocsparse_operation_none, n, nnz, s->dU, v, s->mc_rowptr, s->mc_col, s->mc_info[p],
                                            rocsparse_analysis_policy_reuse, rocsparse_solve_policy_auto, s->mc_buf[p]));
        } else {
            float *v = (float *)s->mc_ilu[p];
            SPCHK(rocsparse_scsrilu0_analysis(s->sp, n, nnz, s->dA, v, s->mc_rowptr, s->mc_col, s->mc_info[p],
                                              rocsparse_analysis_policy_reuse, rocsparse_solve_policy_auto, s->mc_buf[p]));
            SPCHK(rocsparse_scsrsv_analysis(s->sp, rocsparse_operation_none, n, nnz, s->dL, v, s->mc_rowptr, s->mc_col, s->mc_info[p],
                                            rocsparse_analysis_policy_reuse, rocsparse_solve_policy_auto, s->mc_buf[p]));
            SPCHK(rocsparse_scsrsv_analysis(s->sp, rocsparse_operation_none, n, nnz, s->dU, v, s->mc_rowptr, s->mc_col, s->mc_info[p],
                                            rocsparse_analysis_policy_reuse, rocsparse_solve_policy_auto, s->mc_buf[p]));
        }
        s->mc_analysed[p] = 1;
    }
    if (p == IGAP_F64)
        SPCHK(rocsparse_dcsrilu0(s->sp, n, nnz, s->dA, (double *)s->mc_ilu[p], s->mc_rowptr, s->mc_col, s->mc_info[p],
                                 rocsparse_solve_policy_auto, s->mc_buf[p]));
    else
        SPCHK(rocsparse_scsrilu0(s->sp, n, nnz, s->dA, (float *)s->mc_ilu[p], s->mc_rowptr, s->mc_col, s->mc_info[p],
                                 rocsparse_solve_policy_auto, s->mc_buf[p]));
    rocsparse_int zp = -1;
    rocsparse_status st = rocsparse_csrilu0_zero_pivot(s->sp, s->mc_info[p], &zp);
    if (st == rocsparse_status_zero_pivot) return fail(IGAP_ERR_STATE, "ILU(0), coloured order: zero pivot in row %d", (int)zp);
    return IGAP_OK;
}

// ILU(0) numeric factorisation of the current tangent values in precision p
static int pc_ilu_setup(igap_gpu_ctx *c, Solver *s, int p)
{
    int n = (int)c->ndof, nnz = (int)c->nnz;
    if (!s->ilu[p]) {
        size_t es = p == IGAP_F64 ? sizeof(double) : sizeof(float);
        DALLOC(&s->ilu[p], es * (size_t)nnz);
    }
    if (p == IGAP_F64) HIPCHK(hipMemcpyAsync(s->ilu[p], c->d_val64, sizeof(double) * nnz, hipMemcpyDeviceToDevice, c->stream));
    else HIPCHK(hipMemcpyAsync(s->ilu[p], c->d_val32, sizeof(float) * nnz, hipMemcpyDeviceToDevice, c->stream));
    if (!s->analysed[p]) {
        // leftovers of an analysis that failed part way
        if (s->info[p]) { rocsparse_destroy_mat_info(s->info[p]); s->info[p] = nullptr; }
        if (s->buf[p]) { (void)hipFree(s->buf[p]); s->buf[p] = nullptr; }
        SPCHK(rocsparse_create_mat_info(&s->info[p]));
        size_t b0 = 0, b1 = 0, b2 = 0;
        if (p == IGAP_F64) {
            double *v = (double *)s->ilu[p];
            SPCHK(rocsparse_dcsrilu0_buffer_size(s->sp, n, nnz, s->dA, v, s->rowptr32, c->d_col, s->info[p], &b0));
            SPCHK(rocsparse_dcsrsv_buffer_size(s->sp, rocsparse_operation_none, n, nnz, s->dL, v, s->rowptr32, c->d_col, s->info[p], &b1));
            SPCHK(rocsparse_dcsrsv_buffer_size(s->sp, rocsparse_operation_none, n, nnz, s->dU, v, s->rowptr32, c->d_col, s->info[p], &b2));
        } else {
            float *v = (float *)s->ilu[p];
            SPCHK(rocsparse_scsrilu0_buffer_size(s->sp, n, nnz, s->dA, v, s->rowptr32, c->d_col, s->info[p], &b0));
            SPCHK(rocsparse_scsrsv_buffer_size(s->sp, rocsparse_operation_none, n, nnz, s->dL, v, s->rowptr32, c->d_col, s->info[p], &b1));
            SPCHK(rocsparse_scsrsv_buffer_size(s->sp, rocsparse_operation_none, n, nnz, s->dU, v, s->rowptr32, c->d_col, s->info[p], &b2));
        }
        size_t bs = b0 > b1 ? b0 : b1;
        bs = bs > b2 ? bs : b2;
        DALLOC(&s->buf[p], bs > 0 ? bs : 1);
        if (p == IGAP_F64) {
            double *v = (double *)s->ilu[p];
            SPCHK(rocsparse_dcsrilu0_analysis(s->sp, n, nnz, s->dA, v, s->rowptr32, c->d_col, s->info[p], rocsparse_analysis_policy_reuse,
                                              rocsparse_solve_policy_auto, s->buf[p]));
            SPCHK(rocsparse_dcsrsv_analysis(s->sp, rocsparse_operation_none, n, nnz, s->dL, v, s->rowptr32, c->d_col, s->info[p],
                                            rocsparse_analysis_policy_reuse, rocsparse_solve_policy_auto, s->buf[p]));
            SPCHK(rocsparse_dcsrsv_analysis(s->sp, rocsparse_operation_none, n, nnz, s->dU, v, s->rowptr32, c->d_col, s->info[p],
                                            rocsparse_analysis_policy_reuse, rocsparse_solve_policy_auto, s->buf[p]));
        } else {
            float *v = (float *)s->ilu[p];
            SPCHK(rocsparse_scsrilu0_analysis(s->sp, n, nnz, s->dA, v, s->rowptr32, c->d_col, s->info[p], rocsparse_analysis_policy_reuse,
                                              rocsparse_solve_policy_auto, s->buf[p]));
            SPCHK(rocsparse_scsrsv_analysis(s->sp, rocsparse_operation_none, n, nnz, s->dL, v, s->rowptr32, c->d_col, s->info[p],
                                            rocsparse_analysis_policy_reuse, rocsparse_solve_policy_auto, s->buf[p]));
            SPCHK(rocsparse_scsrsv_analysis(s->sp, rocsparse_operation_none, n, nnz, s->dU, v, s->rowptr32, c->d_col, s->info[p],
                                            rocsparse_analysis_policy_reuse, rocsparse_solve_policy_auto, s->buf[p]));
        }
        s->analysed[p] = 1;
    }
    if (p == IGAP_F64)
        SPCHK(rocsparse_dcsrilu0(s->sp, n, nnz, s->dA, (double *)s->ilu[p], s->rowptr32, c->d_col, s->info[p], rocsparse_solve_policy_auto,
                                 s->buf[p]));
    else
        SPCHK(rocsparse_scsrilu0(s->sp, n, nnz, s->dA, (float *)s->ilu[p], s->rowptr32, c->d_col, s->info[p], rocsparse_solve_policy_auto,
                                 s->buf[p]));
    rocsparse_int zp = -1;
    rocsparse_status st = rocsparse_csrilu0_zero_pivot(s->sp, s->info[p], &zp);
    if (st == rocsparse_status_zero_pivot) return fail(IGAP_ERR_STATE, "ILU(0): zero pivot in row %d", (int)zp);
    return IGAP_OK;
}

// y = M^{-1} x (x and y must differ from s->t and s->t2)
static int mg_cycle(igap_gpu_ctx *c, Solver *s, int p, int order, int l, const double *b, double *x);

static int pc_apply(igap_gpu_ctx *c, Solver *s, int pc, int p, int order, const double *x, double *y)
{
    int64_t n = c->ndof;
    if (pc == IGAP_PC_MG) return mg_cycle(c, s, p, order, 0, x, y);
    if (pc == IGAP_PC_ILU0 && order == 1) {
        hipLaunchKernelGGL((k_gather<double, double>), grid1(n), dim3(256), 0, c->stream, n, s->mc_perm, x, s->t);
        if (!s->mc_sweep_ok) {   // generic level-scheduled solves of the same factor
            int nn = (int)n, nnz = (int)c->nnz;
            if (p == IGAP_F64) {
                const double one = 1.0;
                SPCHK(rocsparse_dcsrsv_solve(s->sp, rocsparse_operation_none, nn, nnz, &one, s->dL, (double *)s->mc_ilu[p],
                                             s->mc_rowptr, s->mc_col, s->mc_info[p], s->t, s->t2, rocsparse_solve_policy_auto,
                                             s->mc_buf[p]));
                SPCHK(rocsparse_dcsrsv_solve(s->sp, rocsparse_operation_none, nn, nnz, &one, s->dU, (double *)s->mc_ilu[p],
                                             s->mc_rowptr, s->mc_col, s->mc_info[p], s->t2, s->t, rocsparse_solve_policy_auto,
                                             s->mc_buf[p]));
            } else {
                const float one = 1.0f;
                hipLaunchKernelGGL(k_d2f, grid1(n), dim3(256), 0, c->stream, n, s->t, s->fa);
                SPCHK(rocsparse_scsrsv_solve(s->sp, rocsparse_operation_none, nn, nnz, &one, s->dL, (float *)s->mc_ilu[p],
                                             s->mc_rowptr, s->mc_col, s->mc_info[p], s->fa, s->fb, rocsparse_solve_policy_auto,
                                             s->mc_buf[p]));
                SPCHK(rocsparse_scsrsv_solve(s->sp, rocsparse_operation_none, nn, nnz, &one, s->dU, (float *)s->mc_ilu[p],
                                             s->mc_rowptr, s->mc_col, s->mc_info[p], s->fb, s->fa, rocsparse_solve_policy_auto,
                                             s->mc_buf[p]));
                hipLaunchKernelGGL(k_f2d, grid1(n), dim3(256), 0, c->stream, n, s->fa, s->t);
            }
            hipLaunchKernelGGL(k_scatter, grid1(n), dim3(256), 0, c->stream, n, s->mc_perm, s->t, y);
            HIPCHK(hipGetLastError());
            return IGAP_OK;
        }
        if (s->mc_block_ok) {
            for (int cc = 0; cc < 27; cc++) {
                const int r0 = s->mc_cstart[cc], np = (s->mc_cstart[cc + 1] - r0) / 3;
                if (np == 0) continue;
                if (p == IGAP_F64)
                    hipLaunchKernelGGL(k_mcb_lower<double>, dim3((unsigned)np), dim3(64), 0, c->stream, r0, np, s->mc_rowptr, s->mc_col,
                                       s->mc_dpos, (const double *)s->mc_ilu[p], s->t, s->t2);
                else
                    hipLaunchKernelGGL(k_mcb_lower<float>, dim3((unsigned)np), dim3(64), 0, c->stream, r0, np, s->mc_rowptr, s->mc_col,
                                       s->mc_dpos, (const float *)s->mc_ilu[p], s->t, s->t2);
            }
            for (int cc = 26; cc >= 0; cc--) {
                const int r0 = s->mc_cstart[cc], np = (s->mc_cstart[cc + 1] - r0) / 3;
                if (np == 0) continue;
                if (p == IGAP_F64)
                    hipLaunchKernelGGL(k_mcb_upper<double>, dim3((unsigned)np), dim3(64), 0, c->stream, r0, np, s->mc_rowptr, s->mc_col,
                                       s->mc_dpos, (const double *)s->mc_ilu[p], s->t2, s->t);
                else
                    hipLaunchKernelGGL(k_mcb_upper<float>, dim3((unsigned)np), dim3(64), 0, c->stream, r0, np, s->mc_rowptr, s->mc_col,
                                       s->mc_dpos, (const float *)s->mc_ilu[p], s->t2, s->t);
            }
            hipLaunchKernelGGL(k_scatter, grid1(n), dim3(256), 0, c->stream, n, s->mc_perm, s->t, y);
            HIPCHK(hipGetLastError());
            return IGAP_OK;
        }
        for (int cc = 0; cc < 27; cc++) {
            const int r0 = s->mc_cstart[cc], np = (s->mc_cstart[cc + 1] - r0) / 3;
            if (np == 0) continue;
            const dim3 g((unsigned)((np + 3) / 4));
            if (p == IGAP_F64)
                hipLaunchKernelGGL(k_mc_lower<double>, g, dim3(256), 0, c->stream, r0, np, s->mc_rowptr, s->mc_col, s->mc_dpos,
                                   (const double *)s->mc_ilu[p], s->t, s->t2);
            else
                hipLaunchKernelGGL(k_mc_lower<float>, g, dim3(256), 0, c->stream, r0, np, s->mc_rowptr, s->mc_col, s->mc_dpos,
                                   (const float *)s->mc_ilu[p], s->t, s->t2);
        }
        for (int cc = 26; cc >= 0; cc--) {
            const int r0 = s->mc_cstart[cc], np = (s->mc_cstart[cc + 1] - r0) / 3;
            if (np == 0) continue;
            const dim3 g((unsigned)((np + 3) / 4));
            if (p == IGAP_F64)
                hipLaunchKernelGGL(k_mc_upper<double>, g, dim3(256), 0, c->stream, r0, np, s->mc_rowptr, s->mc_col, s->mc_dpos,
                                   (const double *)s->mc_ilu[p], s->t2, s->t);
            else
                hipLaunchKernelGGL(k_mc_upper<float>, g, dim3(256), 0, c->stream, r0, np, s->mc_rowptr, s->mc_col, s->mc_dpos,
                                   (const float *)s->mc_ilu[p], s->t2, s->t);
        }
        hipLaunchKernelGGL(k_scatter, grid1(n), dim3(256), 0, c->stream, n, s->mc_perm, s->t, y);
        HIPCHK(hipGetLastError());
        return IGAP_OK;
    }
    if (pc == IGAP_PC_JACOBI) {
        hipLaunchKernelGGL(k_mul, grid1(n), dim3(256), 0, c->stream, n, s->diag, x, y);
        HIPCHK(hipGetLastError());
        return IGAP_OK;
    }
    if (pc == IGAP_PC_NONE) {
        HIPCHK(hipMemcpyAsync(y, x, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
        return IGAP_OK;
    }
    int nn = (int)n, nnz = (int)c->nnz;
    if (p == IGAP_F64) {
        const double one = 1.0;
        SPCHK(rocsparse_dcsrsv_solve(s->sp, rocsparse_operation_none, nn, nnz, &one, s->dL, (double *)s->ilu[p], s->rowptr32, c->d_col,
                                     s->info[p], x, s->t, rocsparse_solve_policy_auto, s->buf[p]));
        SPCHK(rocsparse_dcsrsv_solve(s->sp, rocsparse_operation_none, nn, nnz, &one, s->dU, (double *)s->ilu[p], s->rowptr32, c->d_col,
                                     s->info[p], s->t, y, rocsparse_solve_policy_auto, s->buf[p]));
    } else {
        const float one = 1.0f;
        hipLaunchKernelGGL(k_d2f, grid1(n), dim3(256), 0, c->stream, n, x, s->fa);
        SPCHK(rocsparse_scsrsv_solve(s->sp, rocsparse_operation_none, nn, nnz, &one, s->dL, (float *)s->ilu[p], s->rowptr32, c->d_col,
                                     s->info[p], s->fa, s->fb, rocsparse_solve_policy_auto, s->buf[p]));
        SPCHK(rocsparse_scsrsv_solve(s->sp, rocsparse_operation_none, nn, nnz, &one, s->dU, (float *)s->ilu[p], s->rowptr32, c->d_col,
                                     s->info[p], s->fb, s->fa, rocsparse_solve_policy_auto, s->buf[p]));
        hipLaunchKernelGGL(k_f2d, grid1(n), dim3(256), 0, c->stream, n, s->fa, y);
        HIPCHK(hipGetLastError());
    }
    return IGAP_OK;
}

// y = A x (mode 0), y = b - A x (mode 1), y -= A x (mode 2)
static int ensure_pcol(igap_gpu_ctx *c)
{
    const int64_t ncp = c->ndof / 3;
    if (!c->d_pcol) {
        if (c->nnz % 9) return fail(IGAP_ERR_STATE, "pattern is not point-blocked (nnz %lld)", (long long)c->nnz);
        int32_t *pcol = nullptr;
        DALLOC(&pcol, sizeof(int32_t) * (size_t)(c->nnz / 9));
        hipLaunchKernelGGL(k_point_cols, grid1(ncp), dim3(256), 0, c->stream, ncp, (const int64_t *)c->d_rowptr,
                           (const int32_t *)c->d_col, pcol);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            (void)hipFree(pcol);
            return fail(IGAP_ERR_HIP, "k_point_cols: %s", hipGetErrorString(e));
        }
        c->d_pcol = pcol;
    }
    return IGAP_OK;
}

static int spmv(igap_gpu_ctx *c, int p, const double *x, double *y, int mode = 0, const double *b = nullptr)
{
    const int64_t n = c->ndof, ncp = n / 3;
    int rc = ensure_pcol(c);
    if (rc) return rc;
    const dim3 g((unsigned)((ncp + 3) / 4)), bs(256);
    const int64_t *rp = (const int64_t *)c->d_rowptr;
    const int32_t *pc = (const int32_t *)c->d_pcol;
#define IGAP_SPMV(T, V, M) hipLaunchKernelGGL((k_spmv_pb<T, M>), g, bs, 0, c->stream, ncp, rp, pc, (const T *)V, x, b, y)
    if (p == IGAP_F64) {
        if (mode == 0) IGAP_SPMV(double, c->d_val64, 0);
        else if (mode == 1) IGAP_SPMV(double, c->d_val64, 1);
        else IGAP_SPMV(double, c->d_val64, 2);
    } else {
        if (mode == 0) IGAP_SPMV(float, c->d_val32, 0);
        else if (mode == 1) IGAP_SPMV(float, c->d_val32, 1);
        else IGAP_SPMV(float, c->d_val32, 2);
    }
#undef IGAP_SPMV
    HIPCHK(hipGetLastError());
    return IGAP_OK;
}

static int nrm2(igap_gpu_ctx *c, Solver *s, const double *x, double *out)
{
    BLCHK(rocblas_dnrm2(s->bl, (rocblas_int)c->ndof, x, 1, out));
    return IGAP_OK;
}

// h[k] = V[:, k] . w for k < nk (columns of V at stride ld)
static int dots(igap_gpu_ctx *c, Solver *s, const double *V, int64_t ld, int nk, const double *w, double *h)
{
    const int64_t n = c->ndof;
    const int nch = (int)((n + DOT_CHUNK - 1) / DOT_CHUNK);
    hipLaunchKernelGGL(k_dot_part, dim3((unsigned)nch, (unsigned)nk), dim3(256), 0, c->stream, n, ld, V, w, s->part);
    hipLaunchKernelGGL(k_dot_final, dim3((unsigned)nk), dim3(64), 0, c->stream, nch, (const double *)s->part, h);
    HIPCHK(hipGetLastError());
    return IGAP_OK;
}

// ---------------------------------------------------------------- multigrid
// Geometric multigrid preconditioner on the B-spline hierarchy of the patch
// (non-periodic patches with unit weights).  Level l+1 halves every direction
// of level l whose element count is even and >= 4, by removing every other
// interior knot, so its quadratic spline space is a subspace of level l's:
// the prolongation P (coarse control points -> fine) is exact, computed per
// direction by collocation at the fine Greville abscissae and checked at
// sample points.  A coarse level is a context of its own (same physics and
// boundary types) whose geometry and state are the fine ones interpolated at
// the coarse Greville abscissae (T, exact for fields in the coarse space), so
// its tangent is re-assembled by the same kernels at every refactorisation.
// One V-cycle: Chebyshev smoothing of degree MG_CHEB_DEG with point-block
// (3x3) Jacobi on [0.02, 1.1] lambda_max(D^-1 A) (lambda_max from power
// iterations at set-up), residual restricted by P^T (zero on coarse
// Dirichlet rows), coarse correction prolongated by P (zero on fine
// Dirichlet rows), Chebyshev post-smoothing; the coarsest level is solved
// with its explicit inverse (rocSOLVER LU and inversion at set-up, one GEMV
// per cycle) when it has at most MG_DENSE_MAX dofs.  A fixed linear
// operator, so it serves as GMRES's right preconditioner.
namespace igap {
constexpr int MG_MAXLEV = 8;
constexpr int64_t MG_COARSE_NDOF = 1500;
constexpr int64_t MG_DENSE_MAX = 8000;
// Preconditioner precision tag beside IGAP_F64 / IGAP_F32: multigrid levels
// set up in FP32 whose smoothing and residual products use the 16-bit blocks
// of k_pack_h16 (pc_fp32 = 2)
constexpr int PC_F16 = 2;
constexpr int MG_CHEB_DEG = 3;
constexpr int MG_COARSE_DEG = 12;
constexpr int MG_POWER_ITS = 20;        // power iterations of the first set-up (from a hashed vector)
constexpr int MG_POWER_ITS_WARM = 8;    // of later ones, from the previous set-up's vector

struct Op1 {   // 1-D operator in CSR form on the device
    int rows = 0, cols = 0;
    int *rp = nullptr, *ci = nullptr;
    double *w = nullptr;
};

struct MGLevel {
    igap_gpu_ctx *c = nullptr;   // level 0: the caller's context (not owned)
    Solver *s = nullptr;         // level 0: the caller's solver (not owned)
    int halve[3] = {0, 0, 0};    // directions halved towards level l+1
    Op1 P[3], R[3], T[3];        // to level l+1: prolongation, P^T, interpolation
    double *r = nullptr, *t = nullptr, *b = nullptr, *x = nullptr;
    double *Dinv = nullptr;      // 3x3 diagonal-block inverses (smoother)
    double lmax = 1.0;           // lambda_max(D^-1 A) estimate
    double *pv = nullptr;        // the power iteration's last vector, the next set-up's start
    int pv_ready = 0;
    double *dense = nullptr;     // coarsest level: LU factors (column-major)
    rocblas_int *ipiv = nullptr, *info = nullptr;
    uint32_t *h16 = nullptr;     // 16-bit smoothing matrix (k_pack_h16 layout), pc_fp32 = 2
    float *rs = nullptr;         // its row scales
};

struct MGHier {
    int nlev = 0;
    MGLevel lev[MG_MAXLEV];
    double *tmp1 = nullptr, *tmp2 = nullptr;
    int bc_type[36];
};

static void op1_free(Op1 &o)
{
    if (o.rp) (void)hipFree(o.rp);
    if (o.ci) (void)hipFree(o.ci);
    if (o.w) (void)hipFree(o.w);
    o = Op1();
}

void mg_destroy(MGHier *g)
{
    if (!g) return;
    for (int l = 0; l < g->nlev; l++) {
        MGLevel &L = g->lev[l];
        for (int d = 0; d < 3; d++) {
            op1_free(L.P[d]);
            op1_free(L.R[d]);
            op1_free(L.T[d]);
        }
        double *v[] = {L.r, L.t, L.b, L.x, L.pv, L.Dinv, L.dense};
        for (double *q : v)
            if (q) (void)hipFree(q);
        if (L.h16) (void)hipFree(L.h16);
        if (L.rs) (void)hipFree(L.rs);
        if (L.ipiv) (void)hipFree(L.ipiv);
        if (L.info) (void)hipFree(L.info);
        if (l > 0 && L.c) igap_gpu_destroy(L.c);
    }
    if (g->tmp1) (void)hipFree(g->tmp1);
    if (g->tmp2) (void)hipFree(g->tmp2);
    delete g;
}
}  // namespace igap

// quadratic B-splines on an open knot vector kv (nfun + 3 knots): the three
// nonzero values at x (span clamped to the last non-degenerate one, so the
// right end belongs to it) and the index of the first
static void bs2_eval(const std::vector<double> &kv, int nfun, double x, int &first, double N[3])
{
    int m = 2;
    while (m < nfun - 1 && !(x < kv[m + 1])) m++;
    double left[3] = {0, 0, 0}, right[3] = {0, 0, 0};
    N[0] = 1.0;
    for (int j = 1; j <= 2; j++) {
        left[j] = x - kv[m + 1 - j];
        right[j] = kv[m + j] - x;
        double saved = 0.0;
        for (int r = 0; r < j; r++) {
            const double t = N[r] / (right[r + 1] + left[j - r]);
            N[r] = saved + right[r + 1] * t;
            saved = left[j - r] * t;
        }
        N[j] = saved;
    }
    first = m - 2;
}

// A X = B for dense row-major A (n x n) and B (n x k), partial pivoting; B is
// overwritten by X.  Returns false for a singular A.
static bool dense_solve(int n, int k, std::vector<double> A, std::vector<double> &B)
{
    for (int col = 0; col < n; col++) {
        int piv = col;
        for (int r = col + 1; r < n; r++)
            if (std::fabs(A[(size_t)r * n + col]) > std::fabs(A[(size_t)piv * n + col])) piv = r;
        if (A[(size_t)piv * n + col] == 0.0) return false;
        if (piv != col) {
            for (int j = 0; j < n; j++) std::swap(A[(size_t)col * n + j], A[(size_t)piv * n + j]);
            for (int j = 0; j < k; j++) std::swap(B[(size_t)col * k + j], B[(size_t)piv * k + j]);
        }
        const double d = A[(size_t)col * n + col];
        for (int r = col + 1; r < n; r++) {
            const double f = A[(size_t)r * n + col] / d;
            if (f == 0.0) continue;
            for (int j = col; j < n; j++) A[(size_t)r * n + j] -= f * A[(size_t)col * n + j];
            for (int j = 0; j < k; j++) B[(size_t)r * k + j] -= f * B[(size_t)col * k + j];
        }
    }
    for (int r = n - 1; r >= 0; r--) {
        for (int j = 0; j < k; j++) {
            double v = B[(size_t)r * k + j];
            for (int q = r + 1; q < n; q++) v -= A[(size_t)r * n + q] * B[(size_t)q * k + j];
            B[(size_t)r * k + j] = v / A[(size_t)r * n + r];
        }
    }
    return true;
}

// row-major dense (rows x cols) -> device CSR, entries below 1e-15 dropped
static int op1_upload(const std::vector<double> &M, int rows, int cols, Op1 &o)
{
    std::vector<int> rp(rows + 1, 0), ci;
    std::vector<double> w;
    for (int r = 0; r < rows; r++) {
        for (int j = 0; j < cols; j++)
            if (std::fabs(M[(size_t)r * cols + j]) > 1e-15) {
                ci.push_back(j);
                w.push_back(M[(size_t)r * cols + j]);
            }
        rp[r + 1] = (int)ci.size();
    }
    o.rows = rows;
    o.cols = cols;
    DALLOC(&o.rp, sizeof(int) * (rows + 1));
    DALLOC(&o.ci, sizeof(int) * std::max<size_t>(1, ci.size()));
    DALLOC(&o.w, sizeof(double) * std::max<size_t>(1, w.size()));
    HIPCHK(hipMemcpy(o.rp, rp.data(), sizeof(int) * (rows + 1), hipMemcpyHostToDevice));
    if (!ci.empty()) {
        HIPCHK(hipMemcpy(o.ci, ci.data(), sizeof(int) * ci.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(o.w, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice));
    }
    return IGAP_OK;
}

// 1-D operators between the fine open knot vector kf (nel elements, even)
// and its coarsening kc: P (nf x nc), T (nc x nf), nf = nel + 2, nc = nel/2 + 2
static int mg_1d(const std::vector<double> &kf, int nel, std::vector<double> &kc, std::vector<double> &P, std::vector<double> &T)
{
    const int ncel = nel / 2, nf = nel + 2, nc = ncel + 2;
    kc.assign(ncel + 5, 0.0);
    for (int i = 0; i <= 2; i++) kc[i] = kf[2];
    for (int i = 0; i <= ncel; i++) kc[2 + i] = kf[2 + 2 * i];
    kc[ncel + 3] = kc[ncel + 4] = kf[nel + 2];
    auto greville = [](const std::vector<double> &kv, int i) { return 0.5 * (kv[i + 1] + kv[i + 2]); };
    auto row = [](const std::vector<double> &kv, int nfun, double x, std::vector<double> &out, int off, int stride) {
        int f;
        double N[3];
        bs2_eval(kv, nfun, x, f, N);
        for (int a = 0; a < 3; a++) out[(size_t)off + (size_t)(f + a) * stride] = N[a];
    };
    // P = Af^-1 Rc: Af[s][i] = Nf_i(gf_s), Rc[s][j] = Nc_j(gf_s)
    std::vector<double> Af((size_t)nf * nf, 0.0), Rc((size_t)nf * nc, 0.0);
    for (int s_ = 0; s_ < nf; s_++) {
        const double x = greville(kf, s_);
        row(kf, nf, x, Af, s_ * nf, 1);
        row(kc, nc, x, Rc, s_ * nc, 1);
    }
    if (!dense_solve(nf, nc, Af, Rc)) return fail(IGAP_ERR_STATE, "multigrid: singular fine collocation matrix");
    P = Rc;
    // T = Ac^-1 Bf: Ac[s][j] = Nc_j(gc_s), Bf[s][i] = Nf_i(gc_s)
    std::vector<double> Ac((size_t)nc * nc, 0.0), Bf((size_t)nc * nf, 0.0);
    for (int s_ = 0; s_ < nc; s_++) {
        const double x = greville(kc, s_);
        row(kc, nc, x, Ac, s_ * nc, 1);
        row(kf, nf, x, Bf, s_ * nf, 1);
    }
    if (!dense_solve(nc, nf, Ac, Bf)) return fail(IGAP_ERR_STATE, "multigrid: singular coarse collocation matrix");
    T = Bf;
    // check the refinement relation Nc_j = sum_i P[i][j] Nf_i at 5 points per fine span
    double err = 0.0;
    for (int e = 0; e < nel; e++)
        for (int q = 0; q < 5; q++) {
            const double x = kf[2 + e] + (q + 0.5) / 5.0 * (kf[3 + e] - kf[2 + e]);
            std::vector<double> nfv(nf, 0.0), ncv(nc, 0.0);
            row(kf, nf, x, nfv, 0, 1);
            row(kc, nc, x, ncv, 0, 1);
            for (int j = 0; j < nc; j++) {
                double v = 0.0;
                for (int i = 0; i < nf; i++) v += P[(size_t)i * nc + j] * nfv[i];
                err = std::max(err, std::fabs(v - ncv[j]));
            }
        }
    if (err > 1e-10) return fail(IGAP_ERR_STATE, "multigrid: refinement relation violated (%.2e)", err);
    return IGAP_OK;
}

// host: out = (T0 x T1 x T2) in over a 3-component field on (n0,n1,n2)
// points, ops given as row-major dense matrices (nullptr = identity)
// sep: the input is separable (component q depends on the index along q
// alone); component q is then interpolated along q only and copied along the
// other directions, on which it is constant, so the result is separable too
static std::vector<double> apply3_host(const std::vector<double> &in, const int nin[3], const std::vector<double> *op[3],
                                       const int nout[3], bool sep)
{
    std::vector<double> cur = in;
    int dims[3] = {nin[0], nin[1], nin[2]};
    for (int d = 0; d < 3; d++) {
        if (!op[d]) continue;
        int od[3] = {dims[0], dims[1], dims[2]};
        od[d] = nout[d];
        std::vector<double> nxt((size_t)3 * od[0] * od[1] * od[2], 0.0);
        const std::vector<double> &M = *op[d];
        for (int i0 = 0; i0 < od[0]; i0++)
            for (int i1 = 0; i1 < od[1]; i1++)
                for (int i2 = 0; i2 < od[2]; i2++) {
                    const int rowi = d == 0 ? i0 : d == 1 ? i1 : i2;
                    double acc[3] = {0, 0, 0};
                    for (int j = 0; j < dims[d]; j++) {
                        const double wv = M[(size_t)rowi * dims[d] + j];
                        if (wv == 0.0) continue;
                        const int a0 = d == 0 ? j : i0, a1 = d == 1 ? j : i1, a2 = d == 2 ? j : i2;
                        const size_t idx = ((size_t)a0 * dims[1] + a1) * dims[2] + a2;
                        for (int q = 0; q < 3; q++) acc[q] += wv * cur[3 * idx + q];
                    }
                    const size_t o = ((size_t)i0 * od[1] + i1) * od[2] + i2;
                    if (sep) {
                        const int a0 = d == 0 ? 0 : i0, a1 = d == 1 ? 0 : i1, a2 = d == 2 ? 0 : i2;
                        const size_t idx0 = ((size_t)a0 * dims[1] + a1) * dims[2] + a2;
                        for (int q = 0; q < 3; q++)
                            if (q != d) acc[q] = cur[3 * idx0 + q];
                    }
                    for (int q = 0; q < 3; q++) nxt[3 * o + q] = acc[q];
                }
        cur.swap(nxt);
        dims[d] = nout[d];
    }
    return cur;
}

// device: out = (op0 x op1 x op2) in on the control-point grid; which:
// 0 = P (coarse -> fine of level l), 1 = P^T, 2 = T (fine -> coarse)
static int mg_transfer(igap_gpu_ctx *c, MGHier *g, int l, int which, const double *in, double *out,
                       const int32_t *mask = nullptr, int accumulate = 0)
{
    MGLevel &F = g->lev[l];
    const igap_gpu_ctx *cf = F.c, *cc = g->lev[l + 1].c;
    int dims[3];
    for (int d = 0; d < 3; d++) dims[d] = which == 0 ? cc->nb[d] : cf->nb[d];
    int dl[3], nd = 0;
    for (int d = 0; d < 3; d++)
        if (F.halve[d]) dl[nd++] = d;
    if (nd == 0) return fail(IGAP_ERR_STATE, "multigrid: level without coarsening");
    const double *src = in;
    for (int q = 0; q < nd; q++) {
        const int d = dl[q];
        const Op1 &o = which == 0 ? F.P[d] : which == 1 ? F.R[d] : F.T[d];
        int od[3] = {dims[0], dims[1], dims[2]};
        od[d] = o.rows;
        double *dst = q == nd - 1 ? out : (q % 2 == 0 ? g->tmp1 : g->tmp2);
        const int64_t n = 3LL * od[0] * od[1] * od[2];
        int64_t blocks = (n + 255) / 256;
        if (blocks > 8192) blocks = 8192;
        const bool last = q == nd - 1;
        hipLaunchKernelGGL(k_apply1d, dim3((unsigned)blocks), dim3(256), 0, c->stream, d, od[0], od[1], od[2], dims[d], o.rp, o.ci,
                           o.w, src, dst, last ? mask : (const int32_t *)nullptr, last ? accumulate : 0);
        src = dst;
        dims[d] = o.rows;
    }
    HIPCHK(hipGetLastError());
    return IGAP_OK;
}

static int mg_build_levels(igap_gpu_ctx *c, Solver *s, MGHier *g, int maxlev);

// the multigrid hierarchy, built on first use and published only once
// complete: a failed build releases its coarse contexts and buffers
static int mg_build(igap_gpu_ctx *c, Solver *s, int maxlev)
{
    if (s->mg) return IGAP_OK;
    for (int d = 0; d < 3; d++)
        if (c->periodic[d]) return fail(IGAP_ERR_ARG, "IGAP_PC_MG needs a non-periodic patch");
    for (double wv : c->w_h)
        if (wv != 1.0) return fail(IGAP_ERR_ARG, "IGAP_PC_MG needs unit weights (a B-spline patch)");
    if (maxlev <= 0 || maxlev > MG_MAXLEV) maxlev = MG_MAXLEV;
    MGHier *g = new MGHier();
    const int rc = mg_build_levels(c, s, g, maxlev);
    if (rc) {
        mg_destroy(g);
        return rc;
    }
    s->mg = g;
    return IGAP_OK;
}

static int mg_build_levels(igap_gpu_ctx *c, Solver *s, MGHier *g, int maxlev)
{
    std::memcpy(g->bc_type, c->bc_h.bc_type, sizeof g->bc_type);
    g->lev[0].c = c;
    g->lev[0].s = s;
    g->nlev = 1;
    const size_t bytes0 = sizeof(double) * (size_t)c->ndof;
    DALLOC(&g->tmp1, bytes0);
    DALLOC(&g->tmp2, bytes0);
    while (g->nlev < maxlev) {
        const int lf = g->nlev - 1;
        MGLevel &F = g->lev[lf];
        igap_gpu_ctx *cf = F.c;
        if (cf->ndof <= MG_COARSE_NDOF) break;
        int any = 0;
        for (int d = 0; d < 3; d++) {
            F.halve[d] = (cf->nelem[d] % 2 == 0 && cf->nelem[d] >= 4) ? 1 : 0;
            any |= F.halve[d];
        }
        if (!any) break;
        // coarse knots, 1-D operators, coarse geometry
        std::vector<double> kc[3], Pd[3], Td[3];
        int nelc[3], nfd[3], ncd[3];
        const std::vector<double> *Top[3] = {nullptr, nullptr, nullptr};
        for (int d = 0; d < 3; d++) {
            nfd[d] = cf->nb[d];
            if (F.halve[d]) {
                int rc = mg_1d(cf->knot_h[d], cf->nelem[d], kc[d], Pd[d], Td[d]);
                if (rc) return rc;
                nelc[d] = cf->nelem[d] / 2;
                Top[d] = &Td[d];
            } else {
                kc[d] = cf->knot_h[d];
                nelc[d] = cf->nelem[d];
            }
            ncd[d] = nelc[d] + 2;
        }
        std::vector<double> Xc = apply3_host(cf->X_h, nfd, Top, ncd, cf->M.separable != 0);
        std::vector<double> wc((size_t)ncd[0] * ncd[1] * ncd[2], 1.0);
        igap_mesh_desc md;
        std::memset(&md, 0, sizeof md);
        md.porder = 2;
        for (int d = 0; d < 3; d++) {
            md.nelem[d] = nelc[d];
            md.periodic[d] = 0;
            md.knot[d] = kc[d].data();
        }
        md.w = wc.data();
        md.X = Xc.data();
        MGLevel &C = g->lev[lf + 1];
        int rc = igap_gpu_create(&md, &c->phys_h, &c->bc_h, c->device, 0, &C.c);
        if (rc) return rc;
        g->nlev = lf + 2;   // counted from here on, so a failure below releases it
        // every level works on the fine context's stream (its rocSPARSE and
        // rocBLAS handles included), so the cycle needs no cross-stream waits
        HIPCHK(hipStreamSynchronize(C.c->stream));
        HIPCHK(hipStreamDestroy(C.c->stream));
        C.c->stream = c->stream;
        C.c->own_stream = 0;
        rc = solver_get(C.c, 1, &C.s);
        if (rc) return rc;
        for (int d = 0; d < 3; d++) {
            if (!F.halve[d]) continue;
            std::vector<double> Rt((size_t)ncd[d] * nfd[d]);
            for (int i = 0; i < nfd[d]; i++)
                for (int j = 0; j < ncd[d]; j++) Rt[(size_t)j * nfd[d] + i] = Pd[d][(size_t)i * ncd[d] + j];
            if ((rc = op1_upload(Pd[d], nfd[d], ncd[d], F.P[d]))) return rc;
            if ((rc = op1_upload(Rt, ncd[d], nfd[d], F.R[d]))) return rc;
            if ((rc = op1_upload(Td[d], ncd[d], nfd[d], F.T[d]))) return rc;
        }
    }
    for (int l = 0; l < g->nlev; l++) {
        MGLevel &L = g->lev[l];
        const size_t b = sizeof(double) * (size_t)L.c->ndof;
        double **v[] = {&L.r, &L.t, &L.b, &L.x, &L.pv};
        for (double **q : v) DALLOC(q, b);
        DALLOC(&L.Dinv, sizeof(double) * 3 * (size_t)L.c->ndof);
    }
    MGLevel &Lc = g->lev[g->nlev - 1];
    if (g->nlev > 1 && Lc.c->ndof <= MG_DENSE_MAX) {
        const size_t nn = (size_t)Lc.c->ndof;
        DALLOC(&Lc.dense, sizeof(double) * nn * nn);
        DALLOC(&Lc.ipiv, sizeof(rocblas_int) * nn);
        DALLOC(&Lc.info, sizeof(rocblas_int));
    }
    return IGAP_OK;
}

// A x on multigrid level l in the preconditioner's precision (MODE of k_spmv_pb)
static int lvl_spmv(igap_gpu_ctx *c, MGHier *g, int l, int p, const double *x, double *y, int mode = 0, const double *b = nullptr)
{
    MGLevel &L = g->lev[l];
    if (p != PC_F16 || !L.h16) return spmv(L.c, p == PC_F16 ? IGAP_F32 : p, x, y, mode, b);
    const int64_t ncp = L.c->ndof / 3;
    const dim3 gr((unsigned)((ncp + 4 * H16_PPW - 1) / (4 * H16_PPW))), bs(256);
    const uint32_t *hv = (const uint32_t *)L.h16;
    const float *rs = (const float *)L.rs;
    const MeshK &M = L.c->M;
    if (M.periodic[0] || M.periodic[1] || M.periodic[2]) return fail(IGAP_ERR_STATE, "16-bit multigrid level on a periodic patch");
    PointGrid G;
    for (int d = 0; d < 3; d++) G.nb[d] = M.nb[d];
    G.L2 = stencil_prefix(M.nb[2], M.nb[2]);
    G.L12 = stencil_prefix(M.nb[1], M.nb[1]) * G.L2;
    if (mode == 0) hipLaunchKernelGGL((k_spmv_h16<0, H16_PPW>), gr, bs, 0, c->stream, ncp, G, hv, rs, x, b, y);
    else if (mode == 1) hipLaunchKernelGGL((k_spmv_h16<1, H16_PPW>), gr, bs, 0, c->stream, ncp, G, hv, rs, x, b, y);
    else hipLaunchKernelGGL((k_spmv_h16<2, H16_PPW>), gr, bs, 0, c->stream, ncp, G, hv, rs, x, b, y);
    HIPCHK(hipGetLastError());
    return IGAP_OK;
}

// x = Chebyshev(b; x) with point-block Jacobi on level l: deg steps from
// x = 0 (zero_init) or from the given x
static int mg_cheb(igap_gpu_ctx *c, MGHier *g, int l, int p, const double *b, double *x, int zero_init, int deg)
{
    MGLevel &L = g->lev[l];
    igap_gpu_ctx *cl = L.c;
    const int64_t n = cl->ndof, ncp = n / 3;
    const dim3 gn = grid1(n), gp = grid1(ncp), bs(256);
    // lower end 0.02 lambda_max, measured on 64x16x16 and 160x40x40 load steps:
    // 0.1 is 5-7% slower, 0.005 needs more Krylov iterations on the small patch
    const double lo = 0.02 * L.lmax, hi = 1.1 * L.lmax;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    int rc;
    (void)gn;
    if (zero_init) {
        hipLaunchKernelGGL(k_cheb_first, gp, bs, 0, c->stream, ncp, (const double *)L.Dinv, b, x, L.t, L.r, 1.0 / theta);
    } else {
        if ((rc = lvl_spmv(c, g, l, p, x, L.r, 1, b))) return rc;   // r = b - A x
        hipLaunchKernelGGL(k_cheb_step, gp, bs, 0, c->stream, ncp, (const double *)L.Dinv, (const double *)L.r, L.t, x, 0.0, 1.0 / theta,
                           1);
    }
    double rho = 1.0 / sigma;
    for (int k = 1; k < deg; k++) {
        if ((rc = lvl_spmv(c, g, l, p, L.t, L.r, 2))) return rc;     // r -= A d
        const double rn = 1.0 / (2.0 * sigma - rho);
        hipLaunchKernelGGL(k_cheb_step, gp, bs, 0, c->stream, ncp, (const double *)L.Dinv, (const double *)L.r, L.t, x, rn * rho,
                           2.0 * rn / delta, 1);
        rho = rn;
    }
    HIPCHK(hipGetLastError());
    return IGAP_OK;
}

// smoother data of level l: block-diagonal inverses and lambda_max(D^-1 A)
static int mg_smoother_setup(igap_gpu_ctx *c, MGHier *g, int l, int p)
{
    MGLevel &L = g->lev[l];
    igap_gpu_ctx *cl = L.c;
    const int64_t n = cl->ndof, ncp = n / 3;
    const dim3 gn = grid1(n), gp = grid1(ncp), bs(256);
    if (p == IGAP_F64)
        hipLaunchKernelGGL(k_block_diag_inv<double>, gp, bs, 0, c->stream, ncp, (const int64_t *)cl->d_rowptr, (const int32_t *)cl->d_col,
                           (const double *)cl->d_val64, L.Dinv);
    else
        hipLaunchKernelGGL(k_block_diag_inv<float>, gp, bs, 0, c->stream, ncp, (const int64_t *)cl->d_rowptr, (const int32_t *)cl->d_col,
                           (const float *)cl->d_val32, L.Dinv);
    // power iteration on D^-1 A: v <- D^-1 A v / |v|, normalised on the
    // device; lambda = |D^-1 A v| for the last unit v (one copy back).  The
    // first set-up starts from a hashed vector, later ones (a new tangent of
    // the same patch) from the previous set-up's last vector, with fewer
    // iterations
    int its = MG_POWER_ITS;
    if (L.pv_ready) {
        HIPCHK(launch_copy_f64(L.x, L.pv, n, c->stream));
        its = MG_POWER_ITS_WARM;
    } else {
        hipLaunchKernelGGL(k_fill_hash, gn, bs, 0, c->stream, n, L.x);
    }
    int rc;
    for (int it = 0; it < its; it++) {
        if ((rc = dots(cl, L.s, L.x, n, 1, L.x, L.s->h))) return rc;
        hipLaunchKernelGGL(k_scale_rsqrt, gn, bs, 0, c->stream, n, (const double *)L.s->h, L.x);
        if ((rc = spmv(cl, p, L.x, L.r))) return rc;
        hipLaunchKernelGGL(k_cheb_step, gp, bs, 0, c->stream, ncp, (const double *)L.Dinv, (const double *)L.r, L.x, L.t, 0.0, 1.0, 0);
    }
    if ((rc = dots(cl, L.s, L.x, n, 1, L.x, L.s->h))) return rc;
    HIPCHK(launch_copy_f64(L.pv, L.x, n, c->stream));
    double q = 0;
    HIPCHK(hipMemcpyAsync(&q, L.s->h, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const double lam = std::sqrt(q);
    L.lmax = lam > 0 ? lam : 1.0;
    L.pv_ready = lam > 0 ? 1 : 0;
    return IGAP_OK;
}


// coarse states and tangents for the current fine state (U_hist of the fine
// context, level 0 already holding the Newton iterate), smoothers, and the
// coarsest level's LU
static int mg_setup(igap_gpu_ctx *c, Solver *s, int pm)
{
    MGHier *g = s->mg;
    const int p = pm == PC_F16 ? IGAP_F32 : pm;   // precision of the level tangents
    const bool bc_changed = std::memcmp(g->bc_type, c->bc_h.bc_type, sizeof g->bc_type) != 0;
    if (bc_changed) std::memcpy(g->bc_type, c->bc_h.bc_type, sizeof g->bc_type);
    for (int l = 1; l < g->nlev; l++) {
        igap_gpu_ctx *cf = g->lev[l - 1].c, *cc = g->lev[l].c;
        if (bc_changed) {
            int rc = igap_gpu_set_bc(cc, &c->bc_h);
            if (rc) return rc;
        }
        for (int h = 0; h <= cf->torder; h++) {
            int rc = mg_transfer(c, g, l - 1, 2, cf->d_Uh + (size_t)h * cf->ndof, cc->d_Uh + (size_t)h * cc->ndof);
            if (rc) return rc;
        }
        int rc = igap_gpu_tangent(cc, cc->d_Uh, nullptr, p);
        if (rc) return rc;
    }
    for (int l = 0; l < g->nlev; l++) {
        MGLevel &L = g->lev[l];
        if (L.dense) {
            const int64_t n = L.c->ndof;
            HIPCHK(launch_fill_f64(L.dense, 0.0, n * n, c->stream));
            if (p == IGAP_F64)
                hipLaunchKernelGGL(k_csr_to_dense<double>, grid1(n), dim3(256), 0, c->stream, n, (const int64_t *)L.c->d_rowptr,
                                   (const int32_t *)L.c->d_col, (const double *)L.c->d_val64, L.dense);
            else
                hipLaunchKernelGGL(k_csr_to_dense<float>, grid1(n), dim3(256), 0, c->stream, n, (const int64_t *)L.c->d_rowptr,
                                   (const int32_t *)L.c->d_col, (const float *)L.c->d_val32, L.dense);
            BLCHK(rocsolver_dgetrf(L.s->bl, (rocblas_int)n, (rocblas_int)n, L.dense, (rocblas_int)n, L.ipiv, L.info));
            rocblas_int info = 0;
            HIPCHK(hipMemcpyAsync(&info, L.info, sizeof info, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            if (info != 0) return fail(IGAP_ERR_STATE, "multigrid: coarsest matrix singular (getrf info %d)", (int)info);
            // explicit inverse: one dense matrix-vector product per cycle
            BLCHK(rocsolver_dgetri(L.s->bl, (rocblas_int)n, L.dense, (rocblas_int)n, L.ipiv, L.info));
            HIPCHK(hipMemcpyAsync(&info, L.info, sizeof info, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            if (info != 0) return fail(IGAP_ERR_STATE, "multigrid: coarsest inverse failed (getri info %d)", (int)info);
        } else {
            int rc = mg_smoother_setup(c, g, l, p);
            if (rc) return rc;
            if (pm == PC_F16) {
                igap_gpu_ctx *cl = L.c;
                const int64_t ncp = cl->ndof / 3;
                if ((rc = ensure_pcol(cl))) return rc;
                if (!L.h16) {
                    DALLOC(&L.h16, sizeof(uint32_t) * 5 * (size_t)(cl->nnz / 9));
                    DALLOC(&L.rs, sizeof(float) * (size_t)cl->ndof);
                }
                hipLaunchKernelGGL(k_pack_h16, dim3((unsigned)((ncp + 3) / 4)), dim3(256), 0, c->stream, ncp,
                                   (const int64_t *)cl->d_rowptr, (const float *)cl->d_val32, L.h16, L.rs);
                HIPCHK(hipGetLastError());
            }
        }
    }
    return IGAP_OK;
}

// x = V-cycle(b) from level l (x and b distinct from the level's work vectors)
static int mg_cycle(igap_gpu_ctx *c, Solver *s, int p, int order, int l, const double *b, double *x)
{
    (void)order;
    MGHier *g = s->mg;
    MGLevel &L = g->lev[l];
    igap_gpu_ctx *cl = L.c;
    const int64_t n = cl->ndof;
    const dim3 gn = grid1(n), bs(256);
    int rc;
    if (l == g->nlev - 1) {
        if (L.dense) {
            const double one = 1.0, zero = 0.0;
            BLCHK(rocblas_dgemv(L.s->bl, rocblas_operation_none, (rocblas_int)n, (rocblas_int)n, &one, L.dense, (rocblas_int)n, b, 1,
                                &zero, x, 1));
            return IGAP_OK;
        }
        return mg_cheb(c, g, l, p, b, x, 1, MG_COARSE_DEG);
    }
    MGLevel &C = g->lev[l + 1];
    if ((rc = mg_cheb(c, g, l, p, b, x, 1, MG_CHEB_DEG))) return rc;
    // r = b - A x, restricted (zero on coarse Dirichlet rows)
    if ((rc = lvl_spmv(c, g, l, p, x, L.r, 1, b))) return rc;
    if ((rc = mg_transfer(c, g, l, 1, L.r, C.b, C.c->d_mask, 0))) return rc;
    if ((rc = mg_cycle(c, s, p, order, l + 1, C.b, C.x))) return rc;
    // x += P x_c (not on Dirichlet rows), then post-smoothing
    if ((rc = mg_transfer(c, g, l, 0, C.x, x, cl->d_mask, 1))) return rc;
    if ((rc = mg_cheb(c, g, l, p, b, x, 0, MG_CHEB_DEG))) return rc;
    HIPCHK(hipGetLastError());
    return IGAP_OK;
}

// precision of the preconditioner for a tangent of precision p: p itself,
// FP32 (pc_fp32 = 1, or 2 with ILU(0)), or FP32 multigrid levels smoothing
// with 16-bit matrices (pc_fp32 = 2)
static int pc_prec_of(const igap_solver_opts *o, int p)
{
    if (!o->pc_fp32) return p;
    return (o->pc_fp32 == 2 && o->pc == IGAP_PC_MG) ? PC_F16 : IGAP_F32;
}

// GMRES(m), right preconditioned: solve A d = b (d starts at 0)
static int gmres(igap_gpu_ctx *c, Solver *s, const igap_solver_opts *o, int p, const double *b, double *d, int *its, double *relres,
                 int *conv)
{
    *conv = 1;
    const int pp = pc_prec_of(o, p);   // preconditioner precision
    const int64_t n = c->ndof;
    const int m = s->m;
    const rocblas_int N = (rocblas_int)n;
    std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), g(m + 1), hv(m + 2), y(m);
    HIPCHK(launch_fill_f64(d, 0.0, n, c->stream));
    double bnorm = 0;
    int rc = nrm2(c, s, b, &bnorm);
    if (rc) return rc;
    *its = 0;
    *relres = 0;
    if (bnorm == 0.0) return IGAP_OK;
    double rnorm = bnorm;
    int total = 0;
    // r = b - A d; first cycle d = 0 so r = b
    HIPCHK(hipMemcpyAsync(s->w, b, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    while (total < o->ksp_maxit) {
        double beta = rnorm;
        // v0 = r / beta
        hipLaunchKernelGGL(k_axpby, grid1(n), dim3(256), 0, c->stream, n, 1.0 / beta, s->w, 0.0, s->V);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int j = 0;
        for (; j < m && total < o->ksp_maxit; j++) {
            double *vj = s->V + (size_t)j * n, *vn = s->V + (size_t)(j + 1) * n;
            rc = pc_apply(c, s, o->pc, pp, o->pc_order, vj, s->z);
            if (rc) return rc;
            rc = spmv(c, p, s->z, vn);
            if (rc) return rc;
            // classical Gram-Schmidt, twice
            const double one = 1.0, mone = -1.0;
            // (both projections, the norm and the normalisation stay on the
            // device; one copy back per step)
            const int hs = m + 2;
            for (int pass = 0; pass < 2; pass++) {
                rc = dots(c, s, s->V, n, j + 1, vn, s->h + pass * hs);
                if (rc) return rc;
                BLCHK(rocblas_dgemv(s->bl, rocblas_operation_none, N, j + 1, &mone, s->V, N, s->h + pass * hs, 1, &one, vn, 1));
            }
            rc = dots(c, s, vn, n, 1, vn, s->h + 2 * hs);
            if (rc) return rc;
            hipLaunchKernelGGL(k_scale_rsqrt, grid1(n), dim3(256), 0, c->stream, n, s->h + 2 * hs, vn);
            HIPCHK(hipGetLastError());
            std::vector<double> hp(3 * hs);
            HIPCHK(hipMemcpyAsync(hp.data(), s->h, sizeof(double) * 3 * hs, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            for (int k = 0; k <= j; k++) hv[k] = hp[k] + hp[hs + k];
            const double hn = hp[2 * hs] > 0.0 ? std::sqrt(hp[2 * hs]) : 0.0;
            // Givens
            for (int k = 0; k < j; k++) {
                double t = cs[k] * hv[k] + sn[k] * hv[k + 1];
                hv[k + 1] = -sn[k] * hv[k] + cs[k] * hv[k + 1];
                hv[k] = t;
            }
            double den = std::sqrt(hv[j] * hv[j] + hn * hn);
            cs[j] = den > 0 ? hv[j] / den : 1.0;
            sn[j] = den > 0 ? hn / den : 0.0;
            hv[j] = den;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            for (int k = 0; k <= j; k++) H[(size_t)k * m + j] = hv[k];
            total++;
            rnorm = std::fabs(g[j + 1]);
            if (rnorm <= o->ksp_rtol * bnorm || rnorm <= o->ksp_atol || hn == 0.0) { j++; break; }
        }
        // y = H^{-1} g ; d += M^{-1} V y
        for (int k = j - 1; k >= 0; k--) {
            double t = g[k];
            for (int l = k + 1; l < j; l++) t -= H[(size_t)k * m + l] * y[l];
            y[k] = t / H[(size_t)k * m + k];
        }
        HIPCHK(hipMemcpyAsync(s->h, y.data(), sizeof(double) * j, hipMemcpyHostToDevice, c->stream));
        const double one = 1.0, zero = 0.0;
        BLCHK(rocblas_dgemv(s->bl, rocblas_operation_none, N, j, &one, s->V, N, s->h, 1, &zero, s->w, 1));
        rc = pc_apply(c, s, o->pc, pp, o->pc_order, s->w, s->z);
        if (rc) return rc;
        hipLaunchKernelGGL(k_axpby, grid1(n), dim3(256), 0, c->stream, n, 1.0, s->z, 1.0, d);
        if (rnorm <= o->ksp_rtol * bnorm || rnorm <= o->ksp_atol) break;
        // true residual for the restart: w = b - A d
        rc = spmv(c, p, d, s->w);
        if (rc) return rc;
        hipLaunchKernelGGL(k_axpby, grid1(n), dim3(256), 0, c->stream, n, 1.0, b, -1.0, s->w);
        rc = nrm2(c, s, s->w, &rnorm);
        if (rc) return rc;
        if (rnorm <= o->ksp_rtol * bnorm) break;
    }
    *its = total;
    *relres = rnorm / bnorm;
    // KSP_CONVERGED_RTOL / _ATOL, or KSP_DIVERGED_ITS after ksp_maxit
    *conv = (rnorm <= o->ksp_rtol * bnorm || rnorm <= o->ksp_atol) ? 1 : 0;
    return IGAP_OK;
}

// ---------------------------------------------------------------- C ABI
extern "C" int igap_gpu_solver_default_opts(igap_solver_opts *o)
{
    if (!o) return fail(IGAP_ERR_ARG, "null argument");
    std::memset(o, 0, sizeof *o);
    o->snes_atol = 1e-10;      // SNESSetTolerances(abstol 1e-10, ...) src/interface.c:220
    o->snes_rtol = 0.0;
    o->snes_maxit = 50;
    o->ksp_rtol = 1e-5;        // PETSc default rtol (KSPSetTolerances(PETSC_DEFAULT,...) :226)
    o->ksp_atol = 1e-50;
    o->ksp_maxit = 2000;
    o->ksp_restart = 200;
    o->pc = IGAP_PC_ILU0;
    o->tangent_prec = IGAP_F64;
    o->linesearch = 1;
    o->pc_order = 0;
    o->pc_lag = 1;
    return IGAP_OK;
}

extern "C" int igap_gpu_snes_solve(igap_gpu_ctx *c, const igap_solver_opts *o, igap_solver_stats *st)
{
    using clk = std::chrono::steady_clock;
    if (!c || !o || !st) return fail(IGAP_ERR_ARG, "null argument");
    if (o->ksp_restart < 1 || o->ksp_restart > 2000) return fail(IGAP_ERR_ARG, "ksp_restart %d out of [1, 2000]", o->ksp_restart);
    if (o->tangent_prec != IGAP_F64 && o->tangent_prec != IGAP_F32) return fail(IGAP_ERR_ARG, "tangent_prec");
    if (o->pc_order != 0 && o->pc_order != 1) return fail(IGAP_ERR_ARG, "pc_order must be 0 (natural) or 1 (coloured)");
    if (o->pc_lag < 0) return fail(IGAP_ERR_ARG, "pc_lag must be >= 0");
    if (o->pc < IGAP_PC_ILU0 || o->pc > IGAP_PC_MG) return fail(IGAP_ERR_ARG, "pc %d unknown", o->pc);
    if (o->pc_fp32 < 0 || o->pc_fp32 > 2) return fail(IGAP_ERR_ARG, "pc_fp32 must be 0, 1 or 2");
    HIPCHK(hipSetDevice(c->device));
    std::memset(st, 0, sizeof *st);
    auto T0 = clk::now();
    auto ms = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    Solver *s = nullptr;
    int rc = solver_get(c, o->ksp_restart, &s);
    if (rc) return rc;
    const int64_t n = c->ndof;
    const int p = o->tangent_prec;
    if (p == IGAP_F64 && !igap_gpu_csr_val64(c)) return IGAP_ERR_NOMEM;
    if (p == IGAP_F32 && !igap_gpu_csr_val32(c)) return IGAP_ERR_NOMEM;
    // x = current iterate (U_hist level 0, the reference's solv->U)
    HIPCHK(hipMemcpyAsync(s->x, c->d_Uh, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    auto t = clk::now();
    rc = igap_gpu_residual(c, s->x, s->F);
    if (rc) return rc;
    double fnorm = 0;
    rc = nrm2(c, s, s->F, &fnorm);
    if (rc) return rc;
    st->t_residual_ms += ms(t, clk::now());
    st->n_residual++;
    st->fnorm0 = fnorm;
    st->converged = 0;
    double fprev_it = -1.0;   // |F| before the last Newton step
    for (int it = 0;; it++) {
        st->fnorm = fnorm;
        if (!std::isfinite(fnorm)) { st->converged = -4; break; }   // DIVERGED_FNORM_NAN
        if (fnorm <= o->snes_atol) { st->converged = 2; break; }    // CONVERGED_FNORM_ABS
        if (o->snes_rtol > 0 && fnorm <= o->snes_rtol * st->fnorm0) { st->converged = 3; break; }   // CONVERGED_FNORM_RELATIVE
        // IGAPSolvSolve's stagnation test (src/interface.c:318): a Newton step
        // that changes |F| by less than 1e-6 relative ends the solve, reported
        // as SNES_DIVERGED_LINE_SEARCH
        if (fprev_it > 0 && std::fabs(fprev_it - fnorm) / fprev_it < 1e-6) { st->converged = -6; break; }
        if (it >= o->snes_maxit) { st->converged = -5; break; }     // DIVERGED_MAX_IT
        // tangent + preconditioner
        t = clk::now();
        rc = igap_gpu_tangent(c, s->x, nullptr, p);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(c->stream));
        st->t_tangent_ms += ms(t, clk::now());
        t = clk::now();
        const bool refactor = (it % (o->pc_lag > 1 ? o->pc_lag : 1)) == 0;
        const int pp = pc_prec_of(o, p);   // preconditioner precision
        if (refactor && (pp == PC_F16 ? IGAP_F32 : pp) != p && (o->pc == IGAP_PC_ILU0 || o->pc == IGAP_PC_MG)) {
            // FP32 copy of the FP64 tangent for the preconditioner
            float *v32 = igap_gpu_csr_val32(c);
            if (!v32) return IGAP_ERR_NOMEM;
            hipLaunchKernelGGL(k_d2f, grid1(c->nnz), dim3(256), 0, c->stream, c->nnz, (const double *)c->d_val64, v32);
            HIPCHK(hipGetLastError());
        }
        if (o->pc == IGAP_PC_ILU0 && refactor) {
            rc = o->pc_order == 1 ? pc_ilu_setup_mc(c, s, pp) : pc_ilu_setup(c, s, pp);
            if (rc) return rc;
        } else if (o->pc == IGAP_PC_MG && refactor) {
            if ((rc = mg_build(c, s, o->mg_levels))) return rc;
            if ((rc = mg_setup(c, s, pp))) return rc;
        } else if (o->pc == IGAP_PC_JACOBI && refactor) {
            if (p == IGAP_F64) hipLaunchKernelGGL(k_diag_inv<double>, grid1(n), dim3(256), 0, c->stream, n, c->d_rowptr, c->d_col, c->d_val64, s->diag);
            else hipLaunchKernelGGL(k_diag_inv<float>, grid1(n), dim3(256), 0, c->stream, n, c->d_rowptr, c->d_col, c->d_val32, s->diag);
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        st->t_pc_ms += ms(t, clk::now());
        if (refactor) {
            s->pc_ready = 1;
            s->pc_kind = o->pc;
            s->pc_prec = pp;
            s->pc_ord = o->pc_order;
        }
        // solve K d = -F
        t = clk::now();
        hipLaunchKernelGGL(k_axpby, grid1(n), dim3(256), 0, c->stream, n, -1.0, s->F, 0.0, s->Ft);
        int kits = 0, kconv = 1;
        double relres = 0;
        rc = gmres(c, s, o, p, s->Ft, s->d, &kits, &relres, &kconv);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(c->stream));
        st->t_ksp_ms += ms(t, clk::now());
        st->ksp_its += kits;
        st->last_ksp_relres = relres;
        // SNES_DIVERGED_LINEAR_SOLVE: PETSc gives up on the first failed
        // linear solve (maxLinearSolveFailures 1) and keeps the iterate
        if (!kconv) { st->converged = -3; break; }
        // line search
        t = clk::now();
        double lambda = 1.0, f0 = 0.5 * fnorm * fnorm, slope = -fnorm * fnorm;
        double lprev = 0, fprev = 0, ftrial = 0, fn_trial = 0;
        bool ls_ok = true;
        for (int ls = 0;; ls++) {
            hipLaunchKernelGGL(k_waxpy, grid1(n), dim3(256), 0, c->stream, n, lambda, s->d, s->x, s->xt);
            rc = igap_gpu_residual(c, s->xt, s->Ft);
            if (rc) return rc;
            st->n_residual++;
            rc = nrm2(c, s, s->Ft, &fn_trial);
            if (rc) return rc;
            ftrial = 0.5 * fn_trial * fn_trial;
            if (!o->linesearch || (std::isfinite(ftrial) && ftrial <= f0 + 1e-4 * lambda * slope)) break;
            if (lambda < 1e-12 || ls >= 40) { ls_ok = false; break; }
            double lnew;
            if (ls == 0 || !std::isfinite(ftrial)) {
                lnew = std::isfinite(ftrial) ? -slope / (2.0 * (ftrial - f0 - slope)) : 0.1;
            } else {
                double t1 = ftrial - f0 - lambda * slope, t2 = fprev - f0 - lprev * slope;
                double a = (t1 / (lambda * lambda) - t2 / (lprev * lprev)) / (lambda - lprev);
                double b = (-lprev * t1 / (lambda * lambda) + lambda * t2 / (lprev * lprev)) / (lambda - lprev);
                if (a == 0.0) lnew = -slope / (2.0 * b);
                else {
                    double disc = b * b - 3.0 * a * slope;
                    lnew = disc < 0 ? 0.5 * lambda : (-b + std::sqrt(disc)) / (3.0 * a);
                }
            }
            if (!(lnew == lnew)) lnew = 0.5 * lambda;
            if (lnew > 0.5 * lambda) lnew = 0.5 * lambda;
            if (lnew < 0.1 * lambda) lnew = 0.1 * lambda;
            lprev = lambda;
            fprev = ftrial;
            lambda = lnew;
            st->ls_backtracks++;
        }
        st->t_residual_ms += ms(t, clk::now());
        // SNES_DIVERGED_LINE_SEARCH: no sufficient decrease down to the
        // smallest step; the iterate stays where it was
        if (!ls_ok) { st->converged = -6; break; }
        std::swap(s->x, s->xt);
        std::swap(s->F, s->Ft);
        fprev_it = fnorm;
        fnorm = fn_trial;
        st->newton_its++;
        if (o->monitor) {
            std::fprintf(stderr, "  newton %d: |F| %.3e  ksp %d (relres %.1e)  lambda %.3g  %.0f ms\n", st->newton_its, fnorm,
                         kits, relres, lambda, ms(T0, clk::now()));
            std::fflush(stderr);
        }
    }
    HIPCHK(hipMemcpyAsync(c->d_Uh, s->x, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_R, s->F, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    st->t_total_ms = ms(T0, clk::now());
    return IGAP_OK;
}

// PCApply (and, for the multigrid, KSPSolve on its finest level's smoother)
// with the preconditioner the last Newton solve set up
extern "C" int igap_gpu_pc_apply(igap_gpu_ctx *c, int which, const double *b, double *x)
{
    if (!c || !b || !x) return fail(IGAP_ERR_ARG, "null argument");
    if (which != IGAP_PC_APPLY_FULL && which != IGAP_PC_APPLY_SMOOTHER) return fail(IGAP_ERR_ARG, "which = %d", which);
    if (b == x) return fail(IGAP_ERR_ARG, "b and x must differ");
    HIPCHK(hipSetDevice(c->device));
    Solver *s = c->solver;
    if (!s || !s->pc_ready) return fail(IGAP_ERR_STATE, "no preconditioner set up yet (igap_gpu_snes_solve builds it)");
    if (which == IGAP_PC_APPLY_SMOOTHER) {
        if (s->pc_kind != IGAP_PC_MG || !s->mg) return fail(IGAP_ERR_STATE, "the smoother belongs to IGAP_PC_MG");
        // the finest level always smooths (only a coarser one is ever dense)
        return mg_cheb(c, s->mg, 0, s->pc_prec, b, x, 1, MG_CHEB_DEG);
    }
    return pc_apply(c, s, s->pc_kind, s->pc_prec, s->pc_ord, b, x);
}

// MatMult on the context's own tangent values: y = K x (FP64 sums)
extern "C" int igap_gpu_tangent_apply(igap_gpu_ctx *c, int prec, const double *x, double *y)
{
    if (!c || !x || !y) return fail(IGAP_ERR_ARG, "null argument");
    if (prec != IGAP_F64 && prec != IGAP_F32) return fail(IGAP_ERR_ARG, "prec %d (IGAP_F64 or IGAP_F32)", prec);
    if (x == y) return fail(IGAP_ERR_ARG, "x and y must differ");
    if ((prec == IGAP_F64 && !c->d_val64) || (prec == IGAP_F32 && !c->d_val32))
        return fail(IGAP_ERR_STATE, "no %s tangent values assembled yet", prec == IGAP_F64 ? "FP64" : "FP32");
    HIPCHK(hipSetDevice(c->device));
    return spmv(c, prec, x, y);
}

// Self test of the point-block ILU(0): factor the context's current tangent
// values (precision prec) in the coloured order with k_mcb_ilu and with
// rocSPARSE csrilu0; *diff = max |a - b| / max |b| over the factor entries.
// *applicable = 0 (and *diff = 0) where the pattern has no point-block
// colour structure (then the solver uses rocSPARSE anyway).
extern "C" int igap_gpu_selftest_ilu(igap_gpu_ctx *c, int prec, int *applicable, double *diff)
{
    if (!c || !applicable || !diff) return fail(IGAP_ERR_ARG, "null argument");
    if (prec != IGAP_F64 && prec != IGAP_F32) return fail(IGAP_ERR_ARG, "prec must be IGAP_F64 or IGAP_F32");
    HIPCHK(hipSetDevice(c->device));
    if (prec == IGAP_F64 ? !c->d_val64 : !c->d_val32) return fail(IGAP_ERR_STATE, "no tangent of this precision assembled");
    Solver *s = nullptr;
    int rc = solver_get(c, c->solver ? c->solver->m : 200, &s);
    if (rc) return rc;
    rc = mc_build(c, s);
    if (rc) return rc;
    *applicable = s->mc_block_ok;
    *diff = 0.0;
    if (!s->mc_block_ok) return IGAP_OK;
    const int64_t nnz = c->nnz;
    std::vector<double> a(nnz), b(nnz);
    for (int impl = 0; impl < 2; impl++) {
        rc = pc_ilu_setup_mc(c, s, prec, impl);
        if (rc) return rc;
        std::vector<double> &dst = impl == 0 ? a : b;
        if (prec == IGAP_F64) {
            HIPCHK(hipMemcpyAsync(dst.data(), s->mc_ilu[prec], sizeof(double) * nnz, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        } else {
            std::vector<float> t(nnz);
            HIPCHK(hipMemcpyAsync(t.data(), s->mc_ilu[prec], sizeof(float) * nnz, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            for (int64_t j = 0; j < nnz; j++) dst[j] = t[j];
        }
    }
    double md = 0.0, mb = 0.0;
    for (int64_t j = 0; j < nnz; j++) {
        md = std::max(md, std::fabs(a[j] - b[j]));
        mb = std::max(mb, std::fabs(b[j]));
    }
    *diff = mb > 0 ? md / mb : md;
    return IGAP_OK;
}
