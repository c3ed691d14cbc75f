// igap_kernels.hip -- CDNA4 (gfx950) kernels of the IGA residual/tangent assembly.
//
//   k_residual_elem   per-element residual vectors R_e[81]
//                     (volume loop of Residual_iga, src/assmcore_iga.c:118-259)
//   k_residual_gather deterministic per-dof sum of R_e + Neumann load +
//                     Dirichlet rows R_d = U_d - g_d (:258, :273-497)
//   k_tangent_elem    per-element tangents K_e[81x81] on the matrix cores
//                     (volume loop of Tangent_iga, :607-764)
//   k_tangent_gather_rows / k_tangent_gather  deterministic per-nonzero sum of
//                     K_e into the CSR values with MatZeroRowsColumns(diag=1)
//                     fused (:763, :844); row-wise for FP64, column-wise FP32
//   k_csr_cols        column indices of the CSR pattern
//
// Element numbering e = (ex*ny + ey)*nz + ez (the reference's x->y->z loop
// order); local basis b = 9*bx + 3*by + bz; local dof 3*b + i.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <stdint.h>

#include "igap_device.h"

namespace igap {

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// ============================================================================
// residual: one workgroup (4 waves) per element.
//   A  the Gauss-point fields by sum factorisation (as in the tangent):
//      geometry, the physical u-jets of every history level -> LDS
//   P  pointwise residual jets r_q (static: one lane per Gauss point on one
//      wave; transient: four lanes per point, two rule nodes each)
//   T  projection R_e[b][k] = sum_q w_q sum_a r_q[k][a] R_a(q, b), written as
//      w_b sum_q (Q_q^T Phi_q^T w_q r_q[k]) . N_b(q): the transposes of the
//      chain rule (phys_jet_adj) and of the quotient rule (nurbs_param_adj)
//      per Gauss point, then the transposed sum factorisation x -> y -> z.
// ============================================================================
// Element e's Gauss-point fields for the residual and the density integral
// (256 threads): the control points and 1-D tables to LDS, then by sum
// factorisation the physical u-jets of every history level
// (sJet[q][h][k][a], par_periodic added) and the weight, G^-1, d2x/dxi2
// and W jets (sGeo[q][38]).  Ends with a block barrier.
// z and y passes of the Gauss-point field sum factorisation:
//   sA[f][bx*3+by][qz][dz] = sum_bz c_f(b) tz[qz][dz][bz]
//   sB[f][bx][qy][qz][s6]  = sum_by sA[f][bx*3+by][qz][dz(s6)] ty[qy][dy(s6)][by]
// with c = w, w X, w U (per field f) and the y-z jet slots s6 = (00, 10, 01,
// 20, 11, 02) in (dy, dz).  One thread per (f, bx*3+by) for z and per (f, bx,
// qz) for y: the inputs of a thread are loaded once and all its outputs
// (12 and 24) follow, with the reference's summation order per output.
// The NF fields are F0 .. F0 + NF - 1 of that list, stored from field 0 on.
template <int NF, int STR, int F0 = 0>
__device__ __forceinline__ void sf_zy(const double (*sT)[4][3][3], const double *sCP, double *sA, double *sB)
{
    static_assert(NF * 12 <= 256, "one thread per (field, bx, qz) in a 256-thread workgroup");
    const int tid = threadIdx.x;
    if (tid < NF * 9) {
        const int f = tid / 9, bxy = tid - 9 * f, fg = F0 + f;
        double cf[3];
#pragma unroll
        for (int bz = 0; bz < 3; bz++) {
            const double *c = sCP + STR * (3 * bxy + bz);
            cf[bz] = (fg == 0) ? c[3] : c[3] * ((fg < 4) ? c[fg - 1] : c[fg]);
        }
        double *o = sA + 108 * f + 12 * bxy;
#pragma unroll
        for (int qz = 0; qz < 4; qz++)
#pragma unroll
            for (int dz = 0; dz < 3; dz++) {
                const double *tz = &sT[2][qz][dz][0];
                double acc = 0.0;
#pragma unroll
                for (int bz = 0; bz < 3; bz++) acc += cf[bz] * tz[bz];
                o[3 * qz + dz] = acc;
            }
    }
    __syncthreads();
    if (tid < NF * 12) {
        const int f = tid / 12, r = tid - 12 * f, bx = r >> 2, qz = r & 3;
        const double *a = sA + 108 * f + 36 * bx + 3 * qz;   // a[12 by + dz]
        double av[3][3];
#pragma unroll
        for (int by = 0; by < 3; by++)
#pragma unroll
            for (int dz = 0; dz < 3; dz++) av[by][dz] = a[12 * by + dz];
        double *o = sB + 288 * f + 96 * bx + 6 * qz;         // o[24 qy + s6]
#pragma unroll
        for (int qy = 0; qy < 4; qy++)
#pragma unroll
            for (int s6 = 0; s6 < 6; s6++) {
                const int dy = (s6 == 1 || s6 == 4) ? 1 : (s6 == 3 ? 2 : 0);
                const int dz = (s6 == 2 || s6 == 4) ? 1 : (s6 == 5 ? 2 : 0);
                const double *ty = &sT[1][qy][dy][0];
                o[24 * qy + s6] = av[0][dz] * ty[0] + av[1][dz] * ty[1] + av[2][dz] * ty[2];
            }
    }
    __syncthreads();
}

// fields per sum-factorisation pass of gp_fields: static all 7 at once;
// transient the 9 history fields after the 4 geometry fields
template <int NH>
constexpr int gp_fields_per_pass() { return NH == 1 ? 4 + 3 * NH : 3 * NH; }

template <int NH>
__device__ __forceinline__ void gp_fields(const MeshK &M, const double *__restrict__ Uh, int64_t ndof, int e,
                                          double (*sT)[4][3][3], double *sCP, double *sWork, double *sJet, double *sGeo)
{
    constexpr int STR = 4 + 3 * NH;
    constexpr int NF = 4 + 3 * NH;
    constexpr int NFP = gp_fields_per_pass<NH>();
    constexpr int GEO_R = 38;
    const int tid = threadIdx.x;
    const int nz = M.nelem[2], ny = M.nelem[1];
    const int ex = e / (ny * nz), ey = (e / nz) % ny, ez = e % nz;
    const int ed[3] = {ex, ey, ez};
    if (tid < 27) {
        int bx = tid / 9, by = (tid / 3) % 3, bz = tid % 3;
        int ax = ex + bx, ay = ey + by, az = ez + bz;
        int64_t id = ((int64_t)M.na[1] * M.na[2]) * ax + (int64_t)M.na[2] * ay + az;
        sCP[STR * tid + 0] = M.X[3 * id + 0];
        sCP[STR * tid + 1] = M.X[3 * id + 1];
        sCP[STR * tid + 2] = M.X[3 * id + 2];
        sCP[STR * tid + 3] = M.w[id];
        int64_t oc = ((int64_t)M.nb[1] * M.nb[2]) * own1(ax, M.nb[0], M.periodic[0])
                   + (int64_t)M.nb[2] * own1(ay, M.nb[1], M.periodic[1]) + own1(az, M.nb[2], M.periodic[2]);
#pragma unroll
        for (int h = 0; h < NH; h++)
#pragma unroll
            for (int k = 0; k < 3; k++) sCP[STR * tid + 4 + 3 * h + k] = Uh[h * ndof + 3 * oc + k];
    } else if (tid >= 64 && tid < 64 + 12) {
        int d = (tid - 64) / 4, q = (tid - 64) % 4;
        const double *kV = M.knot[d] + ed[d];
        double c = M.gq[q];
        double x = (-c + 1.0) * kV[2] + c * kV[3];
        bspline1d(kV, x, sT[d][q][0], sT[d][q][1], sT[d][q][2]);
    }
    __syncthreads();

    // fields at the Gauss points (z, then y, then x).  Transient: two passes,
    // geometry (w, w X) then the history levels (w U), so that the buffers
    // hold 9 fields instead of 13
    {
        double *sA = sWork, *sB = sWork + NFP * 108;
        sf_zy<(NH == 1 ? NF : 4), STR>(sT, sCP, sA, sB);

        const int q = tid >> 2, part = tid & 3, lq = (tid & 63) & ~3;
        const int qd[3] = {q >> 4, (q >> 2) & 3, q & 3};
        const double *tx = &sT[0][qd[0]][0][0];
        const int boff = (4 * qd[1] + qd[2]) * 6;
        double W[10], Jg[3][10];
        sf_jet(sB + boff, tx, W);
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int f = (part == 0) ? 1 + j : 4 + 3 * j + part - 1;
            if (part == 0 || (NH == 1 && j < NH)) {
                double Pj[10];
                sf_jet(sB + 288 * f + boff, tx, Pj);
                nurbs_param(Pj, 1.0, W, Jg[j]);
            } else {
#pragma unroll
                for (int a = 0; a < 10; a++) Jg[j][a] = (a >= 1 && a <= 3 && a - 1 == j) ? 1.0 : 0.0;
            }
        }
        double G[9], Gi[9], X2[18];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) G[3 * i + j] = Jg[i][1 + j];
        const double detG = inv3(G, Gi);
#pragma unroll
        for (int i = 0; i < 9; i++) Gi[i] = __shfl(Gi[i], lq, 64);
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int s2 = 0; s2 < 6; s2++) X2[6 * d + s2] = __shfl(Jg[d][4 + s2], lq, 64);
        if (NH != 1) {
            __syncthreads();
            sf_zy<3 * NH, STR, 4>(sT, sCP, sA, sB);
            if (part > 0) {
#pragma unroll
                for (int h = 0; h < NH; h++) {
                    double Pj[10];
                    sf_jet(sB + 288 * (3 * h + part - 1) + boff, tx, Pj);
                    nurbs_param(Pj, 1.0, W, Jg[h]);
                }
            }
        }
        if (part > 0) {
            const int k = part - 1;
            double xq[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double *kV = M.knot[d] + ed[d];
                const double c = M.gq[qd[d]];
                xq[d] = (-c + 1.0) * kV[2] + c * kV[3];
            }
#pragma unroll
            for (int h = 0; h < NH; h++) {
                phys_jet(Jg[h], Gi, X2);
                add_periodic1(M, xq, k, Jg[h]);
#pragma unroll
                for (int a = 0; a < 10; a++) sJet[q * 30 * NH + 30 * h + 10 * k + a] = Jg[h][a];
            }
        }
        // sGeo may share memory with the field buffers read above
        __syncthreads();
        if (part == 0) {
            double *g = sGeo + q * GEO_R;
            double dX[3];
#pragma unroll
            for (int d = 0; d < 3; d++) dX[d] = M.knot[d][ed[d] + 3] - M.knot[d][ed[d] + 2];
            g[0] = M.gw[qd[0]] * M.gw[qd[1]] * M.gw[qd[2]] * dX[0] * dX[1] * dX[2] * detG;
#pragma unroll
            for (int i = 0; i < 9; i++) g[1 + i] = Gi[i];
#pragma unroll
            for (int i = 0; i < 18; i++) g[10 + i] = X2[i];
#pragma unroll
            for (int i = 0; i < 10; i++) g[28 + i] = W[i];
        }
    }
    __syncthreads();
}

template <int NH>
__global__ __launch_bounds__(256, (NH == 1 ? 4 : 2)) void k_residual_elem(MeshK M, PhysK P, const double *__restrict__ Uh, int64_t ndof,
                                                       double *__restrict__ Re, int nel)
{
    constexpr int STR = 4 + 3 * NH;            // X(3) w(1) U(3*NH) per control point
    constexpr int WA = gp_fields_per_pass<NH>() * (108 + 288);   // stage-A buffers
    constexpr int GEO_R = 38;                  // wgt, Gi[9], X2[18], W[10]
    constexpr int WG = 64 * GEO_R;             // the Gauss-point geometry (after stage A)
    constexpr int WT = 864 + 324;              // x^T and y^T buffers (after the adjoint; rho lives in sJet)
    constexpr int W1 = WA > WG ? WA : WG;
    __shared__ double sT[3][4][3][3];
    __shared__ double sCP[27 * STR];
    __shared__ double sWork[W1 > WT ? W1 : WT];   // stage A, then sGeo, then x^T / y^T
    __shared__ double sJet[64 * 30 * NH];      // physical u-jets [q][h][k][a]; then r_q / rho_q [q][k][a]
    double *sGeo = sWork;

    const int tid = threadIdx.x;
    const int e = blockIdx.x;
    // ---- A: fields at the Gauss points
    gp_fields<NH>(M, Uh, ndof, e, sT, sCP, sWork, sJet, sGeo);

    // ---- P: pointwise residual jets r_q -> sJet[q * 30 * NH + 10 k + a]
    if (NH == 1) {
        if ((tid >> 6) == (e & 3)) {                            // one wave, lane = Gauss point
            const int q = tid & 63;
            double u[30], r[30];
#pragma unroll
            for (int a = 0; a < 30; a++) u[a] = sJet[q * 30 + a];
            phys_residual_static(P, u, r);
#pragma unroll
            for (int a = 0; a < 30; a++) sJet[q * 30 + a] = r[a];
        }
    } else {
        // lanes 4q..4q+3 evaluate two rule nodes each at s_j: m0 + s_j d with
        // m0 = (u^1 + u^0)/2, d = (u - u^0)/2, then sum over the four lanes
        const int q = tid >> 2, part = tid & 3;
        const double *J = sJet + q * 30 * NH;
        double r[30];
#pragma unroll
        for (int a = 0; a < 30; a++) r[a] = 0.0;
#pragma unroll 1
        for (int gg = 0; gg < 2; gg++) {
            const int g = 2 * part + gg;
            double us[30];
#pragma unroll
            for (int a = 0; a < 30; a++) us[a] = 0.5 * (J[30 + a] + J[60 + a]) + P.s[g] * (0.5 * (J[a] - J[60 + a]));
            State st;
            phys_state(us, P, st);
            phys_residual_acc(st, P.cr[g], r);
        }
#pragma unroll
        for (int a = 0; a < 30; a++) {
            r[a] += __shfl_xor(r[a], 1, 64);
            r[a] += __shfl_xor(r[a], 2, 64);
        }
        if (part == 0) {
#pragma unroll
            for (int k = 0; k < 3; k++)
                r[10 * k] = (P.cii * (J[10 * k] - J[60 + 10 * k])) / (2. * P.dt)
                          + (P.rho * (J[10 * k] + J[60 + 10 * k] - 2 * J[30 + 10 * k])) / (P.dt * P.dt);
#pragma unroll
            for (int a = 0; a < 30; a++) sJet[q * 30 * NH + a] = r[a];
        }
    }
    __syncthreads();

    // ---- T: per (Gauss point, component) adjoint of the basis maps, weighted
    if ((tid & 3) != 0) {
        const int q = tid >> 2, k = (tid & 3) - 1;
        const double *g = sGeo + q * GEO_R;
        double o[10];
        double *rq = sJet + q * 30 * NH + 10 * k;
#pragma unroll
        for (int a = 0; a < 10; a++) o[a] = g[0] * rq[a];
        phys_jet_adj(o, g + 1, g + 10);
        nurbs_param_adj(o, g + 28);
#pragma unroll
        for (int a = 0; a < 10; a++) rq[a] = o[a];
    }
    __syncthreads();
    //      x^T: C1[k][bx][qy][qz][s6] = sum_qx sum_{a: S6(a) = s6} rho[q][k][a] tx[qx][DX(a)][bx]
    double *C1 = sWork, *C2 = sWork + 864;
    for (int i = tid; i < 864; i += 256) {   // s6 slowest: the switch below is mostly wave-uniform
        const int s6 = i / 144, t = i - 144 * s6, qz = t & 3, qy = (t >> 2) & 3, bx = (t >> 4) % 3, k = t / 48;
        double acc = 0.0;
#pragma unroll
        for (int qx = 0; qx < 4; qx++) {
            const double *rq = sJet + (16 * qx + 4 * qy + qz) * 30 * NH + 10 * k;
            const double *tx = &sT[0][qx][0][0];   // tx[3 * dx + bx]
            switch (s6) {
                case 0: acc += rq[0] * tx[bx] + rq[1] * tx[3 + bx] + rq[4] * tx[6 + bx]; break;
                case 1: acc += rq[2] * tx[bx] + rq[5] * tx[3 + bx]; break;
                case 2: acc += rq[3] * tx[bx] + rq[6] * tx[3 + bx]; break;
                case 3: acc += rq[7] * tx[bx]; break;
                case 4: acc += rq[8] * tx[bx]; break;
                default: acc += rq[9] * tx[bx]; break;
            }
        }
        C1[6 * t + s6] = acc;   // [k][bx][qy][qz][s6]
    }
    __syncthreads();
    //      y^T: C2[k][bx][by][qz][dz] = sum_qy sum_{s6: dz(s6) = dz} C1[..][s6] ty[qy][dy(s6)][by]
    for (int i = tid; i < 324; i += 256) {   // dz slowest (mostly wave-uniform branches)
        const int dz = i / 108, t = i - 108 * dz, qz = t & 3, u = t >> 2, by = u % 3, bx = (u / 3) % 3, k = u / 9;
        double acc = 0.0;
#pragma unroll
        for (int qy = 0; qy < 4; qy++) {
            const double *c = C1 + (((k * 3 + bx) * 4 + qy) * 4 + qz) * 6;
            const double *ty = &sT[1][qy][0][0];   // ty[3 * dy + by]
            if (dz == 0) acc += c[0] * ty[by] + c[1] * ty[3 + by] + c[3] * ty[6 + by];
            else if (dz == 1) acc += c[2] * ty[by] + c[4] * ty[3 + by];
            else acc += c[5] * ty[by];
        }
        C2[3 * t + dz] = acc;   // [k][bx][by][qz][dz]
    }
    __syncthreads();
    //      z^T: R_e[3 b + k] = w_b sum_qz sum_dz C2[k][bx][by][qz][dz] tz[qz][dz][bz]
    if (tid < 81) {
        const int k = tid % 3, b = tid / 3, bx = b / 9, by = (b / 3) % 3, bz = b % 3;
        const double *c = C2 + ((k * 3 + bx) * 3 + by) * 12;
        double acc = 0.0;
#pragma unroll
        for (int qz = 0; qz < 4; qz++)
#pragma unroll
            for (int dz = 0; dz < 3; dz++) acc += c[3 * qz + dz] * sT[2][qz][dz][bz];
        Re[(int64_t)e * 81 + tid] = sCP[STR * b + 3] * acc;
    }
}

// ============================================================================
// density integral (iga_intdensity volume part, src/postcore_iga.c:84-234):
// per element sum_q w_q hh(u_q) -> Hd[e][ND]; lane = Gauss point, butterfly
// sums over the 64 lanes
// ============================================================================
template <int NH>
__global__ __launch_bounds__(256, (NH == 1 ? 1 : 2)) void k_density_elem(MeshK M, PhysK P, const double *__restrict__ Uh, int64_t ndof,
                                                      double *__restrict__ Hd, int nel)
{
    constexpr int STR = 4 + 3 * NH;
    constexpr int GEO_R = 38;
    constexpr int ND = (NH == 1) ? ND_STATIC : ND_TRANSIENT;
    __shared__ double sT[3][4][3][3];
    __shared__ double sCP[27 * STR];
    constexpr int WA = gp_fields_per_pass<NH>() * (108 + 288), WG = 64 * GEO_R;
    __shared__ double sWork[WA > WG ? WA : WG];   // stage A, then sGeo
    __shared__ double sJet[64 * 30 * NH];
    double *sGeo = sWork;
    const int e = blockIdx.x;
    gp_fields<NH>(M, Uh, ndof, e, sT, sCP, sWork, sJet, sGeo);
    if (threadIdx.x < 64) {
        const int q = threadIdx.x;
        double hh[ND];
        if (NH == 1) phys_density_static(sJet + q * 30, P, hh);
        else phys_density_transient(sJet + q * 90, P, hh);
        const double w = sGeo[q * GEO_R];
#pragma unroll
        for (int i = 0; i < ND; i++) {
            double v = w * hh[i];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            if (q == 0) Hd[(int64_t)e * ND + i] = v;
        }
    }
}

// H[i] = sum over elements of Hd[e][i], in a fixed order (one block per density)
__global__ __launch_bounds__(256) void k_density_reduce(int nel, int nd, const double *__restrict__ Hd, double *__restrict__ H)
{
    __shared__ double red[256];
    const int i = blockIdx.x;
    double s = 0.0;
    for (int e = threadIdx.x; e < nel; e += 256) s += Hd[(int64_t)e * nd + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) H[i] = red[0];
}

// contributing (element, local index) pairs of owned control point a along one
// direction, sorted by element (the reference accumulates in element order)
__device__ __forceinline__ int elem_pairs(int a, int nelem, int nb, int periodic, int *es, int *ls)
{
    int cnt = 0;
    if (!periodic) {
        for (int l = 2; l >= 0; l--) {
            int e = a - l;
            if (e >= 0 && e < nelem) { es[cnt] = e; ls[cnt] = l; cnt++; }
        }
        return cnt;
    }
    for (int l = 2; l >= 0; l--) {
        int e = a - l;
        if (e < 0) e += nb;
        es[cnt] = e; ls[cnt] = l; cnt++;
    }
    // insertion sort by element
    for (int i = 1; i < cnt; i++)
        for (int j = i; j > 0 && es[j] < es[j - 1]; j--) {
            int te = es[j]; es[j] = es[j - 1]; es[j - 1] = te;
            int tl = ls[j]; ls[j] = ls[j - 1]; ls[j - 1] = tl;
        }
    return cnt;
}

__global__ __launch_bounds__(256) void k_residual_gather(MeshK M, const double *__restrict__ Re, const double *__restrict__ Uh,
                                                         const double *__restrict__ Fneu, const int32_t *__restrict__ mask,
                                                         const double *__restrict__ g, double *__restrict__ R, int64_t ncp)
{
    int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= ncp) return;
    int nb1 = M.nb[1], nb2 = M.nb[2];
    int az = (int)(a % nb2), ay = (int)((a / nb2) % nb1), ax = (int)(a / ((int64_t)nb1 * nb2));
    int exs[3], lxs[3], eys[3], lys[3], ezs[3], lzs[3];
    int nx = elem_pairs(ax, M.nelem[0], M.nb[0], M.periodic[0], exs, lxs);
    int ny = elem_pairs(ay, M.nelem[1], M.nb[1], M.periodic[1], eys, lys);
    int nz = elem_pairs(az, M.nelem[2], M.nb[2], M.periodic[2], ezs, lzs);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int x = 0; x < nx; x++)
        for (int y = 0; y < ny; y++)
            for (int z = 0; z < nz; z++) {
                int64_t e = ((int64_t)exs[x] * M.nelem[1] + eys[y]) * M.nelem[2] + ezs[z];
                int lb = 9 * lxs[x] + 3 * lys[y] + lzs[z];
                const double *r = Re + e * 81 + 3 * lb;
                s0 += r[0]; s1 += r[1]; s2 += r[2];
            }
    int64_t d = 3 * a;
    R[d + 0] = mask[d + 0] ? Uh[d + 0] - g[d + 0] : s0 + Fneu[d + 0];
    R[d + 1] = mask[d + 1] ? Uh[d + 1] - g[d + 1] : s1 + Fneu[d + 1];
    R[d + 2] = mask[d + 2] ? Uh[d + 2] - g[d + 2] : s2 + Fneu[d + 2];
}

// ============================================================================
// tangent: one workgroup (4 waves) per element.
//   stage A  (once)      geometry and u-jets of the 64 Gauss points, 4 lanes
//                        per point splitting the 27 basis functions
//   stage B  (16 chunks of QC = 4 Gauss points)
//     B1  physical basis jets R_q[10][27] -> LDS
//     B2  pointwise tangent D_q = w_q C_q (30x30, the blocks i <= j) -> LDS
//     B3  K^{ij} += sum_{q,a} R_q[a][b] * (sum_b' D_q[(i,a),(j,b')] R_q[b'][c])
//         for the six blocks i <= j (K_e is symmetric; K^{ji} = K^{ij}^T is
//         written by mirroring) on V_MFMA_{F64,F32}_16X16X4: M = b, N = c
//         (27 -> 32), K = (q,a).  Wave w owns column tile w>>1 and blocks
//         3(w&1)..3(w&1)+2 for both row tiles, so each B operand it forms on
//         the VALU from D and R feeds two matrix-core products.  For the
//         static physics the value slot a = 0 of D is zero and is skipped.
// ============================================================================
constexpr int QC = 4;
constexpr int GEO = 38;   // per Gauss point: wgt, Gi[9], X2[18], W[10]
// Operands of the 16x16x4 products: lane l supplies A[l&15][l>>4] and
// B[l>>4][l&15].  Result register r of lane l holds D[row(l>>4, r)][l&15];
// the row map of each instruction is pinned by igap_gpu_selftest_mfma.
template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef double4_t acc_t;
    static __device__ __forceinline__ acc_t run(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int kr, int r) { return 4 * r + kr; }
};
template <> struct Mfma<float> {
    typedef float4_t acc_t;
    static __device__ __forceinline__ acc_t run(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int kr, int r) { return 4 * kr + r; }
};

__constant__ const int kBlkI[6] = {0, 0, 0, 1, 1, 2};
__constant__ const int kBlkJ[6] = {0, 1, 2, 1, 2, 2};


// Stage A of the element tangent (256 threads): the Gauss-point fields by
// sum factorisation from the control points in sCP and the 1-D tables in sT.
// Static physics: the strain-measure state of all 64 points -> sSt[q][STW];
// transient: the physical u-jets of Um and Ud -> sJet[q][h][k][a].  Per point
// sGeo[q][GEOK] = wgt (= Gauss weight x span lengths x det G), then, when
// GEOK allows, G^-1 (9), d2x/dxi2 (18) and the weight jet W (10).  sWork
// holds the field buffers.  Ends with a block barrier.
template <int NH, int GEOK>
__device__ __forceinline__ void tangent_fields(const MeshK &M, const PhysK &P, const int *ed, double (*sT)[4][3][3],
                                               const double *sCP, double *sWork, double *sGeo, double *sJet, double *sSt)
{
    constexpr int NL = (NH == 1) ? 1 : 2;
    constexpr int STR = 4 + 3 * NL;
    constexpr int NF = 4 + 3 * NL;             // fields w, w X (3), w U (3 per held level)
    constexpr int NFP = (NH == 1) ? NF : 6;    // fields per sum-factorisation pass
    const int tid = threadIdx.x;
    // ---- stage A: fields at the 64 Gauss points by sum factorisation.
    //      The fields P_f = sum_b c_{b,f} N_b with c = w, w X (3), w U (3 per
    //      level) are contracted one direction at a time (z, then y, then x:
    //      27 -> 108 -> 288 -> 640 values per field); the rational jets follow
    //      by the quotient rule on P_f / W and the physical jets by the chain
    //      rule, both linear in the basis, so the result is the reference's
    //      sum over b of c_b R_b in another order of summation.
    {
        double *sA = sWork, *sB = sA + NFP * 108;
        sf_zy<(NH == 1 ? NF : 4), STR>(sT, sCP, sA, sB);

        const int q = tid >> 2, part = tid & 3, lq = (tid & 63) & ~3;
        const int qd[3] = {q >> 4, (q >> 2) & 3, q & 3};
        const double *tx = &sT[0][qd[0]][0][0];
        const int boff = (4 * qd[1] + qd[2]) * 6;
        double W[10], Jg[3][10];
        sf_jet(sB + boff, tx, W);
        // part 0: the three geometry fields; part 1 + k: displacement
        // component k of every held level (transient: second pass below)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int f = (part == 0) ? 1 + j : 4 + 3 * j + part - 1;
            if (part == 0 || (NH == 1 && j < NL)) {
                double Pj[10];
                sf_jet(sB + 288 * f + boff, tx, Pj);
                nurbs_param(Pj, 1.0, W, Jg[j]);
            } else {
#pragma unroll
                for (int a = 0; a < 10; a++) Jg[j][a] = (a >= 1 && a <= 3 && a - 1 == j) ? 1.0 : 0.0;
            }
        }
        double G[9], Gi[9], X2[18];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) G[3 * i + j] = Jg[i][1 + j];
        const double detG = inv3(G, Gi);
#pragma unroll
        for (int i = 0; i < 9; i++) Gi[i] = __shfl(Gi[i], lq, 64);
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int s2 = 0; s2 < 6; s2++) X2[6 * d + s2] = __shfl(Jg[d][4 + s2], lq, 64);
        double xq[3], dX[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double *kV = M.knot[d] + ed[d];
            const double c = M.gq[qd[d]];
            xq[d] = (-c + 1.0) * kV[2] + c * kV[3];
            dX[d] = kV[3] - kV[2];
        }
        if (NH != 1) {
            // second pass: the fields w Um (3) and w Ud (3)
            __syncthreads();
            sf_zy<6, STR, 4>(sT, sCP, sA, sB);
            if (part > 0) {
#pragma unroll
                for (int h = 0; h < NL; h++) {
                    double Pj[10];
                    sf_jet(sB + 288 * (3 * h + part - 1) + boff, tx, Pj);
                    nurbs_param(Pj, 1.0, W, Jg[h]);
                }
            }
        }
        if (part > 0) {
            const int k = part - 1;
#pragma unroll
            for (int h = 0; h < NL; h++) {
                phys_jet(Jg[h], Gi, X2);
                // the periodic macro-field shifts every level alike: Um takes
                // it, the difference Ud does not
                if (h == 0) add_periodic1(M, xq, k, Jg[h]);
                if (NH != 1) {
#pragma unroll
                    for (int a = 0; a < 10; a++) sJet[q * 30 * NL + 30 * h + 10 * k + a] = Jg[h][a];
                }
            }
        }
        if (NH == 1) {
            double u[30];
#pragma unroll
            for (int k = 0; k < 3; k++)
#pragma unroll
                for (int a = 0; a < 10; a++) u[10 * k + a] = __shfl(Jg[0][a], lq + 1 + k, 64);
            if (part == 0) {
                State st;
                phys_state(u, P, st);
                double *sp = sSt + q * STW;
#pragma unroll
                for (int i = 0; i < 9; i++) sp[ST_F + i] = st.F[i];
#pragma unroll
                for (int i = 0; i < 27; i++) sp[ST_H + i] = st.H[i];
#pragma unroll
                for (int i = 0; i < 9; i++) { sp[ST_S + i] = st.S[i]; sp[ST_G + i] = st.G[i]; sp[ST_N + i] = st.N[i]; }
            }
        }
        if (part == 0) {
            double *g = sGeo + q * GEOK;
            g[0] = M.gw[qd[0]] * M.gw[qd[1]] * M.gw[qd[2]] * dX[0] * dX[1] * dX[2] * detG;
            if (GEOK > 1) {
#pragma unroll
                for (int i = 0; i < 9; i++) g[1 + i] = Gi[i];
#pragma unroll
                for (int i = 0; i < 18; i++) g[10 + i] = X2[i];
            }
            if (GEOK > 28) {
#pragma unroll
                for (int i = 0; i < 10; i++) g[28 + i] = W[i];
            }
        }
    }
    __syncthreads();
}

// Where k_tangent_elem<..., DIRECT = true> adds its element matrix: the
// elements of one of the 27 colours (ex, ey, ez) = (cx, cy, cz) + 3 (i, j, k)
// of a non-periodic patch -- no two of them share a control point, so none
// share a CSR row -- straight into the CSR values, skipping the entries of
// Dirichlet rows and columns (MatZeroRowsColumns, whose unit diagonal
// k_dirichlet_diag writes afterwards).
struct Scatter {
    const int64_t *rowptr;
    const int32_t *mask;
    int c[3];    // colour
    int nc[3];   // elements of the colour per direction
};

// CSR position of column 3C + j in the row of control point B that starts at
// row_start, for control points B, C within +-2 of each other on a
// non-periodic patch: the column block's slot in the row's sorted stencil
__device__ __forceinline__ int64_t stencil_pos(const MeshK &M, int64_t row_start, int Bx, int By, int Bz, int Cx, int Cy, int Cz,
                                               int j)
{
    const int lx = Bx - 2 < 0 ? 0 : Bx - 2, ly = By - 2 < 0 ? 0 : By - 2, lz = Bz - 2 < 0 ? 0 : Bz - 2;
    const int hy = By + 2 > M.nb[1] - 1 ? M.nb[1] - 1 : By + 2, hz = Bz + 2 > M.nb[2] - 1 ? M.nb[2] - 1 : Bz + 2;
    const int nsy = hy - ly + 1, nsz = hz - lz + 1;
    return row_start + 3 * (((Cx - lx) * nsy + (Cy - ly)) * nsz + (Cz - lz)) + j;
}

// the same position for row 3B + i looked up in rowptr
__device__ __forceinline__ int64_t csr_pos(const MeshK &M, const int64_t *__restrict__ rowptr, int Bx, int By, int Bz, int i,
                                           int Cx, int Cy, int Cz, int j)
{
    const int64_t row = 3 * (((int64_t)M.nb[1] * M.nb[2]) * Bx + (int64_t)M.nb[2] * By + Bz) + i;
    return stencil_pos(M, rowptr[row], Bx, By, Bz, Cx, Cy, Cz, j);
}

template <typename T, int NH, bool DIRECT>
__global__ __launch_bounds__(256, 2) void k_tangent_elem(MeshK M, PhysK P, const double *__restrict__ Uh, int64_t ndof,
                                                         T *__restrict__ Ke, int nel, Scatter S)
{
    // field levels held per control point: static U; transient the two
    // combinations the rule evaluates, Um = (U^1 + U^0)/2 and Ud = (U - U^0)/2
    // (rule node j: Um + s_j Ud)
    constexpr int NL = (NH == 1) ? 1 : 2;
    constexpr int STR = 4 + 3 * NL;
    constexpr int NNODE = (NH == 1) ? 1 : 8;
    constexpr int A0 = (NH == 1) ? 1 : 0;     // first jet slot in the contraction
    constexpr int NA = 10 - A0;               // jet slots per Gauss point in K
    // rows of D are padded to whole 16-byte LDS vectors so that B3 reads them
    // with 128-bit loads at compile-time offsets
    constexpr int VN = 16 / (int)sizeof(T);
    constexpr int RS = (NA + VN - 1) / VN * VN;   // row stride of D
    constexpr int NN = NA * RS;                   // stride of one (i,j) block of D
    static_assert(QC == 4, "the MFMA k-row is the Gauss point of a 4-point chunk");
    // static: the strain-measure state of all 64 Gauss points is built once in
    // stage A; transient: the u-jets of Um and Ud are kept and the eight
    // rule-node states are built per chunk.  Two workgroups share a CU, so
    // the transient keeps its LDS under 80 KB: the field sum factorisation
    // runs in two passes (geometry, then Um/Ud) and the weight jet W is
    // rebuilt per chunk from the basis instead of being stored.
    constexpr int SST = (NH == 1) ? 64 * STW : QC * MW;   // static: states; transient: moment states
    constexpr int GEOK = (NH == 1) ? GEO : GEO - 10;     // per Gauss point: wgt, Gi[9], X2[18] (+ W[10] static)
    __shared__ double sT[3][4][3][3];          // 1-D tables [d][q_d][order][i]
    __shared__ double sCP[27 * STR];
    __shared__ double sGeo[64 * GEOK];
    __shared__ double sJet[(NH == 1) ? 1 : 64 * 30 * NL];
    __shared__ double sSt[SST];
    // stage A's sum-factorisation buffers share memory with stage B's R and D
    constexpr int NF = 4 + 3 * NL;             // fields w, w X (3), w U (3 per held level)
    constexpr int NFP = (NH == 1) ? NF : 6;    // fields per sum-factorisation pass
    constexpr int WORK_A = NFP * (108 + 288) * (int)sizeof(double);
    constexpr int WORK_B = QC * (10 * 32 + 6 * NN) * (int)sizeof(T);
    __shared__ __attribute__((aligned(16))) char sWork[WORK_A > WORK_B ? WORK_A : WORK_B];
    T(*sR)[10][32] = reinterpret_cast<T(*)[10][32]>(sWork);
    T(*sD)[6 * NN] = reinterpret_cast<T(*)[6 * NN]>(sWork + QC * 10 * 32 * sizeof(T));   // blocks i <= j, slots A0..9

    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int e = blockIdx.x;
    int ex, ey, ez;
    if (DIRECT) {
        // element (i, j, k) of the colour's lattice
        const int k = e % S.nc[2], j = (e / S.nc[2]) % S.nc[1], i = e / (S.nc[1] * S.nc[2]);
        ex = S.c[0] + 3 * i;
        ey = S.c[1] + 3 * j;
        ez = S.c[2] + 3 * k;
    } else {
        const int nz = M.nelem[2], ny = M.nelem[1];
        ex = e / (ny * nz);
        ey = (e / nz) % ny;
        ez = e % nz;
    }
    const int ed[3] = {ex, ey, ez};

    // ---- element data, 1-D tables
    if (tid < 27) {
        int bx = tid / 9, by = (tid / 3) % 3, bz = tid % 3;
        int ax = ex + bx, ay = ey + by, az = ez + bz;
        int64_t id = ((int64_t)M.na[1] * M.na[2]) * ax + (int64_t)M.na[2] * ay + az;
        sCP[STR * tid + 0] = M.X[3 * id + 0];
        sCP[STR * tid + 1] = M.X[3 * id + 1];
        sCP[STR * tid + 2] = M.X[3 * id + 2];
        sCP[STR * tid + 3] = M.w[id];
        int64_t oc = ((int64_t)M.nb[1] * M.nb[2]) * own1(ax, M.nb[0], M.periodic[0])
                   + (int64_t)M.nb[2] * own1(ay, M.nb[1], M.periodic[1]) + own1(az, M.nb[2], M.periodic[2]);
        if (NH == 1) {
#pragma unroll
            for (int k = 0; k < 3; k++) sCP[STR * tid + 4 + k] = Uh[3 * oc + k];
        } else {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double u0 = Uh[3 * oc + k], u1 = Uh[ndof + 3 * oc + k], u2 = Uh[2 * ndof + 3 * oc + k];
                sCP[STR * tid + 4 + k] = 0.5 * (u1 + u2);
                sCP[STR * tid + 7 + k] = 0.5 * (u0 - u2);
            }
        }
    } else if (tid >= 64 && tid < 64 + 12) {
        int d = (tid - 64) / 4, q = (tid - 64) % 4;
        const double *kV = M.knot[d] + ed[d];
        double c = M.gq[q];
        double x = (-c + 1.0) * kV[2] + c * kV[3];
        bspline1d(kV, x, sT[d][q][0], sT[d][q][1], sT[d][q][2]);
    }
    __syncthreads();

    tangent_fields<NH, GEOK>(M, P, ed, sT, sCP, reinterpret_cast<double *>(sWork), sGeo, sJet, sSt);

    typedef typename Mfma<T>::acc_t acc_t;
    acc_t acc[3][2];
#pragma unroll
    for (int t = 0; t < 3; t++) { acc[t][0] = (acc_t){0, 0, 0, 0}; acc[t][1] = (acc_t){0, 0, 0, 0}; }
    const int ctile = wv >> 1, blk0 = 3 * (wv & 1), m = lane & 15, kr = lane >> 4;
    int bi[3], bj[3];
#pragma unroll
    for (int t = 0; t < 3; t++) { bi[t] = kBlkI[blk0 + t]; bj[t] = kBlkJ[blk0 + t]; }

    Mom mu = {0.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0};
    if (NH != 1) {
#pragma unroll
        for (int j = 0; j < NNODE; j++) {
            mu.m0 += P.ct[j];
            mu.m1 += P.ct[j] * P.s[j];
            mu.m2 += P.ct[j] * P.s[j] * P.s[j];
        }
        mu.factor();
    }
    T Rn[10];   // static: basis jets of the next chunk's point, built by lanes 32..58 on even chunks
    for (int chunk = 0; chunk < 64 / QC; chunk++) {
        const int q0 = chunk * QC;
        // ---- B1 (transient: + the rule-node states): wave wv builds the basis
        //      jets of Gauss point q0 + wv.  Static: lanes 32..58 are otherwise
        //      idle here, so on even chunks they build the jets of point
        //      q0 + 4 + wv (the next chunk's) and store them on the odd chunk:
        //      one B1 pass per two chunks.
        if (NH == 1) {
            const int ql = wv;
            if ((chunk & 1) == 0) {
                const int hi = lane >> 5, b = lane & 31, q = q0 + ql + 4 * hi;
                if (b < 27) {
                    const int qd[3] = {q >> 4, (q >> 2) & 3, q & 3};
                    const double *t0 = &sT[0][qd[0]][0][0], *t1 = &sT[1][qd[1]][0][0], *t2 = &sT[2][qd[2]][0][0];
                    const double *g = sGeo + q * GEOK;
                    double N[10], R[10];
                    tensor_jet(t0, t1, t2, b, N);
                    nurbs_param(N, sCP[STR * b + 3], g + 28, R);
                    phys_jet(R, g + 1, g + 10);
                    if (hi == 0) {
#pragma unroll
                        for (int a = 0; a < 10; a++) sR[ql][a][b] = (T)R[a];
                    } else {
#pragma unroll
                        for (int a = 0; a < 10; a++) Rn[a] = (T)R[a];
                    }
                } else if (hi == 0) {
#pragma unroll
                    for (int a = 0; a < 10; a++) sR[ql][a][b] = (T)0;
                }
            } else {
                const int b = lane - 32;
                if (b >= 0 && b < 27) {
#pragma unroll
                    for (int a = 0; a < 10; a++) sR[ql][a][b] = Rn[a];
                } else if (lane >= 27 && lane < 32) {
#pragma unroll
                    for (int a = 0; a < 10; a++) sR[ql][a][lane] = (T)0;
                }
            }
        } else {
            const int ql = wv, q = q0 + ql, b = lane;
            if (b < 32) {
                // basis jets; the weight jet W = sum_b w_b N_b is summed over
                // lanes 0..31 (27..31 add zero)
                const int qd[3] = {q >> 4, (q >> 2) & 3, q & 3};
                const double *t0 = &sT[0][qd[0]][0][0], *t1 = &sT[1][qd[1]][0][0], *t2 = &sT[2][qd[2]][0][0];
                const double *g = sGeo + q * GEOK;
                double N[10], W[10], R[10];
                const double wb = (b < 27) ? sCP[STR * b + 3] : 0.0;
                if (b < 27) {
                    tensor_jet(t0, t1, t2, b, N);
                } else {
#pragma unroll
                    for (int a = 0; a < 10; a++) N[a] = 0.0;
                }
#pragma unroll
                for (int a = 0; a < 10; a++) W[a] = wb * N[a];
#pragma unroll
                for (int o = 1; o < 32; o <<= 1)
#pragma unroll
                    for (int a = 0; a < 10; a++) W[a] += __shfl_xor(W[a], o, 64);
                if (b < 27) {
                    nurbs_param(N, wb, W, R);
                    phys_jet(R, g + 1, g + 10);
#pragma unroll
                    for (int a = 0; a < 10; a++) sR[ql][a][b] = (T)R[a];
                } else {
#pragma unroll
                    for (int a = 0; a < 10; a++) sR[ql][a][b] = (T)0;
                }
            } else if (NH != 1 && b < 32 + NNODE) {
                // rule node j: S, G and N at us_j, weighted and summed over
                // the eight lanes 32..39 into the moment state
                const int j = b - 32;
                const double *u = sJet + q * 30 * NL;
                double us[30];
#pragma unroll
                for (int i = 0; i < 30; i++) us[i] = u[i] + P.s[j] * u[30 + i];
                State st;
                phys_state(us, P, st);
                const double w0 = P.ct[j], w1 = w0 * P.s[j], w2 = w1 * P.s[j];
                // node sums over lanes 32..39, one entry at a time (few live
                // registers); lane 32 stores Sb, Gb and the N sums as L^-1 Nm L^-T
                double *ms = sSt + ql * MW;
                const double c12 = 1.0 / (mu.a11 * mu.a22), c22 = 1.0 / mu.d2;
#pragma unroll
                for (int i = 0; i < 9; i++) {
                    double sv = w0 * st.S[i], gv = w0 * st.G[i];
                    double n0 = w0 * st.N[i], n1 = w1 * st.N[i], n2 = w2 * st.N[i];
#pragma unroll
                    for (int o = 1; o < 8; o <<= 1) {
                        sv += __shfl_xor(sv, o, 64);
                        gv += __shfl_xor(gv, o, 64);
                        n0 += __shfl_xor(n0, o, 64);
                        n1 += __shfl_xor(n1, o, 64);
                        n2 += __shfl_xor(n2, o, 64);
                    }
                    if (j == 0) {
                        ms[MS + i] = sv;
                        ms[MG + i] = gv;
                        ms[MN11 + i] = n0 / mu.m0;
                        ms[MN12 + i] = (n1 - mu.lf * n0) * c12;
                        ms[MN22 + i] = ((mu.lf * n0 - n1) * mu.lf + (n2 - mu.lf * n1)) * c22;
                    }
                }
            } else if (NH != 1 && b >= 40) {
                // F^ and H^ (two components each) of the midpoint m0 = (u1+u0)/2
                // and the increment d = (u-u0)/2: X^_1 = a11 (X0 + lf Xd),
                // X^_2 = a22 Xd, for the 9 entries of F and the 27 of H
                const double *u = sJet + q * 30 * NL;
                double *ms = sSt + ql * MW;
                for (int i = b - 40; i < 36; i += 24) {
                    const bool isF = i < 9;
                    const int r = isF ? i : i - 9;
                    int slot;
                    if (isF) slot = 10 * (r / 3) + 1 + (r % 3);
                    else slot = 10 * (r / 9) + sidx((r % 9) / 3, r % 3);
                    double x0 = u[slot];
                    const double xd = u[30 + slot];
                    if (isF && (r / 3) == (r % 3)) x0 += 1.0;
                    const double x1 = mu.a11 * (x0 + mu.lf * xd), x2 = mu.a22 * xd;
                    if (isF) { ms[MF1 + r] = x1; ms[MF2 + r] = x2; }
                    else { ms[MH1 + r] = x1; ms[MH2 + r] = x2; }
                }
            }
        }
        if (NH != 1) __syncthreads();   // node states are read by the whole wave below
        // ---- B2: D_q = w_q C_q on the blocks i <= j for the chunk's points.
        //      Static: the chunk's 216 rows sorted by kind across the waves
        //      (tasks 0..71 first-gradient rows: point t / 18, block
        //      (t % 18) / 3, row t % 3 + 1; 72..215 second-gradient rows:
        //      point u / 36, block (u % 36) / 6, row u % 6 + 4, u = t - 72),
        //      so that three waves run one branch of the one-hot rows.
        //      Transient: likewise with the value rows first.
        if (NH == 1) {
            if (tid < 216) {
                int ql, p, ar;
                if (tid < 72) {
                    ql = tid / 18;
                    const int r = tid - 18 * ql;
                    p = r / 3;
                    ar = r - 3 * p;
                } else {
                    const int u = tid - 72;
                    ql = u / 36;
                    const int r = u - 36 * ql;
                    p = r / 6;
                    ar = 3 + r - 6 * p;
                }
                const int q = q0 + ql;
                const double wgt = sGeo[q * GEOK];
                double row[NA];
                c_row_onehot<false>(sSt + q * STW, P, kBlkI[p], ar + A0, kBlkJ[p], row);
                T *D = sD[ql] + p * NN + ar * RS;
#pragma unroll
                for (int bb = 0; bb < NA; bb++) D[bb] = (T)(wgt * row[bb]);
            }
        } else if (tid < 240) {
            // the chunk's 240 rows sorted by kind: 0..23 value rows (the
            // inertia alone; point t / 6, block t % 6), 24..95 first-gradient
            // rows, 96..239 second-gradient rows
            int ql, p, ar;
            if (tid < 24) {
                ql = tid / 6;
                p = tid - 6 * ql;
                ar = 0;
            } else if (tid < 96) {
                const int u = tid - 24;
                ql = u / 18;
                const int r = u - 18 * ql;
                p = r / 3;
                ar = 1 + r - 3 * p;
            } else {
                const int u = tid - 96;
                ql = u / 36;
                const int r = u - 36 * ql;
                p = r / 6;
                ar = 4 + r - 6 * p;
            }
            const int q = q0 + ql, k = kBlkI[p], l = kBlkJ[p];
            const double wgt = sGeo[q * GEOK];
            double row[NA];
            row[0] = 0.0;   // value column: the inertia alone
            if (ar == 0) {
#pragma unroll
                for (int bb = 1; bb < NA; bb++) row[bb] = 0.0;
                if (k == l) row[0] = P.cii / (2. * P.dt) + P.rho / (P.dt * P.dt);
            } else {
                c_row_momc_onehot<double, false>(sSt + ql * MW, P, k, ar, l, row + 1);
            }
            T *D = sD[ql] + p * NN + ar * RS;
#pragma unroll
            for (int bb = 0; bb < NA; bb++) D[bb] = (T)(wgt * row[bb]);
        }
        __syncthreads();
        // ---- B3: matrix-core contraction over (q, a) of this chunk.  The
        //      4-deep k-row of each MFMA is the chunk's Gauss point (kr) and
        //      the k-step is the jet slot, so every LDS address is a lane base
        //      plus a compile-time offset and the lane's basis column Rc is
        //      read once per chunk.
        {
            const T *Rq = &sR[kr][A0][0];
            const T *Dq = sD[kr] + blk0 * NN;
            T Rc[NA];
#pragma unroll
            for (int bb = 0; bb < NA; bb++) Rc[bb] = Rq[32 * bb + 16 * ctile + m];
#pragma unroll
            for (int a = 0; a < NA; a++) {
                const T Av0 = Rq[32 * a + m], Av1 = Rq[32 * a + 16 + m];
#pragma unroll
                for (int t = 0; t < 3; t++) {
                    typedef T vec_t __attribute__((ext_vector_type(VN)));
                    const vec_t *dv = reinterpret_cast<const vec_t *>(Dq + t * NN + a * RS);
                    T d[RS];
#pragma unroll
                    for (int v = 0; v < RS / VN; v++) {
                        const vec_t x = dv[v];
#pragma unroll
                        for (int c = 0; c < VN; c++) d[VN * v + c] = x[c];
                    }
                    T y = d[0] * Rc[0];
#pragma unroll
                    for (int bb = 1; bb < NA; bb++) y += d[bb] * Rc[bb];
                    acc[t][0] = Mfma<T>::run(Av0, y, acc[t][0]);
                    acc[t][1] = Mfma<T>::run(Av1, y, acc[t][1]);
                }
            }
        }
        __syncthreads();
    }

    const int c = 16 * ctile + m;
    if (DIRECT) {
        // ---- add (row 3b+i, column 3c+j) and the mirrored entry into the CSR
        //      values, Dirichlet rows and columns left at zero.  The elements
        //      of one colour share no row, so no other workgroup of the launch
        //      touches these entries.  The row starts and masks of the
        //      element's 81 rows are staged in LDS (the geometry buffer is
        //      free after the last chunk); then every load of the lane is
        //      issued before its first store, so the epilogue waits for one
        //      round trip.
        int64_t *sRow = reinterpret_cast<int64_t *>(sGeo);
        int32_t *sMsk = reinterpret_cast<int32_t *>(sGeo + 81);
        if (tid < 81) {
            const int b = tid / 3, i = tid - 3 * (tid / 3);
            const int64_t cp = ((int64_t)M.nb[1] * M.nb[2]) * (ex + b / 9) + (int64_t)M.nb[2] * (ey + (b / 3) % 3) + (ez + b % 3);
            sRow[tid] = S.rowptr[3 * cp + i];
            sMsk[tid] = S.mask[3 * cp + i];
        }
        __syncthreads();
        if (c < 27) {
            const int Cx = ex + c / 9, Cy = ey + (c / 3) % 3, Cz = ez + c % 3;
            T vp[3][8], vq[3][8];
#pragma unroll
            for (int t = 0; t < 3; t++)
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const int b = 16 * (k >> 2) + Mfma<T>::row(kr, k & 3);
                    vp[t][k] = vq[t][k] = (T)0;
                    if (b < 27 && sMsk[3 * c + bj[t]] == 0 && sMsk[3 * b + bi[t]] == 0) {
                        const int Bx = ex + b / 9, By = ey + (b / 3) % 3, Bz = ez + b % 3;
                        vp[t][k] = Ke[stencil_pos(M, sRow[3 * b + bi[t]], Bx, By, Bz, Cx, Cy, Cz, bj[t])];
                        if (bi[t] != bj[t]) vq[t][k] = Ke[stencil_pos(M, sRow[3 * c + bj[t]], Cx, Cy, Cz, Bx, By, Bz, bi[t])];
                    }
                }
#pragma unroll
            for (int t = 0; t < 3; t++)
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const int b = 16 * (k >> 2) + Mfma<T>::row(kr, k & 3);
                    if (b < 27 && sMsk[3 * c + bj[t]] == 0 && sMsk[3 * b + bi[t]] == 0) {
                        const int Bx = ex + b / 9, By = ey + (b / 3) % 3, Bz = ez + b % 3;
                        const T v = acc[t][k >> 2][k & 3];
                        Ke[stencil_pos(M, sRow[3 * b + bi[t]], Bx, By, Bz, Cx, Cy, Cz, bj[t])] = vp[t][k] + v;
                        if (bi[t] != bj[t])
                            Ke[stencil_pos(M, sRow[3 * c + bj[t]], Cx, Cy, Cz, Bx, By, Bz, bi[t])] = vq[t][k] + v;
                    }
                }
        }
        return;
    }
    // ---- write K_e (row 3b+i, column 3c+j) and the mirrored block (3c+j, 3b+i)
    T *out = Ke + (int64_t)e * 6561;
    if (c < 27) {
#pragma unroll
        for (int t = 0; t < 3; t++)
#pragma unroll
            for (int bt = 0; bt < 2; bt++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int b = 16 * bt + Mfma<T>::row(kr, r);
                    if (b < 27) {
                        T v = acc[t][bt][r];
                        out[(3 * b + bi[t]) * 81 + 3 * c + bj[t]] = v;
                        if (bi[t] != bj[t]) out[(3 * c + bj[t]) * 81 + 3 * b + bi[t]] = v;
                    }
                }
    }
}

// ============================================================================
// Tangent on a separable patch (k_tangent_sep, static physics): every
// coordinate of the geometry depends on its own parametric direction only,
// x_d = f_d(xi_d), and all control weights are equal, as on the igabeam
// patches (mesh/igabeam.c).  Then G = diag(f_x', f_y', f_z'), the rational
// basis is the B-spline basis, and the physical jet of basis b at Gauss point
// q is a product of three 1-D tables of physical derivatives,
//   R_q[a][b] = Px[qx][dx(a)][bx] Py[qy][dy(a)][by] Pz[qz][dz(a)][bz],
//   P_d[q][0] = N,  P_d[q][1] = N' / g,  P_d[q][2] = (N'' - (g'/g) N') / g^2
// (g = f_d', g' = f_d'' at the point; the chain rule of src/assmcore_iga.c:
// 179-208 with a diagonal G).  The contraction of Tangent_iga
// (src/assmcore_iga.c:723-749),
//   K^ij[b][c] = sum_q sum_{a,a'} R_q[a][b] D_q[(i,a),(j,a')] R_q[a'][c],
// then factors direction by direction (sum factorisation), with s = (dx, dy)
// the x-y part of a jet slot:
//   F_q[a][(s',cz)]   = sum_dz' D_q[a][(s',dz')] Pz[qz][dz'][cz]         per point, a lane per row a
//   E1[(s,bz),(s',cz)] = sum_qz sum_dz Pz[qz][dz][bz] F_q[(s,dz)][(s',cz)]  per (qx, qy) pencil
//   E2[(dx,by,bz),(dx',cy,cz)] += sum_{dy,dy'} Py[qy][dy][by] Py[qy][dy'][cy] E1[((dx,dy),bz),((dx',dy'),cz)]
//   K[(bx,by,bz),(cx,cy,cz)]   += sum_{dx,dx'} Px[qx][dx][bx] Px[qx][dx'][cx] E2   after the fourth qy
// ~0.7 MFLOP per element on the vector ALUs instead of the 27x27 product per
// point and jet slot.  The displacement jets come from the same tables by sum
// factorisation (no quotient or chain rule left to apply), the pointwise
// tangent rows D_q = w_q C_q are those of k_tangent_elem, and the element
// matrix is written in its layout (row 3b+i, column 3c+j).
// ============================================================================
// Per-element set-up on a separable patch (256 threads, no barrier): the 27
// control displacements -> sU[27][3]; per direction d and 1-D Gauss point q
// the 1-D rational basis rho_i = N_i / S (S = sum_i N_i) and its derivatives
// made physical with g = dx_d/dxi_d and g' = d2x_d/dxi_d^2 from the element's
// three control coordinates along d (x_d depends on the index along d alone):
//   sPd[d][q][0][i] = rho_i,  [1][i] = rho_i' / g,  [2][i] = (rho_i'' - (g'/g) rho_i') / g^2
// (the same in the precision T -> sP when given), and g -> sG[d][q].
template <typename T, int NH = 1>
__device__ __forceinline__ void sep_element(const MeshK &M, const double *__restrict__ Uh, const int *ed, double *sU,
                                            double (*sPd)[4][3][3], T (*sP)[4][3][3], double (*sG)[4], int64_t ndof = 0,
                                            double *sUd = nullptr, double *sV = nullptr, double cv1 = 0.0, double cv2 = 0.0)
{
    const int tid = threadIdx.x;
    if (tid < 27) {
        const int bx = tid / 9, by = (tid / 3) % 3, bz = tid % 3;
        const int64_t oc = ((int64_t)M.nb[1] * M.nb[2]) * own1(ed[0] + bx, M.nb[0], M.periodic[0])
                         + (int64_t)M.nb[2] * own1(ed[1] + by, M.nb[1], M.periodic[1]) + own1(ed[2] + bz, M.nb[2], M.periodic[2]);
        if (NH == 1) {
#pragma unroll
            for (int k = 0; k < 3; k++) sU[3 * tid + k] = Uh[3 * oc + k];
        } else {
            // the transient rule's two fields: Um = (U^1 + U^2)/2 -> sU and
            // Ud = (U - U^2)/2 -> sUd (rule node j: Um + s_j Ud); with sV, the
            // nodal cv1 (U - U^2) + cv2 (U + U^2 - 2 U^1)
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double u0 = Uh[3 * oc + k], u1 = Uh[ndof + 3 * oc + k], u2 = Uh[2 * ndof + 3 * oc + k];
                sU[3 * tid + k] = 0.5 * (u1 + u2);
                sUd[3 * tid + k] = 0.5 * (u0 - u2);
                // the inertia combination of the residual's value slot
                if (sV) sV[3 * tid + k] = cv1 * (u0 - u2) + cv2 * (u0 + u2 - 2 * u1);
            }
        }
    } else if (tid >= 64 && tid < 64 + 12) {
        const int d = (tid - 64) >> 2, q = (tid - 64) & 3;
        const double *kV = M.knot[d] + ed[d];
        const double c = M.gq[q];
        const double x = (-c + 1.0) * kV[2] + c * kV[3];
        double t[9];
        bspline1d(kV, x, t, t + 3, t + 6);
        // the rational basis of equal weights is N_b / W with W = S_x S_y S_z,
        // S = sum_i N_i (1 up to the MTOL guard of the B-spline formulas):
        // a product of the 1-D quotients rho_i = N_i / S, whose derivatives
        // follow from the 1-D quotient rule
        const double S0 = t[0] + t[1] + t[2], S1 = t[3] + t[4] + t[5], S2 = t[6] + t[7] + t[8];
        const double iS = 1.0 / S0;
#pragma unroll
        for (int b = 0; b < 3; b++) {
            t[b] *= iS;
            t[3 + b] = (t[3 + b] - t[b] * S1) * iS;
            t[6 + b] = (t[6 + b] - 2.0 * t[3 + b] * S1 - t[b] * S2) * iS;
        }
        double xc[3];
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const int ax = ed[0] + (d == 0 ? b : 0), ay = ed[1] + (d == 1 ? b : 0), az = ed[2] + (d == 2 ? b : 0);
            xc[b] = M.X[3 * (((int64_t)M.na[1] * M.na[2]) * ax + (int64_t)M.na[2] * ay + az) + d];
        }
        const double g = t[3] * xc[0] + t[4] * xc[1] + t[5] * xc[2];
        const double gp = t[6] * xc[0] + t[7] * xc[1] + t[8] * xc[2];
        const double ig = 1.0 / g, r = gp * ig;
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const double p1 = t[3 + b] * ig, p2 = (t[6 + b] - r * t[3 + b]) * (ig * ig);
            sPd[d][q][0][b] = t[b];
            sPd[d][q][1][b] = p1;
            sPd[d][q][2][b] = p2;
            if (sP) {
                sP[d][q][0][b] = (T)t[b];
                sP[d][q][1][b] = (T)p1;
                sP[d][q][2][b] = (T)p2;
            }
        }
        sG[d][q] = g;
    }
}

// Physical displacement jets at the 64 Gauss points of a separable element by
// sum factorisation over the physical tables (z, y, then x; sWork holds 1188
// doubles), after sep_element and a barrier.  Thread 4q + 1 + k returns in J
// the jet of component k at point q (par_periodic's macro-field added unless
// macro is false), thread 4q zeros.
__device__ __forceinline__ void sep_fields(const MeshK &M, const int *ed, const double *sU, const double (*sPd)[4][3][3],
                                           double *sWork, double *J, bool macro = true)
{
    const int tid = threadIdx.x;
    double *sA = sWork, *sB = sA + 3 * 108;
    // z: sA[k][bx*3+by][qz][dz] = sum_bz U_k(b) Pz[qz][dz][bz]
    if (tid < 27) {
        const int k = tid / 9, bxy = tid - 9 * k;
        const double c0 = sU[3 * (3 * bxy) + k], c1 = sU[3 * (3 * bxy + 1) + k], c2 = sU[3 * (3 * bxy + 2) + k];
        double *o = sA + 108 * k + 12 * bxy;
#pragma unroll
        for (int qz = 0; qz < 4; qz++)
#pragma unroll
            for (int dz = 0; dz < 3; dz++) {
                const double *tz = &sPd[2][qz][dz][0];
                o[3 * qz + dz] = c0 * tz[0] + c1 * tz[1] + c2 * tz[2];
            }
    }
    __syncthreads();
    // y: sB[k][bx][qy][qz][s6] = sum_by sA[k][bx*3+by][qz][dz(s6)] Py[qy][dy(s6)][by],
    //    s6 = (dy, dz) in the order (0,0) (1,0) (0,1) (2,0) (1,1) (0,2)
    if (tid < 36) {
        const int k = tid / 12, r = tid - 12 * k, bx = r >> 2, qz = r & 3;
        const double *a = sA + 108 * k + 36 * bx + 3 * qz;   // a[12 by + dz]
        double av[3][3];
#pragma unroll
        for (int by = 0; by < 3; by++)
#pragma unroll
            for (int dz = 0; dz < 3; dz++) av[by][dz] = a[12 * by + dz];
        double *o = sB + 288 * k + 96 * bx + 6 * qz;
#pragma unroll
        for (int qy = 0; qy < 4; qy++)
#pragma unroll
            for (int s6 = 0; s6 < 6; s6++) {
                const int dy = (s6 == 1 || s6 == 4) ? 1 : (s6 == 3 ? 2 : 0);
                const int dz = (s6 == 2 || s6 == 4) ? 1 : (s6 == 5 ? 2 : 0);
                const double *ty = &sPd[1][qy][dy][0];
                o[24 * qy + s6] = av[0][dz] * ty[0] + av[1][dz] * ty[1] + av[2][dz] * ty[2];
            }
    }
    __syncthreads();
    // x
    const int q = tid >> 2, part = tid & 3;
    const int qd[3] = {q >> 4, (q >> 2) & 3, q & 3};
    if (part > 0) {
        const int k = part - 1;
        sf_jet(sB + 288 * k + (4 * qd[1] + qd[2]) * 6, &sPd[0][qd[0]][0][0], J);
        double xq[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double *kV = M.knot[d] + ed[d];
            const double c = M.gq[qd[d]];
            xq[d] = (-c + 1.0) * kV[2] + c * kV[3];
        }
        if (macro) add_periodic1(M, xq, k, J);
    } else {
#pragma unroll
        for (int a = 0; a < 10; a++) J[a] = 0.0;
    }
}

// Gauss weight x span lengths x det G at point q of a separable element
__device__ __forceinline__ double sep_weight(const MeshK &M, const int *ed, int q, const double (*sG)[4])
{
    const int qd[3] = {q >> 4, (q >> 2) & 3, q & 3};
    double dX[3];
#pragma unroll
    for (int d = 0; d < 3; d++) dX[d] = M.knot[d][ed[d] + 3] - M.knot[d][ed[d] + 2];
    return M.gw[qd[0]] * M.gw[qd[1]] * M.gw[qd[2]] * dX[0] * dX[1] * dX[2] * (sG[0][qd[0]] * sG[1][qd[1]] * sG[2][qd[2]]);
}

// Row stride of the pencil's F buffer (18 columns): 19, so that the lanes of a
// row store, each writing its own row, spread over the LDS banks
constexpr int SEP_FS = 19;
// Block stride of the F buffer (9 rows per block): 178, 18 modulo the 32 LDS
// bank pairs, so that the z pass's lanes (18 columns of each of four blocks)
// and the rows' stores put at most two (z pass) or two to three (stores)
// lanes on a bank pair; 171 = 9 x 19 gives three
constexpr int SEP_FB = 178;
// E1 of the pencil in LDS as [block p][bz][cz][s][s'] (s, s' the x-y parts):
// a y-pass thread's 36 values (its bz, cz) contiguous, read two at a time.
// Strides (even: 16-byte aligned pairs): SEP_EC per (bz, cz), SEP_EB per block.
constexpr int SEP_EC = 38;
constexpr int SEP_EB = 342;

// jet slot a = 1..9 of the pointwise rows -> order in z, x-y part s = (dx, dy)
constexpr int SEP_DZ[10] = {0, 0, 0, 1, 0, 0, 1, 0, 1, 2};
constexpr int SEP_XY[10] = {0, 1, 2, 0, 3, 4, 1, 5, 2, 0};

// z pass of one E1 column (block p, column (s', cz)), task zt < 108 of the
// pencil: E1[p][(s,bz)][(s',cz)] = sum_qz sum_{a: s(a) = s} Pz[qz][dz(a)][bz]
// F_qz[p][a][(s',cz)].  HALVES: only the x-y parts s of half h ({0, 1, 3}:
// five jet slots; {2, 4, 5}: four; h uniform over the wave); otherwise all
// six.  Every entry takes its terms in the same order either way.
template <typename T, bool HALVES>
__device__ __forceinline__ void sep_zcol(const T *sF, const T (*sP)[4][3][3], T *sE, int zt, int h)
{
    constexpr int NA = 9;
    constexpr int ZH[6] = {0, 0, 1, 0, 1, 1};              // half of s
    constexpr int ZS[6] = {0, 1, 0, 2, 1, 2};              // place of s in its half
    constexpr int ZI[2][3] = {{0, 1, 3}, {2, 4, 5}};       // s at each place
    constexpr int NS = HALVES ? 3 : 6;
    const int zp = zt / 18, zr = zt - 18 * zp;
    T e1[3 * NS];   // [s][bz]
#pragma unroll
    for (int i = 0; i < 3 * NS; i++) e1[i] = (T)0;
#pragma unroll
    for (int qz = 0; qz < 4; qz++) {
        const T *Fp = sF + (qz * 6 + zp) * SEP_FB + zr;
        const T *pz = &sP[2][qz][0][0];
        if (!HALVES) {
#pragma unroll
            for (int bb = 0; bb < NA; bb++) {
                const int a = bb + 1;
                const T fv = Fp[SEP_FS * bb];
#pragma unroll
                for (int bz = 0; bz < 3; bz++) e1[3 * SEP_XY[a] + bz] += pz[3 * SEP_DZ[a] + bz] * fv;
            }
        } else if (h == 0) {
#pragma unroll
            for (int bb = 0; bb < NA; bb++) {
                const int a = bb + 1;
                if (ZH[SEP_XY[a]] != 0) continue;
                const T fv = Fp[SEP_FS * bb];
#pragma unroll
                for (int bz = 0; bz < 3; bz++) e1[3 * ZS[SEP_XY[a]] + bz] += pz[3 * SEP_DZ[a] + bz] * fv;
            }
        } else {
#pragma unroll
            for (int bb = 0; bb < NA; bb++) {
                const int a = bb + 1;
                if (ZH[SEP_XY[a]] != 1) continue;
                const T fv = Fp[SEP_FS * bb];
#pragma unroll
                for (int bz = 0; bz < 3; bz++) e1[3 * ZS[SEP_XY[a]] + bz] += pz[3 * SEP_DZ[a] + bz] * fv;
            }
        }
    }
    const int sp = zr / 3, cz = zr - 3 * sp;   // column (s', cz)
    T *o = sE + zp * SEP_EB + cz * SEP_EC + sp;
#pragma unroll
    for (int i = 0; i < 3 * NS; i++) o[(i % 3) * 3 * SEP_EC + 6 * (HALVES ? ZI[h][i / 3] : i / 3)] = e1[i];
}

// The pencil loop of the separable tangents, after stage A and a barrier
// (256 threads).  rows(q, p, ar, row) gives row[b - 1] = C_q[(k, ar + 1),
// (l, b)], b = 1..9, of block p = (k, l) at Gauss point q (unweighted; row
// arrives zeroed); sWgt holds the point weights, sF and sE the F and E1
// buffers.  The z pass runs on 108 threads (block p, column (s', cz) of E1;
// FP64: 216, a column per half of the x-y parts), the y/x pass on 162
// (block p, bz, cz, by: E2[dx][dx'][cy], K[bx][cx][cy]), all
// numbered from the rotated thread index rt so that the workgroups sharing a
// CU put them on different SIMDs.  The y pass of pencil c runs beside the
// pointwise stage of pencil c + 1 (F and E1 have buffers of their own), so a
// pencil takes two barriers.  Returns in K the 27 entries [bx][cx][cy] of
// thread rt < 162.
template <typename T, class Rows>
__device__ __forceinline__ void sep_pencils(int rt, const T (*sP)[4][3][3], const double *sWgt, T *sF, T *sE, Rows rows, T *K)
{
    constexpr int NA = 9;    // jet slots 1..9 (no value slot in the pointwise rows)
    constexpr int FR = 18;   // (s', cz) columns of F and E1
    const int yp = rt / 27, yu = rt - 27 * yp;
    const int ybz = yu / 9, ycz = (yu / 3) % 3, yby = yu % 3;
    constexpr int XYX[6] = {0, 1, 0, 2, 1, 0}, XYY[6] = {0, 0, 1, 0, 1, 2};
    T E2[27];   // [dx][dx'][cy]
#pragma unroll
    for (int i = 0; i < 27; i++) { E2[i] = (T)0; K[i] = (T)0; }

    for (int chunk = 0; chunk <= 16; chunk++) {
        if (chunk > 0 && rt < 162) {
            // ---- y pass of the previous pencil, x pass after a qx's fourth qy
            const int pc = chunk - 1, qx = pc >> 2, qy = pc & 3;
            typedef T pair_t __attribute__((ext_vector_type(2)));
            const pair_t *E = reinterpret_cast<const pair_t *>(sE + yp * SEP_EB + (3 * ybz + ycz) * SEP_EC);   // [s1][s2]
            T pyb[3], pyc[9];
#pragma unroll
            for (int d = 0; d < 3; d++) pyb[d] = sP[1][qy][d][yby];
#pragma unroll
            for (int i = 0; i < 9; i++) pyc[i] = sP[1][qy][i / 3][i % 3];   // [dy'][cy]
#pragma unroll
            for (int s1 = 0; s1 < 6; s1++) {
                T ev[6];
#pragma unroll
                for (int v = 0; v < 3; v++) {
                    const pair_t x = E[3 * s1 + v];
                    ev[2 * v] = x[0];
                    ev[2 * v + 1] = x[1];
                }
#pragma unroll
                for (int s2 = 0; s2 < 6; s2++) {
                    const T t1 = pyb[XYY[s1]] * ev[s2];
#pragma unroll
                    for (int cy = 0; cy < 3; cy++) E2[(3 * XYX[s1] + XYX[s2]) * 3 + cy] += t1 * pyc[3 * XYY[s2] + cy];
                }
            }
            if (qy == 3) {
                T px[9], G[27];   // px[dx][b]; G[dx][cx][cy]
#pragma unroll
                for (int i = 0; i < 9; i++) px[i] = sP[0][qx][i / 3][i % 3];
#pragma unroll
                for (int dx = 0; dx < 3; dx++)
#pragma unroll
                    for (int cx = 0; cx < 3; cx++)
#pragma unroll
                        for (int cy = 0; cy < 3; cy++)
                            G[(3 * dx + cx) * 3 + cy] = px[cx] * E2[(3 * dx) * 3 + cy] + px[3 + cx] * E2[(3 * dx + 1) * 3 + cy]
                                                      + px[6 + cx] * E2[(3 * dx + 2) * 3 + cy];
#pragma unroll
                for (int bx = 0; bx < 3; bx++)
#pragma unroll
                    for (int cx = 0; cx < 3; cx++)
#pragma unroll
                        for (int cy = 0; cy < 3; cy++)
                            K[(3 * bx + cx) * 3 + cy] += px[bx] * G[cx * 3 + cy] + px[3 + bx] * G[(3 + cx) * 3 + cy]
                                                       + px[6 + bx] * G[(6 + cx) * 3 + cy];
#pragma unroll
                for (int i = 0; i < 27; i++) E2[i] = (T)0;
            }
        }
        if (chunk == 16) break;
        // ---- rows of D_q and their z contraction F_q: the pencil's 216 rows
        //      sorted by kind, so that a wave mostly takes one branch of the
        //      row function -- tasks 0..71 the first-gradient rows (point
        //      t / 18, block (t % 18) / 3, row t % 3 + 1), 72..215 the
        //      second-gradient rows (point u / 36, block (u % 36) / 6, row
        //      u % 6 + 4, u = t - 72); the mixed wave (tasks 64..127) is the
        //      one without y-pass work (rt 192..255)
        {
            const int t = (rt + 128) & 255;
            if (t < 216) {
                int qz, p, ar;
                if (t < 72) {
                    qz = t / 18;
                    const int r = t - 18 * qz;
                    p = r / 3;
                    ar = r - 3 * p;
                } else {
                    const int u = t - 72;
                    qz = u / 36;
                    const int r = u - 36 * qz;
                    p = r / 6;
                    ar = 3 + r - 6 * p;
                }
                const int q = 4 * chunk + qz;
                double row[NA];
#pragma unroll
                for (int bb = 0; bb < NA; bb++) row[bb] = 0.0;
                rows(q, p, ar, row);
                const double wgt = sWgt[q];
                const T *pz = &sP[2][qz][0][0];
                T f[FR];
#pragma unroll
                for (int i = 0; i < FR; i++) f[i] = (T)0;
#pragma unroll
                for (int bb = 0; bb < NA; bb++) {
                    const int a = bb + 1;
                    const T v = (T)(wgt * row[bb]);
#pragma unroll
                    for (int cz = 0; cz < 3; cz++) f[3 * SEP_XY[a] + cz] += v * pz[3 * SEP_DZ[a] + cz];
                }
                T *o = sF + (qz * 6 + p) * SEP_FB + ar * SEP_FS;
#pragma unroll
                for (int i = 0; i < FR; i++) o[i] = f[i];
            }
        }
        __syncthreads();
        // ---- z pass over the pencil's four points -> E1[p][(s,bz)][(s',cz)].
        //      FP64: on all four waves, each half of rt taking one half of
        //      the x-y parts s (a wave runs one branch); FP32, capped at 168
        //      VGPRs for three workgroups per CU, keeps whole columns on 108
        //      threads (the halves spill there)
        if constexpr (sizeof(T) == 8) {
            if ((rt & 127) < 108) sep_zcol<T, true>(sF, sP, sE, rt & 127, rt >> 7);
        } else if (rt < 108) {
            sep_zcol<T, false>(sF, sP, sE, rt, 0);
        }
        __syncthreads();
    }

}

// Write-out of the separable tangents: thread rt < 162 holds the entries
// K[bx][cx][cy] of block (i, j) = (kBlkI[yp], kBlkJ[yp]) for its (by, bz, cz);
// sWork (at least 972 bytes) is free.
template <typename T, bool DIRECT>
__device__ __forceinline__ void sep_write(const MeshK &M, const Scatter &S, T *__restrict__ Ke, int e, int ex, int ey, int ez, int rt,
                                          const T *K, char *sWork)
{
    const int tid = threadIdx.x;
    const int yp = rt / 27, yu = rt - 27 * yp;
    const int ybz = yu / 9, ycz = (yu / 3) % 3, yby = yu % 3;
    if constexpr (DIRECT) {
        // ---- add (row 3b+i, column 3c+j) and the mirrored entry into the CSR
        //      values, Dirichlet rows and columns left at zero; the elements
        //      of one colour share no row.  Row starts and masks of the
        //      element's 81 rows staged in LDS (the F buffer is free after
        //      the last pencil), then per bx all loads of the nine entries
        //      (and mirrors) before their stores (one round trip for all 27
        //      measured slower: the scattered read-modify-write of the CSR
        //      values, not its latency, is the cost).
        int64_t *sRow = reinterpret_cast<int64_t *>(sWork);
        int32_t *sMsk = reinterpret_cast<int32_t *>(sWork + 81 * sizeof(int64_t));
        if (tid < 81) {
            const int b = tid / 3, i = tid - 3 * (tid / 3);
            const int64_t cp = ((int64_t)M.nb[1] * M.nb[2]) * (ex + b / 9) + (int64_t)M.nb[2] * (ey + (b / 3) % 3) + (ez + b % 3);
            sRow[tid] = S.rowptr[3 * cp + i];
            sMsk[tid] = S.mask[3 * cp + i];
        }
        __syncthreads();
        if (rt < 162) {
            const int i = kBlkI[yp], j = kBlkJ[yp];
#pragma unroll
            for (int bx = 0; bx < 3; bx++) {
                const int b = 9 * bx + 3 * yby + ybz, Bx = ex + bx, By = ey + yby, Bz = ez + ybz;
                T vp[9], vq[9];
                int64_t pp[9], pq[9];
                bool ok[9];
#pragma unroll
                for (int t = 0; t < 9; t++) {
                    const int cx = t / 3, cy = t % 3, c = 9 * cx + 3 * cy + ycz;
                    const int Cx = ex + cx, Cy = ey + cy, Cz = ez + ycz;
                    ok[t] = sMsk[3 * b + i] == 0 && sMsk[3 * c + j] == 0;
                    pp[t] = stencil_pos(M, sRow[3 * b + i], Bx, By, Bz, Cx, Cy, Cz, j);
                    pq[t] = stencil_pos(M, sRow[3 * c + j], Cx, Cy, Cz, Bx, By, Bz, i);
                    vp[t] = vq[t] = (T)0;
                    if (ok[t]) {
                        vp[t] = Ke[pp[t]];
                        if (i != j) vq[t] = Ke[pq[t]];
                    }
                }
#pragma unroll
                for (int t = 0; t < 9; t++) {
                    const T v = K[(3 * bx + t / 3) * 3 + t % 3];
                    if (ok[t]) {
                        Ke[pp[t]] = vp[t] + v;
                        if (i != j) Ke[pq[t]] = vq[t] + v;
                    }
                }
            }
        }
        return;
    }
    // ---- K_e: row 3b+i, column 3c+j, and the mirrored block (3c+j, 3b+i)
    if (rt < 162) {
        const int i = kBlkI[yp], j = kBlkJ[yp];
        T *out = Ke + (int64_t)e * 6561;
#pragma unroll
        for (int bx = 0; bx < 3; bx++)
#pragma unroll
            for (int cx = 0; cx < 3; cx++)
#pragma unroll
                for (int cy = 0; cy < 3; cy++) {
                    const int b = 9 * bx + 3 * yby + ybz, c = 9 * cx + 3 * cy + ycz;
                    const T v = K[(3 * bx + cx) * 3 + cy];
                    out[(3 * b + i) * 81 + 3 * c + j] = v;
                    if (i != j) out[(3 * c + j) * 81 + 3 * b + i] = v;
                }
    }
}

template <typename T, bool DIRECT>
__global__ __launch_bounds__(256, (sizeof(T) == 8) ? 2 : 3) void k_tangent_sep(MeshK M, PhysK P, const double *__restrict__ Uh, int64_t ndof,
                                                        T *__restrict__ Ke, int nel, Scatter S)
{
    constexpr int WORK_A = 3 * (108 + 288) * (int)sizeof(double);   // field passes of the 3 displacement components
    constexpr int WORK_F = 4 * 6 * SEP_FB * (int)sizeof(T);        // F of the pencil's four points (9 rows x 18 columns)
    __shared__ double sPd[3][4][3][3];          // 1-D tables of physical derivatives [d][q_d][order][i]
    __shared__ T sP[3][4][3][3];                // the same in the contraction's precision
    __shared__ double sG[3][4];                 // dx_d / dxi_d at the 1-D Gauss points
    __shared__ double sU[27 * 3];
    __shared__ double sWgt[64];
    // strain-measure states, compact layout, at a stride of 50 doubles: a wave's
    // rows cover the pencil's four points, and 48 would put points two apart
    // on the same LDS banks
    constexpr int SSTR = 50;
    __shared__ double sSt[64 * SSTR];
    __shared__ __attribute__((aligned(16))) T sE[6 * SEP_EB];   // E1 of the current pencil
    __shared__ __attribute__((aligned(16))) char sWork[WORK_A > WORK_F ? WORK_A : WORK_F];
    T *sF = reinterpret_cast<T *>(sWork);

    const int tid = threadIdx.x;
    const int e = blockIdx.x;
    int ex, ey, ez;
    if (DIRECT) {   // element e of the colour S.c
        const int k = e % S.nc[2], j = (e / S.nc[2]) % S.nc[1], i = e / (S.nc[1] * S.nc[2]);
        ex = S.c[0] + 3 * i;
        ey = S.c[1] + 3 * j;
        ez = S.c[2] + 3 * k;
    } else {
        const int nz = M.nelem[2], ny = M.nelem[1];
        ex = e / (ny * nz);
        ey = (e / nz) % ny;
        ez = e % nz;
    }
    const int ed[3] = {ex, ey, ez};

    sep_element<T>(M, Uh, ed, sU, sPd, sP, sG);
    __syncthreads();
    // ---- stage A: displacement jets, strain-measure states, point weights
    {
        double J[10];
        sep_fields(M, ed, sU, sPd, reinterpret_cast<double *>(sWork), J);
        const int q = tid >> 2, part = tid & 3, lq = (tid & 63) & ~3;
        double u[30];
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int a = 0; a < 10; a++) u[10 * k + a] = __shfl(J[a], lq + 1 + k, 64);
        if (part == 0) {
            State st;
            phys_state(u, P, st);
            double *sp = sSt + q * SSTR;
            constexpr int PR[6] = {0, 1, 2, 4, 5, 8};   // index pairs I <= J in slot order
#pragma unroll
            for (int i = 0; i < 9; i++) { sp[i] = st.F[i]; sp[STC_G + i] = st.G[i]; }
#pragma unroll
            for (int k = 0; k < 3; k++)
#pragma unroll
                for (int t = 0; t < 6; t++) sp[STC_H + 6 * k + t] = u[10 * k + 4 + t];
#pragma unroll
            for (int t = 0; t < 6; t++) { sp[STC_S + t] = st.S[PR[t]]; sp[STC_N + t] = st.N[PR[t]]; }
            sWgt[q] = sep_weight(M, ed, q, sG);
        }
    }
    __syncthreads();

    const int rt = (tid + 64 * ((e ^ (e >> 3)) & 3)) & 255;
    T K[27];
    // the branch-light rows (FP32: capped at 168 VGPRs for three workgroups
    // per CU, 5 of them spilled -- faster than the select form of c_row)
    auto rows = [&](int q, int p, int ar, double *row) {
        c_row_onehot<true>(sSt + q * SSTR, P, kBlkI[p], ar + 1, kBlkJ[p], row);
    };
    sep_pencils<T>(rt, sP, sWgt, sF, sE, rows, K);
    sep_write<T, DIRECT>(M, S, Ke, e, ex, ey, ez, rt, K, sWork);
}

// Transient tangent on a separable patch (k_tangent_sep_t, gradelasttime):
// the rule-weighted pointwise rows of k_tangent_elem<..., 3> (c_row_mom: the
// eight rule-node states us_j = Um + s_j Ud collapsed into moments) in the
// pencil loop of k_tangent_sep, plus the inertia term.  The inertia block
// D[(k,0),(k,0)] = w_q (cii / (2 dt) + rho / dt^2) pairs only basis values,
// and w_q and the basis value are products over the directions, so its part
// of K^kk is mc Mx[bx][cx] My[by][cy] Mz[bz][cz] with the 1-D matrices
// M_d[b][c] = sum_q gw_q dX_d g_d(q) rho_b(q) rho_c(q).  Stage A: the jets of
// Um and Ud (two field passes), then the moment states of all 64 points,
// compact and in the contraction's precision (one lane per rule node, sums
// over 8-lane groups), so that the FP32 kernel keeps two workgroups per CU
// (FP64: 95 KB of LDS, one).
template <typename T, bool DIRECT>
__global__ __launch_bounds__(256, (sizeof(T) == 8) ? 1 : 2) void k_tangent_sep_t(MeshK M, PhysK P, const double *__restrict__ Uh,
                                                                               int64_t ndof, T *__restrict__ Ke, int nel, Scatter S)
{
    constexpr int WORK_A = 3 * (108 + 288) * (int)sizeof(double);   // field passes of 3 components
    constexpr int WORK_F = 4 * 6 * SEP_FB * (int)sizeof(T);        // F of the pencil's four points
    constexpr int WORK_J = 64 * 54 * (int)sizeof(T);               // jets of Um and Ud, slots 1..9
    constexpr int WORK = WORK_A > WORK_F ? (WORK_A > WORK_J ? WORK_A : WORK_J) : (WORK_F > WORK_J ? WORK_F : WORK_J);
    __shared__ double sPd[3][4][3][3];
    __shared__ T sP[3][4][3][3];
    __shared__ double sG[3][4];
    __shared__ double sU[27 * 3], sUd[27 * 3];
    __shared__ double sWgt[64];
    __shared__ double sMass[3][3][3];          // 1-D mass matrices M_d[b][c]
    __shared__ T sMom[64 * MWC];               // compact moment states
    __shared__ __attribute__((aligned(16))) T sE[6 * SEP_EB];
    __shared__ __attribute__((aligned(16))) char sWork[WORK];
    T *sF = reinterpret_cast<T *>(sWork);
    T *sJet = reinterpret_cast<T *>(sWork);     // [q][h][k][a - 1], stage A only

    const int tid = threadIdx.x;
    const int e = blockIdx.x;
    int ex, ey, ez;
    if (DIRECT) {
        const int k = e % S.nc[2], j = (e / S.nc[2]) % S.nc[1], i = e / (S.nc[1] * S.nc[2]);
        ex = S.c[0] + 3 * i;
        ey = S.c[1] + 3 * j;
        ez = S.c[2] + 3 * k;
    } else {
        const int nz = M.nelem[2], ny = M.nelem[1];
        ex = e / (ny * nz);
        ey = (e / nz) % ny;
        ez = e % nz;
    }
    const int ed[3] = {ex, ey, ez};

    sep_element<T, 3>(M, Uh, ed, sU, sPd, sP, sG, ndof, sUd);
    __syncthreads();
    // ---- stage A: jets of Um and Ud, point weights, 1-D mass matrices
    {
        double J1[10], J2[10];
        sep_fields(M, ed, sU, sPd, reinterpret_cast<double *>(sWork), J1, true);
        __syncthreads();
        sep_fields(M, ed, sUd, sPd, reinterpret_cast<double *>(sWork), J2, false);
        __syncthreads();
        const int q = tid >> 2, part = tid & 3;
        if (part > 0) {
            const int k = part - 1;
            T *jq = sJet + q * 54;
#pragma unroll
            for (int a = 1; a < 10; a++) {
                jq[9 * k + a - 1] = (T)J1[a];
                jq[27 + 9 * k + a - 1] = (T)J2[a];
            }
        } else {
            sWgt[q] = sep_weight(M, ed, q, sG);
        }
        if (tid < 27) {
            const int d = tid / 9, b = (tid / 3) % 3, c = tid % 3;
            const double dX = M.knot[d][ed[d] + 3] - M.knot[d][ed[d] + 2];
            double m = 0.0;
#pragma unroll
            for (int qq = 0; qq < 4; qq++) m += M.gw[qq] * dX * sG[d][qq] * sPd[d][qq][0][b] * sPd[d][qq][0][c];
            sMass[d][b][c] = m;
        }
    }
    __syncthreads();
    // ---- moment states: lane 8g + j takes rule node j of point g (32 points
    //      per round, two rounds); node sums over the 8-lane group one entry
    //      at a time; F^ and H^ from the jets, 3-4 entries per lane
    {
        Mom mu = {0.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0};
#pragma unroll
        for (int j = 0; j < 8; j++) {
            mu.m0 += P.ct[j];
            mu.m1 += P.ct[j] * P.s[j];
            mu.m2 += P.ct[j] * P.s[j] * P.s[j];
        }
        mu.factor();
        const double c12 = 1.0 / (mu.a11 * mu.a22), c22 = 1.0 / mu.d2;
        const int j = tid & 7;
        for (int round = 0; round < 2; round++) {
            const int q = (tid >> 3) + 32 * round;
            const T *jq = sJet + q * 54;
            T *ms = sMom + q * MWC;
            {
                double us[30];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    us[10 * k] = 0.0;
#pragma unroll
                    for (int a = 1; a < 10; a++) us[10 * k + a] = (double)jq[9 * k + a - 1] + P.s[j] * (double)jq[27 + 9 * k + a - 1];
                }
                State st;
                phys_state(us, P, st);
                const double w0 = P.ct[j], w1 = w0 * P.s[j], w2 = w1 * P.s[j];
                constexpr int PR[6] = {0, 1, 2, 4, 5, 8};   // index pairs I <= J in slot order
#pragma unroll
                for (int t = 0; t < 6; t++) {
                    double sv = w0 * st.S[PR[t]], n0 = w0 * st.N[PR[t]], n1 = w1 * st.N[PR[t]], n2 = w2 * st.N[PR[t]];
#pragma unroll
                    for (int o = 1; o < 8; o <<= 1) {
                        sv += __shfl_xor(sv, o, 64);
                        n0 += __shfl_xor(n0, o, 64);
                        n1 += __shfl_xor(n1, o, 64);
                        n2 += __shfl_xor(n2, o, 64);
                    }
                    if (j == 0) {
                        ms[MC_S + t] = (T)sv;
                        ms[MC_N11 + t] = (T)(n0 / mu.m0);
                        ms[MC_N12 + t] = (T)((n1 - mu.lf * n0) * c12);
                        ms[MC_N22 + t] = (T)(((mu.lf * n0 - n1) * mu.lf + (n2 - mu.lf * n1)) * c22);
                    }
                }
#pragma unroll
                for (int i = 0; i < 9; i++) {
                    double gv = w0 * st.G[i];
#pragma unroll
                    for (int o = 1; o < 8; o <<= 1) gv += __shfl_xor(gv, o, 64);
                    if (j == 0) ms[MC_G + i] = (T)gv;
                }
            }
            // F^ (r < 9: F_kL, k = r / 3, L = r % 3) and H^ (r >= 9: H_k,IK in
            // slot 4 + t of component k = (r - 9) / 6): X^_1 = a11 (X0 + lf Xd),
            // X^_2 = a22 Xd
            for (int r = j; r < 27; r += 8) {
                int src;
                double x0;
                if (r < 9) {
                    src = 9 * (r / 3) + (r % 3);
                    x0 = (double)jq[src] + ((r / 3) == (r % 3) ? 1.0 : 0.0);
                } else {
                    src = 9 * ((r - 9) / 6) + 3 + (r - 9) % 6;
                    x0 = (double)jq[src];
                }
                const double xd = (double)jq[27 + src];
                const double x1 = mu.a11 * (x0 + mu.lf * xd), x2 = mu.a22 * xd;
                if (r < 9) {
                    ms[MC_F1 + r] = (T)x1;
                    ms[MC_F2 + r] = (T)x2;
                } else {
                    ms[MC_H1 + r - 9] = (T)x1;
                    ms[MC_H2 + r - 9] = (T)x2;
                }
            }
        }
    }
    __syncthreads();

    const int rt = (tid + 64 * ((e ^ (e >> 3)) & 3)) & 255;
    T K[27];
    auto rows = [&](int q, int p, int ar, double *row) { c_row_momc_onehot<T>(sMom + q * MWC, P, kBlkI[p], ar + 1, kBlkJ[p], row); };
    sep_pencils<T>(rt, sP, sWgt, sF, sE, rows, K);
    // inertia on the diagonal blocks
    if (rt < 162) {
        const int yp = rt / 27, yu = rt - 27 * yp;
        const int ybz = yu / 9, ycz = (yu / 3) % 3, yby = yu % 3;
        if (kBlkI[yp] == kBlkJ[yp]) {
            const double mc = P.cii / (2. * P.dt) + P.rho / (P.dt * P.dt);
#pragma unroll
            for (int bx = 0; bx < 3; bx++)
#pragma unroll
                for (int cx = 0; cx < 3; cx++)
#pragma unroll
                    for (int cy = 0; cy < 3; cy++)
                        K[(3 * bx + cx) * 3 + cy] += (T)(mc * sMass[0][bx][cx] * sMass[1][yby][cy] * sMass[2][ybz][ycz]);
        }
    }
    sep_write<T, DIRECT>(M, S, Ke, e, ex, ey, ez, rt, K, sWork);
}

// Transposed sum factorisation of the weighted pointwise residual jets
// sR[q * RS + 10 k + a] (after a barrier; sWork holds 1188 doubles): returns
// R_e[tid] = sum_q sum_a r_q[k][a] R_q[a][b] (tid = 3 b + k < 81), the
// reference's projection (src/assmcore_iga.c:230-252) as x -> y -> z passes
// over the physical tables; 0 for tid >= 81.
template <int RS>
__device__ __forceinline__ double sep_project(const double (*sPd)[4][3][3], const double *sR, double *sWork)
{
    // ---- x^T: C1[k][bx][qy][qz][s6] = sum_qx sum_{a: S6(a) = s6} r[q][k][a] Px[qx][DX(a)][bx]
    const int tid = threadIdx.x;
    double *C1 = sWork, *C2 = sWork + 864;
    for (int i = tid; i < 864; i += 256) {   // s6 slowest: the switch below is mostly wave-uniform
        const int s6 = i / 144, t = i - 144 * s6, qz = t & 3, qy = (t >> 2) & 3, bx = (t >> 4) % 3, k = t / 48;
        double acc = 0.0;
#pragma unroll
        for (int qx = 0; qx < 4; qx++) {
            const double *rq = sR + (16 * qx + 4 * qy + qz) * RS + 10 * k;
            const double *tx = &sPd[0][qx][0][0];   // tx[3 * dx + bx]
            switch (s6) {
                case 0: acc += rq[0] * tx[bx] + rq[1] * tx[3 + bx] + rq[4] * tx[6 + bx]; break;
                case 1: acc += rq[2] * tx[bx] + rq[5] * tx[3 + bx]; break;
                case 2: acc += rq[3] * tx[bx] + rq[6] * tx[3 + bx]; break;
                case 3: acc += rq[7] * tx[bx]; break;
                case 4: acc += rq[8] * tx[bx]; break;
                default: acc += rq[9] * tx[bx]; break;
            }
        }
        C1[6 * t + s6] = acc;   // [k][bx][qy][qz][s6]
    }
    __syncthreads();
    // ---- y^T: C2[k][bx][by][qz][dz] = sum_qy sum_{s6: dz(s6) = dz} C1[..][s6] Py[qy][dy(s6)][by]
    for (int i = tid; i < 324; i += 256) {   // dz slowest (mostly wave-uniform branches)
        const int dz = i / 108, t = i - 108 * dz, qz = t & 3, u = t >> 2, by = u % 3, bx = (u / 3) % 3, k = u / 9;
        double acc = 0.0;
#pragma unroll
        for (int qy = 0; qy < 4; qy++) {
            const double *c = C1 + (((k * 3 + bx) * 4 + qy) * 4 + qz) * 6;
            const double *ty = &sPd[1][qy][0][0];   // ty[3 * dy + by]
            if (dz == 0) acc += c[0] * ty[by] + c[1] * ty[3 + by] + c[3] * ty[6 + by];
            else if (dz == 1) acc += c[2] * ty[by] + c[4] * ty[3 + by];
            else acc += c[5] * ty[by];
        }
        C2[3 * t + dz] = acc;   // [k][bx][by][qz][dz]
    }
    __syncthreads();
    // ---- z^T: R_e[3 b + k] = sum_qz sum_dz C2[k][bx][by][qz][dz] Pz[qz][dz][bz]
    double acc = 0.0;
    if (tid < 81) {
        const int k = tid % 3, b = tid / 3, bx = b / 9, by = (b / 3) % 3, bz = b % 3;
        const double *c = C2 + ((k * 3 + bx) * 3 + by) * 12;
#pragma unroll
        for (int qz = 0; qz < 4; qz++)
#pragma unroll
            for (int dz = 0; dz < 3; dz++) acc += c[3 * qz + dz] * sPd[2][qz][dz][bz];
    }
    return acc;
}

// ============================================================================
// Residual on a separable patch (k_residual_sep, static physics; volume loop
// of Residual_iga, src/assmcore_iga.c:118-259): the displacement jets of
// sep_fields, the pointwise residual jets r_q (phys/gradelast.c) times the
// point weight, and the projection
//   R_e[b][k] = sum_q sum_a r_q[k][a] R_q[a][b]
// with R_q the tensor product of the physical tables, as the transposed sum
// factorisation x -> y -> z; no quotient or chain rule to undo.
// ============================================================================
// Six workgroups per CU (26 KB of LDS each, 80 VGPRs; the one-wave pointwise
// stage spills 32 of them): 7.6 ms at 256x64x64 against 8.3 at five and 9.9
// at four, where nothing spills -- the barriers of the field and projection
// passes are what the extra workgroups hide.
__global__ __launch_bounds__(256, 6) void k_residual_sep(MeshK M, PhysK P, const double *__restrict__ Uh, int64_t ndof,
                                                         double *__restrict__ Re, int nel)
{
    __shared__ double sPd[3][4][3][3];
    __shared__ double sG[3][4];
    __shared__ double sU[27 * 3];
    __shared__ double sWork[3 * (108 + 288)];   // field passes, then x^T (864) and y^T (324)
    __shared__ double sJet[64 * 30];            // u-jets [q][k][a], then the weighted r_q
    const int tid = threadIdx.x;
    const int e = blockIdx.x;
    const int nz = M.nelem[2], ny = M.nelem[1];
    const int ed[3] = {e / (ny * nz), (e / nz) % ny, e % nz};

    sep_element<double>(M, Uh, ed, sU, sPd, nullptr, sG);
    __syncthreads();
    {
        double J[10];
        sep_fields(M, ed, sU, sPd, sWork, J);
        const int q = tid >> 2, part = tid & 3;
        if (part > 0) {
#pragma unroll
            for (int a = 0; a < 10; a++) sJet[q * 30 + 10 * (part - 1) + a] = J[a];
        }
    }
    __syncthreads();
    // ---- pointwise residual jets, weighted: one wave, lane = Gauss point
    //      (the wave taking it rotates with the element)
    if ((tid >> 6) == (e & 3)) {
        const int q = tid & 63;
        double u[30], r[30];
#pragma unroll
        for (int a = 0; a < 30; a++) u[a] = sJet[q * 30 + a];
        phys_residual_static(P, u, r);
        const double w = sep_weight(M, ed, q, sG);
#pragma unroll
        for (int a = 0; a < 30; a++) sJet[q * 30 + a] = w * r[a];
    }
    __syncthreads();
    const double v = sep_project<30>(sPd, sJet, sWork);
    if (tid < 81) Re[(int64_t)e * 81 + tid] = v;
}

// Transient residual on a separable patch (k_residual_sep_t, gradelasttime):
// the jets of Um and Ud, the rule's eight nodes per point (one lane per node
// at Um + s_j Ud, weighted by cr_j and summed over the 8-lane group), the
// point weight, and the projection of k_residual_sep.  The value slot's inertia r[10 k] = cii (u - u^2) / (2 dt)
// + rho (u + u^2 - 2 u^1) / dt^2 is linear in the nodal values, so its part of
// R_e is the element mass matrix Mx x My x Mz applied to the nodal
// combination (k_tangent_sep_t's 1-D mass matrices).
__global__ __launch_bounds__(256, 2) void k_residual_sep_t(MeshK M, PhysK P, const double *__restrict__ Uh, int64_t ndof,
                                                           double *__restrict__ Re, int nel)
{
    __shared__ double sPd[3][4][3][3];
    __shared__ double sG[3][4];
    __shared__ double sU[27 * 3], sUd[27 * 3], sV[27 * 3];
    __shared__ double sMass[3][3][3];
    __shared__ double sWork[3 * (108 + 288)];   // field passes, then x^T (864) and y^T (324)
    __shared__ double sJet[64 * 60];            // jets [q][h][k][a] of Um and Ud, then the weighted r_q [q][k][a]
    const int tid = threadIdx.x;
    const int e = blockIdx.x;
    const int nz = M.nelem[2], ny = M.nelem[1];
    const int ed[3] = {e / (ny * nz), (e / nz) % ny, e % nz};

    sep_element<double, 3>(M, Uh, ed, sU, sPd, nullptr, sG, ndof, sUd, sV, P.cii / (2. * P.dt), P.rho / (P.dt * P.dt));
    __syncthreads();
    {
        double J1[10], J2[10];
        sep_fields(M, ed, sU, sPd, sWork, J1, true);
        __syncthreads();
        sep_fields(M, ed, sUd, sPd, sWork, J2, false);
        const int q = tid >> 2, part = tid & 3;
        if (part > 0) {
#pragma unroll
            for (int a = 0; a < 10; a++) {
                sJet[q * 60 + 10 * (part - 1) + a] = J1[a];
                sJet[q * 60 + 30 + 10 * (part - 1) + a] = J2[a];
            }
        }
        if (tid < 27) {
            const int d = tid / 9, b = (tid / 3) % 3, c = tid % 3;
            const double dX = M.knot[d][ed[d] + 3] - M.knot[d][ed[d] + 2];
            double m = 0.0;
#pragma unroll
            for (int qq = 0; qq < 4; qq++) m += M.gw[qq] * dX * sG[d][qq] * sPd[d][qq][0][b] * sPd[d][qq][0][c];
            sMass[d][b][c] = m;
        }
    }
    __syncthreads();
    // ---- rule-weighted residual jets of the stored energy, times the weight:
    //      lane 8g + j takes rule node j of point g + 32 round, the node sums
    //      over the 8-lane group one entry at a time
    {
        const int j = tid & 7;
        for (int round = 0; round < 2; round++) {
            const int q = (tid >> 3) + 32 * round;
            const double *J = sJet + q * 60;
            double r[30];
            {
                double us[30];
#pragma unroll
                for (int a = 0; a < 30; a++) us[a] = J[a] + P.s[j] * J[30 + a];
                State st;
                phys_state(us, P, st);
#pragma unroll
                for (int a = 0; a < 30; a++) r[a] = 0.0;
                phys_residual_acc(st, P.cr[j], r);
            }
            const double w = (j == 0) ? sep_weight(M, ed, q, sG) : 0.0;
#pragma unroll
            for (int a = 0; a < 30; a++) {
                if (a % 10 == 0) continue;   // value slots: the inertia, added below
                double v = r[a];
#pragma unroll
                for (int o = 1; o < 8; o <<= 1) v += __shfl_xor(v, o, 64);
                r[a] = v;
            }
            if (j == 0) {
#pragma unroll
                for (int a = 0; a < 30; a++) sJet[q * 60 + a] = (a % 10 == 0) ? 0.0 : w * r[a];
            }
        }
    }
    __syncthreads();
    double v = sep_project<60>(sPd, sJet, sWork);
    if (tid < 81) {
        const int k = tid % 3, b = tid / 3, bx = b / 9, by = (b / 3) % 3, bz = b % 3;
        double in = 0.0;
#pragma unroll
        for (int cx = 0; cx < 3; cx++)
#pragma unroll
            for (int cy = 0; cy < 3; cy++) {
                const double mxy = sMass[0][bx][cx] * sMass[1][by][cy];
#pragma unroll
                for (int cz = 0; cz < 3; cz++) in += mxy * sMass[2][bz][cz] * sV[3 * (9 * cx + 3 * cy + cz) + k];
            }
        Re[(int64_t)e * 81 + tid] = v + in;
    }
}

// the unit diagonal MatZeroRowsColumns(diag = 1) leaves on the Dirichlet rows
// of a non-periodic patch (their other entries stayed zero)
template <typename TV>
__global__ __launch_bounds__(256) void k_dirichlet_diag(MeshK M, const int64_t *__restrict__ rowptr, const int32_t *__restrict__ mask,
                                                        TV *__restrict__ vals, int64_t ncp)
{
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= ncp) return;
    const int nb1 = M.nb[1], nb2 = M.nb[2];
    const int az = (int)(a % nb2), ay = (int)((a / nb2) % nb1), ax = (int)(a / ((int64_t)nb1 * nb2));
#pragma unroll
    for (int i = 0; i < 3; i++)
        if (mask[3 * a + i]) vals[csr_pos(M, rowptr, ax, ay, az, i, ax, ay, az, i)] = (TV)1.0;
}

// sorted owned control points within +-2 of a (clipped, or wrapped when periodic)
__device__ __forceinline__ int stencil1(int a, int n, int periodic, int *S)
{
    int cnt = 0;
    if (!periodic) {
        int lo = a - 2 < 0 ? 0 : a - 2, hi = a + 2 > n - 1 ? n - 1 : a + 2;
        for (int b = lo; b <= hi; b++) S[cnt++] = b;
        return cnt;
    }
    int tmp[5];
    for (int t = 0; t < 5; t++) { int v = a + t - 2; v = (v % n + n) % n; tmp[t] = v; }
    for (int i = 1; i < 5; i++)
        for (int j = i; j > 0 && tmp[j] < tmp[j - 1]; j--) { int x = tmp[j]; tmp[j] = tmp[j - 1]; tmp[j - 1] = x; }
    for (int t = 0; t < 5; t++)
        if (cnt == 0 || tmp[t] != S[cnt - 1]) S[cnt++] = tmp[t];
    return cnt;
}

// elements along one direction containing owned points a and b, sorted, with
// their local indices
__device__ __forceinline__ int elem_pairs2(int a, int b, int nelem, int nb, int periodic, int *es, int *la, int *lb)
{
    int cnt = 0;
    if (!periodic) {
        int lo = (a > b ? a : b) - 2, hi = a < b ? a : b;
        if (lo < 0) lo = 0;
        if (hi > nelem - 1) hi = nelem - 1;
        for (int e = lo; e <= hi; e++) { es[cnt] = e; la[cnt] = a - e; lb[cnt] = b - e; cnt++; }
        return cnt;
    }
    for (int l = 2; l >= 0; l--) {
        int e = a - l;
        if (e < 0) e += nb;
        int m = b - e;
        if (m < 0) m += nb;
        if (m <= 2) { es[cnt] = e; la[cnt] = l; lb[cnt] = m; cnt++; }
    }
    for (int i = 1; i < cnt; i++)
        for (int j = i; j > 0 && es[j] < es[j - 1]; j--) {
            int t = es[j]; es[j] = es[j - 1]; es[j - 1] = t;
            t = la[j]; la[j] = la[j - 1]; la[j - 1] = t;
            t = lb[j]; lb[j] = lb[j - 1]; lb[j - 1] = t;
        }
    return cnt;
}

// Column-wise variant of the same sums (used for FP32 element matrices, where
// it is faster): one workgroup per owned control point a, thread t <-> the
// t-th control point b of the row's column stencil, summing the (at most 27)
// 3x3 contributions of the elements containing a and b in ascending element
// order -- the order k_tangent_gather_rows uses, so the values are the same.
template <typename TK, typename TV>
__global__ __launch_bounds__(128) void k_tangent_gather(MeshK M, const TK *__restrict__ Ke, const int64_t *__restrict__ rowptr,
                                                        const int32_t *__restrict__ mask, TV *__restrict__ vals)
{
    const int64_t a = blockIdx.x;
    int nb1 = M.nb[1], nb2 = M.nb[2];
    int az = (int)(a % nb2), ay = (int)((a / nb2) % nb1), ax = (int)(a / ((int64_t)nb1 * nb2));
    int Sx[5], Sy[5], Sz[5];
    int nsx = stencil1(ax, M.nb[0], M.periodic[0], Sx);
    int nsy = stencil1(ay, M.nb[1], M.periodic[1], Sy);
    int nsz = stencil1(az, M.nb[2], M.periodic[2], Sz);
    int t = threadIdx.x;
    if (t >= nsx * nsy * nsz) return;
    int sx = t / (nsy * nsz), sy = (t / nsz) % nsy, sz = t % nsz;
    int bx = Sx[sx], by = Sy[sy], bz = Sz[sz];
    int ex[3], lax[3], lbx[3], ey[3], lay[3], lby[3], ez[3], laz[3], lbz[3];
    int nx = elem_pairs2(ax, bx, M.nelem[0], M.nb[0], M.periodic[0], ex, lax, lbx);
    int ny = elem_pairs2(ay, by, M.nelem[1], M.nb[1], M.periodic[1], ey, lay, lby);
    int nz = elem_pairs2(az, bz, M.nelem[2], M.nb[2], M.periodic[2], ez, laz, lbz);
    double s[9];
#pragma unroll
    for (int i = 0; i < 9; i++) s[i] = 0.0;
    for (int x = 0; x < nx; x++)
        for (int y = 0; y < ny; y++)
            for (int z = 0; z < nz; z++) {
                int64_t e = ((int64_t)ex[x] * M.nelem[1] + ey[y]) * M.nelem[2] + ez[z];
                int la = 9 * lax[x] + 3 * lay[y] + laz[z], lb = 9 * lbx[x] + 3 * lby[y] + lbz[z];
                const TK *k = Ke + e * 6561 + (3 * la) * 81 + 3 * lb;
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++) s[3 * i + j] += (double)k[81 * i + j];
            }
    int64_t bcp = ((int64_t)M.nb[1] * M.nb[2]) * bx + (int64_t)M.nb[2] * by + bz;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        int64_t row = 3 * a + i;
        int64_t pos = rowptr[row] + 3 * t;
        bool mr = mask[row] != 0;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            int64_t col = 3 * bcp + j;
            double v = (mr || mask[col]) ? (row == col ? 1.0 : 0.0) : s[3 * i + j];
            vals[pos + j] = (TV)v;
        }
    }
}

// Deterministic per-nonzero sum of the element tangents into the CSR values,
// with MatZeroRowsColumns(diag = 1) applied (src/assmcore_iga.c:763,844):
// one workgroup per owned control point a walks the (at most 27) elements
// containing a in ascending element order; for each it reads a's three rows
// of K_e (243 contiguous values, one per thread, all issued before the sums)
// and adds them into an LDS accumulator indexed by the stencil column point,
// with a barrier between elements.  Every entry gets its contributions in
// ascending element order (bitwise repeatable); the loads are whole-row,
// coalesced segments.  The row starts and the Dirichlet flags of the rows and
// stencil columns are loaded into LDS before the element rounds, so that the
// write-out does not wait on them.
template <typename TK, typename TV>
__global__ __launch_bounds__(256) void k_tangent_gather_rows(MeshK M, const TK *__restrict__ Ke, const int64_t *__restrict__ rowptr,
                                                             const int32_t *__restrict__ mask, TV *__restrict__ vals)
{
    __shared__ double acc[125 * 9];
    // control-point and dof indices fit in 32 bits (the CSR columns are int32)
    const int a = (int)blockIdx.x;
    const int tid = threadIdx.x;
    const int nb1 = M.nb[1], nb2 = M.nb[2];
    const int az = a % nb2, ay = (a / nb2) % nb1, ax = a / (nb1 * nb2);
    int Sx[5], Sy[5], Sz[5];
    const int nsx = stencil1(ax, M.nb[0], M.periodic[0], Sx);
    const int nsy = stencil1(ay, M.nb[1], M.periodic[1], Sy);
    const int nsz = stencil1(az, M.nb[2], M.periodic[2], Sz);
    int ex[3], lax[3], ey[3], lay[3], ez[3], laz[3], dummy[3];
    const int nx = elem_pairs2(ax, ax, M.nelem[0], M.nb[0], M.periodic[0], ex, lax, dummy);
    const int ny = elem_pairs2(ay, ay, M.nelem[1], M.nb[1], M.periodic[1], ey, lay, dummy);
    const int nz = elem_pairs2(az, az, M.nelem[2], M.nb[2], M.periodic[2], ez, laz, dummy);
    const int nst = nsx * nsy * nsz;
    for (int i = tid; i < nst * 9; i += 256) acc[i] = 0.0;
    // this thread's value of a's rows: row 3 la + i, column c = 3 lb + j
    const int i = tid / 81, c = tid - 81 * (tid / 81), lb = c / 3, j = c - 3 * (c / 3);
    const int lbx = lb / 9, lby = (lb / 3) % 3, lbz = lb % 3;
    // stencil slot of the thread's column point in element slot (x, y, z)
    int sxv[3] = {0, 0, 0}, syv[3] = {0, 0, 0}, szv[3] = {0, 0, 0};
#pragma unroll
    for (int x = 0; x < 3; x++) {
        if (x < nx) {
            const int bx = own1(ex[x] + lbx, M.nb[0], M.periodic[0]);
            for (int k = 0; k < nsx; k++) if (Sx[k] == bx) sxv[x] = k;
        }
        if (x < ny) {
            const int by = own1(ey[x] + lby, M.nb[1], M.periodic[1]);
            for (int k = 0; k < nsy; k++) if (Sy[k] == by) syv[x] = k;
        }
        if (x < nz) {
            const int bz = own1(ez[x] + lbz, M.nb[2], M.periodic[2]);
            for (int k = 0; k < nsz; k++) if (Sz[k] == bz) szv[x] = k;
        }
    }
    __shared__ int64_t s_rp[3];
    __shared__ int32_t s_mrow[3], s_mcol[125 * 3];
    double v[27];
#pragma unroll
    for (int x = 0; x < 3; x++)
#pragma unroll
        for (int y = 0; y < 3; y++)
#pragma unroll
            for (int z = 0; z < 3; z++) {
                double t = 0.0;
                if (tid < 243 && x < nx && y < ny && z < nz) {
                    const int64_t e = ((int64_t)ex[x] * M.nelem[1] + ey[y]) * M.nelem[2] + ez[z];
                    const int la = 9 * lax[x] + 3 * lay[y] + laz[z];
                    t = (double)Ke[e * 6561 + (3 * la + i) * 81 + c];
                }
                v[9 * x + 3 * y + z] = t;
            }
    // the output phase's row starts and Dirichlet flags, loaded now so that
    // their latency overlaps the element rounds
    for (int q = tid; q < 3 * nst + 6; q += 256) {
        if (q < 3 * nst) {
            const int t = q / 3, jj = q - 3 * t;
            const int sx = t / (nsy * nsz), sy = (t / nsz) % nsy, sz = t % nsz;
            const int bcp = (nb1 * nb2) * Sx[sx] + nb2 * Sy[sy] + Sz[sz];
            s_mcol[q] = mask[3 * bcp + jj];
        } else if (q < 3 * nst + 3) {
            s_rp[q - 3 * nst] = rowptr[3 * a + (q - 3 * nst)];
        } else {
            s_mrow[q - 3 * nst - 3] = mask[3 * a + (q - 3 * nst - 3)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int x = 0; x < 3; x++)
#pragma unroll
        for (int y = 0; y < 3; y++)
#pragma unroll
            for (int z = 0; z < 3; z++) {
                if (x < nx && y < ny && z < nz) {
                    if (tid < 243) {
                        const int t = (sxv[x] * nsy + syv[y]) * nsz + szv[z];
                        acc[(3 * t + i) * 3 + j] += v[9 * x + 3 * y + z];
                    }
                    __syncthreads();
                }
            }
    // CSR rows 3a+i: positions rowptr + 3t + j, consecutive threads on one row
    for (int idx = tid; idx < nst * 9; idx += 256) {
        const int ii = idx / (3 * nst), r = idx - ii * 3 * nst, t = r / 3, jj = r - 3 * t;
        const int sx = t / (nsy * nsz), sy = (t / nsz) % nsy, sz = t % nsz;
        const int bcp = (nb1 * nb2) * Sx[sx] + nb2 * Sy[sy] + Sz[sz];
        const int row = 3 * a + ii, col = 3 * bcp + jj;
        const double val = (s_mrow[ii] || s_mcol[3 * t + jj]) ? (row == col ? 1.0 : 0.0) : acc[(3 * t + ii) * 3 + jj];
        vals[s_rp[ii] + r] = (TV)val;
    }
}

__global__ __launch_bounds__(128) void k_csr_cols(MeshK M, const int64_t *__restrict__ rowptr, int32_t *__restrict__ col)
{
    const int64_t a = blockIdx.x;
    int nb1 = M.nb[1], nb2 = M.nb[2];
    int az = (int)(a % nb2), ay = (int)((a / nb2) % nb1), ax = (int)(a / ((int64_t)nb1 * nb2));
    int Sx[5], Sy[5], Sz[5];
    int nsx = stencil1(ax, M.nb[0], M.periodic[0], Sx);
    int nsy = stencil1(ay, M.nb[1], M.periodic[1], Sy);
    int nsz = stencil1(az, M.nb[2], M.periodic[2], Sz);
    int t = threadIdx.x;
    if (t >= nsx * nsy * nsz) return;
    int sx = t / (nsy * nsz), sy = (t / nsz) % nsy, sz = t % nsz;
    int64_t bcp = ((int64_t)M.nb[1] * M.nb[2]) * Sx[sx] + (int64_t)M.nb[2] * Sy[sy] + Sz[sz];
    for (int i = 0; i < 3; i++) {
        int64_t pos = rowptr[3 * a + i] + 3 * t;
        for (int j = 0; j < 3; j++) col[pos + j] = (int32_t)(3 * bcp + j);
    }
}

template <typename T>
__global__ void k_fill(T *p, T v, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = v;
}

__global__ void k_copy(double *dst, const double *src, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) dst[i] = src[i];
}

__global__ void k_selftest_mfma(double *o64, float *o32)
{
    int l = threadIdx.x;
    int i = l & 15, k = l >> 4;
    // A[i][k] = 1 + i + 0.25 k ; B[k][j] = 2 - 0.5 j + k
    double a = 1.0 + i + 0.25 * k;
    double b = 2.0 - 0.5 * i + k;   // lane holds B[k = l>>4][j = l&15]
    double4_t d = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, (double4_t){0, 0, 0, 0}, 0, 0, 0);
    float4_t f = __builtin_amdgcn_mfma_f32_16x16x4f32((float)a, (float)b, (float4_t){0, 0, 0, 0}, 0, 0, 0);
    // raw result registers: out[(r*64 + lane)]
    for (int r = 0; r < 4; r++) {
        o64[r * 64 + l] = d[r];
        o32[r * 64 + l] = f[r];
    }
}

// ============================================================================
// Pointwise physics at given jets: the device counterpart of the Phys
// callbacks residual / tangent / density (include/physimpl.h:6-9; gradelast,
// phys/gradelast.c; gradelasttime, phys/gradelasttime.c), through the same
// device functions the assembly kernels use (phys_state, phys_residual_acc,
// c_row, c_row_mom, phys_density_*).  Jet layout of one point:
// u[30 h + 10 k + a] with h = 0 the current level, 1 the previous and 2 the
// one before (the reference's u, u1 = u + 30, u0 = u + 60).  Thread t
// evaluates row (k, a) = t % 30 of point t / 30: r[30 p + row],
// C[900 p + 30 row + 10 l + b] and, on row 0, the densities hh[nd p + i].
// ============================================================================
// the moment state of the transient tangent (c_row_mom) for the jets
// Um = (u1 + u0)/2 and Ud = (u - u0)/2, summed over the rule nodes in order
__device__ void moment_state_seq(const double *Um, const double *Ud, const PhysK &P, const Mom &mu, double *ms)
{
    double S[9], G[9], n0[9], n1[9], n2[9];
    for (int i = 0; i < 9; i++) S[i] = G[i] = n0[i] = n1[i] = n2[i] = 0.0;
    for (int j = 0; j < P.nnode; j++) {
        double us[30];
        for (int i = 0; i < 30; i++) us[i] = Um[i] + P.s[j] * Ud[i];
        State st;
        phys_state(us, P, st);
        const double w0 = P.ct[j], w1 = w0 * P.s[j], w2 = w1 * P.s[j];
        for (int i = 0; i < 9; i++) {
            S[i] += w0 * st.S[i];
            G[i] += w0 * st.G[i];
            n0[i] += w0 * st.N[i];
            n1[i] += w1 * st.N[i];
            n2[i] += w2 * st.N[i];
        }
    }
    const double c12 = 1.0 / (mu.a11 * mu.a22), c22 = 1.0 / mu.d2;
    for (int i = 0; i < 9; i++) {
        ms[MS + i] = S[i];
        ms[MG + i] = G[i];
        ms[MN11 + i] = n0[i] / mu.m0;
        ms[MN12 + i] = (n1[i] - mu.lf * n0[i]) * c12;
        ms[MN22 + i] = ((mu.lf * n0[i] - n1[i]) * mu.lf + (n2[i] - mu.lf * n1[i])) * c22;
    }
    for (int i = 0; i < 36; i++) {
        const bool isF = i < 9;
        const int r = isF ? i : i - 9;
        const int slot = isF ? 10 * (r / 3) + 1 + (r % 3) : 10 * (r / 9) + sidx((r % 9) / 3, r % 3);
        double x0 = Um[slot];
        const double xd = Ud[slot];
        if (isF && (r / 3) == (r % 3)) x0 += 1.0;
        const double x1 = mu.a11 * (x0 + mu.lf * xd), x2 = mu.a22 * xd;
        if (isF) { ms[MF1 + r] = x1; ms[MF2 + r] = x2; }
        else { ms[MH1 + r] = x1; ms[MH2 + r] = x2; }
    }
}

template <int NH>
__global__ __launch_bounds__(128) void k_phys_eval(PhysK P, int64_t npts, const double *__restrict__ u, double *__restrict__ r,
                                                  double *__restrict__ C, double *__restrict__ hh)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npts * 30) return;
    const int64_t p = t / 30;
    const int row = (int)(t - 30 * p), k = row / 10, a = row % 10;
    const double *J = u + p * 30 * NH;
    if (r) {
        double rr[30];
        if (NH == 1) {
            phys_residual_static(P, J, rr);
        } else {
            for (int i = 0; i < 30; i++) rr[i] = 0.0;
            for (int g = 0; g < P.nnode; g++) {
                double us[30];
                for (int i = 0; i < 30; i++) us[i] = 0.5 * (J[30 + i] + J[60 + i]) + P.s[g] * (0.5 * (J[i] - J[60 + i]));
                State st;
                phys_state(us, P, st);
                phys_residual_acc(st, P.cr[g], rr);
            }
            for (int kk = 0; kk < 3; kk++)
                rr[10 * kk] = (P.cii * (J[10 * kk] - J[60 + 10 * kk])) / (2. * P.dt)
                            + (P.rho * (J[10 * kk] + J[60 + 10 * kk] - 2 * J[30 + 10 * kk])) / (P.dt * P.dt);
        }
        r[p * 30 + row] = rr[row];
    }
    if (C) {
        double out[30];
        for (int i = 0; i < 30; i++) out[i] = 0.0;
        if (NH == 1) {
            State st;
            phys_state(J, P, st);
            double sp[STW];
            for (int i = 0; i < 9; i++) sp[ST_F + i] = st.F[i];
            for (int i = 0; i < 27; i++) sp[ST_H + i] = st.H[i];
            for (int i = 0; i < 9; i++) { sp[ST_S + i] = st.S[i]; sp[ST_G + i] = st.G[i]; sp[ST_N + i] = st.N[i]; }
            for (int l = 0; l < 3; l++) c_row<0>(sp, P, k, a, l, 1.0, out + 10 * l);
        } else {
            Mom mu = {0.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0};
            for (int j = 0; j < P.nnode; j++) {
                mu.m0 += P.ct[j];
                mu.m1 += P.ct[j] * P.s[j];
                mu.m2 += P.ct[j] * P.s[j] * P.s[j];
            }
            mu.factor();
            double Um[30], Ud[30], ms[MW];
            for (int i = 0; i < 30; i++) {
                Um[i] = 0.5 * (J[30 + i] + J[60 + i]);
                Ud[i] = 0.5 * (J[i] - J[60 + i]);
            }
            moment_state_seq(Um, Ud, P, mu, ms);
            for (int l = 0; l < 3; l++) c_row_mom<0>(ms, P, k, a, l, out + 10 * l);
            if (a == 0) out[10 * k] += P.cii / (2. * P.dt) + P.rho / (P.dt * P.dt);
        }
        for (int i = 0; i < 30; i++) C[p * 900 + 30 * row + i] = out[i];
    }
    if (hh && row == 0) {
        if (NH == 1) {
            double h[ND_STATIC];
            phys_density_static(J, P, h);
            for (int i = 0; i < ND_STATIC; i++) hh[p * ND_STATIC + i] = h[i];
        } else {
            double h[ND_TRANSIENT];
            phys_density_transient(J, P, h);
            for (int i = 0; i < ND_TRANSIENT; i++) hh[p * ND_TRANSIENT + i] = h[i];
        }
    }
}

// ============================================================================
// launchers (host, C++ linkage; called by the C ABI in igap_capi.cpp)
// ============================================================================
}  // namespace igap

#include "igap_launch.h"

namespace igap {

hipError_t launch_density(const MeshK &M, const PhysK &P, const double *Uh, int64_t ndof, double *Hd, double *H, int nel,
                          hipStream_t s)
{
    dim3 grid(nel), block(256);
    if (P.torder == 0) hipLaunchKernelGGL(k_density_elem<1>, grid, block, 0, s, M, P, Uh, ndof, Hd, nel);
    else hipLaunchKernelGGL(k_density_elem<3>, grid, block, 0, s, M, P, Uh, ndof, Hd, nel);
    const int nd = P.torder == 0 ? ND_STATIC : ND_TRANSIENT;
    hipLaunchKernelGGL(k_density_reduce, dim3(nd), dim3(256), 0, s, nel, nd, Hd, H);
    return hipGetLastError();
}

hipError_t launch_residual_elem(const MeshK &M, const PhysK &P, const double *Uh, int64_t ndof, double *Re, int nel, hipStream_t s)
{
    dim3 grid(nel), block(256);
    if (P.torder == 0 && M.separable) hipLaunchKernelGGL(k_residual_sep, grid, block, 0, s, M, P, Uh, ndof, Re, nel);
    else if (P.torder == 2 && M.separable) hipLaunchKernelGGL(k_residual_sep_t, grid, block, 0, s, M, P, Uh, ndof, Re, nel);
    else if (P.torder == 0) hipLaunchKernelGGL(k_residual_elem<1>, grid, block, 0, s, M, P, Uh, ndof, Re, nel);
    else hipLaunchKernelGGL(k_residual_elem<3>, grid, block, 0, s, M, P, Uh, ndof, Re, nel);
    return hipGetLastError();
}

hipError_t launch_residual_gather(const MeshK &M, const double *Re, const double *Uh, const double *Fneu, const int32_t *mask,
                                  const double *g, double *R, int64_t ncp, hipStream_t s)
{
    dim3 grid((unsigned)((ncp + 255) / 256)), block(256);
    hipLaunchKernelGGL(k_residual_gather, grid, block, 0, s, M, Re, Uh, Fneu, mask, g, R, ncp);
    return hipGetLastError();
}

hipError_t launch_tangent_elem(const MeshK &M, const PhysK &P, const double *Uh, int64_t ndof, void *Ke, int prec, int nel,
                               hipStream_t s)
{
    dim3 grid(nel), block(256);
    const Scatter S{};
    if (P.torder == 0 && M.separable) {
        if (prec == 0) hipLaunchKernelGGL((k_tangent_sep<double, false>), grid, block, 0, s, M, P, Uh, ndof, (double *)Ke, nel, S);
        else hipLaunchKernelGGL((k_tangent_sep<float, false>), grid, block, 0, s, M, P, Uh, ndof, (float *)Ke, nel, S);
        return hipGetLastError();
    }
    if (P.torder == 2 && M.separable) {
        if (prec == 0) hipLaunchKernelGGL((k_tangent_sep_t<double, false>), grid, block, 0, s, M, P, Uh, ndof, (double *)Ke, nel, S);
        else hipLaunchKernelGGL((k_tangent_sep_t<float, false>), grid, block, 0, s, M, P, Uh, ndof, (float *)Ke, nel, S);
        return hipGetLastError();
    }
    if (prec == 0) {
        if (P.torder == 0) hipLaunchKernelGGL((k_tangent_elem<double, 1, false>), grid, block, 0, s, M, P, Uh, ndof, (double *)Ke, nel, S);
        else hipLaunchKernelGGL((k_tangent_elem<double, 3, false>), grid, block, 0, s, M, P, Uh, ndof, (double *)Ke, nel, S);
    } else {
        if (P.torder == 0) hipLaunchKernelGGL((k_tangent_elem<float, 1, false>), grid, block, 0, s, M, P, Uh, ndof, (float *)Ke, nel, S);
        else hipLaunchKernelGGL((k_tangent_elem<float, 3, false>), grid, block, 0, s, M, P, Uh, ndof, (float *)Ke, nel, S);
    }
    return hipGetLastError();
}

// Tangent straight into the CSR values of a non-periodic patch: zero the
// values, one element launch per colour (colours in a fixed order, so every
// entry sums its contributions in the same order on every call), then the
// Dirichlet diagonal.
hipError_t launch_tangent_direct(const MeshK &M, const PhysK &P, const double *Uh, int64_t ndof, const int64_t *rowptr,
                                 const int32_t *mask, void *vals, int prec, int64_t nnz, int64_t ncp, hipStream_t s)
{
    for (int d = 0; d < 3; d++)
        if (M.periodic[d]) return hipErrorInvalidValue;
    const int64_t fb = std::min<int64_t>((nnz + 255) / 256, 65536);
    if (prec == 0) hipLaunchKernelGGL(k_fill<double>, dim3((unsigned)fb), dim3(256), 0, s, (double *)vals, 0.0, nnz);
    else hipLaunchKernelGGL(k_fill<float>, dim3((unsigned)fb), dim3(256), 0, s, (float *)vals, 0.0f, nnz);
    for (int cx = 0; cx < 3; cx++)
        for (int cy = 0; cy < 3; cy++)
            for (int cz = 0; cz < 3; cz++) {
                Scatter S;
                S.rowptr = rowptr;
                S.mask = mask;
                S.c[0] = cx; S.c[1] = cy; S.c[2] = cz;
                for (int d = 0; d < 3; d++) S.nc[d] = (M.nelem[d] - S.c[d] + 2) / 3;
                const int64_t n = (int64_t)S.nc[0] * S.nc[1] * S.nc[2];
                if (n <= 0) continue;
                dim3 grid((unsigned)n), block(256);
                if (P.torder == 0 && M.separable) {
                    if (prec == 0) hipLaunchKernelGGL((k_tangent_sep<double, true>), grid, block, 0, s, M, P, Uh, ndof, (double *)vals, (int)n, S);
                    else hipLaunchKernelGGL((k_tangent_sep<float, true>), grid, block, 0, s, M, P, Uh, ndof, (float *)vals, (int)n, S);
                } else if (P.torder == 2 && M.separable) {
                    if (prec == 0) hipLaunchKernelGGL((k_tangent_sep_t<double, true>), grid, block, 0, s, M, P, Uh, ndof, (double *)vals, (int)n, S);
                    else hipLaunchKernelGGL((k_tangent_sep_t<float, true>), grid, block, 0, s, M, P, Uh, ndof, (float *)vals, (int)n, S);
                } else if (prec == 0) {
                    if (P.torder == 0) hipLaunchKernelGGL((k_tangent_elem<double, 1, true>), grid, block, 0, s, M, P, Uh, ndof, (double *)vals, (int)n, S);
                    else hipLaunchKernelGGL((k_tangent_elem<double, 3, true>), grid, block, 0, s, M, P, Uh, ndof, (double *)vals, (int)n, S);
                } else {
                    if (P.torder == 0) hipLaunchKernelGGL((k_tangent_elem<float, 1, true>), grid, block, 0, s, M, P, Uh, ndof, (float *)vals, (int)n, S);
                    else hipLaunchKernelGGL((k_tangent_elem<float, 3, true>), grid, block, 0, s, M, P, Uh, ndof, (float *)vals, (int)n, S);
                }
            }
    const dim3 g((unsigned)((ncp + 255) / 256)), b(256);
    if (prec == 0) hipLaunchKernelGGL(k_dirichlet_diag<double>, g, b, 0, s, M, rowptr, mask, (double *)vals, ncp);
    else hipLaunchKernelGGL(k_dirichlet_diag<float>, g, b, 0, s, M, rowptr, mask, (float *)vals, ncp);
    return hipGetLastError();
}

hipError_t launch_tangent_gather(const MeshK &M, const void *Ke, int prec, const int64_t *rowptr, const int32_t *mask, void *vals,
                                 int64_t ncp, hipStream_t s)
{
    dim3 grid((unsigned)ncp), block(256);
    if (prec == 0)
        hipLaunchKernelGGL((k_tangent_gather_rows<double, double>), grid, block, 0, s, M, (const double *)Ke, rowptr, mask, (double *)vals);
    else
        hipLaunchKernelGGL((k_tangent_gather<float, float>), grid, dim3(128), 0, s, M, (const float *)Ke, rowptr, mask, (float *)vals);
    return hipGetLastError();
}

hipError_t launch_csr_cols(const MeshK &M, const int64_t *rowptr, int32_t *col, int64_t ncp, hipStream_t s)
{
    dim3 grid((unsigned)ncp), block(128);
    hipLaunchKernelGGL(k_csr_cols, grid, block, 0, s, M, rowptr, col);
    return hipGetLastError();
}

hipError_t launch_fill_f64(double *p, double v, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill<double>, dim3((unsigned)blocks), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_copy_f64(double *dst, const double *src, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------- scratch
// The element vectors of igap_gpu_residual_noalloc live in this file-scope
// device array: the code object carries it, so every GPU the library runs on
// holds its own copy (256 MiB) until the process ends, and using it needs no
// allocation.  Host code must not pass the array's name to a kernel (that is
// a host-side stand-in): scratch_setup looks up each device's address once,
// with that device current, and the kernels take that address.
constexpr int64_t SCRATCH_DOUBLES = (int64_t)32 << 20;   // 81 per element: 414,252 elements
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];
constexpr int SCRATCH_MAX_DEVICES = 64;
static double *g_scratch_dev[SCRATCH_MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

hipError_t scratch_setup(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= SCRATCH_MAX_DEVICES) return hipErrorInvalidDevice;
    if (g_scratch_dev[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;   // the caller makes the device current
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[device] = static_cast<double *>(ptr);
    return hipSuccess;
}

double *scratch_device_ptr(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (device >= 0 && device < SCRATCH_MAX_DEVICES) ? g_scratch_dev[device] : nullptr;
}

int64_t scratch_capacity_doubles() { return SCRATCH_DOUBLES; }

hipError_t launch_phys_eval(const PhysK &P, int64_t npts, const double *u, double *r, double *C, double *hh, hipStream_t s)
{
    if (npts <= 0) return hipSuccess;
    const int64_t nt = 30 * npts;
    dim3 grid((unsigned)((nt + 127) / 128)), block(128);
    if (P.torder == 0) hipLaunchKernelGGL(k_phys_eval<1>, grid, block, 0, s, P, npts, u, r, C, hh);
    else hipLaunchKernelGGL(k_phys_eval<3>, grid, block, 0, s, P, npts, u, r, C, hh);
    return hipGetLastError();
}

hipError_t launch_selftest_mfma(double *o64, float *o32, hipStream_t s)
{
    hipLaunchKernelGGL(k_selftest_mfma, dim3(1), dim3(64), 0, s, o64, o32);
    return hipGetLastError();
}

}  // namespace igap
