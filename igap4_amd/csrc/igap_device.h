// igap_device.h -- device-side building blocks of the IGA assembly kernels.
//
// Per Gauss point: 1-D quadratic B-spline value/1st/2nd derivative with the
// MTOL-guarded denominators of eval_Bspline (src/meshcore_iga.c:243-327), the
// NURBS quotient rule and geometric map of Residual_iga
// (src/assmcore_iga.c:152-208), and the pointwise physics of
// phys/gradelast.c / phys/gradelasttime.c derived from the stored energy
//   Psi = c3 e26(e26^2-3e25^2) + c2 rho^2 + c4 rho^4 + e0 e24^2
//       + e345 (E12^2+E13^2+E23^2) + le^2 sum_K |dev(E_II,K)|^2
// (phys/gradelast.c:10-48,1068) by the chain rule through the strain
// measures E_IJ and EK_IK = dE_II/dX_K.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace igap {

constexpr double MTOL = 1.e-15;   // src/meshcore_iga.c:10
constexpr int NQ = 4;             // Gauss points per direction (src/assmcore_iga.c:113)

// ---------------------------------------------------------------- params
struct MeshK {
    int nelem[3];
    int na[3];        // active control points per direction (nelem + 2)
    int nb[3];        // owned control points (periodic: nelem)
    int periodic[3];
    const double *knot[3];
    const double *w;  // active numbering
    const double *X;  // 3 * active
    double par_periodic[9];
    double gq[4], gw[4];   // Gauss-Legendre nodes/weights on [0,1]
    int separable;    // x_d depends on xi_d alone and the weights are equal (k_tangent_sep applies)
};

struct PhysK {
    double c2, c3, c4, e0, e345, le;
    double rho, cii, dt;
    int torder;   // 0 static, 2 transient
    int nnode;    // 1 static, 8 transient
    double s[8], cr[8], ct[8];   // transient rule in s (nodes, residual, tangent weights)
};

// jet slot of the symmetric second derivative (I,K): 11,12,13,22,23,33 -> 4..9
__device__ __forceinline__ int sidx(int I, int K)
{
    int a = I < K ? I : K, b = I < K ? K : I;
    return (a == 0) ? 4 + b : (a == 1 ? 6 + b : 9);
}

__host__ __device__ __forceinline__ int own1(int a, int n, int periodic) { return periodic ? (a >= n ? a - n : a) : a; }

// ---------------------------------------------------------------- basis
// value/first/second derivative of the three quadratic B-splines of the span
// [kV[2], kV[3]] of the five-knot window kV at x
__device__ __forceinline__ void bspline1d(const double *kV, double x, double *v, double *d1, double *d2)
{
    double h23 = kV[3] - kV[2] + MTOL;
    double n11 = (kV[3] - x) / h23, n12 = (x - kV[2]) / h23;
    double i31 = 1.0 / (kV[3] - kV[1] + MTOL), i42 = 1.0 / (kV[4] - kV[2] + MTOL);
    // i = 0 uses N1[0] = 0, N1[1];  i = 1 uses N1[1], N1[2];  i = 2 uses N1[2], N1[3] = 0
    v[0] = (kV[3] - x) / (kV[3] - kV[1] + MTOL) * n11;
    v[1] = (x - kV[1]) / (kV[3] - kV[1] + MTOL) * n11 + (kV[4] - x) / (kV[4] - kV[2] + MTOL) * n12;
    v[2] = (x - kV[2]) / (kV[4] - kV[2] + MTOL) * n12;
    d1[0] = -2.0 * i31 * n11;
    d1[1] = 2.0 * i31 * n11 - 2.0 * i42 * n12;
    d1[2] = 2.0 * i42 * n12;
    double dn11 = -1.0 / h23, dn12 = 1.0 / h23;
    d2[0] = -2.0 * i31 * dn11;
    d2[1] = 2.0 * i31 * dn11 - 2.0 * i42 * dn12;
    d2[2] = 2.0 * i42 * dn12;
}

// tensor-product jet of basis b = 9i+3j+k from the 1-D tables of the three
// directions, t_d[3*order + i] (value, first, second derivative)
__device__ __forceinline__ void tensor_jet(const double *t0, const double *t1, const double *t2, int b, double *N)
{
    int i = b / 9, j = (b / 3) % 3, k = b % 3;
    double a0 = t0[i], b0 = t0[3 + i], c0 = t0[6 + i];
    double a1 = t1[j], b1 = t1[3 + j], c1 = t1[6 + j];
    double a2 = t2[k], b2 = t2[3 + k], c2 = t2[6 + k];
    N[0] = a0 * a1 * a2;
    N[1] = b0 * a1 * a2;
    N[2] = a0 * b1 * a2;
    N[3] = a0 * a1 * b2;
    N[4] = c0 * a1 * a2;
    N[5] = b0 * b1 * a2;
    N[6] = b0 * a1 * b2;
    N[7] = a0 * c1 * a2;
    N[8] = a0 * b1 * b2;
    N[9] = a0 * a1 * c2;
}

// NURBS quotient rule (src/assmcore_iga.c:153-165): R = N w / W to 2nd order
// One reciprocal of W replaces the reference's 25 divisions (a rounding-level
// difference); the update order of the slots is the reference's.
__device__ __forceinline__ void nurbs_param(const double *N, double wb, const double *W, double *R)
{
    const double iw = 1.0 / W[0], f = wb * iw;
    double w[10];
#pragma unroll
    for (int a = 1; a < 10; a++) w[a] = W[a] * iw;
#pragma unroll
    for (int a = 0; a < 10; a++) R[a] = N[a] * f;
#pragma unroll
    for (int a = 1; a < 10; a++) R[a] -= R[0] * w[a];
    R[4] -= R[1] * w[1] + R[1] * w[1];
    R[5] -= R[2] * w[1] + R[1] * w[2];
    R[6] -= R[3] * w[1] + R[1] * w[3];
    R[7] -= R[2] * w[2] + R[2] * w[2];
    R[8] -= R[3] * w[2] + R[2] * w[3];
    R[9] -= R[3] * w[3] + R[3] * w[3];
}

// parametric -> physical derivatives (src/assmcore_iga.c:179-208).
// Gi = inverse of G_ij = dx_i/dxi_j (row-major), X2[3*6] = d2x_d/dxi dxi.
__device__ __forceinline__ void phys_jet(double *R, const double *Gi, const double *X2)
{
    double p1 = R[1], p2 = R[2], p3 = R[3];
    double r1 = p1 * Gi[0] + p2 * Gi[3] + p3 * Gi[6];
    double r2 = p1 * Gi[1] + p2 * Gi[4] + p3 * Gi[7];
    double r3 = p1 * Gi[2] + p2 * Gi[5] + p3 * Gi[8];
    R[1] = r1; R[2] = r2; R[3] = r3;
    double n[6];
#pragma unroll
    for (int s = 0; s < 6; s++) n[s] = R[4 + s] - r1 * X2[s] - r2 * X2[6 + s] - r3 * X2[12 + s];
    // Hm = [[n0 n1 n2],[n1 n3 n4],[n2 n4 n5]];  R_ij = sum_pq Gi[p][i] Hm[p][q] Gi[q][j]
    const int pi_[6] = {0, 0, 0, 1, 1, 2}, pj_[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int s = 0; s < 6; s++) {
        int i = pi_[s], j = pj_[s];
        double c0 = n[0] * Gi[j] + n[1] * Gi[3 + j] + n[2] * Gi[6 + j];
        double c1 = n[1] * Gi[j] + n[3] * Gi[3 + j] + n[4] * Gi[6 + j];
        double c2 = n[2] * Gi[j] + n[4] * Gi[3 + j] + n[5] * Gi[6 + j];
        R[4 + s] = Gi[i] * c0 + Gi[3 + i] * c1 + Gi[6 + i] * c2;
    }
}

// Transpose of phys_jet (for fixed Gi, X2): given the adjoint o of the
// physical jet, returns in place the adjoint of the parametric jet.  The
// projection sum_a r_a R_a(b) of a residual jet onto a basis function is the
// pairing of the adjoint with the basis function's parametric jet.
__device__ __forceinline__ void phys_jet_adj(double *o, const double *Gi, const double *X2)
{
    const int pi_[6] = {0, 0, 0, 1, 1, 2}, pj_[6] = {0, 1, 2, 1, 2, 2};
    double nb[6];
#pragma unroll
    for (int t = 0; t < 6; t++) {
        const int p = pi_[t], q = pj_[t];
        double acc = 0.0;
#pragma unroll
        for (int s = 0; s < 6; s++) {
            const int i = pi_[s], j = pj_[s];
            double m = Gi[3 * p + i] * Gi[3 * q + j];
            if (p != q) m += Gi[3 * q + i] * Gi[3 * p + j];
            acc += o[4 + s] * m;
        }
        nb[t] = acc;
    }
    double rb[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double acc = o[1 + k];
#pragma unroll
        for (int t = 0; t < 6; t++) acc -= nb[t] * X2[6 * k + t];
        rb[k] = acc;
    }
#pragma unroll
    for (int p = 0; p < 3; p++) o[1 + p] = rb[0] * Gi[3 * p] + rb[1] * Gi[3 * p + 1] + rb[2] * Gi[3 * p + 2];
#pragma unroll
    for (int t = 0; t < 6; t++) o[4 + t] = nb[t];
}

// Transpose of nurbs_param with wb = 1 (for fixed W), in place.
__device__ __forceinline__ void nurbs_param_adj(double *R, const double *W)
{
    const double iw = 1.0 / W[0];
    double w[10];
#pragma unroll
    for (int a = 1; a < 10; a++) w[a] = W[a] * iw;
    R[1] -= 2.0 * w[1] * R[4] + w[2] * R[5] + w[3] * R[6];
    R[2] -= w[1] * R[5] + 2.0 * w[2] * R[7] + w[3] * R[8];
    R[3] -= w[1] * R[6] + w[2] * R[8] + 2.0 * w[3] * R[9];
    double s = 0.0;
#pragma unroll
    for (int a = 1; a < 10; a++) s += w[a] * R[a];
    R[0] -= s;
#pragma unroll
    for (int a = 0; a < 10; a++) R[a] *= iw;
}

__device__ __forceinline__ double inv3(const double *G, double *Gi)
{
    double det = G[0] * (G[4] * G[8] - G[5] * G[7]) - G[1] * (G[3] * G[8] - G[5] * G[6]) + G[2] * (G[3] * G[7] - G[4] * G[6]);
    double id = 1.0 / det;
    Gi[0] = id * (G[4] * G[8] - G[5] * G[7]);
    Gi[1] = id * (G[2] * G[7] - G[1] * G[8]);
    Gi[2] = id * (G[1] * G[5] - G[2] * G[4]);
    Gi[3] = id * (G[5] * G[6] - G[3] * G[8]);
    Gi[4] = id * (G[0] * G[8] - G[2] * G[6]);
    Gi[5] = id * (G[2] * G[3] - G[0] * G[5]);
    Gi[6] = id * (G[3] * G[7] - G[4] * G[6]);
    Gi[7] = id * (G[1] * G[6] - G[0] * G[7]);
    Gi[8] = id * (G[0] * G[4] - G[1] * G[3]);
    return det;
}

// ---------------------------------------------------------------- physics
// Strain-measure state at one jet: F[3k+I], H[9k+3I+K], S[3I+J] (S_II =
// dPsi/dE_II, S_IJ = dPsi/dE_IJ / 2), G[3I+K] = dPsi/dEK_IK, N[3I+J] =
// d2Psi/dE_II dE_JJ.
struct State {
    double F[9], H[27], S[9], G[9], N[9];
};

constexpr double IS2 = 0.70710678118654752440;
constexpr double IS3 = 0.57735026918962576451;
constexpr double IS6 = 0.40824829046386301637;

__device__ __forceinline__ void phys_state(const double *u, const PhysK &P, State &st)
{
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int I = 0; I < 3; I++) {
            st.F[3 * k + I] = (k == I ? 1.0 : 0.0) + u[10 * k + 1 + I];
#pragma unroll
            for (int K = 0; K < 3; K++) st.H[9 * k + 3 * I + K] = u[10 * k + sidx(I, K)];
        }
    double E[9], EK[9];
#pragma unroll
    for (int I = 0; I < 3; I++)
#pragma unroll
        for (int J = 0; J < 3; J++) {
            double s = st.F[I] * st.F[J] + st.F[3 + I] * st.F[3 + J] + st.F[6 + I] * st.F[6 + J];
            E[3 * I + J] = 0.5 * (s - (I == J ? 1.0 : 0.0));
        }
#pragma unroll
    for (int I = 0; I < 3; I++)
#pragma unroll
        for (int K = 0; K < 3; K++)
            EK[3 * I + K] = st.F[I] * st.H[3 * I + K] + st.F[3 + I] * st.H[9 + 3 * I + K] + st.F[6 + I] * st.H[18 + 3 * I + K];
    const double v24[3] = {IS3, IS3, IS3};
    const double v25[3] = {IS2, -IS2, 0.0};
    const double v26[3] = {IS6, IS6, -2 * IS6};
    double e24 = (E[0] + E[4] + E[8]) * IS3;
    double e25 = (E[0] - E[4]) * IS2;
    double e26 = (E[0] + E[4] - 2 * E[8]) * IS6;
    double rho2 = e25 * e25 + e26 * e26;
    double P25 = -6 * P.c3 * e25 * e26 + 2 * P.c2 * e25 + 4 * P.c4 * rho2 * e25;
    double P26 = 3 * P.c3 * (e26 * e26 - e25 * e25) + 2 * P.c2 * e26 + 4 * P.c4 * rho2 * e26;
    double P2525 = -6 * P.c3 * e26 + 2 * P.c2 + 4 * P.c4 * rho2 + 8 * P.c4 * e25 * e25;
    double P2526 = -6 * P.c3 * e25 + 8 * P.c4 * e25 * e26;
    double P2626 = 6 * P.c3 * e26 + 2 * P.c2 + 4 * P.c4 * rho2 + 8 * P.c4 * e26 * e26;
#pragma unroll
    for (int I = 0; I < 3; I++) {
        st.S[4 * I] = 2 * P.e0 * e24 * v24[I] + P25 * v25[I] + P26 * v26[I];
#pragma unroll
        for (int J = 0; J < 3; J++)
            st.N[3 * I + J] = 2 * P.e0 * v24[I] * v24[J] + P2525 * v25[I] * v25[J]
                            + P2526 * (v25[I] * v26[J] + v26[I] * v25[J]) + P2626 * v26[I] * v26[J];
    }
    st.S[1] = st.S[3] = P.e345 * E[1];
    st.S[2] = st.S[6] = P.e345 * E[2];
    st.S[5] = st.S[7] = P.e345 * E[5];
    double le2 = P.le * P.le;
#pragma unroll
    for (int K = 0; K < 3; K++) {
        double tr = (EK[K] + EK[3 + K] + EK[6 + K]) / 3.0;
#pragma unroll
        for (int I = 0; I < 3; I++) st.G[3 * I + K] = 2 * le2 * (EK[3 * I + K] - tr);
    }
}

// r += s * dPsi/du (jet slots 10k + a; value slots untouched)
__device__ __forceinline__ void phys_residual_acc(const State &st, double s, double *r)
{
#pragma unroll
    for (int k = 0; k < 3; k++) {
#pragma unroll
        for (int L = 0; L < 3; L++) {
            double v = st.S[3 * L] * st.F[3 * k] + st.S[3 * L + 1] * st.F[3 * k + 1] + st.S[3 * L + 2] * st.F[3 * k + 2];
            v += st.G[3 * L] * st.H[9 * k + 3 * L] + st.G[3 * L + 1] * st.H[9 * k + 3 * L + 1] + st.G[3 * L + 2] * st.H[9 * k + 3 * L + 2];
            r[10 * k + 1 + L] += s * v;
        }
#pragma unroll
        for (int I = 0; I < 3; I++)
#pragma unroll
            for (int K = I; K < 3; K++) {
                double v = st.G[3 * I + K] * st.F[3 * k + I];
                if (I != K) v += st.G[3 * K + I] * st.F[3 * k + K];
                r[10 * k + sidx(I, K)] += s * v;
            }
    }
}

// affine macro-field of par_periodic (src/assmcore_iga.c:225-229): the value
// gets a_kd * xq_d (knot-space coordinate), the first derivatives a_kd
// par_periodic's affine field on displacement component k (value and first
// derivative slots of one 10-slot jet)
__device__ __forceinline__ void add_periodic1(const MeshK &M, const double *xq, int k, double *u)
{
    u[0] += M.par_periodic[3 * k] * xq[0] + M.par_periodic[3 * k + 1] * xq[1] + M.par_periodic[3 * k + 2] * xq[2];
#pragma unroll
    for (int d = 0; d < 3; d++) u[1 + d] += M.par_periodic[3 * k + d];
}

// Last (x) step of the sum factorisation: the 10-slot parametric jet of one
// field at one Gauss point from its y,z-contracted values
// Bq[96 * bx + s6] (s6 = derivative pair (dy,dz) in the order (0,0) (1,0)
// (0,1) (2,0) (1,1) (0,2)) and the x table tx[3 * dx + bx].
__device__ __forceinline__ void sf_jet(const double *Bq, const double *tx, double *J)
{
    constexpr int DX[10] = {0, 1, 0, 0, 2, 1, 1, 0, 0, 0};
    constexpr int S6[10] = {0, 0, 1, 2, 0, 1, 2, 3, 4, 5};
#pragma unroll
    for (int a = 0; a < 10; a++)
        J[a] = Bq[S6[a]] * tx[3 * DX[a]] + Bq[96 + S6[a]] * tx[3 * DX[a] + 1] + Bq[192 + S6[a]] * tx[3 * DX[a] + 2];
}

// gradelast_residual: r = dPsi/du at the jet u[30]
__device__ __forceinline__ void phys_residual_static(const PhysK &P, const double *u, double *r)
{
#pragma unroll
    for (int i = 0; i < 30; i++) r[i] = 0.0;
    State st;
    phys_state(u, P, st);
    phys_residual_acc(st, 1.0, r);
}

// ---------------------------------------------------------------- densities
// gradelast_density (phys/gradelast.c:1025-1085): hh[0] the stored energy,
// hh[1..6] = 2 dPsi/dE_IJ (IJ = 11,12,13,22,23,33, each off-diagonal E_IJ
// taken once), hh[7..12] the phase indicators of the deviatoric strain
// (e25, e26) and the energy.
constexpr int ND_STATIC = 13, ND_TRANSIENT = 2;

__device__ __forceinline__ double phys_energy(const double *u, const PhysK &P, double &e25, double &e26)
{
    double F[9], E[9], EK[9];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int I = 0; I < 3; I++) F[3 * k + I] = (k == I ? 1.0 : 0.0) + u[10 * k + 1 + I];
#pragma unroll
    for (int I = 0; I < 3; I++)
#pragma unroll
        for (int J = 0; J < 3; J++)
            E[3 * I + J] = 0.5 * (F[I] * F[J] + F[3 + I] * F[3 + J] + F[6 + I] * F[6 + J] - (I == J ? 1.0 : 0.0));
#pragma unroll
    for (int I = 0; I < 3; I++)
#pragma unroll
        for (int K = 0; K < 3; K++)
            EK[3 * I + K] = F[I] * u[sidx(I, K)] + F[3 + I] * u[10 + sidx(I, K)] + F[6 + I] * u[20 + sidx(I, K)];
    const double e24 = (E[0] + E[4] + E[8]) * IS3;
    e25 = (E[0] - E[4]) * IS2;
    e26 = (E[0] + E[4] - 2 * E[8]) * IS6;
    const double rho2 = e25 * e25 + e26 * e26;
    double g2 = 0.0;
#pragma unroll
    for (int K = 0; K < 3; K++) {
        const double a = (EK[K] - EK[3 + K]) * IS2, b = (EK[K] + EK[3 + K] - 2 * EK[6 + K]) * IS6;
        g2 += a * a + b * b;
    }
    return P.c3 * e26 * (e26 * e26 - 3 * e25 * e25) + P.c2 * rho2 + P.c4 * rho2 * rho2 + P.e0 * e24 * e24
         + P.e345 * (E[1] * E[1] + E[2] * E[2] + E[5] * E[5]) + P.le * P.le * g2;
}

__device__ __forceinline__ void phys_density_static(const double *u, const PhysK &P, double *hh)
{
    double e25, e26;
    const double psi = phys_energy(u, P, e25, e26);
    State st;
    phys_state(u, P, st);
    hh[0] = psi;
    hh[1] = 2 * st.S[0];
    hh[2] = 4 * st.S[1];
    hh[3] = 4 * st.S[2];
    hh[4] = 2 * st.S[4];
    hh[5] = 4 * st.S[5];
    hh[6] = 2 * st.S[8];
    hh[7] = (double)(e25 > 0 && e26 > -0.5 * e25);
    hh[8] = (double)(e25 < 0 && e26 > 0.5 * e25);
    hh[9] = (double)(e26 < 0.5 * e25 && e26 < -0.5 * e25);
    hh[10] = (double)(psi < -0.5 && e25 > 0 && e26 > 0);
    hh[11] = (double)(psi < -0.5 && e25 < 0 && e26 > 0);
    hh[12] = (double)(psi < -0.5 && e26 < 0);
}

// gradelasttime_density (phys/gradelasttime.c:9183-9232): the energy at the
// midpoint of the two previous levels plus rho |u1 - u0|^2 / (2 dt^2), and 1.
// J = jets of the levels [u, u1, u0] (90 values).
__device__ __forceinline__ void phys_density_transient(const double *J, const PhysK &P, double *hh)
{
    double m[30];
#pragma unroll
    for (int i = 0; i < 30; i++) m[i] = 0.5 * (J[30 + i] + J[60 + i]);
    double e25, e26, kin = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double v = J[30 + 10 * k] - J[60 + 10 * k];
        kin += v * v;
    }
    hh[0] = phys_energy(m, P, e25, e26) + P.rho * kin / (2.0 * P.dt * P.dt);
    hh[1] = 1.0;
}

// ---------------------------------------------------------------- tangent entries
// The state as stored in LDS: F[0..8], H[9..35], S[36..44], G[45..53], N[54..62]
constexpr int ST_F = 0, ST_H = 9, ST_S = 36, ST_G = 45, ST_N = 54, STW = 63;
// Compact layout (48 values): F[9], the symmetric H_k,IK (= u_k,IK), S and N
// once per index pair (I <= K, in the slot order of sidx), G[9]
constexpr int STC_H = 9, STC_S = 27, STC_G = 33, STC_N = 42, STWC = 48;
template <bool CMP>
__device__ __forceinline__ int st_h(int k, int I, int K) { return CMP ? STC_H + 6 * k + sidx(I, K) - 4 : ST_H + 9 * k + 3 * I + K; }
template <bool CMP>
__device__ __forceinline__ int st_s(int I, int J) { return CMP ? STC_S + sidx(I, J) - 4 : ST_S + 3 * I + J; }
template <bool CMP>
__device__ __forceinline__ int st_g(int I, int K) { return CMP ? STC_G + 3 * I + K : ST_G + 3 * I + K; }
template <bool CMP>
__device__ __forceinline__ int st_n(int I, int J) { return CMP ? STC_N + sidx(I, J) - 4 : ST_N + 3 * I + J; }

// One row of the (k,l) block of C at one state, accumulated as row[b - A0] +=
// s * C[(k,a),(l,b)] for the slots b = A0..9.  The row slot a (1..9) is the
// lane's; all column slots are unrolled at compile time.  Value slots (a or
// b = 0) carry no stored-energy term.  CMP: the state in the compact layout.
template <int A0, bool CMP = false>
__device__ __forceinline__ void c_row(const double *st, const PhysK &P, int k, int a, int l, double s, double *row)
{
    const double *F = st + ST_F;
    const double g2 = 2 * P.le * P.le;
    const double Fl0 = F[3 * l], Fl1 = F[3 * l + 1], Fl2 = F[3 * l + 2];
    const double Fl[3] = {Fl0, Fl1, Fl2};
    if (a >= 1 && a <= 3) {
        const int L = a - 1;
        const double Fk0 = F[3 * k], Fk1 = F[3 * k + 1], Fk2 = F[3 * k + 2];
        const double FkL = F[3 * k + L], FlL = F[3 * l + L];
        const double hk0 = st[st_h<CMP>(k, L, 0)], hk1 = st[st_h<CMP>(k, L, 1)], hk2 = st[st_h<CMP>(k, L, 2)];
        const double fdot = Fk0 * Fl0 + Fk1 * Fl1 + Fk2 * Fl2;
        const bool dia = (k == l);
#pragma unroll
        for (int M = 0; M < 3; M++) {
            const double FkM = (M == 0) ? Fk0 : (M == 1 ? Fk1 : Fk2);
            double v = st[st_n<CMP>(L, M)] * FkL * Fl[M];
            v += 0.5 * P.e345 * ((M == L) ? (fdot - FkL * FlL) : FkM * FlL);
            const double hh = hk0 * st[st_h<CMP>(l, M, 0)] + hk1 * st[st_h<CMP>(l, M, 1)] + hk2 * st[st_h<CMP>(l, M, 2)];
            v += g2 * ((M == L) ? 2.0 / 3.0 : -1.0 / 3.0) * hh;
            if (dia) v += st[st_s<CMP>(L, M)];
            row[1 + M - A0] += s * v;
        }
        const double hk[3] = {hk0, hk1, hk2};
#pragma unroll
        for (int pp = 0; pp < 6; pp++) {
            constexpr int PI[6] = {0, 0, 0, 1, 1, 2}, PK[6] = {0, 1, 2, 1, 2, 2};
            const int I = PI[pp], K = PK[pp];
            double v = g2 * ((L == I) ? 2.0 / 3.0 : -1.0 / 3.0) * hk[K] * Fl[I];
            if (I != K) v += g2 * ((L == K) ? 2.0 / 3.0 : -1.0 / 3.0) * hk[I] * Fl[K];
            if (dia) {
                if (L == I) v += st[st_g<CMP>(I, K)];
                if (I != K && L == K) v += st[st_g<CMP>(K, I)];
            }
            row[4 + pp - A0] += s * v;
        }
    } else if (a >= 4) {
        constexpr int PI[6] = {0, 0, 0, 1, 1, 2}, PK[6] = {0, 1, 2, 1, 2, 2};
        const int I = PI[a - 4], K = PK[a - 4];
        const double FkI = F[3 * k + I], FkK = F[3 * k + K];
        const bool dia = (k == l), off = (I != K);
        // first-gradient columns: C[(k,IK),(l,M)] = C[(l,M),(k,IK)]
#pragma unroll
        for (int M = 0; M < 3; M++) {
            double v = g2 * ((M == I) ? 2.0 / 3.0 : -1.0 / 3.0) * st[st_h<CMP>(l, M, K)] * FkI;
            if (off) v += g2 * ((M == K) ? 2.0 / 3.0 : -1.0 / 3.0) * st[st_h<CMP>(l, M, I)] * FkK;
            if (dia) {
                if (M == I) v += st[st_g<CMP>(I, K)];
                if (off && M == K) v += st[st_g<CMP>(K, I)];
            }
            row[1 + M - A0] += s * v;
        }
        // second-gradient columns (J,K2): pairs (I,K),(K,I) against (J,K2),(K2,J)
#pragma unroll
        for (int pp = 0; pp < 6; pp++) {
            constexpr int PJ[6] = {0, 0, 0, 1, 1, 2}, PK2[6] = {0, 1, 2, 1, 2, 2};
            const int J = PJ[pp], K2 = PK2[pp];
            double v = 0.0;
            if (K == K2) v += ((I == J) ? 2.0 / 3.0 : -1.0 / 3.0) * FkI * Fl[J];
            if (J != K2 && K == J) v += ((I == K2) ? 2.0 / 3.0 : -1.0 / 3.0) * FkI * Fl[K2];
            if (off && I == K2) v += ((K == J) ? 2.0 / 3.0 : -1.0 / 3.0) * FkK * Fl[J];
            if (off && J != K2 && I == J) v += ((K == K2) ? 2.0 / 3.0 : -1.0 / 3.0) * FkK * Fl[K2];
            row[4 + pp - A0] += s * g2 * v;
        }
    }
}

// c_row for the separable kernel (slots 1..9, row[b - 1] = C[(k,a),(l,b)]):
// the same entries with the lane's runtime slot indices turned into one-hot
// vectors and the 2/3 / -1/3 coefficients into per-lane selects, so that
// every column is straight-line arithmetic at compile-time register
// indices, and the second-gradient columns of a second-gradient row come
// from the 3x3 tensor
//   T[J'][K''] = F_lJ' (d(K'',K) F_kI c(I,J') + [I != K] d(K'',I) F_kK c(K,J')),
//   C[(k,IK),(l,JK2)] = g2 (T[J][K2] + [J != K2] T[K2][J]),  c(x,y) = d(x,y) - 1/3.
template <bool CMP>
__device__ __forceinline__ void c_row_onehot(const double *st, const PhysK &P, int k, int a, int l, double *row)
{
    const double g2 = 2 * P.le * P.le;
    const double dd = (k == l) ? 1.0 : 0.0;
    const double *F = st + ST_F;
    double Fl[3], Fk[3];
#pragma unroll
    for (int m = 0; m < 3; m++) { Fl[m] = F[3 * l + m]; Fk[m] = F[3 * k + m]; }
    constexpr double C1 = 2.0 / 3.0, C0 = -1.0 / 3.0;
    if (a <= 3) {
        const int L = a - 1;
        double eL[3], cL[3];
#pragma unroll
        for (int m = 0; m < 3; m++) { eL[m] = (m == L) ? 1.0 : 0.0; cL[m] = (m == L) ? C1 : C0; }
        const double FkL = F[3 * k + L], FlL = F[3 * l + L];
        const double hk[3] = {st[st_h<CMP>(k, L, 0)], st[st_h<CMP>(k, L, 1)], st[st_h<CMP>(k, L, 2)]};
        const double fdot = Fk[0] * Fl[0] + Fk[1] * Fl[1] + Fk[2] * Fl[2];
        const double e1 = 0.5 * P.e345 * (fdot - FkL * FlL), e2 = 0.5 * P.e345 * FlL;
#pragma unroll
        for (int M = 0; M < 3; M++) {
            double v = st[st_n<CMP>(L, M)] * FkL * Fl[M];
            v += eL[M] * e1 + (1.0 - eL[M]) * (e2 * Fk[M]);
            const double hh = hk[0] * st[st_h<CMP>(l, M, 0)] + hk[1] * st[st_h<CMP>(l, M, 1)] + hk[2] * st[st_h<CMP>(l, M, 2)];
            v += g2 * cL[M] * hh;
            v += dd * st[st_s<CMP>(L, M)];
            row[M] = v;
        }
#pragma unroll
        for (int pp = 0; pp < 6; pp++) {
            constexpr int PI[6] = {0, 0, 0, 1, 1, 2}, PK[6] = {0, 1, 2, 1, 2, 2};
            const int I = PI[pp], K = PK[pp];
            double v = g2 * cL[I] * hk[K] * Fl[I];
            double gv = eL[I] * st[st_g<CMP>(I, K)];
            if (I != K) {
                v += g2 * cL[K] * hk[I] * Fl[K];
                gv += eL[K] * st[st_g<CMP>(K, I)];
            }
            row[3 + pp] = v + dd * gv;
        }
    } else {
        const int s = a - 4;
        const int I = (s < 3) ? 0 : (s < 5 ? 1 : 2);
        const int K = (s < 3) ? s : (s < 5 ? s - 2 : 2);
        const double nIK = (I != K) ? 1.0 : 0.0;
        double eI[3], eK[3], cI[3], cK[3];
#pragma unroll
        for (int m = 0; m < 3; m++) {
            eI[m] = (m == I) ? 1.0 : 0.0;
            eK[m] = (m == K) ? 1.0 : 0.0;
            cI[m] = (m == I) ? C1 : C0;
            cK[m] = (m == K) ? C1 : C0;
        }
        const double FkI = F[3 * k + I], FkK = F[3 * k + K];
        const double GIK = st[st_g<CMP>(I, K)], GKI = st[st_g<CMP>(K, I)];
#pragma unroll
        for (int M = 0; M < 3; M++) {
            double v = g2 * cI[M] * st[st_h<CMP>(l, M, K)] * FkI;
            v += nIK * (g2 * cK[M] * st[st_h<CMP>(l, M, I)] * FkK);
            v += dd * (eI[M] * GIK + nIK * eK[M] * GKI);
            row[M] = v;
        }
        const double f1 = FkI, f2 = nIK * FkK;
        double T[3][3];
#pragma unroll
        for (int J = 0; J < 3; J++)
#pragma unroll
            for (int K2 = 0; K2 < 3; K2++) T[J][K2] = Fl[J] * (eK[K2] * f1 * cI[J] + eI[K2] * f2 * cK[J]);
#pragma unroll
        for (int pp = 0; pp < 6; pp++) {
            constexpr int PJ[6] = {0, 0, 0, 1, 1, 2}, PK2[6] = {0, 1, 2, 1, 2, 2};
            const int J = PJ[pp], K2 = PK2[pp];
            double v = T[J][K2];
            if (J != K2) v += T[K2][J];
            row[3 + pp] = g2 * v;
        }
    }
}

// The transient rule weights the pointwise tangent of eight states
// us_j = m0 + s_j d.  F and H are linear in the jet, F_j = F0 + s_j Fd and
// H_j = H0 + s_j Hd, so every product of two of them sums over the nodes to
// the quadratic form
//   sum_j ct_j X_j Y_j = (X0, Xd) M (Y0, Yd)^T,  M = [[mu0, mu1], [mu1, mu2]]
// (mu_p = sum_j ct_j s_j^p), and the N-weighted ones to (X0, Xd) Nm (Y0, Yd)^T
// with Nm = [[N_0, N_1], [N_1, N_2]], N_p = sum_j ct_j s_j^p N_j; S and G
// enter linearly (Sb, Gb).  The rule integrates 1, s and s^2 exactly against
// the positive weight s/2 on [0, 1], so M (= [[1/4, 1/6], [1/6, 1/8]] to
// rounding) is positive definite; with its Cholesky factor M = L L^T and
// X^ = L^T (X0, Xd) the first form is X^_1 Y^_1 + X^_2 Y^_2 and the second
// X^^T (L^-1 Nm L^-T) Y^.  The moment state keeps the transformed quantities:
//   MF1, MF2: F^ (9 each)      MH1, MH2: H^ (27 each)
//   MS, MG:   Sb, Gb           MN11, MN12, MN22: L^-1 Nm L^-T (9 each)
constexpr int MF1 = 0, MF2 = 9, MH1 = 18, MH2 = 45, MS = 72, MG = 81, MN11 = 90, MN12 = 99, MN22 = 108, MW = 117;

// Cholesky factor of M: L = [[a11, 0], [a11 lf, a22]], so that
// X^_1 = a11 (X0 + lf Xd), X^_2 = a22 Xd
struct Mom {
    double m0, m1, m2;
    double lf, a11, a22, d2;
    __device__ __forceinline__ void factor()
    {
        lf = m1 / m0;
        a11 = sqrt(m0);
        d2 = m2 - m1 * lf;
        a22 = sqrt(d2);
    }
};

__device__ __forceinline__ double dot2(double x1, double x2, double y1, double y2) { return x1 * y1 + x2 * y2; }

// c_row for the rule-weighted transient tangent, from the transformed moment
// state ms
template <int A0>
__device__ __forceinline__ void c_row_mom(const double *ms, const PhysK &P, int k, int a, int l, double *row)
{
    const double *F1 = ms + MF1, *F2 = ms + MF2, *H1 = ms + MH1, *H2 = ms + MH2;
    const double g2 = 2 * P.le * P.le;
    const bool dia = (k == l);
    if (a >= 1 && a <= 3) {
        const int L = a - 1;
        const double FkL1 = F1[3 * k + L], FkL2 = F2[3 * k + L];
#pragma unroll
        for (int M = 0; M < 3; M++) {
            const double Fl1 = F1[3 * l + M], Fl2 = F2[3 * l + M];
            double v = ms[MN11 + 3 * L + M] * FkL1 * Fl1 + ms[MN12 + 3 * L + M] * (FkL1 * Fl2 + FkL2 * Fl1)
                     + ms[MN22 + 3 * L + M] * FkL2 * Fl2;
            double e = 0.0;
            if (M == L) {
#pragma unroll
                for (int J = 0; J < 3; J++)
                    if (J != L) e += dot2(F1[3 * k + J], F2[3 * k + J], F1[3 * l + J], F2[3 * l + J]);
            } else {
                e = dot2(F1[3 * k + M], F2[3 * k + M], F1[3 * l + L], F2[3 * l + L]);
            }
            v += 0.5 * P.e345 * e;
            double hh = 0.0;
#pragma unroll
            for (int c = 0; c < 3; c++)
                hh += dot2(H1[9 * k + 3 * L + c], H2[9 * k + 3 * L + c], H1[9 * l + 3 * M + c], H2[9 * l + 3 * M + c]);
            v += g2 * ((M == L) ? 2.0 / 3.0 : -1.0 / 3.0) * hh;
            if (dia) v += ms[MS + 3 * L + M];
            row[1 + M - A0] += v;
        }
#pragma unroll
        for (int pp = 0; pp < 6; pp++) {
            constexpr int PI[6] = {0, 0, 0, 1, 1, 2}, PK[6] = {0, 1, 2, 1, 2, 2};
            const int I = PI[pp], K = PK[pp];
            double v = g2 * ((L == I) ? 2.0 / 3.0 : -1.0 / 3.0)
                     * dot2(H1[9 * k + 3 * L + K], H2[9 * k + 3 * L + K], F1[3 * l + I], F2[3 * l + I]);
            if (I != K)
                v += g2 * ((L == K) ? 2.0 / 3.0 : -1.0 / 3.0)
                   * dot2(H1[9 * k + 3 * L + I], H2[9 * k + 3 * L + I], F1[3 * l + K], F2[3 * l + K]);
            if (dia) {
                if (L == I) v += ms[MG + 3 * I + K];
                if (I != K && L == K) v += ms[MG + 3 * K + I];
            }
            row[4 + pp - A0] += v;
        }
    } else if (a >= 4) {
        constexpr int PI[6] = {0, 0, 0, 1, 1, 2}, PK[6] = {0, 1, 2, 1, 2, 2};
        const int I = PI[a - 4], K = PK[a - 4];
        const bool off = (I != K);
        const double FkI1 = F1[3 * k + I], FkI2 = F2[3 * k + I], FkK1 = F1[3 * k + K], FkK2 = F2[3 * k + K];
#pragma unroll
        for (int M = 0; M < 3; M++) {
            double v = g2 * ((M == I) ? 2.0 / 3.0 : -1.0 / 3.0) * dot2(H1[9 * l + 3 * M + K], H2[9 * l + 3 * M + K], FkI1, FkI2);
            if (off) v += g2 * ((M == K) ? 2.0 / 3.0 : -1.0 / 3.0) * dot2(H1[9 * l + 3 * M + I], H2[9 * l + 3 * M + I], FkK1, FkK2);
            if (dia) {
                if (M == I) v += ms[MG + 3 * I + K];
                if (off && M == K) v += ms[MG + 3 * K + I];
            }
            row[1 + M - A0] += v;
        }
#pragma unroll
        for (int pp = 0; pp < 6; pp++) {
            constexpr int PJ[6] = {0, 0, 0, 1, 1, 2}, PK2[6] = {0, 1, 2, 1, 2, 2};
            const int J = PJ[pp], K2 = PK2[pp];
            double v = 0.0;
            if (K == K2) v += ((I == J) ? 2.0 / 3.0 : -1.0 / 3.0) * dot2(FkI1, FkI2, F1[3 * l + J], F2[3 * l + J]);
            if (J != K2 && K == J) v += ((I == K2) ? 2.0 / 3.0 : -1.0 / 3.0) * dot2(FkI1, FkI2, F1[3 * l + K2], F2[3 * l + K2]);
            if (off && I == K2) v += ((K == J) ? 2.0 / 3.0 : -1.0 / 3.0) * dot2(FkK1, FkK2, F1[3 * l + J], F2[3 * l + J]);
            if (off && J != K2 && I == J) v += ((K == K2) ? 2.0 / 3.0 : -1.0 / 3.0) * dot2(FkK1, FkK2, F1[3 * l + K2], F2[3 * l + K2]);
            row[4 + pp - A0] += g2 * v;
        }
    }
}

// Compact moment state (87 values; the separable transient tangent keeps one
// per Gauss point of the element): F^ (9 + 9), H^ once per symmetric index
// pair (18 + 18, H_k,IK = u_k,IK), Sb and the three N blocks once per
// symmetric pair (6 each), Gb (9).  Pair slots follow sidx.
constexpr int MC_F1 = 0, MC_F2 = 9, MC_H1 = 18, MC_H2 = 36, MC_S = 54, MC_G = 60, MC_N11 = 69, MC_N12 = 75, MC_N22 = 81, MWC = 87;

// c_row_mom (slots 1..9, row[b - 1] = C[(k,a),(l,b)]) from the compact moment
// state, stored in TS, with c_row_onehot's straight-line columns: the lane's
// slot indices as one-hot vectors, the products of two moment-weighted fields as
// dot2 pairs, the second-gradient block of a second-gradient row from the
// 3x3 tensor T[J][K2] = e_K[K2] c(I,J) d1[J] + e_I[K2] c(K,J) d2[J] with
// d1[J] = (F_kI . F_lJ), d2[J] = [I != K] (F_kK . F_lJ).
template <typename TS, bool CMP = true>
__device__ __forceinline__ void c_row_momc_onehot(const TS *ms, const PhysK &P, int k, int a, int l, double *row)
{
    // CMP: the compact layout (MC_*); otherwise the full one (M*) of k_tangent_elem
    constexpr int OF1 = CMP ? MC_F1 : MF1, OF2 = CMP ? MC_F2 : MF2, OH1 = CMP ? MC_H1 : MH1, OH2 = CMP ? MC_H2 : MH2;
    constexpr int OS = CMP ? MC_S : MS, OG = CMP ? MC_G : MG;
    constexpr int ON11 = CMP ? MC_N11 : MN11, ON12 = CMP ? MC_N12 : MN12, ON22 = CMP ? MC_N22 : MN22;
    auto F1 = [&](int i) { return (double)ms[OF1 + i]; };
    auto F2 = [&](int i) { return (double)ms[OF2 + i]; };
    auto hidx = [&](int kk, int I, int K) { return CMP ? 6 * kk + sidx(I, K) - 4 : 9 * kk + 3 * I + K; };
    auto H1 = [&](int kk, int I, int K) { return (double)ms[OH1 + hidx(kk, I, K)]; };
    auto H2 = [&](int kk, int I, int K) { return (double)ms[OH2 + hidx(kk, I, K)]; };
    auto sym = [&](int base, int I, int K) { return (double)ms[base + (CMP ? sidx(I, K) - 4 : 3 * I + K)]; };
    const double g2 = 2 * P.le * P.le;
    const double dd = (k == l) ? 1.0 : 0.0;
    double Fl1[3], Fl2[3], Fk1[3], Fk2[3];
#pragma unroll
    for (int m = 0; m < 3; m++) {
        Fl1[m] = F1(3 * l + m);
        Fl2[m] = F2(3 * l + m);
        Fk1[m] = F1(3 * k + m);
        Fk2[m] = F2(3 * k + m);
    }
    constexpr double C1 = 2.0 / 3.0, C0 = -1.0 / 3.0;
    if (a <= 3) {
        const int L = a - 1;
        double eL[3], cL[3];
#pragma unroll
        for (int m = 0; m < 3; m++) { eL[m] = (m == L) ? 1.0 : 0.0; cL[m] = (m == L) ? C1 : C0; }
        const double FkL1 = F1(3 * k + L), FkL2 = F2(3 * k + L), FlL1 = F1(3 * l + L), FlL2 = F2(3 * l + L);
        double hk1[3], hk2[3];
#pragma unroll
        for (int c = 0; c < 3; c++) { hk1[c] = H1(k, L, c); hk2[c] = H2(k, L, c); }
        const double full = dot2(Fk1[0], Fk2[0], Fl1[0], Fl2[0]) + dot2(Fk1[1], Fk2[1], Fl1[1], Fl2[1]) + dot2(Fk1[2], Fk2[2], Fl1[2], Fl2[2]);
        const double eLL = full - dot2(FkL1, FkL2, FlL1, FlL2);
#pragma unroll
        for (int M = 0; M < 3; M++) {
            double v = sym(ON11, L, M) * FkL1 * Fl1[M] + sym(ON12, L, M) * (FkL1 * Fl2[M] + FkL2 * Fl1[M])
                     + sym(ON22, L, M) * FkL2 * Fl2[M];
            const double e = eL[M] * eLL + (1.0 - eL[M]) * dot2(Fk1[M], Fk2[M], FlL1, FlL2);
            v += 0.5 * P.e345 * e;
            double hh = 0.0;
#pragma unroll
            for (int c = 0; c < 3; c++) hh += dot2(hk1[c], hk2[c], H1(l, M, c), H2(l, M, c));
            v += g2 * cL[M] * hh;
            v += dd * sym(OS, L, M);
            row[M] = v;
        }
#pragma unroll
        for (int pp = 0; pp < 6; pp++) {
            constexpr int PI[6] = {0, 0, 0, 1, 1, 2}, PK[6] = {0, 1, 2, 1, 2, 2};
            const int I = PI[pp], K = PK[pp];
            double v = g2 * cL[I] * dot2(hk1[K], hk2[K], Fl1[I], Fl2[I]);
            double gv = eL[I] * (double)ms[OG + 3 * I + K];
            if (I != K) {
                v += g2 * cL[K] * dot2(hk1[I], hk2[I], Fl1[K], Fl2[K]);
                gv += eL[K] * (double)ms[OG + 3 * K + I];
            }
            row[3 + pp] = v + dd * gv;
        }
    } else {
        const int s = a - 4;
        const int I = (s < 3) ? 0 : (s < 5 ? 1 : 2);
        const int K = (s < 3) ? s : (s < 5 ? s - 2 : 2);
        const double nIK = (I != K) ? 1.0 : 0.0;
        double eI[3], eK[3], cI[3], cK[3];
#pragma unroll
        for (int m = 0; m < 3; m++) {
            eI[m] = (m == I) ? 1.0 : 0.0;
            eK[m] = (m == K) ? 1.0 : 0.0;
            cI[m] = (m == I) ? C1 : C0;
            cK[m] = (m == K) ? C1 : C0;
        }
        const double FkI1 = F1(3 * k + I), FkI2 = F2(3 * k + I), FkK1 = F1(3 * k + K), FkK2 = F2(3 * k + K);
        const double GIK = (double)ms[OG + 3 * I + K], GKI = (double)ms[OG + 3 * K + I];
#pragma unroll
        for (int M = 0; M < 3; M++) {
            double v = g2 * cI[M] * dot2(H1(l, M, K), H2(l, M, K), FkI1, FkI2);
            v += nIK * (g2 * cK[M] * dot2(H1(l, M, I), H2(l, M, I), FkK1, FkK2));
            v += dd * (eI[M] * GIK + nIK * eK[M] * GKI);
            row[M] = v;
        }
        double d1[3], d2[3];
#pragma unroll
        for (int J = 0; J < 3; J++) {
            d1[J] = dot2(FkI1, FkI2, Fl1[J], Fl2[J]);
            d2[J] = nIK * dot2(FkK1, FkK2, Fl1[J], Fl2[J]);
        }
        double T[3][3];
#pragma unroll
        for (int J = 0; J < 3; J++)
#pragma unroll
            for (int K2 = 0; K2 < 3; K2++) T[J][K2] = eK[K2] * cI[J] * d1[J] + eI[K2] * cK[J] * d2[J];
#pragma unroll
        for (int pp = 0; pp < 6; pp++) {
            constexpr int PJ[6] = {0, 0, 0, 1, 1, 2}, PK2[6] = {0, 1, 2, 1, 2, 2};
            const int J = PJ[pp], K2 = PK2[pp];
            double v = T[J][K2];
            if (J != K2) v += T[K2][J];
            row[3 + pp] = g2 * v;
        }
    }
}

__constant__ const int kPairI[6] = {0, 0, 0, 1, 1, 2};
__constant__ const int kPairK[6] = {0, 1, 2, 1, 2, 2};

}  // namespace igap
