/* igacpu.c -- CPU restatement of IGAP4's residual/tangent assembly.
 *
 * TEST INFRASTRUCTURE ONLY: the oracle the HIP path is checked against and
 * the CPU baseline timed beside it.  Nothing in the product links or loads it.
 *
 * Restates, with this project's own code:
 *   Residual_iga / Tangent_iga          src/assmcore_iga.c:16-502, 507-858
 *   dirichlet_iga / neumann_default_iga src/assmcore_iga.c:861-934, 1036-1052
 *   setnz_iga                           src/assmcore_iga.c:1054-1087
 *   eval_Bspline / evalN                src/meshcore_iga.c:243-327, 329-384
 *   matinv (3x3)                        src/mathutil.c:30-41
 *   legendre_handle (n=4)               src/quadrature.c:381-429
 *   igabeam / igabeamtest geometry      mesh/igabeam.c:34-72, mesh/igabeamtest.c:68-70
 * and derives the pointwise physics of phys/gradelast.c / phys/gradelasttime.c
 * from the stored energy by the chain rule (see oc_static_state).
 *
 * The volume loops keep the reference's work profile: the per-Gauss-point
 * naive 81x81x10x10 tangent contraction (assmcore_iga.c:723-749) and a
 * per-element insertion into CSR with a per-row search (the MatSetValues ADD
 * of assmcore_iga.c:763).  OpenMP threads split each element colour (see
 * colour1: elements of one colour share no control point), so the inserts of
 * one colour never touch the same row.
 */
#include "igacpu.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MTOL 1.e-15   /* src/meshcore_iga.c:10 */
#define NQ 4          /* src/assmcore_iga.c:113 */

/* ========================================================================== */
/* physics                                                                     */
/* ========================================================================== */

void oc_coef_from_par(const double *par, double *coef)
{
    /* closed forms of phys/gradelast.c:46-48 */
    double r = par[0], cu = par[1];
    double a = 1.3333333333333333 * r - 1.3333333333333333 * r * cu;
    double D = r * r - 2 * r * r * cu - r * a;
    coef[0] = (2 * cu) / D;            /* c2 */
    coef[1] = -(a / (r * r * D));      /* c3 */
    coef[2] = -(1 / (r * r * D));      /* c4 */
    coef[3] = par[2];                  /* e0 */
    coef[4] = par[3];                  /* e345 */
    coef[5] = par[4];                  /* le */
}

void oc_phys_init(oc_phys *ph, int torder, const double *par9, const double *coef_or_null)
{
    ph->torder = torder;
    if (coef_or_null) memcpy(ph->coef, coef_or_null, 6 * sizeof(double));
    else oc_coef_from_par(par9, ph->coef);
    ph->rho = par9[5];
    ph->cii = par9[6];
    ph->dt = par9[7];
}

/* jet slot of the symmetric second derivative (I,K): 11,12,13,22,23,33 -> 4..9
 * (src/meshcore_iga.c:316-325) */
static const int SIDX[3][3] = {{4, 5, 6}, {5, 7, 8}, {6, 8, 9}};

static const double IS2 = 0.70710678118654752440;  /* 1/sqrt(2) */
static const double IS3 = 0.57735026918962576451;  /* 1/sqrt(3) */
static const double IS6 = 0.40824829046386301637;  /* 1/sqrt(6) */

/* Derivatives of Psi with respect to the strain measures at one state.
 *   E_IJ = (F_kI F_kJ - d_IJ)/2,  EK_IK = d(E_II)/dX_K = F_kI H_kIK,
 *   F_kI = d_kI + u_k,I,  H_kIK = u_k,IK.
 * Outputs: S (symmetric; S_II = dPsi/dE_II, S_IJ = dPsi/dE_IJ / 2),
 *          G_IK = dPsi/dEK_IK, N = d2Psi/dE_nn dE_nn (normal block). */
typedef struct {
    double F[3][3];
    double H[3][3][3];
    double S[3][3];
    double G[3][3];
    double N[3][3];
    double e345x2, le2x2;   /* shear and gradient Hessian scales */
    double psi;
} oc_state;

static void oc_static_state(const double *u, const double *c, oc_state *st)
{
    const double c2 = c[0], c3 = c[1], c4 = c[2], e0 = c[3], e345 = c[4], le = c[5];
    double E[3][3], EK[3][3];
    for (int k = 0; k < 3; k++)
        for (int I = 0; I < 3; I++) {
            st->F[k][I] = (k == I ? 1.0 : 0.0) + u[10 * k + 1 + I];
            for (int K = 0; K < 3; K++) st->H[k][I][K] = u[10 * k + SIDX[I][K]];
        }
    for (int I = 0; I < 3; I++)
        for (int J = 0; J < 3; J++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += st->F[k][I] * st->F[k][J];
            E[I][J] = 0.5 * (s - (I == J ? 1.0 : 0.0));
        }
    for (int I = 0; I < 3; I++)
        for (int K = 0; K < 3; K++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += st->F[k][I] * st->H[k][I][K];
            EK[I][K] = s;
        }
    const double v24[3] = {IS3, IS3, IS3};
    const double v25[3] = {IS2, -IS2, 0.0};
    const double v26[3] = {IS6, IS6, -2 * IS6};
    double e24 = (E[0][0] + E[1][1] + E[2][2]) * IS3;
    double e25 = (E[0][0] - E[1][1]) * IS2;
    double e26 = (E[0][0] + E[1][1] - 2 * E[2][2]) * IS6;
    double rho2 = e25 * e25 + e26 * e26;
    double P25 = -6 * c3 * e25 * e26 + 2 * c2 * e25 + 4 * c4 * rho2 * e25;
    double P26 = 3 * c3 * (e26 * e26 - e25 * e25) + 2 * c2 * e26 + 4 * c4 * rho2 * e26;
    double P2525 = -6 * c3 * e26 + 2 * c2 + 4 * c4 * rho2 + 8 * c4 * e25 * e25;
    double P2526 = -6 * c3 * e25 + 8 * c4 * e25 * e26;
    double P2626 = 6 * c3 * e26 + 2 * c2 + 4 * c4 * rho2 + 8 * c4 * e26 * e26;
    for (int I = 0; I < 3; I++) {
        st->S[I][I] = 2 * e0 * e24 * v24[I] + P25 * v25[I] + P26 * v26[I];
        for (int J = 0; J < 3; J++)
            st->N[I][J] = 2 * e0 * v24[I] * v24[J] + P2525 * v25[I] * v25[J]
                        + P2526 * (v25[I] * v26[J] + v26[I] * v25[J]) + P2626 * v26[I] * v26[J];
    }
    /* dPsi/dE_IJ = 2 e345 E_IJ for I<J; S_IJ carries half of it */
    st->S[0][1] = st->S[1][0] = e345 * E[0][1];
    st->S[0][2] = st->S[2][0] = e345 * E[0][2];
    st->S[1][2] = st->S[2][1] = e345 * E[1][2];
    /* gradient term: le^2 |dev EK_.K|^2 -> G = 2 le^2 (I - 11^T/3) EK */
    double le2 = le * le;
    for (int K = 0; K < 3; K++) {
        double tr = (EK[0][K] + EK[1][K] + EK[2][K]) / 3.0;
        for (int I = 0; I < 3; I++) st->G[I][K] = 2 * le2 * (EK[I][K] - tr);
    }
    st->e345x2 = 2 * e345;
    st->le2x2 = 2 * le2;
    double g2 = 0;
    for (int K = 0; K < 3; K++) {
        double g25 = (EK[0][K] - EK[1][K]) * IS2;
        double g26 = (EK[0][K] + EK[1][K] - 2 * EK[2][K]) * IS6;
        g2 += g25 * g25 + g26 * g26;
    }
    st->psi = c3 * e26 * (e26 * e26 - 3 * e25 * e25) + c2 * rho2 + c4 * rho2 * rho2 + e0 * e24 * e24
            + e345 * (E[0][1] * E[0][1] + E[0][2] * E[0][2] + E[1][2] * E[1][2]) + le2 * g2;
}

double oc_static_energy(const double *u, const double *coef)
{
    oc_state st;
    oc_static_state(u, coef, &st);
    return st.psi;
}

/* gradelast_density, phys/gradelast.c:1025-1085 (energy, 2 dPsi/dE, phases) */
void oc_static_density(double *hh, const double *u, const double *coef)
{
    oc_state st;
    oc_static_state(u, coef, &st);
    double E[3][3];
    for (int I = 0; I < 3; I++)
        for (int J = 0; J < 3; J++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += st.F[k][I] * st.F[k][J];
            E[I][J] = 0.5 * (s - (I == J ? 1.0 : 0.0));
        }
    const double e25 = (E[0][0] - E[1][1]) * IS2;
    const double e26 = (E[0][0] + E[1][1] - 2 * E[2][2]) * IS6;
    hh[0] = st.psi;
    hh[1] = 2 * st.S[0][0];
    hh[2] = 4 * st.S[0][1];
    hh[3] = 4 * st.S[0][2];
    hh[4] = 2 * st.S[1][1];
    hh[5] = 4 * st.S[1][2];
    hh[6] = 2 * st.S[2][2];
    hh[7] = (double)(e25 > 0 && e26 > -0.5 * e25);
    hh[8] = (double)(e25 < 0 && e26 > 0.5 * e25);
    hh[9] = (double)(e26 < 0.5 * e25 && e26 < -0.5 * e25);
    hh[10] = (double)(st.psi < -0.5 && e25 > 0 && e26 > 0);
    hh[11] = (double)(st.psi < -0.5 && e25 < 0 && e26 > 0);
    hh[12] = (double)(st.psi < -0.5 && e26 < 0);
}

void oc_phys_density(const oc_phys *ph, double *hh, const double *u)
{
    if (ph->torder == 0) {
        oc_static_density(hh, u, ph->coef);
        return;
    }
    /* gradelasttime_density, phys/gradelasttime.c:9183-9232 */
    double mid[30];
    for (int i = 0; i < 30; i++) mid[i] = 0.5 * (u[30 + i] + u[60 + i]);
    double kin = 0;
    for (int k = 0; k < 3; k++) {
        const double v = u[30 + 10 * k] - u[60 + 10 * k];
        kin += v * v;
    }
    hh[0] = oc_static_energy(mid, ph->coef) + ph->rho * kin / (2.0 * ph->dt * ph->dt);
    hh[1] = 1.0;
}

/* r = dPsi/du (30 jet slots, slot 10k+a) */
static void oc_static_residual_acc(const oc_state *st, double s, double *r)
{
    for (int k = 0; k < 3; k++) {
        for (int L = 0; L < 3; L++) {
            double v = 0;
            for (int J = 0; J < 3; J++) v += st->S[L][J] * st->F[k][J];
            for (int K = 0; K < 3; K++) v += st->G[L][K] * st->H[k][L][K];
            r[10 * k + 1 + L] += s * v;
        }
        for (int I = 0; I < 3; I++)
            for (int K = I; K < 3; K++) {
                double v = st->G[I][K] * st->F[k][I];
                if (I != K) v += st->G[K][I] * st->F[k][K];
                r[10 * k + SIDX[I][K]] += s * v;
            }
    }
}

void oc_static_residual(double *r, const double *u, const double *coef)
{
    oc_state st;
    oc_static_state(u, coef, &st);
    for (int i = 0; i < 30; i++) r[i] = 0.0;
    oc_static_residual_acc(&st, 1.0, r);
}

/* C = d2Psi/du2 (30x30 row-major, index (10k+a)*30 + 10l+b), scaled by s, accumulated */
static void oc_static_tangent_acc(const oc_state *st, double s, double *C)
{
    const double M[3][3] = {{2. / 3, -1. / 3, -1. / 3}, {-1. / 3, 2. / 3, -1. / 3}, {-1. / 3, -1. / 3, 2. / 3}};
    const double hs = 0.5 * st->e345x2;   /* e345 */
    const double g2 = st->le2x2;          /* 2 le^2 */
    for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++) {
            /* first-gradient x first-gradient */
            for (int L = 0; L < 3; L++)
                for (int Mi = 0; Mi < 3; Mi++) {
                    double v = st->N[L][Mi] * st->F[k][L] * st->F[l][Mi];
                    if (L == Mi) {
                        double t = 0;
                        for (int J = 0; J < 3; J++)
                            if (J != L) t += st->F[k][J] * st->F[l][J];
                        v += 0.5 * hs * t;
                    } else {
                        v += 0.5 * hs * st->F[k][Mi] * st->F[l][L];
                    }
                    double hh = 0;
                    for (int K = 0; K < 3; K++) hh += st->H[k][L][K] * st->H[l][Mi][K];
                    v += g2 * M[L][Mi] * hh;
                    if (k == l) v += st->S[L][Mi];
                    C[(10 * k + 1 + L) * 30 + 10 * l + 1 + Mi] += s * v;
                }
            /* first-gradient (k,L) x second-gradient (l,(I,K)) and its transpose */
            for (int L = 0; L < 3; L++)
                for (int I = 0; I < 3; I++)
                    for (int K = I; K < 3; K++) {
                        double v = g2 * M[L][I] * st->H[k][L][K] * st->F[l][I];
                        if (I != K) v += g2 * M[L][K] * st->H[k][L][I] * st->F[l][K];
                        if (k == l) {
                            if (L == I) v += st->G[I][K];
                            if (I != K && L == K) v += st->G[K][I];
                        }
                        C[(10 * k + 1 + L) * 30 + 10 * l + SIDX[I][K]] += s * v;
                        C[(10 * l + SIDX[I][K]) * 30 + 10 * k + 1 + L] += s * v;
                    }
            /* second x second */
            for (int I = 0; I < 3; I++)
                for (int K = I; K < 3; K++)
                    for (int J = 0; J < 3; J++)
                        for (int K2 = J; K2 < 3; K2++) {
                            int pa[2][2] = {{I, K}, {K, I}}, na_ = (I != K) ? 2 : 1;
                            int pb[2][2] = {{J, K2}, {K2, J}}, nb_ = (J != K2) ? 2 : 1;
                            double v = 0;
                            for (int x = 0; x < na_; x++)
                                for (int y = 0; y < nb_; y++)
                                    if (pa[x][1] == pb[y][1])
                                        v += M[pa[x][0]][pb[y][0]] * st->F[k][pa[x][0]] * st->F[l][pb[y][0]];
                            C[(10 * k + SIDX[I][K]) * 30 + 10 * l + SIDX[J][K2]] += s * g2 * v;
                        }
        }
}

void oc_static_tangent(double *C, const double *u, const double *coef)
{
    oc_state st;
    oc_static_state(u, coef, &st);
    for (int i = 0; i < 900; i++) C[i] = 0.0;
    oc_static_tangent_acc(&st, 1.0, C);
}

/* 4-point Gauss-Legendre on [0,1] from the closed form of the Legendre roots */
void oc_gauss4(double *x, double *w)
{
    double a = sqrt(3.0 / 7.0 - 2.0 / 7.0 * sqrt(6.0 / 5.0));
    double b = sqrt(3.0 / 7.0 + 2.0 / 7.0 * sqrt(6.0 / 5.0));
    double wa = (18.0 + sqrt(30.0)) / 36.0, wb = (18.0 - sqrt(30.0)) / 36.0;
    double t[4] = {-b, -a, a, b}, ww[4] = {wb, wa, wa, wb};
    for (int i = 0; i < 4; i++) {
        x[i] = 0.5 * (t[i] + 1.0);
        w[i] = 0.5 * ww[i];
    }
}

/* Weights of the transient rule.  With p(s) = dPsi(m0 + s d) (degree <= 7 in
 * s, Psi being a degree-8 polynomial of the jet) the reference's stress is
 * the Taylor series of int_0^1 p(s) ds truncated after the d^3 term
 * (phys/gradelasttime.c:4129-4158: coefficients 1, 1/2, 1/(3*2!), 1/(4*3!)),
 * i.e. the functional L[s^k] = 1/(k+1) for k <= 3 and 0 for 4 <= k <= 7.
 * Its tangent (the exact derivative, :4160-9181) is the functional
 * Lt[s^m] = (1/2)/(m+2) for m <= 2 and 0 for 3 <= m <= 7 applied to
 * d2Psi(m0 + s d).  Both are evaluated exactly as weighted sums over eight
 * Chebyshev nodes s_j = 2 cos(pi (j+1/2)/8); the weights solve the 8x8
 * moment systems (long double Gaussian elimination).  */
void oc_trunc_rule(double *s, double *cr, double *ct)
{
    long double V[8][8], br[8], bt[8], ns[8];
    for (int j = 0; j < 8; j++) ns[j] = 2.0L * cosl(3.14159265358979323846264338327950288L * (j + 0.5L) / 8.0L);
    for (int k = 0; k < 8; k++) {
        br[k] = (k <= 3) ? 1.0L / (k + 1) : 0.0L;
        bt[k] = (k <= 2) ? 0.5L / (k + 2) : 0.0L;
        for (int j = 0; j < 8; j++) {
            long double v = 1.0L;
            for (int t = 0; t < k; t++) v *= ns[j];
            V[k][j] = v;   /* row k: sum_j c_j s_j^k = b_k */
        }
    }
    for (int c = 0; c < 8; c++) {   /* partial pivoting */
        int piv = c;
        for (int r = c + 1; r < 8; r++)
            if (fabsl(V[r][c]) > fabsl(V[piv][c])) piv = r;
        for (int j = 0; j < 8; j++) { long double t = V[c][j]; V[c][j] = V[piv][j]; V[piv][j] = t; }
        { long double t = br[c]; br[c] = br[piv]; br[piv] = t; t = bt[c]; bt[c] = bt[piv]; bt[piv] = t; }
        for (int r = c + 1; r < 8; r++) {
            long double f = V[r][c] / V[c][c];
            for (int j = c; j < 8; j++) V[r][j] -= f * V[c][j];
            br[r] -= f * br[c];
            bt[r] -= f * bt[c];
        }
    }
    for (int c = 7; c >= 0; c--) {
        for (int j = c + 1; j < 8; j++) { br[c] -= V[c][j] * br[j]; bt[c] -= V[c][j] * bt[j]; }
        br[c] /= V[c][c];
        bt[c] /= V[c][c];
    }
    for (int j = 0; j < 8; j++) { s[j] = (double)ns[j]; cr[j] = (double)br[j]; ct[j] = (double)bt[j]; }
}

/* Transient scheme of phys/gradelasttime.c: u[90] = {u, u1, u0}.
 * value rows: rho (u + u0 - 2 u1)/dt^2 + cii (u - u0)/(2 dt)       (:4128,4138,4148)
 * other rows: L[dPsi(m0 + s d)], m0 = (u0+u1)/2, d = (u-u0)/2     (:4129-4158)
 * tangent:    value diagonal rho/dt^2 + cii/(2dt) (:8281,8591,8901),
 *             others Lt[d2Psi(m0 + s d)] (see oc_trunc_rule). */
static void oc_transient(const oc_phys *ph, const double *u, double *r, double *C)
{
    const double *u1 = u + 30, *u0 = u + 60;
    double sx[8], cr[8], ct[8];
    oc_trunc_rule(sx, cr, ct);
    if (r) for (int i = 0; i < 30; i++) r[i] = 0.0;
    if (C) for (int i = 0; i < 900; i++) C[i] = 0.0;
    for (int g = 0; g < 8; g++) {
        double us[30];
        for (int i = 0; i < 30; i++) {
            double m0 = 0.5 * (u0[i] + u1[i]), d = 0.5 * (u[i] - u0[i]);
            us[i] = m0 + sx[g] * d;
        }
        oc_state st;
        oc_static_state(us, ph->coef, &st);
        if (r) oc_static_residual_acc(&st, cr[g], r);
        if (C) oc_static_tangent_acc(&st, ct[g], C);
    }
    for (int k = 0; k < 3; k++) {
        int a = 10 * k;
        if (r) r[a] = (ph->cii * (u[a] - u0[a])) / (2. * ph->dt) + (ph->rho * (u[a] + u0[a] - 2 * u1[a])) / (ph->dt * ph->dt);
        if (C) C[a * 30 + a] = ph->cii / (2. * ph->dt) + ph->rho / (ph->dt * ph->dt);
    }
}

void oc_phys_residual(const oc_phys *ph, double *r, const double *u)
{
    if (ph->torder == 0) oc_static_residual(r, u, ph->coef);
    else oc_transient(ph, u, r, NULL);
}

void oc_phys_tangent(const oc_phys *ph, double *C, const double *u)
{
    if (ph->torder == 0) oc_static_tangent(C, u, ph->coef);
    else oc_transient(ph, u, NULL, C);
}

/* ========================================================================== */
/* IGA kernels                                                                 */
/* ========================================================================== */

void oc_matinv3(double *m)
{
    double a[9];
    memcpy(a, m, sizeof a);
    double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6])
               + a[2] * (a[3] * a[7] - a[4] * a[6]);
    double id = 1.0 / det;
    m[0] = id * (a[4] * a[8] - a[5] * a[7]);
    m[1] = id * (a[2] * a[7] - a[1] * a[8]);
    m[2] = id * (a[1] * a[5] - a[2] * a[4]);
    m[3] = id * (a[5] * a[6] - a[3] * a[8]);
    m[4] = id * (a[0] * a[8] - a[2] * a[6]);
    m[5] = id * (a[2] * a[3] - a[0] * a[5]);
    m[6] = id * (a[3] * a[7] - a[4] * a[6]);
    m[7] = id * (a[1] * a[6] - a[0] * a[7]);
    m[8] = id * (a[0] * a[4] - a[1] * a[3]);
}

/* 1-D quadratic B-spline value/1st/2nd derivative on the span [kV[2],kV[3]]
 * of the five-knot window kV, with MTOL added to every denominator. */
static void oc_bspline1d(const double *kV, double x, double *v, double *d1, double *d2)
{
    /* degree 1 on the window: only N1[1], N1[2] are non-zero (N0 = e_2) */
    double n11 = (kV[3] - x) / (kV[3] - kV[2] + MTOL);
    double n12 = (x - kV[2]) / (kV[3] - kV[2] + MTOL);
    double N1[4] = {0.0, n11, n12, 0.0};
    for (int i = 0; i < 3; i++) {
        v[i] = (x - kV[i]) / (kV[i + 2] - kV[i] + MTOL) * N1[i]
             + (kV[i + 3] - x) / (kV[i + 3] - kV[i + 1] + MTOL) * N1[i + 1];
        d1[i] = 2.0 / (kV[i + 2] - kV[i] + MTOL) * N1[i] - 2.0 / (kV[i + 3] - kV[i + 1] + MTOL) * N1[i + 1];
    }
    double dn11 = -1.0 / (kV[3] - kV[2] + MTOL), dn12 = 1.0 / (kV[3] - kV[2] + MTOL);
    double D1[4] = {0.0, dn11, dn12, 0.0};
    for (int i = 0; i < 3; i++)
        d2[i] = 2.0 / (kV[i + 2] - kV[i] + MTOL) * D1[i] - 2.0 / (kV[i + 3] - kV[i + 1] + MTOL) * D1[i + 1];
}

void oc_eval_bspline(double *N, double *const kV[3], const double *xq)
{
    double v[3][3], d1[3][3], d2[3][3];
    for (int d = 0; d < 3; d++) oc_bspline1d(kV[d], xq[d], v[d], d1[d], d2[d]);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                double *n = N + 10 * (9 * i + 3 * j + k);
                n[0] = v[0][i] * v[1][j] * v[2][k];
                n[1] = d1[0][i] * v[1][j] * v[2][k];
                n[2] = v[0][i] * d1[1][j] * v[2][k];
                n[3] = v[0][i] * v[1][j] * d1[2][k];
                n[4] = d2[0][i] * v[1][j] * v[2][k];
                n[5] = d1[0][i] * d1[1][j] * v[2][k];
                n[6] = d1[0][i] * v[1][j] * d1[2][k];
                n[7] = v[0][i] * d2[1][j] * v[2][k];
                n[8] = v[0][i] * d1[1][j] * d1[2][k];
                n[9] = v[0][i] * v[1][j] * d2[2][k];
            }
}

/* Cox-de Boor value with the end-point convention of meshcore_iga.c:348 and
 * the derivative via the divided-difference coefficients of :362-380. */
double oc_evalN(const double *kV, int nk, int k, int i, int p, double xi)
{
    if (k == 0) {
        if (i == nk - p - 2 && fabs(xi - kV[nk - 1]) < 1.e-15 && fabs(kV[nk - 1] - kV[nk - 2]) < 1.e-15) return 1.0;
        if (p == 0) return (xi >= kV[i] && kV[i + 1] > xi) ? 1.0 : 0.0;
        int lz = (kV[i + p] == kV[i]), rz = (kV[i + p + 1] == kV[i + 1]);
        if (lz && rz) return 0.0;
        if (lz) return (kV[i + p + 1] - xi) / (kV[i + p + 1] - kV[i + 1]) * oc_evalN(kV, nk, 0, i + 1, p - 1, xi);
        if (rz) return (xi - kV[i]) / (kV[i + p] - kV[i]) * oc_evalN(kV, nk, 0, i, p - 1, xi);
        return (xi - kV[i]) / (kV[i + p] - kV[i]) * oc_evalN(kV, nk, 0, i, p - 1, xi)
             + (kV[i + p + 1] - xi) / (kV[i + p + 1] - kV[i + 1]) * oc_evalN(kV, nk, 0, i + 1, p - 1, xi);
    }
    if (k < 0 || k > p) return 0.0;
    double al[4][4];
    al[0][0] = 1.0;
    for (int r = 1; r <= k; r++) {
        double den = kV[i + p - r + 1] - kV[i];
        al[r][0] = (den == 0.0) ? 0.0 : al[r - 1][0] / den;
        for (int c = 1; c < r; c++) {
            den = kV[i + p + c - r + 1] - kV[i + c];
            al[r][c] = (den == 0.0) ? 0.0 : (al[r - 1][c] - al[r - 1][c - 1]) / den;
        }
        den = kV[i + p + 1] - kV[i + r];
        al[r][r] = (den == 0.0) ? 0.0 : -al[r - 1][r - 1] / den;
    }
    double s = 0.0;
    for (int j = 0; j <= k; j++) s += al[k][j] * oc_evalN(kV, nk, 0, i + j, p - k, xi);
    double f = 1.0;
    for (int t = p; t > p - k; t--) f *= t;   /* p!/(p-k)! */
    return s * f;
}

/* ========================================================================== */
/* mesh                                                                        */
/* ========================================================================== */

void oc_mesh_igabeam(const double *L, const int *nelem, const int *periodic, int perturbed,
                     double *knot0, double *knot1, double *knot2, double *w, double *X)
{
    double *kv[3] = {knot0, knot1, knot2};
    const int p = 2;
    for (int d = 0; d < 3; d++) {
        double h = 1.0 / nelem[d];
        int n = nelem[d] + 1 + 2 * p, i = 0;
        if (!periodic[d]) {
            while (i < p + 1) kv[d][i++] = 0.;
            while (i < nelem[d] + p) { kv[d][i] = (i - p) * h; i++; }
            while (i < n) kv[d][i++] = 1.;
        } else {
            while (i < n) { kv[d][i] = (i - p) * h; i++; }
        }
    }
    int nb[3] = {nelem[0] + p, nelem[1] + p, nelem[2] + p};
    int ia = 0;
    for (int i = 0; i < nb[0]; i++)
        for (int j = 0; j < nb[1]; j++)
            for (int k = 0; k < nb[2]; k++) {
                w[ia] = 1.0;
                X[3 * ia + 0] = ((double)i - 0.5) / ((double)(nb[0] - 2)) * L[0];
                X[3 * ia + 1] = ((double)j - 0.5) / ((double)(nb[1] - 2)) * L[1];
                X[3 * ia + 2] = ((double)k - 0.5) / ((double)(nb[2] - 2)) * L[2];
                if (perturbed) {
                    double in = (double)(i > 2 && j > 2 && k > 2 && i < nb[0] - 2 && j < nb[1] - 2 && k < nb[2] - 2);
                    X[3 * ia + 0] += 0.2 * L[0] / nb[0] * in * sin(3 * i * k * j);
                    X[3 * ia + 1] += 0.2 * L[1] / nb[1] * in * sin(5 * i * k * j);
                    X[3 * ia + 2] += 0.2 * L[2] / nb[2] * in * sin(7 * i * k * j);
                }
                ia++;
            }
}

int64_t oc_ndof(const oc_mesh *m) { return 3LL * m->nb[0] * m->nb[1] * m->nb[2]; }

/* owned index of active control point a along direction d (periodic wrap) */
static inline int own1(const oc_mesh *m, int d, int a) { return m->periodic[d] ? a % m->nb[d] : a; }

static inline int64_t dof_of_active(const oc_mesh *m, int ax, int ay, int az, int i)
{
    int ox = own1(m, 0, ax), oy = own1(m, 1, ay), oz = own1(m, 2, az);
    return 3LL * (((int64_t)m->nb[1] * m->nb[2]) * ox + (int64_t)m->nb[2] * oy + oz) + i;
}

/* ---- CSR pattern: each row couples to the control points within +-2 per
 * direction (clipped, or wrapped when periodic), three dofs each, sorted. */
static int stencil1(const oc_mesh *m, int d, int a, int *out)
{
    int n = m->nb[d], cnt = 0;
    if (!m->periodic[d]) {
        for (int b = a - 2; b <= a + 2; b++)
            if (b >= 0 && b < n) out[cnt++] = b;
        return cnt;
    }
    int tmp[5];
    for (int t = 0; t < 5; t++) tmp[t] = ((a + t - 2) % n + n) % n;
    /* sort + unique */
    for (int x = 0; x < 5; x++)
        for (int y = x + 1; y < 5; y++)
            if (tmp[y] < tmp[x]) { int s = tmp[x]; tmp[x] = tmp[y]; tmp[y] = s; }
    for (int t = 0; t < 5; t++)
        if (cnt == 0 || tmp[t] != out[cnt - 1]) out[cnt++] = tmp[t];
    return cnt;
}

int64_t oc_csr_nnz(const oc_mesh *m)
{
    int64_t tot[3] = {0, 0, 0};
    int s[5];
    for (int d = 0; d < 3; d++)
        for (int a = 0; a < m->nb[d]; a++) tot[d] += stencil1(m, d, a, s);
    return 9 * tot[0] * tot[1] * tot[2];
}

/* rows [row0, row0 + nrows) of the pattern (nrows < 0: to the last row);
 * rowptr[0] = 0 and positions count from row0 */
static void row_window(const oc_mesh *m, int64_t *row0, int64_t *nrows)
{
    int64_t nd = oc_ndof(m);
    if (*row0 < 0) *row0 = 0;
    if (*row0 > nd) *row0 = nd;
    if (*nrows < 0 || *row0 + *nrows > nd) *nrows = nd - *row0;
}

int64_t oc_csr_nnz_rows(const oc_mesh *m, int64_t row0, int64_t nrows)
{
    row_window(m, &row0, &nrows);
    int64_t tot = 0, r = 0;
    int s[5];
    for (int ax = 0; ax < m->nb[0]; ax++) {
        int nx = stencil1(m, 0, ax, s);
        for (int ay = 0; ay < m->nb[1]; ay++) {
            int ny = stencil1(m, 1, ay, s);
            for (int az = 0; az < m->nb[2]; az++) {
                int nz = stencil1(m, 2, az, s);
                for (int i = 0; i < 3; i++, r++) {
                    if (r >= row0 + nrows) return tot;
                    if (r >= row0) tot += 3 * nx * ny * nz;
                }
            }
        }
    }
    return tot;
}

void oc_csr_pattern_rows(const oc_mesh *m, int64_t row0, int64_t nrows, int64_t *rowptr, int32_t *col)
{
    row_window(m, &row0, &nrows);
    int64_t r = 0, pos = 0;
    int sx[5], sy[5], sz[5];
    rowptr[0] = 0;
    for (int ax = 0; ax < m->nb[0]; ax++) {
        int nx = stencil1(m, 0, ax, sx);
        for (int ay = 0; ay < m->nb[1]; ay++) {
            int ny = stencil1(m, 1, ay, sy);
            for (int az = 0; az < m->nb[2]; az++) {
                int nz = stencil1(m, 2, az, sz);
                for (int i = 0; i < 3; i++, r++) {
                    if (r >= row0 + nrows) return;
                    if (r < row0) continue;
                    for (int x = 0; x < nx; x++)
                        for (int y = 0; y < ny; y++)
                            for (int z = 0; z < nz; z++)
                                for (int j = 0; j < 3; j++)
                                    col[pos++] = (int32_t)dof_of_active(m, sx[x], sy[y], sz[z], j);
                    rowptr[r - row0 + 1] = pos;
                }
            }
        }
    }
}

void oc_csr_pattern(const oc_mesh *m, int64_t *rowptr, int32_t *col) { oc_csr_pattern_rows(m, 0, -1, rowptr, col); }

void oc_setnz(const oc_mesh *m, int32_t *d_nz)
{
    /* setnz_iga (assmcore_iga.c:1078-1080) on one rank: with ps[] = 1 for a
     * periodic direction (the part sends to itself) the stencil is 2p+1 wide. */
    const int p = 2;
    int ia = 0;
    for (int x = 0; x < m->nb[0]; x++)
        for (int y = 0; y < m->nb[1]; y++)
            for (int z = 0; z < m->nb[2]; z++)
                for (int i = 0; i < 3; i++) {
                    int c[3] = {x, y, z}, f = 3;
                    for (int d = 0; d < 3; d++) {
                        int ps = m->periodic[d], n = m->nb[d];
                        int lo = (!ps && c[d] < p) ? c[d] : p;
                        int hi = (!ps && c[d] > n - 1 - p) ? n - 1 - c[d] : p;
                        f *= lo + 1 + hi;
                    }
                    d_nz[ia++] = f;
                }
}

/* ========================================================================== */
/* boundary conditions                                                         */
/* ========================================================================== */

/* face-local frame of boundary face f (assmcore_iga.c:276-313, 418-437):
 * in-plane directions d0,d1 and normal dn */
static void face_dirs(int f, int *d0, int *d1, int *dn)
{
    switch (f / 2) {
    case 0: *d0 = 1; *d1 = 2; *dn = 0; break;
    case 1: *d0 = 0; *d1 = 2; *dn = 1; break;
    default: *d0 = 0; *d1 = 1; *dn = 2; break;
    }
}

static inline int64_t dof_of_dirs(const oc_mesh *m, int d0, int d1, int dn, int a0, int a1, int an, int i)
{
    int a[3];
    a[d0] = a0; a[d1] = a1; a[dn] = an;
    return dof_of_active(m, a[0], a[1], a[2], i);
}

/* dirichlet_iga (assmcore_iga.c:861-934) restated; out[nb(d0)*nb(d1)] */
static void dirichlet_face(const oc_mesh *m, const double *par, int f, int order, int idof, double *out)
{
    const int p = 2;
    int d0, d1, dn;
    face_dirs(f, &d0, &d1, &dn);
    int dd[2] = {d0, d1};
    int nkn = m->nelem[dn] + 5;
    double Z = (f % 2 == 0) ? m->knot[dn][0] : m->knot[dn][nkn - 1] - 1.e-15;
    double pr[3];
    switch (f / 2) {
    case 0: pr[0] = par[3 * idof + 1]; pr[1] = par[3 * idof + 2]; pr[2] = par[3 * idof + 0]; break;
    case 1: pr[0] = par[3 * idof + 0]; pr[1] = par[3 * idof + 2]; pr[2] = par[3 * idof + 1]; break;
    default: pr[0] = par[3 * idof + 0]; pr[1] = par[3 * idof + 1]; pr[2] = par[3 * idof + 2]; break;
    }
    double *d1d[2];
    for (int b = 0; b < 2; b++) {
        int d = dd[b], nb = m->nb[d], nk = m->nelem[d] + 5;
        const double *kv = m->knot[d];
        d1d[b] = (double *)malloc(sizeof(double) * nb);
        double W = kv[p];
        double vec[2] = {pr[b] * W, pr[b]};
        double mat[4];
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) mat[2 * i + j] = oc_evalN(kv, nk, i, j, p, W);
        /* matinv_dot_vec (mathutil.c:6-16, 2x2 inverse of :22-28) */
        double id = 1. / (-(mat[1] * mat[2]) + mat[0] * mat[3]);
        double inv[4] = {id * mat[3], id * (-mat[1]), id * (-mat[2]), id * mat[0]};
        d1d[b][0] = inv[0] * vec[0] + inv[1] * vec[1];
        d1d[b][1] = inv[2] * vec[0] + inv[3] * vec[1];
        for (int i = p; i < nb; i++) d1d[b][i] = d1d[b][i - 1] + (kv[i + p] - kv[i]) / ((double)p) * pr[b];
    }
    int n0 = m->nb[d0], n1 = m->nb[d1];
    for (int i = 0; i < n0; i++)
        for (int j = 0; j < n1; j++) out[n1 * i + j] = d1d[0][i] + d1d[1][j] + pr[2] * Z;
    if (order == 1) {
        int ibz1 = (f % 2 == 0) ? 1 : m->nb[dn] - 2;
        double mm = 0.0 / oc_evalN(m->knot[dn], nkn, 1, ibz1, p, Z);
        for (int i = 0; i < n0 * n1; i++) out[i] += mm;
    }
    free(d1d[0]);
    free(d1d[1]);
}

static int face_present(const oc_mesh *m, int f) { return !m->periodic[f / 2]; }

void oc_dirichlet(const oc_mesh *m, const oc_bc *bc, int32_t *mask, double *g)
{
    int64_t nd = oc_ndof(m);
    for (int64_t i = 0; i < nd; i++) { mask[i] = 0; g[i] = 0.0; }
    for (int f = 0; f < 6; f++) {
        if (!face_present(m, f)) continue;
        int d0, d1, dn;
        face_dirs(f, &d0, &d1, &dn);
        int n0 = m->nb[d0], n1 = m->nb[d1];
        double *vals = (double *)malloc(sizeof(double) * n0 * n1);
        for (int order = 0; order < 2; order++) {
            int an = (f % 2 == 0) ? order : m->nb[dn] - 1 - order;
            for (int idof = 0; idof < 3; idof++) {
                if (bc->bc_type[6 * f + 3 * order + idof] != 0) continue;
                dirichlet_face(m, bc->par_dirichlet, f, order, idof, vals);
                for (int i = 0; i < n0; i++)
                    for (int j = 0; j < n1; j++) {
                        int64_t r = dof_of_dirs(m, d0, d1, dn, i, j, an, idof);
                        mask[r] = 1;
                        g[r] = vals[n1 * i + j];
                    }
            }
        }
        free(vals);
    }
}

void oc_neumann(const oc_mesh *m, const oc_bc *bc, double *F)
{
    const int p = 2;
    double cq[4], wq[4];
    oc_gauss4(cq, wq);
    int64_t nd = oc_ndof(m);
    for (int64_t i = 0; i < nd; i++) F[i] = 0.0;
    for (int f = 0; f < 6; f++) {
        if (!face_present(m, f)) continue;
        int d0, d1, dn;
        face_dirs(f, &d0, &d1, &dn);
        const double *k0 = m->knot[d0], *k1 = m->knot[d1], *kn = m->knot[dn];
        int nk0 = m->nelem[d0] + 5, nk1 = m->nelem[d1] + 5, nkn = m->nelem[dn] + 5;
        double zq = (f % 2 == 0) ? kn[0] : kn[nkn - 1];
        int z0 = (f % 2 == 0) ? 0 : m->nelem[dn] + p - 1;
        int z1 = (f % 2 == 0) ? 1 : m->nelem[dn] + p - 2;
        for (int order = 0; order < 2; order++)
            for (int idof = 0; idof < 3; idof++) {
                if (bc->bc_type[6 * f + 3 * order + idof] != 1) continue;
                double t = bc->neumann[6 * f + 3 * order + idof];
                /* a zero traction adds only -0.0 terms, which leave every sum
                 * as it is (the reference still integrates them) */
                if (t == 0.0) continue;
                for (int ex = 0; ex < m->nelem[d0]; ex++)
                    for (int ey = 0; ey < m->nelem[d1]; ey++) {
                        double X0 = k0[ex + p], X1 = k0[ex + p + 1], Y0 = k1[ey + p], Y1 = k1[ey + p + 1];
                        double val[18];
                        for (int i = 0; i < 18; i++) val[i] = 0.0;
                        for (int qx = 0; qx < NQ; qx++)
                            for (int qy = 0; qy < NQ; qy++) {
                                double x = (-cq[qx] + 1.0) * X0 + cq[qx] * X1;
                                double y = (-cq[qy] + 1.0) * Y0 + cq[qy] * Y1;
                                double wgt = wq[qx] * wq[qy] * (X1 - X0) * (Y1 - Y0);
                                int ia = 0;
                                for (int bx = 0; bx < 3; bx++)
                                    for (int by = 0; by < 3; by++) {
                                        double n2 = oc_evalN(k0, nk0, 0, ex + bx, p, x) * oc_evalN(k1, nk1, 0, ey + by, p, y);
                                        if (order == 0) {
                                            val[ia++] -= wgt * n2 * t;
                                        } else {
                                            val[ia++] -= wgt * n2 * oc_evalN(kn, nkn, 1, z0, p, zq) * t;
                                            val[ia++] -= wgt * n2 * oc_evalN(kn, nkn, 1, z1, p, zq) * t;
                                        }
                                    }
                            }
                        int ia = 0;
                        for (int bx = 0; bx < 3; bx++)
                            for (int by = 0; by < 3; by++) {
                                F[dof_of_dirs(m, d0, d1, dn, ex + bx, ey + by, z0, idof)] += val[ia++];
                                if (order == 1) F[dof_of_dirs(m, d0, d1, dn, ex + bx, ey + by, z1, idof)] += val[ia++];
                            }
                    }
            }
    }
}

/* ========================================================================== */
/* element kernels                                                             */
/* ========================================================================== */

typedef struct {
    double R[27 * 10];   /* physical basis jet */
    double weight;
    double xq[3];
    double detG;
} oc_gp;

/* NURBS/geometry block of assmcore_iga.c:147-208 for one Gauss point */
static void oc_gausspoint(const oc_mesh *m, int ex, int ey, int ez, const double *cq, const double *wq,
                          int qx, int qy, int qz, oc_gp *gp)
{
    const int p = 2;
    int e[3] = {ex, ey, ez}, q[3] = {qx, qy, qz};
    double X0[3], X1[3];
    double *kV[3];
    for (int d = 0; d < 3; d++) {
        X0[d] = m->knot[d][e[d] + p];
        X1[d] = m->knot[d][e[d] + p + 1];
        kV[d] = m->knot[d] + e[d];
        gp->xq[d] = (-cq[q[d]] + 1.0) * X0[d] + cq[q[d]] * X1[d];
    }
    double wV[27], XV[81];
    int ia = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++, ia++) {
                int64_t id = ((int64_t)m->na[1] * m->na[2]) * (ex + i) + (int64_t)m->na[2] * (ey + j) + (ez + k);
                wV[ia] = m->w[id];
                for (int d = 0; d < 3; d++) XV[3 * ia + d] = m->X[3 * id + d];
            }
    double N[270], *R = gp->R, W[10];
    oc_eval_bspline(N, kV, gp->xq);
    for (int a = 0; a < 10; a++) W[a] = 0.0;
    for (int b = 0; b < 27; b++)
        for (int a = 0; a < 10; a++) W[a] += wV[b] * N[10 * b + a];
    for (int b = 0; b < 27; b++)
        for (int a = 0; a < 10; a++) R[10 * b + a] = N[10 * b + a] * wV[b] / W[0];
    for (int b = 0; b < 27; b++)
        for (int a = 1; a < 10; a++) R[10 * b + a] -= R[10 * b] * W[a] / W[0];
    for (int b = 0; b < 27; b++) {
        int s = 10 * b + 4;
        for (int i = 1; i < 4; i++)
            for (int j = i; j < 4; j++) R[s++] -= (R[10 * b + j] * W[i] + R[10 * b + i] * W[j]) / W[0];
    }
    double X[30];
    for (int i = 0; i < 30; i++) X[i] = 0.0;
    for (int b = 0; b < 27; b++)
        for (int d = 0; d < 3; d++)
            for (int a = 0; a < 10; a++) X[10 * d + a] += XV[3 * b + d] * R[10 * b + a];
    double G[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) G[3 * i + j] = X[10 * i + 1 + j];
    gp->detG = G[0] * (G[4] * G[8] - G[5] * G[7]) - G[1] * (G[3] * G[8] - G[5] * G[6]) + G[2] * (G[3] * G[7] - G[4] * G[6]);
    oc_matinv3(G);
    memcpy(N, R, sizeof N);
    for (int b = 0; b < 27; b++)
        for (int i = 0; i < 3; i++) {
            double s = 0.0;
            for (int j = 0; j < 3; j++) s += N[10 * b + 1 + j] * G[3 * j + i];
            R[10 * b + 1 + i] = s;
        }
    for (int b = 0; b < 27; b++)
        for (int a = 4; a < 10; a++)
            for (int d = 0; d < 3; d++) N[10 * b + a] -= R[10 * b + 1 + d] * X[10 * d + a];
    for (int b = 0; b < 27; b++) {
        const double *n = N + 10 * b + 4;
        double Hm[3][3] = {{n[0], n[1], n[2]}, {n[1], n[3], n[4]}, {n[2], n[4], n[5]}};
        double *r = R + 10 * b + 4;
        int s = 0;
        for (int i = 0; i < 3; i++)
            for (int j = i; j < 3; j++) {
                double acc = 0.0;
                for (int pp = 0; pp < 3; pp++)
                    for (int qq = 0; qq < 3; qq++) acc += G[3 * pp + i] * Hm[pp][qq] * G[3 * qq + j];
                r[s++] = acc;
            }
    }
    gp->weight = wq[qx] * wq[qy] * wq[qz] * (X1[0] - X0[0]) * (X1[1] - X0[1]) * (X1[2] - X0[2]) * gp->detG;
}

/* u-jet at a Gauss point: u[(torder+1)*30] (assmcore_iga.c:209-229) */
static void oc_ujet(const oc_mesh *m, const oc_bc *bc, int nh, const double *U_hist,
                    int ex, int ey, int ez, const oc_gp *gp, double *u)
{
    int64_t nd = oc_ndof(m);
    for (int i = 0; i < 30 * nh; i++) u[i] = 0.0;
    for (int bx = 0; bx < 3; bx++)
        for (int by = 0; by < 3; by++)
            for (int bz = 0; bz < 3; bz++) {
                int b = 9 * bx + 3 * by + bz;
                for (int h = 0; h < nh; h++)
                    for (int k = 0; k < 3; k++) {
                        double Uv = U_hist[h * nd + dof_of_active(m, ex + bx, ey + by, ez + bz, k)];
                        for (int a = 0; a < 10; a++) u[30 * h + 10 * k + a] += gp->R[10 * b + a] * Uv;
                    }
            }
    for (int h = 0; h < nh; h++) {
        for (int k = 0; k < 3; k++)
            u[30 * h + 10 * k] += bc->par_periodic[3 * k + 0] * gp->xq[0] + bc->par_periodic[3 * k + 1] * gp->xq[1]
                                + bc->par_periodic[3 * k + 2] * gp->xq[2];
        for (int k = 0; k < 3; k++)
            for (int d = 0; d < 3; d++) u[30 * h + 10 * k + 1 + d] += bc->par_periodic[3 * k + d];
    }
}

void oc_element_residual(const oc_mesh *m, const oc_phys *ph, const oc_bc *bc,
                         const double *U_hist, int ex, int ey, int ez, double *val)
{
    double cq[4], wq[4];
    oc_gauss4(cq, wq);
    int nh = ph->torder + 1;
    for (int i = 0; i < 81; i++) val[i] = 0.0;
    for (int qx = 0; qx < NQ; qx++)
        for (int qy = 0; qy < NQ; qy++)
            for (int qz = 0; qz < NQ; qz++) {
                oc_gp gp;
                oc_gausspoint(m, ex, ey, ez, cq, wq, qx, qy, qz, &gp);
                double u[90], rr[30];
                oc_ujet(m, bc, nh, U_hist, ex, ey, ez, &gp, u);
                oc_phys_residual(ph, rr, u);
                for (int b = 0; b < 27; b++)
                    for (int k = 0; k < 3; k++) {
                        double t = 0.0;
                        for (int a = 0; a < 10; a++) t += gp.R[10 * b + a] * rr[10 * k + a];
                        val[3 * b + k] += gp.weight * t;
                    }
            }
}

void oc_element_tangent(const oc_mesh *m, const oc_phys *ph, const oc_bc *bc,
                        const double *U_hist, int ex, int ey, int ez, double *val)
{
    double cq[4], wq[4];
    oc_gauss4(cq, wq);
    int nh = ph->torder + 1;
    double *tt = (double *)malloc(900 * sizeof(double));
    for (int i = 0; i < 6561; i++) val[i] = 0.0;
    for (int qx = 0; qx < NQ; qx++)
        for (int qy = 0; qy < NQ; qy++)
            for (int qz = 0; qz < NQ; qz++) {
                oc_gp gp;
                oc_gausspoint(m, ex, ey, ez, cq, wq, qx, qy, qz, &gp);
                double u[90];
                oc_ujet(m, bc, nh, U_hist, ex, ey, ez, &gp, u);
                oc_phys_tangent(ph, tt, u);
                /* the naive contraction of assmcore_iga.c:723-749 */
                for (int b = 0; b < 27; b++)
                    for (int k = 0; k < 3; k++) {
                        int row = 3 * b + k;
                        for (int c = 0; c < 27; c++)
                            for (int l = 0; l < 3; l++) {
                                int cl = 3 * c + l;
                                double t = 0.0;
                                for (int a = 0; a < 10; a++)
                                    for (int bb = 0; bb < 10; bb++)
                                        t += gp.R[10 * b + a] * tt[30 * (10 * k + a) + (10 * l + bb)] * gp.R[10 * c + bb];
                                val[81 * row + cl] += gp.weight * t;
                            }
                    }
            }
    free(tt);
}

/* ========================================================================== */
/* assembly                                                                    */
/* ========================================================================== */

/* Element colouring for the threaded loops: elements of one colour share no
 * control point.  Along a direction the colour is e mod 3; in a periodic
 * direction whose element count is not a multiple of 3 the trailing
 * elements (which wrap onto the first control points) get colours 3 and 4. */
static int ncolour1(const oc_mesh *m, int d) { return (m->periodic[d] && m->nelem[d] % 3) ? 3 + m->nelem[d] % 3 : 3; }
static int colour1(const oc_mesh *m, int d, int e)
{
    int reg = m->periodic[d] ? 3 * (m->nelem[d] / 3) : m->nelem[d];
    return e < reg ? e % 3 : 3 + (e - reg);
}

/* the element indices of each direction that lie in the box and have the
 * given colour: ls[d][0..nl[d]) (storage from el, nelem[0]+nelem[1]+nelem[2]) */
static void box_colour_lists(const oc_mesh *m, const int *box, int cx, int cy, int cz, int *el, int *ls[3], int nl[3])
{
    const int cc[3] = {cx, cy, cz};
    int off = 0;
    for (int d = 0; d < 3; d++) {
        int lo = box ? box[2 * d] : 0, hi = box ? box[2 * d + 1] : m->nelem[d];
        if (lo < 0) lo = 0;
        if (hi > m->nelem[d]) hi = m->nelem[d];
        ls[d] = el + off;
        nl[d] = 0;
        for (int e = lo; e < hi; e++)
            if (colour1(m, d, e) == cc[d]) ls[d][nl[d]++] = e;
        off += m->nelem[d];
    }
}

static void set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

void oc_residual(const oc_mesh *m, const oc_phys *ph, const oc_bc *bc, const double *U_hist,
                 double *R, const int *box, int apply_bc, int nthreads)
{
    int64_t nd = oc_ndof(m);
    set_threads(nthreads);
    for (int64_t i = 0; i < nd; i++) R[i] = 0.0;   /* VecSet(RR,0) :105 */
    int ncx = ncolour1(m, 0), ncy = ncolour1(m, 1), ncz = ncolour1(m, 2);
    int *el = (int *)malloc(sizeof(int) * (m->nelem[0] + m->nelem[1] + m->nelem[2]));
    for (int col = 0; col < ncx * ncy * ncz; col++) {
        int cx = col / (ncy * ncz), cy = (col / ncz) % ncy, cz = col % ncz;
        int *ls[3], nl[3];
        box_colour_lists(m, box, cx, cy, cz, el, ls, nl);
        int64_t nt = (int64_t)nl[0] * nl[1] * nl[2];
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t t = 0; t < nt; t++) {
            int ex = ls[0][t / ((int64_t)nl[1] * nl[2])], ey = ls[1][(t / nl[2]) % nl[1]], ez = ls[2][t % nl[2]];
            double val[81];
            oc_element_residual(m, ph, bc, U_hist, ex, ey, ez, val);
            for (int bx = 0; bx < 3; bx++)
                for (int by = 0; by < 3; by++)
                    for (int bz = 0; bz < 3; bz++)
                        for (int k = 0; k < 3; k++)
                            R[dof_of_active(m, ex + bx, ey + by, ez + bz, k)] += val[3 * (9 * bx + 3 * by + bz) + k];
        }
    }
    free(el);
    if (!apply_bc) return;
    double *F = (double *)malloc(sizeof(double) * nd);
    oc_neumann(m, bc, F);
    for (int64_t i = 0; i < nd; i++) R[i] += F[i];
    free(F);
    int32_t *mask = (int32_t *)malloc(sizeof(int32_t) * nd);
    double *g = (double *)malloc(sizeof(double) * nd);
    oc_dirichlet(m, bc, mask, g);
    for (int64_t i = 0; i < nd; i++)
        if (mask[i]) R[i] = U_hist[i] - g[i];
    free(mask);
    free(g);
}

static int64_t find_col(const int64_t *rowptr, const int32_t *col, int64_t row, int32_t c)
{
    int64_t lo = rowptr[row], hi = rowptr[row + 1] - 1;
    while (lo <= hi) {
        int64_t mid = (lo + hi) / 2;
        if (col[mid] == c) return mid;
        if (col[mid] < c) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

void oc_tangent(const oc_mesh *m, const oc_phys *ph, const oc_bc *bc, const double *U_hist,
                const int64_t *rowptr, const int32_t *col, double *val,
                const int *box, int apply_bc, int nthreads, int64_t row0, int64_t nrows)
{
    int64_t nd = oc_ndof(m);
    row_window(m, &row0, &nrows);
    set_threads(nthreads);
    for (int64_t i = 0; i < rowptr[nrows]; i++) val[i] = 0.0;   /* MatZeroEntries :594 */
    int ncx = ncolour1(m, 0), ncy = ncolour1(m, 1), ncz = ncolour1(m, 2);
    int *el = (int *)malloc(sizeof(int) * (m->nelem[0] + m->nelem[1] + m->nelem[2]));
    for (int cc = 0; cc < ncx * ncy * ncz; cc++) {
        int cx = cc / (ncy * ncz), cy = (cc / ncz) % ncy, cz = cc % ncz;
        int *ls[3], nl[3];
        box_colour_lists(m, box, cx, cy, cz, el, ls, nl);
        int64_t nt = (int64_t)nl[0] * nl[1] * nl[2];
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t t = 0; t < nt; t++) {
            int ex = ls[0][t / ((int64_t)nl[1] * nl[2])], ey = ls[1][(t / nl[2]) % nl[1]], ez = ls[2][t % nl[2]];
            double *ke = (double *)malloc(6561 * sizeof(double));
            oc_element_tangent(m, ph, bc, U_hist, ex, ey, ez, ke);
            int64_t idx[81];
            for (int bx = 0; bx < 3; bx++)
                for (int by = 0; by < 3; by++)
                    for (int bz = 0; bz < 3; bz++)
                        for (int k = 0; k < 3; k++)
                            idx[3 * (9 * bx + 3 * by + bz) + k] = dof_of_active(m, ex + bx, ey + by, ez + bz, k);
            /* MatSetValues(...,ADD_VALUES): per row, search each column */
            for (int r = 0; r < 81; r++) {
                if (idx[r] < row0 || idx[r] >= row0 + nrows) continue;   /* rows outside the stored window */
                for (int c = 0; c < 81; c++) {
                    int64_t pos = find_col(rowptr, col, idx[r] - row0, (int32_t)idx[c]);
                    if (pos >= 0) val[pos] += ke[81 * r + c];
                }
            }
            free(ke);
        }
    }
    free(el);
    if (!apply_bc) return;
    /* MatZeroRowsColumns(P, rows, diag=1) (assmcore_iga.c:844) */
    int32_t *mask = (int32_t *)malloc(sizeof(int32_t) * nd);
    double *g = (double *)malloc(sizeof(double) * nd);
    oc_dirichlet(m, bc, mask, g);
    for (int64_t r = 0; r < nrows; r++)
        for (int64_t p = rowptr[r]; p < rowptr[r + 1]; p++)
            if (mask[row0 + r] || mask[col[p]]) val[p] = (col[p] == row0 + r) ? 1.0 : 0.0;
    free(mask);
    free(g);
}

double oc_energy(const oc_mesh *m, const oc_phys *ph, const oc_bc *bc, const double *U_hist)
{
    double cq[4], wq[4];
    oc_gauss4(cq, wq);
    double tot = 0.0;
    for (int ex = 0; ex < m->nelem[0]; ex++)
        for (int ey = 0; ey < m->nelem[1]; ey++)
            for (int ez = 0; ez < m->nelem[2]; ez++)
                for (int qx = 0; qx < NQ; qx++)
                    for (int qy = 0; qy < NQ; qy++)
                        for (int qz = 0; qz < NQ; qz++) {
                            oc_gp gp;
                            oc_gausspoint(m, ex, ey, ez, cq, wq, qx, qy, qz, &gp);
                            double u[90];
                            oc_ujet(m, bc, 1, U_hist, ex, ey, ez, &gp, u);
                            tot += gp.weight * oc_static_energy(u, ph->coef);
                        }
    return tot;
}

int oc_intdensity(const oc_mesh *m, const oc_phys *ph, const oc_bc *bc, const double *U_hist, double *HH)
{
    const int nd = ph->torder == 0 ? OC_NDENSITY_STATIC : OC_NDENSITY_TRANSIENT;
    const int nh = ph->torder + 1;
    double cq[4], wq[4];
    oc_gauss4(cq, wq);
    for (int i = 0; i < nd; i++) HH[i] = 0.0;
    for (int ex = 0; ex < m->nelem[0]; ex++)
        for (int ey = 0; ey < m->nelem[1]; ey++)
            for (int ez = 0; ez < m->nelem[2]; ez++)
                for (int qx = 0; qx < NQ; qx++)
                    for (int qy = 0; qy < NQ; qy++)
                        for (int qz = 0; qz < NQ; qz++) {
                            oc_gp gp;
                            oc_gausspoint(m, ex, ey, ez, cq, wq, qx, qy, qz, &gp);
                            double u[90], hh[OC_NDENSITY_STATIC];
                            oc_ujet(m, bc, nh, U_hist, ex, ey, ez, &gp, u);
                            oc_phys_density(ph, hh, u);
                            for (int i = 0; i < nd; i++) HH[i] += gp.weight * hh[i];
                        }
    /* Neumann faces: F = -int N t (the residual's load), so -int u t = U . F
     * with U the current level (static) or the mean of levels 1 and 2 */
    const int64_t ndof = oc_ndof(m);
    double *F = (double *)malloc(sizeof(double) * ndof);
    oc_neumann(m, bc, F);
    double s = 0;
    for (int64_t i = 0; i < ndof; i++) {
        const double ub = ph->torder == 0 ? U_hist[i] : 0.5 * (U_hist[ndof + i] + U_hist[2 * ndof + i]);
        s += ub * F[i];
    }
    free(F);
    HH[0] += s;
    return nd;
}

double oc_volume(const oc_mesh *m)
{
    double cq[4], wq[4];
    oc_gauss4(cq, wq);
    double tot = 0.0;
    for (int ex = 0; ex < m->nelem[0]; ex++)
        for (int ey = 0; ey < m->nelem[1]; ey++)
            for (int ez = 0; ez < m->nelem[2]; ez++)
                for (int qx = 0; qx < NQ; qx++)
                    for (int qy = 0; qy < NQ; qy++)
                        for (int qz = 0; qz < NQ; qz++) {
                            oc_gp gp;
                            oc_gausspoint(m, ex, ey, ez, cq, wq, qx, qy, qz, &gp);
                            tot += gp.weight;
                        }
    return tot;
}

void oc_element_basis(const oc_mesh *m, int ex, int ey, int ez, int qx, int qy, int qz, double *out)
{
    double cq[4], wq[4];
    oc_gauss4(cq, wq);
    oc_gp gp;
    oc_gausspoint(m, ex, ey, ez, cq, wq, qx, qy, qz, &gp);
    memcpy(out, gp.R, 270 * sizeof(double));
    out[270] = gp.detG;
    out[271] = gp.weight;
}
