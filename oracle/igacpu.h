/* igacpu.h -- CPU restatement of IGAP4's residual/tangent assembly.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle that the HIP path is
 * checked against, and the CPU baseline timed beside it.  It is never linked
 * into, loaded by, or called from the product library.
 *
 * Every function restates a reference function, cited file:line against the
 * reference tree.  The pointwise physics is this project's own derivation of
 * the energy in phys/gradelast.c (not a copy of the generated code); it is
 * checked against the compiled reference callbacks in tests/.
 */
#ifndef IGACPU_H
#define IGACPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- physics -------------------------------------------------------------
 * coef = {c2, c3, c4, e0, e345, le} of
 *   Psi = c3 e26(e26^2-3e25^2) + c2(e25^2+e26^2) + c4(e25^2+e26^2)^2
 *       + e0 e24^2 + e345(E12^2+E13^2+E23^2) + le^2 sum_K(e25_K^2+e26_K^2)
 * (phys/gradelast.c:10-48,1068).  tpar = {rho, cii, dt} for the transient
 * scheme (phys/gradelasttime.c:4128-4158). */
typedef struct {
    int    torder;    /* 0: gradelast, 2: gradelasttime */
    double coef[6];
    double rho, cii, dt;
} oc_phys;

/* (c2,c3,c4,e0,e345,le) from par_mat = {r, curv, e0, e345, le, ...}
 * (phys/gradelast.c:46-48). */
void oc_coef_from_par(const double *par, double *coef);
void oc_phys_init(oc_phys *ph, int torder, const double *par9, const double *coef_or_null);

/* pointwise callbacks with the Phys ABI shapes (include/physimpl.h:6-7) */
void oc_static_residual(double *r, const double *u, const double *coef);
void oc_static_tangent(double *C, const double *u, const double *coef);
void oc_phys_residual(const oc_phys *ph, double *r, const double *u);
void oc_phys_tangent(const oc_phys *ph, double *C, const double *u);
double oc_static_energy(const double *u, const double *coef);
/* Density callbacks of iga_intdensity (include/physimpl.h:9):
 * gradelast_density (phys/gradelast.c:1025-1085): 13 values, hh[0] the stored
 * energy, hh[1..6] = 2 dPsi/dE_IJ for IJ = 11,12,13,22,23,33 (each off-
 * diagonal E_IJ taken once), hh[7..12] the phase indicators;
 * gradelasttime_density (phys/gradelasttime.c:9183-9232): 2 values, the
 * stored energy at the midpoint (u1+u0)/2 plus the kinetic energy
 * rho |u1-u0|^2 / (2 dt^2), and 1 (the volume).  u = [u, u1, u0] jets. */
#define OC_NDENSITY_STATIC 13
#define OC_NDENSITY_TRANSIENT 2
void oc_static_density(double *hh, const double *u, const double *coef);
void oc_phys_density(const oc_phys *ph, double *hh, const double *u);

/* ---- IGA kernels ---------------------------------------------------------- */
void oc_gauss4(double *x, double *w);                     /* quadrature.c:381-429 on [0,1] */
void oc_trunc_rule(double *s, double *cr, double *ct);    /* transient rule (igacpu.c) */
void oc_eval_bspline(double *N, double *const kV[3], const double *xq); /* meshcore_iga.c:243-327 */
double oc_evalN(const double *kV, int nk, int k, int i, int p, double xi); /* meshcore_iga.c:329-384 */
void oc_matinv3(double *m);                               /* mathutil.c:30-41 */

/* ---- mesh ----------------------------------------------------------------- */
typedef struct {
    int     nelem[3];
    int     periodic[3];
    int     na[3];        /* active control points per direction (nelem+2)    */
    int     nb[3];        /* owned control points (periodic: nelem)           */
    double *knot[3];      /* nelem+5 knots each                               */
    double *w;            /* na0*na1*na2                                      */
    double *X;            /* 3*na0*na1*na2                                    */
} oc_mesh;

/* igabeam / igabeamtest with per-direction element counts (mesh/igabeam.c:34-72,
 * mesh/igabeamtest.c:68-70).  Arrays are caller-owned: knots of length
 * nelem[d]+5 each, w of na-product, X of 3*na-product. */
void oc_mesh_igabeam(const double *L, const int *nelem, const int *periodic, int perturbed,
                     double *knot0, double *knot1, double *knot2, double *w, double *X);

/* ---- boundary conditions (include/bcimpl.h:8-15) --------------------------- */
typedef struct {
    int    bc_type[36];       /* [(2*3)*face + 3*order + dof]; 0 Dirichlet, 1 Neumann */
    double par_periodic[9];
    double par_dirichlet[9];
    double neumann[36];       /* constant traction per face/order/dof */
} oc_bc;

int64_t oc_ndof(const oc_mesh *m);

/* CSR pattern of the AIJ matrix: rows are owned dofs, columns sorted */
int64_t oc_csr_nnz(const oc_mesh *m);
void oc_csr_pattern(const oc_mesh *m, int64_t *rowptr, int32_t *col);
/* rows [row0, row0 + nrows) of that pattern (nrows < 0: to the last row),
 * rowptr[0] = 0 */
int64_t oc_csr_nnz_rows(const oc_mesh *m, int64_t row0, int64_t nrows);
void oc_csr_pattern_rows(const oc_mesh *m, int64_t row0, int64_t nrows, int64_t *rowptr, int32_t *col);
void oc_setnz(const oc_mesh *m, int32_t *d_nz);   /* assmcore_iga.c:1054-1087, one rank */

/* Dirichlet rows: mask[ndof] (1 if constrained) and g[ndof] (value inserted,
 * last face wins), restating the INSERT sequence of assmcore_iga.c:416-494. */
void oc_dirichlet(const oc_mesh *m, const oc_bc *bc, int32_t *mask, double *g);
/* Neumann load vector (U independent part of assmcore_iga.c:273-405) */
void oc_neumann(const oc_mesh *m, const oc_bc *bc, double *F);

/* Residual_iga (assmcore_iga.c:16-502) on U_hist[(torder+1)*ndof].
 * box = {x0,x1,y0,y1,z0,z1} restricts the volume loop to elements in the box
 * (NULL: all).  apply_bc = 0 skips Neumann/Dirichlet. */
void oc_residual(const oc_mesh *m, const oc_phys *ph, const oc_bc *bc, const double *U_hist,
                 double *R, const int *box, int apply_bc, int nthreads);
/* Tangent_iga (assmcore_iga.c:507-858) into the CSR pattern of oc_csr_pattern,
 * or of its rows [row0, row0 + nrows) (oc_csr_pattern_rows; nrows < 0: to
 * the last row): contributions to other rows are dropped. */
void oc_tangent(const oc_mesh *m, const oc_phys *ph, const oc_bc *bc, const double *U_hist,
                const int64_t *rowptr, const int32_t *col, double *val,
                const int *box, int apply_bc, int nthreads, int64_t row0, int64_t nrows);

/* per-element kernels (one element, reference local ordering 3*ibpe+idof) */
void oc_element_residual(const oc_mesh *m, const oc_phys *ph, const oc_bc *bc,
                         const double *U_hist, int ex, int ey, int ez, double *val81);
void oc_element_tangent(const oc_mesh *m, const oc_phys *ph, const oc_bc *bc,
                        const double *U_hist, int ex, int ey, int ez, double *val6561);

/* total stored energy  sum_e sum_q w_q Psi(u_q) (static physics; used for
 * R = dPi/dU checks) */
double oc_energy(const oc_mesh *m, const oc_phys *ph, const oc_bc *bc, const double *U_hist);
/* iga_intdensity (src/postcore_iga.c:13-379): HH[i] = sum_e sum_q w_q hh_i(u_q)
 * minus the Neumann face terms (hh[0] = u t, transient (u1+u0)/2 t, the
 * other densities 0 there).  Returns the number of densities. */
int oc_intdensity(const oc_mesh *m, const oc_phys *ph, const oc_bc *bc, const double *U_hist, double *HH);
/* sum_e sum_q w_q detG  (volume, postcore_iga.c with density 1) */
double oc_volume(const oc_mesh *m);

/* pointwise jet at one Gauss point of one element (for basis/geometry tests):
 * out[0..269] = physical basis jet R[27][10], out[270] = detG, out[271] = weight */
void oc_element_basis(const oc_mesh *m, int ex, int ey, int ez, int qx, int qy, int qz, double *out);

#ifdef __cplusplus
}
#endif
#endif
