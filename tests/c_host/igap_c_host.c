/* igap_c_host.c -- a plain C host of the assembly library, as an IGAP4 driver
 * would be: it builds the descriptors of the igabeam patch (mesh/igabeam.c),
 * the St Venant-Kirchhoff + gradient physics and the clamped end + tip
 * displacement, then, through include/igap_gpu.h only,
 *   1. assembles the residual and the FP64 tangent at a smooth state
 *      (the Residual_iga / Tangent_iga callbacks, src/assmcore_iga.c), and
 *   2. solves the load step from zero (IGAPSolvSolve, src/interface.c:302-328),
 * and writes everything to a binary file for the test that runs it:
 *   int64 ndof, int64 nnz, double R[ndof], int64 rowptr[ndof+1],
 *   int32 col[nnz], double vals[nnz], int32 newton_its, int32 converged,
 *   double fnorm, double U[ndof] (the solution).
 *
 *   igap_c_host NX NY NZ OUTFILE
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "igap_gpu.h"

#define CHECK(call)                                                                          \
    do {                                                                                     \
        int rc_ = (call);                                                                    \
        if (rc_ != IGAP_OK) {                                                                \
            fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, igap_gpu_last_error());      \
            exit(1);                                                                         \
        }                                                                                    \
    } while (0)

static void *xmalloc(size_t n)
{
    void *p = malloc(n ? n : 1);
    if (!p) {
        fprintf(stderr, "out of host memory (%zu bytes)\n", n);
        exit(1);
    }
    return p;
}

static void put(FILE *f, const void *p, size_t sz, size_t n)
{
    if (fwrite(p, sz, n, f) != n) {
        fprintf(stderr, "short write\n");
        exit(1);
    }
}

int main(int argc, char **argv)
{
    if (argc != 5) {
        fprintf(stderr, "usage: %s NX NY NZ OUTFILE\n", argv[0]);
        return 2;
    }
    const int nelem[3] = {atoi(argv[1]), atoi(argv[2]), atoi(argv[3])};
    const int periodic[3] = {0, 0, 0};
    const double L[3] = {4.0, 1.0, 1.0};
    const double tip = 0.05, amp = 0.01, pi = 3.14159265358979323846;
    int na[3];
    for (int d = 0; d < 3; d++) {
        if (nelem[d] < 1) {
            fprintf(stderr, "element counts must be positive\n");
            return 2;
        }
        na[d] = nelem[d] + 2;
    }
    const size_t ncp = (size_t)na[0] * na[1] * na[2];

    /* ---- descriptors: igabeam patch, SVK + gradient physics, clamped + tip */
    double *knot[3], *w = xmalloc(sizeof(double) * ncp), *X = xmalloc(sizeof(double) * 3 * ncp);
    for (int d = 0; d < 3; d++) knot[d] = xmalloc(sizeof(double) * (nelem[d] + 5));
    CHECK(igap_mesh_igabeam(L, nelem, periodic, 0, knot[0], knot[1], knot[2], w, X));
    igap_mesh_desc mesh;
    memset(&mesh, 0, sizeof mesh);
    mesh.porder = 2;
    for (int d = 0; d < 3; d++) {
        mesh.nelem[d] = nelem[d];
        mesh.periodic[d] = periodic[d];
        mesh.knot[d] = knot[d];
    }
    mesh.w = w;
    mesh.X = X;

    igap_phys_desc phys;
    igap_phys_gradelast(&phys);
    const double svk[6] = {1.0, 0.0, 0.0, 2.5, 2.0, 0.1};   /* lambda = mu = 1, l = 0.1 */
    phys.use_coef = 1;
    for (int i = 0; i < 6; i++) phys.coef[i] = svk[i];

    igap_bc_desc bc;
    igap_bc_default(&bc);
    for (int order = 0; order < 2; order++)
        for (int dof = 0; dof < 3; dof++) bc.bc_type[6 * 0 + 3 * order + dof] = 0;   /* face 0 clamped */
    bc.bc_type[6 * 1 + 2] = 0;                                                        /* face 1: u_z */
    bc.par_dirichlet[3 * 2] = tip;                                                    /* u_z = tip xi_0 */

    igap_gpu_ctx *ctx = NULL;
    CHECK(igap_gpu_create(&mesh, &phys, &bc, 0, 0, &ctx));
    int64_t ndof, nnz, nel;
    int torder;
    CHECK(igap_gpu_sizes(ctx, &ndof, &nnz, &nel, &torder));
    if ((size_t)ndof != 3 * ncp) {
        fprintf(stderr, "unexpected ndof %lld\n", (long long)ndof);
        return 1;
    }

    /* ---- smooth state U_i = amp sin(pi x/Lx) cos(pi y/Ly) cos(pi z/Lz) (1, 1/2, 1/4)_i */
    double *U = xmalloc(sizeof(double) * ndof);
    for (size_t p = 0; p < ncp; p++) {
        const double f = amp * sin(pi * X[3 * p] / L[0]) * cos(pi * X[3 * p + 1] / L[1]) * cos(pi * X[3 * p + 2] / L[2]);
        U[3 * p] = f;
        U[3 * p + 1] = 0.5 * f;
        U[3 * p + 2] = 0.25 * f;
    }
    void *dU = NULL;
    CHECK(igap_gpu_malloc(ctx, &dU, (int64_t)sizeof(double) * ndof));
    CHECK(igap_gpu_memcpy_htod(ctx, dU, U, (int64_t)sizeof(double) * ndof));

    /* ---- the two callbacks */
    double *R = xmalloc(sizeof(double) * ndof), *vals = xmalloc(sizeof(double) * nnz);
    int64_t *rowptr = xmalloc(sizeof(int64_t) * (ndof + 1));
    int32_t *col = xmalloc(sizeof(int32_t) * nnz);
    CHECK(igap_gpu_residual(ctx, (const double *)dU, NULL));
    CHECK(igap_gpu_memcpy_dtoh(ctx, R, igap_gpu_rvec(ctx), (int64_t)sizeof(double) * ndof));
    CHECK(igap_gpu_tangent(ctx, (const double *)dU, NULL, IGAP_F64));
    CHECK(igap_gpu_memcpy_dtoh(ctx, vals, igap_gpu_csr_val64(ctx), (int64_t)sizeof(double) * nnz));
    CHECK(igap_gpu_memcpy_dtoh(ctx, rowptr, igap_gpu_csr_rowptr(ctx), (int64_t)sizeof(int64_t) * (ndof + 1)));
    CHECK(igap_gpu_memcpy_dtoh(ctx, col, igap_gpu_csr_col(ctx), (int64_t)sizeof(int32_t) * nnz));

    /* ---- the load step from zero: Newton to ||R|| <= 1e-10 ||R0|| */
    memset(U, 0, sizeof(double) * ndof);
    CHECK(igap_gpu_memcpy_htod(ctx, igap_gpu_uhist(ctx), U, (int64_t)sizeof(double) * ndof));
    igap_solver_opts opts;
    CHECK(igap_gpu_solver_default_opts(&opts));
    opts.snes_rtol = 1e-10;
    opts.snes_atol = 1e-30;
    igap_solver_stats st;
    CHECK(igap_gpu_snes_solve(ctx, &opts, &st));
    CHECK(igap_gpu_memcpy_dtoh(ctx, U, igap_gpu_uhist(ctx), (int64_t)sizeof(double) * ndof));
    printf("c host: %lld dofs, %lld nonzeros, newton %d (converged %d), |R| %.3e -> %.3e\n", (long long)ndof,
           (long long)nnz, st.newton_its, st.converged, st.fnorm0, st.fnorm);

    FILE *f = fopen(argv[4], "wb");
    if (!f) {
        fprintf(stderr, "cannot open %s\n", argv[4]);
        return 1;
    }
    put(f, &ndof, sizeof ndof, 1);
    put(f, &nnz, sizeof nnz, 1);
    put(f, R, sizeof(double), (size_t)ndof);
    put(f, rowptr, sizeof(int64_t), (size_t)ndof + 1);
    put(f, col, sizeof(int32_t), (size_t)nnz);
    put(f, vals, sizeof(double), (size_t)nnz);
    const int32_t its = st.newton_its, conv = st.converged;
    put(f, &its, sizeof its, 1);
    put(f, &conv, sizeof conv, 1);
    put(f, &st.fnorm, sizeof st.fnorm, 1);
    put(f, U, sizeof(double), (size_t)ndof);
    if (fclose(f) != 0) {
        fprintf(stderr, "write to %s failed\n", argv[4]);
        return 1;
    }

    CHECK(igap_gpu_free(ctx, dU));
    CHECK(igap_gpu_destroy(ctx));
    for (int d = 0; d < 3; d++) free(knot[d]);
    free(w);
    free(X);
    free(U);
    free(R);
    free(vals);
    free(rowptr);
    free(col);
    return 0;
}
