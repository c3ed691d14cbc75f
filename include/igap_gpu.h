/* igap_gpu.h -- C ABI of the MI355X residual/tangent assembly for IGAP4.
 *
 * One patch lives in one GPU's HBM.  A context is built from descriptors that
 * carry the same data as the reference's Mesh / Phys / BC structs, and its two
 * assembly entry points have the semantics of the Assm callbacks that IGAP4
 * hands to PETSc SNES:
 *
 *   reference                                         this library
 *   ------------------------------------------------  -------------------------------
 *   Residual_ptr  Residual_iga(SNES,Vec U,Vec R,ctx)   igap_gpu_residual(ctx,U,R)
 *     include/assmimpl.h:7, src/assmcore_iga.c:16-502
 *   Tangent_ptr   Tangent_iga(SNES,Vec U,Mat J,Mat P,   igap_gpu_tangent(ctx,U,vals,prec)
 *     ctx)  include/assmimpl.h:8, assmcore_iga.c:507-858
 *   IGAPAssmInit  src/interface.c:131-203               igap_gpu_create
 *   setnz_iga + MatCreateAIJ  assmcore_iga.c:1054-1087  igap_gpu_csr_* (device CSR)
 *   IGAPFinalize  src/interface.c:372-435               igap_gpu_destroy
 *
 * All device pointers are HIP device addresses on the context's GPU; every
 * call is stream-ordered on the context's stream and returns before the
 * kernels finish unless stated otherwise.  Every function returns an int
 * status (IGAP_OK = 0) instead of aborting; igap_gpu_last_error() gives text.
 * A context is used from one host thread at a time.
 */
#ifndef IGAP_GPU_H
#define IGAP_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    IGAP_OK = 0,
    IGAP_ERR_ARG = 1,      /* invalid argument / unsupported configuration */
    IGAP_ERR_HIP = 2,      /* HIP runtime error (message has the HIP text) */
    IGAP_ERR_NOMEM = 3,    /* device allocation failed */
    IGAP_ERR_STATE = 4     /* call out of order (e.g. no pattern built) */
};

enum { IGAP_PHYS_GRADELAST = 0, IGAP_PHYS_GRADELASTTIME = 2 };  /* = torder */
enum { IGAP_F64 = 0, IGAP_F32 = 1 };

/* Mesh of one patch: struct __Mesh (include/meshimpl.h:6-29) for one rank
 * after mpipart_iga (src/meshcore_iga.c:211-228).  Control points use the
 * "active" numbering id = (na1*na2)*i + na2*j + k with na_d = nelem_d + 2
 * (src/assmcore_iga.c:110-111,138); in a periodic direction the last two
 * layers alias the first two.  porder must be 2. */
typedef struct {
    int           porder;
    int           nelem[3];
    int           periodic[3];
    const double *knot[3];   /* knotVector[d], nelem[d]+1+2*porder entries (host) */
    const double *w;         /* wVector, na0*na1*na2 (host)                        */
    const double *X;         /* XVector, 3*na0*na1*na2 (host)                      */
} igap_mesh_desc;

/* Physics: struct __Phys (include/physimpl.h:13-28) restricted to the two
 * reference physics.  par = par_mat {r, curv, e0, e345, le, rho, cii, dt, t}
 * (phys/physgradelast.c:21).  use_coef = 0 derives the energy coefficients
 * (c2,c3,c4,e0,e345,le) from par as phys/gradelast.c:46-48 does; use_coef = 1
 * takes coef[] as given (e.g. St Venant-Kirchhoff: mu,0,0,3lambda/2+mu,2mu,l). */
typedef struct {
    int    kind;             /* IGAP_PHYS_GRADELAST or IGAP_PHYS_GRADELASTTIME */
    double par[9];
    int    use_coef;
    double coef[6];
} igap_phys_desc;

/* Boundary conditions: struct __BC (include/bcimpl.h:8-15).  bc_type index
 * (2*3)*face + 3*order + dof, 0 = Dirichlet, 1 = Neumann.  The Neumann
 * callback is the constant table of neumann_default_iga (assmcore_iga.c:
 * 1036-1052) with user values, neumann[(2*3)*face + 3*order + dof]. */
typedef struct {
    int    bc_type[36];
    double par_periodic[9];
    double par_dirichlet[9];
    double neumann[36];
} igap_bc_desc;

typedef struct igap_gpu_ctx igap_gpu_ctx;

/* ---- set-up ------------------------------------------------------------ */
/* Defaults of the reference constructors: PhysGradElast / PhysGradElastTime
 * (phys/physgradelast.c:9-35, phys/physgradelasttime.c:9-34) and
 * IGAPBCAddDefault (src/interface.c:60-74). */
void igap_phys_gradelast(igap_phys_desc *ph);
void igap_phys_gradelasttime(igap_phys_desc *ph);
void igap_bc_default(igap_bc_desc *bc);

/* igabeam / igabeamtest geometry (mesh/igabeam.c:9-73, mesh/igabeamtest.c:68-70)
 * with per-direction element counts.  knot[d] holds nelem[d]+5 doubles, w
 * na-product, X 3*na-product; all caller-owned host memory. */
int igap_mesh_igabeam(const double L[3], const int nelem[3], const int periodic[3], int perturbed,
                      double *knot0, double *knot1, double *knot2, double *w, double *X);

/* How igap_gpu_tangent turns element tangents into CSR values (flags of
 * igap_gpu_create; 0 chooses at the first tangent call of each precision:
 * element matrices, the faster path, while their buffer fits in free device
 * memory with a tenth of the device to spare, direct scatter otherwise on
 * non-periodic patches).  Both sum every entry in a fixed order (bitwise
 * repeatable) and give the same values up to the order of summation.
 *   IGAP_TANGENT_ELEMENT_MATRICES  the element matrices are written to HBM
 *       (6561 values per element, allocated on the first tangent call) and a
 *       gather sums each entry's contributions in ascending element order
 *   IGAP_TANGENT_DIRECT            the element kernel adds its matrix into the
 *       CSR values, one launch per class of the 27-colouring (the elements of
 *       one class share no control point); non-periodic patches only
 * Element kernels of the static physics: on a separable patch -- every
 * control coordinate x_d depends on the control point's index along d alone
 * and all weights are equal, as on the igabeam patches -- the residual and
 * the element tangents (either way to the CSR values) come from
 * sum-factorised kernels over 1-D tables of physical derivatives; otherwise,
 * and with the flag below, from the general-geometry kernels.  Both give the
 * same values up to rounding.
 *   IGAP_GENERAL_GEOMETRY          the general-geometry kernels on every patch */
enum { IGAP_TANGENT_ELEMENT_MATRICES = 1, IGAP_TANGENT_DIRECT = 2, IGAP_GENERAL_GEOMETRY = 4 };

/* IGAPAssmInit (src/interface.c:131-203) for one GPU: uploads the patch,
 * builds the CSR pattern, the Dirichlet row set and the Neumann load, and
 * allocates U_hist, R and the matrix values.  device < 0 keeps the current
 * device.  flags: 0, or one of IGAP_TANGENT_ELEMENT_MATRICES and
 * IGAP_TANGENT_DIRECT, optionally with IGAP_GENERAL_GEOMETRY. */
int igap_gpu_create(const igap_mesh_desc *mesh, const igap_phys_desc *phys, const igap_bc_desc *bc,
                    int device, int flags, igap_gpu_ctx **ctx);
/* IGAPFinalize (src/interface.c:372-435) for the context: frees everything
 * even after a failure and returns the first failure; NULL is a no-op. */
int igap_gpu_destroy(igap_gpu_ctx *ctx);
/* Message of the last failing call of this thread (no reference
 * counterpart: the reference asserts or exits instead of returning). */
const char *igap_gpu_last_error(void);

/* Replace the boundary conditions without rebuilding the context: what a
 * reference driver does by changing bc->par_dirichlet (the load factor of a
 * load step) between IGAPSolvSolve calls (src/interface.c:302-328). */
int igap_gpu_set_bc(igap_gpu_ctx *ctx, const igap_bc_desc *bc);

/* ---- sizes and device buffers owned by the context ---------------------- */
/* core->nDof, the nonzeros of the setnz_iga pattern (src/assmcore_iga.c:
 * 1054-1087), the element count and phys->torder */
int igap_gpu_sizes(const igap_gpu_ctx *ctx, int64_t *ndof, int64_t *nnz, int64_t *nelem, int *torder);
/* core->U_hist (include/coreimpl.h:17): (torder+1)*ndof doubles, level h at
 * offset h*ndof, 0 = current iterate (src/interface.c:127). */
double *igap_gpu_uhist(igap_gpu_ctx *ctx);
/* iga_intdensity (src/postcore_iga.c:13-379, IGAPPostComputeHH
 * src/interface.c:341-346) of the current U_hist: HH[i] = sum over Gauss
 * points of w_q hh_i(u_q) minus the Neumann face work.  gradelast: 13 values
 * (energy, 2 dPsi/dE_IJ for 11,12,13,22,23,33, six phase fractions);
 * gradelasttime: 2 (energy with the kinetic term, volume).  HH is a host
 * array of 13; *ndensity receives the count.  Synchronous. */
int igap_gpu_intdensity(igap_gpu_ctx *ctx, double *HH, int *ndensity);
/* Solution files (IGAPSolvSaveSolution / IGAPSolvLoadSolution,
 * src/interface.c:271-300, write_solution / read_solution src/ioutil.c:90-159):
 * raw FP64 in the reference's universal dof order (on one GPU the owned order
 * 3*cp + i), holding U_hist level 0 for the static physics and levels
 * 1..torder for the transient one.  Loading a transient file also sets level
 * 0 = level 1 (the next step's initial iterate).  Synchronous; a short or
 * missing file is an error, not an abort. */
int igap_gpu_save_solution(igap_gpu_ctx *ctx, const char *fname);
int igap_gpu_load_solution(igap_gpu_ctx *ctx, const char *fname);
/* device CSR pattern of the AIJ matrix IGAPAssmInit preallocates
 * (src/interface.c:187-198): rows = owned dofs 3*cp+i, columns sorted */
const int64_t *igap_gpu_csr_rowptr(const igap_gpu_ctx *ctx);
const int32_t *igap_gpu_csr_col(const igap_gpu_ctx *ctx);
/* values of that matrix (assm->TT) owned by the context, FP64 and FP32
 * arrays of nnz entries, allocated on first use */
double *igap_gpu_csr_val64(igap_gpu_ctx *ctx);
float *igap_gpu_csr_val32(igap_gpu_ctx *ctx);
/* Dirichlet row mask (int32, 1 = constrained) and the values g that
 * Residual_iga inserts as U_d - g_d (dirichlet_iga, src/assmcore_iga.c:861-934) */
const int32_t *igap_gpu_dirichlet_mask(const igap_gpu_ctx *ctx);
const double *igap_gpu_dirichlet_values(const igap_gpu_ctx *ctx);

/* ---- the assembly callbacks -------------------------------------------- */
/* Residual_iga: copies U (ndof doubles, device) into U_hist level 0, then
 * R = volume integral - Neumann load, Dirichlet rows R_d = U_d - g_d.
 * R may be NULL to write the context's own R buffer (igap_gpu_rvec). */
int igap_gpu_residual(igap_gpu_ctx *ctx, const double *U, double *R);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): igap_gpu_residual with the element vectors in a device
 * array the library's code object carries (256 MiB on every GPU it runs on,
 * kept until the process ends), so the call allocates nothing.  At most
 * 414,252 elements (81 doubles each); a larger patch is IGAP_ERR_ARG.  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU (two
 * streams, two threads) race and must not be issued. */
int igap_gpu_residual_noalloc(igap_gpu_ctx *ctx, const double *U, double *R);
/* the context's residual vector (assm->RR, src/interface.c:187-198) */
double *igap_gpu_rvec(igap_gpu_ctx *ctx);
/* Tangent_iga: copies U into U_hist level 0 (U may be NULL to use U_hist as
 * it is), zeroes the values, assembles the volume tangent deterministically
 * and applies MatZeroRowsColumns(diag = 1) to the Dirichlet rows.
 * prec = IGAP_F64 writes vals as double, IGAP_F32 as float (FP32 MFMA);
 * vals may be NULL for the context's own arrays. */
int igap_gpu_tangent(igap_gpu_ctx *ctx, const double *U, void *vals, int prec);
/* MatMult (PETSc's) on the tangent IGAP4 assembles into assm->TT
 * (src/interface.c:191,219): y = K x with the context's own tangent values
 * of precision prec (those of the last igap_gpu_tangent with vals = NULL,
 * or of the last Newton refresh), FP64 sums.  Device vectors of ndof
 * doubles, x != y; stream-ordered.  IGAP_ERR_STATE when the context holds
 * no tangent values of that precision. */
int igap_gpu_tangent_apply(igap_gpu_ctx *ctx, int prec, const double *x, double *y);

/* Pointwise physics on the device with the context's physics: the Phys
 * callbacks residual / tangent / density (typedefs residual_ptr, tangent_ptr,
 * density_ptr, include/physimpl.h:6-9; set to gradelast_* by PhysGradElast,
 * phys/physgradelast.c:25-28, and to gradelasttime_* by PhysGradElastTime,
 * phys/physgradelasttime.c:25-28), for npts jets in one launch, by the device
 * code the assembly kernels use.
 *   u : npts x 30*(torder+1) doubles, u[30 h + 10 k + a] with h = 0 the
 *       current level, 1 the previous and 2 the one before (the reference's
 *       u, u1 = u + 30, u0 = u + 60; jet slot a: value, 3 first and 6 second
 *       derivatives 11,12,13,22,23,33)
 *   r : npts x 30, as residual_ptr writes it (NULL: not evaluated)
 *   C : npts x 900 row-major, index (10 k + a)*30 + (10 l + b), as tangent_ptr
 *       writes it (NULL: not evaluated)
 *   hh: npts x 13 (gradelast) or npts x 2 (gradelasttime), as density_ptr
 *       writes it (NULL: not evaluated)
 * Device pointers on the context's device; stream-ordered on its stream. */
int igap_gpu_phys_eval(igap_gpu_ctx *ctx, int64_t npts, const double *u, double *r, double *C, double *hh);

/* ---- Newton-Krylov solve (IGAPSolvInit / IGAPSolvSolve, src/interface.c:
 * 207-250, 302-328: SNES NEWTONLS with BT line search, KSP GMRES right
 * preconditioned, PC ILU(0)) with every vector in HBM. ---------------------- */
/* IGAP_PC_MG (no counterpart in the reference, which uses ASM/ILU(0)):
 * geometric multigrid V-cycle on the patch's B-spline hierarchy (every even
 * direction halved per level; exact knot-insertion prolongation; coarse
 * tangents re-assembled at the interpolated state; degree-3 Chebyshev
 * smoothing with 3x3 point-block Jacobi; the coarsest level applied as its
 * explicit inverse).  Non-periodic patches with unit weights only. */
enum { IGAP_PC_ILU0 = 0, IGAP_PC_JACOBI = 1, IGAP_PC_NONE = 2, IGAP_PC_MG = 3 };

typedef struct {
    double snes_atol;     /* stop when ||R|| <= snes_atol (reference: 1e-10, interface.c:220) */
    double snes_rtol;     /* and/or when ||R|| <= snes_rtol ||R0|| (0 disables)             */
    int    snes_maxit;
    double ksp_rtol;      /* GMRES relative tolerance (PETSc default 1e-5)                  */
    double ksp_atol;
    int    ksp_maxit;     /* Krylov iterations per Newton step                              */
    int    ksp_restart;   /* GMRES restart (reference 10000; bounded to fit HBM)            */
    int    pc;            /* IGAP_PC_ILU0 (the reference's ASM/ILU(0) on one part), ...     */
    int    tangent_prec;  /* IGAP_F64 or IGAP_F32 (FP32 matrix-core tangent, FP32 ILU)      */
    int    linesearch;    /* 1: backtracking (SNESLINESEARCHBT, cubic), 0: full steps       */
    int    pc_order;      /* ILU(0) ordering: 0 natural dof order (the reference's), 1     */
                          /* 27 colours of control points (colour 9(i%3)+3(j%3)+k%3, the   */
                          /* three dofs of a point together): ~25x fewer dependent levels  */
                          /* in the triangular solves, a different incomplete factor       */
    int    pc_lag;        /* refactor the preconditioner every pc_lag-th Newton iteration  */
                          /* of a solve (1, the reference: every iteration; k > 1 keeps    */
                          /* the factor of an earlier tangent, SNESSetLagPreconditioner)   */
    int    monitor;       /* 1: one line per Newton iteration on stderr (-snes_monitor)    */
    int    mg_levels;     /* IGAP_PC_MG: at most this many levels (0: up to 8)            */
    int    pc_fp32;       /* 1: preconditioner (ILU(0) factor or multigrid levels) in FP32 */
                          /* with an FP64 tangent: Newton and GMRES keep the FP64 operator */
                          /* 2: as 1, and the multigrid levels smooth with 16-bit copies   */
                          /* of their matrices (binary16 blocks, a power-of-two scale per  */
                          /* row; 20 instead of 36 bytes per 3x3 block); ILU(0) as 1       */
} igap_solver_opts;

typedef struct {
    int    newton_its, ksp_its, ls_backtracks, n_residual;
    int    converged;     /* > 0 converged (2 abs, 3 rel); < 0 diverged, PETSc's codes:      */
                          /* -3 linear solve not converged within ksp_maxit, -4 NaN,        */
                          /* -5 snes_maxit reached, -6 line search failed or the stagnation */
                          /* stop of IGAPSolvSolve (|F| changed < 1e-6 relative in a step,  */
                          /* src/interface.c:318)                                           */
    double fnorm0, fnorm, last_ksp_relres;
    double t_residual_ms, t_tangent_ms, t_pc_ms, t_ksp_ms, t_total_ms;
} igap_solver_stats;

int igap_gpu_solver_default_opts(igap_solver_opts *opts);
/* Solves R(U) = 0 starting from U_hist level 0 (the reference keeps solv->U
 * between solves, src/interface.c:252-259,319); on return U_hist level 0 holds
 * the last iterate and the context's R buffer its residual.  Synchronous. */
int igap_gpu_snes_solve(igap_gpu_ctx *ctx, const igap_solver_opts *opts, igap_solver_stats *stats);

/* PCApply (PETSc's, on the PC of IGAP4's KSP, src/interface.c:224-239) with
 * the preconditioner the last igap_gpu_snes_solve set up (lagged: the
 * tangent of its last refresh): x = M^-1 b.  IGAP_PC_APPLY_SMOOTHER applies
 * only the multigrid's finest-level smoother from x = 0 (KSPSolve on the
 * level-0 smoother of a PCMG): degree-3 Chebyshev with 3x3 point-block
 * Jacobi, which damps the components of b above 2% of the spectrum and
 * leaves the smooth ones.  Device vectors of ndof doubles, b != x;
 * stream-ordered.  IGAP_ERR_STATE before the first solve. */
enum { IGAP_PC_APPLY_FULL = 0, IGAP_PC_APPLY_SMOOTHER = 1 };
int igap_gpu_pc_apply(igap_gpu_ctx *ctx, int which, const double *b, double *x);

/* ---- helpers (no reference counterpart: device memory for callers that
 * hold no HIP runtime of their own; all on the context's device and stream,
 * the copies synchronous) -------------------------------------------------- */
int igap_gpu_synchronize(igap_gpu_ctx *ctx);
int igap_gpu_malloc(igap_gpu_ctx *ctx, void **ptr, int64_t bytes);
int igap_gpu_free(igap_gpu_ctx *ctx, void *ptr);
int igap_gpu_memcpy_htod(igap_gpu_ctx *ctx, void *dst, const void *src, int64_t bytes);
int igap_gpu_memcpy_dtoh(igap_gpu_ctx *ctx, void *dst, const void *src, int64_t bytes);
int igap_gpu_memcpy_dtod(igap_gpu_ctx *ctx, void *dst, const void *src, int64_t bytes);

/* (No reference counterpart.)  Timing of the assembly kernels with HIP
 * events recorded on the context's stream around each launch.
 * which: 0 residual volume kernel, 1 tangent
 * volume kernel, 2 CSR gather kernel, 3 whole residual call, 4 whole tangent
 * call.  Accumulated milliseconds and launch counts since the last reset. */
int igap_gpu_timing_enable(igap_gpu_ctx *ctx, int on);
int igap_gpu_timing_get(igap_gpu_ctx *ctx, int which, double *ms_total, int64_t *count);
int igap_gpu_timing_reset(igap_gpu_ctx *ctx);

/* (No reference counterpart.)  Test hook: the n-th device allocation the
 * library makes from now on (n = 0: the next one) fails as if the device were
 * out of memory, once; n < 0 disarms it.  For tests that check that a call
 * failing part way leaves every object usable (the next call rebuilds what
 * the failed one did not finish). */
int igap_gpu_debug_fail_alloc(int n);

/* (No reference counterpart.)  Self test of the matrix-core operand layout
 * the tangent kernels assume:
 * one 16x16x4 product with V_MFMA_F64_16X16X4F64 and one with
 * V_MFMA_F32_16X16X4F32 of A[i][k] = 1 + i + k/4, B[k][j] = 2 - j/2 + k,
 * lane l supplying A[l&15][l>>4] and B[l>>4][l&15]; out64/out32 receive the
 * 4 x 64 raw result registers (out[r*64 + l]).  Host pointers. */
int igap_gpu_selftest_mfma(double *out64, float *out32);

/* Self test of the solver's own ILU(0) factorisation (no reference
 * counterpart; the reference's factor is PETSc's PCILU): factors the tangent
 * last assembled in precision prec in the 27-colour order once with the
 * library's point-block kernels and once with rocSPARSE csrilu0, and returns
 * *diff = max|a - b| / max|b| over the factor entries.  *applicable = 0 when
 * the pattern lacks the point-block colour structure (a periodic direction
 * whose size is not a multiple of 3), in which case the solver always uses
 * rocSPARSE. */
int igap_gpu_selftest_ilu(igap_gpu_ctx *ctx, int prec, int *applicable, double *diff);

#ifdef __cplusplus
}
#endif
#endif
