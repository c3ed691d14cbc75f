// igap_capi.cpp -- C ABI of the MI355X IGA assembly (include/igap_gpu.h).
//
// Host side of IGAPAssmInit / Residual_iga / Tangent_iga for one patch on one
// GPU: descriptor validation, device layout, the CSR pattern, the Dirichlet
// row set and values (dirichlet_iga, src/assmcore_iga.c:861-934), the
// U-independent Neumann load (:273-405) and stream-ordered launches.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/igap_gpu.h"
#include "igap_ctx.h"
#include "igap_launch.h"

using namespace igap;

static thread_local std::string g_err;
// allocations still to succeed before one fails (< 0: off); a test hook
static std::atomic<int> g_fail_alloc{-1};

hipError_t igap::dmalloc_raw(void **p, size_t bytes)
{
    int n = g_fail_alloc.load();
    while (n >= 0) {
        if (g_fail_alloc.compare_exchange_weak(n, n - 1)) {
            if (n == 0) {
                *p = nullptr;
                return hipErrorOutOfMemory;
            }
            break;
        }
    }
    return hipMalloc(p, bytes);
}

extern "C" int igap_gpu_debug_fail_alloc(int n)
{
    g_fail_alloc.store(n < 0 ? -1 : n);
    return IGAP_OK;
}

int igap::fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// ---------------------------------------------------------------- small math
static void gauss4(double *x, double *w)
{
    double a = std::sqrt(3.0 / 7.0 - 2.0 / 7.0 * std::sqrt(6.0 / 5.0));
    double b = std::sqrt(3.0 / 7.0 + 2.0 / 7.0 * std::sqrt(6.0 / 5.0));
    double wa = (18.0 + std::sqrt(30.0)) / 36.0, wb = (18.0 - std::sqrt(30.0)) / 36.0;
    double t[4] = {-b, -a, a, b}, ww[4] = {wb, wa, wa, wb};
    for (int i = 0; i < 4; i++) {
        x[i] = 0.5 * (t[i] + 1.0);
        w[i] = 0.5 * ww[i];
    }
}

// Nodes and weights of the transient stress rule: the functional that maps
// p(s) = sum_k a_k s^k (degree <= 7) to a_0 + a_1/2 + a_2/3 + a_3/4 (the
// reference's Taylor series of the average-vector-field stress truncated
// after the d^3 term, phys/gradelasttime.c:4129-4158) and its tangent
// counterpart q -> (b_0/2 + b_1/3 + b_2/4)/2 (:4160-9181), as weighted sums
// over eight Chebyshev nodes in [-2, 2].
static void trunc_rule(double *s, double *cr, double *ct)
{
    long double V[8][8], br[8], bt[8], ns[8];
    const long double PI = 3.14159265358979323846264338327950288L;
    for (int j = 0; j < 8; j++) ns[j] = 2.0L * cosl(PI * (j + 0.5L) / 8.0L);
    for (int k = 0; k < 8; k++) {
        br[k] = (k <= 3) ? 1.0L / (k + 1) : 0.0L;
        bt[k] = (k <= 2) ? 0.5L / (k + 2) : 0.0L;
        for (int j = 0; j < 8; j++) {
            long double v = 1.0L;
            for (int t = 0; t < k; t++) v *= ns[j];
            V[k][j] = v;
        }
    }
    for (int c = 0; c < 8; c++) {
        int piv = c;
        for (int r = c + 1; r < 8; r++)
            if (fabsl(V[r][c]) > fabsl(V[piv][c])) piv = r;
        for (int j = 0; j < 8; j++) std::swap(V[c][j], V[piv][j]);
        std::swap(br[c], br[piv]);
        std::swap(bt[c], bt[piv]);
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

// N_{i,p}^{(k)}(xi) with the end-point convention of evalN (src/meshcore_iga.c:329-384)
static double evalN(const double *kV, int nk, int k, int i, int p, double xi)
{
    if (k == 0) {
        if (i == nk - p - 2 && std::fabs(xi - kV[nk - 1]) < 1.e-15 && std::fabs(kV[nk - 1] - kV[nk - 2]) < 1.e-15) return 1.0;
        if (p == 0) return (xi >= kV[i] && kV[i + 1] > xi) ? 1.0 : 0.0;
        bool lz = kV[i + p] == kV[i], rz = kV[i + p + 1] == kV[i + 1];
        if (lz && rz) return 0.0;
        if (lz) return (kV[i + p + 1] - xi) / (kV[i + p + 1] - kV[i + 1]) * evalN(kV, nk, 0, i + 1, p - 1, xi);
        if (rz) return (xi - kV[i]) / (kV[i + p] - kV[i]) * evalN(kV, nk, 0, i, p - 1, xi);
        return (xi - kV[i]) / (kV[i + p] - kV[i]) * evalN(kV, nk, 0, i, p - 1, xi)
             + (kV[i + p + 1] - xi) / (kV[i + p + 1] - kV[i + 1]) * evalN(kV, nk, 0, i + 1, p - 1, xi);
    }
    if (k < 0 || k > p) return 0.0;
    double al[4][4];
    al[0][0] = 1.0;
    for (int r = 1; r <= k; r++) {
        double den = kV[i + p - r + 1] - kV[i];
        al[r][0] = den == 0.0 ? 0.0 : al[r - 1][0] / den;
        for (int c = 1; c < r; c++) {
            den = kV[i + p + c - r + 1] - kV[i + c];
            al[r][c] = den == 0.0 ? 0.0 : (al[r - 1][c] - al[r - 1][c - 1]) / den;
        }
        den = kV[i + p + 1] - kV[i + r];
        al[r][r] = den == 0.0 ? 0.0 : -al[r - 1][r - 1] / den;
    }
    double s = 0.0;
    for (int j = 0; j <= k; j++) s += al[k][j] * evalN(kV, nk, 0, i + j, p - k, xi);
    double f = 1.0;
    for (int t = p; t > p - k; t--) f *= t;
    return s * f;
}

// ---------------------------------------------------------------- descriptors
extern "C" void igap_phys_gradelast(igap_phys_desc *ph)
{
    const double par[9] = {0.25, 0.50, 360., 180., 0.05, 1.0, 1.0, 1.e-3 * (1. / 2.), 0.0};
    std::memset(ph, 0, sizeof *ph);
    ph->kind = IGAP_PHYS_GRADELAST;
    std::memcpy(ph->par, par, sizeof par);
}

extern "C" void igap_phys_gradelasttime(igap_phys_desc *ph)
{
    const double par[9] = {0.25, 0.50, 500., 250., 0.025, 1.0, 1.0, 1.e-3 * (1. / 2.), 0.0};
    std::memset(ph, 0, sizeof *ph);
    ph->kind = IGAP_PHYS_GRADELASTTIME;
    std::memcpy(ph->par, par, sizeof par);
}

extern "C" void igap_bc_default(igap_bc_desc *bc)
{
    std::memset(bc, 0, sizeof *bc);
    for (int i = 0; i < 36; i++) bc->bc_type[i] = 1;
}

extern "C" int igap_mesh_igabeam(const double L[3], const int nelem[3], const int periodic[3], int perturbed, double *knot0,
                                 double *knot1, double *knot2, double *w, double *X)
{
    double *kv[3] = {knot0, knot1, knot2};
    const int p = 2;
    for (int d = 0; d < 3; d++) {
        if (nelem[d] < 1) return fail(IGAP_ERR_ARG, "nelem[%d] = %d < 1", d, nelem[d]);
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
    int64_t ia = 0;
    for (int i = 0; i < nb[0]; i++)
        for (int j = 0; j < nb[1]; j++)
            for (int k = 0; k < nb[2]; k++) {
                w[ia] = 1.0;
                X[3 * ia + 0] = ((double)i - 0.5) / ((double)(nb[0] - 2)) * L[0];
                X[3 * ia + 1] = ((double)j - 0.5) / ((double)(nb[1] - 2)) * L[1];
                X[3 * ia + 2] = ((double)k - 0.5) / ((double)(nb[2] - 2)) * L[2];
                if (perturbed) {
                    double in = (double)(i > 2 && j > 2 && k > 2 && i < nb[0] - 2 && j < nb[1] - 2 && k < nb[2] - 2);
                    X[3 * ia + 0] += 0.2 * L[0] / nb[0] * in * std::sin(3 * i * k * j);
                    X[3 * ia + 1] += 0.2 * L[1] / nb[1] * in * std::sin(5 * i * k * j);
                    X[3 * ia + 2] += 0.2 * L[2] / nb[2] * in * std::sin(7 * i * k * j);
                }
                ia++;
            }
    return IGAP_OK;
}

extern "C" const char *igap_gpu_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------- BC set-up
static void face_dirs(int f, int &d0, int &d1, int &dn)
{
    switch (f / 2) {
    case 0: d0 = 1; d1 = 2; dn = 0; break;
    case 1: d0 = 0; d1 = 2; dn = 1; break;
    default: d0 = 0; d1 = 1; dn = 2; break;
    }
}

static inline int64_t dof_active(const igap_gpu_ctx *c, int ax, int ay, int az, int i)
{
    int ox = own1(ax, c->nb[0], c->periodic[0]), oy = own1(ay, c->nb[1], c->periodic[1]), oz = own1(az, c->nb[2], c->periodic[2]);
    return 3 * (((int64_t)c->nb[1] * c->nb[2]) * ox + (int64_t)c->nb[2] * oy + oz) + i;
}

static inline int64_t dof_dirs(const igap_gpu_ctx *c, int d0, int d1, int dn, int a0, int a1, int an, int i)
{
    int a[3];
    a[d0] = a0; a[d1] = a1; a[dn] = an;
    return dof_active(c, a[0], a[1], a[2], i);
}

// affine Dirichlet data u = sum_d par[3*idof+d] xi_d on a face (dirichlet_iga,
// src/assmcore_iga.c:861-934): control values reproducing the linear function
// along each in-face direction plus the normal coordinate term.
static void dirichlet_face(const igap_gpu_ctx *c, const double *par, int f, int order, int idof, std::vector<double> &out)
{
    const int p = 2;
    int d0, d1, dn;
    face_dirs(f, d0, d1, dn);
    int dd[2] = {d0, d1};
    int nkn = c->nelem[dn] + 5;
    double Z = (f % 2 == 0) ? c->knot_h[dn][0] : c->knot_h[dn][nkn - 1] - 1.e-15;
    double pr[3];
    switch (f / 2) {
    case 0: pr[0] = par[3 * idof + 1]; pr[1] = par[3 * idof + 2]; pr[2] = par[3 * idof + 0]; break;
    case 1: pr[0] = par[3 * idof + 0]; pr[1] = par[3 * idof + 2]; pr[2] = par[3 * idof + 1]; break;
    default: pr[0] = par[3 * idof + 0]; pr[1] = par[3 * idof + 1]; pr[2] = par[3 * idof + 2]; break;
    }
    std::vector<double> d1d[2];
    for (int b = 0; b < 2; b++) {
        int d = dd[b], nb = c->nb[d], nk = c->nelem[d] + 5;
        const double *kv = c->knot_h[d].data();
        d1d[b].assign(nb, 0.0);
        double W = kv[p];
        double vec[2] = {pr[b] * W, pr[b]};
        double mt[4];
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) mt[2 * i + j] = evalN(kv, nk, i, j, p, W);
        double id = 1. / (-(mt[1] * mt[2]) + mt[0] * mt[3]);
        double inv[4] = {id * mt[3], id * (-mt[1]), id * (-mt[2]), id * mt[0]};
        d1d[b][0] = inv[0] * vec[0] + inv[1] * vec[1];
        if (nb > 1) d1d[b][1] = inv[2] * vec[0] + inv[3] * vec[1];
        for (int i = p; i < nb; i++) d1d[b][i] = d1d[b][i - 1] + (kv[i + p] - kv[i]) / ((double)p) * pr[b];
    }
    int n0 = c->nb[d0], n1 = c->nb[d1];
    out.assign((size_t)n0 * n1, 0.0);
    for (int i = 0; i < n0; i++)
        for (int j = 0; j < n1; j++) out[(size_t)n1 * i + j] = d1d[0][i] + d1d[1][j] + pr[2] * Z;
    if (order == 1) {
        int ibz1 = (f % 2 == 0) ? 1 : c->nb[dn] - 2;
        double mm = 0.0 / evalN(c->knot_h[dn].data(), nkn, 1, ibz1, p, Z);
        for (auto &v : out) v += mm;
    }
}

static int setup_bc(igap_gpu_ctx *c, const igap_bc_desc *bc)
{
    c->bc_h = *bc;
    const int p = 2;
    std::vector<int32_t> mask(c->ndof, 0);
    std::vector<double> g(c->ndof, 0.0), F(c->ndof, 0.0), vals;
    for (int i = 0; i < 36; i++)
        if (bc->bc_type[i] != 0 && bc->bc_type[i] != 1) return fail(IGAP_ERR_ARG, "bc_type[%d] = %d (0 or 1)", i, bc->bc_type[i]);
    // Neumann load, U independent (Residual_iga :273-405 with the constant table)
    double cq[4], wq[4];
    gauss4(cq, wq);
    for (int f = 0; f < 6; f++) {
        if (c->periodic[f / 2]) continue;
        int d0, d1, dn;
        face_dirs(f, d0, d1, dn);
        const double *k0 = c->knot_h[d0].data(), *k1 = c->knot_h[d1].data(), *kn = c->knot_h[dn].data();
        int nk0 = c->nelem[d0] + 5, nk1 = c->nelem[d1] + 5, nkn = c->nelem[dn] + 5;
        double zq = (f % 2 == 0) ? kn[0] : kn[nkn - 1];
        int z0 = (f % 2 == 0) ? 0 : c->nelem[dn] + p - 1;
        int z1 = (f % 2 == 0) ? 1 : c->nelem[dn] + p - 2;
        for (int order = 0; order < 2; order++)
            for (int idof = 0; idof < 3; idof++) {
                if (bc->bc_type[6 * f + 3 * order + idof] != 1) continue;
                double t = bc->neumann[6 * f + 3 * order + idof];
                if (t == 0.0) continue;   // -w N 0 adds only signed zeros
                for (int ex = 0; ex < c->nelem[d0]; ex++)
                    for (int ey = 0; ey < c->nelem[d1]; ey++) {
                        double X0 = k0[ex + p], X1 = k0[ex + p + 1], Y0 = k1[ey + p], Y1 = k1[ey + p + 1];
                        double val[18] = {0};
                        for (int qx = 0; qx < 4; qx++)
                            for (int qy = 0; qy < 4; qy++) {
                                double x = (-cq[qx] + 1.0) * X0 + cq[qx] * X1;
                                double y = (-cq[qy] + 1.0) * Y0 + cq[qy] * Y1;
                                double wgt = wq[qx] * wq[qy] * (X1 - X0) * (Y1 - Y0);
                                int ia = 0;
                                for (int bx = 0; bx < 3; bx++)
                                    for (int by = 0; by < 3; by++) {
                                        double n2 = evalN(k0, nk0, 0, ex + bx, p, x) * evalN(k1, nk1, 0, ey + by, p, y);
                                        if (order == 0) {
                                            val[ia++] -= wgt * n2 * t;
                                        } else {
                                            val[ia++] -= wgt * n2 * evalN(kn, nkn, 1, z0, p, zq) * t;
                                            val[ia++] -= wgt * n2 * evalN(kn, nkn, 1, z1, p, zq) * t;
                                        }
                                    }
                            }
                        int ia = 0;
                        for (int bx = 0; bx < 3; bx++)
                            for (int by = 0; by < 3; by++) {
                                F[dof_dirs(c, d0, d1, dn, ex + bx, ey + by, z0, idof)] += val[ia++];
                                if (order == 1) F[dof_dirs(c, d0, d1, dn, ex + bx, ey + by, z1, idof)] += val[ia++];
                            }
                    }
            }
    }
    // Dirichlet rows: the INSERT sequence of Residual_iga :416-494 (last face wins)
    for (int f = 0; f < 6; f++) {
        if (c->periodic[f / 2]) continue;
        int d0, d1, dn;
        face_dirs(f, d0, d1, dn);
        int n0 = c->nb[d0], n1 = c->nb[d1];
        for (int order = 0; order < 2; order++) {
            int an = (f % 2 == 0) ? order : c->nb[dn] - 1 - order;
            for (int idof = 0; idof < 3; idof++) {
                if (bc->bc_type[6 * f + 3 * order + idof] != 0) continue;
                dirichlet_face(c, bc->par_dirichlet, f, order, idof, vals);
                for (int i = 0; i < n0; i++)
                    for (int j = 0; j < n1; j++) {
                        int64_t r = dof_dirs(c, d0, d1, dn, i, j, an, idof);
                        mask[r] = 1;
                        g[r] = vals[(size_t)n1 * i + j];
                    }
            }
        }
    }
    HIPCHK(hipMemcpyAsync(c->d_mask, mask.data(), sizeof(int32_t) * c->ndof, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_g, g.data(), sizeof(double) * c->ndof, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_F, F.data(), sizeof(double) * c->ndof, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < 9; i++) c->M.par_periodic[i] = bc->par_periodic[i];
    return IGAP_OK;
}

// ---------------------------------------------------------------- create/destroy
static int stencil_len(int a, int n, int periodic)
{
    if (!periodic) {
        int lo = a - 2 < 0 ? 0 : a - 2, hi = a + 2 > n - 1 ? n - 1 : a + 2;
        return hi - lo + 1;
    }
    return n < 5 ? n : 5;
}

template <typename T>
static int dalloc(T **p, int64_t n)
{
    if (n <= 0) n = 1;
    hipError_t e = igap::dmalloc(p, sizeof(T) * n);
    if (e != hipSuccess) return fail(IGAP_ERR_NOMEM, "hipMalloc(%lld bytes): %s", (long long)(sizeof(T) * n), hipGetErrorString(e));
    return IGAP_OK;
}

#define TRY(x)                       \
    do {                             \
        int r_ = (x);                \
        if (r_ != IGAP_OK) {         \
            igap_gpu_destroy(c);     \
            return r_;               \
        }                            \
    } while (0)
// A separable patch: every control coordinate x_d depends on the control
// point's (active) index along d alone, and all weights are equal (bitwise).
// Then x_d = f_d(xi_d) on every element, G is diagonal, the rational basis is
// the B-spline basis and k_tangent_sep applies.
static int separable_patch(const igap_gpu_ctx *c)
{
    const int64_t n1 = c->na[1], n2 = c->na[2];
    for (int64_t id = 0; id < c->ncp_active; id++) {
        if (c->w_h[id] != c->w_h[0]) return 0;
        const int64_t a0 = id / (n1 * n2), a1 = (id / n2) % n1, a2 = id % n2;
        const int64_t base[3] = {a0 * n1 * n2, a1 * n2, a2};   // the point with the other two indices 0
        for (int d = 0; d < 3; d++)
            if (c->X_h[3 * id + d] != c->X_h[3 * base[d] + d]) return 0;
    }
    return 1;
}

// a HIP call inside igap_gpu_create: release the partly built context c on failure
#define HIPTRY(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) {                                                                      \
            igap_gpu_destroy(c);                                                                     \
            return fail(IGAP_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_));                          \
        }                                                                                            \
    } while (0)

extern "C" int igap_gpu_create(const igap_mesh_desc *mesh, const igap_phys_desc *phys, const igap_bc_desc *bc, int device, int flags,
                               igap_gpu_ctx **out)
{
    if (!mesh || !phys || !bc || !out) return fail(IGAP_ERR_ARG, "null argument");
    if (flags & ~(IGAP_TANGENT_ELEMENT_MATRICES | IGAP_TANGENT_DIRECT | IGAP_GENERAL_GEOMETRY))
        return fail(IGAP_ERR_ARG, "unknown flags 0x%x", flags);
    if ((flags & IGAP_TANGENT_ELEMENT_MATRICES) && (flags & IGAP_TANGENT_DIRECT))
        return fail(IGAP_ERR_ARG, "IGAP_TANGENT_ELEMENT_MATRICES and IGAP_TANGENT_DIRECT exclude each other");
    *out = nullptr;
    if (mesh->porder != 2) return fail(IGAP_ERR_ARG, "porder = %d: only quadratic patches (porder 2) are supported", mesh->porder);
    if (phys->kind != IGAP_PHYS_GRADELAST && phys->kind != IGAP_PHYS_GRADELASTTIME)
        return fail(IGAP_ERR_ARG, "phys kind %d unknown", phys->kind);
    for (int d = 0; d < 3; d++) {
        if (mesh->nelem[d] < 1) return fail(IGAP_ERR_ARG, "nelem[%d] = %d < 1", d, mesh->nelem[d]);
        if (mesh->periodic[d] && mesh->nelem[d] < 5)
            return fail(IGAP_ERR_ARG, "periodic direction %d needs nelem >= 5 (got %d)", d, mesh->nelem[d]);
        if (!mesh->knot[d]) return fail(IGAP_ERR_ARG, "knot[%d] is null", d);
    }
    if (!mesh->w || !mesh->X) return fail(IGAP_ERR_ARG, "w or X is null");
    if (device >= 0) HIPCHK(hipSetDevice(device));

    igap_gpu_ctx *c = new igap_gpu_ctx();
    HIPTRY(hipGetDevice(&c->device));
    c->torder = phys->kind;
    for (int d = 0; d < 3; d++) {
        c->nelem[d] = mesh->nelem[d];
        c->periodic[d] = mesh->periodic[d] ? 1 : 0;
        c->na[d] = mesh->nelem[d] + 2;
        c->nb[d] = c->periodic[d] ? mesh->nelem[d] : mesh->nelem[d] + 2;
        c->knot_h[d].assign(mesh->knot[d], mesh->knot[d] + mesh->nelem[d] + 5);
    }
    c->nel = (int64_t)c->nelem[0] * c->nelem[1] * c->nelem[2];
    if (c->nel > 0x7fffffff) { delete c; return fail(IGAP_ERR_ARG, "too many elements"); }
    {
        const bool periodic = c->periodic[0] || c->periodic[1] || c->periodic[2];
        if ((flags & IGAP_TANGENT_DIRECT) && periodic) {
            delete c;
            return fail(IGAP_ERR_ARG, "IGAP_TANGENT_DIRECT needs a non-periodic patch");
        }
        // 0 automatic (decided at the first tangent call), 1 element
        // matrices, 2 direct
        c->tangent_mode = (flags & IGAP_TANGENT_DIRECT) ? 2 : (flags & IGAP_TANGENT_ELEMENT_MATRICES) ? 1 : 0;
        c->tangent_periodic = periodic ? 1 : 0;
    }
    c->ncp_active = (int64_t)c->na[0] * c->na[1] * c->na[2];
    c->ncp = (int64_t)c->nb[0] * c->nb[1] * c->nb[2];
    c->ndof = 3 * c->ncp;
    hipError_t he = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (he != hipSuccess) { delete c; return fail(IGAP_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(he)); }
    // the no-allocation residual's device array on this device (once per device)
    HIPTRY(scratch_setup(c->device));
    for (int i = 0; i < 4; i++) {
        he = hipEventCreate(&c->ev[i]);
        if (he != hipSuccess) { igap_gpu_destroy(c); return fail(IGAP_ERR_HIP, "hipEventCreate: %s", hipGetErrorString(he)); }
    }

    // physics
    PhysK &P = c->P;
    std::memset(&P, 0, sizeof P);
    double coef[6];
    if (phys->use_coef) {
        std::memcpy(coef, phys->coef, sizeof coef);
    } else {
        double r = phys->par[0], cu = phys->par[1];
        double a = 1.3333333333333333 * r - 1.3333333333333333 * r * cu;
        double D = r * r - 2 * r * r * cu - r * a;
        coef[0] = (2 * cu) / D;
        coef[1] = -(a / (r * r * D));
        coef[2] = -(1 / (r * r * D));
        coef[3] = phys->par[2];
        coef[4] = phys->par[3];
        coef[5] = phys->par[4];
    }
    P.c2 = coef[0]; P.c3 = coef[1]; P.c4 = coef[2]; P.e0 = coef[3]; P.e345 = coef[4]; P.le = coef[5];
    P.rho = phys->par[5]; P.cii = phys->par[6]; P.dt = phys->par[7];
    P.torder = c->torder;
    P.nnode = c->torder == 0 ? 1 : 8;
    trunc_rule(P.s, P.cr, P.ct);
    if (c->torder != 0 && !(P.dt > 0.0)) { igap_gpu_destroy(c); return fail(IGAP_ERR_ARG, "transient physics needs dt = par[7] > 0"); }

    // mesh upload
    for (int d = 0; d < 3; d++) {
        TRY(dalloc(&c->d_knot[d], c->nelem[d] + 5));
        HIPTRY(hipMemcpy(c->d_knot[d], c->knot_h[d].data(), sizeof(double) * (c->nelem[d] + 5), hipMemcpyHostToDevice));
    }
    TRY(dalloc(&c->d_w, c->ncp_active));
    TRY(dalloc(&c->d_X, 3 * c->ncp_active));
    HIPTRY(hipMemcpy(c->d_w, mesh->w, sizeof(double) * c->ncp_active, hipMemcpyHostToDevice));
    HIPTRY(hipMemcpy(c->d_X, mesh->X, sizeof(double) * 3 * c->ncp_active, hipMemcpyHostToDevice));
    c->w_h.assign(mesh->w, mesh->w + c->ncp_active);
    c->X_h.assign(mesh->X, mesh->X + 3 * c->ncp_active);
    c->phys_h = *phys;
    MeshK &M = c->M;
    std::memset(&M, 0, sizeof M);
    for (int d = 0; d < 3; d++) {
        M.nelem[d] = c->nelem[d]; M.na[d] = c->na[d]; M.nb[d] = c->nb[d]; M.periodic[d] = c->periodic[d];
        M.knot[d] = c->d_knot[d];
    }
    M.w = c->d_w;
    M.X = c->d_X;
    gauss4(M.gq, M.gw);
    M.separable = (flags & IGAP_GENERAL_GEOMETRY) ? 0 : separable_patch(c);

    // CSR pattern: rows 3*cp + i, row length 3 * prod_d stencil_d
    std::vector<int64_t> rowptr(c->ndof + 1);
    rowptr[0] = 0;
    int64_t r = 0;
    for (int ax = 0; ax < c->nb[0]; ax++) {
        int lx = stencil_len(ax, c->nb[0], c->periodic[0]);
        for (int ay = 0; ay < c->nb[1]; ay++) {
            int ly = stencil_len(ay, c->nb[1], c->periodic[1]);
            for (int az = 0; az < c->nb[2]; az++) {
                int64_t len = 3LL * lx * ly * stencil_len(az, c->nb[2], c->periodic[2]);
                for (int i = 0; i < 3; i++, r++) rowptr[r + 1] = rowptr[r] + len;
            }
        }
    }
    c->nnz = rowptr[c->ndof];
    if (c->nnz > 0x7fffffffLL * 4) { igap_gpu_destroy(c); return fail(IGAP_ERR_ARG, "nnz %lld too large", (long long)c->nnz); }
    TRY(dalloc(&c->d_rowptr, c->ndof + 1));
    TRY(dalloc(&c->d_col, c->nnz));
    HIPTRY(hipMemcpy(c->d_rowptr, rowptr.data(), sizeof(int64_t) * (c->ndof + 1), hipMemcpyHostToDevice));
    HIPTRY(launch_csr_cols(M, c->d_rowptr, c->d_col, c->ncp, c->stream));

    // state and BC buffers
    int nh = c->torder + 1;
    TRY(dalloc(&c->d_Uh, nh * c->ndof));
    TRY(dalloc(&c->d_R, c->ndof));
    TRY(dalloc(&c->d_Re, 81 * c->nel));
    TRY(dalloc(&c->d_mask, c->ndof));
    TRY(dalloc(&c->d_g, c->ndof));
    TRY(dalloc(&c->d_F, c->ndof));
    HIPTRY(launch_fill_f64(c->d_Uh, 0.0, nh * c->ndof, c->stream));
    TRY(setup_bc(c, bc));
    HIPTRY(hipStreamSynchronize(c->stream));
    *out = c;
    return IGAP_OK;
}

extern "C" int igap_gpu_destroy(igap_gpu_ctx *c)
{
    if (!c) return IGAP_OK;
    // release everything even after a failure; report the first one
    int rc = IGAP_OK;
    auto note = [&rc](hipError_t e, const char *what) {
        if (e != hipSuccess && rc == IGAP_OK) rc = fail(IGAP_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    };
    note(hipSetDevice(c->device), "hipSetDevice");
    if (c->stream) note(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    for (int d = 0; d < 3; d++)
        if (c->d_knot[d]) note(hipFree(c->d_knot[d]), "hipFree");
    void *bufs[] = {c->d_w, c->d_X, c->d_Uh, c->d_R, c->d_Re, c->d_Ke[0], c->d_Ke[1], c->d_rowptr, c->d_col, c->d_pcol, c->d_val64,
                    c->d_val32, c->d_mask, c->d_g, c->d_F, c->d_Hd, c->d_H};
    for (void *b : bufs)
        if (b) note(hipFree(b), "hipFree");
    if (c->solver) solver_destroy(c->solver);
    for (int i = 0; i < 4; i++)
        if (c->ev[i]) note(hipEventDestroy(c->ev[i]), "hipEventDestroy");
    if (c->stream && c->own_stream) note(hipStreamDestroy(c->stream), "hipStreamDestroy");
    delete c;
    return rc;
}

extern "C" int igap_gpu_set_bc(igap_gpu_ctx *c, const igap_bc_desc *bc)
{
    if (!c || !bc) return fail(IGAP_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    return setup_bc(c, bc);
}

// ---------------------------------------------------------------- accessors
extern "C" int igap_gpu_sizes(const igap_gpu_ctx *c, int64_t *ndof, int64_t *nnz, int64_t *nelem, int *torder)
{
    if (!c) return fail(IGAP_ERR_ARG, "null context");
    if (ndof) *ndof = c->ndof;
    if (nnz) *nnz = c->nnz;
    if (nelem) *nelem = c->nel;
    if (torder) *torder = c->torder;
    return IGAP_OK;
}
extern "C" double *igap_gpu_uhist(igap_gpu_ctx *c) { return c ? c->d_Uh : nullptr; }

extern "C" int igap_gpu_intdensity(igap_gpu_ctx *c, double *HH, int *ndensity)
{
    if (!c || !HH || !ndensity) return fail(IGAP_ERR_ARG, "null argument");
    const int nd = c->torder == 0 ? 13 : 2;
    HIPCHK(hipSetDevice(c->device));
    if (!c->d_Hd) {
        // both buffers or neither: a failed call leaves the context as it was
        double *hd = nullptr, *h = nullptr;
        hipError_t e = igap::dmalloc(&hd, sizeof(double) * (size_t)c->nel * nd);
        if (e == hipSuccess) e = igap::dmalloc(&h, sizeof(double) * nd);
        if (e != hipSuccess) {
            if (hd) (void)hipFree(hd);
            return fail(IGAP_ERR_NOMEM, "density buffers: %s", hipGetErrorString(e));
        }
        c->d_Hd = hd;
        c->d_H = h;
    }
    HIPCHK(launch_density(c->M, c->P, c->d_Uh, c->ndof, c->d_Hd, c->d_H, (int)c->nel, c->stream));
    std::vector<double> H(nd), U((size_t)(c->torder + 1) * c->ndof), F(c->ndof);
    HIPCHK(hipMemcpyAsync(H.data(), c->d_H, sizeof(double) * nd, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(U.data(), c->d_Uh, sizeof(double) * U.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(F.data(), c->d_F, sizeof(double) * c->ndof, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    // Neumann faces (src/postcore_iga.c:302-371): -int u t with F = -int N t
    // the residual's load; u = current level (static) or the mean of the two
    // previous levels (gradelasttime_densityb, phys/gradelasttime.c:9234-9239)
    double s = 0.0;
    const int64_t n = c->ndof;
    for (int64_t i = 0; i < n; i++) s += (c->torder == 0 ? U[i] : 0.5 * (U[n + i] + U[2 * n + i])) * F[i];
    H[0] += s;
    for (int i = 0; i < nd; i++) HH[i] = H[i];
    *ndensity = nd;
    return IGAP_OK;
}

// levels [h0, h1) of U_hist that a solution file holds (src/interface.c:271-300)
static void file_levels(const igap_gpu_ctx *c, int &h0, int &h1)
{
    h0 = c->torder == 0 ? 0 : 1;
    h1 = c->torder + 1;
}

extern "C" int igap_gpu_save_solution(igap_gpu_ctx *c, const char *fname)
{
    if (!c || !fname) return fail(IGAP_ERR_ARG, "null argument");
    int h0, h1;
    file_levels(c, h0, h1);
    const size_t n = (size_t)(h1 - h0) * (size_t)c->ndof;
    std::vector<double> buf(n);
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(buf.data(), c->d_Uh + (size_t)h0 * c->ndof, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    FILE *f = fopen(fname, "wb");
    if (!f) return fail(IGAP_ERR_ARG, "cannot open %s for writing", fname);
    const size_t w = fwrite(buf.data(), sizeof(double), n, f);
    const int cl = fclose(f);
    if (w != n || cl != 0) return fail(IGAP_ERR_ARG, "short write to %s", fname);
    return IGAP_OK;
}

extern "C" int igap_gpu_load_solution(igap_gpu_ctx *c, const char *fname)
{
    if (!c || !fname) return fail(IGAP_ERR_ARG, "null argument");
    int h0, h1;
    file_levels(c, h0, h1);
    const size_t n = (size_t)(h1 - h0) * (size_t)c->ndof;
    std::vector<double> buf(n);
    FILE *f = fopen(fname, "rb");
    if (!f) return fail(IGAP_ERR_ARG, "cannot open %s", fname);
    const size_t r = fread(buf.data(), sizeof(double), n, f);
    fclose(f);
    if (r != n) return fail(IGAP_ERR_ARG, "%s holds %zu of the %zu values of %d level(s)", fname, r, n, h1 - h0);
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(c->d_Uh + (size_t)h0 * c->ndof, buf.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    if (h0 == 1)
        HIPCHK(hipMemcpyAsync(c->d_Uh, c->d_Uh + c->ndof, sizeof(double) * c->ndof, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return IGAP_OK;
}
extern "C" double *igap_gpu_rvec(igap_gpu_ctx *c) { return c ? c->d_R : nullptr; }
extern "C" const int64_t *igap_gpu_csr_rowptr(const igap_gpu_ctx *c) { return c ? c->d_rowptr : nullptr; }
extern "C" const int32_t *igap_gpu_csr_col(const igap_gpu_ctx *c) { return c ? c->d_col : nullptr; }
extern "C" const int32_t *igap_gpu_dirichlet_mask(const igap_gpu_ctx *c) { return c ? c->d_mask : nullptr; }
extern "C" const double *igap_gpu_dirichlet_values(const igap_gpu_ctx *c) { return c ? c->d_g : nullptr; }

extern "C" double *igap_gpu_csr_val64(igap_gpu_ctx *c)
{
    if (!c) return nullptr;
    if (!c->d_val64 && dalloc(&c->d_val64, c->nnz) != IGAP_OK) return nullptr;
    return c->d_val64;
}
extern "C" float *igap_gpu_csr_val32(igap_gpu_ctx *c)
{
    if (!c) return nullptr;
    if (!c->d_val32 && dalloc(&c->d_val32, c->nnz) != IGAP_OK) return nullptr;
    return c->d_val32;
}

// ---------------------------------------------------------------- timing helpers
static int t_begin(igap_gpu_ctx *c, int slot)
{
    if (c->timing) HIPCHK(hipEventRecord(c->ev[slot], c->stream));
    return IGAP_OK;
}
static int t_end(igap_gpu_ctx *c, int slot, int which)
{
    if (!c->timing) return IGAP_OK;
    HIPCHK(hipEventRecord(c->ev[slot + 1], c->stream));
    HIPCHK(hipEventSynchronize(c->ev[slot + 1]));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev[slot], c->ev[slot + 1]));
    c->t_ms[which] += ms;
    c->t_n[which] += 1;
    return IGAP_OK;
}

// ---------------------------------------------------------------- callbacks
extern "C" int igap_gpu_residual(igap_gpu_ctx *c, const double *U, double *R)
{
    if (!c) return fail(IGAP_ERR_ARG, "null context");
    HIPCHK(hipSetDevice(c->device));
    if (!R) R = c->d_R;
    if (c->timing) HIPCHK(hipEventRecord(c->ev[2], c->stream));
    if (U && U != c->d_Uh) HIPCHK(launch_copy_f64(c->d_Uh, U, c->ndof, c->stream));   // U_hist[0] = U (:68)
    int rc = t_begin(c, 0);
    if (rc) return rc;
    HIPCHK(launch_residual_elem(c->M, c->P, c->d_Uh, c->ndof, c->d_Re, (int)c->nel, c->stream));
    rc = t_end(c, 0, 0);
    if (rc) return rc;
    HIPCHK(launch_residual_gather(c->M, c->d_Re, c->d_Uh, c->d_F, c->d_mask, c->d_g, R, c->ncp, c->stream));
    if (c->timing) {
        HIPCHK(hipEventRecord(c->ev[3], c->stream));
        HIPCHK(hipEventSynchronize(c->ev[3]));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
        c->t_ms[3] += ms;
        c->t_n[3] += 1;
    }
    return IGAP_OK;
}

extern "C" int igap_gpu_residual_noalloc(igap_gpu_ctx *c, const double *U, double *R)
{
    if (!c) return fail(IGAP_ERR_ARG, "null context");
    const int64_t cap = scratch_capacity_doubles();
    if (81 * c->nel > cap)
        return fail(IGAP_ERR_ARG, "igap_gpu_residual_noalloc: %lld elements, its device array holds %lld", (long long)c->nel,
                    (long long)(cap / 81));
    HIPCHK(hipSetDevice(c->device));
    double *scr = scratch_device_ptr(c->device);
    if (!scr) return fail(IGAP_ERR_STATE, "no device array on device %d (contexts set it up)", c->device);
    if (!R) R = c->d_R;
    if (U && U != c->d_Uh) HIPCHK(launch_copy_f64(c->d_Uh, U, c->ndof, c->stream));   // U_hist[0] = U (:68)
    HIPCHK(launch_residual_elem(c->M, c->P, c->d_Uh, c->ndof, scr, (int)c->nel, c->stream));
    HIPCHK(launch_residual_gather(c->M, scr, c->d_Uh, c->d_F, c->d_mask, c->d_g, R, c->ncp, c->stream));
    return IGAP_OK;
}

extern "C" int igap_gpu_tangent(igap_gpu_ctx *c, const double *U, void *vals, int prec)
{
    if (!c) return fail(IGAP_ERR_ARG, "null context");
    if (prec != IGAP_F64 && prec != IGAP_F32) return fail(IGAP_ERR_ARG, "prec %d (IGAP_F64 or IGAP_F32)", prec);
    HIPCHK(hipSetDevice(c->device));
    if (!vals) {
        vals = (prec == IGAP_F64) ? (void *)igap_gpu_csr_val64(c) : (void *)igap_gpu_csr_val32(c);
        if (!vals) return IGAP_ERR_NOMEM;
    }
    // automatic choice: element matrices (measured faster: the element
    // kernel's stores stream out and a separate gather sums them) while their
    // buffer fits in free device memory with room to spare, direct scatter
    // into the CSR values otherwise (non-periodic patches)
    if (c->tangent_mode == 0 && c->tangent_choice[prec] == 0) {   // decided once per precision
        int choice = 1;
        if (!c->tangent_periodic) {
            size_t fr = 0, tot = 0;
            HIPCHK(hipMemGetInfo(&fr, &tot));
            const size_t need = (prec == IGAP_F64 ? sizeof(double) : sizeof(float)) * 6561 * (size_t)c->nel;
            if (need + tot / 10 > fr) choice = 2;
        }
        c->tangent_choice[prec] = choice;
    }
    const bool direct = c->tangent_mode == 2 || (c->tangent_mode == 0 && c->tangent_choice[prec] == 2);
    if (direct) {
        // element matrices added straight into the CSR values, colour by colour
        if (c->timing) HIPCHK(hipEventRecord(c->ev[2], c->stream));
        if (U && U != c->d_Uh) HIPCHK(launch_copy_f64(c->d_Uh, U, c->ndof, c->stream));   // U_hist[0] = U (:557)
        int rc = t_begin(c, 0);
        if (rc) return rc;
        HIPCHK(launch_tangent_direct(c->M, c->P, c->d_Uh, c->ndof, c->d_rowptr, c->d_mask, vals, prec, c->nnz, c->ncp, c->stream));
        rc = t_end(c, 0, 1);
        if (rc) return rc;
        if (c->timing) {
            HIPCHK(hipEventRecord(c->ev[3], c->stream));
            HIPCHK(hipEventSynchronize(c->ev[3]));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
            c->t_ms[4] += ms;
            c->t_n[4] += 1;
        }
        return IGAP_OK;
    }
    if (!c->d_Ke[prec]) {
        size_t es = prec == IGAP_F64 ? sizeof(double) : sizeof(float);
        hipError_t e = igap::dmalloc(&c->d_Ke[prec], es * 6561 * (size_t)c->nel);
        if (e != hipSuccess) return fail(IGAP_ERR_NOMEM, "element tangent buffer (%zu bytes): %s", es * 6561 * (size_t)c->nel, hipGetErrorString(e));
    }
    if (c->timing) HIPCHK(hipEventRecord(c->ev[2], c->stream));
    if (U && U != c->d_Uh) HIPCHK(launch_copy_f64(c->d_Uh, U, c->ndof, c->stream));   // U_hist[0] = U (:557)
    int rc = t_begin(c, 0);
    if (rc) return rc;
    HIPCHK(launch_tangent_elem(c->M, c->P, c->d_Uh, c->ndof, c->d_Ke[prec], prec, (int)c->nel, c->stream));
    rc = t_end(c, 0, 1);
    if (rc) return rc;
    rc = t_begin(c, 0);
    if (rc) return rc;
    HIPCHK(launch_tangent_gather(c->M, c->d_Ke[prec], prec, c->d_rowptr, c->d_mask, vals, c->ncp, c->stream));
    rc = t_end(c, 0, 2);
    if (rc) return rc;
    if (c->timing) {
        HIPCHK(hipEventRecord(c->ev[3], c->stream));
        HIPCHK(hipEventSynchronize(c->ev[3]));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
        c->t_ms[4] += ms;
        c->t_n[4] += 1;
    }
    return IGAP_OK;
}

// ---------------------------------------------------------------- pointwise physics
extern "C" int igap_gpu_phys_eval(igap_gpu_ctx *c, int64_t npts, const double *u, double *r, double *C, double *hh)
{
    if (!c) return fail(IGAP_ERR_ARG, "null context");
    if (npts < 0) return fail(IGAP_ERR_ARG, "npts = %lld < 0", (long long)npts);
    if (npts > 0 && !u) return fail(IGAP_ERR_ARG, "null jets");
    if (npts > ((int64_t)1 << 31) / 30) return fail(IGAP_ERR_ARG, "npts = %lld too large for one launch", (long long)npts);
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(launch_phys_eval(c->P, npts, u, r, C, hh, c->stream));
    return IGAP_OK;
}

// ---------------------------------------------------------------- helpers
extern "C" int igap_gpu_synchronize(igap_gpu_ctx *c)
{
    if (!c) return fail(IGAP_ERR_ARG, "null context");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    return IGAP_OK;
}
extern "C" int igap_gpu_malloc(igap_gpu_ctx *c, void **ptr, int64_t bytes)
{
    if (!c || !ptr) return fail(IGAP_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    hipError_t e = igap::dmalloc(ptr, bytes > 0 ? bytes : 1);
    if (e != hipSuccess) return fail(IGAP_ERR_NOMEM, "hipMalloc(%lld): %s", (long long)bytes, hipGetErrorString(e));
    return IGAP_OK;
}
extern "C" int igap_gpu_free(igap_gpu_ctx *c, void *ptr)
{
    if (!c) return fail(IGAP_ERR_ARG, "null context");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipFree(ptr));
    return IGAP_OK;
}
extern "C" int igap_gpu_memcpy_htod(igap_gpu_ctx *c, void *dst, const void *src, int64_t bytes)
{
    if (!c) return fail(IGAP_ERR_ARG, "null context");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return IGAP_OK;
}
extern "C" int igap_gpu_memcpy_dtoh(igap_gpu_ctx *c, void *dst, const void *src, int64_t bytes)
{
    if (!c) return fail(IGAP_ERR_ARG, "null context");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return IGAP_OK;
}
extern "C" int igap_gpu_memcpy_dtod(igap_gpu_ctx *c, void *dst, const void *src, int64_t bytes)
{
    if (!c) return fail(IGAP_ERR_ARG, "null context");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return IGAP_OK;
}

extern "C" int igap_gpu_timing_enable(igap_gpu_ctx *c, int on)
{
    if (!c) return fail(IGAP_ERR_ARG, "null context");
    c->timing = on ? 1 : 0;
    return IGAP_OK;
}
extern "C" int igap_gpu_timing_get(igap_gpu_ctx *c, int which, double *ms_total, int64_t *count)
{
    if (!c || which < 0 || which > 4) return fail(IGAP_ERR_ARG, "bad timing query");
    if (ms_total) *ms_total = c->t_ms[which];
    if (count) *count = c->t_n[which];
    return IGAP_OK;
}
extern "C" int igap_gpu_timing_reset(igap_gpu_ctx *c)
{
    if (!c) return fail(IGAP_ERR_ARG, "null context");
    for (int i = 0; i < 5; i++) { c->t_ms[i] = 0; c->t_n[i] = 0; }
    return IGAP_OK;
}

extern "C" int igap_gpu_selftest_mfma(double *out64, float *out32)
{
    if (!out64 || !out32) return fail(IGAP_ERR_ARG, "null argument");
    double *d64 = nullptr;
    float *d32 = nullptr;
    // one status for the whole sequence so that both buffers are freed on
    // every path
    hipError_t e = hipMalloc(&d64, 256 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&d32, 256 * sizeof(float));
    if (e == hipSuccess) e = launch_selftest_mfma(d64, d32, 0);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(out64, d64, 256 * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(out32, d32, 256 * sizeof(float), hipMemcpyDeviceToHost);
    const hipError_t f64 = d64 ? hipFree(d64) : hipSuccess, f32 = d32 ? hipFree(d32) : hipSuccess;
    if (e == hipSuccess) e = (f64 != hipSuccess) ? f64 : f32;
    if (e != hipSuccess) return fail(IGAP_ERR_HIP, "MFMA self test: %s", hipGetErrorString(e));
    return IGAP_OK;
}
