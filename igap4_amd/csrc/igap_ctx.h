// igap_ctx.h -- the context behind the opaque igap_gpu_ctx handle (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/igap_gpu.h"
#include "igap_device.h"

namespace igap {
struct Solver;
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void solver_destroy(Solver *s);
// hipMalloc of every device buffer the library owns, through the test hook of
// igap_gpu_debug_fail_alloc (an armed countdown makes one allocation fail)
hipError_t dmalloc_raw(void **p, size_t bytes);
template <typename T>
inline hipError_t dmalloc(T **p, size_t bytes)
{
    return dmalloc_raw(reinterpret_cast<void **>(p), bytes);
}
}  // namespace igap

#define HIPCHK(x)                                                                                        \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess) return igap::fail(IGAP_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_));      \
    } while (0)

struct igap_gpu_ctx {
    int device;
    hipStream_t stream;
    int own_stream = 1;             // 0: stream borrowed from another context (multigrid levels)
    int nelem[3], na[3], nb[3], periodic[3];
    int64_t ncp_active, ncp, ndof, nnz, nel;
    int torder;
    std::vector<double> knot_h[3];
    std::vector<double> w_h, X_h;   // host copies of the patch (active numbering)
    igap_phys_desc phys_h;          // descriptors the context was built from
    igap_bc_desc bc_h;
    igap::PhysK P;
    igap::MeshK M;
    // device
    double *d_knot[3] = {nullptr, nullptr, nullptr};
    double *d_w = nullptr, *d_X = nullptr;
    double *d_Uh = nullptr, *d_R = nullptr, *d_Re = nullptr;
    void *d_Ke[2] = {nullptr, nullptr};
    int64_t *d_rowptr = nullptr;
    int32_t *d_col = nullptr;
    int32_t *d_pcol = nullptr;      // neighbour control point of each 3x3 block (nnz/9), built on first use
    double *d_val64 = nullptr;
    float *d_val32 = nullptr;
    int32_t *d_mask = nullptr;
    double *d_g = nullptr, *d_F = nullptr;
    double *d_Hd = nullptr, *d_H = nullptr;   // density integral: per element, total
    // timing
    int timing = 0;
    hipEvent_t ev[4];
    double t_ms[5] = {0, 0, 0, 0, 0};
    int64_t t_n[5] = {0, 0, 0, 0, 0};
    int tangent_mode = 0;           // 0 automatic, 1 element matrices + gather, 2 direct into the CSR
    int tangent_periodic = 0;       // a periodic direction (direct scatter unavailable)
    int tangent_choice[2] = {0, 0}; // automatic mode, per precision: 0 undecided, 1 element matrices, 2 direct
    // Newton-Krylov solver state (created on first use)
    igap::Solver *solver = nullptr;
};
