// igap_launch.h -- host launchers of the assembly kernels (igap_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "igap_device.h"

namespace igap {

hipError_t launch_density(const MeshK &M, const PhysK &P, const double *Uh, int64_t ndof, double *Hd, double *H, int nel,
                          hipStream_t s);
hipError_t launch_residual_elem(const MeshK &M, const PhysK &P, const double *Uh, int64_t ndof, double *Re, int nel, hipStream_t s);
hipError_t launch_residual_gather(const MeshK &M, const double *Re, const double *Uh, const double *Fneu, const int32_t *mask,
                                  const double *g, double *R, int64_t ncp, hipStream_t s);
hipError_t launch_tangent_elem(const MeshK &M, const PhysK &P, const double *Uh, int64_t ndof, void *Ke, int prec, int nel,
                               hipStream_t s);
hipError_t launch_tangent_direct(const MeshK &M, const PhysK &P, const double *Uh, int64_t ndof, const int64_t *rowptr,
                                 const int32_t *mask, void *vals, int prec, int64_t nnz, int64_t ncp, hipStream_t s);
hipError_t launch_tangent_gather(const MeshK &M, const void *Ke, int prec, const int64_t *rowptr, const int32_t *mask, void *vals,
                                 int64_t ncp, hipStream_t s);
hipError_t launch_csr_cols(const MeshK &M, const int64_t *rowptr, int32_t *col, int64_t ncp, hipStream_t s);
hipError_t launch_fill_f64(double *p, double v, int64_t n, hipStream_t s);
hipError_t launch_copy_f64(double *dst, const double *src, int64_t n, hipStream_t s);
hipError_t launch_selftest_mfma(double *o64, float *o32, hipStream_t s);
hipError_t launch_phys_eval(const PhysK &P, int64_t npts, const double *u, double *r, double *C, double *hh, hipStream_t s);
// the no-allocation residual's device array: once per device (current),
// its address, and its size in doubles
hipError_t scratch_setup(int device);
double *scratch_device_ptr(int device);
int64_t scratch_capacity_doubles();

}  // namespace igap
