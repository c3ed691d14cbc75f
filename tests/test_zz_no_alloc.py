"""igap_gpu_residual_noalloc (the residual for callers that cannot allocate
device memory) against igap_gpu_residual, through the C ABI: the same
kernels in the same order, so the results are bitwise equal, on two calls in
a row (the second on a smaller patch, so that the shared device array keeps
the first call's larger contents), and a patch beyond the array's capacity
is refused without launching anything."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _residuals(igap, mesh, phys, bc, U):
    from igap4_amd import _lib
    L = _lib.lib()
    with igap.Assm(mesh, phys, bc, device=0) as A:
        n = A.ndof
        bufs = []
        for _ in range(3):
            p = C.c_void_p()
            _lib.check(L.igap_gpu_malloc(A.ctx, C.byref(p), 8 * n))
            bufs.append(p)
        u_dev, r_ref, r_new = bufs
        try:
            A.htod(u_dev, U)
            _lib.check(L.igap_gpu_residual(A.ctx, u_dev, r_ref))
            _lib.check(L.igap_gpu_residual_noalloc(A.ctx, u_dev, r_new))
            A.synchronize()
            ref = A.dtoh(r_ref, n)
            new = A.dtoh(r_new, n)
        finally:
            for p in bufs:
                L.igap_gpu_free(A.ctx, p)
    return ref, new


def test_noalloc_residual_equals_residual_on_two_calls(igap, oracle):
    # first call: a 24x8x8 clamped beam (REF physics, distorted mesh);
    # second call: a smaller 8x4x4 transient patch with three history levels
    cases = [
        (igap.MeshIGABeam(L=(4, 1, 1), nelem=(24, 8, 8), perturbed=True), igap.PhysGradElast(),
         igap.BCClampedTip(tip=0.03)),
        (igap.MeshIGABeam(L=(2, 1, 1), nelem=(8, 4, 4)), igap.PhysGradElastTime(), igap.BCClampedTip(tip=0.01)),
    ]
    for mesh, phys, bc in cases:
        U = oracle.glibc_rand_state(mesh.ndof)
        ref, new = _residuals(igap, mesh, phys, bc, U)
        print(f"{mesh.nelem}: max |R_noalloc - R| = {np.max(np.abs(new - ref)):.1e}, max |R| = {np.max(np.abs(ref)):.2e}")
        assert np.max(np.abs(ref)) > 0
        assert np.array_equal(new, ref)


def test_noalloc_residual_refuses_patches_beyond_its_array(igap):
    from igap4_amd import _lib
    mesh = igap.MeshIGABeam(L=(4, 1, 1), nelem=(128, 64, 64))   # 524,288 elements > 414,252
    with igap.Assm(mesh, igap.PhysGradElast(), igap.BCDefault(), device=0) as A:
        rc = _lib.lib().igap_gpu_residual_noalloc(A.ctx, None, None)
        msg = _lib.lib().igap_gpu_last_error().decode()
    print(msg)
    assert rc == _lib.IGAP_ERR_ARG and "414252" in msg
