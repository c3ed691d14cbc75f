"""Failure paths of the device solver.

* A device allocation that fails part way through the first solve (the
  solver, its coloured ordering, the multigrid hierarchy with its coarse
  contexts) or the first density integral leaves the context usable: the
  next call rebuilds what the failed one did not finish and reaches the
  answer of an undisturbed run.  The failure is forced from the host
  (igap_gpu_debug_fail_alloc), so every kernel still runs normally.
* The Newton solve reports PETSc's reasons as IGAPSolvSolve sees them
  (src/interface.c:302-328): -3 when the linear solve does not converge,
  -6 when the line search fails or a step changes |F| by less than 1e-6
  relative (the stagnation stop of src/interface.c:318)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _beam(igap, tip=0.01):
    mesh = igap.MeshIGABeam(L=(4, 1, 1), nelem=(16, 4, 4))
    return igap.Assm(mesh, igap.PhysGradElast(coef=igap.SVK_COEF), igap.BCClampedTip(tip=tip), device=0)


def _solve_rc(igap, A, opts):
    from igap4_amd import _lib
    st = _lib.SolverStats()
    rc = _lib.lib().igap_gpu_snes_solve(A.ctx, C.byref(opts), C.byref(st))
    return rc, st.as_dict()


def test_solver_recovers_from_a_failed_allocation(igap):
    from igap4_amd import _lib
    L = _lib.lib()
    opts = igap.assm.default_solver_opts(snes_rtol=1e-10, snes_atol=1e-30, pc="mg", pc_order=1, pc_fp32=1)
    with _beam(igap) as A:
        A.set_uhist(np.zeros(A.ndof))
        st = A.snes_solve(opts)
        U_ref = A.get_uhist()
    assert st["converged"] > 0
    failed = 0
    for k in range(400):
        with _beam(igap) as A:
            A.set_uhist(np.zeros(A.ndof))
            L.igap_gpu_debug_fail_alloc(k)
            rc, _ = _solve_rc(igap, A, opts)
            L.igap_gpu_debug_fail_alloc(-1)
            if rc == 0:
                break              # the k-th allocation was never reached: every earlier one was tried
            failed += 1
            assert rc in (_lib.IGAP_ERR_NOMEM, _lib.IGAP_ERR_HIP), rc
            A.set_uhist(np.zeros(A.ndof))
            st = A.snes_solve(opts)
            assert st["converged"] > 0, (k, st)
            assert np.max(np.abs(A.get_uhist() - U_ref)) <= 1e-9 * np.max(np.abs(U_ref)), k
    print(f"{failed} allocation points of the first multigrid solve failed and recovered")
    assert failed >= 50 and rc == 0


def test_density_recovers_from_a_failed_allocation(igap, oracle):
    from igap4_amd import _lib
    L = _lib.lib()
    with _beam(igap) as A:
        A.set_uhist(oracle.glibc_rand_state(A.ndof))
        H_ref = A.intdensity()
    for k in (0, 1):
        with _beam(igap) as A:
            A.set_uhist(oracle.glibc_rand_state(A.ndof))
            L.igap_gpu_debug_fail_alloc(k)
            HH = np.zeros(13)
            n = C.c_int()
            rc = L.igap_gpu_intdensity(A.ctx, HH.ctypes.data_as(_lib.c_double_p), C.byref(n))
            L.igap_gpu_debug_fail_alloc(-1)
            assert rc == _lib.IGAP_ERR_NOMEM
            assert np.array_equal(A.intdensity(), H_ref)


def test_stagnation_is_reported_as_line_search_divergence(igap):
    """An unreachable tolerance: Newton converges to rounding level and then
    stops on the stagnation test (or a failed line search), not at maxit."""
    opts = igap.assm.default_solver_opts(snes_rtol=0.0, snes_atol=0.0, snes_maxit=50, pc="ilu0", pc_order=1,
                                         ksp_rtol=1e-12)
    with _beam(igap, tip=0.02) as A:
        A.set_uhist(np.zeros(A.ndof))
        st = A.snes_solve(opts)
    print(f"reason {st['converged']}, {st['newton_its']} Newton iterations, |F| {st['fnorm0']:.2e} -> {st['fnorm']:.2e}")
    assert st["converged"] == -6
    assert st["newton_its"] < 50
    assert st["fnorm"] <= 1e-10 * st["fnorm0"]


def test_linear_solve_failure_is_reported(igap):
    opts = igap.assm.default_solver_opts(snes_rtol=1e-10, snes_atol=1e-30, pc="jacobi", ksp_rtol=1e-12, ksp_maxit=1)
    with _beam(igap, tip=0.02) as A:
        # a start whose first linear solve is not trivial (from U = 0 only the
        # Dirichlet rows carry a residual, which GMRES solves in one step)
        A.set_uhist(igap.smooth_state(A.mesh, amp=0.01))
        U0 = A.get_uhist()
        st = A.snes_solve(opts)
        U1 = A.get_uhist()
    assert st["converged"] == -3
    assert st["newton_its"] == 0 and st["ksp_its"] == 1
    assert np.array_equal(U0, U1)   # the iterate is kept
