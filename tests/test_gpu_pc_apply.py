"""igap_gpu_pc_apply (PCApply on the preconditioner of the last solve): the
point-Jacobi preconditioner is the tangent's inverse diagonal; the
multigrid's finest-level smoother and its V-cycle are linear and contract
the error in the tangent's energy norm."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solved(igap, pc):
    mesh = igap.MeshIGABeam(L=(4, 1, 1), nelem=(16, 4, 4))
    A = igap.Assm(mesh, igap.PhysGradElast(coef=igap.SVK_COEF), igap.BCClampedTip(tip=0.02), device=0)
    A.set_uhist(np.zeros(A.ndof))
    st = A.snes_solve(igap.assm.default_solver_opts(snes_rtol=1e-10, snes_atol=1e-30, pc=pc, pc_order=1, pc_lag=1))
    assert st["converged"] > 0
    return mesh, A


def test_pc_apply_needs_a_solve_first(igap):
    from igap4_amd import _lib
    mesh = igap.MeshIGABeam(L=(4, 1, 1), nelem=(8, 2, 2))
    with igap.Assm(mesh, igap.PhysGradElast(), igap.BCDefault(), device=0) as A:
        with pytest.raises(_lib.IgapError):
            A.pc_apply(np.ones(A.ndof))


def _refresh_at(igap, A, U, pc, pc_fp32=0):
    """One Newton iteration from U: the preconditioner is refreshed on the
    tangent at U itself."""
    A.set_uhist(U)
    A.snes_solve(igap.assm.default_solver_opts(snes_rtol=0.0, snes_atol=0.0, snes_maxit=1, pc=pc, pc_order=1, pc_lag=1,
                                               pc_fp32=pc_fp32))


def test_pc_apply_jacobi_is_the_inverse_diagonal(igap):
    mesh, A = _solved(igap, "jacobi")
    with A:
        U = A.get_uhist()
        _refresh_at(igap, A, U, "jacobi")
        rp, col, K = A.tangent(U, "f64")   # the tangent the preconditioner was refreshed on
        b = np.random.default_rng(0).standard_normal(A.ndof)
        x = A.pc_apply(b)
    d = np.array([K[rp[r]:rp[r + 1]][col[rp[r]:rp[r + 1]] == r][0] for r in range(A.ndof)])
    err = np.max(np.abs(x - b / d)) / np.max(np.abs(b / d))
    print(f"Jacobi: max rel |x - b/diag| = {err:.1e}")
    assert err <= 1e-14


def test_pc_apply_smoother_and_cycle_contract_in_the_energy_norm(igap):
    """The finest-level smoother S and the whole V-cycle M^-1 are linear and
    contract the error e -> e - S K e in the energy norm of the (here
    symmetric positive definite) tangent, the smoother less than the cycle."""
    import scipy.sparse as sp
    mesh, A = _solved(igap, "mg")
    with A:
        U = A.get_uhist()
        _refresh_at(igap, A, U, "mg")
        rp, col, K = A.tangent(U, "f64")
        Km = sp.csr_matrix((K, col, rp), shape=(A.ndof, A.ndof))
        mask = A.dirichlet()[0] == 0
        rng = np.random.default_rng(1)
        e = rng.standard_normal(A.ndof) * mask
        b1, b2 = rng.standard_normal(A.ndof), rng.standard_normal(A.ndof)
        lin = {}
        ratio = {}
        for which in ("smoother", "full"):
            S1, S2, S12 = (A.pc_apply(v, which=which) for v in (b1, b2, 2.5 * b1 + b2))
            lin[which] = np.max(np.abs(S12 - (2.5 * S1 + S2))) / np.max(np.abs(S12))
            r = e - A.pc_apply(Km @ e, which=which)
            ratio[which] = np.sqrt((r @ (Km @ r)) / (e @ (Km @ e)))
    print(f"energy-norm error kept: smoother {ratio['smoother']:.3f}, V-cycle {ratio['full']:.3f}; "
          f"linearity {lin['smoother']:.1e}, {lin['full']:.1e}")
    assert lin["smoother"] < 1e-12 and lin["full"] < 1e-12
    assert ratio["full"] < ratio["smoother"] < 1.0
    assert ratio["full"] < 0.5


def test_pc_apply_16bit_smoothing_matrices(igap):
    """pc_fp32 = 2: the multigrid levels smooth with binary16 copies of their
    FP32 matrices (a power-of-two scale per row).  The smoother and the
    V-cycle stay linear, differ from the FP32 ones by the 16-bit rounding of
    the matrix entries (2^-11 relative per entry; measured 6e-4 for the
    smoother and 3.5e-3 for the cycle, whose coarse corrections carry it, and
    measurably, so the 16-bit path is the one that ran) and still contract
    the error in the energy norm as much; pc_fp32 outside 0..2 is rejected."""
    import scipy.sparse as sp
    from igap4_amd import _lib
    mesh, A = _solved(igap, "mg")
    with A:
        U = A.get_uhist()
        rp, col, K = A.tangent(U, "f64")
        Km = sp.csr_matrix((K, col, rp), shape=(A.ndof, A.ndof))
        mask = A.dirichlet()[0] == 0
        rng = np.random.default_rng(2)
        e = rng.standard_normal(A.ndof) * mask
        b1, b2 = rng.standard_normal(A.ndof), rng.standard_normal(A.ndof)
        out = {}
        for f in (1, 2):
            _refresh_at(igap, A, U, "mg", pc_fp32=f)
            out[f] = {w: [A.pc_apply(v, which=w) for v in (b1, b2, 2.5 * b1 + b2, Km @ e)] for w in ("smoother", "full")}
        with pytest.raises(_lib.IgapError):
            A.snes_solve(igap.assm.default_solver_opts(snes_maxit=1, pc="mg", pc_order=1, pc_fp32=3))
    for w in ("smoother", "full"):
        S1, S2, S12, SKe = out[2][w]
        lin = np.max(np.abs(S12 - (2.5 * S1 + S2))) / np.max(np.abs(S12))
        diff = max(np.max(np.abs(out[2][w][i] - out[1][w][i])) / np.max(np.abs(out[1][w][i])) for i in range(3))
        r = e - SKe
        ratio = np.sqrt((r @ (Km @ r)) / (e @ (Km @ e)))
        r1 = e - out[1][w][3]
        ratio1 = np.sqrt((r1 @ (Km @ r1)) / (e @ (Km @ e)))
        print(f"{w}: 16-bit vs FP32 max rel diff {diff:.2e}, linearity {lin:.1e}, "
              f"energy-norm error kept {ratio:.4f} (FP32 {ratio1:.4f})")
        assert lin < 1e-12
        assert 1e-7 < diff < (2e-3 if w == "smoother" else 1e-2)
        assert ratio < 1.0 and abs(ratio - ratio1) < 0.01
