"""The two ways the tangent reaches the CSR values (igap_gpu_create flags):
element matrices in HBM summed by a gather in ascending element order, and
the element kernel adding its matrix straight into the CSR values one
colour class at a time.  Both are checked against the oracle, against each
other (the same sums in another order), for bitwise repeatability, with
Dirichlet rows and columns (MatZeroRowsColumns, src/assmcore_iga.c:844), in
both precisions and for both physics; direct scatter refuses periodic
patches."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    ("ref-distorted", (4, 3, 5), True, 0),
    ("svk-thin", (7, 2, 6), False, 0),
    ("single", (1, 1, 1), False, 0),
    ("transient", (3, 4, 3), True, 2),
]


def _make(igap, oracle, nel, pert, torder):
    mesh = igap.MeshIGABeam(L=(4, 1, 1), nelem=nel, perturbed=pert)
    mo = oracle.Mesh(nel, L=(4, 1, 1), perturbed=pert)
    if torder == 0:
        ph, po = igap.PhysGradElast(), oracle.make_phys(0)
    else:
        ph, po = igap.PhysGradElastTime(), oracle.make_phys(2)
    bc, bo = igap.BCClampedTip(tip=0.02), oracle.clamped_tip_bc(tip=0.02)
    U = oracle.glibc_rand_state((torder + 1) * mo.ndof)
    return mesh, mo, ph, po, bc, bo, U


@pytest.mark.parametrize("case", [c[0] for c in CASES])
def test_direct_scatter_matches_element_matrices_and_oracle(igap, oracle, case):
    from igap4_amd import _lib
    _, nel, pert, torder = [c for c in CASES if c[0] == case][0]
    mesh, mo, ph, po, bc, bo, U = _make(igap, oracle, nel, pert, torder)
    out = {}
    for name, fl in (("ke", _lib.IGAP_TANGENT_ELEMENT_MATRICES), ("direct", _lib.IGAP_TANGENT_DIRECT)):
        with igap.Assm(mesh, ph, bc, device=0, flags=fl) as A:
            rp, col, K = A.tangent(U, prec="f64")
            _, _, K2 = A.tangent(U, prec="f64")
            _, _, K32 = A.tangent(U, prec="f32")
        assert np.array_equal(K, K2)   # bitwise repeatable
        out[name] = (rp, col, K, K32)
    rp_o, col_o, K_o = oracle.tangent(mo, po, bo, U)
    rp, col, Kk, K32k = out["ke"]
    _, _, Kd, K32d = out["direct"]
    assert np.array_equal(col, col_o) and np.array_equal(out["direct"][1], col_o)
    scale = np.max(np.abs(K_o))
    e_d, e_k = np.max(np.abs(Kd - K_o)) / scale, np.max(np.abs(Kk - K_o)) / scale
    e_dk = np.max(np.abs(Kd - Kk)) / scale
    e32 = np.max(np.abs(K32d.astype(np.float64) - Kk)) / scale
    print(f"[{case}] direct vs oracle {e_d:.1e}, element matrices vs oracle {e_k:.1e}, direct vs element matrices "
          f"{e_dk:.1e}, FP32 direct vs FP64 {e32:.1e}")
    assert e_d <= 1e-12 and e_k <= 1e-12 and e_dk <= 1e-13
    assert e32 <= 2e-5
    # Dirichlet rows and columns: zero with a unit diagonal in both paths
    mask, _ = oracle.dirichlet(mo, bo)
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    dm = (mask[rows] != 0) | (mask[col] != 0)
    assert np.array_equal(Kd[dm], np.where(rows[dm] == col[dm], 1.0, 0.0))
    assert np.array_equal(K32d[dm], np.where(rows[dm] == col[dm], 1.0, 0.0).astype(np.float32))


def test_direct_scatter_refuses_periodic_patches(igap):
    from igap4_amd import _lib
    mesh = igap.MeshIGABeam(L=(1, 1, 1), nelem=(5, 5, 5), periodic=(1, 0, 0))
    with pytest.raises(_lib.IgapError):
        igap.Assm(mesh, igap.PhysGradElast(), igap.BCDefault(), device=0, flags=_lib.IGAP_TANGENT_DIRECT)


def test_direct_scatter_newton_matches_element_matrices(igap):
    """The same Newton solve through either scatter reaches the same answer."""
    from igap4_amd import _lib
    mesh = igap.MeshIGABeam(L=(4, 1, 1), nelem=(16, 4, 4))
    sols = []
    for fl in (_lib.IGAP_TANGENT_ELEMENT_MATRICES, _lib.IGAP_TANGENT_DIRECT):
        with igap.Assm(mesh, igap.PhysGradElast(coef=igap.SVK_COEF), igap.BCClampedTip(tip=0.03), device=0, flags=fl) as A:
            A.set_uhist(np.zeros(A.ndof))
            st = A.snes_solve(igap.assm.default_solver_opts(snes_rtol=1e-12, snes_atol=1e-30, pc="mg", pc_order=1))
            assert st["converged"] > 0
            sols.append(A.get_uhist())
    err = np.max(np.abs(sols[0] - sols[1])) / np.max(np.abs(sols[0]))
    print(f"Newton solutions differ by {err:.1e}")
    assert err <= 1e-10
