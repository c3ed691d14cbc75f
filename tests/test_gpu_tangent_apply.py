"""igap_gpu_tangent_apply (MatMult on the assembled tangent): y = K x equals
the product of the CSR the context hands out (igap_gpu_tangent's values on
its rowptr/col pattern), in both precisions, and is refused before any
tangent of that precision exists."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_tangent_apply_needs_a_tangent_first(igap):
    from igap4_amd import _lib
    mesh = igap.MeshIGABeam(L=(4, 1, 1), nelem=(8, 2, 2))
    with igap.Assm(mesh, igap.PhysGradElast(), igap.BCDefault(), device=0) as A:
        with pytest.raises(_lib.IgapError):
            A.tangent_apply(np.ones(A.ndof), "f32")


@pytest.mark.parametrize("prec,tol", [("f64", 1e-14), ("f32", 1e-14)])
def test_tangent_apply_is_the_csr_product(igap, prec, tol):
    import scipy.sparse as sp
    mesh = igap.MeshIGABeam(L=(4, 1, 1), nelem=(16, 4, 4), perturbed=True)
    rng = np.random.default_rng(3)
    with igap.Assm(mesh, igap.PhysGradElast(coef=igap.SVK_COEF), igap.BCClampedTip(tip=0.02), device=0) as A:
        U = 0.01 * rng.standard_normal(A.ndof)
        rp, col, K = A.tangent(U, prec)
        x = rng.standard_normal(A.ndof)
        y = A.tangent_apply(x, prec)
    Km = sp.csr_matrix((K.astype(np.float64), col, rp), shape=(A.ndof, A.ndof))
    ref = Km @ x
    err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
    print(f"{prec}: max rel |Kx - CSR product| = {err:.1e}")
    assert err <= tol
