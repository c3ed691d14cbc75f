"""Edge shapes of the patch on the GPU against the oracle: one element in
some directions (every control point in one element), the smallest periodic
patch the GPU path takes (5 elements per periodic direction), and a long
thin patch; residual and FP64 tangent within 1e-12, the density integral
within 1e-12 (energy and stresses)."""
import numpy as np
import pytest

from test_gpu_parity import TOL64, rowfloor_err

pytestmark = pytest.mark.gpu

EDGES = [
    ("one_element", (1, 1, 1), (0, 0, 0), 0),
    ("flat_2x1x3", (2, 1, 3), (0, 0, 0), 0),
    ("min_periodic_transient", (5, 5, 5), (1, 1, 1), 2),
    ("thin_9x1x1_periodic_x", (9, 1, 1), (1, 0, 0), 0),
]


@pytest.mark.parametrize("name,nel,per,torder", EDGES, ids=[e[0] for e in EDGES])
def test_edge_shapes(igap, oracle, name, nel, per, torder):
    mo = oracle.Mesh(nel, L=(2, 1, 1), periodic=per, perturbed=True)
    mg = igap.MeshIGABeam(L=(2, 1, 1), nelem=nel, periodic=per, perturbed=True)
    ph_o = oracle.make_phys(torder)
    ph_g = igap.PhysGradElast() if torder == 0 else igap.PhysGradElastTime()
    bc_o, bc_g = oracle.make_bc(), igap.BCDefault()
    U = oracle.glibc_rand_state((torder + 1) * mo.ndof)
    R_o = oracle.residual(mo, ph_o, bc_o, U)
    rp, col, K_o = oracle.tangent(mo, ph_o, bc_o, U)
    H_o = oracle.intdensity(mo, ph_o, bc_o, U)
    with igap.Assm(mg, ph_g, bc_g) as A:
        R_g = A.residual(U)
        rp_g, col_g, K_g = A.tangent(U)
        A.set_uhist(U)
        H_g = A.intdensity()
    assert np.array_equal(rp_g, rp) and np.array_equal(col_g, col)
    er, ek = rowfloor_err(R_g, R_o), rowfloor_err(K_g, K_o, rp)
    nc = 7 if torder == 0 else 2
    eh = np.max(np.abs(H_g[:nc] - H_o[:nc])) / np.max(np.abs(H_o[:nc]))
    print(f"[{name}] ndof={mo.ndof} R {er:.2e} K {ek:.2e} HH {eh:.2e}")
    assert er < TOL64 and ek < TOL64 and eh < 1e-12
