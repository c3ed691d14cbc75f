"""The product's boundary data and the periodic macro-field on the GPU against
closed forms that share no code with the library or the oracle
(tests/bc_closed_form.py): affine Dirichlet data at the Greville abscissae,
order-0 and order-1 Neumann loads from B-spline integrals and end
derivatives (src/assmcore_iga.c:273-405, 861-934), and par_periodic
(src/assmcore_iga.c:225-229) through the energy integral, the residual and
the tangent at a periodic patch."""
import os

import numpy as np
import pytest

import bc_closed_form as cf
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

CASES = [((4, 3, 5), (0, 0, 0)), ((5, 4, 6), (0, 0, 1)), ((6, 5, 4), (1, 0, 0))]


def test_dirichlet_rows_closed_form(igap):
    for seed, (nel, per) in enumerate(CASES):
        rng = np.random.default_rng(seed)
        bt = [int(v) for v in rng.integers(0, 2, size=36)]
        par = rng.standard_normal(9)
        mesh = igap.MeshIGABeam(L=(4, 1, 1), nelem=nel, periodic=per)
        with igap.Assm(mesh, igap.PhysGradElast(), igap.BC(bc_type=bt, par_dirichlet=par), device=0) as A:
            mask, g = A.dirichlet()
            # the residual's Dirichlet rows are U - g
            U = 0.01 * rng.standard_normal(A.ndof)
            R = A.residual(U)
        em, eg = cf.dirichlet(mesh.knots, nel, per, bt, par)
        print(f"case {nel} {per}: {int(em.sum())} Dirichlet rows, max |g - closed form| {np.max(np.abs(g - eg)):.1e}")
        assert np.array_equal(mask, em)
        assert np.max(np.abs(g - eg)) < 1e-14
        rows = em == 1
        assert np.max(np.abs(R[rows] - (U[rows] - eg[rows]))) < 1e-14


def test_neumann_load_closed_form(igap):
    """Random tractions of both orders on every non-periodic face: at U = 0
    the static volume term vanishes and the residual is the load itself."""
    for seed, (nel, per) in enumerate(CASES):
        rng = np.random.default_rng(100 + seed)
        t = rng.standard_normal(36)
        mesh = igap.MeshIGABeam(L=(4, 1, 1), nelem=nel, periodic=per)
        with igap.Assm(mesh, igap.PhysGradElast(), igap.BC(bc_type=[1] * 36, neumann=t), device=0) as A:
            R = A.residual(np.zeros(A.ndof))
        eF = cf.neumann(mesh.knots, nel, per, [1] * 36, t)
        err = np.max(np.abs(R - eF)) / np.max(np.abs(eF))
        print(f"case {nel} {per}: max rel |R - closed form| {err:.1e}")
        assert err <= 1e-13


def _periodic_setup(igap, A_field):
    g = np.load(os.path.join(GOLDEN, "pointwise_static.npz"), allow_pickle=False)
    L = (1.5, 1.0, 2.0)
    mesh = igap.MeshIGABeam(L=L, nelem=(5, 6, 7), periodic=(1, 1, 1))
    bc = igap.BC(par_periodic=A_field.reshape(-1))
    return g, L, mesh, igap.PhysGradElast(par=g["par"]), bc


def test_par_periodic_energy_and_zero_residual(igap, oracle):
    """U = 0 with the macro-field gradient A (not symmetric): energy integral
    V Psi(I + A), Psi pinned by the compiled reference's energies (the oracle's
    pointwise energy equals them to 1e-12 in test_oracle_golden), and a
    vanishing residual."""
    A_field = 0.05 * np.random.default_rng(7).standard_normal((3, 3))
    g, L, mesh, phys, bc = _periodic_setup(igap, A_field)
    coef = oracle.coef_from_par(g["par"])
    psi = oracle.static_energy(cf.static_jet_of_gradient(A_field), coef)
    V = float(np.prod(L))
    with igap.Assm(mesh, phys, bc, device=0) as A:
        A.set_uhist(np.zeros(A.ndof))
        H = A.intdensity()
        R0 = A.residual(np.zeros(A.ndof))
        Rr = A.residual(oracle.glibc_rand_state(A.ndof))
    print(f"energy {H[0]:.15e} vs V Psi {V * psi:.15e}; max|R(0)| {np.max(np.abs(R0)):.1e} "
          f"(scale {np.max(np.abs(Rr)):.1e})")
    assert abs(H[0] - V * psi) <= 1e-12 * abs(V * psi)
    assert np.max(np.abs(R0)) <= 1e-11 * np.max(np.abs(Rr))


def test_par_periodic_tangent_is_residual_derivative(igap, oracle):
    """With A != 0, K v equals the central difference of the residual."""
    import scipy.sparse as sp
    A_field = 0.05 * np.random.default_rng(8).standard_normal((3, 3))
    _, _, mesh, phys, bc = _periodic_setup(igap, A_field)
    with igap.Assm(mesh, phys, bc, device=0) as A:
        U = oracle.glibc_rand_state(A.ndof)
        rp, col, K = A.tangent(U)
        v = np.random.default_rng(9).standard_normal(A.ndof)
        eps = 1e-6
        fd = (A.residual(U + eps * v) - A.residual(U - eps * v)) / (2 * eps)
    Kv = sp.csr_matrix((K, col, rp), shape=(A.ndof, A.ndof)) @ v
    err = np.max(np.abs(Kv - fd)) / np.max(np.abs(Kv))
    print(f"K v vs central difference: {err:.1e}")
    assert err < 1e-6
