"""The oracle's boundary branches against closed forms that share no code
with it (tests/bc_closed_form.py): affine Dirichlet data at the Greville
abscissae, the order-0 and order-1 (double traction) Neumann loads from
B-spline integrals and end derivatives, and the periodic macro-field
par_periodic (src/assmcore_iga.c:225-229) through energy and residual at
U = 0 on a periodic patch."""
import os

import numpy as np

import bc_closed_form as cf
from conftest import GOLDEN

CASES = [((4, 3, 5), (0, 0, 0)), ((5, 4, 6), (0, 0, 1)), ((6, 5, 4), (1, 0, 0))]


def _rng_bc(oracle, seed, kind):
    rng = np.random.default_rng(seed)
    if kind == "dirichlet":
        bt = [int(v) for v in rng.integers(0, 2, size=36)]
        return oracle.make_bc(bc_type=bt, par_dirichlet=rng.standard_normal(9)), bt
    return oracle.make_bc(bc_type=[1] * 36, neumann=rng.standard_normal(36)), [1] * 36


def test_oracle_dirichlet_data_closed_form(oracle):
    for seed, (nel, per) in enumerate(CASES):
        m = oracle.Mesh(nel, L=(4, 1, 1), periodic=per)
        bc, bt = _rng_bc(oracle, seed, "dirichlet")
        mask, g = oracle.dirichlet(m, bc)
        em, eg = cf.dirichlet(m.knots, nel, per, bt, np.array(bc.par_dirichlet[:]))
        assert np.array_equal(mask, em)
        assert np.max(np.abs(g - eg)) < 1e-14


def test_oracle_neumann_load_closed_form(oracle):
    """Orders 0 and 1 on every non-periodic face with random tractions."""
    for seed, (nel, per) in enumerate(CASES):
        m = oracle.Mesh(nel, L=(4, 1, 1), periodic=per)
        bc, bt = _rng_bc(oracle, 100 + seed, "neumann")
        F = oracle.neumann(m, bc)
        eF = cf.neumann(m.knots, nel, per, bt, np.array(bc.neumann[:]))
        assert np.max(np.abs(F - eF)) <= 1e-13 * np.max(np.abs(eF))
        # and through the residual: at U = 0 the static volume term vanishes
        R = oracle.residual(m, oracle.make_phys(0), bc, np.zeros(m.ndof))
        assert np.max(np.abs(R - eF)) <= 1e-13 * np.max(np.abs(eF))


def test_oracle_neumann_planted_sign_fails(oracle):
    """A planted mistake in the closed form (order-1 derivative sign) must fail."""
    nel, per = CASES[0]
    m = oracle.Mesh(nel, L=(4, 1, 1), periodic=per)
    t = np.zeros(36)
    t[6 * 0 + 3 + 2] = 1.0   # face 0, order 1, dof z
    bc = oracle.make_bc(bc_type=[1] * 36, neumann=t)
    F = oracle.neumann(m, bc)
    eF = cf.neumann(m.knots, nel, per, [1] * 36, -t)
    assert np.max(np.abs(F - eF)) > 0.5 * np.max(np.abs(F))


def _periodic_A(seed):
    rng = np.random.default_rng(seed)
    return 0.05 * rng.standard_normal((3, 3))


def test_oracle_par_periodic_energy_and_zero_residual(oracle):
    """Periodic patch, U = 0, macro-field gradient A (not symmetric): the
    energy integral is V Psi(I + A) with Psi pinned by the compiled
    reference's energies, and the residual vanishes (uniform stress against
    basis gradients that integrate to zero on a periodic patch)."""
    g = np.load(os.path.join(GOLDEN, "pointwise_static.npz"), allow_pickle=False)
    L = (1.5, 1.0, 2.0)
    m = oracle.Mesh((5, 6, 7), L=L, periodic=(1, 1, 1))
    ph = oracle.make_phys(0, par=g["par"])
    coef = oracle.coef_from_par(g["par"])
    A = _periodic_A(7)
    bc = oracle.make_bc(par_periodic=A.reshape(-1))
    U = np.zeros(m.ndof)
    H = oracle.intdensity(m, ph, bc, U)
    psi = oracle.static_energy(cf.static_jet_of_gradient(A), coef)
    V = float(np.prod(L))
    assert abs(H[0] - V * psi) <= 1e-12 * abs(V * psi)
    # the transposed field gives a different energy (this energy is not
    # isotropic), so a transposed par_periodic index would fail above
    assert abs(oracle.static_energy(cf.static_jet_of_gradient(A.T), coef) - psi) > 1e-6 * abs(psi)
    R = oracle.residual(m, ph, bc, U)
    R_scale = np.max(np.abs(oracle.residual(m, ph, bc, oracle.glibc_rand_state(m.ndof))))
    assert np.max(np.abs(R)) <= 1e-11 * R_scale
