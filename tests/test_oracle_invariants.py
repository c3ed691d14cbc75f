"""Pins the oracle's assembly (the restatement of Residual_iga / Tangent_iga)
where no reference output exists (assmcore_iga.c needs PETSc): by what the
mathematics fixes.  Each check is chosen so that a plausible slip (a dropped
term, a wrong sign or index, a transposed operand) fails one of them."""
import numpy as np
import pytest


def small_state(mesh, nh, seed=0, scale=0.01):
    rng = np.random.default_rng(seed)
    return scale * rng.normal(size=nh * mesh.ndof)


@pytest.fixture(scope="module")
def beam(oracle):
    return oracle.Mesh((3, 2, 2), L=(4.0, 1.0, 1.0), perturbed=True)


def test_partition_of_unity_and_affine_reproduction(oracle):
    """sum_b R_b = 1, sum_b dR_b = 0 and x = sum_b X_b R_b is reproduced with
    grad x = I and grad grad x = 0 on the distorted igabeamtest geometry --
    pins the NURBS quotient rule, G^-1 and the second-derivative correction."""
    m = oracle.Mesh((6, 6, 6), L=(2.0, 1.0, 1.5), perturbed=True)
    Xa = m.X.reshape(m.na[0], m.na[1], m.na[2], 3)
    for e in [(0, 0, 0), (3, 2, 4), (5, 5, 5), (2, 3, 3)]:
        Xe = Xa[e[0]:e[0] + 3, e[1]:e[1] + 3, e[2]:e[2] + 3].reshape(27, 3)
        for q in [(0, 0, 0), (1, 2, 3), (3, 3, 0)]:
            R, detG, w = oracle.element_basis(m, e, q)
            s = R.sum(axis=0)
            assert abs(s[0] - 1) < 1e-13 and np.max(np.abs(s[1:])) < 1e-10
            jet = R.T @ Xe          # [10, 3]: jet of x_j
            assert np.allclose(jet[1:4], np.eye(3), atol=1e-11)
            assert np.max(np.abs(jet[4:])) < 1e-9
            assert detG > 0


def test_quadratic_reproduction_second_derivatives(oracle):
    """On the periodic uniform patch (x = L xi), f = x1^2 has B-spline
    coefficients L^2 t_{i+1} t_{i+2} (blossom); the physical jet must give
    df/dx1 = 2 x1 and d2f/dx1^2 = 2."""
    L = (2.0, 1.0, 1.0)
    m = oracle.Mesh((6, 6, 6), L=L, periodic=(1, 1, 1))
    kv = m.knots[0]
    c1 = np.array([L[0] ** 2 * kv[i + 1] * kv[i + 2] for i in range(m.na[0])])
    for e in [(0, 1, 2), (4, 0, 5)]:
        ce = np.repeat(c1[e[0]:e[0] + 3], 9)
        for q in [(0, 1, 2), (3, 0, 1)]:
            R, _, _ = oracle.element_basis(m, e, q)
            jet = R.T @ ce
            x1 = (1 - oracle.gauss4()[0][q[0]]) * kv[e[0] + 2] + oracle.gauss4()[0][q[0]] * kv[e[0] + 3]
            assert abs(jet[0] - (L[0] * x1) ** 2) < 1e-12
            assert abs(jet[1] - 2 * L[0] * x1) < 1e-12
            assert abs(jet[4] - 2.0) < 1e-10
            assert np.max(np.abs(jet[[2, 3, 5, 6, 7, 8, 9]])) < 1e-10


def test_volume(oracle):
    """sum w_q detG: unit periodic cube -> 1 (the ex1 Volume KAT,
    example/ex1/main.c:50); igabeam -> prod L_d (1 + 1/n_d) (the end control
    points sit at -L/(2n) and L(n+1/2)/n)."""
    assert abs(oracle.volume(oracle.Mesh((5, 6, 7), L=(1, 1, 1), periodic=(1, 1, 1))) - 1.0) < 1e-13
    m = oracle.Mesh((4, 3, 2), L=(4, 1, 1))
    exact = np.prod([Ld * (1 + 1.0 / n) for Ld, n in zip(m.L, m.nelem)])
    assert abs(oracle.volume(m) - exact) < 1e-12 * exact


def test_residual_zero_state(oracle, beam):
    ph = oracle.make_phys(0)
    R = oracle.residual(beam, ph, oracle.make_bc(), np.zeros(beam.ndof))
    assert np.max(np.abs(R)) == 0.0


def test_residual_is_energy_gradient(oracle, beam):
    """R = dPi/dU, Pi = sum_q w_q Psi(u_q): pins the projection, the weights
    and the u-jet assembly (a transposed or dropped term breaks it)."""
    ph = oracle.make_phys(0)
    bc = oracle.make_bc()
    U = small_state(beam, 1, seed=1, scale=0.02)
    R = oracle.residual(beam, ph, bc, U)
    h = 1e-6
    rng = np.random.default_rng(2)
    for j in rng.choice(beam.ndof, 12, replace=False):
        e = np.zeros(beam.ndof)
        e[j] = h
        g = (oracle.energy(beam, ph, bc, U + e) - oracle.energy(beam, ph, bc, U - e)) / (2 * h)
        assert abs(g - R[j]) < 1e-7 * np.max(np.abs(R))


@pytest.mark.parametrize("torder", [0, 2])
def test_tangent_is_residual_jacobian(oracle, beam, torder):
    """K = dR/dU (level 0) by central differences, both physics."""
    ph = oracle.make_phys(torder)
    bc = oracle.make_bc()
    U = small_state(beam, torder + 1, seed=3, scale=0.02)
    rp, col, val = oracle.tangent(beam, ph, bc, U)
    K = oracle.csr_to_dense(rp, col, val, beam.ndof)
    h = 1e-6
    rng = np.random.default_rng(4)
    for j in rng.choice(beam.ndof, 10, replace=False):
        e = np.zeros_like(U)
        e[j] = h
        fd = (oracle.residual(beam, ph, bc, U + e) - oracle.residual(beam, ph, bc, U - e)) / (2 * h)
        assert np.max(np.abs(fd - K[:, j])) < 1e-6 * np.max(np.abs(K))


def test_tangent_symmetric_and_translation_nullspace(oracle, beam):
    ph = oracle.make_phys(0)
    bc = oracle.make_bc()
    U = small_state(beam, 1, seed=5)
    rp, col, val = oracle.tangent(beam, ph, bc, U)
    K = oracle.csr_to_dense(rp, col, val, beam.ndof)
    assert np.max(np.abs(K - K.T)) < 1e-13 * np.max(np.abs(K))
    for k in range(3):
        t = np.zeros(beam.ndof)
        t[k::3] = 1.0
        assert np.max(np.abs(K @ t)) < 1e-10 * np.max(np.abs(K))
        R0 = oracle.residual(beam, ph, bc, U)
        R1 = oracle.residual(beam, ph, bc, U + 0.3 * t)
        assert np.max(np.abs(R1 - R0)) < 1e-12 * max(1.0, np.max(np.abs(R0)))


def test_setnz_counts_equal_pattern_row_lengths(oracle):
    for nel, per in (((3, 2, 2), (0, 0, 0)), ((5, 4, 6), (1, 0, 1)), ((7, 5, 5), (1, 1, 1))):
        m = oracle.Mesh(nel, periodic=per)
        rp, col = oracle.csr_pattern(m)
        assert np.array_equal(oracle.setnz(m), np.diff(rp))
        # closed form nnz = 9 prod_d (5 n_d - 6) for non-periodic directions
        if per == (0, 0, 0):
            assert rp[-1] == 9 * np.prod([5 * (n + 2) - 6 for n in nel])
        for r in range(0, m.ndof, 7):
            c = col[rp[r]:rp[r + 1]]
            assert np.all(np.diff(c) > 0) and r in c


def test_dirichlet_affine_data(oracle):
    """dirichlet_iga reproduces u = a xi on a face: order-0 values are a times
    the Greville abscissae (in knot space) of the in-face direction, and the
    order-1 layer repeats them (zero normal gradient)."""
    m = oracle.Mesh((4, 3, 3), L=(4.0, 1.0, 1.0))
    bt = [1] * 36
    bt[6 * 2 + 0 + 2] = 0       # face 2 (y = 0), order 0, dof z
    bt[6 * 2 + 3 + 2] = 0       # face 2, order 1, dof z
    pd = [0.0] * 9
    pd[3 * 2 + 0] = 0.7         # u_z = 0.7 xi_x
    mask, g = oracle.dirichlet(m, oracle.make_bc(bc_type=bt, par_dirichlet=pd))
    kv = m.knots[0]
    grev = np.array([(kv[i + 1] + kv[i + 2]) / 2 for i in range(m.nb[0])])
    nb = m.nb
    G = g.reshape(nb[0], nb[1], nb[2], 3)
    Mk = mask.reshape(nb[0], nb[1], nb[2], 3)
    assert Mk[:, 0, :, 2].all() and Mk[:, 1, :, 2].all() and not Mk[:, 2:, :, 2].any()
    assert not Mk[..., :2].any()
    for layer in (0, 1):
        assert np.allclose(G[:, layer, :, 2], 0.7 * grev[:, None], atol=1e-15)


def test_clamped_tip_bc_rows(oracle):
    m = oracle.Mesh((4, 2, 2))
    bc = oracle.clamped_tip_bc(tip=0.05, tip_dof=2)
    mask, g = oracle.dirichlet(m, bc)
    nb = m.nb
    Mk = mask.reshape(nb[0], nb[1], nb[2], 3)
    G = g.reshape(nb[0], nb[1], nb[2], 3)
    assert Mk[0].all() and Mk[1].all()                 # clamped: two layers, all dofs
    assert Mk[-1, :, :, 2].all() and not Mk[-1, :, :, :2].any()
    assert np.all(G[0] == 0) and np.all(G[1] == 0)
    assert np.allclose(G[-1, :, :, 2], 0.05 * (1 - 1e-15), rtol=0, atol=1e-17)
    U = small_state(m, 1, seed=6)
    ph = oracle.make_phys(0)
    R = oracle.residual(m, ph, bc, U)
    assert np.array_equal(R[mask == 1], (U - g)[mask == 1])
    rp, col, val = oracle.tangent(m, ph, bc, U)
    K = oracle.csr_to_dense(rp, col, val, m.ndof)
    d = np.where(mask == 1)[0]
    assert np.array_equal(K[d][:, d], np.eye(len(d)))
    assert np.all(K[d][:, mask == 0] == 0) and np.all(K[mask == 0][:, d] == 0)


def test_periodic_wrap_consistency(oracle):
    """A periodic patch assembled whole equals the same physics seen through a
    shifted origin: shifting U by one owned control point in x shifts R."""
    m = oracle.Mesh((5, 5, 5), L=(1, 1, 1), periodic=(1, 1, 1))
    ph = oracle.make_phys(2)
    bc = oracle.make_bc()
    rng = np.random.default_rng(7)
    U = 0.01 * rng.normal(size=3 * m.ndof)
    R = oracle.residual(m, ph, bc, U).reshape(5, 5, 5, 3)
    Us = np.roll(U.reshape(3, 5, 5, 5, 3), 1, axis=1).reshape(-1)
    Rs = oracle.residual(m, ph, bc, Us).reshape(5, 5, 5, 3)
    assert np.max(np.abs(np.roll(R, 1, axis=0) - Rs)) < 1e-10 * np.max(np.abs(R))


def test_neumann_constant_traction(oracle):
    """Order-0 Neumann with unit traction on face 1 integrates the face basis:
    the sum of the load equals minus the parametric face area (1 x 1)."""
    m = oracle.Mesh((3, 4, 2))
    nm = [0.0] * 36
    nm[6 * 1 + 0 + 0] = 1.0
    F = oracle.neumann(m, oracle.make_bc(neumann=nm))
    assert abs(F[0::3].sum() + 1.0) < 1e-14 and np.all(F[1::3] == 0) and np.all(F[2::3] == 0)


def test_intdensity_volume_energy_and_traction_work(oracle):
    """iga_intdensity: the transient's second density integrates to the volume
    (ex1's KAT: 1 on the unit cube), the static energy density integrates to
    the oracle's energy, and a Neumann traction adds -int u.t, which for a
    uniform traction on a face with u constant there is -area * u * t."""
    m = oracle.Mesh((4, 4, 4), L=(1, 1, 1), periodic=(1, 1, 1))
    pht = oracle.make_phys(2)
    U = oracle.glibc_rand_state(3 * m.ndof)
    HH = oracle.intdensity(m, pht, oracle.make_bc(), U)
    assert HH.size == 2 and abs(HH[1] - 1.0) < 1e-12
    mb = oracle.Mesh((5, 3, 3), L=(4, 1, 1))
    ph = oracle.make_phys(0, coef=oracle.SVK_COEF)
    bc = oracle.make_bc()
    U = oracle.smooth_state(mb, amp=0.01)
    HH = oracle.intdensity(mb, ph, bc, U)
    assert HH.size == 13 and abs(HH[0] - oracle.energy(mb, ph, bc, U)) < 1e-12 * abs(HH[0])
    # traction t on face 1 (x = L) in z, u_z = c everywhere (a rigid shift):
    # the boundary term is -area * c * t and the volume terms vanish
    neu = [0.0] * 36
    neu[6 * 1 + 3 * 0 + 2] = 0.3          # face 1, order 0, dof z
    bc_t = oracle.make_bc(neumann=neu)
    c = 0.02
    U = np.zeros(oracle.Mesh((5, 3, 3), L=(4, 1, 1)).ndof)
    U[2::3] = c
    HH = oracle.intdensity(mb, ph, bc_t, U)
    assert abs(HH[0] - (-1.0 * c * 0.3)) < 1e-13
