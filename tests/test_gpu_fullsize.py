"""The HIP assembly at the throughput workload's full size: the 64x16x16
static SVK + gradient beam (16,384 elements, 64,152 dofs, 20.6M nonzeros),
clamped end + tip displacement 0.05, a smooth state with a seeded
perturbation.  Checked against the oracle where it finishes in seconds (the
whole residual; the complete rows of the tip-end control points from a
tangent over the last four element layers) and through properties that do
not depend on size: K v equals the central difference of the residual,
K = K^T, FP32 within its bound of FP64, and bitwise-repeatable results.  The
transient physics (gradelasttime, three history levels) is checked the same
way on the same patch."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NEL = (64, 16, 16)
TIP = 0.05
TOL64 = 1e-12
TOL32 = 2e-5
NTHREADS = min(16, os.cpu_count() or 1)


def rowfloor_err(a, b, rowptr=None):
    """max |a - b| / max(|b|, row inf-norm of b) (global inf-norm without rowptr)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if rowptr is None:
        floor = np.maximum(np.abs(b), np.max(np.abs(b)))
    else:
        rn = np.maximum.reduceat(np.abs(b), rowptr[:-1])
        floor = np.maximum(np.abs(b), np.repeat(rn, np.diff(rowptr)))
    floor = np.where(floor == 0, 1.0, floor)
    return float(np.max(np.abs(a - b) / floor))


@pytest.fixture(scope="module")
def full(igap, oracle):
    mesh_g = igap.MeshIGABeam(L=(4.0, 1.0, 1.0), nelem=NEL)
    mesh_o = oracle.Mesh(NEL, L=(4, 1, 1))
    ph_g, ph_o = igap.PhysGradElast(coef=igap.SVK_COEF), oracle.make_phys(0, coef=oracle.SVK_COEF)
    bc_g, bc_o = igap.BCClampedTip(tip=TIP), oracle.clamped_tip_bc(tip=TIP)
    rng = np.random.default_rng(20)
    U = igap.smooth_state(mesh_g, amp=0.01) + 2e-3 * rng.standard_normal(mesh_o.ndof)
    print(f"[fullsize] {NEL} ndof={mesh_o.ndof}", flush=True)
    with igap.Assm(mesh_g, ph_g, bc_g) as A:
        R = A.residual(U)
        rp, col, K = A.tangent(U, prec="f64")
        _, _, K32 = A.tangent(U, prec="f32")
        mask = A.dirichlet()[0].reshape(-1)
        R2 = A.residual(U)
        _, _, K2 = A.tangent(U, prec="f64")
        # central difference of the residual along a direction that vanishes
        # on the Dirichlet dofs (whose rows and columns the tangent replaces)
        v = rng.standard_normal(mesh_o.ndof) * (mask == 0)
        eps = 1e-6
        dR = (A.residual(U + eps * v) - A.residual(U - eps * v)) / (2 * eps)
    return dict(mesh_o=mesh_o, ph_o=ph_o, bc_o=bc_o, U=U, R=R, rp=rp, col=col, K=K, K32=K32,
                R2=R2, K2=K2, v=v, dR=dR)


def test_fullsize_residual_matches_oracle(oracle, full):
    R_o = oracle.residual(full["mesh_o"], full["ph_o"], full["bc_o"], full["U"], nthreads=NTHREADS)
    err = rowfloor_err(full["R"], R_o)
    print(f"[fullsize residual] max rel err {err:.3e}")
    assert err < TOL64


def test_fullsize_tip_rows_match_oracle(oracle, full):
    """Rows of the control points with x index >= 62 collect only elements
    60..63, so an oracle tangent over those four layers gives them whole; they
    hold the highest element, control-point and CSR indices of the patch and
    the tip's Dirichlet rows."""
    m = full["mesh_o"]
    nb = m.nb
    box = [NEL[0] - 4, NEL[0], 0, NEL[1], 0, NEL[2]]
    rp_o, col_o, val_o = oracle.tangent(m, full["ph_o"], full["bc_o"], full["U"], box=box, nthreads=NTHREADS)
    rp, col, K = full["rp"], full["col"], full["K"]
    assert np.array_equal(rp, rp_o) and np.array_equal(col, col_o)
    r0 = 3 * (nb[0] - 4) * nb[1] * nb[2]
    s0 = int(rp[r0])
    sub_rp = rp[r0:] - s0
    err = rowfloor_err(K[s0:], val_o[s0:], sub_rp)
    print(f"[fullsize tip rows] {len(rp) - 1 - r0} rows, max rel err {err:.3e}")
    assert err < TOL64


def test_fullsize_tangent_is_residual_derivative(full):
    import scipy.sparse as sp
    n = len(full["rp"]) - 1
    K = sp.csr_matrix((full["K"], full["col"], full["rp"]), shape=(n, n))
    Kv = K @ full["v"]
    rel = np.linalg.norm(Kv - full["dR"]) / np.linalg.norm(Kv)
    print(f"[fullsize K v vs central difference] rel err {rel:.3e}")
    assert rel < 1e-7


def test_fullsize_tangent_symmetric(full):
    import scipy.sparse as sp
    n = len(full["rp"]) - 1
    K = sp.csr_matrix((full["K"], full["col"], full["rp"]), shape=(n, n))
    asym = abs(K - K.T).max()
    print(f"[fullsize symmetry] max |K - K^T| / max |K| = {asym / abs(K).max():.3e}")
    assert asym <= 1e-13 * abs(K).max()


def test_fullsize_fp32_bound_and_repeatable(full):
    err = rowfloor_err(full["K32"], full["K"], full["rp"])
    print(f"[fullsize f32] max rel err vs f64 {err:.3e}")
    assert err < TOL32
    assert np.array_equal(full["R"], full["R2"]) and np.array_equal(full["K"], full["K2"])


# ---- the transient physics (gradelasttime, reference parameters with dt = 1
#      so that the inertia diagonal does not hide the stiffness) at the same size
@pytest.fixture(scope="module")
def full_transient(igap, oracle):
    par = list(oracle.REF_PAR_TRANSIENT)
    par[7] = 1.0
    mesh_g = igap.MeshIGABeam(L=(4.0, 1.0, 1.0), nelem=NEL)
    mesh_o = oracle.Mesh(NEL, L=(4, 1, 1))
    ph_g, ph_o = igap.PhysGradElastTime(par=par), oracle.make_phys(2, par=par)
    bc_g, bc_o = igap.BCClampedTip(tip=TIP), oracle.clamped_tip_bc(tip=TIP)
    rng = np.random.default_rng(21)
    n = mesh_o.ndof
    U1 = igap.smooth_state(mesh_g, amp=0.01) + 1e-3 * rng.standard_normal(n)
    U = np.concatenate([U1 + 2e-3 * rng.standard_normal(n), U1, U1 - 1e-3 * rng.standard_normal(n)])
    print(f"[fullsize transient] {NEL} ndof={n}", flush=True)
    with igap.Assm(mesh_g, ph_g, bc_g) as A:
        rp, col, K = A.tangent(U, prec="f64")
        _, _, K32 = A.tangent(U, prec="f32")
        mask = A.dirichlet()[0].reshape(-1)
        v = rng.standard_normal(n) * (mask == 0)
        eps = 1e-6
        Up, Um = U.copy(), U.copy()
        Up[:n] += eps * v
        Um[:n] -= eps * v
        dR = (A.residual(Up) - A.residual(Um)) / (2 * eps)
    return dict(mesh_o=mesh_o, ph_o=ph_o, bc_o=bc_o, U=U, rp=rp, col=col, K=K, K32=K32, v=v, dR=dR)


def test_fullsize_transient_tip_rows_match_oracle(oracle, full_transient):
    """Rows of the control points with x index >= 64 collect only elements 62
    and 63: an oracle tangent over those two layers gives them whole."""
    ft = full_transient
    m = ft["mesh_o"]
    nb = m.nb
    box = [NEL[0] - 2, NEL[0], 0, NEL[1], 0, NEL[2]]
    rp_o, col_o, val_o = oracle.tangent(m, ft["ph_o"], ft["bc_o"], ft["U"], box=box, nthreads=NTHREADS)
    assert np.array_equal(ft["rp"], rp_o) and np.array_equal(ft["col"], col_o)
    r0 = 3 * (nb[0] - 2) * nb[1] * nb[2]
    s0 = int(ft["rp"][r0])
    err = rowfloor_err(ft["K"][s0:], val_o[s0:], ft["rp"][r0:] - s0)
    print(f"[fullsize transient tip rows] {len(ft['rp']) - 1 - r0} rows, max rel err {err:.3e}")
    assert err < TOL64


def test_fullsize_transient_tangent_is_residual_derivative(full_transient):
    import scipy.sparse as sp
    ft = full_transient
    n = len(ft["rp"]) - 1
    K = sp.csr_matrix((ft["K"], ft["col"], ft["rp"]), shape=(n, n))
    Kv = K @ ft["v"]
    rel = np.linalg.norm(Kv - ft["dR"]) / np.linalg.norm(Kv)
    asym = abs(K - K.T).max() / abs(K).max()
    err32 = rowfloor_err(ft["K32"], ft["K"], ft["rp"])
    print(f"[fullsize transient] K v vs central difference {rel:.3e}, asymmetry {asym:.3e}, f32 {err32:.3e}")
    assert rel < 1e-7
    assert asym <= 1e-13
    assert err32 < TOL32
