"""Parity of the HIP assembly (through the C ABI) with the oracle on the same
seeded inputs: entry-wise within 1e-12 relative for FP64, with the row
infinity-norm as floor for cancellation-dominated near-zero entries."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
# FP32 matrix-core tangent against the FP64 one: FP32 rounding of the basis
# jets, the pointwise tangent and the products, accumulated over 64 Gauss
# points and up to 27 elements per entry
TOL32 = 2e-5
NTHREADS = min(16, os.cpu_count() or 1)


def rowfloor_err(a, b, rowptr=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if rowptr is None:
        floor = np.maximum(np.abs(b), np.max(np.abs(b)))
    else:
        rn = np.maximum.reduceat(np.abs(b), rowptr[:-1])
        floor = np.maximum(np.abs(b), np.repeat(rn, np.diff(rowptr)))
    floor = np.where(floor == 0, 1.0, floor)
    return float(np.max(np.abs(a - b) / floor))


def cases():
    # (name, nelem, L, periodic, perturbed, torder, coef, bc kind, state)
    return [
        ("clamped_tip_static_ref", (6, 3, 3), (4, 1, 1), (0, 0, 0), True, 0, None, "tip", "rand"),
        ("ex1_periodic_transient", (8, 8, 8), (1, 1, 1), (1, 1, 1), False, 2, None, "default", "rand"),
        ("svk_beam_smooth", (5, 3, 3), (4, 1, 1), (0, 0, 0), False, 0, "svk", "tip", "smooth"),
        ("mixed_periodic_traction", (5, 3, 5), (1, 1, 1), (1, 0, 1), True, 2, None, "mixed", "rand"),
    ]


def build_case(igap, oracle, c):
    name, nel, L, per, pert, torder, coef, bck, st = c
    mesh_o = oracle.Mesh(nel, L=L, periodic=per, perturbed=pert)
    mesh_g = igap.MeshIGABeam(L=L, nelem=nel, periodic=per, perturbed=pert)
    cf = oracle.SVK_COEF if coef == "svk" else None
    ph_o = oracle.make_phys(torder, coef=cf)
    if torder == 0:
        ph_g = igap.PhysGradElast(coef=cf)
    else:
        ph_g = igap.PhysGradElastTime(coef=cf)
    if bck == "tip":
        bc_o = oracle.clamped_tip_bc(tip=0.02)
        bc_g = igap.BCClampedTip(tip=0.02)
    elif bck == "mixed":
        bt = [1] * 36
        bt[6 * 2 + 0 + 1] = 0      # face 2 (y=0) order 0 dof y
        bt[6 * 3 + 3 + 0] = 0      # face 3 (y=1) order 1 dof x
        pp = [0.01, 0.0, 0.02, 0.0, -0.01, 0.0, 0.0, 0.0, 0.005]
        pd = [0.0, 0.3, 0.0, 0.1, 0.0, 0.0, 0.0, 0.2, 0.0]
        nm = [0.0] * 36
        nm[6 * 3 + 0 + 2] = 0.7     # face 3 order-0 traction on dof z
        nm[6 * 2 + 3 + 0] = -0.4    # face 2 order-1 (double traction) on dof x
        bc_o = oracle.make_bc(bc_type=bt, par_periodic=pp, par_dirichlet=pd, neumann=nm)
        bc_g = igap.BC(bc_type=bt, par_periodic=pp, par_dirichlet=pd, neumann=nm)
    else:
        bc_o = oracle.make_bc()
        bc_g = igap.BCDefault()
    nh = torder + 1
    if st == "rand":
        U = oracle.glibc_rand_state(nh * mesh_o.ndof)
    else:
        U = oracle.smooth_state(mesh_o, amp=0.02, levels=nh)
    return mesh_o, mesh_g, ph_o, ph_g, bc_o, bc_g, U


@pytest.fixture(scope="module")
def built(igap, oracle):
    return {c[0]: build_case(igap, oracle, c) for c in cases()}


def _layout_report(raw):
    import igap4_amd as ig
    ref = ig.selftest_expected()
    out = {}
    for name, fn in (("4r+kr", lambda kr, r: 4 * r + kr), ("4kr+r", lambda kr, r: 4 * kr + r)):
        out[name] = np.array_equal(ig.selftest_unpack(raw.astype(np.float64), fn), ref)
    return out


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_mfma_operand_layout(igap, prec):
    """The 16x16x4 operand/result lane layout the tangent kernels assume."""
    d64, d32 = igap.selftest_mfma()
    raw = d64 if prec == "f64" else d32
    print(f"[mfma {prec}] row maps matching: {_layout_report(raw)}", flush=True)
    D = igap.selftest_unpack(raw.astype(np.float64), igap.MFMA_ROW[prec])
    assert np.array_equal(D, igap.selftest_expected())


@pytest.mark.parametrize("case", [c[0] for c in cases()])
def test_residual_parity(igap, oracle, built, case):
    mesh_o, mesh_g, ph_o, ph_g, bc_o, bc_g, U = built[case]
    print(f"[residual] {case} nelem={mesh_o.nelem}", flush=True)
    R_o = oracle.residual(mesh_o, ph_o, bc_o, U, nthreads=NTHREADS)
    with igap.Assm(mesh_g, ph_g, bc_g) as A:
        R_g = A.residual(U)
        # U_hist level 0 keeps the iterate (assmcore_iga.c:68)
        assert np.array_equal(A.get_uhist(), U)
    err = rowfloor_err(R_g, R_o)
    print(f"  max rel err {err:.3e}")
    assert err < TOL64


@pytest.mark.parametrize("case", [c[0] for c in cases()])
def test_tangent_parity_f64(igap, oracle, built, case):
    mesh_o, mesh_g, ph_o, ph_g, bc_o, bc_g, U = built[case]
    print(f"[tangent f64] {case} nelem={mesh_o.nelem}", flush=True)
    rp_o, col_o, val_o = oracle.tangent(mesh_o, ph_o, bc_o, U, nthreads=NTHREADS)
    with igap.Assm(mesh_g, ph_g, bc_g) as A:
        rp, col, val = A.tangent(U, prec="f64")
    assert np.array_equal(rp, rp_o) and np.array_equal(col, col_o)
    err = rowfloor_err(val, val_o, rp_o)
    print(f"  max rel err {err:.3e}")
    assert err < TOL64


def test_bitwise_repeatable(igap, built):
    _, mesh_g, _, ph_g, _, bc_g, U = built["ex1_periodic_transient"]
    with igap.Assm(mesh_g, ph_g, bc_g) as A:
        R1 = A.residual(U)
        _, _, K1 = A.tangent(U)
        R2 = A.residual(U)
        _, _, K2 = A.tangent(U)
        _, _, F1 = A.tangent(U, prec="f32")
        _, _, F2 = A.tangent(U, prec="f32")
    assert np.array_equal(R1, R2) and np.array_equal(K1, K2) and np.array_equal(F1, F2)


def test_dirichlet_rows_match_oracle(igap, oracle, built):
    mesh_o, mesh_g, ph_o, ph_g, bc_o, bc_g, U = built["mixed_periodic_traction"]
    m_o, g_o = oracle.dirichlet(mesh_o, bc_o)
    with igap.Assm(mesh_g, ph_g, bc_g) as A:
        m_g, g_g = A.dirichlet()
    assert np.array_equal(m_o, m_g)
    assert np.max(np.abs(g_o - g_g)) <= 1e-15


@pytest.mark.parametrize("case", ["clamped_tip_static_ref", "mixed_periodic_traction"])
def test_tangent_f32_error_bound(igap, built, case):
    """FP32 matrix-core tangent against the FP64 path (the FP64 tangent is
    itself pinned to the oracle above)."""
    _, mesh_g, _, ph_g, _, bc_g, U = built[case]
    with igap.Assm(mesh_g, ph_g, bc_g) as A:
        rp, _, v64 = A.tangent(U, prec="f64")
        _, _, v32 = A.tangent(U, prec="f32")
    err = rowfloor_err(v32, v64, rp)
    print(f"[tangent f32] {case} max rel err vs f64 {err:.3e}")
    assert err < TOL32
