"""The sum-factorised residual and tangent of separable patches
(k_residual_sep, k_tangent_sep; transient: k_residual_sep_t, k_tangent_sep_t).

A patch is separable when every control coordinate x_d depends on the control
point's index along d alone and all weights are equal (the test of
igap_gpu_create, restated in ``separable`` below).  There the element vectors
and matrices come from 1-D tables of physical derivatives, and the static
FP64 residual and tangent must match the oracle (the reference's loop nests,
src/assmcore_iga.c:118-259 and 607-764) entry-wise within 1e-12 relative,
and the general-geometry kernels (flag IGAP_GENERAL_GEOMETRY) within the same
bound.  Cases: the igabeam beam (its end elements map quadratically, so the
g' term of the second derivative is exercised), a graded patch (non-uniform
knots, control coordinates a nonlinear function of the index, all weights
2), a fully periodic patch with the par_periodic macro-field, and a truly
distorted igabeamtest patch, which is not separable and takes the general
kernel whatever the flag.  The "_t" cases use the transient physics
(gradelasttime, three history levels of a random state; the tangent's
rule-weighted rows and inertia term, src/assmcore_iga.c:607-764 with
phys/gradelasttime.c) on the same patches: the tangent and the residual
take the sum-factorised kernels (k_tangent_sep_t, k_residual_sep_t)."""
import numpy as np
import pytest

from test_gpu_parity import NTHREADS, TOL32, TOL64, rowfloor_err

pytestmark = pytest.mark.gpu


def separable(mesh):
    """x_d a function of the index along d alone, weights all equal (bitwise)."""
    X = mesh.X.reshape(mesh.na[0], mesh.na[1], mesh.na[2], 3)
    if not np.all(mesh.w == mesh.w[0]):
        return False
    return (np.array_equal(X[..., 0], np.broadcast_to(X[:, :1, :1, 0], X.shape[:3]))
            and np.array_equal(X[..., 1], np.broadcast_to(X[:1, :, :1, 1], X.shape[:3]))
            and np.array_equal(X[..., 2], np.broadcast_to(X[:1, :1, :, 2], X.shape[:3])))


def grade(mesh):
    """Non-uniform open knots t_i = (i/n)^1.3, control coordinates
    L_d (greville + 0.02 sin(1.7 i)) along each direction, weights 2: still
    separable, with a curved map inside every element."""
    idx = np.indices(mesh.na).reshape(3, -1)
    X = mesh.X.reshape(-1, 3)
    for d in range(3):
        n, k = mesh.nelem[d], mesh.knots[d]
        k[2:n + 3] = (np.arange(n + 1) / n) ** 1.3
        i = np.arange(mesh.na[d])
        grev = 0.5 * (k[i + 1] + k[i + 2])
        X[:, d] = mesh.L[d] * (grev + 0.02 * np.sin(1.7 * i))[idx[d]]
    mesh.w[:] = 2.0


def build(igap, oracle, name):
    torder = 2 if name.endswith("_t") else 0
    name = name[:-2] if torder else name
    par_periodic = None
    if name == "beam":
        kw = dict(nelem=(16, 4, 4), L=(4, 1, 1))
    elif name == "graded":
        kw = dict(nelem=(6, 4, 5), L=(3, 1, 1.5))
    elif name == "periodic":
        kw = dict(nelem=(8, 8, 8), L=(1, 1, 1), periodic=(1, 1, 1))
        par_periodic = [0.01, 0.0, 0.02, 0.0, -0.01, 0.0, 0.005, 0.0, 0.004]
    else:   # distorted: nelem >= 4 in every direction, so the perturbation applies
        kw = dict(nelem=(8, 5, 5), L=(4, 1, 1), perturbed=True)
    mo = oracle.Mesh(kw["nelem"], L=kw["L"], periodic=kw.get("periodic", (0, 0, 0)), perturbed=kw.get("perturbed", False))
    mg = igap.MeshIGABeam(**kw)
    if name == "graded":
        grade(mo)
        grade(mg)
    if par_periodic is not None:
        bc_o, bc_g = oracle.make_bc(par_periodic=par_periodic), igap.BC(par_periodic=par_periodic)
    elif any(kw.get("periodic", (0, 0, 0))):
        bc_o, bc_g = oracle.make_bc(), igap.BCDefault()
    else:
        bc_o, bc_g = oracle.clamped_tip_bc(tip=0.02), igap.BCClampedTip(tip=0.02)
    nh = torder + 1
    if name == "beam":
        U = oracle.smooth_state(mo, amp=0.02, levels=nh)
        for h in range(1, nh):   # distinct history levels (the rule's Ud != 0)
            U[h * mo.ndof:(h + 1) * mo.ndof] *= 1.0 - 0.12 * h
    else:
        U = oracle.glibc_rand_state(nh * mo.ndof) * 0.05
    return mo, mg, bc_o, bc_g, U, torder


def phys(igap, oracle, torder):
    if torder == 0:
        return oracle.make_phys(0), igap.PhysGradElast()
    return oracle.make_phys(2), igap.PhysGradElastTime()


CASES = ["beam", "graded", "periodic", "distorted", "beam_t", "graded_t", "periodic_t", "distorted_t"]


@pytest.fixture(scope="module")
def built(igap, oracle):
    return {n: build(igap, oracle, n) for n in CASES}


def test_separable_cases(built):
    for n in CASES:
        assert separable(built[n][1]) == (not n.startswith("distorted")), n
        assert np.array_equal(built[n][0].X, built[n][1].X) and np.array_equal(built[n][0].w, built[n][1].w)


@pytest.mark.parametrize("case", CASES)
def test_tangent_vs_oracle(igap, oracle, built, case):
    mo, mg, bc_o, bc_g, U, torder = built[case]
    ph_o, ph_g = phys(igap, oracle, torder)
    rp_o, col_o, K_o = oracle.tangent(mo, ph_o, bc_o, U, nthreads=NTHREADS)
    with igap.Assm(mg, ph_g, bc_g) as A:
        rp, col, K = A.tangent(U, prec="f64")
    assert np.array_equal(rp, rp_o) and np.array_equal(col, col_o)
    err = rowfloor_err(K, K_o, rp_o)
    print(f"[{case}] separable={separable(mg)} max rel err vs oracle {err:.3e}")
    assert err < TOL64


@pytest.mark.parametrize("case", CASES)
def test_residual_vs_oracle_and_general(igap, oracle, built, case):
    from igap4_amd import _lib
    mo, mg, bc_o, bc_g, U, torder = built[case]
    ph_o, ph_g = phys(igap, oracle, torder)
    R_o = oracle.residual(mo, ph_o, bc_o, U, nthreads=NTHREADS)
    with igap.Assm(mg, ph_g, bc_g) as A:
        R = A.residual(U)
        R2 = A.residual(U)
    with igap.Assm(mg, ph_g, bc_g, flags=_lib.IGAP_GENERAL_GEOMETRY) as A:
        Rg = A.residual(U)
    e_o, e_g = rowfloor_err(R, R_o), rowfloor_err(R, Rg)
    print(f"[{case}] residual: max rel err vs oracle {e_o:.3e}, vs general kernel {e_g:.3e}")
    assert np.array_equal(R, R2)
    if case.startswith("distorted"):   # not separable: the general residual kernel either way
        assert np.array_equal(R, Rg)
    assert e_o < TOL64 and e_g < TOL64


@pytest.mark.parametrize("case", CASES)
def test_sum_factorised_vs_general_kernel(igap, oracle, built, case):
    _, mg, _, bc_g, U, torder = built[case]
    from igap4_amd import _lib
    ph = phys(igap, oracle, torder)[1]
    out = {}
    for name, fl in (("default", 0), ("general", _lib.IGAP_GENERAL_GEOMETRY)):
        with igap.Assm(mg, ph, bc_g, flags=fl) as A:
            rp, _, k64 = A.tangent(U, prec="f64")
            _, _, k32 = A.tangent(U, prec="f32")
            out[name] = (k64, k32)
    e64 = rowfloor_err(out["default"][0], out["general"][0], rp)
    e32 = rowfloor_err(out["default"][1], out["general"][0], rp)
    print(f"[{case}] default vs general FP64 {e64:.3e}; default FP32 vs general FP64 {e32:.3e}")
    if case.startswith("distorted"):
        # not separable: both flags take the general kernel
        assert np.array_equal(out["default"][0], out["general"][0])
        assert np.array_equal(out["default"][1], out["general"][1])
    assert e64 < TOL64
    assert e32 < TOL32


def test_full_patch_and_repeatable(igap):
    """64x16x16 beam: the two kernels agree on every entry; the
    sum-factorised one is bitwise repeatable."""
    from igap4_amd import _lib
    mesh = igap.MeshIGABeam(L=(4, 1, 1), nelem=(64, 16, 16))
    U = igap.smooth_state(mesh, amp=0.02)
    ph, bc = igap.PhysGradElast(coef=igap.SVK_COEF), igap.BCClampedTip(tip=0.05)
    with igap.Assm(mesh, ph, bc) as A:
        rp, _, k1 = A.tangent(U)
        _, _, k2 = A.tangent(U)
        _, _, f1 = A.tangent(U, prec="f32")
        _, _, f2 = A.tangent(U, prec="f32")
    with igap.Assm(mesh, ph, bc, flags=_lib.IGAP_GENERAL_GEOMETRY) as A:
        _, _, kg = A.tangent(U)
    assert np.array_equal(k1, k2) and np.array_equal(f1, f2)
    e64, e32 = rowfloor_err(k1, kg, rp), rowfloor_err(f1, kg, rp)
    print(f"[64x16x16] sum-factorised vs general FP64 {e64:.3e}, FP32 vs general FP64 {e32:.3e}")
    assert e64 < TOL64 and e32 < TOL32


def test_transient_full_patch_direct_and_repeatable(igap):
    """64x16x16 transient beam (three smooth history levels): the
    sum-factorised tangent agrees with the general kernel on every entry in
    both precisions, is bitwise repeatable, and its direct scatter (27 colour
    launches into the CSR values) agrees with the element-matrix path."""
    from igap4_amd import _lib
    mesh = igap.MeshIGABeam(L=(4, 1, 1), nelem=(64, 16, 16))
    U = igap.smooth_state(mesh, amp=0.02, levels=3)
    n = U.size // 3
    U[n:2 * n] *= 0.88   # distinct history levels (the rule's Ud != 0)
    U[2 * n:] *= 0.76
    ph, bc = igap.PhysGradElastTime(), igap.BCClampedTip(tip=0.05)
    with igap.Assm(mesh, ph, bc, flags=_lib.IGAP_TANGENT_ELEMENT_MATRICES) as A:
        rp, _, k1 = A.tangent(U)
        _, _, k2 = A.tangent(U)
        _, _, f1 = A.tangent(U, prec="f32")
    with igap.Assm(mesh, ph, bc, flags=_lib.IGAP_TANGENT_DIRECT) as A:
        _, _, kd = A.tangent(U)
        _, _, kd2 = A.tangent(U)
        _, _, fd = A.tangent(U, prec="f32")
    with igap.Assm(mesh, ph, bc, flags=_lib.IGAP_GENERAL_GEOMETRY) as A:
        _, _, kg = A.tangent(U)
    assert np.array_equal(k1, k2) and np.array_equal(kd, kd2)
    e64, e32 = rowfloor_err(k1, kg, rp), rowfloor_err(f1, kg, rp)
    ed, ed32 = rowfloor_err(kd, k1, rp), rowfloor_err(fd, kg, rp)
    print(f"[64x16x16 transient] sum-factorised vs general FP64 {e64:.3e}, FP32 {e32:.3e}; "
          f"direct vs element matrices {ed:.3e}, direct FP32 vs general FP64 {ed32:.3e}")
    assert e64 < TOL64 and e32 < TOL32 and ed < TOL64 and ed32 < TOL32
