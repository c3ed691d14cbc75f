"""Device Newton-Krylov solve (SNES NEWTONLS / GMRES / ILU(0) on the GPU):
converges to the reference tolerance, reaches the same displacements as a
CPU Newton on the oracle's residual and tangent, and the FP32 matrix-core
tangent reaches the FP64 answer."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

pytestmark = pytest.mark.gpu

NEL = (8, 2, 2)
TIP = 0.02


def cpu_newton(oracle, mesh, ph, bc, U0, tol=1e-13, maxit=30):
    """Newton on the oracle (exact sparse LU solves): the reference answer."""
    U = U0.copy()
    pat = oracle.csr_pattern(mesh)
    R = oracle.residual(mesh, ph, bc, U)
    r0 = np.linalg.norm(R)
    for _ in range(maxit):
        if np.linalg.norm(R) <= tol * max(r0, 1e-300):
            break
        rp, col, val = oracle.tangent(mesh, ph, bc, U, pattern=pat)
        K = sp.csr_matrix((val, col, rp), shape=(mesh.ndof, mesh.ndof))
        U = U - spla.spsolve(K.tocsc(), R)
        R = oracle.residual(mesh, ph, bc, U)
    return U, np.linalg.norm(R) / r0


@pytest.fixture(scope="module")
def problem(igap, oracle):
    mg = igap.MeshIGABeam(L=(4, 1, 1), nelem=NEL)
    mo = oracle.Mesh(NEL, L=(4, 1, 1))
    return mg, mo


@pytest.mark.parametrize("prec,order", [("f64", 0), ("f32", 0), ("f64", 1), ("f32", 1)])
def test_newton_matches_cpu_newton(igap, oracle, problem, prec, order):
    """order 0: ILU(0) in the natural dof order; 1: in the 27-colour order
    (the preconditioner changes, the answer must not)."""
    mg, mo = problem
    ph_o = oracle.make_phys(0, coef=oracle.SVK_COEF)
    bc_o = oracle.clamped_tip_bc(tip=TIP)
    U_ref, rel = cpu_newton(oracle, mo, ph_o, bc_o, np.zeros(mo.ndof))
    assert rel < 1e-12
    with igap.Assm(mg, igap.PhysGradElast(coef=igap.SVK_COEF), igap.BCClampedTip(tip=TIP)) as A:
        A.set_uhist(np.zeros(A.ndof))
        st = A.snes_solve(snes_rtol=1e-10, snes_atol=1e-30, tangent_prec=prec, pc_order=order)
        U = A.get_uhist()
    print(f"[newton {prec} order={order}] its={st['newton_its']} ksp={st['ksp_its']} "
          f"|F|/|F0|={st['fnorm'] / st['fnorm0']:.2e}")
    assert st["converged"] > 0 and st["fnorm"] <= 1e-10 * st["fnorm0"]
    err = np.max(np.abs(U - U_ref)) / np.max(np.abs(U_ref))
    print(f"  displacement rel err vs CPU Newton {err:.2e}")
    assert err < 1e-8


def test_load_steps_reference_criterion(igap):
    """16x4x4 beam, tip 0.05 in 10 load steps, each load step
    to the reference's absolute criterion ||R|| < 1e-10."""
    from igap4_amd.loadstep import run_load_steps
    m = igap.MeshIGABeam(L=(4, 1, 1), nelem=(16, 4, 4))
    with igap.Assm(m, igap.PhysGradElast(coef=igap.SVK_COEF), igap.BCClampedTip(tip=0.0)) as A:
        A.set_uhist(np.zeros(A.ndof))
        opts = igap.assm.default_solver_opts(snes_atol=1e-10, snes_rtol=0.0)
        steps = run_load_steps(A, 10, tip_total=0.05, opts=opts)
        U = A.get_uhist()
    for s in steps:
        print(f"  step {s['step']}: its={s['newton_its']} ksp={s['ksp_its']} |F|={s['fnorm']:.2e} {s['wall_ms']:.1f} ms")
    assert len(steps) == 10 and all(s["converged"] > 0 for s in steps)
    # the tip face carries the full displacement
    nb = m.nb
    Uz = U.reshape(nb[0], nb[1], nb[2], 3)[..., 2]
    assert np.allclose(Uz[-1], 0.05 * (1 - 1e-15), atol=1e-12)


@pytest.mark.parametrize("case", ["beam_f64", "beam_f32", "periodic6_transient", "mixed_periodic6x4x9", "periodic5_fallback"])
def test_own_ilu_matches_rocsparse(igap, oracle, case):
    """The solver's point-block ILU(0) in the 27-colour order equals
    rocSPARSE csrilu0 of the same matrix up to roundoff (FP64 1e-12, FP32
    1e-4 of the largest factor entry), on a clamped beam, a periodic cube
    (periodic offsets taken modulo the grid) and a patch periodic in two
    directions; a periodic direction of 5 points has no colour structure and
    reports the rocSPARSE fallback."""
    prec = "f32" if case == "beam_f32" else "f64"
    if case.startswith("beam"):
        m = igap.MeshIGABeam(L=(4, 1, 1), nelem=(16, 4, 4))
        A = igap.Assm(m, igap.PhysGradElast(coef=igap.SVK_COEF), igap.BCClampedTip(tip=0.02))
        U = igap.smooth_state(m, amp=0.02)
    else:
        nel, per = {"periodic6_transient": ((6, 6, 6), (1, 1, 1)), "mixed_periodic6x4x9": ((6, 4, 9), (1, 0, 1)),
                     "periodic5_fallback": ((5, 5, 5), (1, 1, 1))}[case]
        m = igap.MeshIGABeam(L=(1, 1, 1), nelem=nel, periodic=per)
        A = igap.Assm(m, igap.PhysGradElastTime(), igap.BCDefault())
        U = oracle.glibc_rand_state(3 * A.ndof)
    with A:
        A.tangent(U, prec=prec)
        ok, diff = A.selftest_ilu(prec)
    print(f"  {case}: applicable {ok}, max|own - rocsparse| / max|rocsparse| = {diff:.2e}")
    if case == "periodic5_fallback":
        assert not ok
        return
    assert ok
    assert diff < (1e-4 if prec == "f32" else 1e-12)


@pytest.mark.parametrize("prec", ["f64", "f32", "mixed", "mixed16"])
def test_multigrid_pc_same_answer_fewer_krylov(igap, prec):
    """Geometric multigrid preconditioner (32x8x8 -> 16x4x4 levels,
    Chebyshev smoothing, dense coarsest solve): Newton reaches the
    ILU(0)-preconditioned answer with fewer Krylov iterations, to 1e-9 of the
    displacement scale with an FP64 tangent (also with the preconditioner in
    FP32, "mixed", and with FP32 levels smoothing with 16-bit matrices,
    "mixed16") and 1e-7 with the FP32 tangent (both solves stop at
    ||R|| ~ 1e-12, whose displacement error the soft bending modes amplify)."""
    m = igap.MeshIGABeam(L=(4, 1, 1), nelem=(32, 8, 8))
    out = {}
    tp = "f32" if prec == "f32" else "f64"
    for pc in ("ilu0", "mg"):
        with igap.Assm(m, igap.PhysGradElast(coef=igap.SVK_COEF), igap.BCClampedTip(tip=0.02)) as A:
            A.set_uhist(np.zeros(A.ndof))
            st = A.snes_solve(snes_rtol=1e-11, snes_atol=1e-30, tangent_prec=tp, pc=pc, pc_order=1,
                              pc_fp32={"mixed": 1, "mixed16": 2}.get(prec, 0))
            out[pc] = (A.get_uhist()[:A.ndof].copy(), st)
        print(f"  {pc}: converged {out[pc][1]['converged']} newton {out[pc][1]['newton_its']} "
              f"ksp {out[pc][1]['ksp_its']} |F| {out[pc][1]['fnorm']:.2e}")
    assert out["ilu0"][1]["converged"] > 0 and out["mg"][1]["converged"] > 0
    u0, u1 = out["ilu0"][0], out["mg"][0]
    assert np.max(np.abs(u1 - u0)) <= (1e-7 if prec == "f32" else 1e-9) * np.max(np.abs(u0))
    assert out["mg"][1]["ksp_its"] < out["ilu0"][1]["ksp_its"]


def test_multigrid_pc_rejects_periodic(igap):
    m = igap.MeshIGABeam(L=(1, 1, 1), nelem=(6, 6, 6), periodic=(1, 0, 0))
    with igap.Assm(m, igap.PhysGradElast(coef=igap.SVK_COEF), igap.BCDefault()) as A:
        A.set_uhist(1e-3 * np.random.default_rng(0).standard_normal(A.ndof))
        with pytest.raises(Exception, match="non-periodic"):
            A.snes_solve(pc="mg", pc_order=1)


def test_coloured_fallback_and_fp32_preconditioner(igap, oracle):
    """Newton on a periodic 5x5x5 transient patch (colours coupled through the
    wrap, so the coloured ILU(0) takes rocSPARSE's level-scheduled solves) and
    with an FP32 ILU(0) factor of the FP64 tangent: both reach the
    natural-order FP64 answer to 1e-10 of the displacement scale."""
    m = igap.MeshIGABeam(L=(1, 1, 1), nelem=(5, 5, 5), periodic=(1, 1, 1))
    out = {}
    for name, kw in (("natural", dict(pc_order=0)), ("coloured_fallback", dict(pc_order=1)),
                     ("fp32_pc", dict(pc_order=1, pc_fp32=1))):
        with igap.Assm(m, igap.PhysGradElastTime(), igap.BCDefault()) as A:
            A.set_uhist(oracle.glibc_rand_state(3 * A.ndof))
            st = A.snes_solve(snes_rtol=1e-12, snes_atol=1e-30, **kw)
            out[name] = (A.get_uhist()[:A.ndof].copy(), st)
        print(f"  {name}: converged {st['converged']} newton {st['newton_its']} ksp {st['ksp_its']} |F| {st['fnorm']:.2e}")
        assert st["converged"] > 0
    u0 = out["natural"][0]
    for name in ("coloured_fallback", "fp32_pc"):
        assert np.max(np.abs(out[name][0] - u0)) <= 1e-10 * np.max(np.abs(u0))


@pytest.mark.parametrize("case", ["single_level_5x3x3", "transient_16x4x4"])
def test_multigrid_pc_other_hierarchies(igap, oracle, case):
    """Multigrid where no direction can be halved (one level: the cycle is
    Chebyshev/block-Jacobi smoothing) and on the transient physics (three
    history levels restricted to the coarse levels): Newton reaches the
    ILU(0)-preconditioned answer to 1e-10 of the displacement scale."""
    if case == "single_level_5x3x3":
        m = igap.MeshIGABeam(L=(4, 1, 1), nelem=(5, 3, 3))
        make = lambda: igap.Assm(m, igap.PhysGradElast(coef=igap.SVK_COEF), igap.BCClampedTip(tip=0.02))
        U0 = lambda A: np.zeros(A.ndof)
    else:
        m = igap.MeshIGABeam(L=(4, 1, 1), nelem=(16, 4, 4))
        make = lambda: igap.Assm(m, igap.PhysGradElastTime(), igap.BCDefault())
        U0 = lambda A: oracle.glibc_rand_state(3 * A.ndof)
    out = {}
    for pc in ("ilu0", "mg"):
        with make() as A:
            A.set_uhist(U0(A))
            st = A.snes_solve(snes_rtol=1e-12, snes_atol=1e-30, pc=pc, pc_order=1)
            out[pc] = (A.get_uhist()[:A.ndof].copy(), st)
        print(f"  {case} {pc}: converged {st['converged']} newton {st['newton_its']} ksp {st['ksp_its']} |F| {st['fnorm']:.2e}")
        assert st["converged"] > 0
    u0, u1 = out["ilu0"][0], out["mg"][0]
    assert np.max(np.abs(u1 - u0)) <= 1e-10 * np.max(np.abs(u0))
