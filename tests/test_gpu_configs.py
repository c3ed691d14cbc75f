"""The workloads at their own sizes.

* 256x64x64 static beam (1,048,576 elements, 3,371,544 dofs, 1.21G
  nonzeros; the largest single-GPU case): the residual against the oracle on
  the rows of the tip-end control points (touched only by the last element
  layers, which the oracle assembles), the complete tip rows of the FP64
  tangent likewise, K v against the central difference of the residual,
  K = K^T, and the FP32 matrix-core tangent within its bound of the FP64 one.
* 128x32x32 gradelasttime with the seeded 8-mode initial velocity: three
  time steps through run_time_steps, each solved to ||R|| <= 1e-10 ||R0||,
  the history shift of U_hist (src/assmcore_iga.c:66-80, src/interface.c:127)
  and the transient tangent's last-layer rows against the oracle; and three
  time steps on a 16x4x4 beam whose first step equals a CPU Newton solve on
  the oracle.
* the learned warm start on the 130x34x34 control-point grid (128x32x32
  elements), shortened: every arm reaches the plain FP64 answer to 1e-8.
"""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

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


def tail_rows(A, r0, prec="f64"):
    """rowptr (from 0), col and values of the device CSR rows r0..ndof-1."""
    from igap4_amd import _lib
    L = _lib.lib()
    rp = A.dtoh(L.igap_gpu_csr_rowptr(A.ctx), A.ndof + 1, np.int64)
    p0, p1 = int(rp[r0]), int(rp[-1])
    col = A.dtoh(C.c_void_p(L.igap_gpu_csr_col(A.ctx) + 4 * p0), p1 - p0, np.int32)
    if prec == "f64":
        val = A.dtoh(C.c_void_p(L.igap_gpu_csr_val64(A.ctx) + 8 * p0), p1 - p0)
    else:
        val = A.dtoh(C.c_void_p(L.igap_gpu_csr_val32(A.ctx) + 4 * p0), p1 - p0, np.float32)
    return rp[r0:] - p0, col, val


# ---------------------------------------------------------------- 256x64x64
NEL5 = (256, 64, 64)


@pytest.fixture(scope="module")
def c5(igap, oracle):
    mesh = igap.MeshIGABeam(L=(4.0, 1.0, 1.0), nelem=NEL5)
    mo = oracle.Mesh(NEL5, L=(4, 1, 1))
    ph_o, bc_o = oracle.make_phys(0, coef=oracle.SVK_COEF), oracle.clamped_tip_bc(tip=0.05)
    rng = np.random.default_rng(5)
    U = igap.smooth_state(mesh, amp=0.01) + 1e-3 * rng.standard_normal(mo.ndof)
    nb = mo.nb
    r_res = 3 * (nb[0] - 4) * nb[1] * nb[2]   # control points x >= 254: elements x >= 252 only
    r_tan = 3 * (nb[0] - 2) * nb[1] * nb[2]   # control points x >= 256: elements x >= 254 only
    print(f"[256x64x64] ndof {mo.ndof}", flush=True)
    with igap.Assm(mesh, igap.PhysGradElast(coef=igap.SVK_COEF), igap.BCClampedTip(tip=0.05), device=0) as A:
        R = A.residual(U)
        A.tangent_dev(None, None, 0)
        K = tail_rows(A, r_tan, "f64")
        A.tangent_dev(None, None, 1)
        K32 = tail_rows(A, r_tan, "f32")
        mask = A.dirichlet()[0]
        v = rng.standard_normal(mo.ndof) * (mask == 0)
        eps = 1e-6
        dR = (A.residual(U + eps * v) - A.residual(U - eps * v)) / (2 * eps)
    print("[256x64x64] device done", flush=True)
    return dict(mo=mo, ph_o=ph_o, bc_o=bc_o, U=U, R=R, K=K, K32=K32, v=v, dR=dR, r_res=r_res, r_tan=r_tan)


def test_256_residual_tip_rows_match_oracle(c5, oracle):
    d = c5
    Ro = oracle.residual(d["mo"], d["ph_o"], d["bc_o"], d["U"], box=[252, 256, 0, 64, 0, 64], nthreads=NTHREADS)
    r0 = d["r_res"]
    err = rowfloor_err(d["R"][r0:], Ro[r0:])
    print(f"residual rows of x >= 254: {err:.1e}")
    assert err <= 1e-12


def test_256_tangent_tip_rows_match_oracle(c5, oracle):
    d = c5
    r0 = d["r_tan"]
    pat = oracle.csr_pattern(d["mo"], row0=r0)
    rp, col, Ko = oracle.tangent(d["mo"], d["ph_o"], d["bc_o"], d["U"], pattern=pat, box=[254, 256, 0, 64, 0, 64],
                                 nthreads=NTHREADS, row0=r0)
    grp, gcol, K = d["K"]
    assert np.array_equal(grp, rp) and np.array_equal(gcol, col)
    err = rowfloor_err(K, Ko, rp)
    print(f"tangent rows of x >= 256 ({len(rp) - 1} rows, {len(col)} entries): {err:.1e}")
    assert err <= 1e-12


def test_256_tangent_is_residual_derivative_and_symmetric(c5):
    import scipy.sparse as sp
    d = c5
    r0 = d["r_tan"]
    rp, col, K = d["K"]
    n = len(d["U"])
    Kt = sp.csr_matrix((K, col, rp), shape=(len(rp) - 1, n))
    Kv = Kt @ d["v"]
    err = np.max(np.abs(Kv - d["dR"][r0:])) / np.max(np.abs(Kv))
    S = Kt[:, r0:]
    asym = abs(S - S.T).max() / abs(S).max()
    print(f"K v vs central difference (tip rows): {err:.1e}; asymmetry {asym:.1e}")
    assert err < 1e-6
    assert asym <= 1e-14


def test_256_fp32_tangent_within_bound(c5):
    rp, col, K = c5["K"]
    rp32, col32, K32 = c5["K32"]
    assert np.array_equal(col, col32)
    err = np.max(np.abs(K32.astype(np.float64) - K)) / np.max(np.abs(K))
    print(f"FP32 tangent vs FP64, tip rows: max |dK| / max |K| = {err:.2e}")
    assert err < 2e-5


# ---------------------------------------------------------------- time steps
def _transient(igap, nelem, dt=1e-3):
    from igap4_amd.loadstep import initial_velocity_modes
    mesh = igap.MeshIGABeam(L=(4, 1, 1), nelem=nelem)
    par = list(igap.PhysGradElastTime().par)
    par[7] = dt
    A = igap.Assm(mesh, igap.PhysGradElastTime(par=par), igap.BCDefault(), device=0)
    na, nb = mesh.na, mesh.nb
    Xo = mesh.X.reshape(na[0], na[1], na[2], 3)[:nb[0], :nb[1], :nb[2]].reshape(-1, 3)
    u1, u0 = initial_velocity_modes(Xo, (4, 1, 1), dt)
    A.set_uhist(np.concatenate([u1, u1, u0]))
    return mesh, par, A, u1, u0


def _run_steps_checked(igap, A, nsteps, opts):
    from igap4_amd.loadstep import run_time_steps
    n = A.ndof
    hist = [A.get_uhist().reshape(-1, n).copy()]
    st = run_time_steps(A, nsteps, opts=opts, start_rtol=1e-10,
                        callback=lambda k, H, x: hist.append(np.array(H, copy=True)))
    assert len(st) == nsteps
    for k, x in enumerate(st, 1):
        print(f"  step {k}: newton {x['newton_its']} ksp {x['ksp_its']} |R| {x['r0_start']:.2e} -> {x['fnorm']:.2e} "
              f"({x['wall_ms']:.0f} ms)")
        assert x["converged"] > 0
        assert x["fnorm"] <= 1e-10 * x["r0_start"]
    # history shift: before step k the levels are [guess, U_{k-1}, U_{k-2}]
    for k in range(2, nsteps + 1):
        assert np.array_equal(hist[k][1], hist[k - 1][0])
        assert np.array_equal(hist[k][2], hist[k - 1][1])
    assert np.array_equal(hist[1][1], hist[0][1]) and np.array_equal(hist[1][2], hist[0][2])
    return st, hist


def test_time_steps_small_beam_match_cpu_newton(igap, oracle):
    """Three time steps on a 16x4x4 gradelasttime beam: each converges, the
    history shifts, and the first step's solution equals a Newton solve of
    the oracle's residual and tangent on the CPU."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    mesh, par, A, u1, u0 = _transient(igap, (16, 4, 4))
    opts = igap.assm.default_solver_opts(tangent_prec="f64", pc="ilu0", pc_order=1)
    with A:
        st, hist = _run_steps_checked(igap, A, 3, opts)
    mo = oracle.Mesh((16, 4, 4), L=(4, 1, 1))
    ph = oracle.make_phys(2, par=par)
    bc = oracle.make_bc()
    U = u1.copy()
    R0 = np.linalg.norm(oracle.residual(mo, ph, bc, np.concatenate([U, u1, u0])))
    for _ in range(20):
        H = np.concatenate([U, u1, u0])
        R = oracle.residual(mo, ph, bc, H)
        if np.linalg.norm(R) <= 1e-13 * R0:
            break
        rp, col, K = oracle.tangent(mo, ph, bc, H)
        U = U - spl.spsolve(sp.csr_matrix((K, col, rp), shape=(len(U), len(U))).tocsc(), R)
    err = np.max(np.abs(hist[1][0] - U)) / np.max(np.abs(U))
    print(f"step 1 vs CPU Newton on the oracle: {err:.1e}")
    assert err <= 1e-8


def test_time_steps_128x32x32(igap, oracle):
    """The transient workload at its size: 128x32x32 elements, dt 1e-3, the
    8-mode initial velocity; FP32 matrix-core tangent with point-Jacobi
    GMRES (FP64 residual, Newton update and Krylov vectors)."""
    mesh, par, A, u1, u0 = _transient(igap, (128, 32, 32))
    opts = igap.assm.default_solver_opts(tangent_prec="f32", pc="jacobi", pc_order=1)
    nel = (128, 32, 32)
    with A:
        st, hist = _run_steps_checked(igap, A, 3, opts)
        H = hist[-1].reshape(-1)
        A.set_uhist(H)
        A.tangent_dev(None, None, 0)
        nb = mesh.nb
        r0 = 3 * (nb[0] - 1) * nb[1] * nb[2]   # the last control-point layer: element layer 127 only
        K = tail_rows(A, r0, "f64")
    mo = oracle.Mesh(nel, L=(4, 1, 1))
    pat = oracle.csr_pattern(mo, row0=r0)
    rp, col, Ko = oracle.tangent(mo, oracle.make_phys(2, par=par), oracle.make_bc(), H, pattern=pat,
                                 box=[127, 128, 0, 32, 0, 32], nthreads=NTHREADS, row0=r0)
    assert np.array_equal(K[1], col)
    err = rowfloor_err(K[2], Ko, rp)
    print(f"transient tangent, last control-point layer after step 3: {err:.1e}")
    assert err <= 1e-12


# ---------------------------------------------------------------- warm start grid
def test_warm_start_on_the_130x34x34_grid(igap):
    """Shortened warm-start run on 128x32x32 elements (the 130x34x34
    control-point grid): 12 training steps of the ramp to tip 0.5 in 200
    steps, then 3 steps per arm from the same state; every arm reaches the
    plain FP64 answer to 1e-8."""
    from igap4_amd import warmstart as ws
    s = ws.run(nelem=(128, 32, 32), nsteps=200, ntrain=12, tip_total=0.5, epochs=10, lr=2e-3, verbose=False,
               pc="mg", prec="mixed", neval=3, coarse_filter=2)
    for arm in ("plain", "linear", "quadratic", "learned"):
        r = s[arm]
        print(f"{arm}: {r['wall_ms_per_step']:.0f} ms/step, newton {r['newton_its']}, ksp {r['ksp_its']}, "
              f"diff {r['final_rel_diff_vs_plain']:.1e}, speedup {r.get('speedup_vs_plain', 0):.2f}")
        assert r["all_converged"] and r["steps"] == 3
        assert r["final_rel_diff_vs_plain"] <= 1e-8
    assert s["grid"] == [130, 34, 34]
