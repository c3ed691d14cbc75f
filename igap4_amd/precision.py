"""FP32 tangent against FP64 tangent in Newton's method: how far the inexact
Jacobian of the FP32 matrix-core tangent lets Newton converge at each patch
size.  The reference's solve is one SNES iteration per sub-iteration
(src/interface.c:302-328); this driver calls the device Newton one
iteration at a time the same way and records the history.

  python -m igap4_amd.precision 64,16,16 160,40,40 256,64,64
"""
import json
import sys
import time

import numpy as np

from . import assm as ig
from .loadstep import run_load_steps, tip_bc
from .warmstart import LearnedGuess


def study(nel, nsteps=40, tip=0.1, k=4, iters=6):
    """Load step k of the clamped-tip ramp (tip ``tip`` in ``nsteps`` steps)
    from the quadratic extrapolation of the converged steps before it,
    ``iters`` Newton iterations with the FP64 and with the FP32 matrix-core
    tangent (FP64 residual and update, FP32 multigrid preconditioner, GMRES
    to 1e-8): per iteration the residual over the first residual of the
    reference's iteration and the displacement error against the FP64
    answer; and the FP32 tangent's largest entry error against the FP64
    one.  A linear contraction of the FP32 arm's error is the measured
    ||I - J32^-1 J||, the condition number times the FP32 Jacobian error."""
    mesh = ig.MeshIGABeam(L=(4, 1, 1), nelem=nel)
    out = {"nelem": nel}
    with ig.Assm(mesh, ig.PhysGradElast(coef=ig.SVK_COEF), ig.BCClampedTip(tip=0.0), device=0) as A:
        n = A.ndof
        A.set_uhist(np.zeros(n))
        prev = [np.zeros(n)]
        o = ig.default_solver_opts(snes_rtol=0.0, snes_atol=1e-30, pc="mg", pc_order=1, pc_lag=4, pc_fp32=1,
                                   ksp_rtol=1e-5)
        for s in range(1, k):
            run_load_steps(A, nsteps, tip_total=tip, guess="extrapolate", opts=o, start=s, max_steps=s, prev=prev[-3:],
                           start_rtol=1e-12, callback=lambda kk, u, x: prev.append(u.copy()))
        A.set_bc(tip_bc(tip, k / nsteps))
        g = LearnedGuess(A, mode="quadratic")
        g.set_previous_load(np.zeros(n))
        U0 = g(k, prev)
        R0 = float(np.linalg.norm(A.residual(prev[-1])))
        hist = {}
        for name, tp in (("f64", "f64"), ("f32", "f32")):
            oo = ig.default_solver_opts(snes_rtol=0.0, snes_atol=0.0, snes_maxit=1, pc="mg", pc_order=1, pc_lag=1,
                                        pc_fp32=1, tangent_prec=tp, ksp_rtol=1e-8)
            A.set_uhist(U0)
            h = [float(np.linalg.norm(A.residual(U0))) / R0]
            Us = []
            for it in range(iters):
                st = A.snes_solve(oo)
                h.append(st["fnorm"] / R0)
                Us.append(A.get_uhist()[:n].copy())
            hist[name] = (h, Us)
        Uref = hist["f64"][1][-1]
        sc = float(np.max(np.abs(Uref)))
        for name in ("f64", "f32"):
            h, Us = hist[name]
            out[name] = {"rel_residual": h, "rel_disp_err": [float(np.max(np.abs(u - Uref)) / sc) for u in Us]}
        # FP32 tangent entry error against FP64 at the converged state
        A.set_uhist(Uref)
        rp, col, K64 = A.tangent(Uref, "f64")
        _, _, K32 = A.tangent(Uref, "f32")
        out["tangent_entry_err"] = float(np.max(np.abs(K32.astype(np.float64) - K64)) / np.max(np.abs(K64)))
    return out

def main():
    for nel in [tuple(int(x) for x in a.split(",")) for a in sys.argv[1:]]:
        t = time.time()
        r = study(nel)
        r["s"] = round(time.time() - t, 1)
        print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
