"""End-to-end scenarios of the Newton path on one GPU (JSON summaries).

  transient  gradelasttime on an L = 4:1:1 beam (default 128x32x32 elements),
             zero displacement and the 8-mode seeded initial velocity, time
             steps with history shift, FP32 or FP64 matrix-core tangent,
             FP64 residual and Krylov
  ramp       static SVK + strain-gradient beam, clamped at x = 0 with a tip
             displacement ramped over load steps (the load-step workloads)

  python -m igap4_amd.scenarios transient --nelem 128,32,32 --steps 20
  python -m igap4_amd.scenarios ramp --nelem 64,16,16 --steps 20 --tip 0.05
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import assm as am
from .loadstep import initial_velocity_modes, run_load_steps, run_time_steps


def _owned_X(mesh):
    na, nb = mesh.na, mesh.nb
    return mesh.X.reshape(na[0], na[1], na[2], 3)[:nb[0], :nb[1], :nb[2]].reshape(-1, 3)


def _agg(steps):
    w = [s["wall_ms"] for s in steps]
    return {"steps": len(steps), "all_converged": bool(all(s["converged"] > 0 for s in steps)),
            "newton_its": int(sum(s["newton_its"] for s in steps)), "ksp_its": int(sum(s["ksp_its"] for s in steps)),
            "wall_ms_per_step_mean": float(np.mean(w)) if w else None,
            "wall_ms_per_step_median": float(np.median(w)) if w else None,
            "t_tangent_ms": float(sum(s["t_tangent_ms"] for s in steps)),
            "t_residual_ms": float(sum(s["t_residual_ms"] for s in steps)),
            "t_pc_ms": float(sum(s["t_pc_ms"] for s in steps)), "t_ksp_ms": float(sum(s["t_ksp_ms"] for s in steps))}


def transient(nelem=(128, 32, 32), nsteps=20, prec="f32", pc="jacobi", pc_order=0, dt=1e-3, rtol=1e-10,
              guess="previous", device=0, verbose=True, pc_fp32=0, pc_lag=1):
    """gradelasttime with the reference parameters except dt; relative
    Newton tolerance rtol per step."""
    mesh = am.MeshIGABeam(L=(4, 1, 1), nelem=nelem)
    par = list(am.PhysGradElastTime().par)
    par[7] = dt
    ph = am.PhysGradElastTime(par=par)
    A = am.Assm(mesh, ph, am.BCDefault(), device=device)
    u1, u0 = initial_velocity_modes(_owned_X(mesh), (4, 1, 1), dt)
    A.set_uhist(np.concatenate([u1, u1, u0]))
    opts = am.default_solver_opts(snes_rtol=rtol, snes_atol=1e-30, tangent_prec=prec, pc=pc, pc_order=pc_order,
                                  pc_fp32=pc_fp32, pc_lag=pc_lag)

    def cb(k, H, st):
        if verbose:
            print(f"  step {k}: newton {st['newton_its']} ksp {st['ksp_its']} |F| {st['fnorm0']:.2e} -> "
                  f"{st['fnorm']:.2e}  {st['wall_ms']:.1f} ms", flush=True)

    t0 = time.perf_counter()
    steps = run_time_steps(A, nsteps, opts=opts, guess=guess, callback=cb)
    tt = time.perf_counter() - t0
    out = {"scenario": "transient", "nelem": list(nelem), "ndof": A.ndof, "nnz": A.nnz, "dt": dt, "tangent": prec,
           "pc": pc, "pc_order": pc_order, "pc_fp32": pc_fp32, "pc_lag": pc_lag, "rtol": rtol, "total_s": tt,
           **_agg(steps)}
    A.close()
    return out


def ramp(nelem=(64, 16, 16), nsteps=20, tip=0.05, prec="f64", pc="ilu0", pc_order=0, rtol=1e-10, guess="previous",
         max_steps=None, device=0, verbose=True, atol=1e-30, pc_lag=1, snes_maxit=50, mg_levels=0, pc_fp32=0,
         start_rtol=None):
    """Clamped beam with the tip displacement ramped over ``nsteps`` load
    steps; each step solved to ||R|| <= max(rtol ||R0||, atol) with R0 the
    residual of the initial iterate, or, with ``start_rtol``, to
    ||R|| <= start_rtol ||R(U_prev; load_k)|| (the previous solution under the
    step's load, whatever the initial iterate)."""
    mesh = am.MeshIGABeam(L=(4, 1, 1), nelem=nelem)
    A = am.Assm(mesh, am.PhysGradElast(coef=am.SVK_COEF), am.BCClampedTip(tip=0.0), device=device)
    A.set_uhist(np.zeros(A.ndof))
    opts = am.default_solver_opts(snes_rtol=rtol, snes_atol=atol, tangent_prec=prec, pc=pc, pc_order=pc_order,
                                  pc_lag=pc_lag, monitor=1 if verbose else 0, snes_maxit=snes_maxit,
                                  mg_levels=mg_levels, pc_fp32=pc_fp32)

    def cb(k, U, st):
        if verbose:
            print(f"  step {k}: newton {st['newton_its']} ksp {st['ksp_its']} {st['wall_ms']:.1f} ms", flush=True)

    t0 = time.perf_counter()
    steps = run_load_steps(A, nsteps, tip_total=tip, guess=guess, opts=opts, callback=cb, max_steps=max_steps,
                           start_rtol=start_rtol)
    tt = time.perf_counter() - t0
    out = {"scenario": "ramp", "nelem": list(nelem), "ndof": A.ndof, "nnz": A.nnz, "tip": tip, "load_steps": nsteps,
           "tangent": prec, "pc": pc, "pc_order": pc_order, "pc_lag": pc_lag, "pc_fp32": pc_fp32, "rtol": rtol,
           "atol": atol, "guess": guess, "start_rtol": start_rtol,
           "total_s": tt,
           **_agg(steps)}
    A.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("which", choices=["transient", "ramp"])
    ap.add_argument("--nelem", default=None)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--max-steps", type=int, default=None)
    ap.add_argument("--tip", type=float, default=0.05)
    ap.add_argument("--prec", default=None)
    ap.add_argument("--pc", default=None)
    ap.add_argument("--pc-order", type=int, default=0)
    ap.add_argument("--dt", type=float, default=1e-3)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--atol", type=float, default=1e-30)
    ap.add_argument("--pc-lag", type=int, default=1)
    ap.add_argument("--snes-maxit", type=int, default=50)
    ap.add_argument("--mg-levels", type=int, default=0)
    ap.add_argument("--pc-fp32", type=int, default=0)
    ap.add_argument("--guess", default="previous")
    ap.add_argument("--start-rtol", type=float, default=None,
                    help="ramp: stop at this fraction of the previous solution's residual under the step's load")
    a = ap.parse_args()
    if a.which == "transient":
        nel = tuple(int(x) for x in (a.nelem or "128,32,32").split(","))
        r = transient(nel, a.steps, a.prec or "f32", a.pc or "jacobi", a.pc_order, a.dt, a.rtol, a.guess,
                      pc_fp32=a.pc_fp32, pc_lag=a.pc_lag)
    else:
        nel = tuple(int(x) for x in (a.nelem or "64,16,16").split(","))
        r = ramp(nel, a.steps, a.tip, a.prec or "f64", a.pc or "ilu0", a.pc_order, a.rtol, a.guess, a.max_steps,
                 atol=a.atol, pc_lag=a.pc_lag, snes_maxit=a.snes_maxit, mg_levels=a.mg_levels, pc_fp32=a.pc_fp32,
                 start_rtol=a.start_rtol)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
