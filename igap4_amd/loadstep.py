"""Load-step and time-step drivers around the Newton solve.

IGAP4 has no load-step API: a user changes ``par_dirichlet`` (the load) or
shifts the ``U_hist`` levels (time) between ``IGAPSolvSolve`` calls, and the
next solve starts from the previous solution kept in ``solv->U``
(src/interface.c:252-259, 302-328).  These drivers do that on the device, with
linear extrapolation of the last two converged states as the alternative
initial guess.
"""
from __future__ import annotations

import time

import numpy as np

from .assm import Assm, BCClampedTip


def tip_bc(tip_total, s, tip_dof=2):
    return BCClampedTip(tip=tip_total * s, tip_dof=tip_dof)


def run_load_steps(A: Assm, nsteps, tip_total=0.05, guess="previous", opts=None, tip_dof=2,
                   history=None, callback=None):
    """Ramp the tip displacement over ``nsteps`` load steps, solving each to the
    tolerances of ``opts``.  ``guess`` = "previous" (the reference's behaviour)
    or "extrapolate" (2 U_k - U_{k-1}).  Returns a list of per-step dicts."""
    out = []
    n = A.ndof
    prev = [A.get_uhist()[:n].copy()]
    for k in range(1, nsteps + 1):
        s = k / nsteps
        A.set_bc(tip_bc(tip_total, s, tip_dof))
        if guess == "extrapolate" and len(prev) >= 2:
            U0 = 2.0 * prev[-1] - prev[-2]
        elif callable(guess):
            U0 = guess(k, prev)
        else:
            U0 = prev[-1]
        A.set_uhist(U0)
        t0 = time.perf_counter()
        st = A.snes_solve(opts)
        st["wall_ms"] = (time.perf_counter() - t0) * 1e3
        st["step"] = k
        st["load"] = s
        U = A.get_uhist()[:n].copy()
        prev.append(U)
        if history is not None:
            history.append(U)
        if callback:
            callback(k, U, st)
        out.append(st)
        if st["converged"] <= 0:
            break
    return out
