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
                   history=None, callback=None, max_steps=None, start=1, prev=None, start_rtol=None,
                   ksp_atol_factor=None):
    """Ramp the tip displacement over ``nsteps`` load steps, solving each to the
    tolerances of ``opts``.  ``guess`` = "previous" (the reference's behaviour),
    "extrapolate" (2 U_k - U_{k-1}) or a callable ``guess(k, prev)`` returning
    the initial iterate from the list of previous solutions.  ``start`` and
    ``prev`` (solutions up to step start-1, last one current) resume a ramp;
    ``max_steps`` stops after that step.

    ``start_rtol``: solve every step to ||R|| <= start_rtol ||R0|| with R0 the
    residual of the previous solution under the step's load (the first
    residual of the reference's iteration), whatever the initial iterate, and
    with ``ksp_atol_factor`` stop each Krylov solve at that fraction of the
    threshold.  ``wall_ms`` covers the guess (and R0) and the solve.  Returns
    a list of per-step dicts."""
    out = []
    n = A.ndof
    prev = [A.get_uhist()[:n].copy()] if prev is None else [np.asarray(u)[:n].copy() for u in prev]
    last = nsteps if max_steps is None else min(nsteps, max_steps)
    for k in range(start, last + 1):
        s = k / nsteps
        A.set_bc(tip_bc(tip_total, s, tip_dof))
        t0 = time.perf_counter()
        o = opts
        if start_rtol is not None:
            o = type(opts).from_buffer_copy(opts)
            r0 = float(np.linalg.norm(A.residual(prev[-1])))
            o.snes_atol = start_rtol * r0
            o.snes_rtol = 0.0
            if ksp_atol_factor is not None:
                o.ksp_atol = ksp_atol_factor * o.snes_atol
        if guess == "extrapolate" and len(prev) >= 2:
            U0 = 2.0 * prev[-1] - prev[-2]
        elif callable(guess):
            U0 = guess(k, prev)
        else:
            U0 = prev[-1]
        A.set_uhist(U0)
        st = A.snes_solve(o)
        st["wall_ms"] = (time.perf_counter() - t0) * 1e3
        st["step"] = k
        st["load"] = s
        if start_rtol is not None:
            st["r0_start"] = r0
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


def run_time_steps(A: Assm, nsteps, opts=None, guess="previous", callback=None, start_rtol=None,
                   ksp_atol_factor=None):
    """Transient steps of gradelasttime: before each step the history levels
    shift (two back <- previous <- last converged) and level 0 gets the initial
    iterate ("previous": the last converged state, the reference's behaviour;
    "extrapolate": 2 U_n - U_{n-1}; or a callable ``guess(k, prev)`` on the
    list of converged states, last one current).  U_hist must hold
    [U_n, U_n, U_{n-1}] or any state whose levels 1 and 2 are the last two time
    levels.  ``start_rtol`` / ``ksp_atol_factor`` as in run_load_steps (R0 =
    the residual of the previous state, the reference's first residual);
    ``wall_ms`` covers the guess and the solve.  Returns a list of per-step
    dicts."""
    n = A.ndof
    H = A.get_uhist().reshape(-1, n)
    if H.shape[0] < 3:
        raise ValueError("run_time_steps needs a transient context (three history levels)")
    prev = [H[2].copy(), H[1].copy()]
    out = []
    for k in range(1, nsteps + 1):
        u1, u0 = prev[-1], prev[-2]
        t0 = time.perf_counter()
        o = opts
        if start_rtol is not None:
            o = type(opts).from_buffer_copy(opts)
            r0 = float(np.linalg.norm(A.residual(np.concatenate([u1, u1, u0]))))
            o.snes_atol = start_rtol * r0
            o.snes_rtol = 0.0
            if ksp_atol_factor is not None:
                o.ksp_atol = ksp_atol_factor * o.snes_atol
        if callable(guess):
            U = guess(k, prev)
        elif guess == "previous":
            U = u1.copy()
        else:
            U = 2.0 * u1 - u0
        A.set_uhist(np.concatenate([U, u1, u0]))
        st = A.snes_solve(o)
        st["wall_ms"] = (time.perf_counter() - t0) * 1e3
        st["step"] = k
        if start_rtol is not None:
            st["r0_start"] = r0
        H = A.get_uhist().reshape(-1, n)
        prev.append(H[0].copy())
        if callback:
            callback(k, H, st)
        out.append(st)
        if st["converged"] <= 0:
            break
    return out


def initial_velocity_modes(X, L, dt, amp=0.01, nmodes=8, seed=20261202):
    """The transient workload's start: zero displacement and the velocity
    v0 = amp * sum_m d_m sin(pi k_m . X / L + phi_m) over ``nmodes`` modes with
    integer wave vectors k_m in {1,2}^3, random unit directions d_m and phases
    phi_m (numpy default_rng(seed)); returns the history levels (u1, u0) with
    u1 = 0 and u0 = u1 - dt v0 at the control points X[ncp][3]."""
    rng = np.random.default_rng(seed)
    X = np.asarray(X).reshape(-1, 3)
    L = np.asarray(L, dtype=np.float64)
    v = np.zeros_like(X)
    for _ in range(nmodes):
        kv = rng.integers(1, 3, size=3)
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        phi = rng.uniform(0.0, 2.0 * np.pi)
        v += np.outer(np.sin(np.pi * (X / L) @ kv + phi), d)
    v *= amp
    u1 = np.zeros(X.size)
    u0 = u1 - dt * v.reshape(-1)
    return u1, u0
