"""Writes tests/golden/*.npz from the compiled reference pieces.

Run in the container that has the reference tree (it is not needed, and not
present, where the tests run):

    make -C oracle ref && python tests/golden/gen_golden.py

Inputs are seeded random jets and points; outputs are what the reference's own
functions return for them:
  pointwise_static.npz     gradelast_residual / gradelast_tangent /
                           gradelast_density  (phys/gradelast.c:7-1085)
  pointwise_transient.npz  gradelasttime_residual / gradelasttime_tangent /
                           gradelasttime_density (phys/gradelasttime.c:7-9232)
  basis.npz                eval_Bspline (src/meshcore_iga.c:243-327), evalN
                           (:329-384), legendre_handle n=4 (src/quadrature.c:
                           381-429), matinv 3x3 (src/mathutil.c:17-64)
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
from oracle import oracle as O  # noqa: E402

PAR_STATIC = O.REF_PAR_STATIC
PAR_TRANSIENT = O.REF_PAR_TRANSIENT


def _call(fn, n, u, par):
    out = np.zeros(n)
    u = np.ascontiguousarray(u, dtype=np.float64)
    fn(O._p(out), O._p(u), O._p(np.ascontiguousarray(par)))
    return out


def main():
    R = O.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libigapref.so is not built: make -C oracle ref")
    rng = np.random.default_rng(20261202)

    # ---- F1 static: 24 jets, gradient scale 0.05 .. 0.5
    nS = 24
    scales = np.geomspace(0.05, 0.5, nS)
    U = np.stack([rng.normal(scale=s, size=30) for s in scales])
    r = np.stack([_call(R.gradelast_residual, 30, u, PAR_STATIC) for u in U])
    Cm = np.stack([_call(R.gradelast_tangent, 900, u, PAR_STATIC) for u in U])
    hh = np.stack([_call(R.gradelast_density, 13, u, PAR_STATIC) for u in U])
    psi = hh[:, 0].copy()
    # jets in the energy's wells: symmetric stretch F = sqrt(I + 2E) with a
    # deviatoric E on circles in the (e25, e26) plane, small strain gradients
    UP = []
    rngP = np.random.default_rng(777)   # own stream: the other fixtures keep their jets
    for rad in (0.1, 0.2, 0.3):
        for th in np.linspace(0.0, 2 * np.pi, 8, endpoint=False):
            e25, e26 = rad * np.cos(th), rad * np.sin(th)
            d = np.array([e25 / np.sqrt(2) + e26 / np.sqrt(6), -e25 / np.sqrt(2) + e26 / np.sqrt(6), -2 * e26 / np.sqrt(6)])
            F = np.diag(np.sqrt(1 + 2 * d))
            u = rngP.normal(scale=0.01, size=30)
            for k in range(3):
                for a in range(3):
                    u[10 * k + 1 + a] = F[k, a] - (1.0 if k == a else 0.0)
            UP.append(u)
    UP = np.stack(UP)
    hhP = np.stack([_call(R.gradelast_density, 13, u, PAR_STATIC) for u in UP])
    np.savez_compressed(os.path.join(HERE, "pointwise_static.npz"), u=U, par=PAR_STATIC, r=r, C=Cm, psi=psi, hh=hh,
                        u_phase=UP, hh_phase=hhP)

    # ---- F1 transient: 24 (u, u1, u0) triples with increments 1e-4 .. 0.3
    nT = 24
    inc = np.geomspace(1e-4, 0.3, nT)
    UT = []
    for s in inc:
        base = rng.normal(scale=0.2, size=30)
        UT.append(np.concatenate([base + rng.normal(scale=s, size=30), base + rng.normal(scale=s, size=30), base]))
    UT = np.stack(UT)
    rT = np.stack([_call(R.gradelasttime_residual, 30, u, PAR_TRANSIENT) for u in UT])
    CT = np.stack([_call(R.gradelasttime_tangent, 900, u, PAR_TRANSIENT) for u in UT])
    hT = np.stack([_call(R.gradelasttime_density, 2, u, PAR_TRANSIENT) for u in UT])
    np.savez_compressed(os.path.join(HERE, "pointwise_transient.npz"), u=UT, par=PAR_TRANSIENT, r=rT, C=CT, hh=hT)

    # ---- F2 basis: open and periodic knot vectors, several spans, the 4 Gauss points
    x = np.zeros(4)
    w = np.zeros(4)
    R.legendre_handle(O._p(x), O._p(w), 4, C.c_double(0.0), C.c_double(1.0))
    cases = []
    for nel, per in ((5, 0), (3, 0), (6, 1), (1, 0)):
        h = 1.0 / nel
        kv = np.array([(i - 2) * h for i in range(nel + 5)]) if per else \
            np.array([0.0] * 3 + [(i - 2) * h for i in range(3, nel + 2)] + [1.0] * 3)
        cases.append((nel, per, kv))
    bs_in, bs_out = [], []
    for nel, per, kv in cases:
        for e in range(nel):
            kvs = [np.ascontiguousarray(kv[e:]) for _ in range(3)]
            arr = (C.POINTER(C.c_double) * 3)(*[O._p(k) for k in kvs])
            for q in range(4):
                for q2 in range(4):
                    xq = np.array([(1 - x[q]) * kv[e + 2] + x[q] * kv[e + 3],
                                   (1 - x[q2]) * kv[e + 2] + x[q2] * kv[e + 3],
                                   (1 - x[(q + q2) % 4]) * kv[e + 2] + x[(q + q2) % 4] * kv[e + 3]])
                    N = np.zeros(270)
                    R.eval_Bspline(O._p(N), arr, 2, O._p(xq))
                    bs_in.append(np.concatenate([[nel, per, e], xq]))
                    bs_out.append(N)
    en_in, en_out = [], []
    for nel, per, kv in cases:
        nk = len(kv)
        kvc = np.ascontiguousarray(kv)
        for k in range(3):
            for i in range(nk - 3):
                for xi in (0.0, 0.1234, 0.5, 0.77, 1.0 - 1e-15, 1.0):
                    en_in.append([nel, per, k, i, xi])
                    en_out.append(R.evalN(O._p(kvc), nk, k, i, 2, xi))
    A = rng.normal(size=(16, 9)) + 3 * np.eye(3).reshape(9)
    Ainv = []
    for a in A:
        m = np.ascontiguousarray(a.copy())
        R.matinv(O._p(m), 3)
        Ainv.append(m)
    np.savez_compressed(os.path.join(HERE, "basis.npz"), gl_x=x, gl_w=w, bs_in=np.array(bs_in), bs_out=np.array(bs_out),
                        en_in=np.array(en_in), en_out=np.array(en_out), A=A, Ainv=np.array(Ainv))
    for f in ("pointwise_static.npz", "pointwise_transient.npz", "basis.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
