"""The oracle against the reference's own outputs (tests/golden, written by
tests/golden/gen_golden.py from the compiled reference functions)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def rel_row_err(a, b):
    """max |a-b| / max(|b_ij|, ||b_i||_inf) -- entry-wise relative with the
    row norm as floor for cancellation-dominated near-zero entries."""
    a = np.atleast_2d(a)
    b = np.atleast_2d(b)
    floor = np.maximum(np.abs(b), np.max(np.abs(b), axis=-1, keepdims=True))
    floor = np.where(floor == 0, 1.0, floor)
    return float(np.max(np.abs(a - b) / floor))


def test_static_pointwise_matches_reference(oracle):
    g = _load("pointwise_static.npz")
    ph = oracle.make_phys(0, par=g["par"])
    for u, r, C in zip(g["u"], g["r"], g["C"]):
        assert rel_row_err(oracle.phys_residual(ph, u), r) < 1e-12
        assert rel_row_err(oracle.phys_tangent(ph, u), C.reshape(30, 30)) < 1e-12


def test_static_energy_matches_reference_density(oracle):
    g = _load("pointwise_static.npz")
    coef = oracle.coef_from_par(g["par"])
    for u, psi in zip(g["u"], g["psi"]):
        assert abs(oracle.static_energy(u, coef) - psi) <= 1e-12 * max(1.0, abs(psi))


def test_transient_pointwise_matches_reference(oracle):
    """gradelasttime: value rows (inertia/damping) and the truncated
    average-vector-field stress, at increments from 1e-4 to 0.3."""
    g = _load("pointwise_transient.npz")
    ph = oracle.make_phys(2, par=g["par"])
    for u, r, C in zip(g["u"], g["r"], g["C"]):
        assert rel_row_err(oracle.phys_residual(ph, u), r) < 1e-12
        assert rel_row_err(oracle.phys_tangent(ph, u), C.reshape(30, 30)) < 1e-12


def test_transient_is_not_the_full_avf_integral(oracle):
    """Planted alternative: the untruncated average-vector-field integral
    (4-point Gauss-Legendre in s, exact for the degree-7 integrand) must NOT
    match the reference at large increments -- the truncation is real."""
    g = _load("pointwise_transient.npz")
    par = g["par"]
    coef = oracle.coef_from_par(par)
    s, w = oracle.gauss4()
    u = g["u"][-1]
    u0, u1, uc = u[60:], u[30:60], u[:30]
    m0, d = 0.5 * (u0 + u1), 0.5 * (uc - u0)
    r = sum(wi * oracle.phys_residual(oracle.make_phys(0, par=par), m0 + si * d) for si, wi in zip(s, w))
    mask = np.ones(30, bool)
    mask[[0, 10, 20]] = False
    assert rel_row_err(r[mask], g["r"][-1][mask]) > 1e-8


def test_gauss_legendre_nodes(oracle):
    g = _load("basis.npz")
    x, w = oracle.gauss4()
    assert np.max(np.abs(x - g["gl_x"])) < 4e-16
    assert np.max(np.abs(w - g["gl_w"])) < 4e-16
    # closed form checks: exact for degree 7
    for k in range(8):
        assert abs(np.sum(w * x**k) - 1.0 / (k + 1)) < 1e-15


def test_eval_bspline_matches_reference(oracle):
    import ctypes as C
    g = _load("basis.npz")
    L = oracle.lib()
    meshes = {}
    for inp, out in zip(g["bs_in"], g["bs_out"]):
        nel, per, e = int(inp[0]), int(inp[1]), int(inp[2])
        key = (nel, per)
        if key not in meshes:
            h = 1.0 / nel
            kv = np.array([(i - 2) * h for i in range(nel + 5)]) if per else \
                np.array([0.0] * 3 + [(i - 2) * h for i in range(3, nel + 2)] + [1.0] * 3)
            meshes[key] = kv
        kv = meshes[key]
        kvs = [np.ascontiguousarray(kv[e:]) for _ in range(3)]
        arr = (C.POINTER(C.c_double) * 3)(*[oracle._p(k) for k in kvs])
        N = np.zeros(270)
        L.oc_eval_bspline(oracle._p(N), arr, oracle._p(np.ascontiguousarray(inp[3:6])))
        assert np.array_equal(N, out)


def test_evalN_matches_reference(oracle):
    g = _load("basis.npz")
    for inp, out in zip(g["en_in"], g["en_out"]):
        nel, per, k, i, xi = int(inp[0]), int(inp[1]), int(inp[2]), int(inp[3]), float(inp[4])
        h = 1.0 / nel
        kv = np.array([(j - 2) * h for j in range(nel + 5)]) if per else \
            np.array([0.0] * 3 + [(j - 2) * h for j in range(3, nel + 2)] + [1.0] * 3)
        v = oracle.lib().oc_evalN(oracle._p(np.ascontiguousarray(kv)), len(kv), k, i, 2, xi)
        assert v == out or abs(v - out) <= 1e-15 * max(1, abs(out))


def test_matinv_matches_reference(oracle):
    g = _load("basis.npz")
    for a, ai in zip(g["A"], g["Ainv"]):
        m = np.ascontiguousarray(a.copy())
        oracle.lib().oc_matinv3(oracle._p(m))
        assert np.allclose(m, ai, rtol=1e-14, atol=1e-15)
        assert np.allclose(m.reshape(3, 3) @ a.reshape(3, 3), np.eye(3), atol=1e-13)


def test_static_density_matches_reference(oracle):
    """All 13 gradelast_density outputs: energy, the six stress values and the
    phase indicators (the well jets exercise every indicator)."""
    g = _load("pointwise_static.npz")
    ph = oracle.make_phys(0, par=g["par"])
    for key_u, key_h in (("u", "hh"), ("u_phase", "hh_phase")):
        for u, hh in zip(g[key_u], g[key_h]):
            d = oracle.phys_density(ph, u)
            assert rel_row_err(d[:7], hh[:7]) < 1e-12
            assert np.array_equal(d[7:], hh[7:])
    assert g["hh_phase"][:, 7:].sum(axis=0).min() >= 1   # every indicator is hit


def test_transient_density_matches_reference(oracle):
    g = _load("pointwise_transient.npz")
    ph = oracle.make_phys(2, par=g["par"])
    for u, hh in zip(g["u"], g["hh"]):
        d = oracle.phys_density(ph, u)
        assert abs(d[0] - hh[0]) <= 1e-12 * max(1.0, abs(hh[0])) and d[1] == hh[1] == 1.0
