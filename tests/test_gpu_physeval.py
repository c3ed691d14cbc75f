"""The device pointwise physics (igap_gpu_phys_eval: the Phys callbacks
residual / tangent / density, include/physimpl.h:6-9) against the compiled
reference's own outputs at the committed fixtures (tests/golden, written by
tests/golden/gen_golden.py from phys/gradelast.c and phys/gradelasttime.c)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def rel_row_err(a, b):
    """max |a-b| / max(|b_ij|, ||b_i||_inf) over the rows of b."""
    a = np.atleast_2d(a)
    b = np.atleast_2d(b)
    floor = np.maximum(np.abs(b), np.max(np.abs(b), axis=-1, keepdims=True))
    floor = np.where(floor == 0, 1.0, floor)
    return float(np.max(np.abs(a - b) / floor))


def _ctx(igap, phys):
    mesh = igap.MeshIGABeam(L=(1, 1, 1), nelem=(2, 2, 2))
    return igap.Assm(mesh, phys, igap.BCDefault(), device=0)


def test_static_physics_on_device_matches_reference(igap):
    g = _load("pointwise_static.npz")
    with _ctx(igap, igap.PhysGradElast(par=g["par"])) as A:
        r, Cm, hh = A.phys_eval(g["u"])
        _, _, hp = A.phys_eval(g["u_phase"], residual=False, tangent=False)
    print(f"static: r {rel_row_err(r, g['r']):.2e}  C {rel_row_err(Cm.reshape(-1, 900), g['C']):.2e}  "
          f"hh {rel_row_err(hh[:, :7], g['hh'][:, :7]):.2e}")
    for i in range(len(g["u"])):
        assert rel_row_err(r[i], g["r"][i]) < 1e-12
        assert rel_row_err(Cm[i], g["C"][i].reshape(30, 30)) < 1e-12
        assert rel_row_err(hh[i, :7], g["hh"][i, :7]) < 1e-12
    # the six phase indicators are exact, on jets inside the energy's wells too
    assert np.array_equal(hh[:, 7:], g["hh"][:, 7:])
    assert np.array_equal(hp[:, 7:], g["hh_phase"][:, 7:])
    assert rel_row_err(hp[:, :7], g["hh_phase"][:, :7]) < 1e-12


def test_transient_physics_on_device_matches_reference(igap):
    """gradelasttime: inertia/damping value rows and the truncated
    average-vector-field stress, at increments from 1e-4 to 0.3."""
    g = _load("pointwise_transient.npz")
    with _ctx(igap, igap.PhysGradElastTime(par=g["par"])) as A:
        r, Cm, hh = A.phys_eval(g["u"])
    print(f"transient: r {rel_row_err(r, g['r']):.2e}  C {rel_row_err(Cm.reshape(-1, 900), g['C']):.2e}  "
          f"hh {rel_row_err(hh, g['hh']):.2e}")
    for i in range(len(g["u"])):
        assert rel_row_err(r[i], g["r"][i]) < 1e-12
        assert rel_row_err(Cm[i], g["C"][i].reshape(30, 30)) < 1e-12
        assert rel_row_err(hh[i], g["hh"][i]) < 1e-12


def test_phys_eval_rejects_bad_arguments(igap):
    import ctypes as C
    from igap4_amd import _lib
    with _ctx(igap, igap.PhysGradElast()) as A:
        assert _lib.lib().igap_gpu_phys_eval(A.ctx, -1, None, None, None, None) == _lib.IGAP_ERR_ARG
        assert _lib.lib().igap_gpu_phys_eval(A.ctx, 4, None, None, None, None) == _lib.IGAP_ERR_ARG
        assert _lib.lib().igap_gpu_phys_eval(A.ctx, 0, None, None, None, None) == 0
    assert _lib.lib().igap_gpu_phys_eval(None, 1, C.c_void_p(8), None, None, None) == _lib.IGAP_ERR_ARG
