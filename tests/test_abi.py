"""The C ABI library loads here (no GPU needed), exports every symbol that
the headers under include/ declare, and its host-only entry points agree with the
oracle."""
import ctypes as C
import os
import re

import numpy as np

from conftest import ROOT


def header_symbols():
    """Every function declared by the headers under include/."""
    syms = set()
    inc = os.path.join(ROOT, "include")
    for f in sorted(os.listdir(inc)):
        if f.endswith(".h"):
            txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(inc, f)).read(), flags=re.S)
            syms |= set(re.findall(r"\b(igap_\w+)\s*\(", txt))
    return sorted(syms)


def test_library_exports_every_declared_symbol(igap):
    from igap4_amd import _lib
    L = C.CDLL(_lib.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # and the Python binding types every one of them
    assert sorted(_lib.declared_symbols()) == syms


def test_mesh_generator_matches_oracle(igap, oracle):
    for nel, per, pert in (((3, 2, 4), (0, 0, 0), False), ((5, 6, 5), (1, 0, 1), False), ((7, 6, 6), (0, 0, 0), True)):
        a = igap.MeshIGABeam(L=(4, 1, 1), nelem=nel, periodic=per, perturbed=pert)
        b = oracle.Mesh(nel, L=(4, 1, 1), periodic=per, perturbed=pert)
        assert all(np.array_equal(x, y) for x, y in zip(a.knots, b.knots))
        assert np.array_equal(a.w, b.w) and np.array_equal(a.X, b.X)
        assert a.ndof == b.ndof


def test_phys_defaults_are_the_reference_constructors(igap, oracle):
    s = igap.PhysGradElast()
    t = igap.PhysGradElastTime()
    assert s.kind == 0 and t.kind == 2 and s.use_coef == 0
    assert np.array_equal(np.array(s.par[:]), oracle.REF_PAR_STATIC)
    assert np.array_equal(np.array(t.par[:]), oracle.REF_PAR_TRANSIENT)
    bc = igap.BCDefault()
    assert list(bc.bc_type) == [1] * 36
    assert not any(bc.par_dirichlet) and not any(bc.par_periodic) and not any(bc.neumann)


def test_bad_descriptor_is_an_error_not_a_crash(igap):
    """Validation happens on the host before any device work."""
    from igap4_amd import _lib
    m = igap.MeshIGABeam(nelem=(4, 3, 3), periodic=(1, 0, 0))
    d = m.desc()
    ph = igap.PhysGradElast()
    bc = igap.BCDefault()
    ctx = C.c_void_p()
    rc = _lib.lib().igap_gpu_create(C.byref(d), C.byref(ph), C.byref(bc), -1, 0, C.byref(ctx))
    assert rc == 1 and b"periodic" in _lib.lib().igap_gpu_last_error()
    d.porder = 3
    rc = _lib.lib().igap_gpu_create(C.byref(d), C.byref(ph), C.byref(bc), -1, 0, C.byref(ctx))
    assert rc == 1 and b"porder" in _lib.lib().igap_gpu_last_error()


def test_null_context_is_an_error_not_a_crash(igap):
    """Every entry point that takes a context rejects a null one on the host,
    with IGAP_ERR_ARG and a message, before any device work; destroying a
    null context is a no-op, as free(NULL)."""
    from igap4_amd import _lib
    L = _lib.lib()
    d, f = C.c_double(), C.c_float()
    i, i64 = C.c_int(), C.c_int64()
    calls = {
        "igap_gpu_residual": (None, None, None),
        "igap_gpu_tangent": (None, None, None, 0),
        "igap_gpu_set_bc": (None, None),
        "igap_gpu_sizes": (None, None, None, None, None),
        "igap_gpu_intdensity": (None, C.byref(d), C.byref(i)),
        "igap_gpu_save_solution": (None, b"/nonexistent/x.bin"),
        "igap_gpu_load_solution": (None, b"/nonexistent/x.bin"),
        "igap_gpu_synchronize": (None,),
        "igap_gpu_timing_enable": (None, 1),
        "igap_gpu_timing_get": (None, 0, C.byref(d), C.byref(i64)),
        "igap_gpu_timing_reset": (None,),
        "igap_gpu_snes_solve": (None, None, None),
        "igap_gpu_selftest_ilu": (None, 0, C.byref(i), C.byref(d)),
        "igap_ws_nparams": (None, C.byref(i64)),
        "igap_ws_forward": (None, None, None),
        "igap_ws_train_step": (None, None, None, 0.0, C.byref(d)),
    }
    for name, args in calls.items():
        rc = getattr(L, name)(*args)
        assert rc == 1, name      # IGAP_ERR_ARG
        assert L.igap_gpu_last_error(), name
    assert L.igap_gpu_destroy(None) == _lib.IGAP_OK
    assert L.igap_ws_destroy(None) == _lib.IGAP_OK
