"""ctypes binding of the CPU restatement (oracle/igacpu.c) and of the compiled
reference pieces (oracle/_ref/libigapref.so, when built).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
and by ``bench.py``'s cpu_baseline leg, always as the checker or the CPU
baseline -- never as the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

c_double_p = C.POINTER(C.c_double)
c_int_p = C.POINTER(C.c_int)
c_i64_p = C.POINTER(C.c_int64)
c_i32_p = C.POINTER(C.c_int32)


def _p(a, t=c_double_p):
    return a.ctypes.data_as(t)


class OcPhys(C.Structure):
    _fields_ = [("torder", C.c_int), ("coef", C.c_double * 6),
                ("rho", C.c_double), ("cii", C.c_double), ("dt", C.c_double)]


class OcMesh(C.Structure):
    _fields_ = [("nelem", C.c_int * 3), ("periodic", C.c_int * 3), ("na", C.c_int * 3),
                ("nb", C.c_int * 3), ("knot", c_double_p * 3), ("w", c_double_p),
                ("X", c_double_p)]


class OcBC(C.Structure):
    _fields_ = [("bc_type", C.c_int * 36), ("par_periodic", C.c_double * 9),
                ("par_dirichlet", C.c_double * 9), ("neumann", C.c_double * 36)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(path)
        L.oc_csr_nnz.restype = C.c_int64
        L.oc_csr_nnz_rows.restype = C.c_int64
        L.oc_ndof.restype = C.c_int64
        L.oc_energy.restype = C.c_double
        L.oc_volume.restype = C.c_double
        L.oc_static_energy.restype = C.c_double
        L.oc_evalN.restype = C.c_double
        L.oc_evalN.argtypes = [c_double_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
        _LIB = L
    return _LIB


def ref_lib():
    """The compiled reference pieces, or None when they are not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libigapref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.evalN.restype = C.c_double
        R.evalN.argtypes = [c_double_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
        _REF = R
    return _REF


# ---------------------------------------------------------------- physics
REF_PAR_STATIC = np.array([0.25, 0.50, 360., 180., 0.05, 1.0, 1.0, 1.e-3 * 0.5, 0.0])
REF_PAR_TRANSIENT = np.array([0.25, 0.50, 500., 250., 0.025, 1.0, 1.0, 1.e-3 * 0.5, 0.0])
# St Venant-Kirchhoff lambda = mu = 1 plus a strain-gradient term l = 0.1:
# (c2, c3, c4, e0, e345, le) = (mu, 0, 0, 3 lambda/2 + mu, 2 mu, l)
SVK_COEF = np.array([1.0, 0.0, 0.0, 2.5, 2.0, 0.1])


def coef_from_par(par):
    out = np.zeros(6)
    lib().oc_coef_from_par(_p(np.ascontiguousarray(par, dtype=np.float64)), _p(out))
    return out


def make_phys(torder=0, par=None, coef=None):
    if par is None:
        par = REF_PAR_STATIC if torder == 0 else REF_PAR_TRANSIENT
    par = np.ascontiguousarray(par, dtype=np.float64)
    ph = OcPhys()
    if coef is None:
        lib().oc_phys_init(C.byref(ph), torder, _p(par), None)
    else:
        cf = np.ascontiguousarray(coef, dtype=np.float64)
        lib().oc_phys_init(C.byref(ph), torder, _p(par), _p(cf))
    return ph


def phys_residual(ph, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    r = np.zeros(30)
    lib().oc_phys_residual(C.byref(ph), _p(r), _p(u))
    return r


def phys_tangent(ph, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    Cm = np.zeros(900)
    lib().oc_phys_tangent(C.byref(ph), _p(Cm), _p(u))
    return Cm.reshape(30, 30)


def static_energy(u, coef):
    u = np.ascontiguousarray(u, dtype=np.float64)
    cf = np.ascontiguousarray(coef, dtype=np.float64)
    return lib().oc_static_energy(_p(u), _p(cf))


def gauss4():
    x = np.zeros(4)
    w = np.zeros(4)
    lib().oc_gauss4(_p(x), _p(w))
    return x, w


# ---------------------------------------------------------------- mesh
class Mesh:
    """igabeam patch (mesh/igabeam.c:9-73) with per-direction element counts."""

    def __init__(self, nelem, L=(4.0, 1.0, 1.0), periodic=(0, 0, 0), perturbed=False):
        self.nelem = tuple(int(n) for n in nelem)
        self.periodic = tuple(int(p) for p in periodic)
        self.L = tuple(float(x) for x in L)
        self.na = tuple(n + 2 for n in self.nelem)
        self.nb = tuple(n if p else n + 2 for n, p in zip(self.nelem, self.periodic))
        self.knots = [np.zeros(n + 5) for n in self.nelem]
        ncp = self.na[0] * self.na[1] * self.na[2]
        self.w = np.zeros(ncp)
        self.X = np.zeros(3 * ncp)
        Ls = np.array(self.L)
        ne = np.array(self.nelem, dtype=np.int32)
        pe = np.array(self.periodic, dtype=np.int32)
        lib().oc_mesh_igabeam(_p(Ls), _p(ne, c_int_p), _p(pe, c_int_p), int(perturbed),
                              _p(self.knots[0]), _p(self.knots[1]), _p(self.knots[2]),
                              _p(self.w), _p(self.X))
        self._build()

    def _build(self):
        m = OcMesh()
        for d in range(3):
            m.nelem[d] = self.nelem[d]
            m.periodic[d] = self.periodic[d]
            m.na[d] = self.na[d]
            m.nb[d] = self.nb[d]
            m.knot[d] = _p(self.knots[d])
        m.w = _p(self.w)
        m.X = _p(self.X)
        self.c = m

    @property
    def nelem_total(self):
        return self.nelem[0] * self.nelem[1] * self.nelem[2]

    @property
    def ndof(self):
        return 3 * self.nb[0] * self.nb[1] * self.nb[2]


# ---------------------------------------------------------------- BCs
def make_bc(bc_type=None, par_periodic=None, par_dirichlet=None, neumann=None):
    """Defaults follow IGAPBCAddDefault (src/interface.c:60-74): all Neumann, zero."""
    bc = OcBC()
    bt = [1] * 36 if bc_type is None else list(bc_type)
    for i in range(36):
        bc.bc_type[i] = int(bt[i])
    for i in range(9):
        bc.par_periodic[i] = 0.0 if par_periodic is None else float(par_periodic[i])
        bc.par_dirichlet[i] = 0.0 if par_dirichlet is None else float(par_dirichlet[i])
    for i in range(36):
        bc.neumann[i] = 0.0 if neumann is None else float(neumann[i])
    return bc


def clamped_tip_bc(tip=0.0, tip_dof=2):
    """Face 0 clamped in displacement and gradient (orders 0 and 1, all dofs);
    face 1 order-0 Dirichlet on ``tip_dof`` with u = a * xi_0 (a = tip)."""
    bt = [1] * 36
    for order in (0, 1):
        for dof in range(3):
            bt[6 * 0 + 3 * order + dof] = 0
    bt[6 * 1 + 0 + tip_dof] = 0
    pd = [0.0] * 9
    pd[3 * tip_dof + 0] = tip
    return make_bc(bc_type=bt, par_dirichlet=pd)


# ---------------------------------------------------------------- assembly
def csr_pattern(mesh, nrows=None, row0=0):
    """CSR pattern of the tangent (rowptr, col) for the rows [row0, row0 +
    nrows) (default: every row); a row window serves tangents restricted to
    the elements near one end of the patch."""
    r0 = max(0, min(int(row0), mesh.ndof))
    nr = mesh.ndof - r0 if nrows is None else max(0, min(int(nrows), mesh.ndof - r0))
    nnz = lib().oc_csr_nnz_rows(C.byref(mesh.c), C.c_int64(r0), C.c_int64(nr))
    rowptr = np.zeros(nr + 1, dtype=np.int64)
    col = np.zeros(nnz, dtype=np.int32)
    lib().oc_csr_pattern_rows(C.byref(mesh.c), C.c_int64(r0), C.c_int64(nr), _p(rowptr, c_i64_p), _p(col, c_i32_p))
    return rowptr, col


def setnz(mesh):
    d = np.zeros(mesh.ndof, dtype=np.int32)
    lib().oc_setnz(C.byref(mesh.c), _p(d, c_i32_p))
    return d


def dirichlet(mesh, bc):
    mask = np.zeros(mesh.ndof, dtype=np.int32)
    g = np.zeros(mesh.ndof)
    lib().oc_dirichlet(C.byref(mesh.c), C.byref(bc), _p(mask, c_i32_p), _p(g))
    return mask, g


def neumann(mesh, bc):
    F = np.zeros(mesh.ndof)
    lib().oc_neumann(C.byref(mesh.c), C.byref(bc), _p(F))
    return F


def _box(box):
    if box is None:
        return None
    b = np.ascontiguousarray(box, dtype=np.int32)
    return b


def residual(mesh, ph, bc, U_hist, box=None, apply_bc=True, nthreads=0):
    U = np.ascontiguousarray(U_hist, dtype=np.float64)
    R = np.zeros(mesh.ndof)
    b = _box(box)
    lib().oc_residual(C.byref(mesh.c), C.byref(ph), C.byref(bc), _p(U), _p(R),
                      _p(b, c_int_p) if b is not None else None, int(apply_bc), int(nthreads))
    return R


def tangent(mesh, ph, bc, U_hist, pattern=None, box=None, apply_bc=True, nthreads=0, row0=0):
    """Tangent_iga into ``pattern`` (default: the full pattern); a row-window
    pattern from csr_pattern(mesh, nrows, row0) keeps only those rows (pass
    the same row0)."""
    U = np.ascontiguousarray(U_hist, dtype=np.float64)
    rowptr, col = csr_pattern(mesh) if pattern is None else pattern
    val = np.zeros(len(col))
    b = _box(box)
    lib().oc_tangent(C.byref(mesh.c), C.byref(ph), C.byref(bc), _p(U), _p(rowptr, c_i64_p),
                     _p(col, c_i32_p), _p(val), _p(b, c_int_p) if b is not None else None,
                     int(apply_bc), int(nthreads), C.c_int64(int(row0)), C.c_int64(len(rowptr) - 1))
    return rowptr, col, val


def element_residual(mesh, ph, bc, U_hist, e):
    U = np.ascontiguousarray(U_hist, dtype=np.float64)
    v = np.zeros(81)
    lib().oc_element_residual(C.byref(mesh.c), C.byref(ph), C.byref(bc), _p(U), e[0], e[1], e[2], _p(v))
    return v


def element_tangent(mesh, ph, bc, U_hist, e):
    U = np.ascontiguousarray(U_hist, dtype=np.float64)
    v = np.zeros(6561)
    lib().oc_element_tangent(C.byref(mesh.c), C.byref(ph), C.byref(bc), _p(U), e[0], e[1], e[2], _p(v))
    return v.reshape(81, 81)


def energy(mesh, ph, bc, U):
    U = np.ascontiguousarray(U, dtype=np.float64)
    return lib().oc_energy(C.byref(mesh.c), C.byref(ph), C.byref(bc), _p(U))


def volume(mesh):
    return lib().oc_volume(C.byref(mesh.c))


def phys_density(ph, u):
    """gradelast_density (13 values) or gradelasttime_density (2 values)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    hh = np.zeros(13)
    lib().oc_phys_density(C.byref(ph), _p(hh), _p(u))
    return hh[:13 if ph.torder == 0 else 2]


def intdensity(mesh, ph, bc, U_hist):
    """iga_intdensity (src/postcore_iga.c:13-379)."""
    U = np.ascontiguousarray(U_hist, dtype=np.float64)
    HH = np.zeros(13)
    n = lib().oc_intdensity(C.byref(mesh.c), C.byref(ph), C.byref(bc), _p(U), _p(HH))
    return HH[:n]


def element_basis(mesh, e, q):
    out = np.zeros(272)
    lib().oc_element_basis(C.byref(mesh.c), e[0], e[1], e[2], q[0], q[1], q[2], _p(out))
    return out[:270].reshape(27, 10), out[270], out[271]


def csr_to_dense(rowptr, col, val, n):
    A = np.zeros((n, n))
    for r in range(n):
        A[r, col[rowptr[r]:rowptr[r + 1]]] = val[rowptr[r]:rowptr[r + 1]]
    return A


# ---------------------------------------------------------------- inputs
def glibc_rand_state(n, seed=123456, scale=0.01):
    """U_hist as IGAPSolvRand (src/interface.c:261-269) fills it on one rank:
    scale*(rand()%999999999)/999999999. with glibc rand() seeded by srand(seed)."""
    libc = C.CDLL("libc.so.6")
    libc.srand(C.c_uint(seed))
    libc.rand.restype = C.c_int
    out = np.empty(n)
    for i in range(n):
        out[i] = scale * float(libc.rand() % 999999999) / 999999999.
    return out


def smooth_state(mesh, amp=0.01, levels=1):
    """Throughput input: U_i(X) = amp sin(pi X/Lx) cos(pi Y/Ly) cos(pi Z/Lz) (1, .5, .25)_i
    sampled at the owned control points."""
    na = mesh.na
    X = mesh.X.reshape(na[0], na[1], na[2], 3)[:mesh.nb[0], :mesh.nb[1], :mesh.nb[2]]
    Lx, Ly, Lz = mesh.L
    f = amp * np.sin(np.pi * X[..., 0] / Lx) * np.cos(np.pi * X[..., 1] / Ly) * np.cos(np.pi * X[..., 2] / Lz)
    U = np.stack([f, 0.5 * f, 0.25 * f], axis=-1).reshape(-1)
    return np.tile(U, levels)
