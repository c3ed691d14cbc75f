"""Host-side mirror of IGAP4's assembly interface on top of the C ABI.

Names follow the reference: ``MeshIGABeam`` (mesh/meshigabeam.c:10-27),
``PhysGradElast`` / ``PhysGradElastTime`` (phys/physgradelast*.c),
``BCDefault`` (IGAPBCAddDefault, src/interface.c:60-74) and ``Assm``, whose
``residual`` / ``tangent`` have the semantics of Residual_iga / Tangent_iga
(src/assmcore_iga.c:16-502, 507-858) as registered with SNES by IGAPAssmInit
(src/interface.c:131-203, 218-219).  Arrays crossing this layer are numpy
host arrays; ``*_dev`` methods take device addresses.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import BCDesc, MeshDesc, PhysDesc, check, lib

# St Venant-Kirchhoff lambda = mu = 1 with a strain-gradient length 0.1, as
# energy coefficients (c2, c3, c4, e0, e345, le) = (mu, 0, 0, 3 lambda/2 + mu, 2 mu, l)
SVK_COEF = (1.0, 0.0, 0.0, 2.5, 2.0, 0.1)


def _dp(a):
    return a.ctypes.data_as(_lib.c_double_p)


class MeshIGABeam:
    """Patch of igabeam (mesh/igabeam.c:9-73) with per-direction element counts;
    ``perturbed`` gives the igabeamtest geometry (mesh/igabeamtest.c:68-70)."""

    def __init__(self, L=(4.0, 1.0, 1.0), nelem=(16, 4, 4), periodic=(0, 0, 0), perturbed=False):
        self.L = tuple(float(x) for x in L)
        self.nelem = tuple(int(n) for n in nelem)
        self.periodic = tuple(1 if p else 0 for p in periodic)
        self.na = tuple(n + 2 for n in self.nelem)
        self.nb = tuple(n if p else n + 2 for n, p in zip(self.nelem, self.periodic))
        self.knots = [np.zeros(n + 5) for n in self.nelem]
        ncp = int(np.prod(self.na))
        self.w = np.zeros(ncp)
        self.X = np.zeros(3 * ncp)
        Ls = np.array(self.L, dtype=np.float64)
        ne = (C.c_int * 3)(*self.nelem)
        pe = (C.c_int * 3)(*self.periodic)
        check(lib().igap_mesh_igabeam(_dp(Ls), ne, pe, int(bool(perturbed)), _dp(self.knots[0]),
                                      _dp(self.knots[1]), _dp(self.knots[2]), _dp(self.w), _dp(self.X)))

    @property
    def ndof(self):
        return 3 * int(np.prod(self.nb))

    @property
    def nelem_total(self):
        return int(np.prod(self.nelem))

    def desc(self):
        d = MeshDesc()
        d.porder = 2
        for i in range(3):
            d.nelem[i] = self.nelem[i]
            d.periodic[i] = self.periodic[i]
            d.knot[i] = _dp(self.knots[i])
        d.w = _dp(self.w)
        d.X = _dp(self.X)
        return d


def smooth_state(mesh, amp=0.01, levels=1):
    """Synthetic displacement U_i(X) = amp sin(pi X/Lx) cos(pi Y/Ly) cos(pi Z/Lz) (1, .5, .25)_i
    at the owned control points of ``mesh`` (a smooth, non-trivial state for
    throughput runs), ``levels`` copies for the time levels of U_hist."""
    na, nb = mesh.na, mesh.nb
    X = mesh.X.reshape(na[0], na[1], na[2], 3)[:nb[0], :nb[1], :nb[2]]
    Lx, Ly, Lz = mesh.L
    f = amp * np.sin(np.pi * X[..., 0] / Lx) * np.cos(np.pi * X[..., 1] / Ly) * np.cos(np.pi * X[..., 2] / Lz)
    U = np.stack([f, 0.5 * f, 0.25 * f], axis=-1).reshape(-1)
    return np.tile(U, levels)


def PhysGradElast(par=None, coef=None):
    """gradelast (static) with the defaults of phys/physgradelast.c:21; ``coef``
    overrides the energy coefficients (e.g. ``SVK_COEF``)."""
    ph = PhysDesc()
    lib().igap_phys_gradelast(C.byref(ph))
    return _phys_fill(ph, par, coef)


def PhysGradElastTime(par=None, coef=None):
    """gradelasttime (transient) with the defaults of phys/physgradelasttime.c:21."""
    ph = PhysDesc()
    lib().igap_phys_gradelasttime(C.byref(ph))
    return _phys_fill(ph, par, coef)


def _phys_fill(ph, par, coef):
    if par is not None:
        for i in range(9):
            ph.par[i] = float(par[i])
    if coef is not None:
        ph.use_coef = 1
        for i in range(6):
            ph.coef[i] = float(coef[i])
    return ph


def BCDefault():
    bc = BCDesc()
    lib().igap_bc_default(C.byref(bc))
    return bc


def BC(bc_type=None, par_periodic=None, par_dirichlet=None, neumann=None):
    bc = BCDefault()
    if bc_type is not None:
        for i in range(36):
            bc.bc_type[i] = int(bc_type[i])
    for name, arr, n in (("par_periodic", par_periodic, 9), ("par_dirichlet", par_dirichlet, 9),
                         ("neumann", neumann, 36)):
        if arr is not None:
            dst = getattr(bc, name)
            for i in range(n):
                dst[i] = float(arr[i])
    return bc


def BCClampedTip(tip=0.0, tip_dof=2):
    """Face 0 clamped in displacement and normal gradient (orders 0 and 1, all
    dofs); face 1 has order-0 Dirichlet u_{tip_dof} = tip * xi_0."""
    bt = [1] * 36
    for order in (0, 1):
        for dof in range(3):
            bt[6 * 0 + 3 * order + dof] = 0
    bt[6 * 1 + tip_dof] = 0
    pd = [0.0] * 9
    pd[3 * tip_dof] = tip
    return BC(bc_type=bt, par_dirichlet=pd)


class Assm:
    """IGAPAssmInit + the Residual/Tangent callbacks for one patch on one GPU."""

    def __init__(self, mesh: MeshIGABeam, phys: PhysDesc, bc: BCDesc, device: int = -1, flags: int = 0):
        """``flags``: 0 (automatic), _lib.IGAP_TANGENT_ELEMENT_MATRICES or
        _lib.IGAP_TANGENT_DIRECT (how the tangent reaches the CSR values)."""
        self.mesh = mesh
        self.phys = phys
        self.bc = bc
        ctx = C.c_void_p()
        md = mesh.desc()
        check(lib().igap_gpu_create(C.byref(md), C.byref(phys), C.byref(bc), device, int(flags), C.byref(ctx)))
        self.ctx = ctx
        ndof, nnz, nel, tor = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int()
        check(lib().igap_gpu_sizes(ctx, C.byref(ndof), C.byref(nnz), C.byref(nel), C.byref(tor)))
        self.ndof, self.nnz, self.nel, self.torder = ndof.value, nnz.value, nel.value, tor.value

    # ------------------------------------------------------------ lifetime
    def close(self):
        if self.ctx:
            check(lib().igap_gpu_destroy(self.ctx))
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------ device buffers
    def uhist_ptr(self):
        return lib().igap_gpu_uhist(self.ctx)

    def rvec_ptr(self):
        return lib().igap_gpu_rvec(self.ctx)

    def htod(self, dst_ptr, arr):
        arr = np.ascontiguousarray(arr)
        check(lib().igap_gpu_memcpy_htod(self.ctx, dst_ptr, arr.ctypes.data, arr.nbytes))

    def dtoh(self, src_ptr, n, dtype=np.float64):
        out = np.empty(n, dtype=dtype)
        check(lib().igap_gpu_memcpy_dtoh(self.ctx, out.ctypes.data, src_ptr, out.nbytes))
        return out

    def set_uhist(self, U_hist):
        U_hist = np.ascontiguousarray(U_hist, dtype=np.float64)
        if U_hist.size != (self.torder + 1) * self.ndof:
            raise ValueError(f"U_hist must hold {(self.torder + 1) * self.ndof} values")
        self.htod(self.uhist_ptr(), U_hist)

    def get_uhist(self):
        return self.dtoh(self.uhist_ptr(), (self.torder + 1) * self.ndof)

    def intdensity(self):
        """iga_intdensity / IGAPPostComputeHH of the current U_hist."""
        HH = np.zeros(13)
        n = C.c_int()
        check(lib().igap_gpu_intdensity(self.ctx, _dp(HH), C.byref(n)))
        return HH[:n.value]

    def selftest_ilu(self, prec="f64"):
        """Own point-block ILU(0) against rocSPARSE's on the tangent last
        assembled in ``prec``: (applicable, max|a-b|/max|b|)."""
        ok = C.c_int()
        d = C.c_double()
        check(lib().igap_gpu_selftest_ilu(self.ctx, 0 if prec == "f64" else 1, C.byref(ok), C.byref(d)))
        return bool(ok.value), d.value

    def save_solution(self, fname):
        """IGAPSolvSaveSolution: the reference's solution file (universal dof order)."""
        check(lib().igap_gpu_save_solution(self.ctx, str(fname).encode()))

    def load_solution(self, fname):
        """IGAPSolvLoadSolution (transient: level 0 = level 1 afterwards)."""
        check(lib().igap_gpu_load_solution(self.ctx, str(fname).encode()))

    def set_bc(self, bc):
        self.bc = bc
        check(lib().igap_gpu_set_bc(self.ctx, C.byref(bc)))

    def pattern(self):
        rowptr = self.dtoh(lib().igap_gpu_csr_rowptr(self.ctx), self.ndof + 1, np.int64)
        col = self.dtoh(lib().igap_gpu_csr_col(self.ctx), self.nnz, np.int32)
        return rowptr, col

    def dirichlet(self):
        mask = self.dtoh(lib().igap_gpu_dirichlet_mask(self.ctx), self.ndof, np.int32)
        g = self.dtoh(lib().igap_gpu_dirichlet_values(self.ctx), self.ndof)
        return mask, g

    # ------------------------------------------------------------ callbacks (device)
    def residual_dev(self, U_ptr=None, R_ptr=None):
        check(lib().igap_gpu_residual(self.ctx, U_ptr, R_ptr))

    def tangent_dev(self, U_ptr=None, vals_ptr=None, prec=_lib.IGAP_F64):
        check(lib().igap_gpu_tangent(self.ctx, U_ptr, vals_ptr, prec))

    def synchronize(self):
        check(lib().igap_gpu_synchronize(self.ctx))

    # ------------------------------------------------------------ callbacks (host arrays)
    def residual(self, U_hist):
        """Residual_iga on a host U_hist[(torder+1)*ndof]; returns R (host)."""
        self.set_uhist(U_hist)
        self.residual_dev(None, None)
        return self.dtoh(self.rvec_ptr(), self.ndof)

    def tangent(self, U_hist, prec="f64"):
        """Tangent_iga on a host U_hist; returns (rowptr, col, vals)."""
        self.set_uhist(U_hist)
        p = _lib.IGAP_F64 if prec in ("f64", 0) else _lib.IGAP_F32
        self.tangent_dev(None, None, p)
        rowptr, col = self.pattern()
        if p == _lib.IGAP_F64:
            vals = self.dtoh(lib().igap_gpu_csr_val64(self.ctx), self.nnz)
        else:
            vals = self.dtoh(lib().igap_gpu_csr_val32(self.ctx), self.nnz, np.float32)
        return rowptr, col, vals

    # ------------------------------------------------------------ pointwise physics
    def phys_eval(self, u, residual=True, tangent=True, density=True):
        """The Phys callbacks (residual / tangent / density, include/physimpl.h:6-9)
        of this context's physics on the GPU at host jets u[npts][30*(torder+1)];
        returns (r[npts][30], C[npts][30][30], hh[npts][nd]) (None where not asked)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        w = 30 * (self.torder + 1)
        u = u.reshape(-1, w)
        n = u.shape[0]
        nd = 13 if self.torder == 0 else 2
        sizes = {"u": u.size, "r": 30 * n if residual else 0, "C": 900 * n if tangent else 0,
                 "hh": nd * n if density else 0}
        ptrs = {}
        try:
            for k, m in sizes.items():
                if m:
                    p = C.c_void_p()
                    check(lib().igap_gpu_malloc(self.ctx, C.byref(p), 8 * m))
                    ptrs[k] = p
            self.htod(ptrs["u"], u)
            check(lib().igap_gpu_phys_eval(self.ctx, n, ptrs["u"], ptrs.get("r"), ptrs.get("C"), ptrs.get("hh")))
            self.synchronize()
            r = self.dtoh(ptrs["r"], 30 * n).reshape(n, 30) if residual else None
            Cm = self.dtoh(ptrs["C"], 900 * n).reshape(n, 30, 30) if tangent else None
            hh = self.dtoh(ptrs["hh"], nd * n).reshape(n, nd) if density else None
        finally:
            for p in ptrs.values():
                lib().igap_gpu_free(self.ctx, p)
        return r, Cm, hh

    # ------------------------------------------------------------ Newton-Krylov
    def snes_solve(self, opts=None, **kw):
        """Newton solve from U_hist level 0 (SNESSolve driven by IGAPSolvSolve,
        src/interface.c:302-328); returns the solver statistics as a dict."""
        if opts is None:
            opts = default_solver_opts(**kw)
        st = _lib.SolverStats()
        check(lib().igap_gpu_snes_solve(self.ctx, C.byref(opts), C.byref(st)))
        return st.as_dict()

    def pc_apply(self, b, which="full"):
        """x = M^-1 b with the preconditioner of the last solve (PCApply);
        ``which`` "smoother": the multigrid's finest-level smoother only.
        Host vectors in and out."""
        w = _lib.IGAP_PC_APPLY_FULL if which == "full" else _lib.IGAP_PC_APPLY_SMOOTHER
        b = np.ascontiguousarray(b, dtype=np.float64)
        ptrs = []
        try:
            for _ in range(2):
                p = C.c_void_p()
                check(lib().igap_gpu_malloc(self.ctx, C.byref(p), 8 * self.ndof))
                ptrs.append(p)
            self.htod(ptrs[0], b)
            check(lib().igap_gpu_pc_apply(self.ctx, w, ptrs[0], ptrs[1]))
            self.synchronize()
            return self.dtoh(ptrs[1], self.ndof)
        finally:
            for p in ptrs:
                lib().igap_gpu_free(self.ctx, p)

    def tangent_apply(self, x, prec="f64"):
        """y = K x (MatMult) with the context's own tangent values of ``prec``
        (the last tangent assembled into them).  Host vectors in and out."""
        p = _lib.IGAP_F64 if prec in ("f64", 0) else _lib.IGAP_F32
        x = np.ascontiguousarray(x, dtype=np.float64)
        ptrs = []
        try:
            for _ in range(2):
                q = C.c_void_p()
                check(lib().igap_gpu_malloc(self.ctx, C.byref(q), 8 * self.ndof))
                ptrs.append(q)
            self.htod(ptrs[0], x)
            check(lib().igap_gpu_tangent_apply(self.ctx, p, ptrs[0], ptrs[1]))
            self.synchronize()
            return self.dtoh(ptrs[1], self.ndof)
        finally:
            for q in ptrs:
                lib().igap_gpu_free(self.ctx, q)

    # ------------------------------------------------------------ timing
    def timing(self, on=True):
        check(lib().igap_gpu_timing_enable(self.ctx, 1 if on else 0))

    def timing_reset(self):
        check(lib().igap_gpu_timing_reset(self.ctx))

    def timing_get(self, which):
        ms, n = C.c_double(), C.c_int64()
        check(lib().igap_gpu_timing_get(self.ctx, which, C.byref(ms), C.byref(n)))
        return ms.value, n.value


def default_solver_opts(**kw):
    """igap_gpu_solver_default_opts with keyword overrides, e.g.
    ``default_solver_opts(snes_rtol=1e-10, tangent_prec=1)``."""
    o = _lib.SolverOpts()
    check(lib().igap_gpu_solver_default_opts(C.byref(o)))
    for k, v in kw.items():
        if k in ("tangent_prec",) and isinstance(v, str):
            v = _lib.IGAP_F64 if v == "f64" else _lib.IGAP_F32
        if k == "pc" and isinstance(v, str):
            v = {"ilu0": _lib.IGAP_PC_ILU0, "jacobi": _lib.IGAP_PC_JACOBI, "none": _lib.IGAP_PC_NONE,
                 "mg": _lib.IGAP_PC_MG}[v]
        setattr(o, k, v)
    return o


def selftest_mfma():
    """Raw result registers [r][lane] of the FP64 and FP32 16x16x4 products."""
    o64 = np.zeros(256)
    o32 = np.zeros(256, dtype=np.float32)
    check(lib().igap_gpu_selftest_mfma(_dp(o64), o32.ctypes.data_as(_lib.c_float_p)))
    return o64.reshape(4, 64), o32.reshape(4, 64)


# result row held by register r of lane l, as the tangent kernels read it
MFMA_ROW = {"f64": lambda kr, r: 4 * r + kr, "f32": lambda kr, r: 4 * kr + r}


def selftest_unpack(raw, rowmap):
    D = np.full((16, 16), np.nan)
    for r in range(4):
        for l in range(64):
            D[rowmap(l >> 4, r), l & 15] = raw[r, l]
    return D


def selftest_expected():
    i = np.arange(16)[:, None]
    k = np.arange(4)[None, :]
    A = 1.0 + i + 0.25 * k
    kk = np.arange(4)[:, None]
    j = np.arange(16)[None, :]
    B = 2.0 - 0.5 * j + kk
    return A @ B
