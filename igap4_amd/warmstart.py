"""Learned Newton warm start on the control-point grid.

A 3-D convolutional network (include/igap_ws.h: 4 layers of 3x3x3
convolutions, 16 channels, BF16 matrix-core products with FP32 accumulation,
trained on the GPU with Adam) corrects the quadratic extrapolation of the
next load-step solution.  With the increments dU_j = U_j - U_{j-1} and the
second difference d2 = dU_{k-1} - dU_{k-2}, its inputs are

  * dU_{k-1} / s1,  s1 = rms(dU_{k-1})                      (3 channels)
  * d2 / s2,        s2 = rms(d2)                            (3 channels)
  * dG_k / s1, the load increment g_k - g_{k-1} (Dirichlet data) (3 channels)

so that the small second difference reaches the BF16 products at full
relative precision instead of as the difference of two O(1) inputs.  The
network predicts the third difference (U_k - 3 U_{k-1} + 3 U_{k-2} - U_{k-3})
/ s2, zero on Dirichlet dofs, and the candidate is

  U_ext + d2 + s2 net(x),   U_ext = 2 U_{k-1} - U_{k-2},

i.e. quadratic extrapolation plus the learned term (a skip connection: the
library creates the last layer at zero, so an untrained network gives the
quadratic extrapolation).  The learned term is projected onto the
next-coarser spline space (CoarseFilter: exact knot-insertion prolongation
after interpolation at the coarse Greville abscissae), which keeps its smooth
part and drops the rough error the stiffness would amplify.  Training keeps the epoch with the lowest loss on
the last training steps (held out); optionally (residual_finetune) it then
continues on the candidate's residual, linearised with the assembled tangent
(igap_gpu_tangent_apply), instead of the distance to the next solution.  The
learned term is used at all
only if it lowers the residual norm of the quadratic extrapolation on those
held-out steps.  The candidate is used only if its residual norm is below
the residual norm of the linear extrapolation U_ext, which otherwise stays
the initial iterate; Newton then solves to the same tolerance either way.

IGAP4 itself starts every step from the previous solution
(src/interface.c:252-259,319); extrapolation and the learned correction are
this package's additions around that.

  python -m igap4_amd.warmstart --nelem 64,16,16 --steps 200 --train 100
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import time

import numpy as np

from . import _lib
from ._lib import check, lib

CIN, COUT = 9, 3
DEFAULT_DIL = (1, 2, 4, 1)


def _fp(a):
    return a.ctypes.data_as(_lib.c_float_p)


class WarmStartNet:
    """The network of include/igap_ws.h on one GPU."""

    def __init__(self, grid, cin=CIN, cout=COUT, dil=DEFAULT_DIL, seed=0, device=0):
        self.grid = tuple(int(g) for g in grid)
        self.cin, self.cout = cin, cout
        self.dil = tuple(int(d) for d in dil)
        self.V = int(np.prod(self.grid))
        self.ctx = C.c_void_p()
        g = (C.c_int * 3)(*self.grid)
        d = (C.c_int * 4)(*self.dil)
        check(lib().igap_ws_create(device, g, cin, cout, d, C.c_uint64(seed), C.byref(self.ctx)))
        n = C.c_int64()
        check(lib().igap_ws_nparams(self.ctx, C.byref(n)))
        self.nparams = n.value

    def close(self):
        if self.ctx:
            lib().igap_ws_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _x(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(self.V, self.cin)
        return x

    def forward(self, x):
        x = self._x(x)
        y = np.empty((self.V, self.cout), dtype=np.float32)
        check(lib().igap_ws_forward(self.ctx, _fp(x), _fp(y)))
        return y

    def gradients(self, x, target):
        x = self._x(x)
        t = np.ascontiguousarray(target, dtype=np.float32).reshape(self.V, self.cout)
        g = np.empty(self.nparams, dtype=np.float32)
        loss = C.c_double()
        check(lib().igap_ws_gradients(self.ctx, _fp(x), _fp(t), C.byref(loss), _fp(g)))
        return loss.value, g

    def train_step(self, x, target, lr):
        x = self._x(x)
        t = np.ascontiguousarray(target, dtype=np.float32).reshape(self.V, self.cout)
        loss = C.c_double()
        check(lib().igap_ws_train_step(self.ctx, _fp(x), _fp(t), C.c_float(lr), C.byref(loss)))
        return loss.value

    def get_params(self):
        p = np.empty(self.nparams, dtype=np.float32)
        check(lib().igap_ws_get_params(self.ctx, _fp(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        check(lib().igap_ws_set_params(self.ctx, _fp(p)))


def selftest_mfma_bf16():
    out = np.empty(256, dtype=np.float32)
    check(lib().igap_ws_selftest_mfma_bf16(_fp(out)))
    return out.reshape(4, 64)


# ------------------------------------------------------------------ features
def _rms(a):
    r = float(np.sqrt(np.mean(a * a)))
    return r if r > 0 else 1.0


def features(dU1, dU2, dG):
    """x[v][9] (float32) and the scales (s1, s2) = (rms(dU1), rms(dU1 - dU2))."""
    d2 = dU1 - dU2
    s1, s2 = _rms(dU1), _rms(d2)
    x = np.concatenate([dU1.reshape(-1, 3) / s1, d2.reshape(-1, 3) / s2, dG.reshape(-1, 3) / s1], axis=1)
    return x.astype(np.float32), s1, s2


def samples_from_trajectory(U, G, mask, filt=None):
    """Training pairs from solutions U[0..K] and Dirichlet data G[0..K]:
    (x_k, target_k) for k = 3..K, target = third difference / s2 (projected
    by ``filt`` when given, so that the network fits what the filter keeps)."""
    out = []
    keep = (mask.reshape(-1) == 0)
    for k in range(3, len(U)):
        dU1, dU2 = U[k - 1] - U[k - 2], U[k - 2] - U[k - 3]
        x, s1, s2 = features(dU1, dU2, G[k] - G[k - 1])
        t3 = U[k] - 3 * U[k - 1] + 3 * U[k - 2] - U[k - 3]
        if filt is not None:
            t3 = filt(t3)
        t = (t3 * keep) / s2
        out.append((x, t.reshape(-1, 3).astype(np.float32)))
    return out


def train(net, samples, epochs=60, lr=2e-3, seed=0, log=None):
    """Adam over the samples in a seeded order; returns the mean loss per epoch."""
    rng = np.random.default_rng(seed)
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(samples))
        tot = 0.0
        for i in order:
            x, t = samples[i]
            tot += net.train_step(x, t, lr)
        hist.append(tot / len(samples))
        if log:
            log(ep, hist[-1])
    return hist


def mean_loss(net, samples):
    """Mean over samples of the training loss (mean square error)."""
    return float(np.mean([net.gradients(x, t)[0] for x, t in samples]))


def train_selected(net, samples, epochs=60, lr=2e-3, seed=0, holdout=0.1, log=None):
    """Train on the earlier steps, validate on the last ``holdout`` fraction
    (the steps nearest to the ones the network will be used on), keep the
    epoch with the lowest validation loss, and switch the correction off
    (last layer zero: the skip connection alone) when even that epoch does
    not beat a zero correction on the held-out steps.  Returns a dict."""
    nv = max(1, int(round(holdout * len(samples)))) if len(samples) > 1 else 0
    tr, va = samples[:len(samples) - nv], samples[len(samples) - nv:]
    rng = np.random.default_rng(seed)
    p0 = net.get_params()
    zero_val = float(np.mean([np.mean(t.astype(np.float64) ** 2) for _, t in va])) if va else 0.0
    best, best_p, hist, vhist = np.inf, p0.copy(), [], []
    for ep in range(epochs):
        order = rng.permutation(len(tr))
        hist.append(float(np.mean([net.train_step(*tr[i], lr) for i in order])))
        v = mean_loss(net, va) if va else hist[-1]
        vhist.append(v)
        if v < best:
            best, best_p = v, net.get_params()
        if log:
            log(ep, hist[-1], v)
    enabled = bool(best < zero_val) if va else True
    if enabled:
        net.set_params(best_p)
    else:
        p = p0.copy()
        net.set_params(p)
    return {"train_loss": hist, "val_loss": vhist, "val_loss_best": best, "val_loss_zero_correction": zero_val,
            "train_samples": len(tr), "val_samples": len(va), "correction_enabled": enabled}


def residual_shifted_target(y, rq, kmul, km, alpha):
    """(r, t) for the network's output y of one sample: r = rq + K (km y),
    the candidate's linearised residual over its scale, and the shifted mean
    square target t = y - alpha km (K r).  The gradient of the mean square
    loss |y - t|^2 / N (N = y.size) at t is (2 alpha / N) times the gradient
    of |r|^2 / 2 with respect to y, for K symmetric (kmul(v) = K v) and km
    the 0/1 mask of the dofs the network may move."""
    r = rq + kmul(y * km)
    return r, y - alpha * (km * kmul(r))


def residual_finetune(net, A, samples, cq, s2, keep, bcs, epochs=30, lr=5e-4, seed=0, holdout=0.1, cosine=False,
                      log=None):
    """Train the network on the residual of the start it builds, not on the
    distance to the next solution.  For sample i the candidate is
    c_q + s2_i * y (c_q the quadratic extrapolation, y the network's output)
    and its residual, linearised about c_q with the context's own tangent K
    (assembled by the caller), is s2_i (rq_i + K y) with rq_i = R(c_q)/s2_i;
    ``cq[i]`` and ``s2[i]`` are sample i's c_q and scale.
    The library trains on a mean square error only; with the target shifted
    to y - alpha g, g = m (K (rq_i + K (m y))) with m the mask of the free
    dofs (K symmetric), the gradient it
    back-propagates is alpha g, the gradient of |rq_i + K (m y)|^2 / 2, so Adam
    minimises the linearised residual.  alpha is fixed per call, so the
    samples keep their relative weights (residual_shifted_target).  With
    ``cosine`` the learning rate decays from lr to 0 over the epochs on a
    half cosine, else it stays lr.  The last ``holdout`` fraction is
    scored with the true residual after every epoch, median over the samples
    of |R(c_q + s2 y)| / |R(c_q)| (``bcs[i]()`` sets sample i's load), and
    the epoch with the lowest median is kept.  Returns a dict."""
    nv = max(1, int(round(holdout * len(samples))))
    tr = list(range(len(samples) - nv))
    va = list(range(len(samples) - nv, len(samples)))
    rng = np.random.default_rng(seed)
    km = keep.astype(np.float64)


    rq, r0 = [], []
    for i in range(len(samples)):
        bcs[i]()
        r = A.residual(cq[i])
        rq.append(r / s2[i])
        r0.append(float(np.linalg.norm(r)))

    def val_ratio():
        out = []
        for i in va:
            bcs[i]()
            y = net.forward(samples[i][0]).astype(np.float64).reshape(-1)
            r1 = float(np.linalg.norm(A.residual(cq[i] + s2[i] * y * km)))
            out.append(r1 / r0[i] if r0[i] > 0 else 1.0)
        return float(np.median(out))

    y0 = net.forward(samples[tr[0]][0]).astype(np.float64).reshape(-1)
    g0 = y0 - residual_shifted_target(y0, rq[tr[0]], A.tangent_apply, km, 1.0)[1]
    t0 = samples[tr[0]][1].astype(np.float64).reshape(-1)
    alpha = 0.1 * _rms(t0) / _rms(g0)
    best, best_p, hist, vhist = val_ratio(), net.get_params(), [], []
    v0 = best
    for ep in range(epochs):
        tot = 0.0
        lr_ep = 0.5 * lr * (1.0 + np.cos(np.pi * ep / epochs)) if cosine else lr
        for i in rng.permutation(tr):
            x = samples[i][0]
            y = net.forward(x).astype(np.float64).reshape(-1)
            r, t = residual_shifted_target(y, rq[i], A.tangent_apply, km, alpha)
            tot += float(np.dot(r, r))
            net.train_step(x, t.astype(np.float32).reshape(-1, 3), lr_ep)
        hist.append(tot / len(tr))
        v = val_ratio()
        vhist.append(v)
        if v < best:
            best, best_p = v, net.get_params()
        if log:
            log(ep, hist[-1], v)
    net.set_params(best_p)
    return {"train_residual": hist, "val_ratio": vhist, "val_ratio_start": v0, "val_ratio_best": best,
            "alpha": alpha}


# ------------------------------------------------------- coarse-space filter
def _bs2_rows(kv, nfun, x):
    """Values of the nfun quadratic B-splines of the open knot vector kv at the
    points x (rows), the right end belonging to the last span."""
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    out = np.zeros((x.size, nfun))
    for r, xv in enumerate(x):
        m = 2
        while m < nfun - 1 and not (xv < kv[m + 1]):
            m += 1
        left, right, N = np.zeros(3), np.zeros(3), np.zeros(3)
        N[0] = 1.0
        for j in (1, 2):
            left[j] = xv - kv[m + 1 - j]
            right[j] = kv[m + j] - xv
            saved = 0.0
            for k in range(j):
                t = N[k] / (right[k + 1] + left[j - k])
                N[k] = saved + right[k + 1] * t
                saved = left[j - k] * t
            N[j] = saved
        out[r, m - 2:m + 1] = N
    return out


def coarse_ops_1d(kf, nel, with_knots=False):
    """Prolongation P (nel+2 x nel/2+2, exact knot insertion) and the
    interpolation T at the coarse Greville abscissae (nel/2+2 x nel+2) for an
    open quadratic knot vector with an even number nel >= 4 of elements
    (and the coarse knot vector with ``with_knots``)."""
    kf = np.asarray(kf, dtype=np.float64)
    nc_el = nel // 2
    kc = np.concatenate([[kf[2]] * 3, kf[2 + 2 * np.arange(1, nc_el)], [kf[nel + 2]] * 3])
    nf, nc = nel + 2, nc_el + 2
    gf = 0.5 * (kf[1:nf + 1] + kf[2:nf + 2])
    gc = 0.5 * (kc[1:nc + 1] + kc[2:nc + 2])
    P = np.linalg.solve(_bs2_rows(kf, nf, gf), _bs2_rows(kc, nc, gf))
    T = np.linalg.solve(_bs2_rows(kc, nc, gc), _bs2_rows(kf, nf, gc))
    return (P, T, kc) if with_knots else (P, T)


class CoarseFilter:
    """c -> P (T c): the projection of a control-point field onto a coarser
    spline space (``levels`` halvings of every even direction, the multigrid's
    hierarchy), which keeps the smooth part of a correction and drops what the
    stiffness would amplify (rough error of amplitude e on a mesh of size h
    raises the residual like e / h^4 for this strain-gradient operator, so each
    level cuts that amplification about 16-fold).  Non-periodic directions
    only (identity otherwise)."""

    def __init__(self, mesh, levels=1):
        self.shape = tuple(mesh.nb)
        self.levels = int(levels)
        self.ops = []
        for d in range(3):
            n, kf = mesh.nelem[d], np.asarray(mesh.knots[d], dtype=np.float64)
            P = T = None
            if not mesh.periodic[d]:
                for _ in range(self.levels):
                    if n % 2 or n < 4:
                        break
                    p, t, kf = coarse_ops_1d(kf, n, with_knots=True)
                    P = p if P is None else P @ p
                    T = t if T is None else t @ T
                    n //= 2
            self.ops.append(None if P is None else (P, T))

    def __call__(self, c):
        x = np.asarray(c, dtype=np.float64).reshape(self.shape + (3,))
        for d, op in enumerate(self.ops):
            if op is not None:
                x = np.moveaxis(np.tensordot(op[1], x, axes=(1, d)), 0, d)
        for d, op in enumerate(self.ops):
            if op is not None:
                x = np.moveaxis(np.tensordot(op[0], x, axes=(1, d)), 0, d)
        return x.reshape(-1)


def smooth_candidate(A, c, H):
    """c - S R(H): one application of the multigrid's finest-level smoother
    (igap_gpu_pc_apply, the lagged preconditioner of the last solve) to the
    residual of the candidate state H.  It damps the rough part of a
    candidate's error, which the stiffness amplifies in the residual, and
    leaves the smooth part, where extrapolation and the network differ."""
    d = A.pc_apply(A.residual(H), which="smoother")
    return c - d


class LearnedGuess:
    """``guess(k, prev)`` for run_load_steps.  mode "learned": quadratic
    extrapolation + network, used when its residual norm is below the linear
    extrapolation's; "quadratic": the same candidate without the network,
    same acceptance; "linear": the linear extrapolation always."""

    def __init__(self, A, net=None, mode="learned", use_correction=True, filt=None, levels=None, smooth=False):
        if mode not in ("learned", "quadratic", "linear"):
            raise ValueError(f"unknown mode {mode!r}")
        if mode == "learned" and net is None:
            raise ValueError("mode 'learned' needs a network")
        self.A, self.net, self.mode = A, net, mode
        self.use_correction = use_correction
        self.filt = filt
        self.smooth = smooth
        # U_hist of a candidate state (transient: [c, U_n, U_{n-1}])
        self.levels = levels or (lambda c, prev: c)
        self.mask = A.dirichlet()[0].reshape(-1)
        self.keep = (self.mask == 0)
        self.g_prev = None
        self.log = []

    def set_previous_load(self, g):
        self.g_prev = np.array(g, dtype=np.float64)

    def __call__(self, k, prev):
        g = self.A.dirichlet()[1]
        rec = {"step": k}
        ext = 2.0 * prev[-1] - prev[-2] if len(prev) >= 2 else prev[-1].copy()
        cand = None
        if self.mode != "linear" and len(prev) >= 3 and self.g_prev is not None:
            dU1, dU2 = prev[-1] - prev[-2], prev[-2] - prev[-3]
            cand = ext + (dU1 - dU2) * self.keep
            if self.mode == "learned" and self.use_correction:
                x, s1, s2 = features(dU1, dU2, g - self.g_prev)
                corr = self.net.forward(x).astype(np.float64).reshape(-1)
                if self.filt is not None:
                    corr = self.filt(corr)
                cand = cand + corr * s2 * self.keep
        if self.smooth:
            ext = smooth_candidate(self.A, ext, self.levels(ext, prev))
            if cand is not None:
                cand = smooth_candidate(self.A, cand, self.levels(cand, prev))
        guess = ext
        if cand is not None:
            r_ext = float(np.linalg.norm(self.A.residual(self.levels(ext, prev))))
            r_c = float(np.linalg.norm(self.A.residual(self.levels(cand, prev))))
            rec.update(r_ext=r_ext, r_cand=r_c, accepted=bool(r_c < r_ext))
            if r_c < r_ext:
                guess = cand
        self.g_prev = g.copy()
        self.log.append(rec)
        return guess


# ------------------------------------------------------------------ driver
def run(nelem=(64, 16, 16), nsteps=200, ntrain=100, tip_total=0.5, epochs=60, lr=2e-3, prec="f32", pc_order=1,
        seed=0, device=0, verbose=True, rtol=1e-10, gen_rtol=1e-12, ksp_factor=0.1,
        arms=("plain", "linear", "quadratic", "learned"), pc="ilu0", coarse_filter=1, filter_targets=False,
        neval=None, smooth=False, residual_epochs=0, residual_lr=5e-4, residual_cosine=False):
    """Load-step comparison: steps 1..ntrain generate the training trajectory
    (FP64, linear-extrapolation guesses, solved to gen_rtol so that the
    differences the network learns are not dominated by solver tolerance),
    the network is trained on it, then steps ntrain+1..nsteps are solved once
    per arm from the same state:

      plain      the previous solution as the initial iterate (the reference's
                 behaviour), FP64 tangent and factor, the reference's KSP
                 tolerances: the plain FP64 GPU path
      linear     linear extrapolation                      } tangent and ILU
      quadratic  quadratic extrapolation, accepted against  } in ``prec``;
                 linear extrapolation                       } Krylov stops at
      learned    quadratic extrapolation + network, same    } ksp_factor x the
                 acceptance                                 } Newton threshold

    Every step of every arm is solved to ||R|| <= rtol ||R0||, R0 the residual
    of the previous solution under the step's load (the first residual of the
    reference's iteration), so a better initial iterate is not asked for more;
    the per-step wall time includes building the guess.

    coarse_filter: levels of the coarse-space projection of the learned term
    (0: none); filter_targets: train on targets projected the same way;
    neval: evaluate only the first neval steps after ntrain; smooth: every
    candidate of the extrapolating arms (and of the held-out test) is
    smoothed once by the multigrid's finest-level smoother before the
    residual comparison (pc "mg" only); residual_epochs: after the mean
    square training, this many epochs of residual_finetune (the linearised
    residual of the candidate as the loss) at residual_lr, decaying on a
    half cosine with residual_cosine."""
    if smooth and pc != "mg":
        raise ValueError("smooth needs pc='mg' (the multigrid's smoother)")
    from . import assm as am
    from .loadstep import run_load_steps, tip_bc

    def say(*a):
        if verbose:
            print(*a, flush=True)

    mesh = am.MeshIGABeam(L=(4, 1, 1), nelem=nelem)
    grid = mesh.nb
    mon = 1 if verbose else 0
    opts64 = am.default_solver_opts(tangent_prec="f64", pc=pc, pc_order=pc_order, monitor=mon)
    # prec "f32": FP32 tangent and preconditioner; "mixed": FP64 tangent (Newton
    # and Krylov operator), FP32 preconditioner
    optsx = am.default_solver_opts(tangent_prec="f32" if prec == "f32" else "f64", pc=pc, pc_order=pc_order,
                                   monitor=mon, pc_fp32=1 if prec in ("f32", "mixed") else 0)
    A = am.Assm(mesh, am.PhysGradElast(coef=am.SVK_COEF), am.BCClampedTip(tip=0.0), device=device)
    A.set_uhist(np.zeros(A.ndof))
    U, G = [A.get_uhist()[:A.ndof].copy()], [A.dirichlet()[1]]

    def keep(k, u, st):
        U.append(u.copy())
        G.append(A.dirichlet()[1])
        say(f"  gen step {k}: newton {st['newton_its']} ksp {st['ksp_its']} {st['wall_ms']:.0f} ms")

    t0 = time.perf_counter()
    gen = run_load_steps(A, nsteps, tip_total=tip_total, guess="extrapolate", opts=opts64, callback=keep,
                         max_steps=ntrain, start_rtol=gen_rtol)
    t_gen = time.perf_counter() - t0
    if len(gen) < ntrain or gen[-1]["converged"] <= 0:
        A.close()
        raise RuntimeError(f"training trajectory stopped at load step {len(gen)} of {ntrain} (not converged)")
    mask = A.dirichlet()[0]
    filt = CoarseFilter(mesh, levels=int(coarse_filter)) if coarse_filter else None
    samples = samples_from_trajectory(U, G, mask, filt=filt if filter_targets else None)
    net = WarmStartNet(grid, seed=seed, device=device)
    t0 = time.perf_counter()
    tr = train_selected(net, samples, epochs=epochs, lr=lr, seed=seed,
                        log=lambda ep, l, v: say(f"  epoch {ep}: loss {l:.3e} held-out {v:.3e}")
                        if ep % 10 == 0 else None)
    hist = tr["train_loss"]
    say(f"  held-out loss: best {tr['val_loss_best']:.3e}, zero correction {tr['val_loss_zero_correction']:.3e}; "
        f"correction {'on' if tr['correction_enabled'] else 'off'}")
    # the acceptance metric on the held-out steps: residual norm of the
    # quadratic extrapolation with and without the learned correction
    keepm = (mask.reshape(-1) == 0)
    ft = None
    if residual_epochs:
        # K at the last training state, in the context's own FP64 values
        A.set_uhist(U[-1])
        A.tangent_dev(None, None, _lib.IGAP_F64)
        cqs, s2s, bcs = [], [], []
        for i in range(len(samples)):
            k = i + 3
            dU1, dU2 = U[k - 1] - U[k - 2], U[k - 2] - U[k - 3]
            cqs.append(2.0 * U[k - 1] - U[k - 2] + (dU1 - dU2) * keepm)
            s2s.append(_rms(dU1 - dU2))
            bcs.append(lambda kk=k: A.set_bc(tip_bc(tip_total, kk / nsteps)))
        ft = residual_finetune(net, A, samples, cqs, s2s, keepm, bcs, epochs=residual_epochs, lr=residual_lr,
                               seed=seed, cosine=residual_cosine, log=lambda ep, l, v: say(f"  residual epoch {ep}: |r|^2 {l:.3e} "
                                                                  f"held-out ratio {v:.3g}"))
        del cqs
        say(f"  residual training: held-out ratio {ft['val_ratio_start']:.3g} -> {ft['val_ratio_best']:.3g}")
    corr_on = bool(tr["correction_enabled"] or (ft is not None and ft["val_ratio_best"] < 1.0))
    ratios_val = []
    if corr_on:
        for i in range(len(samples) - tr["val_samples"], len(samples)):
            k = i + 3
            A.set_bc(tip_bc(tip_total, k / nsteps))
            dU1, dU2 = U[k - 1] - U[k - 2], U[k - 2] - U[k - 3]
            cq = 2.0 * U[k - 1] - U[k - 2] + (dU1 - dU2) * keepm
            x, s1, s2 = features(dU1, dU2, G[k] - G[k - 1])
            corr = net.forward(x).astype(np.float64).reshape(-1)
            if filt is not None:
                corr = filt(corr)
            cl = cq + corr * s2 * keepm
            if smooth:
                cq = smooth_candidate(A, cq, cq)
                cl = smooth_candidate(A, cl, cl)
            rq = float(np.linalg.norm(A.residual(cq)))
            ratios_val.append(float(np.linalg.norm(A.residual(cl))) / rq if rq > 0 else 1.0)
    use_corr = bool(corr_on and ratios_val and np.median(ratios_val) < 1.0)
    say(f"  held-out residual ratio learned/quadratic: "
        f"{np.median(ratios_val) if ratios_val else float('nan'):.3g}; correction {'used' if use_corr else 'off'}")
    t_train = time.perf_counter() - t0

    logs, finals = {}, {}
    last = nsteps if neval is None else min(nsteps, ntrain + int(neval))
    for name in arms:
        A.set_bc(tip_bc(tip_total, ntrain / nsteps))
        if name == "plain":
            gs, opts, log, kf = "previous", opts64, [], None
        else:
            gs = LearnedGuess(A, net, mode=name, use_correction=use_corr, filt=filt, smooth=smooth)
            gs.set_previous_load(G[-1])
            opts, log, kf = optsx, gs.log, ksp_factor
        st = run_load_steps(A, nsteps, tip_total=tip_total, guess=gs, opts=opts, start=ntrain + 1, prev=U[-3:],
                            start_rtol=rtol, ksp_atol_factor=kf, max_steps=last,
                            callback=lambda k, u, x, nm=name: say(f"  {nm} step {k}: newton {x['newton_its']} "
                                                                  f"ksp {x['ksp_its']} {x['wall_ms']:.0f} ms"))
        if name == "plain":
            log.extend({"step": x["step"]} for x in st)
        for x, r in zip(st, log):
            r.update(newton_its=x["newton_its"], ksp_its=x["ksp_its"], wall_ms=x["wall_ms"], converged=x["converged"])
        logs[name] = log
        finals[name] = A.get_uhist()[:A.ndof].copy()
        say(f"  {name}: newton {sum(r['newton_its'] for r in log)} ksp {sum(r['ksp_its'] for r in log)} "
            f"wall {sum(r['wall_ms'] for r in log):.0f} ms")
    net.close()
    A.close()

    def agg(log):
        out = {"steps": len(log), "newton_its": int(sum(r["newton_its"] for r in log)),
               "ksp_its": int(sum(r["ksp_its"] for r in log)),
               "wall_ms_per_step": float(np.mean([r["wall_ms"] for r in log])),
               "all_converged": bool(all(r["converged"] > 0 for r in log) and len(log) == last - ntrain)}
        ratios = [r["r_cand"] / r["r_ext"] for r in log if "r_cand" in r and r["r_ext"] > 0]
        if ratios:
            out["accepted_fraction"] = float(np.mean([r.get("accepted", False) for r in log]))
            out["median_residual_ratio_candidate_over_linear"] = float(np.median(ratios))
        return out

    ref = finals.get("plain", finals[arms[0]])
    scale = max(1e-300, float(np.max(np.abs(ref))))
    summary = {
        "workload": f"static SVK + gradient beam {nelem[0]}x{nelem[1]}x{nelem[2]}, tip {tip_total} in {nsteps} load "
                    f"steps, train on 1..{ntrain}, evaluate {ntrain + 1}..{last}",
        "criterion": f"||R|| <= {rtol:g} ||R(U_prev; load_k)|| per step; preconditioner {pc}; arms other than "
                     f"plain: {prec} tangent and preconditioner, Krylov to {ksp_factor:g} x the Newton threshold",
        "grid": list(grid), "train_samples": tr["train_samples"], "val_samples": tr["val_samples"], "epochs": epochs,
        "loss_first": hist[0], "loss_last": hist[-1], "val_loss_best": tr["val_loss_best"],
        "val_loss_zero_correction": tr["val_loss_zero_correction"], "correction_enabled": tr["correction_enabled"],
        "val_residual_ratio_learned_over_quadratic": float(np.median(ratios_val)) if ratios_val else None,
        "correction_used": use_corr, "coarse_filter_levels": int(coarse_filter), "filter_targets": bool(filter_targets),
        "smooth": bool(smooth),
        "residual_training": None if ft is None else {k: v for k, v in ft.items() if k != "train_residual"},
        "t_generate_s": t_gen, "t_train_s": t_train,
        "generate": {"newton_its": int(sum(x["newton_its"] for x in gen))},
    }
    for name in arms:
        summary[name] = agg(logs[name])
        summary[name]["final_rel_diff_vs_plain"] = float(np.max(np.abs(finals[name] - ref)) / scale)
    if "plain" in arms:
        for name in arms:
            summary[name]["speedup_vs_plain"] = summary["plain"]["wall_ms_per_step"] / summary[name]["wall_ms_per_step"]
    return summary


def run_transient(nelem=(128, 32, 32), nsteps=200, ntrain=100, dt=1e-3, epochs=60, lr=2e-3, prec="f32", pc="jacobi",
                  seed=0, device=0, verbose=True, rtol=1e-10, gen_rtol=1e-12, ksp_factor=0.1,
                  arms=("plain", "linear", "quadratic", "learned"), coarse_filter=True):
    """The same comparison on time steps of gradelasttime (the transient
    workload: L = 4:1:1 beam, zero displacement and the 8-mode seeded initial
    velocity, the reference's parameters with time step dt): steps
    1..ntrain generate the training trajectory, steps ntrain+1..nsteps are
    solved once per arm from the same two time levels.  plain = the previous
    state as initial iterate (the reference's), the reference's KSP
    tolerances; the other arms stop the Krylov solve at ksp_factor x the
    Newton threshold.  Tangent and preconditioner as ``prec`` / ``pc`` in every
    arm.  Features: the last increment, the last second difference, and a zero
    load channel (no load varies in time)."""
    from . import assm as am
    from .loadstep import initial_velocity_modes, run_time_steps

    def say(*a):
        if verbose:
            print(*a, flush=True)

    mesh = am.MeshIGABeam(L=(4, 1, 1), nelem=nelem)
    par = list(am.PhysGradElastTime().par)
    par[7] = dt
    A = am.Assm(mesh, am.PhysGradElastTime(par=par), am.BCDefault(), device=device)
    n = A.ndof
    na, nb = mesh.na, mesh.nb
    Xo = mesh.X.reshape(na[0], na[1], na[2], 3)[:nb[0], :nb[1], :nb[2]].reshape(-1, 3)
    u1, u0 = initial_velocity_modes(Xo, (4, 1, 1), dt)
    mon = 1 if verbose else 0
    opts = am.default_solver_opts(tangent_prec=prec, pc=pc, pc_order=1, monitor=0)
    A.set_uhist(np.concatenate([u1, u1, u0]))
    U = [u0.copy(), u1.copy()]

    def keep(k, H, st):
        U.append(H[0].copy())
        say(f"  gen step {k}: newton {st['newton_its']} ksp {st['ksp_its']} {st['wall_ms']:.0f} ms")

    t0 = time.perf_counter()
    gen = run_time_steps(A, ntrain, opts=opts, guess="extrapolate", callback=keep, start_rtol=gen_rtol)
    t_gen = time.perf_counter() - t0
    if len(gen) < ntrain or gen[-1]["converged"] <= 0:
        A.close()
        raise RuntimeError(f"training trajectory stopped at time step {len(gen)} of {ntrain} (not converged)")
    mask = A.dirichlet()[0]
    G = [np.zeros(n) for _ in U]
    samples = samples_from_trajectory(U, G, mask)
    net = WarmStartNet(nb, seed=seed, device=device)
    t0 = time.perf_counter()
    tr = train_selected(net, samples, epochs=epochs, lr=lr, seed=seed,
                        log=lambda ep, l, v: say(f"  epoch {ep}: loss {l:.3e} held-out {v:.3e}")
                        if ep % 10 == 0 else None)
    keepm = (mask.reshape(-1) == 0)
    filt = CoarseFilter(mesh, levels=int(coarse_filter)) if coarse_filter else None
    ratios_val = []
    if tr["correction_enabled"]:
        for i in range(len(samples) - tr["val_samples"], len(samples)):
            k = i + 3
            dU1, dU2 = U[k - 1] - U[k - 2], U[k - 2] - U[k - 3]
            cq = 2.0 * U[k - 1] - U[k - 2] + (dU1 - dU2) * keepm
            x, s1, s2 = features(dU1, dU2, np.zeros(n))
            corr = net.forward(x).astype(np.float64).reshape(-1)
            if filt is not None:
                corr = filt(corr)
            cl = cq + corr * s2 * keepm
            hist = np.concatenate([U[k - 1], U[k - 2]])
            rq = float(np.linalg.norm(A.residual(np.concatenate([cq, hist]))))
            ratios_val.append(float(np.linalg.norm(A.residual(np.concatenate([cl, hist])))) / rq if rq > 0 else 1.0)
    use_corr = bool(corr_on and ratios_val and np.median(ratios_val) < 1.0)
    say(f"  held-out residual ratio learned/quadratic: "
        f"{np.median(ratios_val) if ratios_val else float('nan'):.3g}; correction {'used' if use_corr else 'off'}")
    t_train = time.perf_counter() - t0

    def levels(c, prev):
        return np.concatenate([c, prev[-1], prev[-2]])

    logs, finals = {}, {}
    for name in arms:
        A.set_uhist(np.concatenate([U[-1], U[-1], U[-2]]))
        o = am.default_solver_opts(tangent_prec=prec, pc=pc, pc_order=1, monitor=mon)
        if name == "plain":
            gs, log, kf = "previous", [], None
        else:
            gs = LearnedGuess(A, net, mode=name, use_correction=use_corr, filt=filt, levels=levels)
            gs.set_previous_load(np.zeros(n))
            log, kf = gs.log, ksp_factor
        st = run_time_steps(A, nsteps - ntrain, opts=o, guess=gs, start_rtol=rtol, ksp_atol_factor=kf,
                            callback=lambda k, H, x, nm=name: say(f"  {nm} step {ntrain + k}: newton "
                                                                  f"{x['newton_its']} ksp {x['ksp_its']} "
                                                                  f"{x['wall_ms']:.0f} ms"))
        if name == "plain":
            log.extend({"step": x["step"]} for x in st)
        for x, r in zip(st, log):
            r.update(newton_its=x["newton_its"], ksp_its=x["ksp_its"], wall_ms=x["wall_ms"], converged=x["converged"])
        logs[name] = log
        finals[name] = A.get_uhist()[:n].copy()
        say(f"  {name}: newton {sum(r['newton_its'] for r in log)} ksp {sum(r['ksp_its'] for r in log)} "
            f"wall {sum(r['wall_ms'] for r in log):.0f} ms")
    net.close()
    A.close()

    def agg(log):
        out = {"steps": len(log), "newton_its": int(sum(r["newton_its"] for r in log)),
               "ksp_its": int(sum(r["ksp_its"] for r in log)),
               "wall_ms_per_step": float(np.mean([r["wall_ms"] for r in log])),
               "all_converged": bool(all(r["converged"] > 0 for r in log) and len(log) == nsteps - ntrain)}
        ratios = [r["r_cand"] / r["r_ext"] for r in log if "r_cand" in r and r["r_ext"] > 0]
        if ratios:
            out["accepted_fraction"] = float(np.mean([r.get("accepted", False) for r in log]))
            out["median_residual_ratio_candidate_over_linear"] = float(np.median(ratios))
        return out

    ref = finals.get("plain", finals[arms[0]])
    scale = max(1e-300, float(np.max(np.abs(ref))))
    summary = {
        "workload": f"gradelasttime beam {nelem[0]}x{nelem[1]}x{nelem[2]}, dt {dt:g}, 8-mode initial velocity, "
                    f"{nsteps} time steps, train on 1..{ntrain}, evaluate {ntrain + 1}..{nsteps}",
        "criterion": f"||R|| <= {rtol:g} ||R(U_n)|| per step; tangent {prec}, preconditioner {pc}; arms other than "
                     f"plain: Krylov to {ksp_factor:g} x the Newton threshold",
        "grid": list(nb), "train_samples": tr["train_samples"], "val_samples": tr["val_samples"], "epochs": epochs,
        "loss_first": tr["train_loss"][0], "loss_last": tr["train_loss"][-1], "val_loss_best": tr["val_loss_best"],
        "val_loss_zero_correction": tr["val_loss_zero_correction"],
        "val_residual_ratio_learned_over_quadratic": float(np.median(ratios_val)) if ratios_val else None,
        "correction_used": use_corr, "t_generate_s": t_gen, "t_train_s": t_train,
    }
    for name in arms:
        summary[name] = agg(logs[name])
        summary[name]["final_rel_diff_vs_plain"] = float(np.max(np.abs(finals[name] - ref)) / scale)
    if "plain" in arms:
        for name in arms:
            summary[name]["speedup_vs_plain"] = summary["plain"]["wall_ms_per_step"] / summary[name]["wall_ms_per_step"]
    return summary


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--nelem", default="64,16,16")
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--train", type=int, default=100)
    ap.add_argument("--tip", type=float, default=0.5)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--prec", default="f32", choices=["f64", "f32", "mixed"])
    ap.add_argument("--pc-order", type=int, default=1)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--gen-rtol", type=float, default=1e-12)
    ap.add_argument("--ksp-factor", type=float, default=0.1)
    ap.add_argument("--arms", default="plain,linear,quadratic,learned")
    ap.add_argument("--pc", default="ilu0", choices=["ilu0", "mg", "jacobi"])
    ap.add_argument("--coarse-filter", type=int, default=1, help="levels of the coarse-space projection (0: none)")
    ap.add_argument("--filter-targets", type=int, default=0)
    ap.add_argument("--neval", type=int, default=0, help="evaluate only this many steps after training (0: all)")
    ap.add_argument("--smooth", type=int, default=0, help="smooth every candidate once before the residual test")
    ap.add_argument("--residual-epochs", type=int, default=0, help="epochs of training on the candidate's residual")
    ap.add_argument("--residual-lr", type=float, default=5e-4)
    ap.add_argument("--residual-cosine", type=int, default=0, help="half-cosine decay of the residual training's lr")
    ap.add_argument("--transient", action="store_true", help="time steps of gradelasttime instead of load steps")
    ap.add_argument("--dt", type=float, default=1e-3)
    a = ap.parse_args()
    nel = tuple(int(x) for x in a.nelem.split(","))
    if a.transient:
        s = run_transient(nel, a.steps, a.train, a.dt, a.epochs, a.lr, a.prec if a.prec != "mixed" else "f64",
                          a.pc if a.pc != "ilu0" else "jacobi", rtol=a.rtol, gen_rtol=a.gen_rtol,
                          ksp_factor=a.ksp_factor, arms=tuple(a.arms.split(",")), coarse_filter=a.coarse_filter)
        print(json.dumps(s), flush=True)
        return
    s = run(nel, a.steps, a.train, a.tip, a.epochs, a.lr, a.prec, a.pc_order, rtol=a.rtol, gen_rtol=a.gen_rtol,
            ksp_factor=a.ksp_factor, arms=tuple(a.arms.split(",")), pc=a.pc, coarse_filter=a.coarse_filter,
            filter_targets=bool(a.filter_targets), neval=a.neval or None, smooth=bool(a.smooth),
            residual_epochs=a.residual_epochs, residual_lr=a.residual_lr, residual_cosine=bool(a.residual_cosine))
    print(json.dumps(s), flush=True)


if __name__ == "__main__":
    main()
