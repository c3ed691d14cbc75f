"""Learned Newton warm start on the control-point grid.

A 3-D convolutional network (include/igap_ws.h: 4 layers of 3x3x3
convolutions, 16 channels, BF16 matrix-core products with FP32 accumulation,
trained on the GPU with Adam) predicts a correction to the linear
extrapolation of the next load-step solution from

  * the last two solution increments  dU_{k-1} = U_{k-1} - U_{k-2},
    dU_{k-2} = U_{k-2} - U_{k-3}                            (6 channels)
  * the load increment  dG_k = g_k - g_{k-1}  (Dirichlet data)  (3 channels)

all divided by s_k = rms(dU_{k-1}).  The target is the second difference
(U_k - 2 U_{k-1} + U_{k-2}) / s_k, zero on Dirichlet dofs.  The guess
U_ext + s_k net(x) is used only if its residual norm is below the residual norm
of the extrapolation U_ext = 2 U_{k-1} - U_{k-2}, which otherwise stays the
initial iterate; Newton then solves to the same tolerance either way.

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
def features(dU1, dU2, dG):
    """x[v][9] (float32) and the scale s = rms(dU1)."""
    s = float(np.sqrt(np.mean(dU1 * dU1)))
    s = s if s > 0 else 1.0
    x = np.concatenate([dU1.reshape(-1, 3), dU2.reshape(-1, 3), dG.reshape(-1, 3)], axis=1) / s
    return x.astype(np.float32), s


def samples_from_trajectory(U, G, mask):
    """Training pairs from solutions U[0..K] and Dirichlet data G[0..K]:
    (x_k, target_k) for k = 3..K."""
    out = []
    keep = (mask.reshape(-1) == 0)
    for k in range(3, len(U)):
        dU1, dU2 = U[k - 1] - U[k - 2], U[k - 2] - U[k - 3]
        x, s = features(dU1, dU2, G[k] - G[k - 1])
        t = ((U[k] - 2 * U[k - 1] + U[k - 2]) * keep) / s
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


class LearnedGuess:
    """``guess(k, prev)`` for run_load_steps: extrapolation, replaced by the
    network's guess when that has the smaller residual norm."""

    def __init__(self, A, net, use_net=True):
        self.A, self.net, self.use_net = A, net, use_net
        self.mask = A.dirichlet()[0].reshape(-1)
        self.keep = (self.mask == 0)
        self.g_prev = None
        self.log = []

    def set_previous_load(self, g):
        self.g_prev = np.array(g, dtype=np.float64)

    def __call__(self, k, prev):
        g = self.A.dirichlet()[1]
        rec = {"step": k}
        if len(prev) >= 2:
            ext = 2.0 * prev[-1] - prev[-2]
        else:
            ext = prev[-1].copy()
        guess = ext
        if self.use_net and len(prev) >= 3 and self.g_prev is not None:
            x, s = features(prev[-1] - prev[-2], prev[-2] - prev[-3], g - self.g_prev)
            c = self.net.forward(x).astype(np.float64).reshape(-1) * s * self.keep
            cand = ext + c
            r_ext = float(np.linalg.norm(self.A.residual(ext)))
            r_net = float(np.linalg.norm(self.A.residual(cand)))
            rec.update(r_ext=r_ext, r_net=r_net, accepted=bool(r_net < r_ext))
            if r_net < r_ext:
                guess = cand
        self.g_prev = g.copy()
        self.log.append(rec)
        return guess


# ------------------------------------------------------------------ driver
def run(nelem=(64, 16, 16), nsteps=200, ntrain=100, tip_total=0.5, epochs=60, lr=2e-3, prec="f64", pc_order=0,
        seed=0, device=0, verbose=True, atol=1e-10):
    """Config-4 scenario: steps 1..ntrain generate the training trajectory
    (extrapolation guesses), the network is trained on it, then steps
    ntrain+1..nsteps are solved twice from the same state: with the learned
    guess (accepted against extrapolation) and with extrapolation only."""
    from . import assm as am
    from .loadstep import run_load_steps, tip_bc

    def say(*a):
        if verbose:
            print(*a, flush=True)

    mesh = am.MeshIGABeam(L=(4, 1, 1), nelem=nelem)
    grid = mesh.nb
    # the reference's criterion ||R|| < atol (src/interface.c:220): a relative
    # test against the first residual of a step would demand more of a better
    # initial guess
    opts = am.default_solver_opts(snes_atol=atol, snes_rtol=0.0, tangent_prec=prec, pc_order=pc_order)
    A = am.Assm(mesh, am.PhysGradElast(coef=am.SVK_COEF), am.BCClampedTip(tip=0.0), device=device)
    A.set_uhist(np.zeros(A.ndof))
    U, G = [A.get_uhist()[:A.ndof].copy()], [A.dirichlet()[1]]

    def keep(k, u, st):
        U.append(u.copy())
        G.append(A.dirichlet()[1])
        say(f"  gen step {k}: newton {st['newton_its']} ksp {st['ksp_its']} {st['wall_ms']:.0f} ms")

    t0 = time.perf_counter()
    gen = run_load_steps(A, nsteps, tip_total=tip_total, guess="extrapolate", opts=opts, callback=keep,
                         max_steps=ntrain)
    t_gen = time.perf_counter() - t0
    mask = A.dirichlet()[0]
    samples = samples_from_trajectory(U, G, mask)
    net = WarmStartNet(grid, seed=seed, device=device)
    t0 = time.perf_counter()
    hist = train(net, samples, epochs=epochs, lr=lr, seed=seed,
                 log=lambda ep, l: say(f"  epoch {ep}: loss {l:.3e}") if ep % 10 == 0 else None)
    t_train = time.perf_counter() - t0

    results = {}
    finals = {}
    for name, use in (("learned", True), ("extrapolation", False)):
        A.set_bc(tip_bc(tip_total, ntrain / nsteps))
        gs = LearnedGuess(A, net, use_net=use)
        gs.set_previous_load(G[-1])
        st = run_load_steps(A, nsteps, tip_total=tip_total, guess=gs, opts=opts, start=ntrain + 1, prev=U[-3:])
        for s, r in zip(st, gs.log):
            r.update(newton_its=s["newton_its"], ksp_its=s["ksp_its"], wall_ms=s["wall_ms"], converged=s["converged"])
        results[name] = gs.log
        finals[name] = A.get_uhist()[:A.ndof].copy()
        say(f"  {name}: newton {sum(r['newton_its'] for r in gs.log)} wall {sum(r['wall_ms'] for r in gs.log):.0f} ms")
    net.close()
    A.close()

    def agg(log):
        return {"steps": len(log), "newton_its": int(sum(r["newton_its"] for r in log)),
                "ksp_its": int(sum(r["ksp_its"] for r in log)),
                "wall_ms_per_step": float(np.mean([r["wall_ms"] for r in log])),
                "all_converged": bool(all(r["converged"] > 0 for r in log))}

    lr_log = results["learned"]
    ratios = [r["r_net"] / r["r_ext"] for r in lr_log if "r_net" in r and r["r_ext"] > 0]
    summary = {
        "workload": f"static SVK + gradient beam {nelem[0]}x{nelem[1]}x{nelem[2]}, tip {tip_total} in {nsteps} load "
                    f"steps, train on 1..{ntrain}, evaluate {ntrain + 1}..{nsteps}",
        "grid": list(grid), "train_samples": len(samples), "epochs": epochs,
        "loss_first": hist[0], "loss_last": hist[-1],
        "t_generate_s": t_gen, "t_train_s": t_train,
        "generate": {"newton_its": int(sum(s["newton_its"] for s in gen))},
        "learned": agg(lr_log), "extrapolation": agg(results["extrapolation"]),
        "accepted_fraction": float(np.mean([r.get("accepted", False) for r in lr_log])),
        "median_residual_ratio_net_over_extrapolation": float(np.median(ratios)) if ratios else None,
        "final_rel_diff": float(np.max(np.abs(finals["learned"] - finals["extrapolation"]))
                                / max(1e-300, np.max(np.abs(finals["extrapolation"])))),
    }
    return summary


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--nelem", default="64,16,16")
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--train", type=int, default=100)
    ap.add_argument("--tip", type=float, default=0.5)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--prec", default="f64")
    ap.add_argument("--pc-order", type=int, default=0)
    a = ap.parse_args()
    nel = tuple(int(x) for x in a.nelem.split(","))
    s = run(nel, a.steps, a.train, a.tip, a.epochs, a.lr, a.prec, a.pc_order)
    print(json.dumps(s), flush=True)


if __name__ == "__main__":
    main()
