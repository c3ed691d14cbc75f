"""Benchmark: elements assembled per second (residual + tangent) on one MI355X.

Workload (default): the 64x16x16-element quadratic C1 B-spline beam patch
(L = 4:1:1, 21,384 control points, 64,152 dofs, 20.6M CSR nonzeros), static
gradient elasticity with St Venant-Kirchhoff lambda = mu = 1 and strain-
gradient length 0.1, clamped at x = 0 with a tip displacement on x = L,
4x4x4 Gauss points.  A step is one full residual (volume + Neumann +
Dirichlet rows) plus one full tangent (element tangents on the matrix cores,
deterministic CSR gather, MatZeroRowsColumns) in FP64, on a smooth synthetic
displacement resident in HBM.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--nelem 64,16,16]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# algorithmic flops per element of the tangent: 64 Gauss points x 9 (i,j)
# blocks x (D R: 2*10*10*27 + R^T (D R): 2*10*27*27)
F_K = 64 * 9 * (2 * 10 * 10 * 27 + 2 * 10 * 27 * 27)
# residual: u-jet (2*27*3*10 per level) + projection (2*27*3*10) per Gauss point
F_R_STATIC = 64 * (1620 + 1620)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nelem", default="64,16,16")
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--newton-steps", type=int, default=3,
                    help="load steps of the Newton ramp timed after the assembly bench (0: skip)")
    ap.add_argument("--probe", type=int, default=0,
                    help="profiling run for rocprofv3 --pmc: one 256 MiB device copy (the HBM "
                         "counter calibration) then N residual+tangent calls; prints no JSON")
    ap.add_argument("--pmc-json", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="turn the counter_collection.csv of a FETCH_SIZE and a WRITE_SIZE "
                         "--probe run into profiles/tangent_elem_<nelem>_pmc.json")
    return ap.parse_args()


COPY_BYTES = 1 << 28   # the calibration copy of --probe


def probe(nel, calls):
    import torch
    import igap4_amd as ig
    src = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    mesh = ig.MeshIGABeam(L=(4.0, 1.0, 1.0), nelem=nel)
    with ig.Assm(mesh, ig.PhysGradElast(coef=ig.SVK_COEF), ig.BCClampedTip(tip=0.05)) as A:
        A.set_uhist(ig.smooth_state(mesh, amp=0.01))
        for _ in range(calls):
            A.residual_dev(None, None)
            A.tangent_dev(None, None, 0)
        A.synchronize()
    print(f"probe done: copy {COPY_BYTES} B, {calls} x (residual + tangent) on {nel}", flush=True)


def pmc_json(nel, fetch_csv, write_csv):
    """HBM bytes per launch of the assembly kernels: FETCH_SIZE and WRITE_SIZE
    (KiB per dispatch, averaged over the dispatches of a kernel), each divided
    by what the same counter reads per byte of the calibration copy (its one
    256 MiB dispatch)."""
    import csv

    def per_kernel(path, counter):
        acc = {}
        with open(path) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter:
                    continue
                k = r["Kernel_Name"]
                acc.setdefault(k, {}).setdefault(r["Dispatch_Id"], 0.0)
                acc[k][r["Dispatch_Id"]] += float(r["Counter_Value"])
        return {k: (sum(v.values()) / len(v), max(v.values())) for k, v in acc.items()}

    out = {"workload": f"{nel[0]}x{nel[1]}x{nel[2]} static beam, k_tangent_elem<double,1>",
           "copy_bytes": COPY_BYTES, "kernels": {}}
    total = 0.0
    for name, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        d = per_kernel(path, name)
        # the calibration is the single 256 MiB copy dispatch (the largest of the
        # copy kernels; other copies of the run are a few bytes)
        copy = [mx for k, (_, mx) in d.items() if "copy" in k.lower() or "elementwise" in k.lower()]
        tang = [mean for k, (mean, _) in d.items() if "k_tangent_elem<double, 1>" in k]
        if not copy or not tang:
            raise SystemExit(f"{path}: no copy or tangent dispatch found for {name}")
        factor = max(copy) * 1024.0 / COPY_BYTES
        b = tang[0] * 1024.0 / factor
        out[name] = {"tangent_kib_per_dispatch": tang[0], "copy_kib_per_dispatch": max(copy),
                     "calibration_factor": factor, "tangent_bytes_per_dispatch": b}
        total += b
        # every assembly kernel of the probe, same calibration
        for k, (v, _) in d.items():
            if "igap::" not in k:
                continue
            short = k.split("(")[0].replace("void ", "").strip()
            e = out["kernels"].setdefault(short, {})
            e[name + "_kib_per_dispatch"] = v
            e[name + "_bytes_per_dispatch"] = v * 1024.0 / factor
    out["tangent_elem_bytes_per_launch"] = total
    path = os.path.join(ROOT, "profiles", f"tangent_elem_{nel[0]}x{nel[1]}x{nel[2]}_pmc.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


def traffic_from_profiles(nel):
    """HBM bytes per tangent-kernel launch from the committed rocprofv3 PMC
    summary (FETCH_SIZE / WRITE_SIZE, KiB per dispatch, scaled by the copy-kernel
    calibration recorded there), or None when no summary is committed."""
    path = os.path.join(ROOT, "profiles", f"tangent_elem_{nel[0]}x{nel[1]}x{nel[2]}_pmc.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            d = json.load(f)
        return float(d["tangent_elem_bytes_per_launch"])
    except Exception:
        return None


def measured_bytes(nel, kernel):
    """HBM bytes per dispatch (FETCH_SIZE + WRITE_SIZE, calibrated) of one
    kernel from the committed PMC summary of this workload, or None."""
    path = os.path.join(ROOT, "profiles", f"tangent_elem_{nel[0]}x{nel[1]}x{nel[2]}_pmc.json")
    try:
        with open(path) as f:
            k = json.load(f)["kernels"][kernel]
        return float(k["FETCH_SIZE_bytes_per_dispatch"] + k["WRITE_SIZE_bytes_per_dispatch"])
    except Exception:
        return None


def kernel_rooflines(nel, nelem, nnz, avg, avg32):
    """Every assembly kernel against the ceiling that bounds it.  Tangent:
    F_K per element on the matrix cores; residual: F_R per element on the
    FP64 vector ALUs (counted flops only: basis, geometry and the pointwise
    physics come on top); gathers: the element matrices read once and the
    CSR values written once (8 or 4 bytes each), with the CSR values alone
    and the measured HBM bytes beside them."""
    def flop(kernel_ms, f, peak, bound):
        a = f * nelem / (kernel_ms * 1e-3) / 1e12
        return {"bound": bound, "achieved": a, "peak": peak, "unit": "TFLOP/s", "frac": a / peak, "ms": kernel_ms}

    def hbm(kernel_ms, nbytes, name, csr_bytes):
        a = nbytes / (kernel_ms * 1e-3) / 1e9
        mb = measured_bytes(nel, name)
        return {"bound": "hbm", "achieved": a, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": a / PEAK_HBM_GBS,
                "ms": kernel_ms, "algorithmic_bytes": nbytes, "csr_value_bytes": csr_bytes,
                "csr_value_GBs": csr_bytes / (kernel_ms * 1e-3) / 1e9, "measured_bytes": mb,
                "measured_GBs": None if mb is None else mb / (kernel_ms * 1e-3) / 1e9}

    return {
        "k_tangent_elem<double,1>": flop(avg[1], F_K, PEAK_FP64_TFLOPS, "mfma"),
        "k_tangent_elem<float,1>": flop(avg32[1], F_K, PEAK_FP32_TFLOPS, "mfma"),
        "k_residual_elem<1>": flop(avg[0], F_R_STATIC, PEAK_FP64_TFLOPS, "valu"),
        "k_tangent_gather_rows<double,double>": hbm(avg[2], 8 * (6561 * nelem + nnz),
                                                    "igap::k_tangent_gather_rows<double, double>", 8 * nnz),
        "k_tangent_gather<float,float>": hbm(avg32[2], 4 * (6561 * nelem + nnz), "igap::k_tangent_gather<float, float>",
                                             4 * nnz),
    }


def cpu_baseline(threads, nel, nx=None):
    """The oracle (CPU restatement with the reference's loop nest and naive
    81x81x100 contraction) on the same workload: one residual and one tangent
    over the elements x < nx (default: all) of the same patch, physics and
    state (about 14 s with 16 threads at 64x16x16); elements/s (R+K)."""
    from oracle import oracle as O
    nx = nel[0] if nx is None else min(nx, nel[0])
    m = O.Mesh(nel, L=(4, 1, 1))
    ph = O.make_phys(0, coef=O.SVK_COEF)
    bc = O.clamped_tip_bc(tip=0.05)
    U = O.smooth_state(m, amp=0.01)
    pat = O.csr_pattern(m)
    box = [0, nx, 0, nel[1], 0, nel[2]]
    t0 = time.perf_counter()
    O.residual(m, ph, bc, U, box=box, nthreads=threads)
    t1 = time.perf_counter()
    O.tangent(m, ph, bc, U, pattern=pat, box=box, nthreads=threads)
    t2 = time.perf_counter()
    ns = nx * nel[1] * nel[2]
    return {"value": ns / (t2 - t0), "unit": "elements/s", "cores": threads, "kind": "port",
            "sample": f"one residual + one tangent over the {ns} elements x < {nx} of the {nel[0]}x{nel[1]}x{nel[2]} "
                      f"patch (same physics and state, FP64, {threads} OpenMP threads): t_R = {t1 - t0:.2f} s, "
                      f"t_K = {t2 - t1:.2f} s"}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only: the whole patch lives in one "
                                   "MI355X's HBM and there is no multi-GPU path; run with --gpus 1"}))
        sys.exit(2)
    nel = tuple(int(x) for x in a.nelem.split(","))
    if a.pmc_json:
        pmc_json(nel, *a.pmc_json)
        return
    if a.probe:
        probe(nel, a.probe)
        return
    import igap4_amd as ig

    mesh = ig.MeshIGABeam(L=(4.0, 1.0, 1.0), nelem=nel)
    phys = ig.PhysGradElast(coef=ig.SVK_COEF)
    bc = ig.BCClampedTip(tip=0.05)
    U = ig.smooth_state(mesh, amp=0.01)

    A = ig.Assm(mesh, phys, bc, device=0)
    A.set_uhist(U)
    nelem = A.nel

    def step(prec):
        A.residual_dev(None, None)
        A.tangent_dev(None, None, prec)

    for _ in range(a.warmup):
        step(0)
        step(1)
    A.synchronize()
    # ---- timed region: K steps of FP64 R+K
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step(0)
    A.synchronize()
    t1 = time.perf_counter()
    # ---- same with the FP32 matrix-core tangent (residual stays FP64)
    for _ in range(a.steps):
        step(1)
    A.synchronize()
    t2 = time.perf_counter()
    ms64 = (t1 - t0) / a.steps * 1e3
    ms32 = (t2 - t1) / a.steps * 1e3

    # ---- per-kernel durations with HIP events on the library's stream
    A.timing(True)
    A.timing_reset()
    for _ in range(3):
        step(0)
    kt = {w: A.timing_get(w) for w in range(5)}
    A.timing_reset()
    for _ in range(3):
        step(1)
    kt32 = {w: A.timing_get(w) for w in range(5)}
    A.timing(False)
    avg = {w: kt[w][0] / max(1, kt[w][1]) for w in kt}
    avg32 = {w: kt32[w][0] / max(1, kt32[w][1]) for w in kt32}
    tk = avg[1] * 1e-3
    achieved = F_K * nelem / tk / 1e12
    traffic = traffic_from_profiles(nel)

    line = {
        "metric": "elements assembled/s (R+K)",
        "value": nelem / (ms64 * 1e-3),
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (smooth displacement field on the igabeam patch)",
        "config": {"workload": f"static gradient elasticity, {nel[0]}x{nel[1]}x{nel[2]}-element quadratic "
                               f"C1 B-spline beam 4:1:1, SVK lambda=mu=1 + gradient l=0.1, clamped + tip "
                               f"Dirichlet, 4x4x4 Gauss points",
                   "nelem": nelem, "ndof": A.ndof, "nnz": A.nnz},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                     "kernel": "k_tangent_elem<double,1>", "kernel_ms": avg[1]},
        "kernels_ms": {"residual_elem": avg[0], "tangent_elem_f64": avg[1], "tangent_gather_f64": avg[2],
                       "residual_call": avg[3], "tangent_call_f64": avg[4],
                       "tangent_elem_f32": avg32[1], "tangent_gather_f32": avg32[2],
                       "tangent_call_f32": avg32[4]},
        "kernel_rooflines": kernel_rooflines(nel, nelem, A.nnz, avg, avg32),
        "fp32_tangent": {"value": nelem / (ms32 * 1e-3), "ms_per_step": ms32,
                         "tangent_elem_tflops": F_K * nelem / (avg32[1] * 1e-3) / 1e12,
                         "peak": PEAK_FP32_TFLOPS},
    }
    A.close()
    if a.newton_steps > 0:
        line["newton_load_step"] = newton_load_steps(nel, a.newton_steps)
    if a.cpu_baseline:
        th = a.cpu_threads or min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(th, nel)
    print(json.dumps(line), flush=True)


def newton_load_steps(nel, k, nsteps=20, tip=0.05):
    """The second measure of the workload: wall time per Newton load step of
    the clamped-tip ramp (tip 0.05 in 20 load steps) to ||R||/||R0|| <= 1e-10
    from the previous solution (the reference's start), device Newton with
    GMRES preconditioned by the geometric multigrid cycle (set up once per
    load step: every step here takes 4 Newton iterations), FP64 tangent; the
    first k steps of the ramp (the first includes one-time set-up)."""
    import igap4_amd as ig
    from igap4_amd.loadstep import run_load_steps
    mesh = ig.MeshIGABeam(L=(4.0, 1.0, 1.0), nelem=nel)
    with ig.Assm(mesh, ig.PhysGradElast(coef=ig.SVK_COEF), ig.BCClampedTip(tip=0.0)) as A:
        A.set_uhist(np.zeros(A.ndof))
        opts = ig.assm.default_solver_opts(snes_rtol=1e-10, snes_atol=1e-30, tangent_prec="f64", pc="mg", pc_order=1,
                                           pc_lag=4)
        st = run_load_steps(A, nsteps, tip_total=tip, opts=opts, max_steps=k)
    w = [s["wall_ms"] for s in st]
    return {"ms_per_step_median": float(np.median(w)), "ms_per_step": w, "steps_timed": len(st),
            "newton_its": [s["newton_its"] for s in st], "ksp_its": [s["ksp_its"] for s in st],
            "converged": bool(all(s["converged"] > 0 for s in st)),
            "solver": "NEWTONLS+BT, GMRES(200) right-PC, geometric multigrid on the B-spline hierarchy "
                      "(Chebyshev/point-block-Jacobi smoothing, dense coarsest LU) set up every 4th Newton "
                      "iteration, FP64 tangent, ||R|| <= 1e-10 ||R0||",
            "ramp": f"tip {tip} in {nsteps} load steps"}


if __name__ == "__main__":
    main()
