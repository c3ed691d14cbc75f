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
    return ap.parse_args()


def traffic_from_profiles():
    """HBM bytes per tangent-kernel launch from the committed rocprofv3 PMC
    summary (FETCH_SIZE / WRITE_SIZE, KiB per dispatch, scaled by the copy-kernel
    calibration recorded there), or None when no summary is committed."""
    path = os.path.join(ROOT, "profiles", "tangent_elem_64x16x16_pmc.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            d = json.load(f)
        return float(d["tangent_elem_bytes_per_launch"])
    except Exception:
        return None


def cpu_baseline(threads):
    """The oracle (CPU restatement with the reference's loop nest and naive
    contraction) on the 16x4x4 beam of the same physics: elements/s (R+K)."""
    from oracle import oracle as O
    nel = (16, 4, 4)
    m = O.Mesh(nel, L=(4, 1, 1))
    ph = O.make_phys(0, coef=O.SVK_COEF)
    bc = O.clamped_tip_bc(tip=0.05)
    U = O.smooth_state(m, amp=0.01)
    pat = O.csr_pattern(m)
    t0 = time.perf_counter()
    O.residual(m, ph, bc, U, nthreads=threads)
    t1 = time.perf_counter()
    O.tangent(m, ph, bc, U, pattern=pat, nthreads=threads)
    t2 = time.perf_counter()
    nelem = m.nelem_total
    return {"value": nelem / (t2 - t0), "unit": "elements/s", "cores": threads, "kind": "port",
            "sample": f"one residual + one tangent of the 16x4x4-element beam (256 elements, FP64, "
                      f"{threads} OpenMP threads): t_R = {t1 - t0:.3f} s, t_K = {t2 - t1:.3f} s"}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only: the whole patch lives in one "
                                   "MI355X's HBM and there is no multi-GPU path; run with --gpus 1"}))
        sys.exit(2)
    import igap4_amd as ig

    nel = tuple(int(x) for x in a.nelem.split(","))
    mesh = ig.MeshIGABeam(L=(4.0, 1.0, 1.0), nelem=nel)
    phys = ig.PhysGradElast(coef=ig.SVK_COEF)
    bc = ig.BCClampedTip(tip=0.05)
    from oracle import oracle as O  # smooth synthetic state only (numpy)
    U = O.smooth_state(O.Mesh(nel, L=(4, 1, 1)), amp=0.01)

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
    traffic = traffic_from_profiles()

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
        "fp32_tangent": {"value": nelem / (ms32 * 1e-3), "ms_per_step": ms32,
                         "tangent_elem_tflops": F_K * nelem / (avg32[1] * 1e-3) / 1e12,
                         "peak": PEAK_FP32_TFLOPS},
    }
    if a.cpu_baseline:
        th = a.cpu_threads or min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(th)
    A.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
