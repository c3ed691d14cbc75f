"""Benchmark: elements assembled per second (residual + tangent) on one MI355X,
and wall time per Newton load step.

Workload (default): the largest single-GPU patch, 256x64x64 elements of the
quadratic C1 B-spline beam (L = 4:1:1, 1,123,848 control points, 3,371,544
dofs, 1.21G CSR nonzeros), static gradient elasticity with St Venant-Kirchhoff
lambda = mu = 1 and strain-gradient length 0.1, clamped at x = 0 with a tip
displacement on x = L, 4x4x4 Gauss points.  A step is one full residual
(volume + Neumann + Dirichlet rows) plus one full tangent (element tangents on
the matrix cores, deterministic CSR gather, MatZeroRowsColumns) in FP64, on a
smooth synthetic displacement resident in HBM.

The second measure is the wall time per Newton load step of the 40-step ramp
to tip 0.1, solved to ||R|| <= 1e-10 ||R0|| (R0: the residual of the previous
solution under the step's load), for several arms on the same steps and with
the same Krylov stopping rule (the reference's GMRES rtol 1e-5).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--nelem 256,64,64]
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
# floating-point operations the sum-factorised tangent of a separable patch
# (k_tangent_sep) executes per element in its contraction (the pointwise
# tangent rows and the field stage come on top, as for F_K):
#   F_q  54 rows x 64 points x (9 MUL + 27 FMA)
#   z    108 items x 16 pencils x 108 FMA
#   y    162 items x 16 pencils x (36 MUL + 108 FMA)
#   x    162 items x 4 x-slabs x (27 MUL + 135 FMA)
F_SEP = (54 * 64 * (9 + 2 * 27) + 108 * 16 * 2 * 108 + 162 * 16 * (36 + 2 * 108) + 162 * 4 * (27 + 2 * 135))


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nelem", default="256,64,64")
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--newton-steps", type=int, default=2,
                    help="load steps of the Newton ramp timed per arm after the assembly bench (0: skip)")
    ap.add_argument("--ramp-steps", type=int, default=40, help="load steps of the whole ramp")
    ap.add_argument("--ramp-tip", type=float, default=0.1, help="tip displacement at the end of the ramp")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the cpu_baseline sample")
    ap.add_argument("--general-geometry", action="store_true",
                    help="--probe: run the general-geometry kernels (context flag IGAP_GENERAL_GEOMETRY) instead "
                         "of the sum-factorised ones the separable beam takes")
    ap.add_argument("--probe", type=int, default=0,
                    help="profiling run for rocprofv3 --pmc: one 256 MiB device copy (the HBM "
                         "counter calibration) then N x (residual + FP64 and FP32 tangent); prints no JSON")
    ap.add_argument("--pmc-json", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="turn the counter_collection.csv of a FETCH_SIZE and a WRITE_SIZE "
                         "--probe run into profiles/tangent_elem_<nelem>_pmc.json")
    return ap.parse_args()


COPY_BYTES = 1 << 28   # the calibration copy of --probe


def say(msg):
    """Progress on stderr (stdout carries only the JSON line)."""
    print(msg, file=sys.stderr, flush=True)


def probe(nel, calls, general=False):
    """Counter-collection run: one 256 MiB device-to-device copy (the HBM
    counter calibration), then ``calls`` x (residual + FP64 tangent + FP32
    tangent) on the workload (general-geometry kernels with ``general``)."""
    import ctypes as C
    import igap4_amd as ig
    from igap4_amd import _lib
    mesh = ig.MeshIGABeam(L=(4.0, 1.0, 1.0), nelem=nel)
    flags = _lib.IGAP_GENERAL_GEOMETRY if general else 0
    with ig.Assm(mesh, ig.PhysGradElast(coef=ig.SVK_COEF), ig.BCClampedTip(tip=0.05), flags=flags) as A:
        L = _lib.lib()
        src, dst = C.c_void_p(), C.c_void_p()
        _lib.check(L.igap_gpu_malloc(A.ctx, C.byref(src), COPY_BYTES))
        _lib.check(L.igap_gpu_malloc(A.ctx, C.byref(dst), COPY_BYTES))
        _lib.check(L.igap_gpu_memcpy_dtod(A.ctx, dst, src, COPY_BYTES))
        A.synchronize()
        _lib.check(L.igap_gpu_free(A.ctx, src))
        _lib.check(L.igap_gpu_free(A.ctx, dst))
        A.set_uhist(ig.smooth_state(mesh, amp=0.01))
        for _ in range(calls):
            A.residual_dev(None, None)
            A.tangent_dev(None, None, 0)
            A.tangent_dev(None, None, 1)
        A.synchronize()
    say(f"probe done: copy {COPY_BYTES} B, {calls} x (residual + FP64 and FP32 tangents) on {nel}")


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

    out = {"workload": f"{nel[0]}x{nel[1]}x{nel[2]} static beam, FP64 element tangent kernel",
           "copy_bytes": COPY_BYTES, "kernels": {}}
    total = 0.0
    for name, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        d = per_kernel(path, name)
        # the calibration is the single 256 MiB copy dispatch (the largest of the
        # copy kernels; other copies of the run are a few bytes)
        copy = [mx for k, (_, mx) in d.items() if "copy" in k.lower() or "elementwise" in k.lower()]
        # the FP64 element tangent of the probe: the sum-factorised kernel on
        # the (separable) beam, the general-geometry one otherwise
        tang = [mean for k, (mean, _) in d.items() if "k_tangent_sep<double, false>" in k] or \
               [mean for k, (mean, _) in d.items() if "k_tangent_elem<double, 1, false>" in k]
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


def kernel_rooflines(nel, nelem, nnz, avg, avg32, gen, gen32, genr):
    """Every assembly kernel against the ceiling that bounds it.  Tangents:
    the counted F_K per element against the FP64 (FP32) peak -- the
    sum-factorised kernel also with the operations it executes (F_SEP), the
    general-geometry kernel on the matrix cores; residual: F_R per element on
    the FP64 vector ALUs (counted flops only: basis, geometry and the
    pointwise physics come on top); gathers: the element matrices read once
    and the CSR values written once (8 or 4 bytes each), with the CSR values
    alone and the measured HBM bytes beside them."""
    def flop(kernel_ms, f, peak, bound, executed=None):
        a = f * nelem / (kernel_ms * 1e-3) / 1e12
        r = {"bound": bound, "achieved": a, "peak": peak, "unit": "TFLOP/s", "frac": a / peak, "ms": kernel_ms}
        if executed:
            x = executed * nelem / (kernel_ms * 1e-3) / 1e12
            r.update({"executed_flop_per_element": executed, "executed_tflops": x, "executed_frac": x / peak})
        return r

    def hbm(kernel_ms, nbytes, name, csr_bytes):
        a = nbytes / (kernel_ms * 1e-3) / 1e9
        mb = measured_bytes(nel, name)
        return {"bound": "hbm", "achieved": a, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": a / PEAK_HBM_GBS,
                "ms": kernel_ms, "algorithmic_bytes": nbytes, "csr_value_bytes": csr_bytes,
                "csr_value_GBs": csr_bytes / (kernel_ms * 1e-3) / 1e9, "measured_bytes": mb,
                "measured_GBs": None if mb is None else mb / (kernel_ms * 1e-3) / 1e9}

    return {
        "k_tangent_sep<double, false>": flop(avg[1], F_K, PEAK_FP64_TFLOPS, "valu", F_SEP),
        "k_tangent_sep<float, false>": flop(avg32[1], F_K, PEAK_FP32_TFLOPS, "valu", F_SEP),
        "k_tangent_elem<double, 1, false>": flop(gen, F_K, PEAK_FP64_TFLOPS, "mfma"),
        "k_tangent_elem<float, 1, false>": flop(gen32, F_K, PEAK_FP32_TFLOPS, "mfma"),
        "k_residual_sep": flop(avg[0], F_R_STATIC, PEAK_FP64_TFLOPS, "valu"),
        "k_residual_elem<1>": flop(genr, F_R_STATIC, PEAK_FP64_TFLOPS, "valu"),
        "k_tangent_gather_rows<double,double>": hbm(avg[2], 8 * (6561 * nelem + nnz),
                                                    "igap::k_tangent_gather_rows<double, double>", 8 * nnz),
        "k_tangent_gather<float,float>": hbm(avg32[2], 4 * (6561 * nelem + nnz), "igap::k_tangent_gather<float, float>",
                                             4 * nnz),
    }


def host_cpu():
    """The host's CPU model, logical CPU count and the CPUs this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, os.cpu_count() or 1, avail


def cpu_threads():
    """Threads for the CPU baseline: every CPU this process may run on, capped
    by OMP_NUM_THREADS where the host sets it (the CPU share of one GPU slot)."""
    _, _, avail = host_cpu()
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        return min(avail, int(omp))
    return avail


def cpu_baseline(threads, nel, target_s=15.0):
    """The oracle (CPU restatement with the reference's loop nest, the naive
    81x81x100 element contraction and a per-row column search for every
    MatSetValues entry) on a slab of the same workload: one residual and one
    tangent over the elements x < nx of the same patch, physics and state,
    the tangent into the CSR rows of the control points those elements touch.
    Volume integrals only: the boundary terms (face integrals of the whole
    patch, Dirichlet rows) are a per-face cost that a slab does not scale.
    nx is sized from a one-layer calibration run so that the sample takes
    about target_s seconds; elements/s (R+K)."""
    from oracle import oracle as O
    m = O.Mesh(nel, L=(4, 1, 1))
    ph = O.make_phys(0, coef=O.SVK_COEF)
    bc = O.clamped_tip_bc(tip=0.05)
    U = O.smooth_state(m, amp=0.01)
    per_layer = nel[1] * nel[2]

    def run(nx):
        rows = 3 * (nx + 2) * m.nb[1] * m.nb[2]
        pat = O.csr_pattern(m, nrows=rows)
        box = [0, nx, 0, nel[1], 0, nel[2]]
        t0 = time.perf_counter()
        O.residual(m, ph, bc, U, box=box, apply_bc=False, nthreads=threads)
        t1 = time.perf_counter()
        O.tangent(m, ph, bc, U, pattern=pat, box=box, apply_bc=False, nthreads=threads)
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1

    tr, tk = run(1)
    nx = 1
    if tr + tk < 0.5 * target_s:
        nx = int(max(1, min(nel[0], round(target_s / max(1e-3, tr + tk)))))
        if nx > 1:
            tr, tk = run(nx)
    ns = nx * per_layer
    model, ncpu, avail = host_cpu()
    return {"value": ns / (tr + tk), "unit": "elements/s", "cores": threads, "kind": "port",
            "cpu_model": model, "host_logical_cpus": ncpu, "cpus_available": avail,
            "sample": f"one residual + one tangent over the {ns} elements x < {nx} of the {nel[0]}x{nel[1]}x{nel[2]} "
                      f"patch, volume integrals (same physics and state, FP64, {threads} OpenMP threads on {model}; "
                      f"{avail} CPUs available to the process, {ncpu} on the host): t_R = {tr:.2f} s, "
                      f"t_K = {tk:.2f} s"}


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
        probe(nel, a.probe, a.general_geometry)
        return
    import igap4_amd as ig

    mesh = ig.MeshIGABeam(L=(4.0, 1.0, 1.0), nelem=nel)
    phys = ig.PhysGradElast(coef=ig.SVK_COEF)
    bc = ig.BCClampedTip(tip=0.05)
    U = ig.smooth_state(mesh, amp=0.01)

    A = ig.Assm(mesh, phys, bc, device=0)
    A.set_uhist(U)
    nelem = A.nel
    say(f"bench: context of {nelem} elements, {A.nnz} nonzeros")

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
    # the general-geometry element kernels (what a non-separable patch runs)
    # on the same workload, for comparison
    from igap4_amd import _lib
    with ig.Assm(mesh, phys, bc, device=0, flags=_lib.IGAP_GENERAL_GEOMETRY) as Gc:
        Gc.set_uhist(U)
        gen = {}
        for prec in (0, 1):
            Gc.tangent_dev(None, None, prec)
            Gc.synchronize()
            Gc.timing(True)
            Gc.timing_reset()
            for _ in range(3):
                Gc.tangent_dev(None, None, prec)
            gen[prec] = Gc.timing_get(1)[0] / max(1, Gc.timing_get(1)[1])
            Gc.timing(False)
        Gc.residual_dev(None, None)
        Gc.synchronize()
        Gc.timing(True)
        Gc.timing_reset()
        for _ in range(3):
            Gc.residual_dev(None, None)
        gen["residual"] = Gc.timing_get(0)[0] / max(1, Gc.timing_get(0)[1])
        Gc.timing(False)

    line = {
        "metric": "elements assembled/s (R+K); wall time per Newton load step at ||R||/||R0||<=1e-10",
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
                               f"Dirichlet, 4x4x4 Gauss points; load steps: tip {a.ramp_tip} in {a.ramp_steps} "
                               f"steps",
                   "nelem": nelem, "ndof": A.ndof, "nnz": A.nnz},
        # achieved = the counted F_K of the element tangent per element (the
        # block-form contraction) / kernel time, against the FP64 peak; the
        # sum-factorised kernel executes F_SEP of them, so frac can pass 1
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                     "kernel": "k_tangent_sep<double, false>", "kernel_ms": avg[1],
                     "executed_flop_per_element": F_SEP,
                     "executed_tflops": F_SEP * nelem / tk / 1e12},
        "kernels_ms": {"residual_elem": avg[0], "tangent_elem_f64": avg[1], "tangent_gather_f64": avg[2],
                       "residual_call": avg[3], "tangent_call_f64": avg[4],
                       "tangent_elem_f32": avg32[1], "tangent_gather_f32": avg32[2],
                       "tangent_call_f32": avg32[4],
                       "tangent_elem_general_f64": gen[0], "tangent_elem_general_f32": gen[1],
                       "residual_elem_general": gen["residual"]},
        "kernel_rooflines": kernel_rooflines(nel, nelem, A.nnz, avg, avg32, gen[0], gen[1], gen["residual"]),
        "fp32_tangent": {"value": nelem / (ms32 * 1e-3), "ms_per_step": ms32,
                         "tangent_elem_tflops": F_K * nelem / (avg32[1] * 1e-3) / 1e12,
                         "peak": PEAK_FP32_TFLOPS},
    }
    A.close()
    say(f"bench: R+K {line['value']:.4g} elements/s, {ms64:.1f} ms per step")
    # the other scatter (element tangents added straight into the CSR values,
    # no element-matrix buffers): its whole tangent call, for comparison
    with ig.Assm(mesh, phys, bc, device=0, flags=_lib.IGAP_TANGENT_DIRECT) as D:
        D.set_uhist(U)
        for prec, key in ((0, "tangent_call_direct_f64"), (1, "tangent_call_direct_f32")):
            D.tangent_dev(None, None, prec)
            D.synchronize()
            D.timing(True)
            D.timing_reset()
            for _ in range(3):
                D.tangent_dev(None, None, prec)
            ms, n = D.timing_get(4)
            D.timing(False)
            line["kernels_ms"][key] = ms / max(1, n)
    say(f"bench: direct-scatter tangent {line['kernels_ms']['tangent_call_direct_f64']:.1f} ms (FP64)")
    if a.newton_steps > 0:
        line["newton_load_step"] = newton_load_steps(nel, a.newton_steps, nsteps=a.ramp_steps, tip=a.ramp_tip)
    if a.cpu_baseline:
        th = a.cpu_threads or cpu_threads()
        line["cpu_baseline"] = cpu_baseline(th, nel, target_s=a.cpu_seconds)
        say(f"bench: cpu baseline {line['cpu_baseline']['value']:.4g} elements/s")
    print(json.dumps(line), flush=True)


def newton_load_steps(nel, timed=2, nsteps=40, tip=0.1, ramp_in=2, fp32_maxit=15):
    """The second measure of the workload: wall time per Newton load step of
    the clamped-tip ramp (tip ``tip`` in ``nsteps`` load steps), every step
    solved to ||R|| <= 1e-10 ||R0|| with R0 the residual of the previous
    solution under the step's load (the first residual of the reference's
    iteration).  After ``ramp_in`` steps, each of the next ``timed`` steps is
    solved from the same converged history by every arm:

      plain         the previous solution as the initial iterate (the
                    reference's start), FP64 tangent and preconditioner
      quadratic     the warm start without its network: quadratic
                    extrapolation of the last three solutions, accepted
                    against linear extrapolation by residual norm; FP64
      fast          the same start, FP64 tangent, multigrid levels in FP32
                    smoothing with 16-bit copies of their matrices
      fp32_tangent  the same start, Newton on the FP32 matrix-core tangent
                    (the preconditioner as in fast), at most fp32_maxit
                    iterations

    Every arm uses the same Newton test and the same Krylov rule (GMRES(200)
    to the reference's rtol 1e-5) and the same preconditioner (geometric
    multigrid on the B-spline hierarchy, set up every 4th Newton iteration);
    wall time includes building the guess.  The history continues from the
    plain arm's solution.  Displacements of each arm are compared with the
    plain arm's at every timed step."""
    import igap4_amd as ig
    from igap4_amd.loadstep import run_load_steps, tip_bc
    from igap4_amd.warmstart import LearnedGuess

    mesh = ig.MeshIGABeam(L=(4.0, 1.0, 1.0), nelem=nel)
    common = dict(snes_rtol=0.0, snes_atol=1e-30, pc="mg", pc_order=1, pc_lag=4, ksp_rtol=1e-5)
    arms = {
        "plain": dict(tangent_prec="f64", pc_fp32=0),
        "quadratic": dict(tangent_prec="f64", pc_fp32=0),
        "fast": dict(tangent_prec="f64", pc_fp32=2),
        "fp32_tangent": dict(tangent_prec="f32", pc_fp32=2, snes_maxit=fp32_maxit),
    }
    out = {"ramp": f"tip {tip} in {nsteps} load steps", "criterion": "||R|| <= 1e-10 ||R(U_prev; load_k)||",
           "krylov": "GMRES(200), right preconditioned, rtol 1e-5 (every arm)",
           "preconditioner": "geometric multigrid on the B-spline hierarchy (Chebyshev/point-block-Jacobi "
                             "smoothing, dense coarsest inverse), set up every 4th Newton iteration; FP64 levels "
                             "in plain and quadratic, FP32 levels smoothing with 16-bit matrices in fast and "
                             "fp32_tangent",
           "arms": {k: {"ms_per_step": [], "newton_its": [], "ksp_its": [], "converged": [],
                        "final_rel_residual": [], "max_rel_diff_vs_plain": []} for k in arms}}
    with ig.Assm(mesh, ig.PhysGradElast(coef=ig.SVK_COEF), ig.BCClampedTip(tip=0.0)) as A:
        n = A.ndof
        A.set_uhist(np.zeros(n))
        prev, G = [np.zeros(n)], [A.dirichlet()[1]]
        opts = {k: ig.assm.default_solver_opts(**common, **v) for k, v in arms.items()}
        ramp = []

        def keep(k, u, st):
            prev.append(u.copy())
            G.append(A.dirichlet()[1])

        t0 = time.perf_counter()
        for k in range(1, ramp_in + 1):
            st = run_load_steps(A, nsteps, tip_total=tip, guess="extrapolate", opts=opts["fast"], start=k,
                                max_steps=k, prev=prev[-3:], start_rtol=1e-10, callback=keep)
            ramp.append({"step": k, "ms": st[0]["wall_ms"], "newton_its": st[0]["newton_its"],
                         "ksp_its": st[0]["ksp_its"], "converged": st[0]["converged"]})
            say(f"bench: ramp-in step {k}: {st[0]['wall_ms']:.0f} ms, newton {st[0]['newton_its']}, "
                f"ksp {st[0]['ksp_its']}")
            if st[0]["converged"] <= 0:
                out["ramp_in"] = ramp
                out["error"] = f"ramp-in step {k} did not converge"
                return out
        out["ramp_in"] = ramp
        out["ramp_in_s"] = time.perf_counter() - t0
        for k in range(ramp_in + 1, ramp_in + timed + 1):
            sols = {}
            for name in arms:
                if name == "plain":
                    guess = "previous"
                else:
                    guess = LearnedGuess(A, mode="quadratic")
                    guess.set_previous_load(G[-1])
                got = []
                st = run_load_steps(A, nsteps, tip_total=tip, guess=guess, opts=opts[name], start=k, max_steps=k,
                                    prev=prev[-3:], start_rtol=1e-10,
                                    callback=lambda kk, u, x: got.append(u.copy()))
                x = st[0]
                sols[name] = got[0] if got else A.get_uhist()[:n].copy()
                r = out["arms"][name]
                r["ms_per_step"].append(x["wall_ms"])
                r["newton_its"].append(x["newton_its"])
                r["ksp_its"].append(x["ksp_its"])
                r["converged"].append(x["converged"])
                r["final_rel_residual"].append(x["fnorm"] / x["r0_start"] if x["r0_start"] > 0 else 0.0)
                say(f"bench: step {k} {name}: {x['wall_ms']:.0f} ms, newton {x['newton_its']}, ksp {x['ksp_its']}, "
                    f"converged {x['converged']}")
            ref = sols["plain"]
            scale = max(1e-300, float(np.max(np.abs(ref))))
            for name in arms:
                out["arms"][name]["max_rel_diff_vs_plain"].append(float(np.max(np.abs(sols[name] - ref)) / scale))
            prev.append(ref)
            A.set_bc(tip_bc(tip, k / nsteps))
            G.append(A.dirichlet()[1])
    for name, r in out["arms"].items():
        r["ms_per_step_median"] = float(np.median(r["ms_per_step"])) if r["ms_per_step"] else None
        r["all_converged"] = bool(r["converged"] and all(c > 0 for c in r["converged"]))
    p = out["arms"]["plain"]["ms_per_step_median"]
    for name, r in out["arms"].items():
        r["speedup_vs_plain"] = (p / r["ms_per_step_median"]) if p and r["ms_per_step_median"] else None
    out["steps_timed"] = list(range(ramp_in + 1, ramp_in + timed + 1))
    return out


if __name__ == "__main__":
    main()
