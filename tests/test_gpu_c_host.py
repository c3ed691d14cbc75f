"""The library driven by a plain C host (tests/c_host/igap_c_host.c, built by
igap4_amd/build.sh), the way an IGAP4 driver would call it: descriptors of the
igabeam patch, the residual and FP64 tangent callbacks at a smooth state, then
a Newton load step from zero.  Its results must equal those of the Python
host on the same inputs and match the oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "c_host", "igap_c_host")
NEL = (8, 4, 4)
TIP = 0.05


def rowfloor_err(a, b, rowptr=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if rowptr is None:
        floor = np.maximum(np.abs(b), np.max(np.abs(b)))
    else:
        rn = np.maximum.reduceat(np.abs(b), rowptr[:-1])
        floor = np.maximum(np.abs(b), np.repeat(rn, np.diff(rowptr)))
    floor = np.where(floor == 0, 1.0, floor)
    return float(np.max(np.abs(a - b) / floor))


def read_output(path):
    raw = open(path, "rb").read()
    off = 0

    def take(dtype, n):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=n, offset=off)
        off += a.nbytes
        return a

    ndof, nnz = (int(v) for v in take(np.int64, 2))
    out = {"R": take(np.float64, ndof), "rowptr": take(np.int64, ndof + 1), "col": take(np.int32, nnz),
           "vals": take(np.float64, nnz)}
    out["newton_its"], out["converged"] = (int(v) for v in take(np.int32, 2))
    out["fnorm"] = float(take(np.float64, 1)[0])
    out["U"] = take(np.float64, ndof)
    assert off == len(raw)
    return out


def test_c_host_matches_python_host_and_oracle(igap, oracle, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built: run igap4_amd/build.sh")
    path = str(tmp_path / "c_host.bin")
    print(f"[c host] {NEL}", flush=True)
    r = subprocess.run([EXE, *map(str, NEL), path], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    c = read_output(path)

    mesh = igap.MeshIGABeam(L=(4.0, 1.0, 1.0), nelem=NEL)
    U = igap.smooth_state(mesh, amp=0.01)
    with igap.Assm(mesh, igap.PhysGradElast(coef=igap.SVK_COEF), igap.BCClampedTip(tip=TIP)) as A:
        R = A.residual(U)
        rp, col, K = A.tangent(U, prec="f64")
        A.set_uhist(np.zeros(A.ndof))
        st = A.snes_solve(snes_rtol=1e-10, snes_atol=1e-30)
        Usol = A.get_uhist()
    # the two hosts evaluate the smooth state with different sin/cos
    # implementations, so the inputs agree to rounding, not bit for bit
    assert np.array_equal(c["rowptr"], rp) and np.array_equal(c["col"], col)
    assert rowfloor_err(c["R"], R) < 1e-12
    assert rowfloor_err(c["vals"], K, rp) < 1e-12
    # against the oracle
    m_o = oracle.Mesh(NEL, L=(4, 1, 1))
    ph_o, bc_o = oracle.make_phys(0, coef=oracle.SVK_COEF), oracle.clamped_tip_bc(tip=TIP)
    assert rowfloor_err(c["R"], oracle.residual(m_o, ph_o, bc_o, U)) < 1e-12
    rp_o, col_o, K_o = oracle.tangent(m_o, ph_o, bc_o, U)
    assert np.array_equal(rp_o, c["rowptr"]) and rowfloor_err(c["vals"], K_o, rp_o) < 1e-12
    # the same Newton solve
    assert c["converged"] > 0 and c["newton_its"] == st["newton_its"]
    assert np.linalg.norm(c["U"] - Usol) <= 1e-10 * np.linalg.norm(Usol)
